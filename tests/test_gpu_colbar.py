"""Collision barrier on the GPU: cell lists bit-exact, pair sets exact,
forces against the oracle (which replays the reference's traversal)."""
import numpy as np
import pytest

import _oracle as O
from filsim_amd import api, params
from helpers import rel, to_dev

pytestmark = pytest.mark.gpu


def _system(fp, nfil, nw, seed, min_sep=1.9):
    box = params.box(fp)
    X, q = api.make_filaments(box, [nw] * nfil, dL=2.2, min_sep=min_sep,
                              seed=seed)
    sf = np.repeat(np.arange(nfil), nw).astype(np.int32)
    si = np.tile(np.arange(nw), nfil).astype(np.int32)
    return X, sf, si


@pytest.mark.parametrize("grid,nfil,nw,nt,ifs", [
    (64, 16, 15, 5, 0), (64, 40, 15, 1, 2), (96, 120, 20, 3, 0)])
def test_lists_pairs_forces(cuda, grid, nfil, nw, nt, ifs):
    torch = cuda
    fp = params.fcm_params(grid)
    cp = params.colbar_params(fp, initial_friction_steps=ifs)
    X, sf, si = _system(fp, nfil, nw, seed=grid + nfil)
    # perturb so that some pairs sit inside the cutoff and under the cap
    rng = np.random.default_rng(1)
    Xs = X + rng.normal(0, 0.6, X.shape)
    Y = np.fmod(Xs, np.array(params.box(fp)))
    ob = O.Colbar(cp, sf, si)
    ob.link(Xs)
    h_o, l_o = ob.lists()
    gb = api.Colbar(cp, sf, si)
    dXs, dY = to_dev(torch, Xs), to_dev(torch, Y)
    gb.link(dXs)
    h_g, l_g = gb.lists()
    assert np.array_equal(h_g, h_o) and np.array_equal(l_g, l_o)
    p_o = ob.pairs(Y, nt)
    p_g = gb.pairs(dY, nt)
    p_b = O.pairs_brute(cp, Y, sf, si, nt)
    assert len(p_o) > 0
    assert np.array_equal(p_g, p_o) and np.array_equal(p_o, p_b)
    F0 = rng.standard_normal(Y.shape)
    Fo, md_o = ob.apply(Y, nt, F0)
    dF = to_dev(torch, F0)
    md_g = gb.apply(dY, nt, dF)
    Fg = dF.cpu().numpy()
    assert md_g == md_o
    assert rel(Fg - F0, Fo - F0) <= 1e-12
    # Newton's third law on the barrier part
    assert np.abs((Fg - F0).sum(0)).max() <= 1e-9 * np.abs(Fg - F0).max()
    gb.close()


def test_more_than_a_million_cells(cuda):
    """A 512^3 box has 110^3 barrier cells: the cell-count scan runs over
    more than 1024 blocks of 1024. A dense cluster (a 64^3 box's worth of
    filaments) sits inside it so pairs exist."""
    torch = cuda
    fp = params.fcm_params(512)
    cp = params.colbar_params(fp)
    X, sf, si = _system(params.fcm_params(64), 40, 15, seed=3)
    rng = np.random.default_rng(2)
    Xs = X + rng.normal(0, 0.6, X.shape) + 100.0
    Y = np.fmod(Xs, np.array(params.box(fp)))
    ob = O.Colbar(cp, sf, si)
    ob.link(Xs)
    h_o, l_o = ob.lists()
    gb = api.Colbar(cp, sf, si)
    gb.link(to_dev(torch, Xs))
    h_g, l_g = gb.lists()
    assert len(h_o) > 1024 * 1024
    assert np.array_equal(h_g, h_o) and np.array_equal(l_g, l_o)
    p_o = ob.pairs(Y, 1)
    assert len(p_o) > 0
    assert np.array_equal(gb.pairs(to_dev(torch, Y), 1), p_o)
    gb.close()


def test_stale_list(cuda):
    """Between links the list is reused even after segments move
    (main.cpp:394 vs :565)."""
    torch = cuda
    fp = params.fcm_params(64)
    cp = params.colbar_params(fp)
    X, sf, si = _system(fp, 20, 15, seed=4)
    Y0 = np.fmod(X, np.array(params.box(fp)))
    ob = O.Colbar(cp, sf, si)
    gb = api.Colbar(cp, sf, si)
    ob.link(X)
    gb.link(to_dev(torch, X))
    Y1 = Y0 + np.random.default_rng(2).normal(0, 1.0, Y0.shape)
    assert np.array_equal(gb.pairs(to_dev(torch, Y1), 5), ob.pairs(Y1, 5))
    gb.close()


@pytest.mark.parametrize("mask", [1, 2])
def test_masked_variants(cuda, mask):
    """apply_collision_barrier_not_swimmer (mask 1) and
    _not_test_particles (mask 2) against the oracle: pair sets exact, forces
    to 1e-12, minimum distance exact; the mask removes pairs inside the
    cutoff (filament 0 beside filament 1, the last two segments beside
    segments 3 and 14)."""
    from test_oracle_colbar import _crowded, _kept
    torch = cuda
    fp, cp, X, sf, si = _crowded()
    sf, si = sf.astype(np.int32), si.astype(np.int32)
    Y = np.fmod(X, np.array(params.box(fp)))
    nt = 3
    allp = O.pairs_brute(cp, Y, sf, si, nt)
    keep = _kept(mask, allp, sf)
    assert 0 < keep.sum() < len(allp)
    ob = O.Colbar(cp, sf, si)
    ob.link(X)
    ob.set_mask(mask)
    gb = api.Colbar(cp, sf, si)
    dX, dY = to_dev(torch, X), to_dev(torch, Y)
    gb.link(dX)
    gb.set_mask(mask)
    p_g = gb.pairs(dY, nt)
    assert np.array_equal(p_g, ob.pairs(Y, nt))
    assert np.array_equal(p_g, allp[keep])
    F0 = np.random.default_rng(2).standard_normal(Y.shape)
    Fo, md_o = ob.apply(Y, nt, F0)
    dF = to_dev(torch, F0)
    md_g = gb.apply(dY, nt, dF)
    Fg = dF.cpu().numpy()
    assert md_g == md_o
    assert rel(Fg - F0, Fo - F0) <= 1e-12
    with pytest.raises(Exception):
        gb.set_mask(3)
    gb.close()
