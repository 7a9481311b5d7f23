"""Pins the oracle barrier: linked-list pairs = brute force, Newton's third
law, closed-form pair force on both sides of the cap, LIFO lists."""
import numpy as np
import pytest

import _oracle as O
from filsim_amd import abi, api, params
from helpers import rel

PI = 3.14159265358979323846


def _setup(grid, nfil, nw, seed):
    fp = params.fcm_params(grid)
    cp = params.colbar_params(fp)
    X, _ = api.make_filaments(params.box(fp), [nw] * nfil, min_sep=1.9,
                              seed=seed)
    X = X + np.random.default_rng(seed).normal(0, 0.6, X.shape)
    sf = np.repeat(np.arange(nfil), nw)
    si = np.tile(np.arange(nw), nfil)
    return fp, cp, X, sf, si


@pytest.mark.parametrize("grid,nfil,nw,nt", [(64, 16, 15, 3), (48, 10, 8, 0),
                                             (96, 60, 20, 2)])
def test_pairs_match_brute_force(grid, nfil, nw, nt):
    fp, cp, X, sf, si = _setup(grid, nfil, nw, grid)
    Y = np.fmod(X, np.array(params.box(fp)))
    cb = O.Colbar(cp, sf, si)
    cb.link(X)
    p1 = cb.pairs(Y, nt)
    p2 = O.pairs_brute(cp, Y, sf, si, nt)
    assert len(p1) > 0 and np.array_equal(p1, p2)


def test_lifo_lists():
    fp, cp, X, sf, si = _setup(64, 16, 15, 1)
    cb = O.Colbar(cp, sf, si)
    cb.link(X)
    head, lst = cb.lists()
    seen = []
    for c in range(len(head)):
        j = head[c]
        prev = None
        while j >= 0:
            if prev is not None:
                assert j < prev  # LIFO: later segments first
            seen.append(j)
            prev = j
            j = lst[j]
    assert sorted(seen) == list(range(len(sf)))


def test_newton_third_law():
    fp, cp, X, sf, si = _setup(64, 30, 15, 2)
    Y = np.fmod(X, np.array(params.box(fp)))
    cb = O.Colbar(cp, sf, si)
    cb.link(X)
    F, md = cb.apply(Y, 3, np.zeros_like(Y))
    assert np.abs(F).max() > 0
    assert np.abs(F.sum(0)).max() <= 1e-10 * np.abs(F).max()
    assert md < (1.1 * 2) ** 2


@pytest.mark.parametrize("r", [1.0, 1.5, 1.99, 2.05, 2.15, 2.3])
def test_closed_form_pair(r):
    """F = FS/2a ((chi R^2 - r^2)/(chi R^2 - R^2))^4 r_vec, R = 2a,
    chi = 1.21, r capped below Barrier_distance_cap
    (CollisionBarrierFilament.cpp:109-128)."""
    fp = params.fcm_params(64)
    cp = params.colbar_params(fp)
    d = np.array([0.6, -0.48, 0.64])
    Y = np.array([[20.0, 20.0, 20.0], [20.0, 20.0, 20.0]])
    Y[1] += r * d
    cb = O.Colbar(cp, [0, 1], [0, 0])
    cb.link(Y)
    F, md = cb.apply(Y, 3, np.zeros_like(Y))
    R2 = 4.0
    chi = 1.21 * R2
    rr = max(r, cp.cap)
    expect = 0.0
    if r * r < chi:
        expect = 15 * PI * ((chi - rr * rr) / (chi - R2)) ** 4 * rr
    assert np.allclose(F[0], -expect * d, rtol=1e-13, atol=1e-300)
    assert np.allclose(F[1], expect * d, rtol=1e-13, atol=1e-300)
    assert np.isclose(md, r * r, rtol=1e-14)


def test_adjacent_exclusion_and_friction_steps():
    """Neighbours in one filament are skipped unless nt < initialFrictionSteps
    (CollisionBarrierFilament.cpp:80)."""
    fp = params.fcm_params(64)
    Y = np.array([[20.0, 20, 20], [21.5, 20, 20], [23.0, 20, 20]])
    for ifs, nt, expect in [(0, 1, 0), (5, 1, 2), (5, 5, 0)]:
        cp = params.colbar_params(fp, initial_friction_steps=ifs)
        cb = O.Colbar(cp, [0, 0, 0], [0, 1, 2])
        cb.link(Y)
        assert len(cb.pairs(Y, nt)) == expect


def _crowded(nfil=24, nw=10, seed=7):
    """Filaments of a 48^3 box with filament 0 laid 1.5 beside filament 1
    and the last two segments moved next to segments 3 and 14, so that both
    masks remove pairs inside the cutoff."""
    fp, cp, X, sf, si = _setup(48, nfil, nw, seed)
    X[:nw] = X[nw:2 * nw] + np.array([1.5, 0.0, 0.0])
    X[-2] = X[3] + np.array([0.0, 1.6, 0.0])
    X[-1] = X[14] + np.array([0.0, 0.0, 1.7])
    return fp, cp, X, sf, si


def _kept(mask, pairs, sf):
    n = len(sf)
    if mask == abi.FS_BARRIER_NOT_SWIMMER:
        return (sf[pairs[:, 0]] == 0) == (sf[pairs[:, 1]] == 0)
    return (pairs[:, 0] < n - 2) & (pairs[:, 1] < n - 2)


@pytest.mark.parametrize("mask", [abi.FS_BARRIER_NOT_SWIMMER,
                                  abi.FS_BARRIER_NOT_TEST_PARTICLES])
def test_masked_variants(mask):
    """apply_collision_barrier_not_swimmer / _not_test_particles
    (CollisionBarrierFilament.cpp:441-945): the masked pair set is the
    brute-force set filtered by the mask's rule, the forces are the
    closed-form pair forces summed over the kept pairs only, and the minimum
    distance is taken over the kept pairs."""
    fp, cp, X, sf, si = _crowded()
    box = np.array(params.box(fp))
    Y = np.fmod(X, box)
    nt = 3
    allp = O.pairs_brute(cp, Y, sf, si, nt)
    keep = _kept(mask, allp, sf)
    assert 0 < keep.sum() < len(allp)
    cb = O.Colbar(cp, sf, si)
    cb.link(X)
    cb.set_mask(mask)
    assert np.array_equal(cb.pairs(Y, nt), allp[keep])
    F, md = cb.apply(Y, nt, np.zeros_like(Y))
    R2 = (2 * cp.a) ** 2
    chi = 1.21 * R2
    Fe = np.zeros_like(Y)
    r2min = np.inf
    for i, j in allp[keep]:
        d = Y[i] - Y[j]
        d = d - box * np.floor(d / box + 0.5)
        r2 = float(d @ d)
        r2min = min(r2min, r2)
        if r2 < cp.cap ** 2:
            d = d / np.sqrt(r2) * cp.cap
            r2 = cp.cap ** 2
        f = cp.fs_over_2a * ((chi - r2) / (chi - R2)) ** 4 * d
        Fe[i] += f
        Fe[j] -= f
    assert rel(F, Fe) <= 1e-12
    assert np.isclose(md, r2min, rtol=1e-14)
    # the full barrier sees the removed pairs
    cb.set_mask(abi.FS_BARRIER_ALL)
    F_all, md_all = cb.apply(Y, nt, np.zeros_like(Y))
    assert rel(F_all, F) > 1e-3 and md_all <= md
