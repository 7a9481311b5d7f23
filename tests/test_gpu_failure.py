"""Failure paths on the GPU, each forced from the host with inputs every
kernel handles in bounds: a Broyden loop stopped at broyden_maxiter, an
exploding residual, non-finite and far-off coordinates through the FCM and
the barrier, a non-finite state reaching the Jacobian inverse (both
Gauss-Jordan kernels), and filaments too long for the register-resident
inverse against the oracle."""
import numpy as np
import pytest

import _oracle as O
from filsim_amd import api, params
from helpers import random_segments, rel, to_dev
from sysfix import config, pair

pytestmark = pytest.mark.gpu


def test_broyden_cap_matches_oracle(cuda):
    """TOL 1e-14 cannot be met in two iterations: both stop at
    broyden_maxiter and commit the step (main.cpp:503, 676-679); the GPU
    reports the step as not converged."""
    sp, nws, props, st = config(grid=48, nfil=5, nw=8, swim=True, seed=4,
                                tol=1e-14, maxiter=2)
    o, g = pair(sp, nws, props, st)
    Lb = np.abs(st["X"]).max()
    for nt in (1, 2, 3):
        io, ig = o.timestep(nt), g.timestep(nt)
        print("nt", nt, "iterations oracle", io, "gpu", ig, flush=True)
        assert io == ig == 2
        s = g.step_status()
        assert not s["converged"] and not s["exploded"]
        assert not s["nonfinite"] and s["res_inf"] > 1e-14
    so, sg = o.get_state(), g.get_state()
    assert np.abs(sg["X"] - so["X"]).max() <= 1e-9 * Lb
    g.close()


def test_converged_step_status(cuda):
    sp, nws, props, st = config(grid=48, nfil=5, nw=8, swim=True, seed=4,
                                tol=1e-8)
    g = api.System(sp, nws, props)
    g.set_state(**st)
    it = g.timestep(1)
    s = g.step_status()
    assert 0 < it < sp.broyden_maxiter
    assert s["converged"] and not s["exploded"] and not s["nonfinite"]
    assert s["res_inf"] <= 1e-8
    g.close()


def test_exploding_residual_is_flagged(cuda):
    """A gravity of 1e10 per unit length moves the extrapolated guess far off:
    the first residual exceeds 1e6, which the reference reports as "Prepare
    for explosion" (MobilityConstraints.cpp:764-767)."""
    sp, nws, props, st = config(grid=48, nfil=4, nw=8, seed=3, tol=1e-6,
                                maxiter=3, weight=(0.0, 0.0, -1e10))
    g = api.System(sp, nws, props)
    g.set_state(**st)
    g.timestep(1)
    s = g.step_status()
    assert s["exploded"]
    g.close()


def test_nonfinite_and_far_coordinates_through_fcm(cuda):
    """A NaN coordinate takes stencil cell = segment index (no conversion out
    of range) and its NaN weights reach every grid point through the FFT;
    the solve returns normally. A coordinate 1e20 out has all-zero Gaussian
    weights: it adds nothing to the grid and receives V = W = 0, and every
    other segment moves as if it had no force."""
    for prec, tol in ((0, 1e-12), (1, 1e-5)):
        fp = params.fcm_params(48)
        Y, F, T = random_segments(fp, 60, seed=11)
        with api.Fcm(fp, 60, precision=prec) as fcm:
            Yn = Y.copy()
            Yn[7, 1] = np.nan
            V, W = fcm.mobility_host(Yn, F, T)
            assert np.isnan(V).all() and np.isnan(W).all()
            Yf = Y.copy()
            Yf[7] = [1e20, -1e20, 1e20]
            V, W = fcm.mobility_host(Yf, F, T)
            F0, T0 = F.copy(), T.copy()
            F0[7] = 0
            T0[7] = 0
            Vr, Wr = fcm.mobility_host(Y, F0, T0)
        assert np.all(V[7] == 0) and np.all(W[7] == 0)
        keep = np.arange(60) != 7
        assert rel(V[keep], Vr[keep]) <= tol
        assert rel(W[keep], Wr[keep]) <= tol
        print("precision", prec, "far segment: V", rel(V[keep], Vr[keep]),
              flush=True)


def test_nonfinite_coordinate_through_barrier(cuda):
    """A NaN segment is binned into cell 0 of its axes and forms no pair:
    the pair set equals the brute-force set (which drops it by the same
    comparison) and every other segment's barrier force is bit for bit the
    force with that segment parked far from all others."""
    torch = cuda
    fp = params.fcm_params(64)
    cp = params.colbar_params(fp)
    nfil, nw = 16, 15
    X, q = api.make_filaments(params.box(fp), [nw] * nfil, dL=2.2,
                              min_sep=1.9, seed=80)
    sf = np.repeat(np.arange(nfil), nw).astype(np.int32)
    si = np.tile(np.arange(nw), nfil).astype(np.int32)
    rng = np.random.default_rng(1)
    Xs = X + rng.normal(0, 0.6, X.shape)
    L = np.array(params.box(fp))
    k = 37
    # a parking spot at least 6a from every other segment (periodic)
    for _ in range(10000):
        p = rng.uniform(0, 1, 3) * L
        d = Xs - p
        d -= L * np.floor(d / L + 0.5)
        dist = np.linalg.norm(d, axis=1)
        dist[k] = np.inf
        if dist.min() > 6.0:
            break
    else:
        pytest.skip("no free spot")
    Xnan, Xpark = Xs.copy(), Xs.copy()
    Xnan[k] = [np.nan, 1.0, np.nan]
    Xpark[k] = p
    Ynan, Ypark = np.fmod(Xnan, L), np.fmod(Xpark, L)
    nt = 5
    gb = api.Colbar(cp, sf, si)
    F0 = rng.standard_normal(Xs.shape)
    out = []
    for Xa, Ya in ((Xnan, Ynan), (Xpark, Ypark)):
        gb.link(to_dev(torch, Xa))
        pg = gb.pairs(to_dev(torch, Ya), nt)
        pb = O.pairs_brute(cp, Ya, sf, si, nt)
        assert np.array_equal(pg, pb)
        assert not np.any(pg == k)
        dF = to_dev(torch, F0)
        md = gb.apply(to_dev(torch, Ya), nt, dF)
        assert np.isfinite(md)
        out.append((dF.cpu().numpy(), len(pg)))
    (Fn, npn), (Fp, npp) = out
    assert npn == npp > 0
    assert np.array_equal(Fn, Fp)
    assert np.array_equal(Fn[k], F0[k])
    gb.close()


@pytest.mark.parametrize("nw", [15, 34])
def test_nonfinite_state_flags_singular_jacobian(cuda, nw):
    """A NaN bead orientation makes the approximate Jacobian non-finite. Both
    Gauss-Jordan kernels (register-resident for 6 Nw <= 192, the global
    memory one above) keep every index in range and raise the singular
    flag, so the timestep stops with an error instead of committing NaN.
    The same system then steps a clean state exactly like a fresh one."""
    sp, nws, props, st = config(grid=160, nfil=3, nw=nw, seed=5, tol=1e-8,
                                ifs=1000)
    g = api.System(sp, nws, props)
    bad = {k: v.copy() for k, v in st.items()}
    bad["q"][nw + 3] = np.nan
    bad["qt"][nw + 3] = np.nan
    g.set_state(**bad)
    with pytest.raises(Exception, match="singular"):
        g.timestep(1)
    g.set_state(**st)
    ia = g.timestep(1)
    ref = api.System(sp, nws, props)
    ref.set_state(**st)
    ib = ref.timestep(1)
    assert ia == ib
    assert np.array_equal(g.get_state()["X"], ref.get_state()["X"])
    g.close()
    ref.close()


@pytest.mark.parametrize("grid,nw", [(160, 34), (270, 66)])
def test_long_filaments_jacobian_and_step_match_oracle(cuda, grid, nw):
    """Long filaments against the oracle: J, its inverse and timesteps under
    the friction mobility. 34 segments (J is 204 x 204) use the
    block-tridiagonal factorisation with a partly filled last block; 66
    segments (396 x 396) are past its 64-segment limit and take the dense
    Gauss-Jordan inverse in global memory."""
    sp, nws, props, st = config(grid=grid, nfil=3, nw=nw, seed=5, tol=1e-9,
                                ifs=1000, swim=True)
    o, g = pair(sp, nws, props, st)
    nt = 14
    for obj in (o, g):
        obj.initial_guess()
        obj.forces(nt)
        obj.build_jacobian(nt)
    for f in range(3):
        Jo, Jg = o.jacobian(f), g.jacobian(f)
        assert rel(Jg, Jo) <= 1e-14, f
        Io, Ig = o.jinv(f), g.jinv(f)
        assert rel(Ig, Io) <= 1e-11, f
    o2, g2 = pair(sp, nws, props, st)
    Lb = np.abs(st["X"]).max()
    for nt in (1, 2):
        io, ig = o2.timestep(nt), g2.timestep(nt)
        print("nt", nt, "iterations oracle", io, "gpu", ig, flush=True)
        assert abs(io - ig) <= 1
    assert np.abs(g2.get_state()["X"] - o2.get_state()["X"]).max() <= 1e-8 * Lb
    g.close()
    g2.close()
