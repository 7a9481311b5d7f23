"""Filament mechanics, residual, Jacobian, Broyden and whole timesteps on
the GPU against the FP64 oracle, stage by stage on identical states."""
import numpy as np
import pytest

from helpers import rel, to_dev
from sysfix import config, d2h, pair, rich_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def warm():
    sp, nws, props, st0 = config(swim=True, seed=5)
    return sp, nws, props, rich_state(st0, nws, seed=1)


def _close(a, b, tol):
    return rel(a, b) <= tol


def test_guess_forces_assign(cuda, warm):
    sp, nws, props, st = warm
    o, g = pair(sp, nws, props, st)
    for obj in (o, g):
        obj.initial_guess()
        obj.forces(14)
    so, sg = o.get_state(), g.get_state()
    for k in ("Xs", "Us", "qs", "F", "TAU"):
        assert _close(sg[k], so[k], 1e-14), k
    Yo, Fo, To, _, _ = o.fcm_arrays()
    Yp, Fp, Tp, _, _ = g.fcm_ptrs()
    n = 3 * g.np
    assert _close(d2h(Yp, n), Yo.ravel(), 1e-15)
    assert _close(d2h(Fp, n), Fo.ravel(), 1e-14)
    assert _close(d2h(Tp, n), To.ravel(), 1e-14)


def test_jacobian_and_inverse(cuda, warm):
    sp, nws, props, st = warm
    o, g = pair(sp, nws, props, st)
    for obj in (o, g):
        obj.initial_guess()
        obj.forces(14)
        obj.build_jacobian(14)
    for f in (0, 7, len(nws) - 1):
        Jo, Jg = o.jacobian(f), g.jacobian(f)
        assert _close(Jg, Jo, 1e-14), f
        Io, Ig = o.jinv(f), g.jinv(f)
        assert _close(Ig, Io, 1e-11), f
        # inverse property on the GPU result (lambda rows carry 1/dt)
        D = np.ones(Jg.shape[0])
        D[3 * nws[f] + 3:] = sp.dt
        assert np.abs(np.diag(D) @ Ig @ Jg - np.eye(len(D))).max() < 1e-9


def test_mobility_residual_broyden(cuda, warm):
    torch = cuda
    sp, nws, props, st = warm
    o, g = pair(sp, nws, props, st)
    nt = 14
    for obj in (o, g):
        obj.initial_guess()
        obj.forces(nt)
        obj.build_jacobian(nt)
        obj.barrier(nt, 1)
        obj.mobility(nt)
    _, _, _, Vo, Wo = o.fcm_arrays()
    _, _, _, Vp, Wp = g.fcm_ptrs()
    n = 3 * g.np
    assert _close(d2h(Vp, n), Vo.ravel(), 1e-12)
    assert _close(d2h(Wp, n), Wo.ravel(), 1e-12)
    Eo, chk_o, mx_o = o.residual(nt, Vo, Wo)
    Vd = to_dev(torch, d2h(Vp, n))
    Wd = to_dev(torch, d2h(Wp, n))
    E = torch.zeros(6 * g.np, dtype=torch.float64, device="cuda")
    chk_g, mx_g = g.residual(nt, Vd, Wd, E)
    assert chk_g == chk_o == 1
    assert abs(mx_g - mx_o) <= 1e-12 * mx_o
    assert _close(E.cpu().numpy(), Eo, 1e-12)
    # first Broyden iteration: solve, update, re-evaluate, secant column
    Uo = o.jsolve(Eo, 0)
    U = torch.zeros_like(E)
    g.jsolve(E, U, 0)
    assert _close(U.cpu().numpy(), Uo, 1e-11)
    o.update(Uo)
    g.update(to_dev(torch, Uo))
    for obj in (o, g):
        obj.forces(nt)
        obj.barrier(nt, 0)
        obj.mobility(nt)
    _, _, _, Vo, Wo = o.fcm_arrays()
    Eno, _, _ = o.residual(nt, Vo, Wo)
    En = torch.zeros_like(E)
    g.residual(nt, to_dev(torch, d2h(Vp, n)), to_dev(torch, d2h(Wp, n)), En)
    assert _close(En.cpu().numpy(), Eno, 1e-11)
    o.broyden_update(Eo, Eno, Uo, 0)
    g.broyden_update(to_dev(torch, Eo), to_dev(torch, Eno), to_dev(torch, Uo),
                     0)
    x_o = o.jsolve(Eno, 1)
    x = torch.zeros_like(E)
    g.jsolve(to_dev(torch, Eno), x, 1)
    assert _close(x.cpu().numpy(), x_o, 1e-11)
    # secant condition of the updated inverse: H1 dE = -update
    dE = Eno - Eo
    h = o.jsolve(dE, 1)
    assert _close(h, -Uo, 1e-9)


def test_step(cuda, warm):
    sp, nws, props, st = warm
    o, g = pair(sp, nws, props, st)
    for obj in (o, g):
        obj.initial_guess()
        obj.step(14)
    so, sg = o.get_state(), g.get_state()
    for k in ("X", "Xt", "U", "Ut", "q", "qt", "lam", "lam1", "lam2"):
        assert _close(sg[k], so[k], 1e-14), k


def test_whole_timesteps_friction(cuda):
    """Broyden machinery over 16 steps (BDF2 switch at nt=2, lambda
    extrapolation switches at nt=12, 13) with the friction mobility."""
    sp, nws, props, st = config(swim=True, seed=8, tol=1e-9, ifs=100)
    o, g = pair(sp, nws, props, st)
    for nt in range(1, 17):
        io = o.timestep(nt)
        ig = g.timestep(nt)
        assert abs(io - ig) <= 1, (nt, io, ig)
    so, sg = o.get_state(), g.get_state()
    Lb = np.abs(so["X"]).max()
    assert np.abs(sg["X"] - so["X"]).max() <= 1e-8 * Lb
    assert np.abs(sg["q"] - so["q"]).max() <= 1e-8
    # the residual pins lambda only to ~ TOL * 9 pi mu a / dt (the BlockD
    # sensitivity, Filament.hpp:913-946)
    assert np.abs(sg["lam"] - so["lam"]).max() <= 30 * sp.tol / sp.dt


def test_whole_timesteps_fcm(cuda):
    sp, nws, props, st = config(grid=48, nfil=6, nw=10, swim=True, seed=8,
                                tol=1e-8)
    o, g = pair(sp, nws, props, st)
    for nt in range(1, 3):
        io = o.timestep(nt)
        ig = g.timestep(nt)
        assert abs(io - ig) <= 1, (nt, io, ig)
    so, sg = o.get_state(), g.get_state()
    Lb = np.abs(so["X"]).max()
    assert np.abs(sg["X"] - so["X"]).max() <= 1e-8 * Lb
    assert np.abs(sg["q"] - so["q"]).max() <= 1e-8
