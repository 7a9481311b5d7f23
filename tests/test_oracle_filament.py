"""Pins the oracle filament mechanics, residual, Jacobian and Broyden."""
import numpy as np
import pytest

import _oracle as O
from filsim_amd import params
from sysfix import config, rich_state

PI = 3.14159265358979323846


def rotmat(q):
    s, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - s * z), 2 * (x * z + s * y)],
        [2 * (x * y + s * z), 1 - 2 * (x * x + z * z), 2 * (y * z - s * x)],
        [2 * (x * z - s * y), 2 * (y * z + s * x), 1 - 2 * (x * x + y * y)]])


def qmul(p, q):
    return np.concatenate([[p[0] * q[0] - p[1:] @ q[1:]],
                           p[0] * q[1:] + q[0] * p[1:] + np.cross(p[1:], q[1:])])


def unit_q(rng):
    q = rng.standard_normal(4)
    return q / np.linalg.norm(q)


def test_quaternion_identities():
    rng = np.random.default_rng(0)
    L = O.lib()
    for _ in range(20):
        q = unit_q(rng)
        v = rng.standard_normal(3) * 0.7
        # expm: q <- exp(v) q = (cos |v|/2, sin(|v|/2) v/|v|) * q
        qe = q.copy()
        L.or_quat_expm(O.dp(qe), O.dp(v))
        th = np.linalg.norm(v)
        ex = np.concatenate([[np.cos(th / 2)], np.sin(th / 2) * v / th])
        assert np.allclose(qe, qmul(ex, q), atol=1e-15)
        assert abs(np.linalg.norm(qe) - 1) < 1e-14
        # rotation by exp(v) = Rodrigues rotation about v by |v|
        k = v / th
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        Rv = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
        assert np.allclose(rotmat(qe), Rv @ rotmat(q), atol=1e-14)
        # tangent = R(q) e_x; RotateVector = R(q) v
        t = np.zeros(3)
        L.or_quat_tangent(O.dp(q), O.dp(t))
        assert np.allclose(t, rotmat(q)[:, 0], atol=1e-15)
        w = rng.standard_normal(3)
        w2 = w.copy()
        L.or_quat_rotate(O.dp(q), O.dp(w2))
        assert np.allclose(w2, rotmat(q) @ w, atol=1e-14)
        # midpoint: m^2 relates p1 to p2 (m p1^* m = p2 up to sign)
        p1 = unit_q(rng)
        p2 = qmul(ex, p1)
        m = np.zeros(4)
        L.or_quat_midpoint(O.dp(p1), O.dp(p2), O.dp(m))
        rel_half = qmul(m, np.concatenate([[p1[0]], -p1[1:]]))
        assert np.allclose(rotmat(qmul(rel_half, rel_half)) @ rotmat(p1),
                           rotmat(p2), atol=1e-13)


def test_dexpinv_is_inverse_of_exp_derivative():
    """u' = dexpinv(u, w) means exp(u + h u') q = exp(h w) exp(u) q + O(h^2)
    (both branches of MobilityConstraints.cpp:109-143)."""
    rng = np.random.default_rng(1)
    L = O.lib()
    for scale in (0.9, 1e-13):
        u = rng.standard_normal(3) * scale
        w = rng.standard_normal(3)
        ud = np.zeros(3)
        L.or_dexpinv(O.dp(u), O.dp(w), O.dp(ud))
        q0 = unit_q(rng)
        h = 1e-6

        def ex(v, q):
            qq = q.copy()
            L.or_quat_expm(O.dp(qq), O.dp(np.ascontiguousarray(v)))
            return qq
        a = ex(u + h * ud, q0)
        b = ex(h * w, ex(u, q0))
        if scale < 1e-6:  # exp of a sub-threshold vector is the identity
            b = ex(h * w, q0)
        assert np.abs(rotmat(a) - rotmat(b)).max() < 50 * h * h + 1e-12


def test_straight_filament_no_torque_constraint_sum():
    sp, nws, props, st = config(grid=64, nfil=4, nw=10, seed=2,
                                weight=(0, 0, -1.0))
    s = O.Sim(sp, nws, props)
    st = dict(st)
    st["lam"] = np.random.default_rng(0).standard_normal((4 * 9, 3))
    s.set_state(**st)
    s.forces(1)
    out = s.get_state()
    # straight, untwisted, no preferred curvature: elastic torque vanishes;
    # the constraint torques are -dL/2 t x lam on parallel tangents
    F = out["F"].reshape(4, 10, 3)
    # constraint forces telescope: sum over a filament = total gravity
    assert np.allclose(F.sum(1), np.array([0, 0, -1.0 * 2.2 * 10]), atol=1e-12)
    s2 = O.Sim(sp, nws, props)
    s2.set_state(**dict(st, lam=np.zeros((36, 3))))
    s2.forces(1)
    assert np.abs(s2.get_state()["TAU"]).max() < 1e-14


def test_residual_zero_on_exact_translation():
    """BDF2 integrates uniform sedimentation under the friction mobility
    exactly (MobilityConstraints.cpp:616-760), so the extrapolated guess
    already solves the step: zero Broyden iterations."""
    sp, nws, props, st = config(grid=64, nfil=3, nw=8, seed=4, ifs=1000,
                                weight=(0.3, -0.2, -1.0), barrier=0)
    V = np.array([0.3, -0.2, -1.0]) * sp.dL / (6 * PI)
    X = st["X"]
    st = dict(st, Xt=X - V * sp.dt, Xs=X)
    s = O.Sim(sp, nws, props)
    s.set_state(**st)
    assert s.timestep(5) == 0
    out = s.get_state()
    assert np.abs(out["X"] - (X + V * sp.dt)).max() < 1e-13


def _fd_jacobian(sp, nws, props, st, nt, h=1e-6):
    s = O.Sim(sp, nws, props)
    n = 6 * sum(nws)

    def resid(u):
        s.set_state(**st)
        s.initial_guess()  # qs, Xs from the state (Us, U history)
        s.set_state(Us=st["Us"], qs=st["qs"], Xs=st["Xs"])
        s.update(u)
        s.forces(nt)
        s.mobility(nt)
        _, _, _, V, W = s.fcm_arrays()
        e, _, _ = s.residual(nt, V.copy(), W.copy())
        return e
    cols = []
    for c in range(n):
        u = np.zeros(n)
        u[c] = h
        cols.append(-(resid(u) - resid(-u)) / (2 * h))
    return np.array(cols).T


def test_jacobian_exact_blocks_against_finite_differences():
    """Blocks of buildApproximateJacobian that are exact derivatives of the
    residual under the friction mobility at Us = 0 (Filament.hpp:811-853):
    first-bead rows, BlockC (positions vs Lie elements), BlockD (positions vs
    tensions, scaled by the 10x pre-factor) and BlockF (Lie vs tensions).
    Tension columns are in units of lambda*dt, as the reference's J."""
    sp, nws, props, st0 = config(grid=64, nfil=1, nw=5, seed=6, ifs=1000,
                                 barrier=0)
    st = rich_state(st0, nws, seed=3)
    st["Us"] = np.zeros_like(st["Us"])
    st["qs"] = st["q"]
    nt = 5
    nw = nws[0]
    s = O.Sim(sp, nws, props)
    s.set_state(**st)
    s.forces(nt)
    s.build_jacobian(nt)
    J = s.jacobian(0)
    FD = _fd_jacobian(sp, nws, props, st, nt)
    lam = slice(3 * nw + 3, 6 * nw)
    FD[:, lam] /= sp.dt
    U = slice(3, 3 * nw + 3)
    tol = 1e-6 * np.abs(J).max()
    assert np.abs(J[0:3] - FD[0:3]).max() < tol
    assert np.abs(J[3:3 * nw, U] - FD[3:3 * nw, U]).max() < tol
    assert np.abs(J[3:3 * nw, lam] - 10 * FD[3:3 * nw, lam]).max() < 10 * tol
    assert np.abs(J[3 * nw:, lam] - FD[3 * nw:, lam]).max() < tol


def test_inverse_and_broyden_secant():
    sp, nws, props, st0 = config(grid=64, nfil=3, nw=6, seed=7, ifs=1000)
    st = rich_state(st0, nws, seed=2)
    s = O.Sim(sp, nws, props)
    s.set_state(**st)
    s.initial_guess()
    s.forces(14)
    s.build_jacobian(14)
    for f in range(3):
        J, Ji = s.jacobian(f), s.jinv(f)
        D = np.ones(len(J))
        D[3 * nws[f] + 3:] = sp.dt
        assert np.abs(np.diag(D) @ Ji @ J - np.eye(len(J))).max() < 1e-10
    n = 6 * sum(nws)
    rng = np.random.default_rng(0)
    E, En, upd = (rng.standard_normal(n) for _ in range(3))
    s.broyden_update(E, En, upd, 0)
    assert np.allclose(s.jsolve(En - E, 1), -upd, atol=1e-10)


def test_timestep_converges_to_tol():
    sp, nws, props, st = config(grid=48, nfil=4, nw=8, seed=9, tol=1e-8,
                                swim=True)
    s = O.Sim(sp, nws, props)
    s.set_state(**st)
    its = [s.timestep(nt) for nt in (1, 2)]
    assert all(0 < i < sp.broyden_maxiter for i in its)
    s.forces(2)


def test_spring_links_hookes_law():
    """Spring links (spring_link_functions.cpp:24-45) on hand-placed beads:
    k = 5, natural length 2. Stretched along z (|d| = 3): the beads attract
    with 5 each. Compressed across the periodic boundary (|d| = 1 via the
    minimum image): they repel through the boundary with 5 each. A bead in
    two links gets both. Forces come from the previous positions Xs."""
    sp, nws, props, st = config(grid=64, nfil=3, nw=8, seed=2)
    Lz = sp.fcm.nz * sp.fcm.dx
    st = dict(st)
    Xs = st["Xs"].copy()
    Xs[0] = [1.0, 1.0, 1.0]          # filament 0, bead 0
    Xs[8] = [1.0, 1.0, 4.0]          # filament 1, bead 0: 3 above
    Xs[16] = [5.0, 5.0, 0.5]         # filament 2, bead 0
    Xs[23] = [5.0, 5.0, Lz - 0.5]    # filament 2, bead 7: 1 below, wrapped
    st["Xs"] = Xs
    s0 = O.Sim(sp, nws, props)
    s0.set_state(**st)
    s0.forces(1)
    F0 = s0.get_state()["F"]
    s1 = O.Sim(sp, nws, props)
    s1.set_state(**st)
    s1.set_spring_links([[0, 0, 1, 0], [2, 0, 2, 7], [1, 0, 2, 0]], 5.0, 2.0)
    s1.forces(1)
    dF = s1.get_state()["F"] - F0
    # the third link (8 -> 16): d = (-4, -4, 3.5), |d| = sqrt(44.25)
    d3 = np.array([-4.0, -4.0, 3.5])
    f3 = -5.0 * (np.sqrt(44.25) - 2.0) * d3 / np.sqrt(44.25)
    want = np.zeros_like(dF)
    want[0] = [0, 0, 5.0]
    want[8] = [0, 0, -5.0]
    want[8] += f3
    want[16] = [0, 0, 5.0]
    want[23] = [0, 0, -5.0]
    want[16] -= f3
    assert np.abs(dF - want).max() < 1e-12
    assert np.abs(dF.sum(0)).max() < 1e-12
    # the FCM forces carry them (assign_filament_data after the springs)
    Fa = s1.fcm_arrays()[1]
    assert np.array_equal(Fa, s1.get_state()["F"])


def test_trial_step_is_the_step_without_step():
    """A gait-reset trial pass (main_gaitreset.cpp:310-812, gait_reset_i < 2)
    runs the whole solve and skips Filament::step (:775-781): the committed
    state and the lambda history stay bit for bit, the trial's iterate is
    exactly the position the committed step moves to, and lam carries the
    trial's result (Filament.hpp:226-236 does not reset it)."""
    sp, nws, props, st = config(grid=48, nfil=5, nw=8, swim=True, seed=4,
                                tol=1e-9, ifs=100)
    a, b = O.Sim(sp, nws, props), O.Sim(sp, nws, props)
    for s in (a, b):
        s.set_state(**st)
        s.timestep(1)
        s.timestep(2)
    s0 = a.get_state()
    ia = a.timestep_trial(3)
    s1 = a.get_state()
    for k in ("X", "Xt", "U", "Ut", "q", "qt", "lam1", "lam2"):
        assert np.array_equal(s0[k], s1[k]), k
    ib = b.timestep(3)
    s2 = b.get_state()
    assert ia == ib
    for k in ("Xs", "Us", "qs", "lam"):
        assert np.array_equal(s1[k], s2[k]), k
    assert np.array_equal(s1["Xs"], s2["X"]) and not np.array_equal(s1["Xs"],
                                                                     s1["X"])


def test_barrier_switch_equals_barrier_off_system():
    """set_barrier(0) steps exactly like a system built without the barrier;
    set_barrier(1, FS_BARRIER_ALL) exactly like the default."""
    from filsim_amd import abi
    sp, nws, props, st = config(grid=48, nfil=5, nw=8, swim=True, seed=4,
                                tol=1e-9, ifs=100)
    sp_off, *_ = config(grid=48, nfil=5, nw=8, swim=True, seed=4, tol=1e-9,
                        ifs=100, barrier=0)
    on, off, sw = (O.Sim(sp, nws, props), O.Sim(sp_off, nws, props),
                   O.Sim(sp, nws, props))
    for s in (on, off, sw):
        s.set_state(**st)
    for nt in (1, 2):
        sw.set_barrier(0)
        i_sw, i_off = sw.timestep(nt), off.timestep(nt)
        assert i_sw == i_off
        assert np.array_equal(sw.get_state()["X"], off.get_state()["X"])
    sw.set_barrier(1, abi.FS_BARRIER_ALL)
    on.set_state(**{k: v for k, v in sw.get_state().items()})
    assert sw.timestep(3) == on.timestep(3)
    assert np.array_equal(sw.get_state()["X"], on.get_state()["X"])
