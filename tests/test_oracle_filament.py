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


def _chain(q0, axes, thetas):
    """Bead frames q_{i+1} = q_i * (cos th_i/2, sin th_i/2 e_i): each link
    turns the next bead by th_i about the body-frame axis e_i."""
    qs = [q0]
    for e, th in zip(axes, thetas):
        qs.append(qmul(qs[-1], np.concatenate([[np.cos(th / 2)],
                                               np.sin(th / 2) * e])))
    return np.array(qs)


def _elastic_only(sp, nws, props, q):
    st = dict(X=np.zeros((len(q), 3)), Xs=np.zeros((len(q), 3)),
              Xt=np.zeros((len(q), 3)), q=q, qs=q, qt=q,
              lam=np.zeros((len(q) - 1, 3)))
    s = O.Sim(sp, nws, props)
    s.set_state(**st)
    s.forces(1)
    return s.get_state()["TAU"]


def test_elastic_torque_of_a_uniformly_strained_link():
    """applyElasticTorques (Filament.hpp:289-342) on a link whose second
    bead is the first turned by theta about body axis e_k: the midpoint frame
    is q_i * (cos th/4, sin th/4 e_k), the discrete strain 2 vec(q_m^*
    (q_j - q_i)/dL) is exactly (4/dL) sin(th/4) e_k, and the torque is
    R(q_m) diag(C, kap, kap) strain, added to bead i and taken from bead j.
    C != kap, so twist (k = x) and bend (k = y, z) tell the moduli apart; a
    sign error, a swapped modulus or a wrong midpoint frame fails."""
    nw, C, kap = 6, 2.0, 3.0
    fp = params.fcm_params(48)
    sp = params.sim_params(fp, barrier_on=0)
    props = params.fil_props(1, kap=kap, c=C)
    rng = np.random.default_rng(1)
    q0 = unit_q(rng)
    D = np.diag([C, kap, kap])
    for k in range(3):
        e = np.eye(3)[k]
        for th in (0.3, -0.3, 1.7):
            q = _chain(q0, [e] * (nw - 1), [th] * (nw - 1))
            tau = _elastic_only(sp, [nw], props, q)
            want = np.zeros((nw, 3))
            for i in range(nw - 1):
                qm = qmul(q[i], np.concatenate([[np.cos(th / 4)],
                                                np.sin(th / 4) * e]))
                t = rotmat(qm) @ D @ (4 / sp.dL * np.sin(th / 4) * e)
                want[i] += t
                want[i + 1] -= t
            assert np.abs(tau - want).max() < 1e-14 * max(1, np.abs(want).max())
            # interior beads of a uniform chain feel no net torque; the ends
            # feel +-R(q_m) D strain
            assert np.abs(tau[1:-1]).max() < 1e-13
            assert np.abs(want[0]).max() > 0.1


def test_elastic_torque_vanishes_at_the_preferred_strain():
    """The swimming drive (Filament.hpp:301-326): preferred strain-twist
    (STRAINTWISTZ, beta facPref cos(k s - w t + phi), alpha facPref sin(...))
    with facPref = T0 (L - s)/L. Turning each link by the rotation whose
    discrete strain equals its own preferred vector p_i (axis p_i/|p_i|,
    angle 4 asin(dL |p_i|/4)) gives zero torque on every bead. Straight and
    turned the other way, the torque is R(q_m) D ((4/dL) sin(th/4) e - p)
    with q_m the half-turned frame."""
    nw = 8
    fp = params.fcm_params(48)
    sp = params.sim_params(fp, barrier_on=0, strain_twist0=0.05)
    L = nw * sp.dL
    props = params.fil_props(1, kap=3.0, c=2.0, T0=1.5 / L, omega=2 * PI,
                             wavenumber=2 * PI / L, phase=0.7, alpha=0.4,
                             beta=1.0)
    nt = 3
    tt = nt * sp.dt
    Lf = (nw - 1) * sp.dL
    pref = []
    for i in range(nw - 1):
        s = sp.dL * i + 0.5 * sp.dL
        fac = props[0].T0 * (Lf - s) / Lf
        ph = props[0].wavenumber * s - props[0].omega * tt + props[0].phase
        pref.append([sp.strain_twist0, props[0].beta * fac * np.cos(ph),
                     props[0].alpha * fac * np.sin(ph)])
    pref = np.array(pref)
    mag = np.linalg.norm(pref, axis=1)
    axes = pref / mag[:, None]
    th = 4 * np.arcsin(sp.dL * mag / 4)
    q0 = unit_q(np.random.default_rng(5))

    def tau_of(q):
        st = dict(X=np.zeros((nw, 3)), Xs=np.zeros((nw, 3)),
                  Xt=np.zeros((nw, 3)), q=q, qs=q, qt=q,
                  lam=np.zeros((nw - 1, 3)))
        s = O.Sim(sp, [nw], props)
        s.set_state(**st)
        s.forces(nt)
        return s.get_state()["TAU"]
    D = np.diag([2.0, 3.0, 3.0])

    def closed_form(q, sgn):
        want = np.zeros((nw, 3))
        for i in range(nw - 1):
            h = sgn * th[i] / 4
            qm = qmul(q[i], np.concatenate([[np.cos(h)], np.sin(h) * axes[i]]))
            strain = 4 / sp.dL * np.sin(h) * axes[i]
            t = rotmat(qm) @ D @ (strain - pref[i])
            want[i] += t
            want[i + 1] -= t
        return want
    t_pref = tau_of(_chain(q0, axes, th))
    q_straight = _chain(q0, axes, 0 * th)
    q_back = _chain(q0, axes, -th)
    t_straight, t_back = tau_of(q_straight), tau_of(q_back)
    scale = np.abs(t_straight).max()
    assert scale > 0.01
    assert np.abs(t_pref).max() < 1e-13 * scale
    assert np.abs(t_straight - closed_form(q_straight, 0)).max() < 1e-13 * scale
    assert np.abs(t_back - closed_form(q_back, -1)).max() < 1e-13 * scale


def _blockE_vs_fd(kap, c, straight, nw=6):
    """Rows [3Nw, 6Nw) x the U columns of the oracle's J and of a central
    finite difference of the residual, at Us = 0 under the friction
    mobility, both minus the identity."""
    sp, nws, props, st0 = config(grid=64, nfil=1, nw=nw, seed=6, ifs=1000,
                                 barrier=0, weight=(0, 0, 0), kap=kap)
    props[0].c = c
    st = rich_state(st0, nws, seed=3)
    if straight:
        st.update(q=st0["q"], qt=st0["q"], lam=np.zeros_like(st["lam"]))
    st["Us"] = np.zeros_like(st["Us"])
    st["qs"] = st["q"]
    nt = 5
    s = O.Sim(sp, nws, props)
    s.set_state(**st)
    s.forces(nt)
    s.build_jacobian(nt)
    J = s.jacobian(0)
    FD = _fd_jacobian(sp, nws, props, st, nt)
    U = slice(3, 3 * nw + 3)
    eye = np.eye(3 * nw)
    return J[3 * nw:, U] - eye, FD[3 * nw:, U] - eye


def test_blockE_is_exact_without_elastic_moduli():
    """With kap = C = 0 the Lie rows of the residual depend on U only through
    the constraint torques -dL/2 t_i x (lam_{i-1} + lam_i) and dexpinv; at
    Us = 0 their exact derivative is I - 2dt/3 (dOmega + [Omega x]/2) with
    dOmega = -dL/2 Tfac [lam x][t x], which is what BlockE
    (Filament.hpp:949-1095) evaluates. Bent, twisted filament, nonzero lam
    and TAU: BlockE equals the finite-difference Jacobian."""
    E, Efd = _blockE_vs_fd(0.0, 0.0, straight=False)
    assert np.abs(Efd).max() > 1e-5
    assert np.abs(E - Efd).max() < 1e-7 * np.abs(Efd).max()


def test_blockE_elastic_terms_on_a_straight_filament():
    """On a straight, untwisted filament with no preferred strain and lam = 0
    the exact derivative of the Lie residual is the linearised beam
    stiffness. The reference's approximate blocks relate to it in a fixed
    pattern, which pins the restatement's signs and moduli:
    twist (kap = 0): the diagonal and right off-diagonal blocks are exact and
    the left off-diagonal blocks (bead k w.r.t. k-1, Filament.hpp:1026-1034,
    1075-1081) carry the opposite sign;
    bend (C = 0): every block is exactly half the derivative
    (0.5*myKAP*rcross(t_+-1)*rcross(t)/DL where both links would give
    myKAP)."""
    nw = 6
    E, Efd = _blockE_vs_fd(0.0, 2.0, straight=True, nw=nw)
    scale = np.abs(Efd).max()
    assert scale > 1e-5
    for i in range(nw):
        for j in range(max(0, i - 1), min(nw, i + 2)):
            a = E[3 * i:3 * i + 3, 3 * j:3 * j + 3]
            b = Efd[3 * i:3 * i + 3, 3 * j:3 * j + 3]
            sign = -1.0 if j == i - 1 else 1.0
            assert np.abs(a - sign * b).max() < 1e-7 * scale, (i, j)
    E, Efd = _blockE_vs_fd(3.0, 0.0, straight=True, nw=nw)
    scale = np.abs(Efd).max()
    assert scale > 1e-5
    assert np.abs(E - 0.5 * Efd).max() < 1e-7 * scale


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
