"""Builds matching oracle and GPU filament systems from one seeded state."""
import ctypes as C

import numpy as np

import _oracle as O
from filsim_amd import api, params


def d2h(ptr, n):
    """Copy n doubles from a raw device pointer to a numpy array."""
    hip = C.CDLL("libamdhip64.so")
    out = np.zeros(n)
    rc = hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), ptr,
                       C.c_size_t(8 * n), 2)
    assert rc == 0
    return out


def config(grid=64, nfil=16, nw=15, seed=3, tol=1e-10, maxiter=100,
           precision=0, weight=(0.0, 0.0, -1.0), swim=False, dt=0.01,
           barrier=1, ifs=0, kap=1.0):
    fp = params.fcm_params(grid)
    sp = params.sim_params(fp, dt=dt, tol=tol, maxiter=maxiter,
                           precision=precision, weight=weight,
                           barrier_on=barrier, initial_friction_steps=ifs)
    nws = [nw] * nfil
    if swim:
        rng = np.random.default_rng(seed)
        props = params.swimming_props(nfil, nw, kap=kap,
                                      phases=list(rng.uniform(0, 2 * np.pi,
                                                              nfil)))
    else:
        props = params.fil_props(nfil, kap=kap)
    X, q = api.make_filaments(params.box(fp), nws, dL=2.2, min_sep=2.2,
                              seed=seed)
    state = dict(X=X, Xs=X, Xt=X, q=q, qs=q, qt=q)
    return sp, nws, props, state


def pair(sp, nws, props, state):
    o = O.Sim(sp, nws, props)
    g = api.System(sp, nws, props)
    o.set_state(**state)
    g.set_state(**state)
    return o, g


def _qmul(p, q):
    s = p[..., 0] * q[..., 0] - (p[..., 1:] * q[..., 1:]).sum(-1)
    v = (p[..., :1] * q[..., 1:] + q[..., :1] * p[..., 1:] +
         np.cross(p[..., 1:], q[..., 1:]))
    return np.concatenate([s[..., None], v], -1)


def rich_state(state, nws, seed=0, bend=0.15, mag=0.02):
    """A non-trivial state from a straight one: every bead rotated by its own
    small random rotation (bent, twisted filaments), past positions shifted,
    nonzero Lie-algebra history and Lagrange multipliers."""
    rng = np.random.default_rng(seed)
    X, q = state["X"], state["q"]
    npt = X.shape[0]
    nlam = npt - len(nws)

    def small_rot(n, amp):
        v = rng.normal(0, amp, (n, 3))
        th = np.linalg.norm(v, axis=1, keepdims=True)
        return np.concatenate([np.cos(th / 2), np.sin(th / 2) * v / th], 1)

    qn = _qmul(small_rot(npt, bend), q)
    qt = _qmul(small_rot(npt, bend * 0.2), qn)
    qs = _qmul(small_rot(npt, bend * 0.2), qn)
    return dict(X=X, Xs=X + rng.normal(0, mag, X.shape),
                Xt=X - rng.normal(0, mag, X.shape),
                U=rng.normal(0, mag, (npt, 3)), Us=rng.normal(0, mag, (npt, 3)),
                Ut=rng.normal(0, mag, (npt, 3)), q=qn, qs=qs, qt=qt,
                lam=rng.normal(0, 0.3, (nlam, 3)),
                lam1=rng.normal(0, 0.3, (nlam, 3)),
                lam2=rng.normal(0, 0.3, (nlam, 3)))


def advance_oracle_state(sp, nws, props, state, steps):
    """Runs the oracle a few steps to reach a non-trivial state."""
    o = O.Sim(sp, nws, props)
    o.set_state(**state)
    for nt in range(1, steps + 1):
        o.timestep(nt)
    s = o.get_state()
    return {k: s[k] for k in ("X", "Xs", "Xt", "U", "Us", "Ut", "q", "qs",
                              "qt", "lam", "lam1", "lam2")}
