"""The gait-reset driver's passes (main_gaitreset.cpp:310-812) on the GPU
against the oracle: per timestep a trial pass without the barrier, a trial
pass without swimmer/network pairs (apply_collision_barrier_not_swimmer,
CollisionBarrierFilament.cpp:441-696), then the committed step with the
full barrier. Filament 0 (the swimmer) lies 1.8 beside filament 1, inside
the barrier's cutoff, so the passes differ."""
import numpy as np
import pytest

import _oracle as O
from filsim_amd import abi, api
from sysfix import config

pytestmark = pytest.mark.gpu


def _crowded_state(st, nw):
    X = st["X"].copy()
    q = st["q"].copy()
    t = X[nw + 1] - X[nw]
    t /= np.linalg.norm(t)
    e = np.cross(t, [0.0, 0.0, 1.0])
    if np.linalg.norm(e) < 0.1:
        e = np.cross(t, [1.0, 0.0, 0.0])
    e /= np.linalg.norm(e)
    X[:nw] = X[nw:2 * nw] + 1.8 * e
    q[:nw] = q[nw:2 * nw]
    return dict(X=X, Xs=X, Xt=X, q=q, qs=q, qt=q)


def _passes(s, nt):
    its, xs = [], []
    for i in range(3):
        s.set_barrier(i >= 1, abi.FS_BARRIER_NOT_SWIMMER if i == 1
                      else abi.FS_BARRIER_ALL)
        its.append(s.timestep_trial(nt) if i < 2 else s.timestep(nt))
        xs.append(s.get_state()["Xs"].copy())
    return its, xs


@pytest.mark.parametrize("ifs,steps", [(100, 2), (0, 1)])
def test_gait_reset_passes_match_oracle(cuda, ifs, steps):
    nw = 8
    sp, nws, props, st = config(grid=48, nfil=5, nw=nw, swim=True, seed=4,
                                tol=1e-9, ifs=ifs)
    st = _crowded_state(st, nw)
    o, g = O.Sim(sp, nws, props), api.System(sp, nws, props)
    for s in (o, g):
        s.set_state(**st)
    Lb = np.abs(st["X"]).max()
    for nt in range(1, steps + 1):
        before = g.get_state()
        io, xo = _passes(o, nt)
        ig, xg = _passes(g, nt)
        print("nt", nt, "iterations oracle", io, "gpu", ig, flush=True)
        for a, b in zip(io, ig):
            assert abs(a - b) <= 1
        for a, b in zip(xo, xg):
            assert np.abs(a - b).max() <= 1e-8 * Lb
        # the swimmer/network mask changes the solution; the trial passes
        # left the committed state for the last pass
        assert np.abs(xg[1] - xg[2]).max() > 1e-6
        assert not np.array_equal(before["X"], g.get_state()["X"])
    so, sg = o.get_state(), g.get_state()
    assert np.abs(sg["X"] - so["X"]).max() <= 1e-8 * Lb
    assert np.abs(sg["q"] - so["q"]).max() <= 1e-8
    g.close()


def test_trial_step_keeps_committed_state(cuda):
    sp, nws, props, st = config(grid=48, nfil=5, nw=8, swim=True, seed=4,
                                tol=1e-9, ifs=100)
    a, b = api.System(sp, nws, props), api.System(sp, nws, props)
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
    assert np.array_equal(s1["Xs"], s2["X"])
    # a trial pass fills the profiler's buckets like a committed step
    a.set_profiling(True)
    a.timestep_trial(3)
    pr = a.profile()
    assert pr["Total"] > 0 and pr["Bld Jinv"] > 0 and pr["ColBarri"] > 0
    a.close()
    b.close()


def test_barrier_switch_rules(cuda):
    sp, nws, props, st = config(grid=48, nfil=3, nw=6, seed=2, barrier=0)
    s = api.System(sp, nws, props)
    s.set_barrier(0)
    with pytest.raises(Exception):
        s.set_barrier(1)           # built without the barrier
    s.close()
    sp, nws, props, st = config(grid=48, nfil=3, nw=6, seed=2)
    s = api.System(sp, nws, props)
    s.set_state(**st)
    with pytest.raises(Exception):
        s.set_barrier(1, 7)        # unknown mask
    batch = api.Batch([s])
    # every call that would change what the captured launches read, or free
    # it, is refused while the system belongs to a batch
    with pytest.raises(Exception):
        s.set_barrier(0)
    with pytest.raises(Exception):
        s.set_spring_links([[0, 0, 1, 0]], 1.0, 2.2)
    with pytest.raises(Exception):
        s.set_tuning(abi.FcmTuning(-1, -1, -1, -1, -1))
    with pytest.raises(Exception):
        s.autotune(1)
    with pytest.raises(Exception):
        s.set_profiling(True)
    n_in, n_out = api.system_warmstart_dims(s)
    m = api.Mlp(n_in, n_out, width=64)
    with pytest.raises(Exception):
        api.system_set_warmstart(s, m)
    with pytest.raises(Exception):
        s.close()
    assert s.h
    its = batch.timestep(1)        # the batch still steps the system
    assert its[0] >= 0
    batch.close()
    s.set_barrier(0)
    s.set_spring_links([[0, 0, 1, 0]], 1.0, 2.2)
    m.close()
    s.close()
