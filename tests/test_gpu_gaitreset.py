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
from sysfix import config, d2h

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


def _swimmer_row(state, arrays, nw):
    """Filament 0's Xs, qs and V, W, F rows (the swimmer velocity file)."""
    Y, F, T, V, W = arrays
    return tuple(a[:nw].copy() for a in (state["Xs"], state["qs"], V, W, F))


def _read_rows(path):
    out = []
    for line in open(path).read().split("\n")[1:]:
        if line:
            t = line.split(" ")[:-1]
            out.append((int(t[0]), np.array([float(x) for x in t[1:]])))
    return out


def test_gait_reset_and_remove_passes_match_oracle(cuda, tmp_path):
    """main_gaitresetandremove.cpp:328-339, 458-480, 775-820 under the FCM
    mobility: per timestep a trial pass of every filament without the
    barrier, a trial pass of the swimmer alone (a one-filament system that
    takes filament 0's state, fs_system_copy_filaments), then the committed
    step of every filament with the full barrier from the state the first
    pass left. Each pass's iteration count, the swimmer velocity rows of
    all three passes, the barrier force (linked_list.F_total) and the force
    rows of the committed step against the oracle; the state copy bit for
    bit; the swimmer's COM and mean velocity as the reference computes
    them."""
    nw = 8
    sp, nws, props, st = config(grid=48, nfil=5, nw=nw, swim=True, seed=4,
                                tol=1e-9, ifs=0)
    st = _crowded_state(st, nw)
    p1 = (abi.FilProps * 1)()
    p1[0] = props[0]
    o, o1 = O.Sim(sp, nws, props), O.Sim(sp, [nw], p1)
    g, g1 = api.System(sp, nws, props), api.System(sp, [nw], p1)
    for s in (o, g):
        s.set_state(**st)
    L = np.abs(st["X"]).max()
    files = {k: str(tmp_path / (k + ".dat"))
             for k in ("reset", "remove", "swimmer", "collision", "forces")}
    for k, kind in (("reset", "velocity"), ("remove", "velocity"),
                    ("swimmer", "velocity"), ("collision", "collision"),
                    ("forces", "forces")):
        api.io_write_names(files[k], kind, nw)
    expect = {k: [] for k in files}
    for nt in (1, 2):
        # pass 0: every filament, no barrier, trial
        o.set_barrier(0, abi.FS_BARRIER_ALL)
        g.set_barrier(0, abi.FS_BARRIER_ALL)
        i0 = (o.timestep_trial(nt), g.timestep_trial(nt))
        g.write_swimmer_velocity(files["reset"], nt)
        expect["reset"].append(_swimmer_row(o.get_state(), o.fcm_arrays(), nw))
        # pass 1: the swimmer alone from the state pass 0 left
        so = o.get_state()
        o1.set_state(**{k: v[:nw - 1] if k.startswith("lam") else v[:nw]
                        for k, v in so.items()})
        g1.copy_filaments(g, 0)
        sg, s1 = g.get_state(), g1.get_state()
        for k, v in s1.items():
            assert np.array_equal(v, sg[k][:len(v)]), k
        o1.set_barrier(0, abi.FS_BARRIER_ALL)
        g1.set_barrier(0, abi.FS_BARRIER_ALL)
        i1 = (o1.timestep_trial(nt), g1.timestep_trial(nt))
        g1.write_swimmer_velocity(files["remove"], nt)
        expect["remove"].append(_swimmer_row(o1.get_state(),
                                             o1.fcm_arrays(), nw))
        # the full system still holds pass 0's iterate
        assert np.array_equal(g.get_state()["Xs"], sg["Xs"])
        # pass 2: every filament, full barrier, committed step
        o.set_barrier(1, abi.FS_BARRIER_ALL)
        g.set_barrier(1, abi.FS_BARRIER_ALL)
        i2 = (o.timestep(nt), g.timestep(nt))
        g.write_swimmer_velocity(files["swimmer"], nt)
        g.write_swimmer_collision(files["collision"], nt)
        g.write_forces(files["forces"], nt)
        so, ao = o.get_state(), o.fcm_arrays()
        expect["swimmer"].append(_swimmer_row(so, ao, nw))
        expect["collision"].append((o.barrier_force()[:nw],))
        expect["forces"].append((np.hstack([ao[1], ao[2]]),))
        print("nt", nt, "iterations (oracle, gpu) per pass", i0, i1, i2,
              flush=True)
        for a, b in (i0, i1, i2):
            assert abs(a - b) <= 1
        Fb = g.barrier_force(nt)
        Fo = o.barrier_force()
        assert np.abs(Fo).max() > 0
        assert np.abs(Fb - Fo).max() <= 1e-6 * np.abs(Fo).max()
        # COM (Filament::getCOM) and the mean segment velocity printout
        X = g.get_state()["X"]
        com = np.zeros(3)
        for n in range(nw):
            com += X[n]
        assert np.array_equal(g.com(0), com / nw)
        V = d2h(g.fcm_ptrs()[3], 3 * g.np).reshape(-1, 3)
        v = [np.float32(0.0)] * 3
        for n in range(nw):
            for k in range(3):
                v[k] = np.float32(float(v[k]) + V[n, k] / nw)
        assert np.array_equal(g.mean_velocity(0), np.array(v, np.float64))
    # removing the other filaments changes the swimmer's solution
    r0 = np.concatenate([a.ravel() for a in expect["reset"][-1]])
    r1 = np.concatenate([a.ravel() for a in expect["remove"][-1]])
    assert np.abs(r0 - r1).max() > 1e-6
    for k, rows in expect.items():
        got = _read_rows(files[k])
        assert [nt for nt, _ in got] == [1, 2], k
        for (nt, vals), ref in zip(got, rows):
            if k == "forces":
                # every filament's rows (their first segments are multiples
                # of 8: rows (8 f) // 6 + i)
                F, T = ref[0][:, :3], ref[0][:, 3:]
                idx = np.concatenate([np.arange(nw) + (nw * f) // 6
                                      for f in range(len(nws))])
                ref = (np.hstack([F[idx], T[idx]]),)
            if k in ("reset", "remove", "swimmer"):
                # per bead Xs qs, then per segment V W F
                pos = vals[:7 * nw].reshape(nw, 7)
                vel = vals[7 * nw:].reshape(nw, 9)
                got_parts = (pos[:, :3], pos[:, 3:], vel[:, :3], vel[:, 3:6],
                             vel[:, 6:])
            else:
                got_parts = (vals.reshape(ref[0].shape),)
            assert len(vals) == sum(a.size for a in ref), k
            for a, b in zip(got_parts, ref):
                err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
                print(k, "nt", nt, "relative error %.3e" % err, flush=True)
                assert err <= 1e-6, k
    for s in (g, g1):
        s.close()
