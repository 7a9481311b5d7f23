"""Broyden history carried across timesteps (fs_system_set_history_carry;
no reference counterpart): every step still ends at the residual
tolerance, so trajectories agree with the restarting solver to the
tolerance's level; the carried corrections cut the iteration count; the
column cap keeps the newest half when exceeded; setting the state clears
the history."""
import numpy as np
import pytest

from filsim_amd import api
from sysfix import config

pytestmark = pytest.mark.gpu


def _swimmers(precision=1):
    return config(grid=64, nfil=12, nw=15, seed=21, tol=1e-6, swim=True,
                  dt=0.005, kap=1000.0, weight=(0, 0, 0), precision=precision)


def _run(cap, steps, precision=1):
    sp, nws, props, st = _swimmers(precision)
    s = api.System(sp, nws, props)
    s.set_state(**st)
    api.system_set_history_carry(s, cap)
    its = [s.timestep(nt) for nt in range(1, steps + 1)]
    return s, its


@pytest.mark.parametrize("cap", [8, 96])
def test_carry_converges_to_the_same_trajectory(cuda, cap):
    """cap 8 exceeds the cap every few steps (the newest 4 columns kept),
    cap 96 carries everything it may."""
    steps = 24
    s0, i0 = _run(0, steps)
    s1, i1 = _run(cap, steps)
    x0, x1 = s0.get_state()["X"], s1.get_state()["X"]
    err = np.abs(x1 - x0).max() / np.abs(x0).max()
    print("restart iters", i0, "carry", cap, "iters", i1, "rel pos diff", err)
    sp = _swimmers()[0]
    assert max(i1) < sp.broyden_maxiter
    assert err < 1e-6
    assert i1[0] == i0[0]  # nothing to carry into the first step
    if cap == 96:
        assert np.mean(i1[4:]) < np.mean(i0[4:])


def test_set_state_clears_carried_history(cuda):
    sp, nws, props, st = _swimmers()
    a = api.System(sp, nws, props)
    a.set_state(**st)
    full = a.get_state()  # every field, so the restore below is complete
    api.system_set_history_carry(a, 96)
    for nt in range(1, 6):
        a.timestep(nt)
    a.set_state(**full)  # back to the start: the next step restarts
    b = api.System(sp, nws, props)
    b.set_state(**full)
    api.system_set_history_carry(b, 96)
    assert a.timestep(1) == b.timestep(1)
    assert np.array_equal(a.get_state()["X"], b.get_state()["X"])


def test_carry_bounds():
    sp, nws, props, st = _swimmers()
    s = api.System(sp, nws, props)
    with pytest.raises(api.abi.FsError):
        api.system_set_history_carry(s, -1)
    with pytest.raises(api.abi.FsError):
        api.system_set_history_carry(s, 193)


def test_batch_refuses_carried_history(cuda):
    sp, nws, props, st = _swimmers()
    s = api.System(sp, nws, props)
    s.set_state(**st)
    api.system_set_history_carry(s, 16)
    with pytest.raises(api.abi.FsError):
        api.Batch([s])
    api.system_set_history_carry(s, 0)
    b = api.Batch([s])
    b.close()
