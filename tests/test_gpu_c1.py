"""The CPU-runnable reference case run in full on the GPU against the
oracle: 16 filaments x 15 segments (a = 1, spacing 2.2a, KB = 1, mu = 1),
periodic 64^3 grid with dx = a/1.86 and ngd = 12, overlap-rejected random
placement (2.2a), constant gravity along -z, collision barrier on, BDF2 +
Broyden with TOL 1e-4, dt = 0.01, 100 steps, plain FP64. The filaments
sediment into contact late in the run (6-9 Broyden iterations per step)."""
import numpy as np
import pytest

import _oracle as O
from filsim_amd import api, params
from sysfix import config

pytestmark = pytest.mark.gpu


def test_sedimenting_filaments_100_steps(cuda):
    sp, nws, props, st = config(grid=64, nfil=16, nw=15, seed=20261202,
                                tol=1e-4, dt=0.01, weight=(0.0, 0.0, -1.0))
    o, g = O.Sim(sp, nws, props), api.System(sp, nws, props)
    for s in (o, g):
        s.set_state(**st)
    its_o, its_g = [], []
    for nt in range(1, 101):
        its_o.append(o.timestep(nt))
        its_g.append(g.timestep(nt))
        if nt % 20 == 0:
            print("step", nt, "oracle", its_o[-20:], "gpu", its_g[-20:],
                  flush=True)
    diff = np.abs(np.array(its_o) - np.array(its_g))
    assert diff.max() <= 1 and (diff > 0).sum() <= 5, (its_o, its_g)
    assert sum(its_o) > 100  # the run reaches the colliding regime
    so, sg = o.get_state(), g.get_state()
    box = np.array(params.box(params.fcm_params(64)))
    err = np.abs(sg["X"] - so["X"]).max() / box.max()
    print("final positions: max |dX| / box = %.3e" % err, flush=True)
    assert err <= 1e-8 if diff.max() == 0 else err <= 1e-5
    assert np.abs(sg["q"] - so["q"]).max() <= (1e-8 if diff.max() == 0
                                               else 1e-5)
    g.close()
