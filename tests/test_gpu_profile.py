"""Per-timestep profile in the reference profiler's buckets: profiling adds
event records only (results bit-identical), the FCM sub-buckets add up to
at most the FCM bucket, and the total covers the top-level buckets."""
import numpy as np
import pytest

from sysfix import config
from filsim_amd import api

pytestmark = pytest.mark.gpu


def test_profile_buckets(cuda):
    sp, nws, props, st = config(grid=48, nfil=10, nw=15, seed=6, tol=1e-8,
                                swim=True, dt=0.005, kap=100.0,
                                weight=(0, 0, 0))
    a = api.System(sp, nws, props)
    b = api.System(sp, nws, props)
    a.set_state(**st)
    b.set_state(**st)
    a.set_profiling(True)
    for nt in range(1, 4):
        assert a.timestep(nt) == b.timestep(nt)
        pr = a.profile()
        print(nt, {k: round(v, 3) for k, v in pr.items()})
        assert all(v >= 0 for v in pr.values())
        sub = sum(pr[k] for k in ("FCMMakeZ", "FCMGauss", "FCMForcD",
                                  "FCMFFTFw", "FCMFFTBk", "FCMFlowS",
                                  "FCMParVR"))
        assert 0 < sub <= pr["FCM"] * 1.001 + 1e-3
        top = sum(pr[k] for k in ("FCM", "Assign F", "Bld Jinv", "ErrorChk",
                                  "Slv Jinv", "ColBarri"))
        assert 0 < top <= pr["Total"] * 1.001 + 1e-3
    assert np.array_equal(a.get_state()["X"], b.get_state()["X"])
    a.set_profiling(False)
