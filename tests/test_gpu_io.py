"""Restart of a system on the GPU: a system rebuilt from the binary
checkpoint continues bit for bit; one rebuilt from the reference's .bak text
continues within its 16-decimal rounding."""
import numpy as np
import pytest

from sysfix import config
from filsim_amd import api

pytestmark = pytest.mark.gpu


def test_restart_from_checkpoint_and_bak(cuda, tmp_path):
    sp, nws, props, st = config(grid=32, nfil=6, nw=12, seed=4, tol=1e-8,
                                swim=True, dt=0.005, kap=100.0,
                                weight=(0, 0, 0))
    a = api.System(sp, nws, props)
    a.set_state(**st)
    for nt in range(1, 4):
        a.timestep(nt)
    ck, bak, dat = (str(tmp_path / n) for n in ("r.fsck", "r.bak", "r.dat"))
    api.system_write_checkpoint(a, ck, 3)
    api.system_write_bak(a, bak, 3)
    api.system_write_dat(a, dat, 3)
    b = api.System(sp, nws, props)
    assert api.system_read_checkpoint(b, ck) == 3
    c = api.System(sp, nws, props)
    assert api.system_read_bak(c, bak) == 3
    for nt in range(4, 7):
        ia, ib, ic = a.timestep(nt), b.timestep(nt), c.timestep(nt)
        assert ia == ib
        assert abs(ia - ic) <= 1
    xa, xb, xc = (s.get_state()["X"] for s in (a, b, c))
    assert np.array_equal(xa, xb)
    assert np.abs(xa - xc).max() <= 1e-9
    line = open(dat).read().split("\n")[0].split(" ")
    assert line[0] == "3" and len(line) - 2 == 7 * sum(nws)
