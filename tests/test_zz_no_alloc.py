"""The no-allocation barrier (fs_colbar_barrier_noalloc) against the handle
path (fs_colbar_create + fs_colbar_link + fs_colbar_apply) on the same
inputs: two calls in a row, the second with fewer segments in a smaller box
and another pair mask, forces and minimum distance bit for bit. Also the
size limit of its scratch array."""
import ctypes as C

import numpy as np
import pytest

from filsim_amd import abi, api, params
from helpers import to_dev

pytestmark = pytest.mark.gpu


def _case(grid, nfil, nw, seed):
    fp = params.fcm_params(grid)
    cp = params.colbar_params(fp)
    X, _ = api.make_filaments(params.box(fp), [nw] * nfil, dL=2.2,
                              min_sep=1.9, seed=seed)
    rng = np.random.default_rng(seed)
    Xs = X + rng.normal(0, 0.6, X.shape)
    Y = np.fmod(Xs, np.array(params.box(fp)))
    sf = np.repeat(np.arange(nfil), nw).astype(np.int32)
    si = np.tile(np.arange(nw), nfil).astype(np.int32)
    F0 = rng.standard_normal(X.shape)
    return cp, Xs, Y, sf, si, F0


def _handle_path(torch, cp, Xs, Y, sf, si, F0, nt, mask):
    gb = api.Colbar(cp, sf, si)
    abi.check(gb.lib.fs_colbar_set_mask(gb.h, mask))
    gb.link(to_dev(torch, Xs))
    dF = to_dev(torch, F0)
    md = gb.apply(to_dev(torch, Y), nt, dF)
    out = dF.cpu().numpy()
    gb.close()
    return out, md


def test_noalloc_barrier_matches_handle_path(cuda):
    torch = cuda
    cases = [(_case(96, 120, 20, 3), 4, abi.FS_BARRIER_ALL),
             (_case(64, 16, 15, 5), 7, abi.FS_BARRIER_NOT_SWIMMER)]
    assert cases[1][0][1].shape[0] < cases[0][0][1].shape[0]
    for (cp, Xs, Y, sf, si, F0), nt, mask in cases:
        print("no-alloc barrier: np", Xs.shape[0], "mask", mask, flush=True)
        Fh, mdh = _handle_path(torch, cp, Xs, Y, sf, si, F0, nt, mask)
        dF = to_dev(torch, F0)
        dmd = torch.zeros(1, dtype=torch.float64, device="cuda")
        api.colbar_barrier_noalloc(
            cp, torch.tensor(sf, device="cuda"), torch.tensor(si, device="cuda"),
            to_dev(torch, Xs), to_dev(torch, Y), nt, dF, dmd, mask=mask)
        torch.cuda.synchronize()
        Fn = dF.cpu().numpy()
        assert np.abs(Fh - F0).max() > 0  # pairs inside the cutoff
        assert np.array_equal(Fn, Fh)
        assert float(dmd.item()) == mdh


def test_noalloc_barrier_rejects_oversized_inputs(cuda):
    """A 1,024^3 box has 10.6M barrier cells: more than the 32 MiB array
    holds. The call fails before touching the GPU."""
    fp = params.fcm_params(1024)
    cp = params.colbar_params(fp)
    lib = abi.load()
    rc = lib.fs_colbar_barrier_noalloc(C.byref(cp), 0, None, None, None, None,
                                       1, abi.FS_BARRIER_ALL, None, None, None)
    assert rc == abi.FS_ERR_SIZE
    assert b"32 MiB" in lib.fs_last_error() or b"scratch" in lib.fs_last_error()
