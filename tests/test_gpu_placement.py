"""Initial configurations placed on the GPU (fs_make_filaments_gpu) are the
host generator's (fs_make_filaments) bit for bit: same trials drawn, same
trials accepted, including windows cut short where the segment count
changes and the failure after max_tries rejections."""
import time

import numpy as np
import pytest

from filsim_amd import abi, api, params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grid,nws,seed", [
    (64, [15] * 16, 3),
    (64, [8, 15, 10, 20, 15, 3] * 8, 11),      # mixed lengths, dense
    (96, [20] * 135, 5),                       # many rejections per filament
    (256, [30] * 1000, 20261202),               # the 256^3 workload
])
def test_gpu_placement_matches_host(cuda, grid, nws, seed):
    box = params.box(params.fcm_params(grid))
    Xh, qh = api.make_filaments(box, nws, seed=seed)
    Xg, qg = api.make_filaments(box, nws, seed=seed, gpu=True)
    assert np.array_equal(Xg, Xh) and np.array_equal(qg, qh)


def test_gpu_placement_c5_scale(cuda):
    """8,000 filaments x 30 segments in the 512^3 box."""
    box = params.box(params.fcm_params(512))
    t0 = time.perf_counter()
    Xg, qg = api.make_filaments(box, [30] * 8000, seed=20261202, gpu=True)
    tg = time.perf_counter() - t0
    t0 = time.perf_counter()
    Xh, qh = api.make_filaments(box, [30] * 8000, seed=20261202)
    th = time.perf_counter() - t0
    print("placement 8000 x 30 at 512^3: gpu %.2f s, host %.2f s" % (tg, th))
    assert np.array_equal(Xg, Xh) and np.array_equal(qg, qh)


def test_gpu_placement_gives_up_like_host(cuda):
    box = params.box(params.fcm_params(48))
    for gpu in (False, True):
        with pytest.raises(abi.FsError):
            api.make_filaments(box, [15] * 400, seed=1, max_tries=50, gpu=gpu)
    # a count that fits with few tries succeeds identically
    Xh, qh = api.make_filaments(box, [15] * 10, seed=1, max_tries=50)
    Xg, qg = api.make_filaments(box, [15] * 10, seed=1, max_tries=50, gpu=True)
    assert np.array_equal(Xg, Xh) and np.array_equal(qg, qh)
