"""FCM mobility on the GPU against the FP64 oracle (same seeded inputs)."""
import numpy as np
import pytest

import _oracle as O
from filsim_amd import abi, api, params
from helpers import random_segments, rel, to_dev

pytestmark = pytest.mark.gpu

# (grid, ngd, dx) -- radix 2, 3 and 5 sizes, non-cubic boxes, reference dx
CASES = [
    ((32, 32, 32), 12, None),
    ((48, 32, 40), 12, None),
    ((64, 64, 64), 12, None),
    ((30, 36, 50), 12, None),
    ((54, 24, 64), 12, None),
    ((40, 40, 40), 20, "ref"),
    # stencils wrapping onto both ends of a spread tile
    ((16, 16, 16), 12, None),
    ((24, 24, 24), 20, "ref"),
    # matrix-core tiles: fewer tiles per axis than a stencil can meet (the
    # stencil meets one tile at both ends), reference resolution (ngd 20)
    ((16, 16, 32), 12, None),
    ((24, 24, 48), 20, "ref"),
    ((48, 48, 48), 20, "ref"),
    # 512-point lines (the 512^3 suspension's FFT length)
    ((512, 16, 24), 12, None),
    # 128/256/512-point lines take compile-time power-of-two plans
    ((256, 128, 32), 12, None),
    ((32, 256, 128), 12, None),
    ((16, 24, 256), 12, None),
    ((24, 16, 512), 12, None),
]


def _fp(grid, ngd, dx):
    if dx == "ref":
        return params.reference_fcm_params(grid)
    return params.fcm_params(grid, ngd=ngd)


@pytest.mark.parametrize("mfma", [1, 0])
@pytest.mark.parametrize("grid,ngd,dx", CASES)
def test_mobility_fp64_matches_oracle(cuda, grid, ngd, dx, mfma):
    """mfma=1 (the default) spreads on the matrix cores where the grid
    takes 8 x 8 x 16 tiles, mfma=0 always with the vector gather."""
    fp = _fp(grid, ngd, dx)
    Y, F, T = random_segments(fp, 200, seed=sum(grid))
    Vo, Wo = O.mobility(fp, Y, F, T)
    with api.Fcm(fp, 200) as fcm:
        fcm.set_tuning(spread_mfma=mfma)
        V, W = fcm.mobility_host(Y, F, T)
    print("grid", grid, "mfma", mfma, "relV", rel(V, Vo), "relW", rel(W, Wo))
    assert rel(V, Vo) <= 1e-12
    assert rel(W, Wo) <= 1e-12


@pytest.mark.parametrize("mfma", [1, 0])
@pytest.mark.parametrize("grid", [64, (256, 128, 64), (512, 32, 48)])
def test_mobility_fp32_error_bound(cuda, grid, mfma):
    """FP32 spread on the matrix cores (mfma=1, the default) and with the
    vector gather (mfma=0)."""
    fp = params.fcm_params(grid)
    Y, F, T = random_segments(fp, 400, seed=5)
    Vo, Wo = O.mobility(fp, Y, F, T)
    with api.Fcm(fp, 400, precision=abi.FS_FP32) as fcm:
        fcm.set_tuning(spread_mfma=mfma)
        V, W = fcm.mobility_host(Y, F, T)
    ev, ew = rel(V, Vo), rel(W, Wo)
    print("fp32 relV %.3e relW %.3e" % (ev, ew))
    assert ev < 1e-5 and ew < 1e-5


def test_negative_positions_truncation(cuda):
    """Y from fmod can be negative: the stencil centre truncates toward 0
    (FCMfunctions.cpp:565-567)."""
    fp = params.fcm_params(32)
    Y, F, T = random_segments(fp, 64, seed=11)
    L = np.array(params.box(fp))
    Y = Y - L * (np.arange(64)[:, None] % 2)  # half the segments in (-L, 0)
    Vo, Wo = O.mobility(fp, Y, F, T)
    with api.Fcm(fp, 64) as fcm:
        V, W = fcm.mobility_host(Y, F, T)
    assert rel(V, Vo) <= 1e-12 and rel(W, Wo) <= 1e-12


def test_device_pointers_and_determinism(cuda):
    torch = cuda
    fp = params.fcm_params(48)
    Y, F, T = random_segments(fp, 300, seed=3)
    dY, dF, dT = (to_dev(torch, a) for a in (Y, F, T))
    V1 = torch.zeros_like(dY)
    W1 = torch.zeros_like(dY)
    V2 = torch.zeros_like(dY)
    W2 = torch.zeros_like(dY)
    with api.Fcm(fp, 300) as fcm:
        fcm.mobility(dY, dF, dT, V1, W1)
        fcm.mobility(dY, dF, dT, V2, W2)
        torch.cuda.synchronize()
    assert torch.equal(V1, V2) and torch.equal(W1, W2)
    Vo, Wo = O.mobility(fp, Y, F, T)
    assert rel(V1.cpu().numpy(), Vo) <= 1e-12


@pytest.mark.parametrize("batch,tile,zrun", [(1, 4, 4), (2, 8, 8), (8, 16, 16),
                                             (16, 8, 4), (4, 16, 8)])
def test_tunings_agree(cuda, batch, tile, zrun):
    fp = params.fcm_params(64)
    Y, F, T = random_segments(fp, 500, seed=2)
    Vo, Wo = O.mobility(fp, Y, F, T)
    with api.Fcm(fp, 500) as fcm:
        fcm.set_tuning(fft_batch=batch, spread_tile=tile, spread_zrun=zrun,
                       spread_mfma=0)
        V, W = fcm.mobility_host(Y, F, T)
    assert rel(V, Vo) <= 1e-12 and rel(W, Wo) <= 1e-12


def test_friction(cuda):
    torch = cuda
    fp = params.fcm_params(32)
    _, F, T = random_segments(fp, 50, seed=9)
    dF, dT = to_dev(torch, F), to_dev(torch, T)
    V = torch.zeros_like(dF)
    W = torch.zeros_like(dF)
    with api.Fcm(fp, 50) as fcm:
        fcm.friction(dF, dT, V, W)
        torch.cuda.synchronize()
    Vo, Wo = O.friction(fp, F, T)
    assert np.array_equal(V.cpu().numpy(), Vo)
    assert np.array_equal(W.cpu().numpy(), Wo)


@pytest.mark.parametrize("mfma", [1, 0])
@pytest.mark.parametrize("grid", [64, (96, 64, 48)])
def test_mobility_fp32_reference_resolution(cuda, grid, mfma):
    """FP32 at the reference's own resolution (dx = a/3.3, ngd = 20): the
    matrix-core spread's staging and tables at another support."""
    fp = params.reference_fcm_params(grid)
    Y, F, T = random_segments(fp, 300, seed=11)
    Vo, Wo = O.mobility(fp, Y, F, T)
    with api.Fcm(fp, 300, precision=abi.FS_FP32) as fcm:
        fcm.set_tuning(spread_mfma=mfma)
        V, W = fcm.mobility_host(Y, F, T)
    ev, ew = rel(V, Vo), rel(W, Wo)
    print("fp32 ref grid", grid, "mfma", mfma, "relV %.3e relW %.3e" % (ev, ew))
    assert ev < 1e-5 and ew < 1e-5
