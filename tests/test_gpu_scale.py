"""FCM at the largest configuration (384^3 grid, 240,000 segments), checked
through properties that hold at any size: the FP64 mobility operator is
symmetric (<F1, M F2> = <F2, M F1> with forces and torques) and linear, and
the FP32 path agrees with FP64 to its stated bound."""
import numpy as np
import pytest

from helpers import to_dev
from filsim_amd import abi, api, params

pytestmark = pytest.mark.gpu


def test_mobility_symmetric_and_linear_384(cuda):
    torch = cuda
    fp = params.fcm_params(384)
    n = 240000
    rng = np.random.default_rng(11)
    L = np.array(params.box(fp))
    Y = to_dev(torch, rng.uniform(0, 1, (n, 3)) * L)
    F1, T1, F2, T2 = (to_dev(torch, rng.standard_normal((n, 3)))
                      for _ in range(4))
    V = torch.zeros_like(Y)
    W = torch.zeros_like(Y)
    with api.Fcm(fp, n) as fcm:
        fcm.mobility(Y, F1, T1, V, W)
        V1, W1 = V.clone(), W.clone()
        fcm.mobility(Y, F2, T2, V, W)
        V2, W2 = V.clone(), W.clone()
        fcm.mobility(Y, 2.5 * F1, 2.5 * T1, V, W)
        V3, W3 = V.clone(), W.clone()
    a = float((F1 * V2).sum() + (T1 * W2).sum())
    b = float((F2 * V1).sum() + (T2 * W1).sum())
    print("symmetry", a, b)
    assert abs(a - b) <= 1e-10 * (abs(a) + abs(b))
    lin = float(((V3 - 2.5 * V1).abs().max() + (W3 - 2.5 * W1).abs().max()) /
                (V1.abs().max() + W1.abs().max()))
    assert lin <= 1e-13
    with api.Fcm(fp, n, precision=abi.FS_FP32) as f32:
        f32.mobility(Y, F1, T1, V, W)
    e = float((V - V1).norm() / V1.norm())
    print("fp32 vs fp64 at 384^3", e)
    assert e <= 1e-5


def test_mobility_256_30k_matches_oracle(cuda):
    """The mobility microbenchmark's configuration itself: 30,000 segments
    uniform in the periodic 256^3 box (dx = a/1.86, ngd = 12), F and T
    i.i.d. N(0, 1), seed 20261202; FP64 against the oracle to 1e-12
    normwise (element maximum reported), FP32 to 1e-5."""
    import _oracle as O
    from helpers import random_segments, rel
    fp = params.fcm_params(256)
    Y, F, T = random_segments(fp, 30000, seed=20261202)
    Vo, Wo = O.mobility(fp, Y, F, T)
    with api.Fcm(fp, 30000) as fcm:
        V, W = fcm.mobility_host(Y, F, T)
    ev, ew = rel(V, Vo), rel(W, Wo)
    mv = float(np.abs(V - Vo).max() / np.abs(Vo).max())
    mw = float(np.abs(W - Wo).max() / np.abs(Wo).max())
    print("256^3 / 30k FP64: relV %.2e relW %.2e, element max %.2e %.2e"
          % (ev, ew, mv, mw), flush=True)
    assert ev <= 1e-12 and ew <= 1e-12
    with api.Fcm(fp, 30000, precision=abi.FS_FP32) as f32:
        V32, W32 = f32.mobility_host(Y, F, T)
    print("256^3 / 30k FP32: relV %.2e relW %.2e"
          % (rel(V32, Vo), rel(W32, Wo)), flush=True)
    assert rel(V32, Vo) <= 1e-5 and rel(W32, Wo) <= 1e-5
