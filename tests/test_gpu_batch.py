"""Batched sweep and autotuner on the GPU.

A batched suspension must reproduce its unbatched run bit for bit (same
kernels, launch order and reductions; only the loop control moves to the
device), whatever the other suspensions in the batch do. The autotuner may
only change launch shapes, so the FP64 solve stays within the oracle bar."""
import numpy as np
import pytest

import _oracle as O
from helpers import random_segments, rel
from filsim_amd import abi, api, params

pytestmark = pytest.mark.gpu

FIELDS = ("X", "Xt", "U", "Ut", "q", "qt", "lam", "lam1", "lam2")


def _points(n, seed, precision):
    pts = []
    for k in range(n):
        fp, sp, nws, props, st, info = params.sweep_point(
            seed + k, grids=(32, 48, 64), nfil_range=(3, 14),
            nw_range=(6, 20), precision=precision, tol=1e-9, dt=0.01)
        pts.append((sp, nws, props, st, info))
    return pts


@pytest.mark.parametrize("precision", [abi.FS_FP32, abi.FS_FP64])
def test_batched_equals_unbatched(cuda, precision):
    pts = _points(4, 100 + precision, precision)
    print([p[4] for p in pts])
    solo, bat = [], []
    for sp, nws, props, st, info in pts:
        for group in (solo, bat):
            s = api.System(sp, nws, props)
            s.set_state(**st)
            group.append(s)
    batch = api.Batch(bat)
    for nt in range(1, 5):
        # fused FCM stages on 1 stream, per-system graphs on 3 streams, one
        # parent graph, fused on 3 streams
        batch.set_fused(nt in (1, 4))
        batch.set_lanes(0 if nt in (1, 3) else 3)
        ib = batch.timestep(nt)
        isolo = [s.timestep(nt) for s in solo]
        print("step", nt, "batched iters", ib, "solo", isolo)
        assert ib == isolo
    for a, b in zip(solo, bat):
        sa, sb = a.get_state(), b.get_state()
        for k in FIELDS:
            assert np.array_equal(sa[k], sb[k]), k
    batch.close()
    # after the batch is gone the systems step on their own as before
    for a, b in zip(solo, bat):
        assert a.timestep(5) == b.timestep(5)
        assert np.array_equal(a.get_state()["X"], b.get_state()["X"])


@pytest.mark.parametrize("precision", [abi.FS_FP32, abi.FS_FP64])
def test_fused_groups_equal_unbatched(cuda, precision):
    """Fused FCM stages with groups of several members: three suspensions
    on 48^3 grids (run-time FFT plans) and two on 64^3 (compile-time plans),
    different filament and segment counts, stepped as one batch with fused
    stages and one after another: same iteration counts, same bits."""
    pts = []
    for k, grid in enumerate((48, 64, 48, 64, 48)):
        fp, sp, nws, props, st, info = params.sweep_point(
            300 + k, grids=(grid,), nfil_range=(3, 14), nw_range=(6, 20),
            precision=precision, tol=1e-9, dt=0.01)
        pts.append((sp, nws, props, st, info))
    print([(p[4]["grid"], p[4]["segments"]) for p in pts])
    solo, bat = [], []
    for sp, nws, props, st, info in pts:
        for group in (solo, bat):
            s = api.System(sp, nws, props)
            s.set_state(**st)
            group.append(s)
    batch = api.Batch(bat)
    batch.set_fused(True)
    for nt in range(1, 4):
        ib = batch.timestep(nt)
        isolo = [s.timestep(nt) for s in solo]
        print("step", nt, "fused iters", ib, "solo", isolo)
        assert ib == isolo
    for a, b in zip(solo, bat):
        sa, sb = a.get_state(), b.get_state()
        for k in FIELDS:
            assert np.array_equal(sa[k], sb[k]), k
    batch.close()


def test_batch_rejects_friction_steps(cuda):
    sp, nws, props, st, info = _points(1, 7, abi.FS_FP32)[0]
    sp.initial_friction_steps = 1
    s = api.System(sp, nws, props)
    s.set_state(**st)
    with pytest.raises(abi.FsError):
        api.Batch([s])


def test_autotune_keeps_parity(cuda):
    fp = params.fcm_params(64)
    Y, F, T = random_segments(fp, 400, seed=5)
    Vo, Wo = O.mobility(fp, Y, F, T)
    with api.Fcm(fp, 400) as fcm:
        t = fcm.autotune(400, reps=2)
        print("tuned", t.fft_batch, t.spread_tile, t.spread_zrun, t.fft_xlines)
        got = fcm.get_tuning()
        assert (got.fft_batch, got.spread_tile, got.spread_zrun,
                got.fft_xlines) == (t.fft_batch, t.spread_tile,
                                    t.spread_zrun, t.fft_xlines)
        V, W = fcm.mobility_host(Y, F, T)
    assert rel(V, Vo) <= 1e-12 and rel(W, Wo) <= 1e-12
