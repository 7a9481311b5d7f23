"""Warm-start network: forward and gradients of the BF16 MFMA kernels
against the FP64 oracle (same BF16 rounding points), training reduces the
loss, and the system-level fallback/gain."""
import os
import sys

import numpy as np
import pytest

from filsim_amd import api, params, warmstart
from sysfix import config

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _net(torch, n_in=446, n_out=180, width=256, B=96, seed=0):
    rng = np.random.default_rng(seed)
    mlp = api.Mlp(n_in, n_out, width=width, seed=seed + 1)
    mu = rng.standard_normal(n_in).astype(np.float32) * 0.1
    sig = (1 + rng.random(n_in)).astype(np.float32)
    tau = (0.5 + rng.random(n_out)).astype(np.float32)
    mlp.set_norm(mu, sig, tau)
    X = rng.standard_normal((B, n_in)).astype(np.float32)
    T = rng.standard_normal((B, n_out)).astype(np.float32)
    dX = torch.tensor(X, device="cuda")
    dT = torch.tensor(T, device="cuda")
    return mlp, X, T, dX, dT, mu, sig, tau


def test_forward_matches_oracle(cuda):
    torch = cuda
    mlp, X, T, dX, dT, mu, sig, tau = _net(torch)
    Y = torch.zeros((X.shape[0], 180), dtype=torch.float32, device="cuda")
    mlp.forward(dX, Y)
    p = mlp.get_params()
    yn, _, _, _ = M.forward(p, mlp.dims, X, mu, sig, tau)
    Yo = yn * tau
    Yg = Y.cpu().numpy()
    err = np.linalg.norm(Yg - Yo) / np.linalg.norm(Yo)
    print("forward rel err", err)
    # same BF16 rounding points; the rest is FP32 vs FP64 accumulation and
    # the rare hidden value that rounds to the other BF16 neighbour
    assert err < 2e-3
    # with the fixed linear term y_n += x_n S^T
    S = np.random.default_rng(5).standard_normal((180, X.shape[1])) * 0.05
    mlp.set_skip(S)
    mlp.forward(dX, Y)
    yn2, _, _, _ = M.forward(p, mlp.dims, X, mu, sig, tau, skip=S)
    Yo2 = yn2 * tau
    err2 = np.linalg.norm(Y.cpu().numpy() - Yo2) / np.linalg.norm(Yo2)
    print("forward with linear term rel err", err2)
    assert err2 < 2e-3
    assert np.linalg.norm(Yo2 - Yo) > 0.1 * np.linalg.norm(Yo)


def test_gradients_match_oracle(cuda):
    torch = cuda
    mlp, X, T, dX, dT, mu, sig, tau = _net(torch, seed=3)
    loss, g = mlp.gradients(dX, dT)
    p = mlp.get_params()
    lo, go = M.loss_and_grad(p, mlp.dims, X, T, mu, sig, tau)
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    cos = g @ go / np.linalg.norm(g) / np.linalg.norm(go)
    print("loss", loss, lo, "grad rel err", err, "cos", cos)
    assert abs(loss - lo) < 1e-5 * lo
    assert err < 1e-4 and cos > 0.99999
    # per-layer: every weight block agrees
    o = 0
    for l in range(4):
        n = mlp.dims[l] * mlp.dims[l + 1] + mlp.dims[l + 1]
        e = np.linalg.norm(g[o:o + n] - go[o:o + n]) / np.linalg.norm(go[o:o + n])
        assert e < 1e-3, (l, e)
        o += n


def test_training_reduces_loss(cuda):
    torch = cuda
    rng = np.random.default_rng(4)
    n_in, n_out = 40, 24
    A = rng.standard_normal((n_in, n_out)) / np.sqrt(n_in)
    X = rng.standard_normal((4096, n_in)).astype(np.float32)
    T = (np.tanh(X @ A) + 0.1 * X[:, :n_out]).astype(np.float32)
    dX = torch.tensor(X, device="cuda")
    dT = torch.tensor(T, device="cuda")
    mlp, hist = warmstart.train(dX, dT, torch, width=64, epochs=15,
                                batch=256, lr=3e-3)
    first, last = hist[0][2], hist[-1][2]
    print("holdout loss", first, "->", last)
    assert last < 0.2 * first


def _swimmers(precision=1):
    sp, nws, props, st = config(grid=64, nfil=12, nw=15, seed=21, tol=1e-6,
                                swim=True, dt=0.005, kap=1000.0,
                                weight=(0, 0, 0), precision=precision)
    return sp, nws, props, st


def test_fallback_is_exact_plain_path(cuda):
    """A network whose guess is far off is rejected every step, and the
    trajectory is then bit-identical to the plain extrapolated path."""
    sp, nws, props, st = _swimmers()
    plain = api.System(sp, nws, props)
    plain.set_state(**st)
    ws = api.System(sp, nws, props)
    ws.set_state(**st)
    n_in, n_out = api.system_warmstart_dims(ws)
    bad = api.Mlp(n_in, n_out, width=64, seed=9)
    bad.set_norm(np.zeros(n_in), np.ones(n_in), 1e3 * np.ones(n_out))
    api.system_set_warmstart(ws, bad)
    for nt in range(1, 4):
        a = plain.timestep(nt)
        b = ws.timestep(nt)
        used, rp, rm = api.system_last_step_info(ws)
        assert a == b and used == 0 and rm > rp
    sa, sb = plain.get_state(), ws.get_state()
    assert np.array_equal(sa["X"], sb["X"]) and np.array_equal(sa["q"], sb["q"])


def test_trained_warmstart_gain(cuda):
    """Train on the plain solver's own trajectory (the network with its ridge
    linear term), then continue from the recorded state with and without
    it for 30 steps (same states, same tolerance). The recording system
    itself continues with the network, so its velocity history is the
    trajectory's own. The network must save at least 2 Broyden iterations
    per step on average (measured: about 5 of 11), and its holdout loss
    must not exceed the ridge fit it starts from."""
    torch = cuda
    sp, nws, props, st = _swimmers()
    rec = api.System(sp, nws, props)
    rec.set_state(**st)
    X, T, its_rec = warmstart.record(rec, 60, torch=torch, skip=5)
    mlp, hist = warmstart.train(X, T, torch, epochs=40, batch=128, lr=1e-3,
                                linear=True)
    state = rec.get_state()
    keep = {k: state[k] for k in ("X", "Xt", "U", "Ut", "q", "qt", "lam",
                                  "lam1", "lam2")}
    keep.update(Xs=state["X"], Us=state["U"], qs=state["q"])
    base = api.System(sp, nws, props)
    base.set_state(**keep)
    api.system_set_warmstart(rec, mlp)
    ib, il, used = [], [], 0
    for nt in range(66, 96):
        ib.append(base.timestep(nt))
        il.append(rec.timestep(nt))
        used += api.system_last_step_info(rec)[0]
    print("plain iters", ib, "learned", il, "mlp used", used,
          "holdout network %.3f ridge %.3f extrapolation %.3f" %
          (hist[-1][2], mlp.ridge_holdout, mlp.zero_holdout))
    assert hist[-1][2] <= mlp.ridge_holdout_net
    assert mlp.ridge_holdout < 0.5 * mlp.zero_holdout
    assert np.mean(il) <= np.mean(ib) - 2.0


def test_check_mode_fallback_is_exact_plain_path(cuda):
    """With the extrapolated guess solved only on every third step, a
    network whose guess is far off is still rejected on every step (then the
    extrapolated guess is solved after all), and the trajectory is
    bit-identical to the plain path."""
    sp, nws, props, st = _swimmers()
    plain = api.System(sp, nws, props)
    plain.set_state(**st)
    ws = api.System(sp, nws, props)
    ws.set_state(**st)
    n_in, n_out = api.system_warmstart_dims(ws)
    bad = api.Mlp(n_in, n_out, width=64, seed=9)
    bad.set_norm(np.zeros(n_in), np.ones(n_in), 1e3 * np.ones(n_out))
    api.system_set_warmstart(ws, bad)
    api.system_set_warmstart_check(ws, 3)
    for nt in range(1, 6):
        a = plain.timestep(nt)
        b = ws.timestep(nt)
        used, rp, rm = api.system_last_step_info(ws)
        assert a == b and used == 0 and np.isfinite(rp) and rm > rp
    sa, sb = plain.get_state(), ws.get_state()
    assert np.array_equal(sa["X"], sb["X"]) and np.array_equal(sa["q"], sb["q"])


def test_check_mode_matches_every_step_mode(cuda):
    """Two recorders with the same trajectory and history continue with the
    same trained network, one solving the extrapolated guess every step, the
    other only every 4th step (or when the network's guess does not beat the
    last measured extrapolation residual). Skipping that solve changes
    nothing else: while the two runs have made the same choices their
    network residuals, iteration counts and states agree bit for bit, and
    the check mode solves the extrapolated guess on fewer steps."""
    torch = cuda
    sp, nws, props, st = _swimmers()
    runs = []
    for _ in range(2):
        s = api.System(sp, nws, props)
        s.set_state(**st)
        X, T, _ = warmstart.record(s, 60, torch=torch, skip=5)
        runs.append(s)
    mlp, _ = warmstart.train(X, T, torch, epochs=40, batch=128, lr=1e-3,
                             linear=True)
    a, b = runs
    api.system_set_warmstart(a, mlp)
    api.system_set_warmstart(b, mlp)
    api.system_set_warmstart_check(b, 4)
    same, agreed, solved, used_b = True, 0, 0, 0
    for nt in range(66, 96):
        ia = a.timestep(nt)
        ua, pa, ma = api.system_last_step_info(a)
        ib = b.timestep(nt)
        ub, pb, mb = api.system_last_step_info(b)
        assert np.isfinite(pa)
        solved += int(np.isfinite(pb))
        used_b += ub
        if same:
            assert ma == mb  # same state: same network guess and residual
            if ua != ub:
                same = False
                continue
            assert ia == ib
            assert np.array_equal(a.get_state()["X"], b.get_state()["X"])
            agreed += 1
    print("steps agreeing", agreed, "extrapolated solves", solved,
          "network used", used_b)
    assert agreed >= 5
    assert solved <= 20 and used_b >= 15
