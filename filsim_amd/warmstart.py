"""Training driver of the learned warm start.

Record (features, target) pairs from the plain solver, normalise them, and
train the network with the library's own forward/backward/Adam kernels.
torch only allocates device buffers and draws minibatch indices.
"""
import math
import time

import numpy as np

from . import api


def record(system, steps, nt0=1, torch=None, skip=0, log=None):
    """Runs `skip` plain timesteps from nt0 (the start-up transient, kept out
    of the data), then `steps` more while recording one sample per filament
    per step. Returns device tensors X (rows x n_in), T and the Broyden
    iteration counts of the recorded steps. `log` gets one line per step."""
    for nt in range(nt0, nt0 + skip):
        it = system.timestep(nt)
        if log:
            log("start-up step %d: %d iterations" % (nt, it))
    nt0 += skip
    n_in, n_out = api.system_warmstart_dims(system)
    cap = system.nfil * steps
    X = torch.zeros((cap, n_in), dtype=torch.float32, device="cuda")
    T = torch.zeros((cap, n_out), dtype=torch.float32, device="cuda")
    api.system_record(system, X, T, cap)
    its = []
    for nt in range(nt0, nt0 + steps):
        its.append(system.timestep(nt))
        if log:
            log("recorded step %d: %d iterations" % (nt, its[-1]))
    n = api.system_record_count(system)
    api.system_record(system, None, None, 0)
    return X[:n], T[:n], its


def normalisation(X, T, torch):
    mu = X.mean(0)
    sigma = X.std(0).clamp_min(1e-6)
    tau = T.abs().mean(0).clamp_min(1e-12)
    return mu.cpu().numpy(), sigma.cpu().numpy(), tau.cpu().numpy()


def _bf16(x):
    """Round-to-nearest-even to bfloat16 (the network's input rounding)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(
        np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def ridge(X, T, mu, sigma, tau, train, hold, alpha=1e-3, max_rows=40000,
          seed=0):
    """Ridge regression of the normalised targets T/tau on the normalised
    features (BF16, as the network sees them), fitted in FP64 on the host
    over at most max_rows training rows: returns the map S (n_out x n_in),
    the intercept c and the holdout loss (mean squared normalised error, the
    network's own metric) of x_n S^T + c. alpha scales the penalty by the
    mean diagonal of the Gram matrix."""
    rng = np.random.default_rng(seed)
    if len(train) > max_rows:
        train = rng.choice(train, max_rows, replace=False)

    def norm_rows(idx):
        xi = X[idx].cpu().numpy()
        ti = T[idx].cpu().numpy()
        return (_bf16((xi - mu) / sigma).astype(np.float64),
                ti.astype(np.float64) / tau)
    Xn, Tn = norm_rows(np.sort(train))
    xm, tm = Xn.mean(0), Tn.mean(0)
    Xc = Xn - xm
    G = Xc.T @ Xc
    G[np.diag_indices_from(G)] += alpha * np.trace(G) / G.shape[0]
    S = np.linalg.solve(G, Xc.T @ (Tn - tm)).T
    c = tm - S @ xm
    Xh, Th = norm_rows(np.sort(hold))
    loss = float(np.mean((Xh @ S.T + c - Th) ** 2)) if len(hold) else None
    return S, c, loss


def train(X, T, torch, width=256, epochs=20, batch=1024, lr=1e-3, seed=1,
          holdout=0.1, log=None, lr_final=None, linear=False):
    """Trains a fresh network on (X, T); returns (mlp, history). With
    lr_final the step size follows a cosine from lr down to lr_final over
    all steps. With linear=True a ridge regression is fitted first and set
    as the network's fixed linear term (fs_mlp_set_skip, its intercept as
    the output bias, the last layer zero), so the layers learn what it
    leaves, and the parameters of the epoch with the lowest holdout loss
    are kept (the last history entry); the ridge's own holdout loss is kept
    in mlp.ridge_holdout (FP64) and mlp.ridge_holdout_net (through the
    network's BF16 arithmetic)."""
    n = X.shape[0]
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    nh = int(n * holdout)
    hold, tr = perm[:nh], perm[nh:]
    mlp = api.Mlp(X.shape[1], T.shape[1], width=width, seed=seed)
    mu, sigma, tau = normalisation(X[torch.as_tensor(tr, device="cuda")],
                                   T[torch.as_tensor(tr, device="cuda")],
                                   torch)
    mlp.set_norm(mu, sigma, tau)
    # holdout loss of a zero correction (the plain extrapolated guess)
    mlp.zero_holdout = float(np.mean(
        (T[torch.as_tensor(hold, device="cuda")].cpu().numpy() / tau) ** 2)) \
        if nh else None
    mlp.ridge_holdout = None

    def holdout_loss():
        if not nh:
            return None
        idx = torch.as_tensor(hold, device="cuda")
        hl, _ = mlp.gradients(X[idx].contiguous(), T[idx].contiguous())
        return hl
    if linear:
        S, c, rl = ridge(X, T, mu, sigma, tau, tr, hold, seed=seed)
        mlp.set_skip(S)
        # last layer zero: the network starts as the ridge fit exactly (its
        # intercept in the output bias b4, last in the packed vector)
        p = mlp.get_params()
        n_out, w = T.shape[1], mlp.width
        p[-n_out - n_out * w:-n_out] = 0.0
        p[-n_out:] = c
        mlp.set_params(p)
        mlp.ridge_holdout = rl
        mlp.ridge_holdout_net = holdout_loss()
        if log:
            log("ridge holdout %.4g (as evaluated by the network %.4g)" %
                (rl, mlp.ridge_holdout_net))
    best = (holdout_loss(), mlp.get_params()) if linear else None
    hist = []
    t0 = time.time()
    per_epoch = max(1, (len(tr) - batch) // batch + 1)
    total, step = epochs * per_epoch, 0
    for ep in range(epochs):
        order = rng.permutation(tr)
        losses = []
        for i in range(0, len(order) - batch + 1, batch):
            idx = torch.as_tensor(order[i:i + batch], device="cuda")
            lr_t = lr
            if lr_final is not None:
                lr_t = lr_final + 0.5 * (lr - lr_final) * (
                    1.0 + math.cos(math.pi * step / total))
            step += 1
            losses.append(mlp.train_step(X[idx].contiguous(),
                                         T[idx].contiguous(), lr_t,
                                         want_loss=(i == 0)))
        hl = holdout_loss()
        if best is not None and hl is not None and hl < best[0]:
            best = (hl, mlp.get_params())
        hist.append((ep, losses[0] if losses else None, hl))
        if log:
            log("epoch %d train %.4g holdout %s (%.1fs)" %
                (ep, losses[0] if losses else float("nan"),
                 "%.4g" % hl if hl is not None else "-", time.time() - t0))
    if best is not None and best[0] is not None and hist and \
            (hist[-1][2] is None or best[0] < hist[-1][2]):
        # keep the parameters of the best holdout epoch (the ridge start
        # counts as epoch -1), so the network never ends worse than it
        mlp.set_params(best[1])
        hist.append((len(hist), None, best[0]))
    return mlp, hist


class OnlineRefresh:
    """Keeps an attached network current on a long run: records the
    learned path's own converged steps (features, converged - extrapolated
    target) into a window of `window` steps; each time the window is full,
    `epochs` passes of Adam at `lr` over it, then a new window starts. The
    normalisation and the linear term stay those of the first training.
    Call after_step() after every timestep."""

    def __init__(self, system, mlp, torch, window=100, epochs=1, lr=2e-4,
                 batch=1024, seed=3):
        self.system, self.mlp, self.torch = system, mlp, torch
        self.epochs, self.lr, self.batch = epochs, lr, batch
        self.rng = np.random.default_rng(seed)
        n_in, n_out = api.system_warmstart_dims(system)
        self.cap = system.nfil * window
        self.X = torch.zeros((self.cap, n_in), dtype=torch.float32,
                             device="cuda")
        self.T = torch.zeros((self.cap, n_out), dtype=torch.float32,
                             device="cuda")
        self.refreshes = 0
        api.system_record(system, self.X, self.T, self.cap)

    def after_step(self):
        if api.system_record_count(self.system) < self.cap:
            return False
        api.system_record(self.system, None, None, 0)
        for _ in range(self.epochs):
            order = self.rng.permutation(self.cap)
            for i in range(0, self.cap - self.batch + 1, self.batch):
                idx = self.torch.as_tensor(order[i:i + self.batch],
                                           device="cuda")
                self.mlp.train_step(self.X[idx].contiguous(),
                                    self.T[idx].contiguous(), self.lr,
                                    want_loss=False)
        self.refreshes += 1
        api.system_record(self.system, self.X, self.T, self.cap)
        return True

    def close(self):
        api.system_record(self.system, None, None, 0)
