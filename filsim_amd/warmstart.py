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


def train(X, T, torch, width=256, epochs=20, batch=1024, lr=1e-3, seed=1,
          holdout=0.1, log=None, lr_final=None):
    """Trains a fresh network on (X, T); returns (mlp, history). With
    lr_final the step size follows a cosine from lr down to lr_final over
    all steps."""
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
        hl = None
        if nh:
            idx = torch.as_tensor(hold, device="cuda")
            hl, _ = mlp.gradients(X[idx].contiguous(), T[idx].contiguous())
        hist.append((ep, losses[0] if losses else None, hl))
        if log:
            log("epoch %d train %.4g holdout %s (%.1fs)" %
                (ep, losses[0] if losses else float("nan"),
                 "%.4g" % hl if hl is not None else "-", time.time() - t0))
    return mlp, hist
