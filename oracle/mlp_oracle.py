"""FP64 CPU oracle of the warm-start MLP (test infrastructure only).

Restates the network of filsim_amd/csrc/mlp.hip -- n_in -> W -> W -> W ->
n_out with ReLU, BF16 weights/activations, FP32/FP64 accumulation -- with
the BF16 rounding applied at the same points (inputs after normalisation,
weights, hidden activations, backward signals), so the GPU's MFMA results
can be compared against it. With `bf16=False` it is the exact FP64 network
whose gradients tests pin by central differences.

No reference counterpart (the reference has no learned component).
"""
import numpy as np


def bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params, dims):
    """Split the packed vector (W1,b1,...,W4,b4; W row-major [out][in])."""
    out, o = [], 0
    for l in range(4):
        n_in, n_out = dims[l], dims[l + 1]
        W = params[o:o + n_in * n_out].reshape(n_out, n_in)
        o += n_in * n_out
        b = params[o:o + n_out]
        o += n_out
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


def forward(params, dims, X, mu, sigma, tau, rounding=True, skip=None):
    """Normalised output y_n (and the activations); skip: the optional fixed
    linear map S (n_out x n_in), y_n += x_n S^T."""
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params, dims)
    if rounding:  # the GPU normalises in FP32, then rounds to BF16
        h = bf16((np.asarray(X, np.float32) - np.asarray(mu, np.float32)) /
                 np.asarray(sigma, np.float32))
    else:
        h = (np.asarray(X, np.float64) - mu) / sigma
    zs, hs = [], [h]
    for l in range(3):
        W, b = layers[l]
        z = h @ r(W).T + b
        zs.append(z)
        h = r(np.maximum(z, 0.0))
        hs.append(h)
    W, b = layers[3]
    yn = h @ r(W).T + b
    if skip is not None:
        yn = yn + hs[0] @ r(skip).T
    return yn, zs, hs, layers


def loss_and_grad(params, dims, X, T, mu, sigma, tau, rounding=True,
                  skip=None):
    """Mean squared normalised error and its gradient (packed like params);
    the fixed skip map has no gradient."""
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    yn, zs, hs, layers = forward(params, dims, X, mu, sigma, tau, rounding,
                                 skip)
    tn = np.asarray(T, np.float64) / tau
    d = yn - tn
    n = d.size
    loss = (d * d).sum() / n
    g = 2.0 * d / n
    grads = [None] * 4
    for l in range(3, -1, -1):
        W, _ = layers[l]
        gb = r(g)
        dW = gb.T @ hs[l]
        db = g.sum(0)
        grads[l] = (dW, db)
        if l > 0:
            dh = gb @ r(W)
            g = dh * (zs[l - 1] > 0)
    flat = []
    for dW, db in grads:
        flat.append(dW.ravel())
        flat.append(db.ravel())
    return loss, np.concatenate(flat)
