"""Pins the oracle FCM against closed forms and exact discrete identities."""
import numpy as np
import pytest

import _oracle as O
from filsim_amd import params
from helpers import random_segments, rel


def test_projection_closed_form():
    """The reference's own dormant check (FCMfunctions.cpp:252-286): body
    force f_z = sin(2 pi x / L) gives u_z = f_z / (mu k^2), u_x = u_y = 0."""
    for mu in (1.0, 2.5):
        fp = params.fcm_params((32, 16, 24), mu=mu)
        L = 32 * fp.dx
        x = np.arange(32) * fp.dx
        f = np.zeros((3, 32, 16, 24))
        f[2] = np.sin(2 * np.pi * x / L)[:, None, None]
        c3 = np.stack([O.rfft3d(f[i]) for i in range(3)])
        u3 = O.project(fp, c3)
        u = np.stack([O.irfft3d(u3[i], 24) for i in range(3)])
        k = 2 * np.pi / L
        assert np.abs(u[2] - f[2] / (mu * k * k)).max() < 1e-12 / k ** 2
        assert np.abs(u[:2]).max() < 1e-12


def test_projection_divergence_free_and_mean_zero():
    fp = params.fcm_params((16, 20, 12))
    rng = np.random.default_rng(0)
    f = rng.standard_normal((3, 16, 20, 12))
    c3 = np.stack([O.rfft3d(f[i]) for i in range(3)])
    u3 = O.project(fp, c3)
    # the reference's convention: the Nyquist index carries +N/2
    # (FCMfunctions.cpp:133-153)
    def wn(n):
        i = np.arange(n)
        return 2 * np.pi / (n * fp.dx) * np.where(i <= n // 2, i, i - n)
    kx, ky = wn(16), wn(20)
    kz = 2 * np.pi * np.fft.rfftfreq(12, fp.dx)
    K = np.meshgrid(kx, ky, kz, indexing="ij")
    div = sum(K[i] * u3[i] for i in range(3))
    assert np.abs(div).max() < 1e-12 * np.abs(u3).max()
    assert np.abs(u3[:, 0, 0, 0]).max() == 0.0


def _mob(fp, Y, F, T):
    return O.mobility(fp, Y, F, T)


def test_adjointness_and_linearity():
    """<F1, M F2> = <F2, M F1> including torques (same Gaussians and dx^3
    weights on both sides of the FCM operator), and M is linear."""
    fp = params.fcm_params(24)
    Y, F1, T1 = random_segments(fp, 30, seed=1)
    _, F2, T2 = random_segments(fp, 30, seed=2)
    V1, W1 = _mob(fp, Y, F1, T1)
    V2, W2 = _mob(fp, Y, F2, T2)
    a = (F1 * V2).sum() + (T1 * W2).sum()
    b = (F2 * V1).sum() + (T2 * W1).sum()
    assert abs(a - b) <= 1e-12 * abs(a)
    V3, W3 = _mob(fp, Y, 2 * F1 - 3 * F2, 2 * T1 - 3 * T2)
    assert rel(V3, 2 * V1 - 3 * V2) < 1e-13
    assert rel(W3, 2 * W1 - 3 * W2) < 1e-13


def test_translation_by_whole_cells():
    fp = params.fcm_params((24, 20, 28))
    Y, F, T = random_segments(fp, 25, seed=4)
    V, W = _mob(fp, Y, F, T)
    # positive shift: positions stay positive, so the truncated stencil
    # centre (FCMfunctions.cpp:565-567) moves by exactly the same cells
    shift = np.array([3, 5, 7]) * fp.dx
    V2, W2 = _mob(fp, Y + shift, F, T)
    assert rel(V2, V) < 1e-12 and rel(W2, W) < 1e-12


def test_grid_sum_normalisation():
    """sum_grid f dx^3 = sum F for the monopole (Gaussian normalisation at
    discretisation accuracy) and the dipole part integrates to zero."""
    fp = params.reference_fcm_params(32)
    Y, F, T = random_segments(fp, 10, seed=6)
    f3 = O.spread(fp, Y, F, 0 * T)
    tot = f3.reshape(3, -1).sum(1) * fp.dx ** 3
    assert np.abs(tot - F.sum(0)).max() < 1e-6 * np.abs(F).sum()
    f3t = O.spread(fp, Y, 0 * F, T)
    assert np.abs(f3t.reshape(3, -1).sum(1)).max() * fp.dx ** 3 < 1e-8


def test_spread_brute_force_tiny_grid():
    """Per-grid-point brute force of the truncated-Gaussian spread."""
    fp = params.fcm_params((12, 14, 16), ngd=6)
    Y, F, T = random_segments(fp, 4, seed=8)
    f3 = O.spread(fp, Y, F, T)
    n = np.array([12, 14, 16])
    s2 = fp.sigma ** 2
    sd2 = fp.sigma_dip ** 2
    ref = np.zeros((3, 12, 14, 16))
    for s in range(4):
        c = np.trunc(Y[s] / fp.dx).astype(int)
        offs = [np.arange(c[d] - 3 + 1, c[d] - 3 + 7) for d in range(3)]
        for ix in offs[0]:
            for iy in offs[1]:
                for iz in offs[2]:
                    r = np.array([ix, iy, iz]) * fp.dx - Y[s]
                    g = np.exp(-(r ** 2).sum() / (2 * s2)) / (2 * np.pi * s2) ** 1.5
                    gd = np.exp(-(r ** 2).sum() / (2 * sd2)) / (2 * np.pi * sd2) ** 1.5
                    # dipole of a torque: 1/2 curl(T g_D) = 1/2 grad g_D x T
                    grad = -r / sd2 * gd
                    w = np.array([ix, iy, iz]) % n
                    ref[:, w[0], w[1], w[2]] += F[s] * g + 0.5 * np.cross(grad, T[s])
    assert rel(f3, ref) < 1e-13


@pytest.mark.parametrize("n", [48, 64])
def test_self_mobility_hasimoto(n):
    """One segment in a periodic box: translational mobility approaches
    1/(6 pi mu a) (1 - 2.8373 a/L + 4.19 (a/L)^3) and rotational
    1/(8 pi mu a^3) (1 - 4 pi a^3 / (3 L^3)) at the reference resolution."""
    fp = params.reference_fcm_params(n)
    L = n * fp.dx
    Y = np.array([[0.31 * L, 0.52 * L, 0.77 * L]])
    V, _ = _mob(fp, Y, np.array([[1.0, 0, 0]]), np.zeros((1, 3)))
    has = (1 - 2.837297 / L + 4.19 / L ** 3) / (6 * np.pi)
    assert abs(V[0, 0] - has) < 2e-3 * has
    _, W = _mob(fp, Y, np.zeros((1, 3)), np.array([[0, 0, 1.0]]))
    rot = (1 - 4 * np.pi / (3 * L ** 3)) / (8 * np.pi)
    assert abs(W[0, 2] - rot) < 1e-6 * rot
