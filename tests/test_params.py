"""Host-side suspension generators: the cubic rod packing's geometry."""
import numpy as np
from scipy.spatial import cKDTree

from filsim_amd import params


def _tangents(q):
    w, x, y, z = q.T
    return np.stack([1 - 2 * y * y - 2 * z * z, 2 * (x * y + w * z),
                     2 * (x * z - w * y)], 1)


def test_rod_packing_geometry():
    fp = params.fcm_params(256)
    box = np.array(params.box(fp))
    nfil, nw = 2000, 30
    X, q = params.rod_packing_filaments(box, nfil, nw, seed=9)
    assert X.shape == (nfil * nw, 3) and q.shape == (nfil * nw, 4)
    assert np.abs((q * q).sum(1) - 1).max() < 1e-14
    # consecutive segments are dL apart along the quaternion's tangent
    t = _tangents(q).reshape(nfil, nw, 3)
    dX = np.diff(X.reshape(nfil, nw, 3), axis=1)
    assert np.abs(dX - 2.2 * t[:, :-1]).max() < 1e-12
    # one axis per filament, a third of the filaments on each
    axis = np.abs(t[:, 0]).argmax(1)
    counts = np.bincount(axis, minlength=3)
    assert counts.min() > 0.28 * nfil
    assert np.abs(np.abs(t[:, 0]).max(1) - 1).max() < 1e-14
    # no mean orientation order
    Q = np.einsum("ni,nj->ij", t[:, 0], t[:, 0]) / nfil - np.eye(3) / 3
    assert np.abs(Q).max() < 0.02
    # segments of different filaments at least min_sep apart (periodic)
    Xw = np.mod(X, box)
    tree = cKDTree(Xw, boxsize=box)
    pairs = tree.query_pairs(2.2, output_type="ndarray")
    fil = np.arange(nfil * nw) // nw
    assert not np.any(fil[pairs[:, 0]] != fil[pairs[:, 1]])


def test_rod_packing_rejects_overfull_box():
    fp = params.fcm_params(64)
    try:
        params.rod_packing_filaments(params.box(fp), 500, 15)
    except ValueError:
        return
    raise AssertionError("500 filaments fit a 64^3 box")
