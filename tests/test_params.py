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


def test_frame_quaternion_maps_the_frame():
    rng = np.random.default_rng(0)
    for _ in range(50):
        t = rng.normal(size=3)
        t /= np.linalg.norm(t)
        n = np.cross(t, [1.0, 0.0, 0.0])
        n /= np.linalg.norm(n)
        q = params.frame_quaternion(t, n)
        assert abs(np.dot(q, q) - 1) < 1e-12
        assert np.abs(_tangents(q[None])[0] - t).max() < 1e-12
        w, x, y, z = q
        ey = [2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)]
        assert np.abs(np.array(ey) - n).max() < 1e-12


def test_network_filaments():
    fp = params.fcm_params(96)
    box = np.array(params.box(fp))
    nws, X, q, links, nodes = params.network_filaments(
        box, n_nodes=40, min_node_spacing=8.0, connection_radius=20.0,
        connection_probability=0.5, spacing_away_from_node=2.0, seed=3)
    nf = len(nws)
    assert nf > 10 and X.shape == (sum(nws), 3) and q.shape == (sum(nws), 4)
    # nodes respect their spacing (minimum image)
    d = nodes[:, None] - nodes[None]
    d -= box * np.floor(d / box + 0.5)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(len(nodes)) * 1e9
    assert r.min() >= 8.0
    off = np.concatenate([[0], np.cumsum(nws)])
    ends = {}
    for f in range(nf):
        Xf, qf = X[off[f]:off[f + 1]], q[off[f]:off[f + 1]]
        t = _tangents(qf)
        step = np.diff(Xf, axis=0)
        assert np.abs(step - 2.2 * t[:-1]).max() < 1e-12
        # each end lies on the line to its node, bead n at
        # CoM + (n - nb/2) dL t: bead 0 is s + [0, dL/2) before node 1, the
        # last bead s + [dL, 3 dL/2) before node 2 (s = 2)
        for end, sgn, lo in ((0, -1.0, 2.0), (nws[f] - 1, 1.0, 4.2)):
            v = nodes - Xf[end]
            v -= box * np.floor(v / box + 0.5)
            along = v @ t[0] * sgn
            perp = np.linalg.norm(v - np.outer(v @ t[0], t[0]), axis=1)
            k = int(np.argmin(perp + 1e3 * (along < 0)))
            assert perp[k] < 1e-9 and lo - 1e-9 <= along[k] < lo + 1.1
            ends.setdefault(k, []).append((f, end))
    # beads of different filaments keep 1.05 * 2a apart
    fil = np.repeat(np.arange(nf), nws)
    dd = X[:, None] - X[None]
    dd -= box * np.floor(dd / box + 0.5)
    rr = np.sqrt((dd ** 2).sum(-1))
    assert rr[fil[:, None] != fil[None]].min() >= 2.1
    # spring links: every pair of ends at a node, nothing else
    want = sorted(tuple(sorted([a, b])) for e in ends.values()
                  for i, a in enumerate(e) for b in e[i + 1:])
    got = sorted(tuple(sorted([(r[0], r[1]), (r[2], r[3])])) for r in links)
    assert got == want and len(links) > 0
