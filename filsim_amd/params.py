"""Parameter sets for the FCM and the filament timestep.

The defaults restate multi_filament_header.hpp:47-60 (sigma = a/sqrt(pi),
sigma_dip = a/(36 pi)^(1/6), dx = sigma_dip/1.5) and the workloads the
benchmarks and tests run on (grid spacing a/1.86 with 12 support points).
"""
import math

import numpy as np

from . import abi

PI = 3.14159265358979323846264338327950288


def fcm_params(n, a=1.0, mu=1.0, dx=None, ngd=12):
    """FcmParams for an n^3 (or (nx,ny,nz)) periodic grid."""
    if isinstance(n, int):
        n = (n, n, n)
    sigma_dip = a / (36.0 * PI) ** (1.0 / 6.0)
    if dx is None:
        dx = a / 1.86
    p = abi.FcmParams()
    p.nx, p.ny, p.nz = n
    p.ngd = ngd
    p.a = a
    p.mu = mu
    p.dx = dx
    p.sigma = a / math.sqrt(PI)
    p.sigma_dip = sigma_dip
    return p


def reference_fcm_params(n, a=1.0, mu=1.0):
    """The reference's own resolution: dx = sigma_dip/1.5, ngd = 20."""
    sigma_dip = a / (36.0 * PI) ** (1.0 / 6.0)
    return fcm_params(n, a=a, mu=mu, dx=sigma_dip / 1.5, ngd=20)


def box(p):
    return (p.nx * p.dx, p.ny * p.dx, p.nz * p.dx)


def colbar_params(fp, fs_over_2a=15 * PI, cap=2.0, initial_friction_steps=0):
    c = abi.ColbarParams()
    b = box(fp)
    for d in range(3):
        c.box[d] = b[d]
    c.a = fp.a
    c.fs_over_2a = fs_over_2a
    c.cap = cap
    c.initial_friction_steps = initial_friction_steps
    return c


def sim_params(fp, dL=2.2, dt=0.01, tol=1e-4, maxiter=100, precision=0,
               impl_euler_steps=1, initial_friction_steps=0,
               weight=(0.0, 0.0, 0.0), barrier_on=1, fs_over_2a=15 * PI,
               cap=2.0, sperm_decay=0, strain_twist0=0.0):
    s = abi.SimParams()
    s.fcm = fp
    s.precision = precision
    s.dL = dL
    s.dt = dt
    s.tol = tol
    s.broyden_maxiter = maxiter
    s.impl_euler_steps = impl_euler_steps
    s.initial_friction_steps = initial_friction_steps
    for d in range(3):
        s.weight_per_length[d] = weight[d]
    s.barrier_on = barrier_on
    s.barrier_fs_over_2a = fs_over_2a
    s.barrier_cap = cap
    s.sperm_decay = sperm_decay
    s.strain_twist0 = strain_twist0
    return s


def fil_props(nfil, kap=1.0, c=None, T0=0.0, omega=0.0, wavenumber=0.0,
              phase=0.0, alpha=0.0, beta=1.0, radius=1.0):
    """Array of FilProps; scalar or per-filament sequence arguments."""
    arr = (abi.FilProps * nfil)()

    def pick(v, f):
        return v[f] if hasattr(v, "__len__") else v

    for f in range(nfil):
        pr = arr[f]
        pr.kap = pick(kap, f)
        pr.c = pick(c if c is not None else kap, f)
        pr.T0 = pick(T0, f)
        pr.omega = pick(omega, f)
        pr.wavenumber = pick(wavenumber, f)
        pr.phase = pick(phase, f)
        pr.alpha = pick(alpha, f)
        pr.beta = pick(beta, f)
        pr.radius = pick(radius, f)
    return arr


def swimming_props(nfil, nw, dL=2.2, amplitude_times_L=0.5, period=1.0,
                   kap=1.0, phases=None):
    """Travelling-wave swimmers: preferred curvature amplitude 0.5/L,
    wavenumber 2 pi/L, period 1 time unit (L = Nw*dL, the reference's
    L_filament_long)."""
    L = nw * dL
    ph = phases if phases is not None else 0.0
    return fil_props(nfil, kap=kap, T0=amplitude_times_L / L,
                     omega=2 * PI / period, wavenumber=2 * PI / L, phase=ph,
                     alpha=0.0, beta=1.0)


SWEEP_GRIDS = (32, 48, 64, 96, 128)


def sweep_point(seed, grids=SWEEP_GRIDS, nfil_range=(8, 200),
                nw_range=(15, 40), kb_range=(0.1, 10.0), g_sigma=0.2,
                precision=1, tol=1e-4, dt=0.01, maxiter=100,
                max_volume_fraction=0.05, max_length_fraction=0.9):
    """One suspension of a batched parameter sweep: grid drawn from `grids`,
    filament count uniform in nfil_range with segment counts uniform in
    nw_range, bending/twist modulus log-uniform in kb_range, gravity per
    unit length ~ N(0, g_sigma^2) along z, barrier on, no initial friction
    steps. Draws that cannot fit the box are capped and the caps returned:
    segments per filament so a filament spans at most max_length_fraction
    of the box, filament count so segments (spheres of radius a) fill at
    most max_volume_fraction of it, and fewer filaments (10% at a time)
    when the overlap-free placement still fails. Returns (fp, sp, nws,
    props, state, info)."""
    import numpy as np
    from . import api
    rng = np.random.default_rng(seed)
    grid = int(rng.choice(grids))
    fp = fcm_params(grid)
    L = min(box(fp))
    nfil = int(rng.integers(nfil_range[0], nfil_range[1] + 1))
    nws = rng.integers(nw_range[0], nw_range[1] + 1, nfil)
    nw_cap = max(2, int(max_length_fraction * L / 2.2))
    nws = np.minimum(nws, nw_cap)
    seg_vol = 4.0 / 3.0 * PI * fp.a ** 3
    seg_cap = int(max_volume_fraction * float(np.prod(box(fp))) / seg_vol)
    while nfil > 1 and int(nws[:nfil].sum()) > seg_cap:
        nfil -= 1
    nws = [int(n) for n in nws[:nfil]]
    kb = float(np.exp(rng.uniform(np.log(kb_range[0]), np.log(kb_range[1]))))
    g = float(rng.normal(0.0, g_sigma))
    sp = sim_params(fp, dt=dt, tol=tol, maxiter=maxiter, precision=precision,
                    weight=(0.0, 0.0, g), barrier_on=1,
                    initial_friction_steps=0)
    place_seed = int(rng.integers(1, 2 ** 31))
    while True:
        try:
            X, q = api.make_filaments(box(fp), nws, dL=2.2, min_sep=2.2,
                                      seed=place_seed)
            break
        except abi.FsError:
            if nfil <= 1:
                raise
            nfil = max(1, int(nfil * 0.9))
            nws = nws[:nfil]
    props = fil_props(nfil, kap=kb)
    state = dict(X=X, Xs=X, Xt=X, q=q, qs=q, qt=q)
    info = dict(grid=grid, nfil=nfil, segments=int(sum(nws)), kb=kb,
                gravity=g, nw_cap=nw_cap, segment_cap=seg_cap)
    return fp, sp, nws, props, state, info


def _qmul(a, b):
    """Hamilton product of (w, x, y, z) quaternions (UnitQuaternion.hpp
    convention: tangent = R(q) e_x)."""
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return (aw * bw - ax * bx - ay * by - az * bz,
            aw * bx + ax * bw + ay * bz - az * by,
            aw * by - ax * bz + ay * bw + az * bx,
            aw * bz + ax * by - ay * bx + az * bw)


def rod_packing_filaments(box_len, nfil, nw, dL=2.2, seed=1, jitter=0.3,
                          min_sep=2.2, roll=True):
    """Dense straight filaments with no net alignment, for suspensions too
    dense for random placement: a cubic rod packing with m x m lines along
    each of x, y and z, per period cell (c_x, c_y, c_z) = box / m

        x lines at (y, z) = (j c_y,         k c_z + c_z/2)
        y lines at (x, z) = (j c_x + c_x/2, k c_z)
        z lines at (x, y) = (j c_x,         k c_y + c_y/2)

    (all offset by c/4), so lines of different axes are c/2 apart and
    parallel lines c apart. Each line holds as many end-to-end filaments as
    fit with a gap of min_sep + 2 jitter + dL between them; nfil of these
    slots are drawn at random, each
    filament pointing either way along its line, rolled about its own axis
    by a random angle (roll=True) and its base jittered by up to `jitter`
    per axis. Equal numbers of lines along the three axes cancel the mean
    orientation order, whose coherent stresslets make an aligned lattice
    of stiff swimmers at this density unstable. m is the smallest count
    that holds nfil. Returns (X, q) like api.make_filaments; raises
    ValueError when nfil does not fit with min_sep between filaments."""
    import numpy as np
    L = np.asarray(box_len, dtype=float)
    length = (nw - 1) * dL
    per_line = [int(L[a] // (length + min_sep + 2 * jitter + dL))
                for a in range(3)]
    if min(per_line) < 1:
        raise ValueError("filament longer than the box")
    m = 1
    while sum(per_line) * m * m < nfil:
        m += 1
    c = L / m
    if 0.5 * c.min() - 2 * jitter < min_sep:
        raise ValueError("%d filaments do not fit at separation %g" %
                         (nfil, min_sep))
    # line offsets in the two transverse axes, per line axis (in cells)
    trans = {0: ((1, 0.0), (2, 0.5)), 1: ((0, 0.5), (2, 0.0)),
             2: ((0, 0.0), (1, 0.5))}
    h = np.sqrt(0.5)
    q_axis = {0: (1.0, 0.0, 0.0, 0.0), 1: (h, 0.0, 0.0, h),
              2: (h, 0.0, -h, 0.0)}
    slots = [(a, j, k, n) for a in range(3) for j in range(m)
             for k in range(m) for n in range(per_line[a])]
    rng = np.random.default_rng(seed)
    chosen = sorted(rng.choice(len(slots), nfil, replace=False))
    X = np.zeros((nfil * nw, 3))
    q = np.zeros((nfil * nw, 4))
    i = np.arange(nw)[:, None]
    for f, si in enumerate(chosen):
        a, j, k, n = slots[si]
        d = 1.0 if rng.random() < 0.5 else -1.0
        theta = rng.uniform(0.0, 2.0 * np.pi) if roll else 0.0
        base = np.zeros(3)
        (t1, o1), (t2, o2) = trans[a]
        base[t1] = (j + o1 + 0.25) * c[t1]
        base[t2] = (k + o2 + 0.25) * c[t2]
        base[a] = n * L[a] / per_line[a] + 0.25 * c[a] + \
            (0.0 if d > 0 else length)
        base += rng.uniform(-jitter, jitter, 3)
        e = np.zeros(3)
        e[a] = d
        X[f * nw:(f + 1) * nw] = base + i * dL * e
        qf = q_axis[a]
        if d < 0:  # pi about the body z axis: tangent -> -tangent
            qf = _qmul(qf, (0.0, 0.0, 0.0, 1.0))
        qf = _qmul(qf, (np.cos(0.5 * theta), np.sin(0.5 * theta), 0.0, 0.0))
        q[f * nw:(f + 1) * nw] = qf
    return X, q


def _qmul4(a, b):
    return np.array(_qmul(tuple(a), tuple(b)))


def _qsqrt(p):
    """quaSquareRootInPlace (UnitQuaternion.hpp:388-409)."""
    s, v = p[0], np.asarray(p[1:], dtype=float)
    if s < -.9999999999999:
        return np.array([0.0, 0.0, 0.0, 1.0])
    if s > .9999999999999:
        return np.array([1.0, 0.0, 0.0, 0.0])
    return np.concatenate([[np.sqrt(.5 * (1 + s))], v / np.sqrt(2. * (1 + s))])


def _qrotate(q, v):
    """RotateVector (UnitQuaternion.hpp:323-338)."""
    w, x, y, z = q
    R = np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return R @ np.asarray(v, dtype=float)


def frame_quaternion(tangent, normal):
    """Bead orientation of Bead::initialBeadSetup (Bead.hpp:151-166): the
    half-angle rotation taking e_x to the tangent, then the one taking the
    rotated e_y to the normal."""
    t = np.asarray(tangent, dtype=float)
    qt = _qsqrt(np.array([t[0], 0.0, -t[2], t[1]]))
    v = _qrotate(qt, [0.0, 1.0, 0.0])
    p = _qsqrt(np.concatenate([[np.dot(v, normal)], np.cross(v, normal)]))
    return _qmul4(p, qt)


def _periodic(d, L):
    """VecPeriodicX (spring_link_functions.cpp:131-138)."""
    return d - L * np.floor(d / L + 0.5)


def network_filaments(box_len, n_nodes, min_node_spacing, connection_radius,
                      connection_probability, spacing_away_from_node,
                      dL=2.2, a=1.0, allow_overlap=False, seed=1,
                      node_box=None, max_node_tries=100000):
    """Connected-nodes network (filament_initialisation_functions.cpp:
    259-410, spring_link_functions.cpp:8-22, 145-177), with numpy's seeded
    generator in place of Armadillo's:
      * nodes uniform in node_box (default: the periodic box), each redrawn
        while closer than min_node_spacing to an earlier node;
      * every node pair nn < mm draws u ~ U(0,1) and is joined when
        u < connection_probability and the pair is closer than
        connection_radius (minimum image);
      * a joining filament has int((d - 2 spacing_away_from_node) / dL)
        beads, centred between the nodes along the unit vector node1 ->
        node2 with bead n at CoM + (n - nbeads/2) dL t, and the frame of
        Bead::initialBeadSetup with normal t x e_x; it is skipped when any
        bead is closer than sep (2a) to a bead of a placed filament, with
        sep = 0.01 when allow_overlap else 1.05 (the reference also skips
        pairs whose centres are 2.5 filament lengths apart; here every pair
        of beads is checked). Pairs giving fewer than 2 beads are skipped;
      * every two filament ends at one node are joined by a spring link
        (spring_links::decide_spring_links).
    Returns (nws, X, q, links, nodes): segment counts, positions (np, 3),
    orientations (np, 4), spring links (n, 4) as (filament A, bead B,
    filament C, bead D) rows and the node positions (n_nodes, 3)."""
    rng = np.random.default_rng(seed)
    L = np.asarray(box_len, dtype=float)
    NB = L if node_box is None else np.asarray(node_box, dtype=float)
    nodes = []
    for _ in range(n_nodes):
        for _t in range(max_node_tries):
            p = rng.uniform(size=3) * NB
            if not nodes or np.sqrt(((_periodic(np.array(nodes) - p, L)) ** 2)
                                    .sum(1)).min() >= min_node_spacing:
                break
        else:
            raise ValueError("could not place node %d at spacing %g" %
                             (len(nodes), min_node_spacing))
        nodes.append(p)
    nodes = np.array(nodes).reshape(-1, 3)
    pairs = []
    for nn in range(n_nodes):
        for mm in range(nn + 1, n_nodes):
            if rng.uniform() < connection_probability:
                d = np.linalg.norm(_periodic(nodes[nn] - nodes[mm], L))
                if d < connection_radius:
                    pairs.append((nn, mm))
    sep = (0.01 if allow_overlap else 1.05) * 2 * a
    nws, Xs, qs = [], [], []
    at_nodes = [[] for _ in range(n_nodes)]
    placed = np.zeros((0, 3))
    ex = np.array([1.0, 0.0, 0.0])
    for n1, n2 in pairs:
        v = _periodic(nodes[n2] - nodes[n1], L)
        d = np.linalg.norm(v)
        nb = int((d - 2 * spacing_away_from_node) / dL)
        if nb < 2:
            continue
        t = v / d
        normal = np.cross(t, ex)
        normal = normal / np.linalg.norm(normal)
        com = nodes[n1] + v / 2
        X = np.array([com + (n * dL) * t - nb * dL / 2 * t for n in range(nb)])
        if len(placed):
            dd = _periodic(X[:, None, :] - placed[None, :, :], L)
            if np.sqrt((dd ** 2).sum(-1)).min() < sep:
                continue
        f = len(nws)
        nws.append(nb)
        Xs.append(X)
        qs.append(np.tile(frame_quaternion(t, normal), (nb, 1)))
        placed = np.concatenate([placed, X])
        at_nodes[n1].append((f, 0))
        at_nodes[n2].append((f, nb - 1))
    links = [[e1[0], e1[1], e2[0], e2[1]] for ends in at_nodes
             for i, e1 in enumerate(ends) for e2 in ends[i + 1:]]
    X = np.concatenate(Xs) if Xs else np.zeros((0, 3))
    q = np.concatenate(qs) if qs else np.zeros((0, 4))
    return (nws, X, q, np.array(links, dtype=np.int32).reshape(-1, 4),
            nodes)
