"""Spring-link networks (spring_link_functions.cpp:24-45) on the GPU against
the oracle: forces after one force assembly, whole timesteps with the
friction and FCM mobilities, and the batched sweep with links."""
import numpy as np
import pytest

from filsim_amd import api
from helpers import rel
from sysfix import config, d2h, pair, rich_state

pytestmark = pytest.mark.gpu


def _links(nws, seed=0):
    """Filament ends joined head to tail, plus bead 0 of filament 0 in two
    more links (a node with three springs), plus a few random pairs."""
    rng = np.random.default_rng(seed)
    nf = len(nws)
    rows = [[f, nws[f] - 1, f + 1, 0] for f in range(nf - 1)]
    rows += [[0, 0, 1, 0], [0, 0, 2, 0]]
    for _ in range(4):
        a, c = rng.choice(nf, 2, replace=False)
        rows.append([int(a), int(rng.integers(nws[a])), int(c),
                     int(rng.integers(nws[c]))])
    return np.array(rows, dtype=np.int32)


def test_spring_forces_match_oracle(cuda):
    sp, nws, props, st0 = config(swim=True, seed=5)
    st = rich_state(st0, nws, seed=1)
    links = _links(nws)
    o, g = pair(sp, nws, props, st)
    o.set_spring_links(links, 3.0, 2.2)
    g.set_spring_links(links, 3.0, 2.2)
    o.forces(14)
    g.forces(14)
    Fo, Fg = o.get_state()["F"], g.get_state()["F"]
    assert rel(Fg, Fo) <= 1e-14
    _, Fp, _, _, _ = g.fcm_ptrs()
    assert rel(d2h(Fp, 3 * g.np), Fo.ravel()) <= 1e-14
    # removing them restores the plain forces
    g.set_spring_links(np.zeros((0, 4), np.int32), 3.0, 2.2)
    g.forces(14)
    o2, _ = pair(sp, nws, props, st)
    o2.forces(14)
    assert rel(g.get_state()["F"], o2.get_state()["F"]) <= 1e-14


def test_spring_link_validation(cuda):
    sp, nws, props, st = config(nfil=3, nw=8, seed=2)
    g = api.System(sp, nws, props)
    with pytest.raises(Exception):
        g.set_spring_links([[0, 8, 1, 0]], 1.0, 1.0)   # bead out of range
    with pytest.raises(Exception):
        g.set_spring_links([[3, 0, 1, 0]], 1.0, 1.0)   # no filament 3
    with pytest.raises(Exception):
        g.set_spring_links([[1, 2, 1, 2]], 1.0, 1.0)   # bead to itself
    g.close()


@pytest.mark.parametrize("mob", ["friction", "fcm"])
def test_whole_timesteps_with_springs(cuda, mob):
    if mob == "friction":
        sp, nws, props, st = config(swim=True, seed=8, tol=1e-9, ifs=100)
        steps = 8
    else:
        sp, nws, props, st = config(grid=48, nfil=6, nw=10, swim=True,
                                    seed=8, tol=1e-8)
        steps = 2
    links = _links(nws, seed=3)
    o, g = pair(sp, nws, props, st)
    o.set_spring_links(links, 2.0, 3.0)
    g.set_spring_links(links, 2.0, 3.0)
    for nt in range(1, steps + 1):
        io = o.timestep(nt)
        ig = g.timestep(nt)
        assert abs(io - ig) <= 1, (nt, io, ig)
    so, sg = o.get_state(), g.get_state()
    Lb = np.abs(so["X"]).max()
    assert np.abs(sg["X"] - so["X"]).max() <= 1e-8 * Lb
    assert np.abs(sg["q"] - so["q"]).max() <= 1e-8
    g.close()


def test_batch_with_springs_is_bitwise(cuda):
    """A batched system with spring links reproduces its own unbatched
    timesteps bit for bit."""
    sp, nws, props, st = config(grid=48, nfil=6, nw=10, swim=True, seed=8,
                                tol=1e-6)
    links = _links(nws, seed=4)
    a = api.System(sp, nws, props)
    b = api.System(sp, nws, props)
    for s in (a, b):
        s.set_state(**st)
        s.set_spring_links(links, 2.0, 3.0)
    batch = api.Batch([b])
    for nt in range(1, 4):
        ia = a.timestep(nt)
        ib = batch.timestep(nt)[0]
        assert ia == ib, (nt, ia, ib)
    sa, sb = a.get_state(), b.get_state()
    for k in ("X", "q", "lam", "U"):
        assert np.array_equal(sa[k], sb[k]), k
    batch.close()
    a.close()
    b.close()


@pytest.mark.parametrize("ifs", [100, 0])
def test_network_timesteps_match_oracle(cuda, ifs):
    """A connected-nodes network (filaments of different lengths between
    random nodes, spring links at the nodes) stepped under the friction
    (ifs=100) and FCM (ifs=0) mobilities."""
    from filsim_amd import params
    import _oracle as O
    fp = params.fcm_params(48)
    box = params.box(fp)
    nws, X, q, links, _ = params.network_filaments(
        box, n_nodes=14, min_node_spacing=7.0, connection_radius=16.0,
        connection_probability=0.6, spacing_away_from_node=1.5, seed=4)
    assert len(nws) >= 4 and len(links) >= 2
    sp = params.sim_params(fp, dt=0.01, tol=1e-9, maxiter=100,
                           weight=(0.0, 0.0, -0.5), barrier_on=1,
                           initial_friction_steps=ifs)
    props = params.fil_props(len(nws), kap=5.0)
    st = dict(X=X, Xs=X, Xt=X, q=q, qs=q, qt=q)
    o = O.Sim(sp, nws, props)
    g = api.System(sp, nws, props)
    for s in (o, g):
        s.set_state(**st)
        s.set_spring_links(links, 4.0, 2.2)
    for nt in range(1, 4):
        io, ig = o.timestep(nt), g.timestep(nt)
        assert abs(io - ig) <= 1, (nt, io, ig)
    so, sg = o.get_state(), g.get_state()
    Lb = np.abs(so["X"]).max()
    assert np.abs(sg["X"] - so["X"]).max() <= 1e-8 * Lb
    assert np.abs(sg["q"] - so["q"]).max() <= 1e-8
    g.close()
