"""Trajectory and restart files (fs_io_*): the reference's text layouts
(file_functions.cpp:6-81, 150-161, 211-319; Filament.hpp:547-562, 658-759)
and the bit-faithful binary checkpoint. Host arrays only: no GPU."""
import re

import numpy as np
import pytest

from filsim_amd import abi, api, params

NUM = re.compile(r"^-?\d+\.\d{16}$")


def _state(nws, seed=0):
    rng = np.random.default_rng(seed)
    npt = sum(nws)
    nlam = npt - len(nws)
    st = {k: rng.normal(0, 50, (npt, 3)) for k in ("X", "Xt", "U", "Ut")}
    for k in ("q", "qt"):
        v = rng.normal(0, 1, (npt, 4))
        st[k] = v / np.linalg.norm(v, axis=1, keepdims=True)
    for k in ("lam", "lam1", "lam2"):
        st[k] = rng.normal(0, 3, (nlam, 3))
    st["Xs"], st["Us"], st["qs"] = st["X"] + 1, st["U"] + 1, st["q"]
    return st


def test_bak_layout_and_round_trip(tmp_path):
    nws = [5, 3, 7]
    st = _state(nws)
    path = str(tmp_path / "run.bak")
    api.io_write_bak(path, 42, nws, st)
    text = open(path).read()
    assert not text.endswith("\n")
    tok = text.split(" ")
    assert tok[-1] == ""  # every value followed by one space
    tok = tok[:-1]
    npt, nlam = sum(nws), sum(nws) - len(nws)
    assert tok[0] == "42" and len(tok) == 1 + 20 * npt + 9 * nlam
    assert all(NUM.match(t) for t in tok[1:])
    # per filament: beads (X Xt q qt U Ut) then links (lam lam1 lam2)
    first = [float(t) for t in tok[1:21]]
    assert np.allclose(first, np.concatenate(
        [st["X"][0], st["Xt"][0], st["q"][0], st["qt"][0], st["U"][0],
         st["Ut"][0]]), rtol=0, atol=1e-15 * 200)
    link0 = [float(t) for t in tok[1 + 20 * 5:1 + 20 * 5 + 9]]
    assert np.allclose(link0, np.concatenate(
        [st["lam"][0], st["lam1"][0], st["lam2"][0]]), rtol=0, atol=1e-14)
    nt, back = api.io_read_bak(path, nws)
    assert nt == 42
    for k in ("X", "Xt", "q", "qt", "U", "Ut", "lam", "lam1", "lam2"):
        assert np.abs(back[k] - st[k]).max() <= 1e-15 * max(
            1.0, np.abs(st[k]).max()), k
    # continue_from_checkpoint: Xs = X, qs = q, Us = U
    assert np.array_equal(back["Xs"], back["X"])
    assert np.array_equal(back["qs"], back["q"])
    assert np.array_equal(back["Us"], back["U"])


def test_bak_truncated_rejected(tmp_path):
    nws = [4, 4]
    path = str(tmp_path / "t.bak")
    api.io_write_bak(path, 3, nws, _state(nws))
    text = open(path).read()
    open(path, "w").write(text[: len(text) // 2])
    with pytest.raises(abi.FsError):
        api.io_read_bak(path, nws)


def test_dat_appends_lines(tmp_path):
    nws = [4, 6]
    st = _state(nws, 1)
    path = str(tmp_path / "run.dat")
    for nt in (10, 20):
        api.io_write_dat(path, nt, nws, st["X"], st["q"])
    lines = open(path).read().split("\n")
    assert lines[-1] == "" and len(lines) == 3
    for nt, line in zip((10, 20), lines[:2]):
        tok = line.split(" ")[:-1]
        assert tok[0] == str(nt) and len(tok) == 1 + 7 * sum(nws)
        v = np.array([float(t) for t in tok[1:]]).reshape(-1, 7)
        assert np.abs(v[:, :3] - st["X"]).max() <= 1e-13
        assert np.abs(v[:, 3:] - st["q"]).max() <= 1e-15


def test_par_round_trip(tmp_path):
    nws = [15, 20, 30, 2]
    fp = params.fcm_params(64)
    sp = params.sim_params(fp, dt=0.005, tol=1e-6, maxiter=100)
    path = str(tmp_path / "run.par")
    api.io_write_par(path, nws, sp, kap=1000.0, time_steps=2000)
    lines = open(path).read().split("\n")
    assert lines[0].split()[:4] == ["SimulationName", "SimulationTag", "Nsw",
                                    "Nworm"]
    assert lines[1].split()[2:4] == ["4", "30"] and lines[2] == ""
    assert lines[3] == "BeadNumbers" and lines[4] == "15 20 30 2"
    assert lines[6] == "BendingModuli are KAP*"
    assert api.io_read_par(path) == nws


def test_binary_checkpoint_bitwise(tmp_path):
    nws = [7, 9]
    st = _state(nws, 2)
    path = str(tmp_path / "run.fsck")
    api.io_write_checkpoint(path, 77, nws, st)
    nt, back = api.io_read_checkpoint(path, nws)
    assert nt == 77
    for k in api._IO_FIELDS:
        assert np.array_equal(back[k].reshape(st[k].shape), st[k]), k
    with pytest.raises(abi.FsError):
        api.io_read_checkpoint(path, [8, 8])


def test_spring_link_file(tmp_path):
    """The reference's -springlinks.dat layout (spring_link_functions.cpp:
    50-101): header line, count, 'A B C D' rows."""
    p = tmp_path / "Test-springlinks.dat"
    p.write_text("Spring link data file. Number of links written first, then "
                 "springs between {filament A, bead B} and {filament C, bead "
                 "D} in the form 'A B C D'.\n3\n0 14 1 0\n1 14 2 0\n"
                 "2 14 0 0\n")
    links = api.io_read_spring_links(str(p))
    assert links.tolist() == [[0, 14, 1, 0], [1, 14, 2, 0], [2, 14, 0, 0]]
    q = tmp_path / "out-springlinks.dat"
    api.io_write_spring_links(str(q), links)
    assert q.read_text() == p.read_text()
    assert api.io_read_spring_links(str(q)).tolist() == links.tolist()
    api.io_write_spring_links(str(q), np.zeros((0, 4), np.int32))
    assert api.io_read_spring_links(str(q)).shape == (0, 4)
    bad = tmp_path / "bad-springlinks.dat"
    bad.write_text("header\n2\n0 1 2 3\n")
    with pytest.raises(abi.FsError):
        api.io_read_spring_links(str(bad))


def test_swimmer_and_force_file_layouts(tmp_path):
    """The extended output files (file_functions.cpp:102-145, 164-208;
    Filament.hpp:509-650): column names of filament 0 after "it"; at nt = 0
    the positions of filament 0 and the integer 0 for every velocity and
    force; later rows in 16-decimal fixed notation."""
    nws = [4, 3, 13]
    npt = sum(nws)
    rng = np.random.default_rng(5)
    Xs, qs = rng.normal(0, 9, (npt, 3)), rng.normal(0, 1, (npt, 4))
    V, W, F, T, Ft = (rng.normal(0, 1, (npt, 3)) for _ in range(5))
    pv, pc, pf = (str(tmp_path / n) for n in ("v.dat", "c.dat", "f.dat"))
    api.io_write_names(pv, "velocity", nws[0])
    api.io_write_names(pc, "collision", nws[0])
    api.io_write_names(pf, "forces", nws[0])
    api.io_write_swimmer_velocity(pv, 0, nws[0], Xs, qs)
    api.io_write_swimmer_collision(pc, 0, nws[0])
    api.io_write_forces(pf, 0, nws)
    api.io_write_swimmer_velocity(pv, 7, nws[0], Xs, qs, V, W, F)
    api.io_write_swimmer_collision(pc, 7, nws[0], Ft)
    api.io_write_forces(pf, 7, nws, F, T)
    lv, lc, lf = (open(p).read().split("\n") for p in (pv, pc, pf))
    for lines in (lv, lc, lf):
        assert len(lines) == 4 and lines[-1] == ""
    assert lv[0] == "it " + "X Y Z qr qi qj qk " * 4 + \
        "Vx Vy Vz Wx Wy Wz Fx Fy Fz " * 4
    assert lc[0] == "it " + "Fx Fy Fz " * 4
    assert lf[0] == "it " + "Fx Fy Fz Tx Ty Tz " * 4
    # nt = 0: positions, then integer zeros
    t = lv[1].split(" ")[:-1]
    assert t[0] == "0" and len(t) == 1 + 7 * 4 + 9 * 4
    assert all(NUM.match(x) for x in t[1:29]) and t[29:] == ["0"] * 36
    assert lc[1] == "0 " + "0 " * 12
    assert lf[1] == "0 " + "0 " * (6 * npt)
    # nt = 7: filament 0's Xs qs then V W F per segment
    t = [float(x) for x in lv[2].split(" ")[1:-1]]
    assert lv[2].startswith("7 ")
    pos = np.hstack([Xs[:4], qs[:4]]).ravel()
    vel = np.hstack([V[:4], W[:4], F[:4]]).ravel()
    assert np.abs(np.array(t) - np.concatenate([pos, vel])).max() <= 1e-14
    t = [float(x) for x in lc[2].split(" ")[1:-1]]
    assert np.abs(np.array(t) - Ft[:4].ravel()).max() <= 1e-15
    # forces: filament f reads rows (first segment of f) // 6 + i, so
    # filament 1 (first segment 4) rows 0..2 and filament 2 (first segment
    # 7) rows 1..13 -- the reference's forceLengthFilled / 6
    t = np.array([float(x) for x in lf[2].split(" ")[1:-1]]).reshape(-1, 6)
    rows = list(range(4)) + list(range(0, 3)) + list(range(1, 14))
    assert np.abs(t - np.hstack([F[rows], T[rows]])).max() <= 1e-15


def test_swimmer_files_reject_missing_arrays(tmp_path):
    p = str(tmp_path / "v.dat")
    with pytest.raises(abi.FsError):
        api.io_write_swimmer_velocity(p, 3, 2, np.zeros((2, 3)),
                                      np.zeros((2, 4)))
    with pytest.raises(abi.FsError):
        api.io_write_forces(p, 3, [2, 2])
    with pytest.raises(abi.FsError):
        api.io_write_names(p, "velocity", 0)
