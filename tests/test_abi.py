"""C-ABI checks that need no GPU: the library loads, exports every function
include/*.h declares, rejects bad arguments before touching the device, and
the host-side generator is deterministic and overlap-free."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from filsim_amd import abi, api, params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\**\s*(fs_\w+)\s*\(",
                             src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_exports_every_declared_symbol():
    lib = C.CDLL(abi.LIB_PATH)
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers them all
    bound = set(abi.PROTOS) | {"fs_last_error"}
    assert not (set(names) - bound), set(names) - bound


def test_missing_library_fails_loudly(tmp_path):
    saved = abi._lib
    abi._lib = None
    try:
        with pytest.raises(abi.FsError):
            abi.load(str(tmp_path / "nope.so"))
    finally:
        abi._lib = saved


def test_bad_arguments_rejected():
    lib = abi.load()
    h = C.c_void_p()
    fp = params.fcm_params((64, 64, 66 + 1))  # odd size
    assert lib.fs_fcm_create(C.byref(fp), 10, 0, C.byref(h)) == -1
    assert b"grid" in lib.fs_last_error()
    fp = params.fcm_params((64, 64, 14 * 7))  # factor 7
    assert lib.fs_fcm_create(C.byref(fp), 10, 0, C.byref(h)) == -1
    fp = params.fcm_params(64, ngd=30)
    assert lib.fs_fcm_create(C.byref(fp), 10, 0, C.byref(h)) == -1
    fp = params.fcm_params(64)
    assert lib.fs_fcm_create(C.byref(fp), 10, 7, C.byref(h)) == -1
    cp = params.colbar_params(params.fcm_params(8))  # 4.3a box: < 3 cells
    sf = np.zeros(2, dtype=np.int32)
    assert lib.fs_colbar_create(C.byref(cp), 2, abi.iptr(sf), abi.iptr(sf),
                                C.byref(h)) == -1
    sp = params.sim_params(params.fcm_params(64))
    nw = np.array([1], dtype=np.int32)
    props = params.fil_props(1)
    assert lib.fs_system_create(C.byref(sp), 1, abi.iptr(nw), props,
                                C.byref(h)) == -1


def test_device_count_never_faults():
    assert api.device_count() >= 0


def test_make_filaments_deterministic_no_overlap():
    fp = params.fcm_params(96)
    box = params.box(fp)
    nw = [15] * 40
    X1, q1 = api.make_filaments(box, nw, min_sep=2.2, seed=11)
    X2, q2 = api.make_filaments(box, nw, min_sep=2.2, seed=11)
    assert np.array_equal(X1, X2) and np.array_equal(q1, q2)
    X3, _ = api.make_filaments(box, nw, min_sep=2.2, seed=12)
    assert not np.array_equal(X1, X3)
    L = np.array(box)
    fil = np.repeat(np.arange(40), 15)
    d = X1[:, None, :] - X1[None, :, :]
    d -= L * np.floor(d / L + 0.5)
    r = np.sqrt((d ** 2).sum(-1))
    other = fil[:, None] != fil[None, :]
    assert r[other].min() >= 2.2
    # straight filaments: consecutive beads 2.2 apart along R(q) e_x
    seg = np.linalg.norm(np.diff(X1.reshape(40, 15, 3), axis=1), axis=-1)
    assert np.allclose(seg, 2.2)
    assert np.allclose(np.linalg.norm(q1, axis=1), 1)


def test_shim_compiles_reference_call_sites():
    """include/filsim_shim.hpp with the reference driver's own calls on a
    filament container (a stand-in for class Filament: beads[i].Xs(d), .F(d),
    .TAU(d), length()) compiles with the host compiler alone."""
    import subprocess
    src = os.path.join(ROOT, "filsim_amd", "host", "driver.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror",
                        "-fsyntax-only", src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
