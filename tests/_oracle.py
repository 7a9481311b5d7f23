"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: used only by tests, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from filsim_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
P = C.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        build()
    L = C.CDLL(ORACLE_LIB)
    FP = C.POINTER(abi.FcmParams)
    protos = {
        "or_fft1d": (None, [C.c_int, _DP, C.c_int]),
        "or_rfft3d": (None, [C.c_int] * 3 + [_DP, _DP]),
        "or_irfft3d": (None, [C.c_int] * 3 + [_DP, _DP]),
        "or_fcm_mobility": (None, [FP, C.c_int, _DP, _DP, _DP, _DP, _DP]),
        "or_fcm_spread": (None, [FP, C.c_int, _DP, _DP, _DP, _DP]),
        "or_fcm_project": (None, [FP, _DP]),
        "or_fcm_interp": (None, [FP, C.c_int, _DP, _DP, _DP, _DP]),
        "or_fcm_friction": (None, [FP, C.c_int, _DP, _DP, _DP, _DP]),
        "or_fcm_last_times": (None, [_DP]),
        "or_colbar_create": (P, [C.POINTER(abi.ColbarParams), C.c_int, _IP,
                                 _IP]),
        "or_colbar_destroy": (None, [P]),
        "or_colbar_dims": (None, [P, _IP]),
        "or_colbar_link": (None, [P, _DP]),
        "or_colbar_get_lists": (None, [P, _IP, _IP]),
        "or_colbar_apply": (C.c_double, [P, _DP, C.c_int, _DP]),
        "or_colbar_set_mask": (None, [P, C.c_int]),
        "or_colbar_pairs": (C.c_int, [P, _DP, C.c_int, C.c_int, _IP]),
        "or_colbar_pairs_brute": (C.c_int, [C.POINTER(abi.ColbarParams),
                                            C.c_int, _DP, _IP, _IP, C.c_int,
                                            C.c_int, _IP]),
        "or_sim_create": (P, [C.POINTER(abi.SimParams), C.c_int, _IP,
                              C.POINTER(abi.FilProps)]),
        "or_sim_destroy": (None, [P]),
        "or_sim_sizes": (None, [P, _IP, _IP, _IP]),
        "or_sim_set_state": (None, [P, C.POINTER(abi.StatePtrs)]),
        "or_sim_get_state": (None, [P, C.POINTER(abi.StatePtrs)]),
        "or_sim_timestep": (C.c_int, [P, C.c_int]),
        "or_sim_timestep_trial": (C.c_int, [P, C.c_int]),
        "or_sim_set_barrier": (None, [P, C.c_int, C.c_int]),
        "or_sim_barrier_force": (None, [P, _DP]),
        "or_sim_bucket_times": (None, [P, _DP]),
        "or_num_threads": (C.c_int, []),
        "or_sim_initial_guess": (None, [P]),
        "or_sim_forces": (None, [P, C.c_int]),
        "or_sim_set_spring_links": (C.c_int, [P, C.c_int, _IP, C.c_double,
                                              C.c_double]),
        "or_sim_fcm_arrays": (None, [P] + [C.POINTER(_DP)] * 5),
        "or_sim_build_jacobian": (None, [P, C.c_int]),
        "or_sim_get_jacobian": (None, [P, C.c_int, _DP]),
        "or_sim_get_jinv": (None, [P, C.c_int, _DP]),
        "or_sim_jsolve": (None, [P, _DP, _DP, C.c_int]),
        "or_sim_broyden_update": (None, [P, _DP, _DP, _DP, C.c_int]),
        "or_sim_update": (None, [P, _DP]),
        "or_sim_residual": (C.c_int, [P, C.c_int, _DP, _DP, _DP, _DP]),
        "or_sim_step": (None, [P, C.c_int]),
        "or_sim_mobility": (None, [P, C.c_int]),
        "or_sim_barrier": (C.c_double, [P, C.c_int, C.c_int]),
        "or_quat_expm": (None, [_DP, _DP]),
        "or_quat_midpoint": (None, [_DP, _DP, _DP]),
        "or_quat_tangent": (None, [_DP, _DP]),
        "or_quat_rotate": (None, [_DP, _DP]),
        "or_dexpinv": (None, [_DP, _DP, _DP]),
        "or_inverse": (C.c_int, [C.c_int, _DP, _DP]),
    }
    for name, (res, args) in protos.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def dp(a):
    return a.ctypes.data_as(_DP)


def ip(a):
    return a.ctypes.data_as(_IP)


def fft1d(x, sign=-1):
    n = len(x)
    buf = np.empty(2 * n)
    buf[0::2] = x.real
    buf[1::2] = x.imag
    lib().or_fft1d(n, dp(buf), sign)
    return buf[0::2] + 1j * buf[1::2]


def rfft3d(r):
    nx, ny, nz = r.shape
    r = np.ascontiguousarray(r, dtype=np.float64)
    c = np.empty(2 * nx * ny * (nz // 2 + 1))
    lib().or_rfft3d(nx, ny, nz, dp(r), dp(c))
    c = c[0::2] + 1j * c[1::2]
    return c.reshape(nx, ny, nz // 2 + 1)


def irfft3d(c, nz):
    nx, ny, nzh = c.shape
    buf = np.empty(2 * c.size)
    buf[0::2] = c.real.ravel()
    buf[1::2] = c.imag.ravel()
    r = np.empty((nx, ny, nz))
    lib().or_irfft3d(nx, ny, nz, dp(buf), dp(r))
    return r


def num_threads():
    """Threads of the oracle's OpenMP loops."""
    return lib().or_num_threads()


def mobility(fp, Y, F, T):
    np_ = Y.shape[0]
    V = np.zeros((np_, 3))
    W = np.zeros((np_, 3))
    Y, F, T = (np.ascontiguousarray(a, dtype=np.float64) for a in (Y, F, T))
    lib().or_fcm_mobility(C.byref(fp), np_, dp(Y), dp(F), dp(T), dp(V), dp(W))
    return V, W


def last_times():
    t = np.zeros(5)
    lib().or_fcm_last_times(dp(t))
    return t


def spread(fp, Y, F, T):
    n = fp.nx * fp.ny * fp.nz
    f3 = np.zeros(3 * n)
    Y, F, T = (np.ascontiguousarray(a, dtype=np.float64) for a in (Y, F, T))
    lib().or_fcm_spread(C.byref(fp), Y.shape[0], dp(Y), dp(F), dp(T), dp(f3))
    return f3.reshape(3, fp.nx, fp.ny, fp.nz)


def project(fp, c3):
    """c3: complex array (3, nx, ny, nz//2+1); returns projected copy."""
    buf = np.empty(2 * c3.size)
    buf[0::2] = c3.real.ravel()
    buf[1::2] = c3.imag.ravel()
    lib().or_fcm_project(C.byref(fp), dp(buf))
    return (buf[0::2] + 1j * buf[1::2]).reshape(c3.shape)


def interp(fp, Y, u3):
    np_ = Y.shape[0]
    V = np.zeros((np_, 3))
    W = np.zeros((np_, 3))
    u3 = np.ascontiguousarray(u3, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    lib().or_fcm_interp(C.byref(fp), np_, dp(Y), dp(u3), dp(V), dp(W))
    return V, W


def friction(fp, F, T):
    np_ = F.shape[0]
    V = np.zeros((np_, 3))
    W = np.zeros((np_, 3))
    lib().or_fcm_friction(C.byref(fp), np_, dp(F), dp(T), dp(V), dp(W))
    return V, W


class Colbar:
    def __init__(self, cp, seg_fil, seg_idx):
        self.cp = cp
        self.np = len(seg_fil)
        self.sf = np.ascontiguousarray(seg_fil, dtype=np.int32)
        self.si = np.ascontiguousarray(seg_idx, dtype=np.int32)
        self.h = lib().or_colbar_create(C.byref(cp), self.np, ip(self.sf),
                                        ip(self.si))
        m = np.zeros(3, dtype=np.int32)
        lib().or_colbar_dims(self.h, ip(m))
        self.M = tuple(int(v) for v in m)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_colbar_destroy(self.h)
            self.h = None

    def link(self, Xs):
        Xs = np.ascontiguousarray(Xs, dtype=np.float64)
        lib().or_colbar_link(self.h, dp(Xs))

    def lists(self):
        nc = self.M[0] * self.M[1] * self.M[2]
        head = np.zeros(nc, dtype=np.int32)
        lst = np.zeros(self.np, dtype=np.int32)
        lib().or_colbar_get_lists(self.h, ip(head), ip(lst))
        return head, lst

    def apply(self, Y, nt, F):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.float64).copy()
        md = lib().or_colbar_apply(self.h, dp(Y), nt, dp(F))
        return F, md

    def set_mask(self, mask):
        lib().or_colbar_set_mask(self.h, mask)

    def pairs(self, Y, nt, max_pairs=1 << 20):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        out = np.zeros(2 * max_pairs, dtype=np.int32)
        n = lib().or_colbar_pairs(self.h, dp(Y), nt, max_pairs, ip(out))
        assert n <= max_pairs
        return out[:2 * n].reshape(n, 2)


def pairs_brute(cp, Y, seg_fil, seg_idx, nt, max_pairs=1 << 20):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    sf = np.ascontiguousarray(seg_fil, dtype=np.int32)
    si = np.ascontiguousarray(seg_idx, dtype=np.int32)
    out = np.zeros(2 * max_pairs, dtype=np.int32)
    n = lib().or_colbar_pairs_brute(C.byref(cp), Y.shape[0], dp(Y), ip(sf),
                                    ip(si), nt, max_pairs, ip(out))
    assert n <= max_pairs
    return out[:2 * n].reshape(n, 2)


STATE_FIELDS = ("X", "Xs", "Xt", "U", "Us", "Ut", "q", "qs", "qt", "lam",
                "lam1", "lam2", "F", "TAU")


def _width(name):
    return 4 if name in ("q", "qs", "qt") else 3


class Sim:
    """Oracle filament system."""

    def __init__(self, sp, nw, props):
        self.sp = sp
        self.nw = np.ascontiguousarray(nw, dtype=np.int32)
        self.props = props
        self.h = lib().or_sim_create(C.byref(sp), len(nw), ip(self.nw), props)
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        lib().or_sim_sizes(self.h, C.byref(a), C.byref(b), C.byref(c))
        self.nfil, self.np, self.nlam = a.value, b.value, c.value
        self.boff = np.concatenate([[0], np.cumsum(self.nw)]).astype(int)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_sim_destroy(self.h)
            self.h = None

    def _rows(self, name):
        return self.nlam if name.startswith("lam") else self.np

    def set_state(self, **kw):
        s = abi.StatePtrs()
        keep = []
        for k, v in kw.items():
            a = np.ascontiguousarray(v, dtype=np.float64)
            assert a.size == self._rows(k) * _width(k), k
            keep.append(a)
            setattr(s, k, dp(a))
        lib().or_sim_set_state(self.h, C.byref(s))

    def get_state(self):
        s = abi.StatePtrs()
        out = {}
        for k in STATE_FIELDS:
            a = np.zeros((self._rows(k), _width(k)))
            out[k] = a
            setattr(s, k, dp(a))
        lib().or_sim_get_state(self.h, C.byref(s))
        return out

    def fcm_arrays(self):
        ptrs = [_DP() for _ in range(5)]
        lib().or_sim_fcm_arrays(self.h, *[C.byref(p) for p in ptrs])
        return [np.ctypeslib.as_array(p, shape=(self.np, 3)) for p in ptrs]

    def timestep(self, nt):
        return lib().or_sim_timestep(self.h, nt)

    def timestep_trial(self, nt):
        return lib().or_sim_timestep_trial(self.h, nt)

    def set_barrier(self, on, mask=abi.FS_BARRIER_ALL):
        lib().or_sim_set_barrier(self.h, 1 if on else 0, mask)

    def initial_guess(self):
        lib().or_sim_initial_guess(self.h)

    def forces(self, nt):
        lib().or_sim_forces(self.h, nt)

    def set_spring_links(self, links, k, length):
        """links: (n, 4) rows (filament A, bead B, filament C, bead D)."""
        a = np.ascontiguousarray(links, dtype=np.int32).reshape(-1, 4)
        rc = lib().or_sim_set_spring_links(self.h, len(a), ip(a), k, length)
        assert rc == 0, "spring link outside the system"

    def build_jacobian(self, nt):
        lib().or_sim_build_jacobian(self.h, nt)

    def jacobian(self, f):
        n = 6 * int(self.nw[f])
        out = np.zeros((n, n))
        lib().or_sim_get_jacobian(self.h, f, dp(out))
        return out

    def jinv(self, f):
        n = 6 * int(self.nw[f])
        out = np.zeros((n, n))
        lib().or_sim_get_jinv(self.h, f, dp(out))
        return out

    def jsolve(self, b, it):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(6 * self.np)
        lib().or_sim_jsolve(self.h, dp(b), dp(x), it)
        return x

    def broyden_update(self, E, En, upd, it):
        E, En, upd = (np.ascontiguousarray(a, dtype=np.float64)
                      for a in (E, En, upd))
        lib().or_sim_broyden_update(self.h, dp(E), dp(En), dp(upd), it)

    def update(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        lib().or_sim_update(self.h, dp(u))

    def residual(self, nt, V, W):
        err = np.zeros(6 * self.np)
        mx = C.c_double()
        V = np.ascontiguousarray(V, dtype=np.float64)
        W = np.ascontiguousarray(W, dtype=np.float64)
        chk = lib().or_sim_residual(self.h, nt, dp(V), dp(W), dp(err),
                                    C.byref(mx))
        return err, chk, mx.value

    def step(self, nt):
        lib().or_sim_step(self.h, nt)

    def mobility(self, nt):
        lib().or_sim_mobility(self.h, nt)

    def barrier(self, nt, link):
        return lib().or_sim_barrier(self.h, nt, link)

    def bucket_times(self):
        """Seconds of the last timestep per profiler bucket (oracle.h)."""
        out = np.zeros(12)
        lib().or_sim_bucket_times(self.h, dp(out))
        return out

    def barrier_force(self):
        """linked_list.F_total of the last barrier application."""
        out = np.zeros((self.np, 3))
        lib().or_sim_barrier_force(self.h, dp(out))
        return out
