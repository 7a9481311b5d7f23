"""Thin Python handles over the C-ABI (include/filsim.h).

Device arrays are torch.cuda float64 tensors (torch is only the allocator);
host arrays are numpy. Every call goes through libfilsim_amd.so and raises
FsError on a non-zero status.
"""
import ctypes as C

import numpy as np

from . import abi


def _ptr(t):
    """Device pointer of a contiguous torch tensor (or None)."""
    if t is None:
        return None
    assert t.is_contiguous(), "tensor must be contiguous"
    return C.c_void_p(t.data_ptr())


def _stream(s):
    if s is None:
        return None
    return C.c_void_p(s.cuda_stream if hasattr(s, "cuda_stream") else int(s))


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Fcm:
    """FCM mobility solver on one GPU (struct Fcm, FCMfunctions.hpp:7-59)."""

    def __init__(self, fp, max_np, precision=abi.FS_FP64):
        self.lib = abi.load()
        self.fp = fp
        self.max_np = max_np
        self.precision = precision
        self.h = C.c_void_p()
        abi.check(self.lib.fs_fcm_create(C.byref(fp), max_np, precision,
                                         C.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.fs_fcm_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def mobility(self, Y, F, T, V, W, stream=None):
        n = Y.shape[0]
        abi.check(self.lib.fs_fcm_mobility(self.h, n, _ptr(Y), _ptr(F),
                                           _ptr(T), _ptr(V), _ptr(W),
                                           _stream(stream)))

    def mobility_host(self, Y, F, T):
        Y, F, T = (_f64(a, (-1, 3)) for a in (Y, F, T))
        n = Y.shape[0]
        V = np.zeros((n, 3))
        W = np.zeros((n, 3))
        abi.check(self.lib.fs_fcm_mobility_host(
            self.h, n, abi.dptr(Y), abi.dptr(F), abi.dptr(T), abi.dptr(V),
            abi.dptr(W)))
        return V, W

    def friction(self, F, T, V, W, stream=None):
        abi.check(self.lib.fs_fcm_friction(self.h, F.shape[0], _ptr(F),
                                           _ptr(T), _ptr(V), _ptr(W),
                                           _stream(stream)))

    def set_timing(self, on=True):
        abi.check(self.lib.fs_fcm_set_timing(self.h, 1 if on else 0))

    def stage_times(self):
        t = np.zeros(7, dtype=np.float32)
        abi.check(self.lib.fs_fcm_stage_times(self.h, abi.fptr(t)))
        return t

    KERNELS = ("tables_bin_sort", "tile_lists", "spread", "fft_z_r2c",
               "fft_y_forward", "fft_x_project", "fft_y_inverse",
               "fft_z_c2r", "interp", "interp_sum")

    def kernel_times(self):
        """Per-kernel device times (ms) of the last timed solve, in the
        order of KERNELS."""
        t = np.zeros(10, dtype=np.float32)
        abi.check(self.lib.fs_fcm_kernel_times(self.h, abi.fptr(t)))
        return t

    def get_tuning(self):
        t = abi.FcmTuning()
        abi.check(self.lib.fs_fcm_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, fft_batch=-1, spread_tile=-1, spread_zrun=-1,
                   fft_xlines=-1, spread_mfma=-1):
        t = abi.FcmTuning(fft_batch, spread_tile, spread_zrun, fft_xlines,
                          spread_mfma)
        abi.check(self.lib.fs_fcm_set_tuning(self.h, C.byref(t)))

    def autotune(self, np_, reps=3):
        """Fastest launch shapes for np_ segments (fs_fcm_autotune); they
        stay set. Returns the tuning."""
        t = abi.FcmTuning()
        abi.check(self.lib.fs_fcm_autotune(self.h, np_, reps, C.byref(t)))
        return t


class Colbar:
    """Collision barrier (struct LinkedList,
    CollisionBarrierFilament.hpp:19-57)."""

    def __init__(self, cp, seg_fil, seg_idx):
        self.lib = abi.load()
        self.sf = np.ascontiguousarray(seg_fil, dtype=np.int32)
        self.si = np.ascontiguousarray(seg_idx, dtype=np.int32)
        self.np = len(self.sf)
        self.h = C.c_void_p()
        abi.check(self.lib.fs_colbar_create(C.byref(cp), self.np,
                                            abi.iptr(self.sf),
                                            abi.iptr(self.si),
                                            C.byref(self.h)))
        m = np.zeros(3, dtype=np.int32)
        abi.check(self.lib.fs_colbar_dims(self.h, abi.iptr(m)))
        self.M = tuple(int(v) for v in m)

    def close(self):
        if self.h:
            self.lib.fs_colbar_destroy(self.h)
            self.h = None

    __del__ = close

    def link(self, Xs, stream=None):
        abi.check(self.lib.fs_colbar_link(self.h, _ptr(Xs), _stream(stream)))

    def apply(self, Y, nt, F, stream=None):
        md = C.c_double()
        abi.check(self.lib.fs_colbar_apply(self.h, _ptr(Y), nt, _ptr(F),
                                           C.byref(md), _stream(stream)))
        return md.value

    def lists(self):
        nc = self.M[0] * self.M[1] * self.M[2]
        head = np.zeros(nc, dtype=np.int32)
        lst = np.zeros(max(self.np, 1), dtype=np.int32)
        abi.check(self.lib.fs_colbar_get_lists(self.h, abi.iptr(head),
                                               abi.iptr(lst)))
        return head, lst[:self.np]

    def set_mask(self, mask):
        """Pair mask of the barrier variants (abi.FS_BARRIER_*)."""
        abi.check(self.lib.fs_colbar_set_mask(self.h, mask))

    def pairs(self, Y, nt, max_pairs=1 << 20):
        out = np.zeros(2 * max_pairs, dtype=np.int32)
        n = C.c_int()
        abi.check(self.lib.fs_colbar_pairs(self.h, _ptr(Y), nt, max_pairs,
                                           abi.iptr(out), C.byref(n)))
        assert n.value <= max_pairs
        return out[:2 * n.value].reshape(n.value, 2)


def colbar_barrier_noalloc(cp, seg_fil, seg_idx, Xs, Y, nt, F,
                           min_dist2=None, mask=abi.FS_BARRIER_ALL,
                           stream=None):
    """Link + apply without device allocation (fs_colbar_barrier_noalloc):
    every array argument is a device tensor; F += barrier forces, and
    min_dist2 (a one-element float64 device tensor, optional) receives the
    minimum squared distance."""
    lib = abi.load()
    abi.check(lib.fs_colbar_barrier_noalloc(
        C.byref(cp), int(Xs.shape[0]), _ptr(seg_fil), _ptr(seg_idx), _ptr(Xs),
        _ptr(Y), nt, mask, _ptr(F),
        _ptr(min_dist2) if min_dist2 is not None else None, _stream(stream)))


STATE_FIELDS = ("X", "Xs", "Xt", "U", "Us", "Ut", "q", "qs", "qt", "lam",
                "lam1", "lam2", "F", "TAU")


def _width(name):
    return 4 if name in ("q", "qs", "qt") else 3


class System:
    """Filament system with its FCM solver and barrier, resident on the GPU
    (class Filament + filamentJsolver + main.cpp time loop)."""

    def __init__(self, sp, nw, props):
        self.lib = abi.load()
        self.sp = sp
        self.nw = np.ascontiguousarray(nw, dtype=np.int32)
        self.props = props
        self.h = C.c_void_p()
        abi.check(self.lib.fs_system_create(C.byref(sp), len(self.nw),
                                            abi.iptr(self.nw), props,
                                            C.byref(self.h)))
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        abi.check(self.lib.fs_system_sizes(self.h, C.byref(a), C.byref(b),
                                           C.byref(c)))
        self.nfil, self.np, self.nlam = a.value, b.value, c.value

    def close(self):
        """Frees the system; raises (and keeps it) while a live Batch holds
        it, since the batch's captured launches use its buffers."""
        if self.h:
            abi.check(self.lib.fs_system_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # still held by a batch at interpreter exit
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _rows(self, k):
        return self.nlam if k.startswith("lam") else self.np

    def set_state(self, **kw):
        s = abi.StatePtrs()
        keep = []
        for k, v in kw.items():
            a = _f64(v)
            assert a.size == self._rows(k) * _width(k), k
            keep.append(a)
            setattr(s, k, abi.dptr(a))
        abi.check(self.lib.fs_system_set_state(self.h, C.byref(s)))

    def get_state(self):
        s = abi.StatePtrs()
        out = {}
        for k in STATE_FIELDS:
            a = np.zeros((self._rows(k), _width(k)))
            out[k] = a
            setattr(s, k, abi.dptr(a))
        abi.check(self.lib.fs_system_get_state(self.h, C.byref(s)))
        return out

    def fcm_ptrs(self):
        ps = [C.c_void_p() for _ in range(5)]
        abi.check(self.lib.fs_system_fcm_arrays(self.h,
                                                *[C.byref(p) for p in ps]))
        return ps

    def timestep(self, nt):
        it = C.c_int()
        abi.check(self.lib.fs_system_timestep(self.h, nt, C.byref(it)))
        return it.value

    def timestep_trial(self, nt):
        """The step without Filament::step (a gait-reset trial pass)."""
        it = C.c_int()
        abi.check(self.lib.fs_system_timestep_trial(self.h, nt, C.byref(it)))
        return it.value

    def step_status(self):
        """Outcome of the last timestep (fs_system_step_status): dict of
        converged, exploded, nonfinite (bools) and res_inf."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        r = C.c_double()
        abi.check(self.lib.fs_system_step_status(
            self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(r)))
        return dict(converged=bool(a.value), exploded=bool(b.value),
                    nonfinite=bool(c.value), res_inf=r.value)

    def set_barrier(self, on, mask=abi.FS_BARRIER_ALL):
        abi.check(self.lib.fs_system_set_barrier(self.h, 1 if on else 0, mask))

    def initial_guess(self):
        abi.check(self.lib.fs_fil_initial_guess(self.h))

    def forces(self, nt):
        abi.check(self.lib.fs_fil_forces(self.h, nt))

    def build_jacobian(self, nt):
        abi.check(self.lib.fs_fil_build_jacobian(self.h, nt))

    def jacobian(self, f):
        n = 6 * int(self.nw[f])
        out = np.zeros((n, n))
        abi.check(self.lib.fs_fil_get_jacobian(self.h, f, abi.dptr(out)))
        return out

    def jinv(self, f):
        n = 6 * int(self.nw[f])
        out = np.zeros((n, n))
        abi.check(self.lib.fs_fil_get_jinv(self.h, f, abi.dptr(out)))
        return out

    def jsolve(self, b, x, it):
        abi.check(self.lib.fs_fil_jsolve(self.h, _ptr(b), _ptr(x), it))

    def broyden_update(self, E, En, upd, it):
        abi.check(self.lib.fs_fil_broyden_update(self.h, _ptr(E), _ptr(En),
                                                 _ptr(upd), it))

    def update(self, u):
        abi.check(self.lib.fs_fil_update(self.h, _ptr(u)))

    def residual(self, nt, V, W, err):
        chk = C.c_int()
        mx = C.c_double()
        abi.check(self.lib.fs_fil_residual(self.h, nt, _ptr(V), _ptr(W),
                                           _ptr(err), C.byref(chk),
                                           C.byref(mx)))
        return chk.value, mx.value

    def step(self, nt):
        abi.check(self.lib.fs_fil_step(self.h, nt))

    def mobility(self, nt):
        abi.check(self.lib.fs_system_mobility(self.h, nt))

    def barrier(self, nt, link):
        md = C.c_double()
        abi.check(self.lib.fs_system_barrier(self.h, nt, 1 if link else 0,
                                             C.byref(md)))
        return md.value

    def barrier_force(self, nt):
        """linked_list.F_total of the last step (fs_system_barrier_force):
        the barrier force alone, (np, 3) on the host."""
        import torch
        out = torch.zeros(3 * self.np, dtype=torch.float64, device="cuda")
        abi.check(self.lib.fs_system_barrier_force(self.h, nt, _ptr(out)))
        return out.cpu().numpy().reshape(-1, 3)

    def copy_filaments(self, src, first=0):
        """This system's filaments take the state of src's filaments first,
        first + 1, ... (fs_system_copy_filaments)."""
        abi.check(self.lib.fs_system_copy_filaments(self.h, src.h, first))

    def com(self, fil=0):
        """Filament::getCOM of filament fil."""
        c = np.zeros(3)
        abi.check(self.lib.fs_system_com(self.h, fil, abi.dptr(c)))
        return c

    def mean_velocity(self, fil=0):
        """Mean V over filament fil's segments, accumulated in float."""
        v = np.zeros(3)
        abi.check(self.lib.fs_system_mean_velocity(self.h, fil, abi.dptr(v)))
        return v

    def write_swimmer_velocity(self, path, nt):
        abi.check(self.lib.fs_system_write_swimmer_velocity(
            self.h, path.encode(), nt))

    def write_swimmer_collision(self, path, nt):
        abi.check(self.lib.fs_system_write_swimmer_collision(
            self.h, path.encode(), nt))

    def write_forces(self, path, nt):
        abi.check(self.lib.fs_system_write_forces(self.h, path.encode(), nt))

    def autotune(self, reps=3):
        """Autotune the system's FCM launch shapes (fs_system_autotune)."""
        t = abi.FcmTuning()
        abi.check(self.lib.fs_system_autotune(self.h, reps, C.byref(t)))
        return t

    def set_tuning(self, t):
        """Launch shapes of the system's FCM solver (fs_system_set_tuning)."""
        abi.check(self.lib.fs_system_set_tuning(self.h, C.byref(t)))

    PROFILE_BUCKETS = ("FCM", "Assign F", "Bld Jinv", "ErrorChk", "Slv Jinv",
                       "Total", "ColBarri", "FCMMakeZ", "FCMGauss",
                       "FCMForcD", "FCMFFTFw", "FCMFFTBk", "FCMFlowS",
                       "FCMParVR")

    def set_profiling(self, on=True):
        abi.check(self.lib.fs_system_set_profiling(self.h, 1 if on else 0))

    def set_spring_links(self, links, k, natural_length):
        """Spring links as (n, 4) rows (filament A, bead B, filament C,
        bead D) with spring constant k and natural length
        (fs_system_set_spring_links); an empty array removes them."""
        a = np.ascontiguousarray(links, dtype=np.int32).reshape(-1, 4)
        abi.check(self.lib.fs_system_set_spring_links(
            self.h, len(a), abi.iptr(a) if len(a) else None, float(k),
            float(natural_length)))

    def profile(self):
        """ms per bucket of the last profiled timestep (fs_system_profile)."""
        ms = np.zeros(len(self.PROFILE_BUCKETS))
        abi.check(self.lib.fs_system_profile(self.h, abi.dptr(ms), len(ms)))
        return dict(zip(self.PROFILE_BUCKETS, ms.tolist()))


class Batch:
    """Independent systems advanced together through captured hipGraphs
    (fs_batch_*); each reproduces its unbatched timesteps bit for bit. The
    systems must outlive the batch."""

    def __init__(self, systems):
        self.lib = abi.load()
        self.systems = list(systems)
        arr = (C.c_void_p * len(self.systems))(*[s.h for s in self.systems])
        self.h = C.c_void_p()
        abi.check(self.lib.fs_batch_create(arr, len(self.systems),
                                           C.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.fs_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def timestep(self, nt):
        it = (C.c_int * len(self.systems))()
        abi.check(self.lib.fs_batch_timestep(self.h, nt, it))
        return list(it)

    def set_lanes(self, lanes):
        """0: one parent graph per phase; k: per-system graphs on k streams."""
        abi.check(self.lib.fs_batch_set_lanes(self.h, lanes))

    def set_fused(self, on=True):
        """Fused FCM stages: one launch per stage for each group of systems
        with equal launch shapes (off by default)."""
        abi.check(self.lib.fs_batch_set_fused(self.h, 1 if on else 0))


def make_filaments(box, nw, dL=2.2, min_sep=2.2, seed=1, max_tries=100000,
                   gpu=False):
    """Seeded straight filaments (heads uniform, orientations uniform),
    overlap-rejected at min_sep. Returns X (np,3), q (np,4). gpu=True tests
    the trials on the GPU (fs_make_filaments_gpu); the result is the same
    bit for bit."""
    lib = abi.load()
    nw = np.ascontiguousarray(nw, dtype=np.int32)
    npt = int(nw.sum())
    X = np.zeros((npt, 3))
    q = np.zeros((npt, 4))
    b = np.ascontiguousarray(box, dtype=np.float64)
    fn = lib.fs_make_filaments_gpu if gpu else lib.fs_make_filaments
    abi.check(fn(abi.dptr(b), len(nw), abi.iptr(nw), dL, min_sep, seed,
                 max_tries, abi.dptr(X), abi.dptr(q)))
    return X, q


def device_count():
    lib = abi.load()
    n = C.c_int()
    abi.check(lib.fs_device_count(C.byref(n)))
    return n.value


class Mlp:
    """The learned warm start's network (fs_mlp_*)."""

    def __init__(self, n_in, n_out, width=256, seed=1):
        self.lib = abi.load()
        self.n_in, self.n_out, self.width = n_in, n_out, width
        self.h = C.c_void_p()
        abi.check(self.lib.fs_mlp_create(n_in, width, n_out, seed,
                                         C.byref(self.h)))
        n = C.c_int()
        abi.check(self.lib.fs_mlp_num_params(self.h, C.byref(n)))
        self.nparams = n.value
        self.dims = [n_in, width, width, width, n_out]

    def close(self):
        if self.h:
            self.lib.fs_mlp_destroy(self.h)
            self.h = None

    __del__ = close

    def set_norm(self, mu, sigma, tau):
        mu, sigma, tau = (np.ascontiguousarray(a, dtype=np.float32)
                          for a in (mu, sigma, tau))
        abi.check(self.lib.fs_mlp_set_norm(self.h, abi.fptr(mu),
                                           abi.fptr(sigma), abi.fptr(tau)))

    def set_skip(self, S):
        """Fixed linear term y_n += x_n S^T (S: n_out x n_in); None
        removes it (fs_mlp_set_skip)."""
        if S is None:
            abi.check(self.lib.fs_mlp_set_skip(self.h, None))
            return
        S = np.ascontiguousarray(S, dtype=np.float32)
        assert S.shape == (self.n_out, self.n_in)
        abi.check(self.lib.fs_mlp_set_skip(self.h, abi.fptr(S)))

    def forward(self, X, Y, stream=None):
        abi.check(self.lib.fs_mlp_forward(self.h, X.shape[0], _ptr(X),
                                          _ptr(Y), _stream(stream)))

    def train_step(self, X, T, lr, want_loss=True, stream=None):
        loss = C.c_float()
        abi.check(self.lib.fs_mlp_train_step(
            self.h, X.shape[0], _ptr(X), _ptr(T), lr,
            C.byref(loss) if want_loss else None, _stream(stream)))
        return loss.value

    def gradients(self, X, T):
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_float()
        abi.check(self.lib.fs_mlp_gradients(self.h, X.shape[0], _ptr(X),
                                            _ptr(T), abi.fptr(g),
                                            self.nparams, C.byref(loss)))
        return loss.value, g

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        abi.check(self.lib.fs_mlp_get_params(self.h, abi.fptr(p),
                                             self.nparams))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        abi.check(self.lib.fs_mlp_set_params(self.h, abi.fptr(p),
                                             self.nparams))


def _sys_ws_dims(system):
    a, b = C.c_int(), C.c_int()
    abi.check(system.lib.fs_system_warmstart_dims(system.h, C.byref(a),
                                                  C.byref(b)))
    return a.value, b.value


def system_warmstart_dims(system):
    return _sys_ws_dims(system)


def system_set_warmstart(system, mlp):
    abi.check(system.lib.fs_system_set_warmstart(
        system.h, mlp.h if mlp is not None else None))


def system_set_warmstart_check(system, every):
    """Solve the extrapolated guess beside the network's on every
    `every`-th step only (1: every step); in between the network's guess is
    kept while its residual beats the last measured extrapolation residual
    (fs_system_set_warmstart_check)."""
    abi.check(system.lib.fs_system_set_warmstart_check(system.h, every))


def system_set_history_carry(system, max_cols):
    """Carry the Broyden history across timesteps while it holds at most
    max_cols columns (0: restart every step, the reference's behaviour;
    fs_system_set_history_carry)."""
    abi.check(system.lib.fs_system_set_history_carry(system.h, max_cols))


def system_record(system, X, T, capacity):
    abi.check(system.lib.fs_system_ws_record(
        system.h, _ptr(X) if X is not None else None,
        _ptr(T) if T is not None else None, capacity))


def system_record_count(system):
    n = C.c_int()
    abi.check(system.lib.fs_system_ws_count(system.h, C.byref(n)))
    return n.value


def system_last_step_info(system):
    u = C.c_int()
    a, b = C.c_double(), C.c_double()
    abi.check(system.lib.fs_system_last_step_info(system.h, C.byref(u),
                                                  C.byref(a), C.byref(b)))
    return u.value, a.value, b.value


# ---- trajectory files (fs_io_*): state dicts of host arrays in the
# get_state layout
_IO_FIELDS = ("X", "Xs", "Xt", "U", "Us", "Ut", "q", "qs", "qt", "lam",
              "lam1", "lam2")


def _state_ptrs(state, nws, fill=False):
    npt = int(sum(nws))
    nlam = npt - len(nws)
    s = abi.StatePtrs()
    keep = {}
    for k in _IO_FIELDS:
        rows = nlam if k.startswith("lam") else npt
        w = _width(k)
        if fill:
            a = np.zeros((max(rows, 1), w))
        else:
            a = _f64(state[k]).reshape(rows, w)
        keep[k] = a
        setattr(s, k, abi.dptr(a))
    return s, keep


def io_write_bak(path, nt, nws, state):
    lib = abi.load()
    nw = np.ascontiguousarray(nws, dtype=np.int32)
    s, _keep = _state_ptrs(state, nws)
    abi.check(lib.fs_io_write_bak(path.encode(), nt, len(nw), abi.iptr(nw),
                                  C.byref(s)))


def io_read_bak(path, nws):
    lib = abi.load()
    nw = np.ascontiguousarray(nws, dtype=np.int32)
    s, keep = _state_ptrs(None, nws, fill=True)
    nt = C.c_int()
    abi.check(lib.fs_io_read_bak(path.encode(), len(nw), abi.iptr(nw),
                                 C.byref(s), C.byref(nt)))
    return nt.value, keep


def io_write_dat(path, nt, nws, X, q):
    lib = abi.load()
    nw = np.ascontiguousarray(nws, dtype=np.int32)
    X, q = _f64(X), _f64(q)
    abi.check(lib.fs_io_write_dat(path.encode(), nt, len(nw), abi.iptr(nw),
                                  abi.dptr(X), abi.dptr(q)))


def io_write_names(path, kind, nw0):
    """Column names of the swimmer velocity ("velocity"), swimmer
    collision ("collision") or force ("forces") file."""
    lib = abi.load()
    fn = {"velocity": lib.fs_io_write_swimmer_velocity_names,
          "collision": lib.fs_io_write_swimmer_collision_names,
          "forces": lib.fs_io_write_force_names}[kind]
    abi.check(fn(path.encode(), nw0))


def _opt(a):
    return None if a is None else abi.dptr(a)


def io_write_swimmer_velocity(path, nt, nw0, Xs, qs, V=None, W=None, F=None):
    lib = abi.load()
    arrs = [_f64(a) if a is not None else None for a in (Xs, qs, V, W, F)]
    abi.check(lib.fs_io_write_swimmer_velocity(path.encode(), nt, nw0,
                                               *[_opt(a) for a in arrs]))


def io_write_swimmer_collision(path, nt, nw0, F_total=None):
    lib = abi.load()
    a = _f64(F_total) if F_total is not None else None
    abi.check(lib.fs_io_write_swimmer_collision(path.encode(), nt, nw0,
                                                _opt(a)))


def io_write_forces(path, nt, nws, F=None, T=None):
    lib = abi.load()
    nw = np.ascontiguousarray(nws, dtype=np.int32)
    F = _f64(F) if F is not None else None
    T = _f64(T) if T is not None else None
    abi.check(lib.fs_io_write_forces(path.encode(), nt, len(nw),
                                     abi.iptr(nw), _opt(F), _opt(T)))


def io_write_par(path, nws, sp, kap=1.0, c=1.0, time_steps=0, plot_steps=1,
                 bending_factors=None, name="filsim", tag="run"):
    lib = abi.load()
    nw = np.ascontiguousarray(nws, dtype=np.int32)
    bf = _f64(bending_factors if bending_factors is not None
              else np.ones(len(nw)))
    abi.check(lib.fs_io_write_par(path.encode(), name.encode(), tag.encode(),
                                  len(nw), abi.iptr(nw), abi.dptr(bf),
                                  C.byref(sp), kap, c, time_steps,
                                  plot_steps))


def io_read_par(path):
    lib = abi.load()
    nf = C.c_int()
    abi.check(lib.fs_io_read_par(path.encode(), C.byref(nf), None, 0))
    nw = np.zeros(nf.value, dtype=np.int32)
    abi.check(lib.fs_io_read_par(path.encode(), C.byref(nf), abi.iptr(nw),
                                 len(nw)))
    return [int(v) for v in nw]


def io_write_spring_links(path, links):
    lib = abi.load()
    a = np.ascontiguousarray(links, dtype=np.int32).reshape(-1, 4)
    abi.check(lib.fs_io_write_spring_links(path.encode(), len(a),
                                           abi.iptr(a) if len(a) else None))


def io_read_spring_links(path):
    """(n, 4) int array of (filament A, bead B, filament C, bead D)."""
    lib = abi.load()
    n = C.c_int()
    abi.check(lib.fs_io_read_spring_links(path.encode(), C.byref(n), None, 0))
    a = np.zeros((n.value, 4), dtype=np.int32)
    if n.value:
        abi.check(lib.fs_io_read_spring_links(path.encode(), C.byref(n),
                                              abi.iptr(a), n.value))
    return a


def io_write_checkpoint(path, nt, nws, state):
    lib = abi.load()
    nw = np.ascontiguousarray(nws, dtype=np.int32)
    s, _keep = _state_ptrs(state, nws)
    abi.check(lib.fs_io_write_checkpoint(path.encode(), nt, len(nw),
                                         abi.iptr(nw), C.byref(s)))


def io_read_checkpoint(path, nws):
    lib = abi.load()
    nw = np.ascontiguousarray(nws, dtype=np.int32)
    s, keep = _state_ptrs(None, nws, fill=True)
    nt = C.c_int()
    abi.check(lib.fs_io_read_checkpoint(path.encode(), len(nw), abi.iptr(nw),
                                        C.byref(s), C.byref(nt)))
    return nt.value, keep


def system_write_dat(system, path, nt):
    abi.check(system.lib.fs_system_write_dat(system.h, path.encode(), nt))


def system_write_bak(system, path, nt):
    abi.check(system.lib.fs_system_write_bak(system.h, path.encode(), nt))


def system_read_bak(system, path):
    nt = C.c_int()
    abi.check(system.lib.fs_system_read_bak(system.h, path.encode(),
                                            C.byref(nt)))
    return nt.value


def system_write_checkpoint(system, path, nt):
    abi.check(system.lib.fs_system_write_checkpoint(system.h, path.encode(),
                                                    nt))


def system_read_checkpoint(system, path):
    nt = C.c_int()
    abi.check(system.lib.fs_system_read_checkpoint(system.h, path.encode(),
                                                   C.byref(nt)))
    return nt.value
