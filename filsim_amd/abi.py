"""ctypes binding of the C-ABI declared in include/filsim.h.

The structures here mirror the C structs field for field. ``load()`` opens
the in-tree ``libfilsim_amd.so`` and raises if it is missing: there is no
fallback path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfilsim_amd.so")

FS_FP64 = 0
FS_FP32 = 1
# status codes (include/filsim.h)
FS_OK = 0
FS_ERR_ARG = -1
FS_ERR_HIP = -2
FS_ERR_SIZE = -3
FS_ERR_STATE = -4
# barrier pair masks (fs_colbar_set_mask, fs_system_set_barrier)
FS_BARRIER_ALL = 0
FS_BARRIER_NOT_SWIMMER = 1
FS_BARRIER_NOT_TEST_PARTICLES = 2


class FcmParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("ngd", C.c_int), ("a", C.c_double), ("mu", C.c_double),
                ("dx", C.c_double), ("sigma", C.c_double),
                ("sigma_dip", C.c_double)]


class FcmTuning(C.Structure):
    _fields_ = [("fft_batch", C.c_int), ("spread_tile", C.c_int),
                ("spread_zrun", C.c_int), ("fft_xlines", C.c_int),
                ("spread_mfma", C.c_int)]


class ColbarParams(C.Structure):
    _fields_ = [("box", C.c_double * 3), ("a", C.c_double),
                ("fs_over_2a", C.c_double), ("cap", C.c_double),
                ("initial_friction_steps", C.c_int)]


class SimParams(C.Structure):
    _fields_ = [("fcm", FcmParams), ("precision", C.c_int),
                ("dL", C.c_double), ("dt", C.c_double), ("tol", C.c_double),
                ("broyden_maxiter", C.c_int), ("impl_euler_steps", C.c_int),
                ("initial_friction_steps", C.c_int),
                ("weight_per_length", C.c_double * 3),
                ("barrier_on", C.c_int), ("barrier_fs_over_2a", C.c_double),
                ("barrier_cap", C.c_double), ("sperm_decay", C.c_int),
                ("strain_twist0", C.c_double)]


class FilProps(C.Structure):
    _fields_ = [("kap", C.c_double), ("c", C.c_double), ("T0", C.c_double),
                ("omega", C.c_double), ("wavenumber", C.c_double),
                ("phase", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("radius", C.c_double)]


_DP = C.POINTER(C.c_double)


class StatePtrs(C.Structure):
    _fields_ = [(n, _DP) for n in ("X", "Xs", "Xt", "U", "Us", "Ut", "q",
                                   "qs", "qt", "lam", "lam1", "lam2", "F",
                                   "TAU")]


P = C.c_void_p
I = C.c_int
D = C.c_double
IP = C.POINTER(C.c_int)
FP = C.POINTER(C.c_float)

# name -> argtypes (all return int status)
PROTOS = {
    "fs_version": [],
    "fs_device_count": [IP],
    "fs_fcm_create": [C.POINTER(FcmParams), I, I, C.POINTER(P)],
    "fs_fcm_destroy": [P],
    "fs_fcm_mobility": [P, I, P, P, P, P, P, P],
    "fs_fcm_mobility_host": [P, I, _DP, _DP, _DP, _DP, _DP],
    "fs_fcm_friction": [P, I, P, P, P, P, P],
    "fs_fcm_set_timing": [P, I],
    "fs_fcm_stage_times": [P, FP],
    "fs_fcm_kernel_times": [P, FP],
    "fs_fcm_get_tuning": [P, C.POINTER(FcmTuning)],
    "fs_fcm_set_tuning": [P, C.POINTER(FcmTuning)],
    "fs_fcm_autotune": [P, C.c_int, C.c_int, C.POINTER(FcmTuning)],
    "fs_colbar_create": [C.POINTER(ColbarParams), I, IP, IP, C.POINTER(P)],
    "fs_colbar_destroy": [P],
    "fs_colbar_dims": [P, IP],
    "fs_colbar_link": [P, P, P],
    "fs_colbar_apply": [P, P, I, P, _DP, P],
    "fs_colbar_get_lists": [P, IP, IP],
    "fs_colbar_link_host": [P, _DP],
    "fs_colbar_apply_host": [P, _DP, I, _DP, _DP],
    "fs_colbar_pairs": [P, P, I, I, IP, IP],
    "fs_colbar_set_mask": [P, I],
    "fs_system_create": [C.POINTER(SimParams), I, IP, C.POINTER(FilProps),
                         C.POINTER(P)],
    "fs_system_destroy": [P],
    "fs_system_sizes": [P, IP, IP, IP],
    "fs_system_set_state": [P, C.POINTER(StatePtrs)],
    "fs_system_get_state": [P, C.POINTER(StatePtrs)],
    "fs_system_fcm_arrays": [P, C.POINTER(P), C.POINTER(P), C.POINTER(P),
                             C.POINTER(P), C.POINTER(P)],
    "fs_system_timestep": [P, I, IP],
    "fs_fil_initial_guess": [P],
    "fs_fil_forces": [P, I],
    "fs_fil_build_jacobian": [P, I],
    "fs_fil_get_jinv": [P, I, _DP],
    "fs_fil_get_jacobian": [P, I, _DP],
    "fs_fil_jsolve": [P, P, P, I],
    "fs_fil_broyden_update": [P, P, P, P, I],
    "fs_fil_update": [P, P],
    "fs_fil_residual": [P, I, P, P, P, IP, _DP],
    "fs_fil_step": [P, I],
    "fs_system_mobility": [P, I],
    "fs_system_barrier": [P, I, I, _DP],
    "fs_system_set_barrier": [P, I, I],
    "fs_system_timestep_trial": [P, I, IP],
    "fs_colbar_barrier_noalloc": [C.POINTER(ColbarParams), I, P, P, P, P, I,
                                  I, P, P, P],
    "fs_system_step_status": [P, IP, IP, IP, _DP],
    "fs_make_filaments": [_DP, I, IP, D, D, C.c_ulonglong, I, _DP, _DP],
    "fs_make_filaments_gpu": [_DP, I, IP, D, D, C.c_ulonglong, I, _DP, _DP],
    "fs_mlp_create": [I, I, I, C.c_ulonglong, C.POINTER(P)],
    "fs_mlp_destroy": [P],
    "fs_mlp_num_params": [P, IP],
    "fs_mlp_set_norm": [P, FP, FP, FP],
    "fs_mlp_set_skip": [P, FP],
    "fs_mlp_forward": [P, I, P, P, P],
    "fs_mlp_train_step": [P, I, P, P, C.c_float, FP, P],
    "fs_mlp_gradients": [P, I, P, P, FP, I, FP],
    "fs_mlp_get_params": [P, FP, I],
    "fs_mlp_set_params": [P, FP, I],
    "fs_system_set_warmstart": [P, P],
    "fs_system_set_warmstart_check": [P, C.c_int],
    "fs_system_set_history_carry": [P, C.c_int],
    "fs_system_warmstart_dims": [P, IP, IP],
    "fs_system_ws_record": [P, P, P, I],
    "fs_system_ws_count": [P, IP],
    "fs_system_last_step_info": [P, IP, _DP, _DP],
    "fs_system_autotune": [P, C.c_int, C.POINTER(FcmTuning)],
    "fs_system_set_tuning": [P, C.POINTER(FcmTuning)],
    "fs_system_set_profiling": [P, C.c_int],
    "fs_system_profile": [P, _DP, C.c_int],
    "fs_io_write_dat": [C.c_char_p, C.c_int, C.c_int, IP, _DP, _DP],
    "fs_io_write_bak": [C.c_char_p, C.c_int, C.c_int, IP, C.POINTER(StatePtrs)],
    "fs_io_read_bak": [C.c_char_p, C.c_int, IP, C.POINTER(StatePtrs), IP],
    "fs_io_write_par": [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, IP, _DP,
                        C.POINTER(SimParams), C.c_double, C.c_double,
                        C.c_int, C.c_int],
    "fs_io_read_par": [C.c_char_p, IP, IP, C.c_int],
    "fs_system_set_spring_links": [P, C.c_int, IP, C.c_double, C.c_double],
    "fs_io_write_spring_links": [C.c_char_p, C.c_int, IP],
    "fs_io_read_spring_links": [C.c_char_p, IP, IP, C.c_int],
    "fs_io_write_checkpoint": [C.c_char_p, C.c_int, C.c_int, IP,
                               C.POINTER(StatePtrs)],
    "fs_io_read_checkpoint": [C.c_char_p, C.c_int, IP, C.POINTER(StatePtrs),
                              IP],
    "fs_system_write_dat": [P, C.c_char_p, C.c_int],
    "fs_system_write_bak": [P, C.c_char_p, C.c_int],
    "fs_system_read_bak": [P, C.c_char_p, IP],
    "fs_system_write_checkpoint": [P, C.c_char_p, C.c_int],
    "fs_system_read_checkpoint": [P, C.c_char_p, IP],
    "fs_system_barrier_force": [P, I, P],
    "fs_system_copy_filaments": [P, P, I],
    "fs_system_com": [P, I, _DP],
    "fs_system_mean_velocity": [P, I, _DP],
    "fs_io_write_swimmer_velocity_names": [C.c_char_p, I],
    "fs_io_write_swimmer_collision_names": [C.c_char_p, I],
    "fs_io_write_force_names": [C.c_char_p, I],
    "fs_io_write_swimmer_velocity": [C.c_char_p, I, I, _DP, _DP, _DP, _DP,
                                     _DP],
    "fs_io_write_swimmer_collision": [C.c_char_p, I, I, _DP],
    "fs_io_write_forces": [C.c_char_p, I, I, IP, _DP, _DP],
    "fs_system_write_swimmer_velocity": [P, C.c_char_p, I],
    "fs_system_write_swimmer_collision": [P, C.c_char_p, I],
    "fs_system_write_forces": [P, C.c_char_p, I],
    "fs_batch_create": [C.POINTER(P), C.c_int, C.POINTER(P)],
    "fs_batch_destroy": [P],
    "fs_batch_size": [P, IP],
    "fs_batch_timestep": [P, C.c_int, IP],
    "fs_batch_set_lanes": [P, C.c_int],
    "fs_batch_set_fused": [P, C.c_int],
}

_lib = None


class FsError(RuntimeError):
    pass


def load(path=None):
    """Open libfilsim_amd.so (once) and set every prototype."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FsError("HIP library missing: %s (run __graft_entry__.build())"
                      % path)
    lib = C.CDLL(path)
    for name, args in PROTOS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.fs_last_error.argtypes = []
    lib.fs_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = _lib.fs_last_error().decode() if _lib is not None else ""
        raise FsError("filsim error %d: %s" % (rc, msg))
    return rc


def dptr(a):
    """ctypes double* of a contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_DP)


def iptr(a):
    return a.ctypes.data_as(IP)


def fptr(a):
    return a.ctypes.data_as(FP)
