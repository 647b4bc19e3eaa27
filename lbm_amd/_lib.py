"""ctypes binding of liblbm_d2q9.so (the C ABI declared in include/lbm_d2q9.h).

The library is built in-tree by ``lbm_amd/csrc/Makefile`` (``__graft_entry__.build``).
Loading fails loudly: there is no CPU fallback for any kernel of this package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_d2q9.so")

LBM_OK = 0
LBM_ERR_ARG = -1
LBM_ERR_HIP = -2
LBM_ERR_UNSUPPORTED = -3

STORE = {"f64": 0, "f32": 1, "f16": 2, "bf16": 3, "f16e": 4, "q16": 5, "f16e8": 6}
STORE_BYTES = {"f64": 8, "f32": 4, "f16": 2, "bf16": 2, "f16e": 2, "q16": 2, "f16e8": 2}
BC_NONE, BC_VELOCITY, BC_PRESSURE = 0, 1, 2

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_dp = ctypes.POINTER(ctypes.c_double)


class FusedDesc(ctypes.Structure):
    """Mirror of lbm_fused_desc."""
    _fields_ = [
        ("nx", c_i32), ("ny", c_i32), ("storage", c_i32),
        ("om_p", c_dbl), ("om_m", c_dbl),
        ("bc", c_i32 * 4), ("corners", c_i32), ("obstacles_first", c_i32),
        ("n_links", c_i32),
        ("links_host", ctypes.POINTER(c_i64)),
        ("ibb_host", c_dp),
        ("link_ibb_host", ctypes.POINTER(ctypes.c_uint8)),
        ("n_obs", c_i32),
        ("obs_offsets_host", ctypes.POINTER(c_i32)),
        ("ring", c_i32),
        ("cells_per_thread", c_i32),
        ("vector_width", c_i32),
        ("q16_range_host", c_dp),
    ]


class CaseDesc(ctypes.Structure):
    """Mirror of lbm_case_desc."""
    _fields_ = [
        ("nx", c_i32), ("ny", c_i32), ("om_p", c_dbl), ("om_m", c_dbl),
        ("bc", c_i32 * 4),
        ("profile_host", c_dp * 5),
        ("n_links", c_i32),
        ("links_host", ctypes.POINTER(c_i64)),
        ("ibb_host", c_dp),
        ("link_ibb_host", ctypes.POINTER(ctypes.c_uint8)),
        ("n_obs", c_i32),
        ("obs_offsets_host", ctypes.POINTER(c_i32)),
    ]


class BatchDesc(ctypes.Structure):
    """Mirror of lbm_batch_desc."""
    _fields_ = [
        ("n_cases", c_i32), ("cases", ctypes.POINTER(CaseDesc)), ("storage", c_i32),
        ("cells_per_thread", c_i32), ("vector_width", c_i32), ("ring", c_i32), ("grid", c_i32),
        ("q16_range_host", c_dp),
    ]


class PbbDesc(ctypes.Structure):
    """Mirror of lbm_pbb_desc."""
    _fields_ = [
        ("nx", c_i32), ("ny", c_i32), ("storage", c_i32), ("om_p", c_dbl), ("om_m", c_dbl),
        ("rho_out", c_dbl), ("inlet_profile_host", c_dp), ("ramp_host", c_dp),
        ("iterations", c_i32), ("window", c_i32), ("checkpoint_every", c_i32),
    ]


# (name, argtypes) of every exported entry point, in header order.
_SIGNATURES = {
    "lbm_abi_version": ([], c_int),
    "lbm_last_error": ([], ctypes.c_char_p),
    "lbm_macro": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_equilibrium": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_col_str": ([c_int, c_int, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_vp], c_int),
    "lbm_drag_lift": ([c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_bounce_back_obstacle": ([c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_zou_he_left_wall_velocity": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_zou_he_right_wall_velocity": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_zou_he_right_wall_pressure": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_zou_he_top_wall_velocity": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_zou_he_bottom_wall_velocity": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_zou_he_bottom_left_corner": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_zou_he_top_left_corner": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_zou_he_top_right_corner": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_zou_he_bottom_right_corner": ([c_int, c_int, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_fused_create": ([ctypes.POINTER(FusedDesc), ctypes.POINTER(c_vp)], c_int),
    "lbm_fused_destroy": ([c_vp], None),
    "lbm_fused_set_inlets": ([c_vp, c_int, c_int, c_vp, c_int, c_vp, c_vp], c_int),
    "lbm_fused_set_profiles": ([c_vp, c_vp, c_vp], c_int),
    "lbm_fused_set_amplitudes": ([c_vp, c_i64, c_int, c_vp, c_vp], c_int),
    "lbm_init": ([c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_step": ([c_vp, c_i64, c_vp, c_vp, c_vp], c_int),
    "lbm_steps": ([c_vp, c_i64, c_int, c_vp], c_int),
    "lbm_steps_prepare": ([c_vp, c_i64, c_int, c_vp], c_int),
    "lbm_epilogue": ([c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_drag_only": ([c_vp, c_i64, c_vp], c_int),
    "lbm_drag_read": ([c_vp, c_i64, c_int, c_vp, c_vp], c_int),
    "lbm_drag_read_noalloc": ([c_vp, c_i64, c_int, c_vp, c_vp], c_int),
    "lbm_fused_save": ([c_vp, c_i64, c_vp], c_int),
    "lbm_fused_restore": ([c_vp, c_i64, c_vp], c_int),
    "lbm_mass": ([c_vp, c_i64, c_vp, c_vp], c_int),
    "lbm_fused_device_bytes": ([c_vp], c_i64),
    "lbm_mv_avg_chunk": ([c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_vp], c_int),
    "lbm_batch_create": ([ctypes.POINTER(BatchDesc), ctypes.POINTER(c_vp)], c_int),
    "lbm_batch_destroy": ([c_vp], None),
    "lbm_batch_case": ([c_vp, c_int, ctypes.POINTER(c_vp)], c_int),
    "lbm_batch_info": ([c_vp, c_vp], c_int),
    "lbm_batch_set_amplitudes": ([c_vp, c_i64, c_int, c_vp, c_vp], c_int),
    "lbm_batch_init": ([c_vp, c_vp], c_int),
    "lbm_batch_steps": ([c_vp, c_i64, c_int, c_vp], c_int),
    "lbm_batch_steps_prepare": ([c_vp, c_i64, c_int, c_vp], c_int),
    "lbm_pbb_create": ([ctypes.POINTER(PbbDesc), ctypes.POINTER(c_vp)], c_int),
    "lbm_pbb_destroy": ([c_vp], None),
    "lbm_pbb_forward": ([c_vp, c_vp, c_dbl, c_vp, c_vp], c_int),
    "lbm_pbb_backward": ([c_vp, c_vp, c_vp, c_vp], c_int),
    "lbm_pbb_drag_history": ([c_vp, c_vp], c_int),
    "lbm_pbb_final_state": ([c_vp, c_vp, c_vp], c_int),
    "lbm_copy_probe": ([c_vp, c_vp, c_i64, c_int, c_vp], c_int),
    "lbm_copy_probe_nt": ([c_vp, c_vp, c_i64, c_int, c_vp], c_int),
}

_lib = None


class LbmError(RuntimeError):
    """A non-zero status returned by the D2Q9 library."""


class Unsupported(LbmError):
    """LBM_ERR_UNSUPPORTED: the fused path cannot express a boundary plan."""


def load():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C lbm_amd/csrc` "
            "(or __graft_entry__.build()); lbm_amd has no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.lbm_abi_version() != 1:
        raise ImportError("liblbm_d2q9.so ABI version mismatch")
    _lib = lib
    return lib


def exported_names():
    return list(_SIGNATURES)


def check(status, what=""):
    """Raise on a non-zero library status."""
    if status == LBM_OK:
        return
    msg = load().lbm_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if status == LBM_ERR_UNSUPPORTED:
        raise Unsupported(text)
    raise LbmError(text)


def stream_ptr():
    """The current torch stream as a hipStream_t handle."""
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
