"""Loader of the in-tree HIP library parrelray_amd/lib/libprr.so (C ABI of include/prr.h).

There is no fallback: if the library is missing or fails to load, every call
raises.  Build it with `python -c "import __graft_entry__ as g; g.build()"` or
`make -C parrelray_amd/csrc`.
"""
import ctypes as C
import os

from . import _abi as A

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
# PRR_LIB selects another build of the same library (e.g. a different register budget)
LIB_PATH = os.environ.get("PRR_LIB") or os.path.join(LIB_DIR, "libprr.so")

_lib = None


class PrrError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{A.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_d = C.c_double
_dp = C.POINTER(C.c_double)


def _sig(L, name, res, args):
    f = getattr(L, name)
    f.restype = res
    f.argtypes = args


def lib():
    """The loaded libprr.so (raises when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    # One HIP runtime per process: torch ships its own libamdhip64 (same soname).
    # Loading torch first makes libprr.so bind to that runtime instead of
    # bringing in a second copy, which would hide the GPU from torch.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    st = C.c_int
    _sig(L, "prr_abi_version", C.c_int, [])
    _sig(L, "prr_global_error", C.c_char_p, [])
    _sig(L, "prr_last_error", C.c_char_p, [_vp])
    _sig(L, "prr_create", st, [C.POINTER(_vp), C.c_int])
    _sig(L, "prr_destroy", None, [_vp])
    _sig(L, "prr_set_spacetime", st, [_vp, _i32, _d, _d, _d, _dp])
    _sig(L, "prr_set_tracer", st, [_vp, C.POINTER(A.Tracer)])
    _sig(L, "prr_set_texture", st, [_vp, C.c_int, _vp, C.c_int, C.c_int, _i32])
    _sig(L, "prr_set_launch", st, [_vp, C.POINTER(A.Launch)])
    _sig(L, "prr_render_noalloc", st, [_vp, C.POINTER(A.Frame), C.c_int, C.POINTER(A.Image), _vp])
    _sig(L, "prr_get_launch", st, [_vp, C.POINTER(A.Launch)])
    _sig(L, "prr_render", st, [_vp, C.POINTER(A.Frame), C.c_int, _vp, C.POINTER(A.RayOut), _vp])
    _sig(L, "prr_get_stats", st, [_vp, C.POINTER(A.Stats)])
    _sig(L, "prr_trace", st, [_vp, _vp, C.c_int, _vp, _vp, _vp, _vp])
    _sig(L, "prr_set_spheres", st, [_vp, C.POINTER(A.Sphere), C.c_int])
    _sig(L, "prr_graph_create", st, [_vp, C.POINTER(A.Frame), C.c_int, _vp, C.POINTER(_vp)])
    _sig(L, "prr_graph_launch", st, [_vp, _vp])
    _sig(L, "prr_graph_destroy", None, [_vp])
    _sig(L, "prr_params_read", st, [C.POINTER(_vp), C.c_char_p])
    _sig(L, "prr_params_read_into", st, [_vp, C.c_char_p])
    _sig(L, "prr_params_load_sim", st, [C.POINTER(_vp), C.c_char_p])
    _sig(L, "prr_params_free", None, [_vp])
    _sig(L, "prr_params_header_exists", C.c_int, [_vp, C.c_char_p])
    _sig(L, "prr_params_par_exists", C.c_int, [_vp, C.c_char_p, C.c_char_p])
    _sig(L, "prr_params_par", C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int])
    _sig(L, "prr_params_pari", C.c_int, [_vp, C.c_char_p, C.c_char_p])
    _sig(L, "prr_params_parf", _d, [_vp, C.c_char_p, C.c_char_p, _d])
    _sig(L, "prr_params_frame_count", C.c_int, [_vp])
    _sig(L, "prr_params_spacetime", st, [_vp, C.POINTER(_i32), _dp, _dp, _dp, _dp])
    _sig(L, "prr_params_tracer", st, [_vp, C.POINTER(A.Tracer)])
    _sig(L, "prr_params_frame", st, [_vp, C.c_int, C.POINTER(A.Frame)])
    _sig(L, "prr_texture_checker", st, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64])
    _sig(L, "prr_texture_value_noise", st, [_vp, C.c_int, C.c_int, C.c_int, C.c_uint64])
    _sig(L, "prr_image_read", st, [C.c_char_p, _vp, C.POINTER(C.c_int), C.POINTER(C.c_int)])
    _sig(L, "prr_debug_state_size", C.c_int, [_i32])
    _sig(L, "prr_debug_rhs", st, [_i32, _d, _d, _d, _d, _d, _vp, _i32, _vp])
    _sig(L, "prr_debug_kerr_delta_rhs", st, [_d, _d, _d, _d, _d, _vp, _vp, _vp])
    _sig(L, "prr_debug_kerr_stages", st, [_d, _d, _d, _d, _d, _vp, _d, _i32, _vp, _vp])
    _sig(L, "prr_debug_rhs_vjp", st, [_d, _d, _d, _d, _d, _vp, _vp, _vp, _vp])
    _sig(L, "prr_debug_orbit_init", st, [_d, _d, _d, _d, _d, _d, _d, _d, _d, _vp, _vp])
    _sig(L, "prr_debug_kerr_step", st, [_d, _d, _d, _d, _d, _vp, _d, _vp])
    _sig(L, "prr_debug_kerr_step_vjp", st, [_d, _d, _d, _d, _d, _vp, _d, _vp])
    _sig(L, "prr_debug_escape_color", st, [_d, _d, _d, _d, _d, _vp, _vp, C.c_int, C.c_int, _vp, _vp])
    _sig(L, "prr_debug_state_from_xu", st, [_i32, _d, _d, _vp, _vp, _vp, _vp])
    _sig(L, "prr_debug_xu_from_state", st, [_i32, _d, _d, _d, _d, _d, _vp, _vp, _vp])
    _sig(L, "prr_debug_init_ray", st, [_i32, _d, _d, _vp, _vp, _vp, _vp])
    _sig(L, "prr_debug_camera_ray", st, [C.POINTER(A.Frame), _i64, _vp])
    _sig(L, "prr_debug_queue_rays", st, [C.POINTER(A.Frame), _vp, _i64, _vp])
    _sig(L, "prr_debug_trace_rays", st, [_i32, C.POINTER(A.Tracer), C.POINTER(A.Frame), _vp, _vp, C.c_int, _vp, _vp,
                                         _vp, _vp, _vp])
    _sig(L, "prr_debug_trace_log", st, [_i32, C.POINTER(A.Tracer), C.POINTER(A.Frame), _i64, C.c_int, _vp,
                                        C.POINTER(C.c_int)])
    _sig(L, "prr_frame_from_orbit", st, [C.POINTER(A.Frame), C.POINTER(A.Orbit), C.POINTER(A.Frame)])
    _sig(L, "prr_adjoint", st, [_vp, C.POINTER(A.Frame), C.POINTER(A.Orbit), _vp, _i32, _dp, _vp])
    if L.prr_abi_version() != A.PRR_ABI_VERSION:
        raise RuntimeError("libprr.so ABI version mismatch")
    _lib = L
    return L


def check(status, ctx=None):
    if status != A.PRR_OK:
        L = lib()
        msg = L.prr_last_error(ctx) if ctx else L.prr_global_error()
        raise PrrError(status, (msg or b"").decode(errors="replace"))
