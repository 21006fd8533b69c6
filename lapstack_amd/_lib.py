"""ctypes binding of lapstack_amd/liblsh.so (declarations: include/lsh.h)."""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LSH_LIB", os.path.join(HERE, "liblsh.so"))
LSH_HOST, LSH_DEVICE = 0, 1
ABI_VERSION = 3

_p = C.c_void_p
_u32 = C.c_uint32


class HipError(RuntimeError):
    """An error string returned by the library."""


class FuseParamsC(C.Structure):
    _fields_ = [("levels", C.c_int32), ("energy_radius", _u32), ("top_rule", _u32), ("top_radius", _u32),
                ("entropy_bins", _u32), ("use_chroma", _u32), ("depth_level", _u32), ("halo", C.c_float)]


class DepthParamsC(C.Structure):
    _fields_ = [("scale", _u32), ("focus_kind", _u32), ("r_in", _u32), ("r_out", _u32), ("sml_step", _u32),
                ("agg_radius", _u32), ("agg_eps", C.c_float), ("lambda_", C.c_float), ("sigma_c", C.c_float),
                ("cg_iters", _u32), ("median", _u32), ("upsample_kind", _u32), ("up_radius", _u32),
                ("up_eps", C.c_float), ("gate", C.c_float), ("robust", C.c_float), ("tol", C.c_float),
                ("precision", _u32), ("warm", _p)]


class DepthInfoC(C.Structure):
    _fields_ = [("dw", _u32), ("dh", _u32), ("iters1", _u32), ("iters2", _u32), ("rel1", C.c_float),
                ("rel2", C.c_float), ("p90", C.c_float), ("floor", C.c_float), ("true_rel1", C.c_double),
                ("true_rel2", C.c_double)]


# include/lsh.h lsh_slice_fn / lsh_log_fn (the error return is a char pointer the
# callback keeps alive; c_void_p, since ctypes cannot return c_char_p from a callback safely)
SLICE_FN = C.CFUNCTYPE(_p, _p, _u32, C.POINTER(_p), C.POINTER(C.c_uint64), C.POINTER(C.c_int))
LOG_FN = C.CFUNCTYPE(None, _p, C.c_char_p)


class WlsReportC(C.Structure):
    _fields_ = [("iters", _u32), ("outer", _u32), ("rel", C.c_double), ("true_rel", C.c_double),
                ("err_fp64", C.c_double), ("ms", C.c_double)]


class BatchStackC(C.Structure):
    _fields_ = [("w", _u32), ("h", _u32), ("n", _u32), ("frames", _p)]


class WavParamsC(C.Structure):
    _fields_ = [("power", C.c_float), ("smooth", _u32), ("gate", C.c_float), ("edge", _u32)]


class BatchOptionsC(C.Structure):
    _fields_ = [("tune", _u32), ("tile_h", _u32), ("blocks_per_cu", _u32), ("frames_per_round", _u32)]


class BatchReportC(C.Structure):
    _fields_ = [("ms", C.c_double), ("mpix_frames", C.c_double), ("rounds", _u32), ("max_levels", _u32),
                ("tile_h", _u32), ("blocks_per_cu", _u32), ("resident_blocks", _u32), ("frames_per_round", _u32),
                ("from_cache", _u32),
                ("ntried", _u32), ("tried_cfg", _u32 * 8), ("tried_us_per_mpf", C.c_float * 8)]


# (name, restype, argtypes) of every entry point include/lsh.h declares
SIGNATURES = [
    ("lsh_abi", _u32, []),
    ("lsh_available", _p, []),
    ("lsh_free_str", None, [_p]),
    ("lsh_stream", _p, [C.POINTER(_p)]),
    ("lsh_sync", _p, []),
    ("lsh_mem_stats", _p, [_p, _p]),
    ("lsh_mem_release", _p, []),
    ("lsh_prof_enable", _p, [_u32]),
    ("lsh_prof_read", _p, [_p, C.c_size_t, _p, _p, _p, _u32, C.POINTER(_u32)]),
    ("lsh_prof_reset", _p, []),
    ("lsh_frames_new", _p, [_u32, _u32, _u32, C.POINTER(_p)]),
    ("lsh_frames_process", _p, [_p, _p, C.c_int, _p, _p, _u32, _p, _p, _u32, _p, _p, _p, _p, C.c_int]),
    ("lsh_frames_ingest", _p, [_p, _p, C.c_int, _u32, _p, _p, _p, C.c_int]),
    ("lsh_frames_planes", _p, [_p, C.POINTER(_p)]),
    ("lsh_frames_download", _p, [_p, _p, C.c_int]),
    ("lsh_frames_last_evals", C.c_uint64, [_p]),
    ("lsh_frames_search_stats", _p, [_p, _p, _p, _u32]),
    ("lsh_frames_search_trace", _p, [_p, _u32]),
    ("lsh_frames_search_costs", _p, [_p, _p, _u32, C.POINTER(_u32)]),
    ("lsh_frames_search_blocks", _p, [_p, _p, _u32, C.POINTER(_u32)]),
    ("lsh_frames_free", None, [_p]),
    ("lsh_fuser_new", _p, [_u32, _u32, C.POINTER(FuseParamsC), C.POINTER(_p)]),
    ("lsh_fuser_levels", _p, [_p, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
    ("lsh_fuser_push", _p, [_p, _p, C.c_int]),
    ("lsh_fuser_push_frames", _p, [_p, _p]),
    ("lsh_fuser_shares", _p, [_p, _p, _u32, C.POINTER(_u32)]),
    ("lsh_fuser_finish", _p, [_p, _p, C.c_int, _p, C.c_int, _p]),
    ("lsh_fuser_free", None, [_p]),
    ("lsh_depth_from_slices", _p, [_p, C.c_int, _u32, _p, C.c_int, _u32, _u32, C.POINTER(DepthParamsC), _p, _p,
                                   C.c_int, _p, C.POINTER(DepthInfoC)]),
    ("lsh_depth_from_source", _p, [SLICE_FN, _p, _u32, _p, C.c_int, _u32, _u32, C.POINTER(DepthParamsC), _p, _p,
                                   C.c_int, _p, C.POINTER(DepthInfoC), LOG_FN, _p]),
    ("lsh_depth_grid", _p, [_p, C.c_int, _u32, _p, C.c_int, _u32, _u32, C.POINTER(DepthParamsC), _p, _p, _p, C.c_int,
                            C.POINTER(DepthInfoC)]),
    ("lsh_warm_new", _p, [C.c_uint64, C.POINTER(_p)]),
    ("lsh_warm_free", None, [_p]),
    ("lsh_warm_nparams", _p, [_p, C.POINTER(_u32)]),
    ("lsh_warm_get", _p, [_p, _p, _u32]),
    ("lsh_warm_set", _p, [_p, _p, _u32]),
    ("lsh_warm_set_sigma", _p, [_p, C.c_float]),
    ("lsh_warm_predict", _p, [_p, _p, _p, _p, C.c_int, _u32, _u32, _u32, C.c_float, _p, C.c_int]),
    ("lsh_warm_grads", _p, [_p, _p, _p, _p, _p, C.c_int, _u32, _u32, _u32, C.c_float, C.POINTER(C.c_double), _p,
                            _u32]),
    ("lsh_warm_train_step", _p, [_p, _p, _p, _p, _p, C.c_int, _u32, _u32, _u32, C.c_float, C.c_float,
                                 C.POINTER(C.c_double)]),
    ("lsh_stack_stream", _p, [_p, _p, _p, _u32, C.c_int, _p, _u32, _u32, C.POINTER(DepthParamsC), _p, C.c_int, _u32,
                              _p, _p, _p]),
    ("lsh_stack_stream_files", _p, [_p, _p, _p, _u32, _u32, _p, _u32, _u32, C.POINTER(DepthParamsC), _p, C.c_int,
                                    _u32, _p, _p, _p]),
    ("lsh_image_info", _p, [C.c_char_p, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
    ("lsh_image_load", _p, [C.c_char_p, _p, C.c_size_t, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
    ("lsh_image_save", _p, [C.c_char_p, _p, C.c_int, _u32, _u32, _u32]),
    ("lsh_image_save_u16", _p, [C.c_char_p, _p, C.c_int, _u32, _u32]),
    ("lsh_save_gray16", _p, [C.c_char_p, _p, C.c_int, _u32, _u32, C.c_float]),
    ("lsh_save_gray", _p, [C.c_char_p, _p, C.c_int, _u32, _u32]),
    ("lsh_batch_fuse", _p, [_p, _p, _u32, C.POINTER(BatchOptionsC), C.POINTER(BatchReportC)]),
    ("lsh_wav_new", _p, [_u32, _u32, _p, C.POINTER(WavParamsC), _p, _u32, _p, _u32, C.POINTER(_p)]),
    ("lsh_wav_push", _p, [_p, _p, _u32, _p, _u32]),
    ("lsh_wav_push_frames", _p, [_p, _p]),
    ("lsh_wav_finish", _p, [_p, _p, _u32]),
    ("lsh_wav_free", None, [_p]),
    ("lsh_slab_ranges", _u32, [_u32, _u32, _u32, _p, _u32]),
    ("lsh_gauss_levels", _p, [_p, _u32, _u32, _u32, _p, _u32, C.POINTER(_u32)]),
    ("lsh_common_area", _p, [_p, _u32, _u32, _u32, _u32, _p]),
    ("lsh_crop", _p, [_p, _u32, _u32, _u32, _u32, _p, _p, _u32]),
    ("lsh_resize", _p, [_p, _u32, _u32, _u32, _u32, _u32, _u32, _u32, _p, _u32]),
    ("lsh_to_u16", _p, [_p, _u32, _u32, _u32, _p, _u32]),
    ("lsh_wls_solve", _p, [_p, _p, _p, _p, C.c_int, _u32, _u32, C.c_float, C.c_float, _u32, C.c_double, _u32, _u32,
                           _p, _p, C.c_int, C.POINTER(WlsReportC)]),
    ("lsh_wls_solve_noalloc", _p, [_p, _p, _p, _p, _u32, _u32, C.c_float, C.c_float, C.c_double, _u32, _p,
                                   C.POINTER(WlsReportC)]),
]

_lock = threading.Lock()
_lib = None


def _start_torch_first():
    """torch ships a HIP runtime of its own beside the system one this library
    links.  Both can run in one process (device pointers pass between them),
    but torch finds no device when the library's runtime started first, so
    torch's starts here, before the library loads, when torch is present."""
    try:
        import torch
    except ImportError:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except RuntimeError:
        pass


def load(path: str = LIB_PATH):
    """Load the library (raises OSError when it is not built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        _start_torch_first()
        if not os.path.exists(path):
            raise OSError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        lib = C.CDLL(path)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.lsh_abi() != ABI_VERSION:
            raise OSError(f"{path}: ABI {lib.lsh_abi()} but this binding speaks {ABI_VERSION}")
        _lib = lib
        return lib


def take_err(ptr):
    """The library's error string (released), or None."""
    if not ptr:
        return None
    s = C.cast(ptr, C.c_char_p).value.decode("utf-8", "replace")
    _lib.lsh_free_str(ptr)
    return s


def check(ptr):
    msg = take_err(ptr)
    if msg is not None:
        raise HipError(msg)


def stream_ptr() -> int:
    s = C.c_void_p()
    check(load().lsh_stream(C.byref(s)))
    return s.value or 0


def sync():
    check(load().lsh_sync())


def mem_stats():
    """(live, cached) bytes of the library's device memory on the current device."""
    live, cached = C.c_uint64(), C.c_uint64()
    check(load().lsh_mem_stats(C.byref(live), C.byref(cached)))
    return live.value, cached.value


def mem_release():
    """Return the library's kept device blocks to the device."""
    check(load().lsh_mem_release())


def prof_enable(on: bool):
    check(load().lsh_prof_enable(int(on)))


def prof_reset():
    check(load().lsh_prof_reset())


def prof_read():
    """{kernel: (total ms, total algorithmic bytes, launches)}, timed with HIP events on the library's stream."""
    lib = load()
    cap = 256
    names = C.create_string_buffer(1 << 16)
    ms = (C.c_double * cap)()
    by = (C.c_double * cap)()
    cnt = (C.c_uint64 * cap)()
    n = C.c_uint32()
    check(lib.lsh_prof_read(names, len(names), ms, by, cnt, cap, C.byref(n)))
    keys = names.value.decode().split("\n")[: n.value]
    return {k: (ms[i], by[i], int(cnt[i])) for i, k in enumerate(keys)}
