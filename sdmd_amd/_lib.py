"""Loader for the native library sdmd_amd/libsdmd.so (C host + HIP engine).

There is no Python or CPU fallback: if the library is missing the import of
any engine entry point raises.
"""
import ctypes as C
import os

from ._ctypes_types import Pair, SystemDesc

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsdmd.so")

_lib = None


class PopRec(C.Structure):
    class _R(C.Structure):
        _fields_ = [("atom", C.c_int32), ("partner", C.c_int32), ("type", C.c_int32),
                    ("sub", C.c_int32), ("counter", C.c_int64), ("pcounter", C.c_int64),
                    ("time", C.c_double), ("pot", C.c_double)]
    _fields_ = [("target", C.c_int32), ("partner", C.c_int32), ("type", C.c_int32),
                ("sub", C.c_int32), ("exec", C.c_int32), ("nrenew", C.c_int32),
                ("dt", C.c_double), ("pot", C.c_double), ("rand_calls", C.c_int64),
                ("r", _R * 2)]


class RandState(C.Structure):
    """sdmd_rand_state (sdmd_host.h): the glibc TYPE_3 rand() state."""
    _fields_ = [("r", C.c_int32 * 34), ("f", C.c_int), ("b", C.c_int), ("calls", C.c_int64)]


class RemdTemp(C.Structure):
    """sdmd_remd_temp (sdmd_remd.h): struct REMDTempStr."""
    _fields_ = [("num", C.c_int32), ("T", C.c_double)]


class GpuOpts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("block", C.c_int32), ("trace", C.c_int32),
                ("profile", C.c_int32), ("trace_capacity", C.c_int64), ("hash_every", C.c_int64),
                ("validate", C.c_int32), ("max_batch", C.c_int32), ("seed", C.c_uint64),
                ("helpers", C.c_int32), ("pad", C.c_int32)]


class FastStats(C.Structure):
    _fields_ = [("batches", C.c_int64), ("deferred", C.c_int64), ("single", C.c_int64),
                ("screens", C.c_int64), ("confirms", C.c_int64), ("validated", C.c_int64),
                ("agree", C.c_int64), ("rebases", C.c_int64), ("causal", C.c_int64),
                ("max_screen_rel_err", C.c_double), ("max_time_rel_err", C.c_double), ("clock", C.c_double),
                ("helpers", C.c_int64)]


class GpuStats(C.Structure):
    _fields_ = [("pops", C.c_int64), ("frame", C.c_int64), ("recal", C.c_int64),
                ("per_type", C.c_int64 * 16), ("pair_evals", C.c_int64), ("candidates", C.c_int64),
                ("normals_used", C.c_int64), ("currenttime", C.c_double), ("error", C.c_int32),
                ("pad", C.c_int32)]


class GpuEvent(C.Structure):
    _fields_ = [("target", C.c_int32), ("partner", C.c_int32), ("type", C.c_int32),
                ("sub", C.c_int32), ("time", C.c_double), ("potential", C.c_double),
                ("stale", C.c_int32), ("pad", C.c_int32)]


# every symbol declared in include/sdmd_host.h and include/sdmd_gpu.h
HOST_SYMBOLS = ["sdmd_host_load", "sdmd_host_prepare", "sdmd_host_desc", "sdmd_host_normals",
                "sdmd_normal_stream", "sdmd_host_energy",
                "sdmd_rand_seed", "sdmd_rand_next", "sdmd_host_free", "sdmd_strerror",
                "sdmd_remd_rearrange"]
GPU_SYMBOLS = ["sdmd_gpu_open", "sdmd_gpu_close", "sdmd_gpu_set_normals", "sdmd_gpu_prepare",
               "sdmd_gpu_run", "sdmd_gpu_predict", "sdmd_gpu_create_tree", "sdmd_gpu_update_tree",
               "sdmd_gpu_pop", "sdmd_gpu_advance", "sdmd_gpu_do_event", "sdmd_gpu_cell_move",
               "sdmd_gpu_download", "sdmd_gpu_get_tree", "sdmd_gpu_stats_get", "sdmd_gpu_get_trace",
               "sdmd_gpu_get_hashes", "sdmd_gpu_phase_times", "sdmd_gpu_last_kernel_ms",
               "sdmd_gpu_run_for", "sdmd_gpu_run_until", "sdmd_gpu_fast_stats_get", "sdmd_gpu_autotune",
               "sdmd_gpu_run_graph", "sdmd_gpu_run_events", "sdmd_gpu_energy", "sdmd_gpu_set_temperature",
               "sdmd_gpu_energy_noalloc"]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("sdmd_amd native library not built: %s missing "
                          "(run python -c 'import __graft_entry__ as g; g.build()')" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P, I, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.sdmd_host_load.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, C.POINTER(P), C.c_char_p, C.c_int]
    L.sdmd_host_prepare.argtypes = [P, C.c_long]
    L.sdmd_host_desc.restype = C.POINTER(SystemDesc)
    L.sdmd_host_desc.argtypes = [P]
    L.sdmd_host_normals.argtypes = [P, C.c_void_p, C.c_long, C.c_void_p]
    L.sdmd_normal_stream.argtypes = [C.c_uint32, C.c_int64, C.c_void_p, C.c_long, C.c_void_p]
    L.sdmd_host_energy.argtypes = [C.POINTER(SystemDesc), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.sdmd_rand_seed.argtypes = [C.c_void_p, C.c_uint32]
    L.sdmd_rand_next.restype = C.c_int32
    L.sdmd_rand_next.argtypes = [C.c_void_p]
    L.sdmd_host_free.argtypes = [P]
    L.sdmd_strerror.restype = C.c_char_p
    L.sdmd_strerror.argtypes = [C.c_int]
    L.sdmd_gpu_open.argtypes = [C.POINTER(C.POINTER(SystemDesc)), I, C.POINTER(GpuOpts), C.POINTER(P)]
    L.sdmd_gpu_close.argtypes = [P]
    L.sdmd_gpu_set_normals.argtypes = [P, I, C.c_void_p, I64]
    L.sdmd_gpu_prepare.argtypes = [P]
    L.sdmd_gpu_run.argtypes = [P, I64]
    L.sdmd_gpu_run_for.argtypes = [P, I64, C.c_double]
    L.sdmd_gpu_run_until.argtypes = [P, C.c_double, I64]
    L.sdmd_gpu_autotune.argtypes = [P, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.sdmd_gpu_run_graph.argtypes = [P, C.c_int32, C.c_double]
    L.sdmd_gpu_run_events.argtypes = [P, I64, C.c_double]
    L.sdmd_gpu_energy.argtypes = [P, C.c_void_p, C.c_void_p]
    L.sdmd_gpu_energy_noalloc.argtypes = [P, C.c_void_p, C.c_void_p]
    L.sdmd_gpu_set_temperature.argtypes = [P, I, C.c_double]
    L.sdmd_remd_rearrange.argtypes = [C.c_void_p, C.c_void_p, I, C.c_void_p, C.c_void_p]
    L.sdmd_gpu_fast_stats_get.argtypes = [P, C.c_int32, C.POINTER(FastStats)]
    L.sdmd_gpu_predict.argtypes = [P, I, C.c_void_p]
    L.sdmd_gpu_create_tree.argtypes = [P, I]
    L.sdmd_gpu_update_tree.argtypes = [P, I, C.c_void_p]
    L.sdmd_gpu_pop.argtypes = [P, I, C.POINTER(GpuEvent)]
    L.sdmd_gpu_advance.argtypes = [P, I, D]
    L.sdmd_gpu_do_event.argtypes = [P, I, I, I]
    L.sdmd_gpu_cell_move.argtypes = [P, I, I]
    L.sdmd_gpu_download.argtypes = [P, I, C.POINTER(SystemDesc)]
    L.sdmd_gpu_get_tree.argtypes = [P, I, C.c_void_p, C.c_void_p]
    L.sdmd_gpu_stats_get.argtypes = [P, I, C.POINTER(GpuStats)]
    L.sdmd_gpu_get_trace.argtypes = [P, I, C.c_void_p, I64, C.POINTER(I64)]
    L.sdmd_gpu_get_hashes.argtypes = [P, I, C.c_void_p, I64, C.POINTER(I64)]
    L.sdmd_gpu_phase_times.argtypes = [P, I, C.c_void_p, I]
    L.sdmd_gpu_last_kernel_ms.argtypes = [P, C.POINTER(D)]
    _lib = L
    return L


class SdmdError(RuntimeError):
    pass


def check(rc, what=""):
    if rc != 0:
        msg = lib().sdmd_strerror(rc).decode()
        raise SdmdError("%s failed: %s (%d)" % (what, msg, rc))
    return rc
