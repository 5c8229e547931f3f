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
                ("helpers", C.c_int32), ("window", C.c_int32)]


class DatInfo(C.Structure):
    """sdmd_dat_info (include/sdmd_io.h)"""
    _fields_ = [("frame", C.c_int64), ("event_sums", C.c_int64 * 7), ("timestep", C.c_double),
                ("outputrate", C.c_double), ("hb_formed", C.c_int32), ("alpha_hb_formed", C.c_int32),
                ("remd_flag", C.c_int32), ("remd_port", C.c_int32), ("remd_output_rate", C.c_int32),
                ("remd_server", C.c_char * 50), ("remd_extra", C.c_char * 50), ("remd_num", C.c_int32),
                ("remd_temperature", C.c_double)]


class Resume(C.Structure):
    """sdmd_resume (include/sdmd_io.h)"""
    _fields_ = [("natom", C.c_int32), ("seed", C.c_uint32), ("rand_calls", C.c_int64), ("pops", C.c_int64),
                ("recal", C.c_int64), ("node", C.c_void_p), ("leaf", C.c_void_p)]


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
                ("last_exec", C.c_int32)]


class GpuEvent(C.Structure):
    _fields_ = [("target", C.c_int32), ("partner", C.c_int32), ("type", C.c_int32),
                ("sub", C.c_int32), ("time", C.c_double), ("potential", C.c_double),
                ("stale", C.c_int32), ("pad", C.c_int32)]


# every symbol declared in include/sdmd_host.h and include/sdmd_gpu.h
HOST_SYMBOLS = ["sdmd_host_load", "sdmd_host_prepare", "sdmd_host_desc", "sdmd_host_normals",
                "sdmd_normal_stream", "sdmd_host_energy",
                "sdmd_rand_seed", "sdmd_rand_next", "sdmd_host_free", "sdmd_strerror",
                "sdmd_remd_rearrange", "sdmd_dat_info_init", "sdmd_write_gro", "sdmd_write_cnt",
                "sdmd_save_dat", "sdmd_host_load_dat", "sdmd_save_resume", "sdmd_load_resume",
                "sdmd_remd_run"]
GPU_SYMBOLS = ["sdmd_gpu_open", "sdmd_gpu_close", "sdmd_gpu_set_normals", "sdmd_gpu_prepare",
               "sdmd_gpu_run", "sdmd_gpu_predict", "sdmd_gpu_create_tree", "sdmd_gpu_update_tree",
               "sdmd_gpu_pop", "sdmd_gpu_advance", "sdmd_gpu_do_event", "sdmd_gpu_cell_move",
               "sdmd_gpu_download", "sdmd_gpu_get_tree", "sdmd_gpu_stats_get", "sdmd_gpu_get_trace",
               "sdmd_gpu_get_hashes", "sdmd_gpu_phase_times", "sdmd_gpu_last_kernel_ms",
               "sdmd_gpu_run_for", "sdmd_gpu_run_until", "sdmd_gpu_fast_stats_get", "sdmd_gpu_autotune",
               "sdmd_gpu_run_graph", "sdmd_gpu_run_events", "sdmd_gpu_energy", "sdmd_gpu_set_temperature",
               "sdmd_gpu_energy_noalloc", "sdmd_gpu_run_to_output", "sdmd_gpu_set_tree"]


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
    L.sdmd_dat_info_init.restype = None
    L.sdmd_dat_info_init.argtypes = [C.POINTER(DatInfo), P, C.POINTER(SystemDesc)]
    L.sdmd_write_gro.argtypes = [P, C.POINTER(SystemDesc), C.c_char_p, C.c_int32]
    L.sdmd_write_cnt.argtypes = [C.POINTER(SystemDesc), C.c_char_p, C.c_int32]
    L.sdmd_save_dat.argtypes = [P, C.POINTER(SystemDesc), C.POINTER(DatInfo), C.c_char_p]
    L.sdmd_host_load_dat.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(P), C.POINTER(DatInfo), C.c_char_p, C.c_int32]
    L.sdmd_save_resume.argtypes = [C.POINTER(Resume), C.c_char_p]
    L.sdmd_load_resume.argtypes = [C.POINTER(Resume), C.c_char_p]
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
    L.sdmd_gpu_run_to_output.argtypes = [P, C.c_double, C.c_double, I64]
    L.sdmd_gpu_set_tree.argtypes = [P, C.c_int32, C.c_void_p, C.c_void_p]
    L.sdmd_gpu_set_temperature.argtypes = [P, I, C.c_double]
    L.sdmd_remd_rearrange.argtypes = [C.c_void_p, C.c_void_p, I, C.c_void_p, C.c_void_p]
    L.sdmd_remd_run.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, I64, C.c_void_p, C.c_void_p]
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
