"""ctypes binding of libclblm_hip.so (the C-ABI declared in include/clblm_hip.h).

The HIP library is the only implementation: if it is missing or no GPU is
visible, every call raises.  There is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclblm_hip.so")
MAXMOL = 47
FP64, FP32C = 0, 1

EXPORTS = ["clblm_version", "clblm_last_error", "clblm_ctx_create", "clblm_ctx_destroy",
           "clblm_linelist_upload", "clblm_linelist_free", "clblm_odlay", "clblm_od_tangent", "clblm_od_batch", "clblm_sync",
           "clblm_rt_up", "clblm_level2layer", "clblm_rt_jac", "clblm_level_jac", "clblm_fft_scan",
           "clblm_timing", "clblm_work", "clblm_tuning", "clblm_autotune", "clblm_scanfn",
           "clblm_linelist_broadening", "clblm_rt_down", "clblm_tape3_read", "clblm_linelist_load_tape3",
           "clblm_od_graph_create", "clblm_od_graph_launch", "clblm_od_graph_check", "clblm_od_graph_destroy",
           "clblm_rt_jac_noalloc", "clblm_calculate_dv", "clblm_set_dv", "clblm_interp_spectrum",
           "clblm_interp_to_finest", "clblm_rt_up_grids", "clblm_mtckd_load", "clblm_mtckd_absrb",
           "clblm_rt_mode1", "clblm_filter", "clblm_mtckd_read_nc", "clblm_mtckd_load_nc",
           "clblm_dev_alloc", "clblm_dev_free", "clblm_dev_put", "clblm_dev_get", "clblm_contnm_scales",
           "clblm_contnm_absrb", "clblm_rt_flux", "clblm_rt_jac_paths"]
GRID_EXACT, GRID_FIXED_RATIO, GRID_UNIFORM = -1, 0, 1
OK, ERR_ARG, ERR_UNSUPPORTED, ERR_HIP, ERR_CAPACITY = 0, 1, 2, 3, 4
OD_SPECIES_BROAD = 1
OD_MTCKD = 2
OD_CONTNM = 4
TUNE_F4_XCD_CHUNK = 1
TUNE_F4_FAR_FIELD = 2
TUNE_F4_STREAM = 3
TUNE_F4_POINTS_FP32C = 4
TUNE_F4_POINTS_FP64 = 5
TUNE_CONV_UNROLL = 6
TUNE_CONV_SPLIT = 7


class Layer(ctypes.Structure):
    """clblm_layer"""
    _fields_ = [("pave", ctypes.c_double), ("tave", ctypes.c_double), ("wtot", ctypes.c_double),
                ("zbot", ctypes.c_double), ("ztop", ctypes.c_double),
                ("w", ctypes.c_double * (MAXMOL + 1)), ("sel", ctypes.c_int32 * (MAXMOL + 1))]


class Mode1Params(ctypes.Structure):
    """clblm_mode1_params"""
    _fields_ = [("nlay", ctypes.c_int32), ("obs_level", ctypes.c_int32), ("lin_in_tau", ctypes.c_int32),
                ("thermal", ctypes.c_int32), ("solar", ctypes.c_int32), ("v1", ctypes.c_double),
                ("v2", ctypes.c_double), ("surf_emis", ctypes.c_double), ("t_skin", ctypes.c_double),
                ("surf_refl", ctypes.c_double)]


class JacPathsParams(ctypes.Structure):
    """clblm_jac_paths_params"""
    _fields_ = [("nlay", ctypes.c_int32), ("obs_level", ctypes.c_int32), ("uplook", ctypes.c_int32),
                ("lin_in_tau", ctypes.c_int32), ("solar", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("surf_emis", ctypes.c_double), ("t_skin", ctypes.c_double), ("rsfc_sun", ctypes.c_double),
                ("amf_view", ctypes.c_void_p), ("amf_down", ctypes.c_void_p), ("amf_sun", ctypes.c_void_p)]


class FluxParams(ctypes.Structure):
    """clblm_flux_params"""
    _fields_ = [("nlay", ctypes.c_int32), ("nang", ctypes.c_int32), ("lin_in_tau", ctypes.c_int32),
                ("lambertian", ctypes.c_int32), ("v1", ctypes.c_double), ("v2", ctypes.c_double),
                ("dv", ctypes.c_double), ("dv_flux", ctypes.c_double), ("surf_emis", ctypes.c_double),
                ("t_skin", ctypes.c_double)]


class ODParams(ctypes.Structure):
    """clblm_od_params"""
    _fields_ = [("v1", ctypes.c_double), ("v2", ctypes.c_double), ("dvnormal", ctypes.c_double),
                ("dvnarrow", ctypes.c_double), ("sample", ctypes.c_double), ("dvout", ctypes.c_double),
                ("precision", ctypes.c_int32), ("contnm", ctypes.c_int32),
                ("continuum", ctypes.c_void_p), ("nptabs", ctypes.c_int32), ("flags", ctypes.c_int32)]


class ODDir(ctypes.Structure):
    """clblm_od_dir"""
    _fields_ = [("dt", ctypes.c_double), ("dlnwtot", ctypes.c_double), ("dlnw", ctypes.c_double * (MAXMOL + 1))]


_lib = None


class ClblmError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ClblmError("libclblm_hip.so not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    # PyTorch-ROCm ships its own libamdhip64.so.7; loading it first makes the
    # library resolve to that one runtime, so torch tensors and this library
    # share devices, streams and memory.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    lib.clblm_last_error.restype = ctypes.c_char_p
    lib.clblm_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.clblm_ctx_destroy.argtypes = [ctypes.c_void_p]
    P = ctypes.c_void_p
    lib.clblm_linelist_upload.argtypes = [P, ctypes.c_int64] + [P] * 9 + [ctypes.POINTER(ctypes.c_int)]
    lib.clblm_linelist_free.argtypes = [P, ctypes.c_int]
    lib.clblm_linelist_broadening.argtypes = [P, ctypes.c_int, P, P, P, P]
    lib.clblm_rt_down.argtypes = [P, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double, P, P, P,
                                  ctypes.c_int, P, P, P, P]
    lib.clblm_tape3_read.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                     ctypes.POINTER(ctypes.c_int64)] + [P] * 13
    lib.clblm_linelist_load_tape3.argtypes = [P, ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)]
    lib.clblm_odlay.argtypes = [P, ctypes.c_int, ctypes.POINTER(Layer), ctypes.POINTER(ODParams),
                                ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(ctypes.c_double), P]
    lib.clblm_od_batch.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P, ctypes.POINTER(ctypes.c_int32),
                                   P, P, P]
    lib.clblm_sync.argtypes = [P]
    lib.clblm_rt_up.argtypes = [P, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                P, P, P, P, P, P]
    lib.clblm_level2layer.argtypes = [ctypes.c_int, ctypes.c_int] + [P] * 8
    lib.clblm_rt_jac.argtypes = [P, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double, P, P, P, P, P,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double, P, P, P, P, P, P]
    lib.clblm_rt_jac_noalloc.argtypes = lib.clblm_rt_jac.argtypes
    D, I64 = ctypes.c_double, ctypes.c_int64
    lib.clblm_calculate_dv.argtypes = [ctypes.c_int, P, P, P, P, D, D, D, P]
    lib.clblm_set_dv.argtypes = [ctypes.c_int, P, P, P, P, D, D, D, D, D, ctypes.c_int, D, P, P, P]
    lib.clblm_interp_spectrum.argtypes = [P, P, I64, D, D, P, I64, D, P]
    lib.clblm_interp_to_finest.argtypes = [P, ctypes.c_int, D, P, P, P, P, ctypes.POINTER(D),
                                           ctypes.POINTER(I64), P]
    lib.clblm_rt_up_grids.argtypes = [P, ctypes.c_int, D, P, P, P, P, P, P, D, D, D, I64, ctypes.c_int, D, P,
                                      ctypes.POINTER(D), ctypes.POINTER(I64), P]
    lib.clblm_level_jac.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int64, P, P, P, P, P, ctypes.c_double,
                                    ctypes.c_double, P, P, P, P, P, P, P]
    lib.clblm_fft_scan.argtypes = [P, P, ctypes.c_int64] + [ctypes.c_double] * 5 + [ctypes.c_int] + \
        [ctypes.c_double] * 3 + [ctypes.c_int, ctypes.c_int, P, ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), P]
    lib.clblm_timing.argtypes = [P, ctypes.c_int] + [ctypes.POINTER(ctypes.c_double)] * 4 + \
        [ctypes.POINTER(ctypes.c_int64)]
    lib.clblm_work.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.clblm_tuning.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.clblm_scanfn.argtypes = [P, P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64] + [ctypes.c_double] * 5 + \
        [ctypes.c_int, ctypes.c_double, P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), P]
    lib.clblm_mtckd_load.argtypes = [P, ctypes.c_int, P, P, P, P] + [ctypes.c_double] * 4
    lib.clblm_mtckd_absrb.argtypes = [P, ctypes.c_int, P, P, P, P, P, ctypes.c_int64, P, P]
    lib.clblm_contnm_scales.argtypes = [P, D, D, D, D, D]
    lib.clblm_rt_flux.argtypes = [P, P, P, ctypes.c_int64, P, P, P, P, P, P, P, P, P]
    lib.clblm_rt_jac_paths.argtypes = [P, P, ctypes.c_int64, D, D] + [P] * 12
    lib.clblm_contnm_absrb.argtypes = [P, ctypes.c_int, P, P, P, P, ctypes.c_int32, P, ctypes.c_int64, P, P]
    lib.clblm_rt_mode1.argtypes = [P, P, P, P, P, P, P, P, P, P, P, P, P, P, I64, D, P, P, ctypes.POINTER(D),
                                   ctypes.POINTER(I64), P]
    lib.clblm_mtckd_read_nc.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), P, P, P, P,
                                        ctypes.POINTER(D), ctypes.POINTER(D)]
    lib.clblm_mtckd_load_nc.argtypes = [P, ctypes.c_char_p, D, D]
    lib.clblm_filter.argtypes = [P, P, I64, D, D, ctypes.c_int, P, P, P, P, P, P, P]
    lib.clblm_autotune.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double)]
    lib.clblm_od_graph_create.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P, ctypes.POINTER(ctypes.c_int32),
                                          P, P, ctypes.POINTER(ctypes.c_void_p)]
    lib.clblm_od_graph_launch.argtypes = [P, P, P]
    lib.clblm_od_graph_check.argtypes = [P, P]
    lib.clblm_od_graph_destroy.argtypes = [P, P]
    lib.clblm_od_tangent.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int,
                                     P, P, P, P, P]
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise ClblmError("clblm error %d: %s" % (rc, load().clblm_last_error().decode()))
