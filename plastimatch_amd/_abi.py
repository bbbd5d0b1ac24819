"""ctypes binding of libplmregisterhip.so (the C ABI in include/plm_hip_bspline.h).

This is the Python-side counterpart of the ``LOAD_LIBRARY_SAFE`` /
``LOAD_SYMBOL`` plugin loader of the reference (src/sys/cuda_delayload.h:26-59):
it loads the HIP flavor library from the package directory and declares
every exported symbol.  There is no fallback: if the library is missing,
``load()`` raises.
"""
import ctypes
import os

LIB_NAME = "libplmregisterhip.so"
VOPT_KEYS = 1000      # PLM_HIP_VOPT_KEYS
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

PLM_HIP_OK = 0
PLM_HIP_WARN_NO_OVERLAP = 1
PLM_HIP_ERR_ARG = -1
PLM_HIP_ERR_HIP = -2
PLM_HIP_ERR_GEOMETRY = -3
PLM_HIP_ERR_UNSUPPORTED = -4
PLM_HIP_ERR_NO_DEVICE = -5

PLM_METRIC_MSE = 0
PLM_METRIC_MI = 1
PLM_GRAD_REF_NN = 0
PLM_GRAD_EXACT = 1
PLM_GRAD_NONE = 2
PLM_MEM_HOST = 0
PLM_MEM_DEVICE = 1
PLM_STORE_F32 = 0
PLM_STORE_BF16 = 1
PLM_STORE_U8 = 2
PLM_RESAMPLE = {"voxel_rate": 0, "voxel_legacy": 1, "mm": 2, "pct": 3}

c_float3 = ctypes.c_float * 3
c_int64_3 = ctypes.c_int64 * 3
c_float9 = ctypes.c_float * 9


class PlmBspGeom(ctypes.Structure):
    _fields_ = [
        ("img_origin", c_float3),
        ("img_spacing", c_float3),
        ("img_dim", c_int64_3),
        ("roi_offset", c_int64_3),
        ("roi_dim", c_int64_3),
        ("vox_per_rgn", c_int64_3),
        ("direction_cosines", c_float9),
    ]


class PlmImgDesc(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("location", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("dim", c_int64_3),
        ("origin", c_float3),
        ("spacing", c_float3),
        ("direction_cosines", c_float9),
    ]


class PlmRegParms(ctypes.Structure):
    """Regularization weights (Regularization_parms)."""
    _fields_ = [(n, ctypes.c_float) for n in (
        "total_displacement_penalty", "diffusion_penalty", "curvature_penalty", "lame_coefficient_1",
        "lame_coefficient_2", "linear_elastic_multiplier", "third_order_penalty",
        "curvature_mixed_weight")]


class PlmLbfgsParms(ctypes.Structure):
    """L-BFGS parameters (plm_lbfgs_parms: liblbfgs's lbfgs_parameter_t
    fields of the More-Thuente path plus the reference's max_its and
    max_feval)."""
    _fields_ = [("m", ctypes.c_int), ("epsilon", ctypes.c_double), ("past", ctypes.c_int),
                ("delta", ctypes.c_double), ("max_iterations", ctypes.c_int),
                ("max_linesearch", ctypes.c_int), ("min_step", ctypes.c_double),
                ("max_step", ctypes.c_double), ("ftol", ctypes.c_double), ("gtol", ctypes.c_double),
                ("xtol", ctypes.c_double), ("max_its", ctypes.c_int), ("max_feval", ctypes.c_int)]


class PlmLbfgsResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int), ("fevals", ctypes.c_int),
                ("fx", ctypes.c_double)]


# progress callback: (instance, d_x, d_g, fx, xnorm, gnorm, step, n, k, ls) -> int
class PlmLbfgsReport(ctypes.Structure):
    _fields_ = [("fx", ctypes.c_double), ("xnorm", ctypes.c_double), ("gnorm", ctypes.c_double),
                ("step", ctypes.c_double), ("k", ctypes.c_int), ("ls", ctypes.c_int)]


PLM_LBFGS_PROGRESS = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int)

PLM_LBFGS_SUCCESS = 0
PLM_LBFGS_STOP = 1
PLM_LBFGS_ALREADY_MINIMIZED = 2
PLM_LBFGSERR_MAXIMUMLINESEARCH = -998


class PlmHistDesc(ctypes.Structure):
    _fields_ = [
        ("fixed_bins", ctypes.c_int64),
        ("moving_bins", ctypes.c_int64),
        ("fixed_offset", ctypes.c_float),
        ("fixed_delta", ctypes.c_float),
        ("moving_offset", ctypes.c_float),
        ("moving_delta", ctypes.c_float),
    ]


# every symbol the header declares: name -> (restype, argtypes)
_P = ctypes.c_void_p
SIGNATURES = {
    "plm_hip_abi_version": (ctypes.c_int, []),
    "plm_hip_last_error": (ctypes.c_char_p, []),
    "plm_hip_bspline_create": (ctypes.c_int, [
        ctypes.POINTER(PlmBspGeom), ctypes.POINTER(PlmImgDesc), ctypes.POINTER(PlmImgDesc),
        ctypes.POINTER(PlmImgDesc), ctypes.POINTER(PlmHistDesc), ctypes.c_int,
        ctypes.POINTER(_P)]),
    "plm_hip_bspline_score": (ctypes.c_int, [
        _P, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P]),
    "plm_hip_bspline_score_device": (ctypes.c_int, [
        _P, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, _P]),
    "plm_hip_bspline_score_noalloc": (ctypes.c_int, [
        ctypes.POINTER(PlmBspGeom), ctypes.POINTER(PlmImgDesc), ctypes.POINTER(PlmImgDesc),
        ctypes.POINTER(PlmImgDesc), ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P]),
    "plm_hip_bspline_destroy": (ctypes.c_int, [_P]),
    "plm_hip_bspline_set_roi": (ctypes.c_int, [_P, ctypes.POINTER(PlmImgDesc), ctypes.POINTER(PlmImgDesc)]),
    "plm_hip_volume_gradient": (ctypes.c_int, [ctypes.POINTER(PlmImgDesc), _P, ctypes.c_int,
                                               ctypes.c_int]),
    "plm_hip_bspline_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "plm_hip_bspline_set_tuning": (ctypes.c_int, [_P, ctypes.c_int]),
    "plm_hip_bspline_set_occupancy": (ctypes.c_int, [_P, ctypes.c_int]),
    "plm_hip_bspline_get_tuning": (ctypes.c_int, [_P, _P, _P]),
    "plm_hip_bspline_autotune": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int,
                                                ctypes.POINTER(ctypes.c_int),
                                                ctypes.POINTER(ctypes.c_double)]),
    "plm_hip_bspline_get_timing": (ctypes.c_int, [_P, _P, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_int64)]),
    "plm_hip_bspline_interpolate_vf": (ctypes.c_int, [_P, _P, _P, _P]),
    "plm_hip_bspline_score_vf_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "plm_hip_bspline_warp": (ctypes.c_int, [_P, _P, _P, ctypes.c_float, ctypes.c_int, _P]),
    "plm_hip_bspline_regularize": (ctypes.c_int, [_P, ctypes.POINTER(PlmRegParms), _P, _P, _P, ctypes.c_int,
                                                  _P]),
    "plm_hip_bspline_num_knots": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "plm_hip_bspline_num_tiles": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "plm_hip_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.POINTER(_P)]),
    "plm_hip_batch_set_buffers": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P), _P, _P]),
    "plm_hip_batch_score_device": (ctypes.c_int, [_P, ctypes.c_int, _P]),
    "plm_hip_batch_capture": (ctypes.c_int, [_P, ctypes.c_int]),
    "plm_hip_batch_launch": (ctypes.c_int, [_P, _P]),
    "plm_hip_batch_set_tuning": (ctypes.c_int, [_P, ctypes.c_int]),
    "plm_hip_batch_set_occupancy": (ctypes.c_int, [_P, ctypes.c_int]),
    "plm_hip_batch_get_tuning": (ctypes.c_int, [_P, _P, _P]),
    "plm_hip_batch_autotune": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_double)]),
    "plm_hip_batch_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "plm_hip_batch_get_timing": (ctypes.c_int, [_P, _P, ctypes.c_int,
                                                ctypes.POINTER(ctypes.c_int64)]),
    "plm_hip_batch_destroy": (ctypes.c_int, [_P]),
    "plm_hip_resample_geometry": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(PlmImgDesc), _P, _P, _P, _P,
                                                 ctypes.POINTER(ctypes.c_int)]),
    "plm_hip_volume_resample": (ctypes.c_int, [ctypes.POINTER(PlmImgDesc), _P, _P, _P, ctypes.c_int, _P,
                                               ctypes.c_int, ctypes.c_int]),
    "plm_hip_stage_resample": (ctypes.c_int, [ctypes.c_int, _P, ctypes.POINTER(PlmImgDesc), ctypes.c_int, _P,
                                              ctypes.c_int, ctypes.c_int]),
    "plm_hip_hist_eqsp": (ctypes.c_int, [ctypes.POINTER(PlmImgDesc), ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                         ctypes.c_int]),
    "plm_hip_hist_vopt": (ctypes.c_int, [ctypes.POINTER(PlmImgDesc), ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _P,
                                         ctypes.c_int]),
    "plm_hip_bspline_set_hist_keys": (ctypes.c_int, [_P, _P, _P]),
    "plm_hip_lbfgs_parms_init": (None, [ctypes.POINTER(PlmLbfgsParms)]),
    "plm_hip_bspline_optimize_lbfgs": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(PlmRegParms),
                                                      ctypes.POINTER(PlmLbfgsParms), _P, _P, PLM_LBFGS_PROGRESS,
                                                      _P, ctypes.POINTER(PlmLbfgsResult)]),
    "plm_hip_bspline_optimize_lbfgs_resident": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int,
                                                               ctypes.POINTER(PlmRegParms),
                                                               ctypes.POINTER(PlmLbfgsParms), _P, _P, _P,
                                                               ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                               ctypes.POINTER(PlmLbfgsResult)]),
}

_lib = None


class PlmHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"plm_hip error {code}: {msg}")
        self.code = code


def load():
    """Load the HIP flavor library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    # A ROCm torch wheel carries its own libamdhip64.so.7, which it loads
    # under a name the dynamic loader does not match against an already
    # loaded copy.  Loading torch first makes this library bind to torch's
    # runtime (same soname), so the process holds one HIP runtime; the
    # other order would leave torch with a second runtime that sees no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc < 0:
        msg = load().plm_hip_last_error()
        raise PlmHipError(rc, msg.decode() if msg else "")
    return rc
