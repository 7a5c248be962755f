"""Loader of libcsx.so (the HIP kernels + C ABI declared in include/csx.h).

There is no fallback: if the in-tree library is missing or fails to load,
every entry point raises.  Build it with ``python -m crystalscatter_amd.build``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcsx.so")

CSX_PREC_REF64 = 0
CSX_PREC_MIXED = 1
NCOEF = 151
NIM2 = 131


class CsxError(RuntimeError):
    pass


class Scalars(ctypes.Structure):
    """Mirror of csx_scalars (include/csx.h)."""
    _fields_ = [
        ("combo_key", ctypes.c_uint32), ("zmax", ctypes.c_int32), ("centerbeam", ctypes.c_int32),
        ("orcase", ctypes.c_int32), ("fix_orcase", ctypes.c_int32), ("peakmax1", ctypes.c_int32),
        ("peakmax2", ctypes.c_int32), ("dsch", ctypes.c_int32), ("para", ctypes.c_int32),
        ("twin", ctypes.c_int32), ("ordis", ctypes.c_int32),
    ] + [(n, ctypes.c_double) for n in (
        "qmax", "pixnox", "pixnoy", "pixx_m", "pixy_m", "det", "wave", "beamx", "beamy",
        "radius", "sigma", "length", "sigmal", "dbeta", "polTheta", "polPhi",
        "cosphi", "sinphi", "costheta", "sintheta", "cosphic", "sinphic")] + [
        ("limq", ctypes.c_double * 18), ("norm", ctypes.c_double),
    ] + [(n, ctypes.c_double) for n in (
        "base", "i0", "iso", "ifluc", "rfluc", "dwfactor", "uca", "ucb", "domainsize",
        "width_zuf", "reldis", "dist", "phiwidth", "ceff", "por", "sphno_cubevol")] + [
        ("xmin", ctypes.c_int32), ("xmax", ctypes.c_int32), ("ymin", ctypes.c_int32),
        ("ymax", ctypes.c_int32), ("iimin", ctypes.c_int32), ("cs", ctypes.c_int32),
        ("use1d", ctypes.c_int32), ("qmin", ctypes.c_double), ("axmin", ctypes.c_int32),
        ("axmax", ctypes.c_int32), ("aymin", ctypes.c_int32), ("aymax", ctypes.c_int32),
        ("peak_eta", ctypes.c_double), ("peak_beta", ctypes.c_double), ("peak_c", ctypes.c_double * 5),
        ("peak_gamma_b", ctypes.c_double), ("peak_gamma_bh", ctypes.c_double),
    ]


class Grid(ctypes.Structure):
    _fields_ = [("xmin", ctypes.c_int32), ("xmax", ctypes.c_int32), ("ymin", ctypes.c_int32),
                ("ymax", ctypes.c_int32)]


class Mask(ctypes.Structure):
    _fields_ = [("rects", ctypes.c_int32 * 16), ("bstop", ctypes.c_int32), ("border", ctypes.c_int32)]


NM_MAXPAR = 30
TUNE_VARIANTS, TUNE_SIZES = 4, 4
PASS_RING = 64  # CSX_PASS_RING
NM_REFERENCE, NM_CORRECTED = 0, 1
OBJECTIVE = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double))


class NMResult(ctypes.Structure):
    """Mirror of csx_nm_result."""
    _fields_ = [("x", ctypes.c_double * NM_MAXPAR), ("f", ctypes.c_double), ("iters", ctypes.c_int32),
                ("evals", ctypes.c_int32)]


class NMStats(ctypes.Structure):
    """Mirror of csx_nm_stats."""
    _fields_ = [("rounds", ctypes.c_int64), ("evals", ctypes.c_int64), ("prep_ms", ctypes.c_double),
                ("stage_ms", ctypes.c_double), ("run_ms", ctypes.c_double), ("pruned", ctypes.c_int64),
                ("screen_ms", ctypes.c_double), ("launch_ms", ctypes.c_double)]


SCREEN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double))


class NMScreenOpts(ctypes.Structure):
    """Mirror of csx_nm_screen_opts."""
    _fields_ = [("fn", SCREEN), ("user", ctypes.c_void_p), ("margin", ctypes.c_double)]


class NMFit(ctypes.Structure):
    """Mirror of csx_nm_fit."""
    _fields_ = [("base", ctypes.c_void_p), ("d_data", ctypes.c_void_p), ("w", ctypes.c_int32), ("h", ctypes.c_int32),
                ("mask", Mask), ("x0", ctypes.c_double * NM_MAXPAR), ("lo", ctypes.c_double * NM_MAXPAR),
                ("hi", ctypes.c_double * NM_MAXPAR)]


_LIB = None

# (name, restype, argtypes) of every exported entry point of include/csx.h.
_VP, _CP, _D, _I = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_int
_PD, _PI32, _PF = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float)
_PI64 = ctypes.POINTER(ctypes.c_int64)
SIGNATURES = [
    ("csx_create", _I, [_I, ctypes.POINTER(_VP)]),
    ("csx_destroy", None, [_VP]),
    ("csx_set_stream", _I, [_VP, _VP]),
    ("csx_synchronize", _I, [_VP]),
    ("csx_upload", _I, [_VP, ctypes.POINTER(Scalars), _PD, _PD, _PI32, _PI32, _PF]),
    ("csx_render", _I, [_VP, ctypes.POINTER(Grid), _I, _VP]),
    ("csx_render_stats", _I, [_VP, ctypes.POINTER(Grid), _I, _VP, ctypes.POINTER(ctypes.c_uint64)]),
    ("csx_qgrid", _I, [_VP, ctypes.POINTER(Grid), _VP]),
    ("csx_fit_fqs", _I, [_VP, _I, _I, ctypes.POINTER(Mask), _I, _VP, _PD, _PI64, _PI64, _VP]),
    ("csx_fit_fqs_noalloc", _I, [_VP, _I, _I, ctypes.POINTER(Mask), _I, _VP, _PD, _PI64, _PI64, _VP]),
    ("csx_fit_batch_begin", _I, [_VP]),
    ("csx_fit_batch_stage", _I, [_VP, ctypes.POINTER(Scalars), _PD, _PI32, _PI32, _PF, _I, _I,
                                 ctypes.POINTER(Mask), _VP]),
    ("csx_fit_batch_run", _I, [_VP, _I, _PD, _PI64]),
    ("csx_render_batch_run", _I, [_VP, _I]),
    ("csx_tune_set", _I, [ctypes.c_uint32, _I, _I]),
    ("csx_tune_get", _I, [ctypes.c_uint32, _I]),
    ("csx_tune_load", _I, [_CP]),
    ("csx_tune_force", _I, [_VP, _I]),
    ("csx_set_batch_graph", _I, [_VP, _I]),
    ("csx_batch_graph_launches", ctypes.c_longlong, [_VP]),
    ("csx_set_device_prep", _I, [_VP, _I]),
    ("csx_engine_stage_points", _I, [_VP, _VP, _I, _I, ctypes.POINTER(_CP), _PD, _VP, ctypes.c_int64, _I]),
    ("csx_device_prep_counts", _I, [_VP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    ("csx_fit_batch_table", _I, [_VP, _I, _I, _VP, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("csx_nm_run_cb", _I, [_I, _I, _I, OBJECTIVE, _VP, _PD, _PD, _PD, _D, _I, _D, ctypes.POINTER(NMResult)]),
    ("csx_nm_batch", _I, [_VP, _I, ctypes.POINTER(NMFit), _I, _I, _I, ctypes.POINTER(_CP), _D, _I, _D, _I, _I,
                          ctypes.POINTER(NMResult), ctypes.POINTER(NMStats)]),
    ("csx_nm_batch_screened", _I, [_VP, _I, ctypes.POINTER(NMFit), _I, _I, _I, ctypes.POINTER(_CP), _D, _I, _D, _I,
                                   _I, ctypes.POINTER(NMScreenOpts), ctypes.POINTER(NMResult),
                                   ctypes.POINTER(NMStats)]),
    ("csx_nm_eval", _I, [_VP, _I, ctypes.POINTER(NMFit), _I, ctypes.POINTER(_CP), _I, _PD, _I, _I, _PD]),
    ("csx_last_error", _I, [_VP, _CP, _I]),
    ("csx_last_kernel_ms", _D, [_VP]),
    ("csx_set_timing", _I, [_VP, _I]),
    ("csx_pass_times", _I, [_VP, _I, _PD]),
    ("csx_engine_new", _VP, [_I]),
    ("csx_engine_free", None, [_VP]),
    ("csx_engine_set", _I, [_VP, _CP, _D]),
    ("csx_engine_get", _D, [_VP, _CP]),
    ("csx_engine_prepare", _I, [_VP]),
    ("csx_engine_calc", _I, [_VP, _I, _PD, _PI32]),
    ("csx_engine_calc_device", _I, [_VP, _I, _VP, _PI32]),
    ("csx_engine_set_fit_data", _I, [_VP, _I, _I, _PD]),
    ("csx_engine_set_nofit_rect", _I, [_VP, _I, _I, _I, _I, _I]),
    ("csx_engine_fit", _I, [_VP, _I, _I, _I, _PD, _PI64, _PI64]),
    ("csx_engine_table", _VP, [_VP, _CP, _PI64, ctypes.POINTER(_I)]),
    ("csx_engine_scalars", ctypes.POINTER(Scalars), [_VP]),
    ("csx_engine_ctx", _VP, [_VP]),
    ("csx_engine_last_error", _I, [_VP, _CP, _I]),
    ("csx_engine_stage_image", _I, [_VP, _VP, _VP, ctypes.c_int64]),
    ("csx_rand_new", _VP, [ctypes.c_uint32]),
    ("csx_rand_free", None, [_VP]),
    ("csx_rand_next", ctypes.c_int32, [_VP]),
    ("csx_rand_zuf", _I, [_VP, _PD, ctypes.c_int64]),
    ("csx_rand_skip", _I, [_VP, ctypes.c_int64]),
    ("csx_rand_zuf_device", _I, [_VP, _VP, ctypes.c_int64, _VP]),
    ("csx_engine_tpv", _I, [_VP, _CP, _VP, _CP, _I]),
    ("csx_engine_tpv_value", _D, [_VP, _CP]),
    ("csx_engine_tpv_batch", _I, [_VP, _VP, _CP, _VP, _I, _VP, ctypes.c_int64, _VP, _I, _PD, _I]),
    ("csx_post_noise", _I, [_VP, _VP, _VP, ctypes.c_int64, _I]),
    ("csx_post_zero", _I, [_VP, _VP, _VP, ctypes.c_int64]),
    ("csx_post_zero_rects", _I, [_VP, _VP, ctypes.c_int64, _I, _VP, _I, _I, _I, _I, _I]),
    ("csx_post_scale_intensity", _I, [_VP, _VP, ctypes.c_int64, _I, _I]),
    ("csx_post_log_scale", _I, [_VP, _VP, ctypes.c_int64, _I, _I, _I, _I, _I]),
    ("csx_post_scale_clip", _I, [_VP, _VP, ctypes.c_int64, _I, _I, _I, _I, _I]),
    ("csx_post_rphi", _I, [_VP, _VP, _I, _I, _I, _I, _I, _I, _I, _I, _VP, _I, _I, _I]),
    ("csx_post_rphi_many", _I, [_VP, _VP, _I, _I, _I, _I, _VP, _I, _I, _VP, _I, _I, _I]),
    ("csx_post_ifft", _I, [_VP, _VP, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _VP]),
    ("csx_mlp_nparams", ctypes.c_int64, []),
    ("csx_mlp_create", _I, [_I, _I, ctypes.c_uint64, ctypes.POINTER(_VP)]),
    ("csx_mlp_destroy", None, [_VP]),
    ("csx_mlp_set_stream", _I, [_VP, _VP]),
    ("csx_mlp_last_error", _I, [_VP, _CP, _I]),
    ("csx_mlp_get_params", _I, [_VP, _PF]),
    ("csx_mlp_set_params", _I, [_VP, _PF]),
    ("csx_mlp_get_features", _I, [_VP, _PF, _PI32, _PF]),
    ("csx_mlp_set_features", _I, [_VP, _PF, _PI32, _PF]),
    ("csx_mlp_forward", _I, [_VP, _I, _I, _VP, _VP, _VP, _VP]),
    ("csx_mlp_predict", _I, [_VP, _I, _I, _VP, _VP, _VP, _VP]),
    ("csx_mlp_gradient", _I, [_VP, _I, _I, _VP, _VP, _VP, _VP, _PF]),
    ("csx_mlp_get_grads", _I, [_VP, _PF]),
    ("csx_mlp_adam", _I, [_VP, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("csx_mlp_train_step", _I, [_VP, _I, _I, _VP, _VP, _VP, _VP, ctypes.c_float, _PF]),
    ("csx_mlp_synchronize", _I, [_VP]),
]
MLP_NPAR, MLP_NFREQ, MLP_IN, MLP_IN_PAD, MLP_HIDDEN, MLP_LAYERS = 6, 32, 70, 128, 256, 4


def lib():
    """Load libcsx.so; raises CsxError if it is missing (no fallback path exists)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise CsxError("libcsx.so is not built: run `python -m crystalscatter_amd.build`")
        # One HIP runtime per process.  PyTorch ships its own libamdhip64 with the same
        # soname as ROCm's: loaded first, it also serves libcsx.so; loaded after libcsx.so,
        # a second runtime comes up and torch finds no GPU.  So torch, when present, goes first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB
