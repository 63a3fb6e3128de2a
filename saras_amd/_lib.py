"""ctypes binding of the C ABI in include/*.h (libsaras_mg.so, gfx950).

The library is loaded from saras_amd/_build; a missing library is an error (there is no CPU
fallback for any solver entry point)."""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libsaras_mg.so")

SR_OK, SR_ERR_ARG, SR_ERR_HIP, SR_ERR_STATE, SR_ERR_NOMEM, SR_ERR_NOT_CONVERGED, SR_ERR_IO = 0, 1, 2, 3, 4, 5, 6
SR_PRESSURE, SR_RESIDUAL, SR_TMP = 0, 1, 2
SR_SEM_REFERENCE, SR_SEM_PROJECTED = 0, 1
SR_PREC_FP64, SR_PREC_MIXED, SR_PREC_MIXED_GCR = 0, 1, 2

_dp = C.POINTER(C.c_double)


class SrMgDesc(C.Structure):
    _fields_ = [
        ("xInd", C.c_int), ("yInd", C.c_int), ("zInd", C.c_int),
        ("vcDepth", C.c_int), ("vcCount", C.c_int),
        ("preSmooth", C.c_int), ("postSmooth", C.c_int),
        ("gsSmooth", C.c_int), ("resType", C.c_int),
        ("xPer", C.c_int), ("yPer", C.c_int), ("zPer", C.c_int),
        ("testPoisson", C.c_int), ("semantics", C.c_int), ("precision", C.c_int),
        ("padMode", C.c_int), ("coarseProject", C.c_int), ("coarseDirect", C.c_int),
        ("mgTolerance", C.c_double), ("relTolerance", C.c_double),
        ("dXi", C.c_double), ("dEt", C.c_double), ("dZt", C.c_double),
        ("xix2", _dp), ("xixx", _dp), ("ety2", _dp), ("etyy", _dp), ("ztz2", _dp), ("ztzz", _dp),
        ("wall", _dp * 6),
        ("innerPreSmooth", C.c_int), ("innerPostSmooth", C.c_int),
    ]


class SrMgStats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("converged", C.c_int), ("residual", C.c_double), ("solve_ms", C.c_double),
                ("sweep0_ms", C.c_double), ("residual0", C.c_double)]


SR_TS_VX, SR_TS_VY, SR_TS_VZ, SR_TS_P, SR_TS_PP, SR_TS_MGRHS = 0, 1, 2, 3, 4, 5
SR_TS_HX, SR_TS_HY, SR_TS_HZ, SR_TS_DIV, SR_TS_T, SR_TS_TRHS = 6, 7, 8, 12, 13, 14
SR_TS_PP1, SR_TS_PP2 = 15, 16


class SrTsDesc(C.Structure):
    _fields_ = [
        ("xInd", C.c_int), ("yInd", C.c_int), ("zInd", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("xGrid", C.c_int), ("yGrid", C.c_int), ("zGrid", C.c_int),
        ("betaX", C.c_double), ("betaY", C.c_double), ("betaZ", C.c_double),
        ("xPer", C.c_int), ("yPer", C.c_int), ("zPer", C.c_int),
        ("probType", C.c_int),
        ("Re", C.c_double), ("dt", C.c_double), ("cnTolerance", C.c_double),
        ("vcDepth", C.c_int), ("vcCount", C.c_int), ("preSmooth", C.c_int), ("postSmooth", C.c_int),
        ("resType", C.c_int), ("gsSmooth", C.c_int),
        ("mgTolerance", C.c_double),
        ("scalar", C.c_int), ("rbcType", C.c_int), ("Ra", C.c_double), ("Pr", C.c_double),
        ("semantics", C.c_int), ("relTolerance", C.c_double), ("precision", C.c_int),
        ("coarseDirect", C.c_int),
        ("innerPreSmooth", C.c_int), ("innerPostSmooth", C.c_int), ("padMode", C.c_int),
    ]


class SrTsStats(C.Structure):
    _fields_ = [("jacobi_iters", C.c_int * 4), ("mg_cycles", C.c_int), ("step_ms", C.c_double),
                ("solve_ms", C.c_double), ("guess_ms", C.c_double), ("guess_choice", C.c_int),
                ("rel_guess", C.c_double), ("train_loss", C.c_double), ("sweep0_ms", C.c_double),
                ("mg_converged", C.c_int), ("mg_residual", C.c_double),
                ("guess_rel_linear", C.c_double), ("guess_rel_learned", C.c_double),
                ("guess_predict_ms", C.c_double)]


SR_GUESS_LAYERS, SR_GUESS_IN_CH = 12, 8
_fp = C.POINTER(C.c_float)


class SrGuessDesc(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("fine", C.c_int * 3), ("seed", C.c_ulonglong), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("history", C.c_int)]


class SrGuessStats(C.Structure):
    _fields_ = [("rel_warm", C.c_double), ("rel_extrap", C.c_double), ("rel_learned", C.c_double),
                ("choice", C.c_int), ("predict_ms", C.c_double), ("rel_linear", C.c_double), ("net_used", C.c_int)]


SR_LOC_NODE, SR_LOC_VX, SR_LOC_VY, SR_LOC_VZ = 0, 1, 2, 3
SR_BC_PERIODIC, SR_BC_DIRICHLET, SR_BC_NEUMANN, SR_BC_NONE = 0, 1, 2, 3


class SrFgridDesc(C.Structure):
    _fields_ = [("xInd", C.c_int), ("yInd", C.c_int), ("zInd", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
                ("xGrid", C.c_int), ("yGrid", C.c_int), ("zGrid", C.c_int),
                ("betaX", C.c_double), ("betaY", C.c_double), ("betaZ", C.c_double),
                ("xPer", C.c_int), ("yPer", C.c_int), ("zPer", C.c_int), ("iScheme", C.c_int)]


class SrBc(C.Structure):
    _fields_ = [("kind", C.c_int * 6), ("value", C.c_double * 6)]


class SrJacobiDesc(C.Structure):
    _fields_ = [("form", C.c_int), ("dt", C.c_double), ("Re", C.c_double), ("coef", C.c_double),
                ("cnTolerance", C.c_double), ("zCollocMetric", C.c_int)]


# (name, restype, argtypes) of every entry point declared in include/*.h
_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
SIGNATURES = [
    ("sr_mg_create", C.c_int, [C.POINTER(SrMgDesc), C.POINTER(C.c_void_p)]),
    ("sr_mg_destroy", C.c_int, [_vp]),
    ("sr_mg_solve", C.c_int, [_vp, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(SrMgStats)]),
    ("sr_mg_history", C.c_int, [_vp, _dp, C.c_int]),
    ("sr_mg_residual", C.c_int, [_vp, C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    ("sr_mg_set_level", C.c_int, [_vp, C.c_int]),
    ("sr_mg_get_level", C.c_int, [_vp, _ip]),
    ("sr_mg_set_zero_bc", C.c_int, [_vp, C.c_int]),
    ("sr_mg_upload", C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    ("sr_mg_download", C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    ("sr_mg_level_dims", C.c_int, [_vp, C.c_int, _ip, _ip, _ip]),
    ("sr_mg_op_smooth", C.c_int, [_vp, C.c_int]),
    ("sr_mg_op_compute_residual", C.c_int, [_vp]),
    ("sr_mg_op_coarsen", C.c_int, [_vp]),
    ("sr_mg_op_prolong", C.c_int, [_vp]),
    ("sr_mg_op_solve_coarsest", C.c_int, [_vp]),
    ("sr_mg_op_impose_bc", C.c_int, [_vp]),
    ("sr_mg_op_update_pads", C.c_int, [_vp, C.c_int]),
    ("sr_mg_op_compute_error", C.c_int, [_vp, C.c_int, _dp]),
    ("sr_mg_op_vcycle", C.c_int, [_vp]),
    ("sr_mg_time_kernel", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    ("sr_mg_kernel_bytes", C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    ("sr_mg_info", C.c_int, [_vp, C.c_int, C.POINTER(C.c_long)]),
    ("sr_grid_axis", C.c_int, [C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp]),
    ("sr_device_alloc", C.c_int, [C.c_ulong, C.POINTER(C.c_void_p)]),
    ("sr_device_free", C.c_int, [C.c_void_p]),
    ("sr_memcpy", C.c_int, [C.c_void_p, C.c_void_p, C.c_ulong, C.c_int]),
    ("sr_last_error", C.c_char_p, []),
    ("sr_ts_create", C.c_int, [C.POINTER(SrTsDesc), C.POINTER(C.c_void_p)]),
    ("sr_ts_destroy", C.c_int, [_vp]),
    ("sr_ts_init_tgv", C.c_int, [_vp]),
    ("sr_ts_upload", C.c_int, [_vp, C.c_int, _dp]),
    ("sr_ts_download", C.c_int, [_vp, C.c_int, _dp]),
    ("sr_ts_field_dims", C.c_int, [_vp, C.c_int, _ip, _ip, _ip]),
    ("sr_ts_time_advance", C.c_int, [_vp, C.POINTER(SrTsStats)]),
    ("sr_ts_deriv", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    ("sr_ts_build_rhs", C.c_int, [_vp]),
    ("sr_ts_solve_v", C.c_int, [_vp, C.c_int, _ip]),
    ("sr_ts_impose_vbc", C.c_int, [_vp, C.c_int]),
    ("sr_ts_divergence", C.c_int, [_vp]),
    ("sr_ts_energy", C.c_int, [_vp, _dp]),
    ("sr_ts_max_div", C.c_int, [_vp, _dp]),
    ("sr_ts_init_rbc", C.c_int, [_vp, C.c_ulonglong, C.c_double]),
    ("sr_ts_impose_tbc", C.c_int, [_vp]),
    ("sr_ts_build_trhs", C.c_int, [_vp]),
    ("sr_ts_solve_t", C.c_int, [_vp, _ip]),
    ("sr_ts_nusselt", C.c_int, [_vp, _dp]),
    ("sr_ts_coefficients", C.c_int, [_vp, _dp, _dp]),
    ("sr_ts_attach_guess", C.c_int, [_vp, _vp, C.c_int]),
    ("sr_guess_create", C.c_int, [C.POINTER(SrGuessDesc), C.POINTER(C.c_void_p)]),
    ("sr_guess_destroy", C.c_int, [_vp]),
    ("sr_guess_layer_info", C.c_int, [C.c_int, _ip, _ip, _ip]),
    ("sr_guess_set_layer", C.c_int, [_vp, C.c_int, _fp, _fp]),
    ("sr_guess_get_layer", C.c_int, [_vp, C.c_int, _fp, _fp]),
    ("sr_guess_get_grad", C.c_int, [_vp, C.c_int, _fp, _fp]),
    ("sr_guess_forward", C.c_int, [_vp, _fp, _fp]),
    ("sr_guess_backward", C.c_int, [_vp, _fp]),
    ("sr_guess_adam", C.c_int, [_vp]),
    ("sr_guess_tensor", C.c_int, [_vp, C.c_int, _fp, C.POINTER(C.c_long)]),
    ("sr_guess_predict", C.c_int, [_vp, _vp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(SrGuessStats)]),
    ("sr_guess_train", C.c_int, [_vp, C.c_void_p, _dp]),
    ("sr_ts_pressure_solver", C.c_int, [_vp, C.POINTER(C.c_void_p)]),
    ("sr_guess_observe", C.c_int, [_vp, _vp, C.c_void_p, C.c_void_p]),
    ("sr_guess_kappa", C.c_int, [_vp, _dp, _ip]),
    ("sr_guess_extrapolate", C.c_int, [_vp, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("sr_guess_set_policy", C.c_int, [_vp, C.c_int]),
    ("sr_guess_time_layer", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    ("sr_fgrid_create", C.c_int, [C.POINTER(SrFgridDesc), C.POINTER(C.c_void_p)]),
    ("sr_fgrid_destroy", C.c_int, [_vp]),
    ("sr_fgrid_dims", C.c_int, [_vp, C.c_int, _ip, _ip, _ip]),
    ("sr_fgrid_stream", C.c_int, [_vp, C.POINTER(C.c_void_p)]),
    ("sr_fgrid_sync", C.c_int, [_vp]),
    ("sr_deriv1", C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp]),
    ("sr_deriv2", C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp]),
    ("sr_vf_diff", C.c_int, [_vp] + [_vp] * 6),
    ("sr_sf_diff", C.c_int, [_vp, _vp, _vp]),
    ("sr_vf_nlin", C.c_int, [_vp] + [_vp] * 9),
    ("sr_sf_nlin", C.c_int, [_vp] + [_vp] * 5),
    ("sr_divergence", C.c_int, [_vp] + [_vp] * 4),
    ("sr_gradient", C.c_int, [_vp] + [_vp] * 4),
    ("sr_bc_fill", C.c_int, [_vp, C.c_int, C.POINTER(SrBc), _vp]),
    ("sr_jacobi_visc", C.c_int, [_vp, C.c_int, C.POINTER(SrJacobiDesc), C.POINTER(SrBc), _vp, _vp, _vp, _ip]),
    ("sr_probe_gather", C.c_int, [C.c_int, C.POINTER(C.c_void_p), _ip, C.c_int, _ip, _dp]),
    ("sr_probe_gather_noalloc", C.c_int, [C.c_int, C.POINTER(C.c_void_p), _ip, C.c_int, _ip, _dp]),
    ("sr_ts_field_ptr", C.c_int, [_vp, C.c_int, C.POINTER(C.c_void_p)]),
    ("sr_ts_kernel_bytes", C.c_int, [_vp, C.c_int, _dp]),
    ("sr_ts_time_kernel", C.c_int, [_vp, C.c_int, C.c_int, _dp]),
]

_lib = None


class SarasError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"saras_mg error {code}: {msg}")
        self.code = code


def load():
    """Load libsaras_mg.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code):
    if code != SR_OK:
        msg = load().sr_last_error()
        raise SarasError(code, msg.decode() if msg else "")
    return code


class DeviceArray:
    """A float64 buffer in HBM owned through the C ABI (no torch involved)."""

    def __init__(self, shape):
        import numpy as np
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        check(load().sr_device_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape
        check(load().sr_memcpy(self.ptr, a.ctypes.data, self.nbytes, 1))

    def download(self):
        import numpy as np
        a = np.empty(self.shape)
        check(load().sr_memcpy(a.ctypes.data, self.ptr, self.nbytes, 2))
        return a

    def free(self):
        if self.ptr:
            load().sr_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
