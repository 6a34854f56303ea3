"""ctypes binding of libigap4h.so (include/igap4h.h) and the reference struct
layouts it takes (include/phys.h:11-26, include/nurbs.h:4-24,
include/bc.h:4-11, include/iga.h:8-25).

The HIP library is the product path: importing this module does not fall back
to anything if the library is missing; `lib()` raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "csrc", "libigap4h.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_i32_p = ctypes.POINTER(ctypes.c_int32)
c_i64_p = ctypes.POINTER(ctypes.c_int64)
_fnp = ctypes.c_void_p  # host function pointers (never called by the library)


class Phys(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int), ("ndof", ctypes.c_int), ("torder", ctypes.c_int),
                ("sorder", ctypes.c_int), ("par_mat", c_double_p), ("ndensity", ctypes.c_int),
                ("nfield", ctypes.c_int), ("residual", _fnp), ("tangent", _fnp), ("field", _fnp),
                ("density", _fnp), ("densityb", _fnp), ("assert_par_mat", _fnp)]


class Nurbs(ctypes.Structure):
    _fields_ = [("porder", ctypes.c_int), ("nelem", c_int_p), ("nbasis", c_int_p), ("nknot", c_int_p),
                ("knotVector", ctypes.POINTER(c_double_p)), ("nboun", ctypes.c_int), ("boun", c_int_p),
                ("nsendPart", ctypes.c_int), ("sendPart", c_int_p), ("nrecvPart", ctypes.c_int),
                ("recvPart", c_int_p), ("sendBlock", c_int_p), ("recvBlock", c_int_p),
                ("universalIdx_temp", c_int_p), ("bc_periodic", c_int_p), ("nelem_global", c_int_p),
                ("knotVector_global", ctypes.POINTER(c_double_p))]


class BC(ctypes.Structure):
    _fields_ = [("bc_type", c_int_p), ("par_periodic", c_double_p), ("par_dirichlet", c_double_p),
                ("par_neumann", c_double_p)]


class IGA(ctypes.Structure):
    _fields_ = [("phys", ctypes.POINTER(Phys)), ("nurbs", ctypes.POINTER(Nurbs)), ("bc", ctypes.POINTER(BC)),
                ("U_hist", c_double_p), ("nselfIdx", ctypes.c_int), ("selfIdx", c_int_p),
                ("nsendIdx", ctypes.c_int), ("sendIdx", c_int_p), ("sendPtr", c_int_p),
                ("nrecvIdx", ctypes.c_int), ("recvIdx", c_int_p), ("recvPtr", c_int_p),
                ("globalIdx", c_int_p)]


NEUMANN_FN = ctypes.CFUNCTYPE(None, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_double_p)


class Opts(ctypes.Structure):
    _fields_ = [("physics", ctypes.c_int), ("nquad", ctypes.c_int), ("elem_coef", c_double_p),
                ("neumann", NEUMANN_FN)]


IGAP4H_OK, IGAP4H_EINVAL, IGAP4H_ENOMEM, IGAP4H_EHIP, IGAP4H_ENODEV = 0, -1, -2, -3, -4
IGAP4H_FP64, IGAP4H_FP32 = 0, 1
IGAP4H_VALS_F64, IGAP4H_VALS_F32 = 0, 1
IGAP4H_PHYS_GE, IGAP4H_PHYS_GETIME, IGAP4H_PHYS_FROM_TORDER = 0, 1, -1
IGAP4H_PC_NONE, IGAP4H_PC_BJACOBI3, IGAP4H_PC_SGS, IGAP4H_PC_MG = 0, 1, 2, 3
IGAP4H_KSP_GMRES, IGAP4H_KSP_CG = 0, 1
IGAP4H_MG_GS, IGAP4H_MG_CHEBYSHEV, IGAP4H_MG_CHEB4 = 0, 1, 2
IGAP4H_SMOOTH_SAME, IGAP4H_SMOOTH_BF16, IGAP4H_SMOOTH_F16 = 0, 1, 2
IGAP4H_TANGENT_AUTO, IGAP4H_TANGENT_TSTRIP, IGAP4H_TANGENT_SV = 0, 1, 2
IGAP4H_BATCH_MAX_VARIANTS = 8


class NewtonOpts(ctypes.Structure):
    """igap4h_newton_opts (include/igap4h.h)."""
    _fields_ = [("atol", ctypes.c_double), ("rtol", ctypes.c_double), ("max_newton", ctypes.c_int),
                ("tangent_prec", ctypes.c_int), ("vals_type", ctypes.c_int), ("ksp_type", ctypes.c_int),
                ("restart", ctypes.c_int), ("ksp_rtol", ctypes.c_double), ("ksp_max_it", ctypes.c_int),
                ("ksp_guess", ctypes.c_int), ("pc", ctypes.c_int), ("line_search", ctypes.c_int),
                ("stagnation", ctypes.c_int), ("dirichlet_lift", ctypes.c_int),
                ("mg_pre", ctypes.c_int), ("mg_post", ctypes.c_int), ("mg_graph", ctypes.c_int),
                ("mg_smoother", ctypes.c_int), ("lag_jacobian", ctypes.c_int), ("lag_persist", ctypes.c_int),
                ("mg_smooth_vals", ctypes.c_int), ("mg_coarse_deg", ctypes.c_int),
                ("mg_cheb_ratio", ctypes.c_double)]


class NewtonInfo(ctypes.Structure):
    """igap4h_newton_info (include/igap4h.h)."""
    _fields_ = [("converged", ctypes.c_int), ("newton_its", ctypes.c_int), ("ksp_its", ctypes.c_int),
                ("fnorm0", ctypes.c_double), ("fnorm", ctypes.c_double), ("t_residual", ctypes.c_double),
                ("t_tangent", ctypes.c_double), ("t_solve", ctypes.c_double), ("t_total", ctypes.c_double),
                ("fnorm_hist", ctypes.c_double * 64), ("ksp_its_hist", ctypes.c_int * 64), ("tangents", ctypes.c_int)]

    def as_dict(self):
        k = min(self.newton_its + 1, 64)
        return {"converged": self.converged, "newton_its": self.newton_its, "ksp_its": self.ksp_its,
                "fnorm0": self.fnorm0, "fnorm": self.fnorm, "t_residual": self.t_residual,
                "t_tangent": self.t_tangent, "t_solve": self.t_solve, "t_total": self.t_total,
                "fnorm_hist": list(self.fnorm_hist[:k]), "ksp_its_hist": list(self.ksp_its_hist[:k - 1]),
                "tangents": self.tangents}

# every entry point declared in include/igap4h.h
SYMBOLS = ("igap4h_create", "igap4h_destroy", "igap4h_last_error", "igap4h_sizes", "igap4h_pattern",
           "igap4h_pattern_dev", "igap4h_dirichlet_rows", "igap4h_residual", "igap4h_tangent",
           "igap4h_residual_dev", "igap4h_residual_dev_noalloc", "igap4h_tangent_dev", "igap4h_set_coef", "igap4h_energy_dev", "igap4h_intdensity_dev",
           "igap4h_newton_default_opts", "igap4h_newton_solve_dev", "igap4h_spmv_dev", "igap4h_linear_solve_dev",
           "igap4h_residual_vjp_coef_dev", "igap4h_batch_create", "igap4h_batch_destroy", "igap4h_batch_sizes",
           "igap4h_batch_tune", "igap4h_batch_set_variant", "igap4h_batch_assemble", "igap4h_write_solution",
           "igap4h_read_solution", "igap4h_knot_insert", "igap4h_update_bc", "igap4h_update_neumann",
           "igap4h_set_tangent_form")

_lib = None


def lib():
    """Load libigap4h.so (build it with __graft_entry__.build() or
    `make -C igap4standard_amd/csrc`).  Raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libigap4h.so is not built: %s (run __graft_entry__.build())" % LIB_PATH)
        # torch ships its own libamdhip64.so.7 (same soname as /opt/rocm's).
        # Loading torch first makes libigap4h.so bind to that one runtime;
        # loading this library first would map two HIP runtimes into the
        # process and torch could no longer see the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.igap4h_create.argtypes = [ctypes.POINTER(IGA), ctypes.POINTER(Opts), ctypes.POINTER(vp)]
        L.igap4h_destroy.argtypes = [vp]
        L.igap4h_destroy.restype = None
        L.igap4h_last_error.restype = ctypes.c_char_p
        L.igap4h_sizes.argtypes = [vp, c_i64_p, c_i64_p, c_i64_p, c_i64_p, c_int_p]
        L.igap4h_pattern.argtypes = [vp, c_i64_p, c_i32_p]
        L.igap4h_pattern_dev.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]
        L.igap4h_dirichlet_rows.argtypes = [vp, c_i32_p]
        L.igap4h_residual.argtypes = [vp, c_double_p, c_double_p, c_double_p]
        L.igap4h_tangent.argtypes = [vp, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, vp]
        L.igap4h_residual_dev.argtypes = [vp, vp, vp, vp]
        L.igap4h_residual_dev_noalloc.argtypes = [vp, vp, vp, vp]
        L.igap4h_tangent_dev.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.igap4h_set_coef.argtypes = [vp, c_double_p]
        L.igap4h_energy_dev.argtypes = [vp, vp, c_double_p, vp]
        L.igap4h_intdensity_dev.argtypes = [vp, vp, c_double_p, c_int_p, vp]
        L.igap4h_residual_vjp_coef_dev.argtypes = [vp, vp, vp, vp, vp]
        L.igap4h_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                          ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
        L.igap4h_batch_destroy.argtypes = [vp]
        L.igap4h_batch_destroy.restype = None
        L.igap4h_batch_sizes.argtypes = [vp, c_i64_p, c_int_p]
        L.igap4h_batch_tune.argtypes = [vp, ctypes.c_int, c_int_p, c_double_p, c_int_p]
        L.igap4h_batch_set_variant.argtypes = [vp, ctypes.c_int]
        L.igap4h_batch_assemble.argtypes = [vp, vp]
        L.igap4h_update_bc.argtypes = [vp, ctypes.POINTER(BC)]
        L.igap4h_update_neumann.argtypes = [vp]
        L.igap4h_set_tangent_form.argtypes = [vp, ctypes.c_int]
        L.igap4h_write_solution.argtypes = [ctypes.c_char_p, c_double_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.igap4h_read_solution.argtypes = [ctypes.c_char_p, c_double_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.igap4h_knot_insert.argtypes = [c_double_p, c_int_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         c_int_p]
        L.igap4h_newton_default_opts.argtypes = [ctypes.POINTER(NewtonOpts)]
        L.igap4h_newton_solve_dev.argtypes = [vp, vp, ctypes.POINTER(NewtonOpts), ctypes.POINTER(NewtonInfo), vp]
        L.igap4h_spmv_dev.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp]
        L.igap4h_linear_solve_dev.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.POINTER(NewtonOpts), c_int_p,
                                              c_double_p, vp]
        for name in SYMBOLS:
            if name not in ("igap4h_destroy", "igap4h_last_error", "igap4h_batch_destroy"):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


class IgapError(RuntimeError):
    """A negative igap4h status; .code holds it (IGAP4H_EINVAL, ...)."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


def check(rc):
    if rc != 0:
        msg = lib().igap4h_last_error()
        raise IgapError("igap4h error %d: %s" % (rc, msg.decode() if msg else ""), rc)
    return rc
