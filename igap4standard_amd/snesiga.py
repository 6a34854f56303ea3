"""Host-side mirror of the reference's assembly interface on top of the
C-ABI library (libigap4h.so):

    snesiga_init(pb)                    src/snesiga_init.c:12-132
    snesiga_Residual(snes, U, R, app)   src/snesiga_Residual.c:13-417
    snesiga_Tangent(snes, U, J, P, app) src/snesiga_Tangent.c:13-284

Same argument meaning: U is the current iterate, R is overwritten, P (a CSR
container on the AIJ pattern) is zeroed and filled, and app.U_hist[0:ndof]
is set to U as a side effect.  Errors raise IgapError (the reference
asserts/exits); the return value is 0 like the PETSc callbacks.

App also exposes the device-resident form (HBM pointers, HIP stream) used by
bench.py and the GPU tests: residual_dev / tangent_dev / pattern_dev.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .problem import GE, GETIME, Problem


def _dp(a):
    return a.ctypes.data_as(L.c_double_p)


def _ip(a):
    return a.ctypes.data_as(L.c_int_p)


@dataclass
class CSR:
    """AIJ values on the fixed pattern (MatCreateAIJ + KEEP_NONZERO_PATTERN)."""
    rowptr: np.ndarray
    col: np.ndarray
    vals: np.ndarray

    def to_dense(self):
        n = len(self.rowptr) - 1
        K = np.zeros((n, n), dtype=self.vals.dtype)
        for r in range(n):
            K[r, self.col[self.rowptr[r]:self.rowptr[r + 1]]] = self.vals[self.rowptr[r]:self.rowptr[r + 1]]
        return K

    def matvec(self, x):
        rows = np.repeat(np.arange(len(self.rowptr) - 1), np.diff(self.rowptr))
        return np.bincount(rows, weights=self.vals.astype(np.float64) * x[self.col], minlength=len(self.rowptr) - 1)


class App:
    """IGA application context (include/iga.h:8-25) bound to a device context.

    rank: a partition.RankBlock to make the context of that rank's block of
    the partitioned patch pb (its local elements, knot window, owned basis
    functions, boundary faces, selfIdx/recvIdx/globalIdx, the coefficients of
    its elements); its device vectors are in the local ghosted numbering."""

    def __init__(self, pb: Problem, physics=None, rank=None):
        self.pb = pb
        self.rank = rank
        lib = L.lib()
        nh = pb.nhist
        nb = pb.nbasis if rank is None else rank.nbasis
        # --- Phys
        self._par = np.ascontiguousarray(pb.par_mat, dtype=np.float64)
        self.phys = L.Phys(ndim=3, ndof=3, torder=pb.torder, sorder=2, par_mat=_dp(self._par),
                           ndensity=1, nfield=0)
        # --- Nurbs (non-periodic; one rank owning the patch, or one rank's block)
        self._nelem = np.array(pb.nelem if rank is None else rank.nelem, dtype=np.int32)
        self._nbasis = np.array(nb, dtype=np.int32)
        self._knots = [np.ascontiguousarray(k, dtype=np.float64) for k in (pb.knots if rank is None else rank.knots)]
        self._nknot = np.array([len(k) for k in self._knots], dtype=np.int32)
        self._kv = (L.c_double_p * 3)(*[_dp(k) for k in self._knots])
        self._boun = np.array(pb.boun if rank is None else [f for f in rank.boun if f in pb.boun], dtype=np.int32)
        self.nurbs = L.Nurbs(porder=2, nelem=_ip(self._nelem), nbasis=_ip(self._nbasis), nknot=_ip(self._nknot),
                             knotVector=ctypes.cast(self._kv, ctypes.POINTER(L.c_double_p)),
                             nboun=len(self._boun), boun=_ip(self._boun))
        # --- BC
        self._bc_type = np.ascontiguousarray(pb.bc_type, dtype=np.int32)
        self._pp = np.ascontiguousarray(pb.par_periodic, dtype=np.float64)
        self._pd = np.ascontiguousarray(pb.par_dirichlet, dtype=np.float64)
        self._pn = np.ascontiguousarray(pb.par_neumann if pb.par_neumann is not None else np.zeros(1),
                                        dtype=np.float64)
        self.bc = L.BC(bc_type=_ip(self._bc_type), par_periodic=_dp(self._pp), par_dirichlet=_dp(self._pd),
                       par_neumann=_dp(self._pn))
        # --- IGA
        if rank is None:
            self.ndof = pb.ndof
            self.U_hist = np.zeros(nh * self.ndof)
            self._gidx = np.arange(self.ndof, dtype=np.int32)
            self.iga = L.IGA(phys=ctypes.pointer(self.phys), nurbs=ctypes.pointer(self.nurbs),
                             bc=ctypes.pointer(self.bc), U_hist=_dp(self.U_hist), nselfIdx=self.ndof,
                             globalIdx=_ip(self._gidx))
        else:
            self.ndof = rank.nlocal                          # the local ghosted window
            self.U_hist = np.zeros(nh * rank.nself)
            self._gidx = np.ascontiguousarray(rank.globalIdx, dtype=np.int32)
            self._self = np.ascontiguousarray(rank.selfIdx, dtype=np.int32)
            self._send = np.ascontiguousarray(rank.sendIdx, dtype=np.int32)
            self._sptr = np.ascontiguousarray(rank.sendPtr, dtype=np.int32)
            self._recv = np.ascontiguousarray(rank.recvIdx, dtype=np.int32)
            self._rptr = np.ascontiguousarray(rank.recvPtr, dtype=np.int32)
            self.iga = L.IGA(phys=ctypes.pointer(self.phys), nurbs=ctypes.pointer(self.nurbs),
                             bc=ctypes.pointer(self.bc), U_hist=_dp(self.U_hist), nselfIdx=rank.nself,
                             selfIdx=_ip(self._self), nsendIdx=len(self._send), sendIdx=_ip(self._send),
                             sendPtr=_ip(self._sptr), nrecvIdx=len(self._recv), recvIdx=_ip(self._recv),
                             recvPtr=_ip(self._rptr), globalIdx=_ip(self._gidx))
        # --- options
        coef = pb.elem_coef
        if coef is not None and rank is not None:       # the rank's elements, in its local order
            ne = pb.nelem
            c3 = np.asarray(coef, dtype=np.float64).reshape(ne[0], ne[1], ne[2], 6)
            d, n = rank.displ, rank.nelem
            coef = c3[d[0]:d[0] + n[0], d[1]:d[1] + n[1], d[2]:d[2] + n[2]].reshape(-1, 6)
        self._coef = None if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
        if pb.neumann is not None:
            fn = pb.neumann

            def cb(t, par, face, order, idof, xq):
                t[0] = float(fn(face, order, idof, (xq[0], xq[1], xq[2])))

            self._cb = L.NEUMANN_FN(cb)
        else:
            self._cb = L.NEUMANN_FN()
        phys_id = {GE: L.IGAP4H_PHYS_GE, GETIME: L.IGAP4H_PHYS_GETIME}[pb.physics] if physics is None else physics
        self.opts = L.Opts(physics=phys_id, nquad=pb.nquad,
                           elem_coef=_dp(self._coef) if self._coef is not None else L.c_double_p(),
                           neumann=self._cb)
        ctx = ctypes.c_void_p()
        L.check(lib.igap4h_create(ctypes.byref(self.iga), ctypes.byref(self.opts), ctypes.byref(ctx)))
        self.ctx = ctx
        a, b, c, d, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        L.check(lib.igap4h_sizes(ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d),
                                 ctypes.byref(e)))
        self.nnz, self.ndirichlet, self.nelem, self.nhist = b.value, c.value, d.value, e.value
        assert a.value == self.ndof

    def close(self):
        if getattr(self, "ctx", None):
            L.lib().igap4h_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- index sets
    def pattern(self):
        rowptr = np.zeros(self.ndof + 1, np.int64)
        col = np.zeros(self.nnz, np.int32)
        L.check(L.lib().igap4h_pattern(self.ctx, rowptr.ctypes.data_as(L.c_i64_p), col.ctypes.data_as(L.c_i32_p)))
        return rowptr, col

    def dirichlet_rows(self):
        rows = np.zeros(max(self.ndirichlet, 1), np.int32)
        L.check(L.lib().igap4h_dirichlet_rows(self.ctx, rows.ctypes.data_as(L.c_i32_p)))
        return rows[:self.ndirichlet]

    def pattern_dev(self):
        r, c = ctypes.c_void_p(), ctypes.c_void_p()
        L.check(L.lib().igap4h_pattern_dev(self.ctx, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    def update_bc(self):
        """Push pb.bc_type / pb.par_dirichlet (e.g. a new load step) to the context."""
        self._bc_type[:] = np.asarray(self.pb.bc_type, dtype=np.int32)
        self._pd[:] = np.asarray(self.pb.par_dirichlet, dtype=np.float64)
        L.check(L.lib().igap4h_update_bc(self.ctx, ctypes.byref(self.bc)))
        nd = ctypes.c_int64()
        L.check(L.lib().igap4h_sizes(self.ctx, None, None, ctypes.byref(nd), None, None))
        self.ndirichlet = nd.value

    def update_neumann(self):
        """Evaluate the traction hook again (its values changed)."""
        L.check(L.lib().igap4h_update_neumann(self.ctx))

    def set_coef(self, coef):
        self._coef = np.ascontiguousarray(coef, dtype=np.float64)
        L.check(L.lib().igap4h_set_coef(self.ctx, _dp(self._coef)))

    def set_tangent_form(self, form):
        """IGAP4H_TANGENT_AUTO / _TSTRIP / _SV (include/igap4h.h)."""
        L.check(L.lib().igap4h_set_tangent_form(self.ctx, form))

    # ---- host-array assembly
    def residual(self, U, R=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if R is None:
            R = np.zeros(self.ndof)
        L.check(L.lib().igap4h_residual(self.ctx, _dp(U), _dp(self.U_hist), _dp(R)))
        return R

    def tangent(self, U, vals=None, prec=L.IGAP4H_FP64, vals_type=L.IGAP4H_VALS_F64):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if vals is None:
            vals = np.zeros(self.nnz, np.float64 if vals_type == L.IGAP4H_VALS_F64 else np.float32)
        L.check(L.lib().igap4h_tangent(self.ctx, _dp(U), _dp(self.U_hist), prec, vals_type,
                                       vals.ctypes.data_as(ctypes.c_void_p)))
        return vals

    # ---- device-array assembly (pointers are integers, e.g. tensor.data_ptr())
    def residual_dev(self, d_Uh, d_R, stream=None):
        L.check(L.lib().igap4h_residual_dev(self.ctx, ctypes.c_void_p(d_Uh), ctypes.c_void_p(d_R),
                                            ctypes.c_void_p(stream or 0)))

    def residual_dev_noalloc(self, d_Uh, d_R, stream=None):
        """residual_dev staged through the library's fixed device array (no
        allocation)."""
        L.check(L.lib().igap4h_residual_dev_noalloc(self.ctx, ctypes.c_void_p(d_Uh), ctypes.c_void_p(d_R),
                                                    ctypes.c_void_p(stream or 0)))

    def tangent_dev(self, d_Uh, d_vals, prec=L.IGAP4H_FP64, vals_type=L.IGAP4H_VALS_F64, stream=None):
        L.check(L.lib().igap4h_tangent_dev(self.ctx, ctypes.c_void_p(d_Uh), prec, vals_type,
                                           ctypes.c_void_p(d_vals), ctypes.c_void_p(stream or 0)))

    # ---- Newton-Krylov driver (device arrays)
    @staticmethod
    def newton_opts(**kw):
        """igap4h_newton_opts with the library defaults (reference SNES/KSP
        settings, src/snes_init.c:18-39) overridden by keyword."""
        o = L.NewtonOpts()
        L.check(L.lib().igap4h_newton_default_opts(ctypes.byref(o)))
        for k, v in kw.items():
            if not hasattr(o, k):
                raise TypeError("unknown Newton option %r" % k)
            setattr(o, k, v)
        return o

    def newton_solve_dev(self, d_Uh, opts=None, stream=None, **kw):
        """Solve R(U) = 0 from slice 0 of the device array d_Uh (updated in
        place).  Returns the igap4h_newton_info fields as a dict."""
        o = opts if opts is not None else self.newton_opts(**kw)
        info = L.NewtonInfo()
        L.check(L.lib().igap4h_newton_solve_dev(self.ctx, ctypes.c_void_p(d_Uh), ctypes.byref(o),
                                                ctypes.byref(info), ctypes.c_void_p(stream or 0)))
        return info.as_dict()

    def spmv_dev(self, d_vals, vals_type, d_x, d_y, stream=None):
        L.check(L.lib().igap4h_spmv_dev(self.ctx, ctypes.c_void_p(d_vals), vals_type, ctypes.c_void_p(d_x),
                                        ctypes.c_void_p(d_y), ctypes.c_void_p(stream or 0)))

    def linear_solve_dev(self, d_vals, vals_type, d_b, d_x, opts=None, stream=None, **kw):
        """K x = b on the device; returns (iterations, relative residual)."""
        o = opts if opts is not None else self.newton_opts(**kw)
        its = ctypes.c_int()
        rel = ctypes.c_double()
        L.check(L.lib().igap4h_linear_solve_dev(self.ctx, ctypes.c_void_p(d_vals), vals_type, ctypes.c_void_p(d_b),
                                                ctypes.c_void_p(d_x), ctypes.byref(o), ctypes.byref(its),
                                                ctypes.byref(rel), ctypes.c_void_p(stream or 0)))
        return its.value, rel.value

    def residual_vjp_coef_dev(self, d_U, d_lam, d_g, stream=None):
        """d_g (n_elem x 6) = lambda^T dR/dcoef at U (phys/ge)."""
        L.check(L.lib().igap4h_residual_vjp_coef_dev(self.ctx, ctypes.c_void_p(d_U), ctypes.c_void_p(d_lam),
                                                     ctypes.c_void_p(d_g), ctypes.c_void_p(stream or 0)))

    def intdensity_dev(self, d_Uh, stream=None):
        """iga_intdensity: the ndensity integrated density outputs (13 for
        phys/ge, 2 for phys/getime), Neumann boundary terms included."""
        HH = (ctypes.c_double * 13)()
        nd = ctypes.c_int()
        L.check(L.lib().igap4h_intdensity_dev(self.ctx, ctypes.c_void_p(d_Uh), HH, ctypes.byref(nd),
                                              ctypes.c_void_p(stream or 0)))
        return np.array(HH[:nd.value])

    def energy_dev(self, d_U, stream=None):
        e = ctypes.c_double()
        L.check(L.lib().igap4h_energy_dev(self.ctx, ctypes.c_void_p(d_U), ctypes.byref(e),
                                          ctypes.c_void_p(stream or 0)))
        return e.value


def snesiga_init(pb: Problem):
    """Builds the context, the AIJ pattern and the residual vector
    (src/snesiga_init.c:12-132)."""
    app = App(pb)
    rowptr, col = app.pattern()
    P = CSR(rowptr, col, np.zeros(app.nnz))
    R = np.zeros(app.ndof)
    return app, R, P


def snesiga_Residual(snes, U, R, app: App):
    """R <- residual at U (src/snesiga_Residual.c:13-417); app.U_hist[0:ndof] = U."""
    app.residual(U, R)
    return 0


def snesiga_Tangent(snes, U, J, P: CSR, app: App, prec=L.IGAP4H_FP64):
    """P.vals <- tangent at U on the AIJ pattern (src/snesiga_Tangent.c:13-284);
    J is not touched unless it is a different matrix, which is then given
    the same values (the reference only re-assembles J, :275-279)."""
    vt = L.IGAP4H_VALS_F32 if P.vals.dtype == np.float32 else L.IGAP4H_VALS_F64
    app.tangent(U, P.vals, prec=prec, vals_type=vt)
    if J is not None and J is not P:
        J.vals[:] = P.vals
    return 0
