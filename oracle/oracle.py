"""ctypes front end of the CPU oracle (liboracle.so) and of the reference's own
libm-only sources built into oracle/_ref (libigap4ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg; the product never imports this module.

A "problem" is any object with the attributes
    nelem (3 ints), knots (3 float64 arrays of nelem[d]+5), nquad, physics
    (0 ge / 1 getime), par_mat (9), elem_coef (None or (n_elem, 6)), boun,
    bc_type (36), par_dirichlet (9), par_periodic (9), par_neumann,
    neumann (None or a Python callable t(face, order, idof, xq) -> float)
which is what igap4standard_amd.problem.Problem provides.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libigap4ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
NEUMANN_FN = ctypes.CFUNCTYPE(None, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp)


class OrProblem(ctypes.Structure):
    _fields_ = [
        ("nelem", ctypes.c_int * 3),
        ("knot", _dp * 3),
        ("nquad", ctypes.c_int),
        ("physics", ctypes.c_int),
        ("par_mat", _dp),
        ("elem_coef", _dp),
        ("nboun", ctypes.c_int),
        ("boun", _ip),
        ("bc_type", _ip),
        ("par_dirichlet", _dp),
        ("par_periodic", _dp),
        ("par_neumann", _dp),
        ("neumann", NEUMANN_FN),
    ]


def _d(a):
    return a.ctypes.data_as(_dp)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.or_evalN.restype = ctypes.c_double
        L.or_ge_energy.restype = ctypes.c_double
        L.or_ndof.restype = ctypes.c_int64
        L.or_nnz.restype = ctypes.c_int64
        L.or_energy.restype = ctypes.c_double
        L.or_tangent_elements.argtypes = [ctypes.POINTER(OrProblem), _dp, _i64p, _i32p, _dp,
                                          ctypes.c_int64, ctypes.c_int64]
        L.or_residual_elements.argtypes = [ctypes.POINTER(OrProblem), _dp, _dp, ctypes.c_int64,
                                           ctypes.c_int64]
        L.or_evalN.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_double]
        L.or_residual_mt.argtypes = [ctypes.POINTER(OrProblem), _dp, _dp, ctypes.c_int]
        L.or_tangent_window.argtypes = [ctypes.POINTER(OrProblem), _dp, _ip, _ip, _dp]
        _lib = L
    return _lib


class Bound:
    """A problem bound to the C struct; keeps every referenced array alive."""

    def __init__(self, pb):
        self.pb = pb
        self.knots = [np.ascontiguousarray(k, dtype=np.float64) for k in pb.knots]
        self.par_mat = np.ascontiguousarray(pb.par_mat, dtype=np.float64)
        self.elem_coef = None if pb.elem_coef is None else np.ascontiguousarray(pb.elem_coef, dtype=np.float64)
        self.boun = np.ascontiguousarray(pb.boun, dtype=np.int32)
        self.bc_type = np.ascontiguousarray(pb.bc_type, dtype=np.int32)
        self.par_dirichlet = np.ascontiguousarray(pb.par_dirichlet, dtype=np.float64)
        self.par_periodic = np.ascontiguousarray(pb.par_periodic, dtype=np.float64)
        self.par_neumann = np.ascontiguousarray(
            pb.par_neumann if pb.par_neumann is not None else np.zeros(1), dtype=np.float64)
        s = OrProblem()
        for d in range(3):
            s.nelem[d] = int(pb.nelem[d])
            s.knot[d] = _d(self.knots[d])
        s.nquad = int(pb.nquad)
        s.physics = int(pb.physics)
        s.par_mat = _d(self.par_mat)
        s.elem_coef = _d(self.elem_coef) if self.elem_coef is not None else _dp()
        s.nboun = len(self.boun)
        s.boun = self.boun.ctypes.data_as(_ip)
        s.bc_type = self.bc_type.ctypes.data_as(_ip)
        s.par_dirichlet = _d(self.par_dirichlet)
        s.par_periodic = _d(self.par_periodic)
        s.par_neumann = _d(self.par_neumann)
        if pb.neumann is not None:
            fn = pb.neumann

            def cb(t, par, face, order, idof, xq):
                t[0] = float(fn(face, order, idof, (xq[0], xq[1], xq[2])))

            self._cb = NEUMANN_FN(cb)
            s.neumann = self._cb
        else:
            s.neumann = NEUMANN_FN()
        self.s = s

    @property
    def ptr(self):
        return ctypes.byref(self.s)


def ndof(pb):
    return int(lib().or_ndof(Bound(pb).ptr))


def nnz(pb):
    return int(lib().or_nnz(Bound(pb).ptr))


def pattern(pb):
    b = Bound(pb)
    n, z = lib().or_ndof(b.ptr), lib().or_nnz(b.ptr)
    rowptr = np.zeros(n + 1, np.int64)
    col = np.zeros(z, np.int32)
    lib().or_pattern(b.ptr, rowptr.ctypes.data_as(_i64p), col.ctypes.data_as(_i32p))
    return rowptr, col


def dirichlet_rows(pb):
    b = Bound(pb)
    n = lib().or_dirichlet_count(b.ptr)
    rows = np.zeros(max(n, 1), np.int32)
    m = lib().or_dirichlet_rows(b.ptr, rows.ctypes.data_as(_i32p))
    assert m == n
    return rows[:n]


def residual(pb, U_hist):
    b = Bound(pb)
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    R = np.zeros(lib().or_ndof(b.ptr))
    rc = lib().or_residual(b.ptr, _d(U), _d(R))
    if rc:
        raise RuntimeError("or_residual failed: %d" % rc)
    return R


def residual_mt(pb, U_hist, nthreads=None):
    """residual() with the element loop on nthreads threads (default: the
    CPUs this process may use, at most 16)."""
    if nthreads is None:
        nthreads = min(16, len(os.sched_getaffinity(0)))
    b = Bound(pb)
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    R = np.zeros(lib().or_ndof(b.ptr))
    rc = lib().or_residual_mt(b.ptr, _d(U), _d(R), int(nthreads))
    if rc:
        raise RuntimeError("or_residual_mt failed: %d" % rc)
    return R


def window_nodes(pb, lo, hi):
    """(row node ranges, column node ranges) of tangent_window: per
    direction the nodes whose support lies in the element window [lo, hi)
    and the window's nodes [lo, hi + 1]."""
    rows, cols = [], []
    for d in range(3):
        rlo = 0 if lo[d] == 0 else lo[d] + 2
        rhi = pb.nelem[d] + 1 if hi[d] == pb.nelem[d] else hi[d] - 1
        rows.append(np.arange(rlo, rhi + 1))
        cols.append(np.arange(lo[d], hi[d] + 2))
    return rows, cols


def tangent_window(pb, U_hist, lo, hi):
    """Element sums of the tangent over the element window [lo, hi) per
    direction (no Dirichlet step) on the complete rows of that window, as a
    dense (3 * rows, 3 * cols) block (node-major, dof fastest; see
    window_nodes)."""
    b = Bound(pb)
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    rows, cols = window_nodes(pb, lo, hi)
    nr = 3 * int(np.prod([len(r) for r in rows]))
    nc = 3 * int(np.prod([len(c) for c in cols]))
    out = np.zeros((nr, nc))
    lo_a = np.ascontiguousarray(lo, dtype=np.int32)
    hi_a = np.ascontiguousarray(hi, dtype=np.int32)
    rc = lib().or_tangent_window(b.ptr, _d(U), lo_a.ctypes.data_as(_ip), hi_a.ctypes.data_as(_ip), _d(out))
    if rc:
        raise RuntimeError("or_tangent_window failed: %d" % rc)
    return out


def tangent(pb, U_hist, rowptr=None, col=None):
    b = Bound(pb)
    if rowptr is None:
        rowptr, col = pattern(pb)
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    vals = np.zeros(len(col))
    rc = lib().or_tangent(b.ptr, _d(U), rowptr.ctypes.data_as(_i64p), col.ctypes.data_as(_i32p), _d(vals))
    if rc:
        raise RuntimeError("or_tangent failed: %d" % rc)
    return rowptr, col, vals


def energy(pb, U_hist):
    b = Bound(pb)
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    return float(lib().or_energy(b.ptr, _d(U)))


def density(physics, u, coef, tp=None):
    """Density outputs at one point: 13 (ge, u of 30 slots) or 2 (getime, u
    of 90 slots, tp = (rho, c, dt))."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    if physics == 1:
        h = np.zeros(2)
        lib().or_getime_density(_d(h), _d(u), _d(c), _d(np.ascontiguousarray(tp, dtype=np.float64)))
    else:
        h = np.zeros(13)
        lib().or_ge_density(_d(h), _d(u), _d(c))
    return h


def intdensity(pb, U_hist):
    """iga_intdensity: the ndensity integrated outputs (Neumann terms included)."""
    b = Bound(pb)
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    HH = np.zeros(13)
    nd = lib().or_intdensity(b.ptr, _d(U), _d(HH))
    if nd < 0:
        raise RuntimeError("or_intdensity failed: %d" % nd)
    return HH[:nd]


def gauss(n):
    x, w = np.zeros(n), np.zeros(n)
    if lib().or_gauss(n, _d(x), _d(w)):
        raise ValueError(n)
    return x, w


def eval_bspline(windows, xq):
    """windows: three 6-knot arrays; xq: 3 coords -> (27, 10)."""
    ws = [np.ascontiguousarray(w, dtype=np.float64) for w in windows]
    arr = (_dp * 3)(*[_d(w) for w in ws])
    N = np.zeros(270)
    x = np.ascontiguousarray(xq, dtype=np.float64)
    lib().or_eval_bspline(_d(N), arr, _d(x))
    return N.reshape(27, 10)


def evalN(kv, k, i, p, xi):
    kv = np.ascontiguousarray(kv, dtype=np.float64)
    return float(lib().or_evalN(_d(kv), len(kv), k, i, p, xi))


def coef_from_par(par):
    c = np.zeros(6)
    lib().or_coef_from_par(_d(np.ascontiguousarray(par, dtype=np.float64)), _d(c))
    return c


def ge_energy(u, coef):
    return float(lib().or_ge_energy(_d(np.ascontiguousarray(u, dtype=np.float64)),
                                    _d(np.ascontiguousarray(coef, dtype=np.float64))))


def ge_residual(u, coef):
    r = np.zeros(30)
    lib().or_ge_residual(_d(r), _d(np.ascontiguousarray(u, dtype=np.float64)),
                         _d(np.ascontiguousarray(coef, dtype=np.float64)))
    return r


def ge_tangent(u, coef):
    D = np.zeros(900)
    lib().or_ge_tangent(_d(D), _d(np.ascontiguousarray(u, dtype=np.float64)),
                        _d(np.ascontiguousarray(coef, dtype=np.float64)))
    return D.reshape(30, 30)


def ge_energy_coef_grad(u):
    g = np.zeros(6)
    lib().or_ge_energy_coef_grad(_d(g), _d(np.ascontiguousarray(u, dtype=np.float64)))
    return g


def getime_residual(u90, coef, tp):
    r = np.zeros(30)
    lib().or_getime_residual(_d(r), _d(np.ascontiguousarray(u90, dtype=np.float64)),
                             _d(np.ascontiguousarray(coef, dtype=np.float64)),
                             _d(np.ascontiguousarray(tp, dtype=np.float64)))
    return r


def getime_tangent(u90, coef, tp):
    D = np.zeros(900)
    lib().or_getime_tangent(_d(D), _d(np.ascontiguousarray(u90, dtype=np.float64)),
                            _d(np.ascontiguousarray(coef, dtype=np.float64)),
                            _d(np.ascontiguousarray(tp, dtype=np.float64)))
    return D.reshape(30, 30)


def iga_dirichlet(par, face_id, bc_order, idof, kv3, nb3):
    kvs = [np.ascontiguousarray(k, dtype=np.float64) for k in kv3]
    arr = (_dp * 3)(*[_d(k) for k in kvs])
    nk = (ctypes.c_int * 3)(*[len(k) for k in kvs])
    nb = (ctypes.c_int * 3)(*nb3)
    out = np.zeros(nb3[0] * nb3[1])
    lib().or_iga_dirichlet(_d(out), _d(np.ascontiguousarray(par, dtype=np.float64)), face_id, bc_order,
                           idof, arr, nk, nb)
    return out


def tangent_elements_timed(pb, U_hist, rowptr, col, vals, e_begin, e_end):
    """CPU baseline: the reference's per-element tangent loop over a range."""
    b = Bound(pb)
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    rc = lib().or_tangent_elements(b.ptr, _d(U), rowptr.ctypes.data_as(_i64p), col.ctypes.data_as(_i32p),
                                   _d(vals), e_begin, e_end)
    if rc:
        raise RuntimeError("or_tangent_elements failed: %d" % rc)


def residual_elements_timed(pb, U_hist, R, e_begin, e_end):
    b = Bound(pb)
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    rc = lib().or_residual_elements(b.ptr, _d(U), _d(R), e_begin, e_end)
    if rc:
        raise RuntimeError("or_residual_elements failed: %d" % rc)


# --------------------------------------------------------------------------
# The reference's own libm-only sources (oracle/_ref), for pinning the oracle.
# --------------------------------------------------------------------------
_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_PATH)
        L.evalN.restype = ctypes.c_double
        L.evalN.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        _ref = L
    return _ref


def ref_ge_residual(u, par):
    r = np.zeros(30)
    ref().ge_residual(_d(r), _d(np.ascontiguousarray(u, dtype=np.float64)),
                      _d(np.ascontiguousarray(par, dtype=np.float64)))
    return r


def ref_ge_tangent(u, par):
    D = np.zeros(900)
    ref().ge_tangent(_d(D), _d(np.ascontiguousarray(u, dtype=np.float64)),
                     _d(np.ascontiguousarray(par, dtype=np.float64)))
    return D.reshape(30, 30)


def ref_ge_energy(u, par):
    h = np.zeros(13)
    ref().ge_density(_d(h), _d(np.ascontiguousarray(u, dtype=np.float64)),
                     _d(np.ascontiguousarray(par, dtype=np.float64)))
    return h[0]


def ref_ge_density(u, par):
    """phys/ge/density.c:6-66, all 13 outputs."""
    h = np.zeros(13)
    ref().ge_density(_d(h), _d(np.ascontiguousarray(u, dtype=np.float64)),
                     _d(np.ascontiguousarray(par, dtype=np.float64)))
    return h


def ref_getime_density(u90, par):
    """phys/getime/density.c:6-53, both outputs."""
    h = np.zeros(2)
    ref().getime_density(_d(h), _d(np.ascontiguousarray(u90, dtype=np.float64)),
                         _d(np.ascontiguousarray(par, dtype=np.float64)))
    return h


def ref_densityb(physics, ub, traction, par, order):
    """phys/ge/density.c:68-84 (13 outputs) or phys/getime/density.c:57-62
    (2 outputs) at the face values ub (one per history slice)."""
    n = 2 if physics == 1 else 13
    h = np.zeros(n)
    fn = ref().getime_densityb if physics == 1 else ref().ge_densityb
    fn(_d(h), _d(np.ascontiguousarray(ub, dtype=np.float64)), ctypes.c_double(traction),
       _d(np.ascontiguousarray(par, dtype=np.float64)), ctypes.c_int(order))
    return h


def ref_getime_energy(u90, par):
    """phys/getime/density.c:6-50 (hh[0])."""
    h = np.zeros(2)
    ref().getime_density(_d(h), _d(np.ascontiguousarray(u90, dtype=np.float64)),
                         _d(np.ascontiguousarray(par, dtype=np.float64)))
    return float(h[0])


def ref_getime_residual(u90, par):
    r = np.zeros(30)
    ref().getime_residual(_d(r), _d(np.ascontiguousarray(u90, dtype=np.float64)),
                          _d(np.ascontiguousarray(par, dtype=np.float64)))
    return r


def ref_getime_tangent(u90, par):
    D = np.zeros(900)
    ref().getime_tangent(_d(D), _d(np.ascontiguousarray(u90, dtype=np.float64)),
                         _d(np.ascontiguousarray(par, dtype=np.float64)))
    return D.reshape(30, 30)


def ref_eval_bspline(windows, xq):
    ws = [np.ascontiguousarray(w, dtype=np.float64) for w in windows]
    arr = (_dp * 3)(*[_d(w) for w in ws])
    N = np.zeros(270)
    ref().eval_Bspline(_d(N), arr, 2, _d(np.ascontiguousarray(xq, dtype=np.float64)))
    return N.reshape(27, 10)


def ref_evalN(kv, k, i, p, xi):
    kv = np.ascontiguousarray(kv, dtype=np.float64)
    return float(ref().evalN(_d(kv), len(kv), k, i, p, xi))


def ref_gauss(n):
    x, w = np.zeros(n), np.zeros(n)
    ref().legendre_handle(_d(x), _d(w), ctypes.c_int(n), ctypes.c_double(0.0), ctypes.c_double(1.0))
    return x, w


def ref_iga_dirichlet(par, face_id, bc_order, idof, kv3, nb3):
    kvs = [np.ascontiguousarray(k, dtype=np.float64) for k in kv3]
    arr = (_dp * 3)(*[_d(k) for k in kvs])
    nk = (ctypes.c_int * 3)(*[len(k) for k in kvs])
    nb = (ctypes.c_int * 3)(*nb3)
    out = np.zeros(nb3[0] * nb3[1])
    ref().iga_dirichlet(_d(out), _d(np.ascontiguousarray(par, dtype=np.float64)), face_id, bc_order, idof,
                        arr, nk, nb, 3, 2)
    return out


def knotinsert(U_universal, nelem, nref, ndof=3):
    """Restatement of mesh/nurbs/nurbs_knotinsert.c:7-107 (porder 2,
    non-periodic): the coarse control values (nelem / 2^nref) refined nref
    times with the per-index masks of :41-45, scattered to fine indices
    2 i - 2 + j (:87-91).  Returns the refined array (fine size)."""
    ne = [n >> nref for n in nelem]
    nb = [n + 2 for n in ne]
    U = np.asarray(U_universal, dtype=np.float64)[:ndof * nb[0] * nb[1] * nb[2]].reshape(nb[0], nb[1], nb[2], ndof)

    def mask(i, n):   # n = coarse element count
        if i == 0:
            return (0.0, 0.0, 1.0, 0.5)
        if i == 1:
            return (0.0, 0.5, 0.75, 0.25)
        if i < n:
            return (0.25, 0.75, 0.75, 0.25)
        if i == n:
            return (0.25, 0.75, 0.5, 0.0)
        return (0.5, 1.0, 0.0, 0.0)

    for _ in range(nref):
        nbn = [2 * n + 2 for n in ne]
        out = np.zeros((nbn[0], nbn[1], nbn[2], ndof))
        for ix in range(nb[0]):
            vx = mask(ix, ne[0])
            for iy in range(nb[1]):
                vy = mask(iy, ne[1])
                for iz in range(nb[2]):
                    vz = mask(iz, ne[2])
                    for jx in range(4):
                        for jy in range(4):
                            for jz in range(4):
                                w = vx[jx] * vy[jy] * vz[jz]
                                if w != 0.0:
                                    out[2 * ix - 2 + jx, 2 * iy - 2 + jy, 2 * iz - 2 + jz] += U[ix, iy, iz] * w
        U = out
        ne = [2 * n for n in ne]
        nb = [n + 2 for n in ne]
    return U.reshape(-1)

