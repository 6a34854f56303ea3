"""Solution files and knot-insertion refinement through the C-ABI, with the
reference's names and argument meaning (src/ioutil.c:89-158,
mesh/nurbs/nurbs_knotinsert.c:7-107; one rank, so no IOcomm)."""
import ctypes

import numpy as np

from . import _lib as L


def write_solution(fname, U_hist, ndof, ihist0, ihist1):
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    L.check(L.lib().igap4h_write_solution(str(fname).encode(), U.ctypes.data_as(L.c_double_p), int(ndof), ihist0, ihist1))


def read_solution(fname, U_hist, ndof, ihist0, ihist1):
    """Fills U_hist[ihist0*ndof : ihist1*ndof] (a contiguous float64 array) in place."""
    if not (isinstance(U_hist, np.ndarray) and U_hist.dtype == np.float64 and U_hist.flags.c_contiguous):
        raise TypeError("U_hist must be a contiguous float64 array")
    L.check(L.lib().igap4h_read_solution(str(fname).encode(), U_hist.ctypes.data_as(L.c_double_p), int(ndof), ihist0, ihist1))


def nurbs_knotinsert(U_universal, nelem, nref, porder=2, ndim=3, ndof=3, bc_periodic=(0, 0, 0)):
    """In place: the coarse solution (nelem / 2^nref per direction) in the
    leading part of U_universal becomes the refined one (nelem)."""
    if not (isinstance(U_universal, np.ndarray) and U_universal.dtype == np.float64 and U_universal.flags.c_contiguous):
        raise TypeError("U_universal must be a contiguous float64 array")
    need = ndof * int(np.prod([n + 2 for n in nelem]))
    if U_universal.size < need:
        raise ValueError("U_universal holds %d values, the refined mesh needs %d" % (U_universal.size, need))
    ne = (ctypes.c_int * 3)(*nelem)
    per = (ctypes.c_int * 3)(*bc_periodic)
    L.check(L.lib().igap4h_knot_insert(U_universal.ctypes.data_as(L.c_double_p), ne, nref, porder, ndim, ndof, per))
