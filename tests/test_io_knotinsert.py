"""Host utilities of the C-ABI (no GPU): solution files (src/ioutil.c:89-158)
and knot-insertion refinement (mesh/nurbs/nurbs_knotinsert.c:7-107).

Pins: read(write(U)) is bitwise U and the file is the raw slices; the
library's refinement equals the oracle's restatement of the reference masks,
and both leave the spline unchanged (coarse and fine tensor-product splines
agree at random points, evaluated with scipy's independent B-spline design
matrices)."""
import numpy as np
import pytest
from scipy.interpolate import BSpline

from igap4standard_amd import _lib as L
from igap4standard_amd import problem as P
from igap4standard_amd.ioutil import nurbs_knotinsert, read_solution, write_solution


def test_solution_file_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    ndof = 3 * 5 * 6 * 7
    U = rng.standard_normal(3 * ndof)
    f = tmp_path / "sol.bin"
    write_solution(f, U, ndof, 1, 3)
    raw = np.fromfile(f, dtype=np.float64)
    assert np.array_equal(raw, U[ndof:3 * ndof])
    V = np.zeros(3 * ndof)
    read_solution(f, V, ndof, 1, 3)
    assert np.array_equal(V[ndof:], U[ndof:]) and not V[:ndof].any()
    with pytest.raises(L.IgapError):
        read_solution(f, V, ndof, 0, 3)          # file too short
    with pytest.raises(L.IgapError):
        read_solution(f, V, ndof - 1, 1, 3)      # file too long for the request


def spline_values(U, ne, pts, ndof=3):
    knots = [P.uniform_knots(n) for n in ne]
    B = [BSpline.design_matrix(pts[:, d], knots[d], 2).toarray() for d in range(3)]
    C = U.reshape(ne[0] + 2, ne[1] + 2, ne[2] + 2, ndof)
    return np.einsum("pa,pb,pc,abci->pi", B[0], B[1], B[2], C)


@pytest.mark.parametrize("nelem,nref", [((4, 4, 4), 1), ((8, 4, 16), 2), ((6, 12, 6), 1)])
def test_knot_insert(oracle, nelem, nref):
    rng = np.random.default_rng(1)
    ne_c = [n >> nref for n in nelem]
    nc = 3 * int(np.prod([n + 2 for n in ne_c]))
    nf = 3 * int(np.prod([n + 2 for n in nelem]))
    Uc = rng.standard_normal(nc)
    U = np.zeros(nf)
    U[:nc] = Uc
    nurbs_knotinsert(U, nelem, nref)
    ref = oracle.knotinsert(Uc, nelem, nref)
    pts = rng.uniform(0.0, 1.0, (200, 3))
    fc = spline_values(Uc, ne_c, pts)
    ff = spline_values(U, list(nelem), pts)
    fo = spline_values(ref, list(nelem), pts)
    assert np.max(np.abs(ff - fc)) < 1e-13          # the library refines exactly
    if min(ne_c) >= 2:
        assert np.max(np.abs(U - ref)) < 1e-14      # and equals the reference's masks
        assert np.max(np.abs(fo - fc)) < 1e-13
    else:
        # a coarse direction with one element: the reference takes the i == 1
        # branch (nurbs_knotinsert.c:42) before i == nelem_old (:44), and its
        # result no longer reproduces the coarse spline; the library does
        assert np.max(np.abs(fo - fc)) > 1e-3


def test_knot_insert_rejects():
    U = np.zeros(3 * 6 ** 3)
    with pytest.raises(L.IgapError):
        nurbs_knotinsert(U, (4, 4, 4), 3)         # 4 is not a multiple of 8
    with pytest.raises(L.IgapError):
        nurbs_knotinsert(U, (4, 4, 4), 1, bc_periodic=(1, 0, 0))
