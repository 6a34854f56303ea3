"""Parity of the HIP path (through the C-ABI) with the oracle.

Bar: integer index sets bit-exact; FP64 residual and tangent within 1e-12
relative (2-norm / Frobenius) at non-equilibrium states; FP32 tangent within
1e-5 relative."""
import sys

import numpy as np
import pytest

from igap4standard_amd import _lib as L
from igap4standard_amd import problem as P
from igap4standard_amd.snesiga import App, snesiga_init, snesiga_Residual, snesiga_Tangent

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def state(pb, seed=1, amp=0.02):
    rng = np.random.default_rng(seed)
    base = P.smooth_state(pb, amp=amp * 5, noise=0.0)
    return base + rng.uniform(-1, 1, base.size) * amp


@pytest.mark.parametrize("nelem", [(2, 2, 2), (3, 2, 4), (8, 8, 8)])
def test_pattern_bit_exact(oracle, nelem):
    say("pattern", nelem)
    pb = P.Problem(nelem=nelem, knots=[P.uniform_knots(n) for n in nelem], bc_type=P.bc_ex1())
    app = App(pb)
    rowptr, col = app.pattern()
    r0, c0 = oracle.pattern(pb)
    assert np.array_equal(rowptr, r0) and np.array_equal(col, c0)
    assert np.array_equal(app.dirichlet_rows(), oracle.dirichlet_rows(pb))


def test_dirichlet_rows_stretch(oracle):
    pb = P.stretch_problem(5)
    app = App(pb)
    assert np.array_equal(app.dirichlet_rows(), oracle.dirichlet_rows(pb))


@pytest.mark.parametrize("case", ["ex1_ref", "stretch_lame", "stretch_ref_q3", "nonuniform"])
def test_residual_ge(oracle, case):
    say("residual ge", case)
    if case == "ex1_ref":
        pb = P.reference_example(4)
    elif case == "stretch_lame":
        pb = P.stretch_problem((4, 3, 5), a=0.01, noise=0.1)
    elif case == "stretch_ref_q3":
        pb = P.stretch_problem(3, a=0.01, material="reference", nquad=3)
    else:
        pb = P.stretch_problem(3, a=0.01, noise=0.1)
        rng = np.random.default_rng(3)
        for d in range(3):
            k = pb.knots[d]
            inner = np.sort(rng.uniform(0.05, 0.95, len(k) - 6))
            pb.knots[d] = np.concatenate([[0, 0, 0], inner, [1, 1, 1]])
    U = state(pb)
    app = App(pb)
    R = app.residual(U)
    R0 = oracle.residual(pb, U)
    assert rel(R, R0) < TOL64, rel(R, R0)
    assert np.array_equal(app.U_hist[:app.ndof], U)


def test_residual_neumann_and_periodic_F0(oracle):
    pb = P.stretch_problem(3, a=0.01, noise=0.1)
    pb.par_periodic = np.array([0.001, 0.0, 0.0, 0.0, -0.002, 0.0, 0.0, 0.0, 0.0005])
    pb.neumann = lambda face, order, idof, xq: (0.1 + xq[0] * xq[1] - 0.3 * xq[2]) * (1 + idof) * (1 + order)
    U = state(pb)
    app = App(pb)
    R = app.residual(U)
    R0 = oracle.residual(pb, U)
    assert rel(R, R0) < TOL64, rel(R, R0)


@pytest.mark.parametrize("case", ["ex1_ref_2", "stretch_lame_3", "stretch_ref_q3"])
def test_tangent_ge(oracle, case):
    say("tangent ge", case)
    if case == "ex1_ref_2":
        pb = P.reference_example(2)
    elif case == "stretch_lame_3":
        pb = P.stretch_problem((3, 2, 3), a=0.01, noise=0.1)
    else:
        pb = P.stretch_problem(2, a=0.01, material="reference", nquad=3)
    U = state(pb)
    app, R, Pm = snesiga_init(pb)
    rowptr, col, v0 = oracle.tangent(pb, U, Pm.rowptr, Pm.col)
    assert snesiga_Tangent(None, U, None, Pm, app) == 0
    assert rel(Pm.vals, v0) < TOL64, rel(Pm.vals, v0)
    v32 = app.tangent(U, prec=L.IGAP4H_FP32)
    assert rel(v32, v0) < TOL32, rel(v32, v0)
    v32f = app.tangent(U, prec=L.IGAP4H_FP32, vals_type=L.IGAP4H_VALS_F32)
    assert v32f.dtype == np.float32 and rel(v32f.astype(np.float64), v0) < TOL32
    assert snesiga_Residual(None, U, R, app) == 0
    assert rel(R, oracle.residual(pb, U)) < TOL64


def test_getime(oracle):
    say("getime")
    pb = P.reference_example(2, physics=P.GETIME)
    pb.bc_type = P.bc_all_clamped()
    pb.bc_type[0:3] = 1   # face 0 free standard-order: exercises Neumann-free rows
    rng = np.random.default_rng(5)
    n = pb.ndof
    U0 = rng.uniform(-1, 1, n) * 0.02
    U1 = U0 + rng.uniform(-1, 1, n) * 0.002
    U = U1 + rng.uniform(-1, 1, n) * 0.002
    Uh = np.concatenate([U, U1, U0])
    app = App(pb)
    app.U_hist[:] = Uh
    R = app.residual(U)
    R0 = oracle.residual(pb, Uh)
    assert rel(R, R0) < TOL64, rel(R, R0)
    rowptr, col = app.pattern()
    _, _, v0 = oracle.tangent(pb, Uh, rowptr, col)
    v = app.tangent(U)
    assert rel(v, v0) < TOL64, rel(v, v0)
    v32 = app.tangent(U, prec=L.IGAP4H_FP32)
    assert rel(v32, v0) < TOL32, rel(v32, v0)


def test_getime_colours(oracle):
    """4^3 getime (several elements per colour launch), nq = 4: residual and
    the three tangent forms against the oracle."""
    say("getime 4^3")
    pb = P.reference_example(4, physics=P.GETIME)
    rng = np.random.default_rng(11)
    n = pb.ndof
    U0 = P.smooth_state(pb, amp=0.05, noise=0.0, nhist=1) + rng.uniform(-1, 1, n) * 0.01
    U1 = U0 + rng.uniform(-1, 1, n) * 0.002
    U = U1 + rng.uniform(-1, 1, n) * 0.002
    Uh = np.concatenate([U, U1, U0])
    app = App(pb)
    app.U_hist[:] = Uh
    assert rel(app.residual(U), oracle.residual(pb, Uh)) < TOL64
    rowptr, col = app.pattern()
    _, _, v0 = oracle.tangent(pb, Uh, rowptr, col)
    assert rel(app.tangent(U), v0) < TOL64, rel(app.tangent(U), v0)
    v32 = app.tangent(U, prec=L.IGAP4H_FP32)
    assert rel(v32, v0) < TOL32, rel(v32, v0)
    v32f = app.tangent(U, prec=L.IGAP4H_FP32, vals_type=L.IGAP4H_VALS_F32)
    assert rel(v32f.astype(np.float64), v0) < TOL32


def test_energy_gradient_on_gpu(oracle):
    """R = dPi/dU on interior rows, Pi from the device energy kernel."""
    import torch
    pb = P.reference_example(3)
    pb.bc_type = np.ones(36, np.int32)
    U = state(pb, amp=0.01)
    app = App(pb)
    R = app.residual(U)
    dev = torch.device("cuda:0")
    h = 1e-6
    rng = np.random.default_rng(2)
    for k in rng.choice(pb.ndof, 6, replace=False):
        Up, Um = U.copy(), U.copy()
        Up[k] += h
        Um[k] -= h
        ep = app.energy_dev(torch.tensor(Up, device=dev).data_ptr())
        em = app.energy_dev(torch.tensor(Um, device=dev).data_ptr())
        torch.cuda.synchronize()
        assert abs((ep - em) / (2 * h) - R[k]) < 1e-6 * np.abs(R).max()
    assert abs(app.energy_dev(torch.tensor(U, device=dev).data_ptr()) - oracle.energy(pb, U)) < 1e-12 * abs(oracle.energy(pb, U))


@pytest.mark.parametrize("nelem,nquad", [((1, 3, 2), 4), ((2, 1, 1), 2), ((1, 1, 1), 1), ((3, 2, 1), 3)])
def test_ragged_meshes_and_quadrature(oracle, nelem, nquad):
    """One-element directions and every Gauss order 1..4: pattern, residual
    and both tangent forms against the oracle (chunks of Gauss points and
    colour grids degenerate here)."""
    say("ragged", nelem, nquad)
    pb = P.stretch_problem(nelem, a=0.01, noise=0.1, nquad=nquad)
    U = state(pb)
    app = App(pb)
    rowptr, col = app.pattern()
    r0, c0 = oracle.pattern(pb)
    assert np.array_equal(rowptr, r0) and np.array_equal(col, c0)
    R = app.residual(U)
    assert rel(R, oracle.residual(pb, U)) < TOL64
    _, _, v0 = oracle.tangent(pb, U, rowptr, col)
    assert rel(app.tangent(U), v0) < TOL64
    assert rel(app.tangent(U, prec=L.IGAP4H_FP32), v0) < TOL32


def test_getime_ragged_q2(oracle):
    pb = P.reference_example(2, physics=P.GETIME, nquad=2)
    pb.nelem = (2, 1, 2)
    pb.knots = [P.uniform_knots(n) for n in pb.nelem]
    rng = np.random.default_rng(7)
    n = pb.ndof
    U0 = rng.uniform(-1, 1, n) * 0.01
    U1 = U0 + rng.uniform(-1, 1, n) * 0.001
    U = U1 + rng.uniform(-1, 1, n) * 0.001
    Uh = np.concatenate([U, U1, U0])
    app = App(pb)
    app.U_hist[:] = Uh
    R = app.residual(U)
    assert rel(R, oracle.residual(pb, Uh)) < TOL64
    rowptr, col = app.pattern()
    _, _, v0 = oracle.tangent(pb, Uh, rowptr, col)
    assert rel(app.tangent(U), v0) < TOL64


@pytest.mark.parametrize("physics", [P.GE, P.GETIME])
def test_tangent_periodic_F0(oracle, physics):
    """The tangent with a par_periodic F0 added to every slice's slots
    (snesiga_Tangent.c:125-144) against the oracle, FP64 and FP32."""
    say("tangent with F0", physics)

    def problem():
        pb = P.stretch_problem((3, 2, 4), a=0.01, noise=0.1)
        pb.physics = physics
        pb.par_mat = pb.par_mat.copy()
        pb.par_mat[7] = 0.5            # dt: a mass term rho/dt^2 that does not hide the elastic part
        return pb

    pb = problem()
    pb.par_periodic = np.array([0.001, 0.0, 0.0, 0.0, -0.002, 0.0, 0.0, 0.0, 0.0005])
    U = state(pb) if physics == P.GE else P.smooth_state(pb, amp=0.05, noise=1e-3)
    app = App(pb)
    rowptr, col = app.pattern()
    _, _, v0 = oracle.tangent(pb, U, rowptr, col)
    pb0 = problem()
    _, _, vz = oracle.tangent(pb0, U, rowptr, col)
    assert rel(vz, v0) > 1e-6                       # F0 changes the tangent
    app.U_hist[:] = U                               # history slices (getime) as the oracle has them
    v = app.tangent(U[:pb.ndof], prec=L.IGAP4H_FP64)
    assert rel(v, v0) < TOL64, rel(v, v0)
    v32 = app.tangent(U[:pb.ndof], prec=L.IGAP4H_FP32)
    assert rel(v32, v0) < TOL32, rel(v32, v0)


def _gt_state(pb, seed):
    rng = np.random.default_rng(seed)
    n = pb.ndof
    U0 = P.smooth_state(pb, amp=0.05, noise=0.0, nhist=1) + rng.uniform(-1, 1, n) * 0.01
    U1 = U0 + rng.uniform(-1, 1, n) * 0.002
    U = U1 + rng.uniform(-1, 1, n) * 0.002
    return np.concatenate([U, U1, U0])


@pytest.mark.parametrize("form", [L.IGAP4H_TANGENT_TSTRIP, L.IGAP4H_TANGENT_SV])
@pytest.mark.parametrize("case", ["ge_323_q4", "ge_132_q4", "ge_211_q2", "ge_111_q1", "ge_321_q3", "ge_F0",
                                  "gt_444_q4", "gt_212_q2", "gt_F0", "gt_321_q3"])
def test_tangent_forms(oracle, form, case):
    """Both element-tangent kernels (T strip; strain variation, chunks of 1,
    2 or 4 Gauss points with ragged last chunks) against the oracle, FP64 and
    FP32 (FP64 and FP32 values, and FP64 into FP32 values), ge and getime,
    Gauss orders 1..4, with and without the periodic F0."""
    say("tangent form", form, case)
    phys, dims, nq = case.split("_")[0], case.split("_")[1], case.split("_")[-1]
    if dims == "F0":
        pb = P.stretch_problem((3, 2, 4), a=0.01, noise=0.1)
        pb.par_periodic = np.array([0.001, 0.0, 0.0, 0.0, -0.002, 0.0, 0.0, 0.0, 0.0005])
    else:
        pb = P.stretch_problem(tuple(int(x) for x in dims), a=0.01, noise=0.1, nquad=int(nq[1]))
    if phys == "gt":
        pb.physics = P.GETIME
        pb.par_mat = pb.par_mat.copy()
        pb.par_mat[7] = 0.5
        Uh = _gt_state(pb, 13)
    else:
        Uh = state(pb, seed=17)
    app = App(pb)
    app.set_tangent_form(form)
    app.U_hist[:] = Uh
    rowptr, col = app.pattern()
    _, _, v0 = oracle.tangent(pb, Uh, rowptr, col)
    U = Uh[:pb.ndof]
    v = app.tangent(U, prec=L.IGAP4H_FP64)
    assert rel(v, v0) < TOL64, rel(v, v0)
    v32 = app.tangent(U, prec=L.IGAP4H_FP32)
    assert rel(v32, v0) < TOL32, rel(v32, v0)
    v32f = app.tangent(U, prec=L.IGAP4H_FP32, vals_type=L.IGAP4H_VALS_F32)
    assert v32f.dtype == np.float32 and rel(v32f.astype(np.float64), v0) < TOL32
    # FP64 contraction into FP32 values (the colours' contributions add in FP32)
    v64f = app.tangent(U, prec=L.IGAP4H_FP64, vals_type=L.IGAP4H_VALS_F32)
    assert v64f.dtype == np.float32 and rel(v64f.astype(np.float64), v0) < TOL32
