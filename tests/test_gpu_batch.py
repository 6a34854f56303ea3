"""Batched assembly (igap4h_batch_*): every instance's R and K from one graph
replay equal the single-instance entry points bit for bit when the batch uses
the single-instance kernel variant, and match the oracle-checked FP64 path to
the usual bars for every variant the autotuner may pick; graph replays are
repeatable."""
import sys

import numpy as np
import pytest

from igap4standard_amd import _lib as L
from igap4standard_amd import problem as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_batch():
    from igap4standard_amd.batch import batch_problems
    pbs = batch_problems(6, seed=0x5EED0004)
    # small sizes: keep the test fast but ragged
    for k, n in enumerate((2, 3, 4, 2, 5, 3)):
        old = pbs[k]
        pbs[k] = P.stretch_problem((n, n + 1, n), a=0.005, nquad=3)
        pbs[k].elem_coef = P.coef_from_lame(*[np.full(pbs[k].n_elem, v) for v in (1.0 + k, 0.5 + 0.1 * k, 0.05)])
    return pbs


@pytest.mark.parametrize("prec,vt", [(L.IGAP4H_FP64, L.IGAP4H_VALS_F64), (L.IGAP4H_FP32, L.IGAP4H_VALS_F32)])
def test_batch_equals_single_instance(prec, vt):
    import torch
    from igap4standard_amd.batch import Batch
    say("batch", prec, vt)
    pbs = small_batch()
    B = Batch(pbs, prec, vt)
    default_qc = B.variant()
    for rep in range(2):
        B.assemble()
        torch.cuda.synchronize()
        Rb = [r.cpu().numpy().copy() for r in B.R]
        Vb = [v.cpu().numpy().copy() for v in B.vals]
        for a, U, R, V, rb, vb in zip(B.apps, B.U, B.R, B.vals, Rb, Vb):
            a.residual_dev(U.data_ptr(), R.data_ptr())
            a.tangent_dev(U.data_ptr(), V.data_ptr(), prec, vt)
            torch.cuda.synchronize()
            assert np.array_equal(R.cpu().numpy(), rb)
            assert np.array_equal(V.cpu().numpy(), vb)
    # every variant of the tuning space against the default one
    best, table = B.tune(2)
    say("  tuner", table, best)
    assert best in table and all(t > 0 for t in table.values())
    ref = [v.astype(np.float64) for v in Vb]
    tol = 1e-12 if prec == L.IGAP4H_FP64 else 1e-5
    for qc in table:
        B.set_variant(qc)
        B.assemble()
        torch.cuda.synchronize()
        for v, r in zip(B.vals, ref):
            x = v.cpu().numpy().astype(np.float64)
            assert np.linalg.norm(x - r) / np.linalg.norm(r) < tol, (qc, np.linalg.norm(x - r) / np.linalg.norm(r))
    if prec == L.IGAP4H_FP32:
        # form 2 is form 1 with a smaller LDS footprint: the same values bit for bit
        assert 22 in table
        out = {}
        for qc in (12, 22):
            B.set_variant(qc)
            B.assemble()
            torch.cuda.synchronize()
            out[qc] = [v.cpu().numpy().copy() for v in B.vals]
        assert all(np.array_equal(a, b) for a, b in zip(out[12], out[22]))
    B.set_variant(default_qc)
    B.close()


def test_batch_compact_variant_needs_nq3():
    """Variant 22 sizes its point data for 3^3 Gauss points: a batch with an
    nq = 4 instance leaves it out of the tuning space and refuses it."""
    from igap4standard_amd.batch import Batch
    pbs = [P.stretch_problem(2, a=0.005, nquad=3), P.stretch_problem(3, a=0.005, nquad=4)]
    B = Batch(pbs, L.IGAP4H_FP32, L.IGAP4H_VALS_F32)
    best, table = B.tune(1)
    assert 22 not in table and 12 in table
    with pytest.raises(L.IgapError):
        B.set_variant(22)
    B.close()


def test_batch_matches_oracle(oracle):
    import torch
    from igap4standard_amd.batch import Batch
    pbs = small_batch()[:3]
    B = Batch(pbs, L.IGAP4H_FP64, L.IGAP4H_VALS_F64)
    B.assemble()
    torch.cuda.synchronize()
    for pb, U, R, V, a in zip(pbs, B.U, B.R, B.vals, B.apps):
        Un = U.cpu().numpy()
        R0 = oracle.residual(pb, Un)
        rowptr, col = a.pattern()
        _, _, v0 = oracle.tangent(pb, Un, rowptr, col)
        assert np.linalg.norm(R.cpu().numpy() - R0) / np.linalg.norm(R0) < 1e-12
        assert np.linalg.norm(V.cpu().numpy() - v0) / np.linalg.norm(v0) < 1e-12
    B.close()


def test_batch_follows_update_bc():
    """igap4h_update_bc on a batched context (a Dirichlet row list of another
    length, the Neumann load switched off and on again) is picked up by the
    next batch assembly, which then equals the single-instance calls bit for
    bit."""
    import torch
    from igap4standard_amd.batch import Batch
    say("batch after update_bc")
    pbs = small_batch()[:3]
    pbs[1].neumann = lambda face, order, idof, xq: 0.01 * (1 + idof) * (xq[0] + 0.5 * xq[1] - xq[2])
    B = Batch(pbs, L.IGAP4H_FP64, L.IGAP4H_VALS_F64)

    def check():
        B.assemble()
        torch.cuda.synchronize()
        Rb = [r.cpu().numpy().copy() for r in B.R]
        Vb = [v.cpu().numpy().copy() for v in B.vals]
        for a, U, R, V, rb, vb in zip(B.apps, B.U, B.R, B.vals, Rb, Vb):
            a.residual_dev(U.data_ptr(), R.data_ptr())
            a.tangent_dev(U.data_ptr(), V.data_ptr())
            torch.cuda.synchronize()
            assert np.array_equal(R.cpu().numpy(), rb)
            assert np.array_equal(V.cpu().numpy(), vb)
        return Rb

    R0 = check()
    n0 = B.apps[0].ndirichlet
    pbs[0].bc_type[12:18] = 0            # face 2 clamped too: more Dirichlet rows
    B.apps[0].update_bc()
    pbs[1].bc_type = P.bc_all_clamped()  # no Neumann face left: the load is dropped
    B.apps[1].update_bc()
    assert B.apps[0].ndirichlet > n0
    R1 = check()
    assert not np.array_equal(R0[1], R1[1])
    pbs[1].bc_type = P.bc_clamp_stretch()  # Neumann faces again: the load is rebuilt
    B.apps[1].update_bc()
    R2 = check()
    assert np.array_equal(R2[1], R0[1])
    B.close()
