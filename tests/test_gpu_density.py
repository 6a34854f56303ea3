"""iga_intdensity on the GPU (igap4h_intdensity_dev) against the oracle
(or_intdensity, src/iga_intdensity.c:10-293): every density output of phys/ge
(13) and phys/getime (2), volume integrals minus the Neumann boundary terms.
Bar: 1e-12 relative (norm over the continuous outputs; the indicator
integrals are sums of the same weights)."""
import sys

import numpy as np
import pytest

from igap4standard_amd import problem as P
from igap4standard_amd.snesiga import App

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def traction(face, order, idof, xq):
    return (0.3 + xq[0] - 0.5 * xq[1] * xq[2]) * (1 + idof) * (1 - 0.4 * order)


def cases():
    ref = P.reference_example(4)                       # reference par_mat, ex1 boundary conditions
    ref.neumann = traction
    lame = P.stretch_problem((3, 4, 5), a=0.01, noise=0.1)
    lame.bc_type[18 + 3:18 + 6] = 1
    lame.par_periodic = np.array([0.001, 0.0, 0.0, 0.0, -0.002, 0.0, 0.0, 0.0, 0.0005])
    lame.neumann = traction
    gt = P.transient_problem(3)
    gt.bc_type = P.bc_ex1()
    gt.neumann = traction
    gt_free = P.transient_problem(4)                   # all clamped: no boundary term
    return {"ge_reference": (ref, 0.15), "ge_lame": (lame, 0.05), "getime": (gt, 0.05), "getime_clamped": (gt_free, 0.05)}


@pytest.mark.parametrize("name", ["ge_reference", "ge_lame", "getime", "getime_clamped"])
def test_intdensity_matches_oracle(oracle, name):
    import torch
    pb, amp = cases()[name]
    say("intdensity", name)
    U = P.smooth_state(pb, amp=amp, noise=amp * 0.2, seed=0x5EED0201)
    app = App(pb)
    d = torch.tensor(U, device="cuda:0")
    HH = app.intdensity_dev(d.data_ptr())
    HH2 = app.intdensity_dev(d.data_ptr())
    H0 = oracle.intdensity(pb, U)
    say("  gpu", HH, "\n  oracle", H0)
    assert len(HH) == len(H0) == (13 if pb.physics == P.GE else 2)
    assert np.array_equal(HH, HH2)                     # fixed-order sums
    cont = slice(0, 7) if pb.physics == P.GE else slice(0, 2)
    assert np.linalg.norm(HH[cont] - H0[cont]) <= 1e-12 * np.linalg.norm(H0[cont])
    if pb.physics == P.GE:
        assert np.allclose(HH[7:], H0[7:], rtol=1e-12, atol=1e-15)
        if name == "ge_reference":
            assert np.count_nonzero(H0[7:10]) >= 2       # several variants present
    e = app.energy_dev(d.data_ptr())
    assert abs(e - oracle.energy(pb, U)) <= 1e-12 * abs(e)
    app.close()
