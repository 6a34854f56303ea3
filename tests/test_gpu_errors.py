"""Failure paths, forced from the host (every kernel still exits normally):
a non-finite residual is reported as a divergence code, the stagnation rule
of src/snes_solve.c:30-35 and a line-search failure stop the Newton loop
with their codes, refused device allocations (IGAP4H_DEBUG_ALLOC_LIMIT)
return IGAP4H_ENOMEM and leave nothing half-built, and the traction hook is
evaluated again on request (the reference calls it in every callback)."""
import os
import sys

import numpy as np
import pytest

from igap4standard_amd import _lib as L
from igap4standard_amd import problem as P
from igap4standard_amd.snesiga import App

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture
def dev():
    import torch
    return torch, torch.device("cuda:0")


def test_nan_in_U_reported_as_divergence(dev, oracle):
    torch, d = dev
    say("NaN in U")
    pb = P.stretch_problem(3, a=0.01, noise=0.1)
    app = App(pb)
    U = P.affine_guess(pb)
    nb = pb.nbasis
    U[3 * ((2 * nb[1] + 2) * nb[2] + 2)] = np.nan       # an interior node (not a Dirichlet row)
    Uh = torch.tensor(U, device=d)
    info = app.newton_solve_dev(Uh.data_ptr(), rtol=1e-10, atol=0.0, dirichlet_lift=1)
    torch.cuda.synchronize()
    assert info["converged"] == -4 and info["newton_its"] == 0 and not np.isfinite(info["fnorm0"])
    # the context stays usable
    Uh = torch.tensor(P.affine_guess(pb), device=d)
    info = app.newton_solve_dev(Uh.data_ptr(), rtol=1e-10, atol=0.0, dirichlet_lift=1)
    torch.cuda.synchronize()
    assert info["converged"] == 1
    assert np.linalg.norm(oracle.residual(pb, Uh.cpu().numpy())) < 1e-10 * info["fnorm0"] * 10


def test_stagnation_rule(dev):
    """A linear tolerance above 1 makes every Newton update zero: with full
    steps the second step leaves |F| unchanged and the stagnation rule
    (|dF| / F < 1e-6, fold starting at 999 as snes_solve.c:23) stops the solve
    with -1; with the backtracking line search no step length decreases |F|
    and the solve stops with -3, keeping the last accepted iterate."""
    torch, d = dev
    pb = P.stretch_problem(3, a=0.01, noise=0.1)
    app = App(pb)
    U0 = P.affine_guess(pb) + 1e-3
    kw = dict(atol=1e-30, rtol=0.0, ksp_rtol=10.0, ksp_guess=0)
    Uh = torch.tensor(U0, device=d)
    info = app.newton_solve_dev(Uh.data_ptr(), line_search=0, **kw)
    torch.cuda.synchronize()
    say("  stagnation", info["converged"], info["newton_its"], info["fnorm_hist"])
    assert info["converged"] == -1 and info["newton_its"] == 2
    assert info["fnorm_hist"][0] == info["fnorm_hist"][1] == info["fnorm_hist"][2]
    assert np.array_equal(Uh.cpu().numpy(), U0)
    Uh = torch.tensor(U0, device=d)
    info = app.newton_solve_dev(Uh.data_ptr(), line_search=1, **kw)
    torch.cuda.synchronize()
    assert info["converged"] == -3 and info["newton_its"] == 1
    assert np.array_equal(Uh.cpu().numpy(), U0)


def test_forced_allocation_failure(dev, monkeypatch):
    torch, d = dev
    pb = P.stretch_problem(6, a=0.01, noise=0.1)
    monkeypatch.setenv("IGAP4H_DEBUG_ALLOC_LIMIT", str(1 << 20))     # the CSR columns need 1.7 MB
    with pytest.raises(L.IgapError) as e:
        App(pb)
    assert e.value.code == L.IGAP4H_ENOMEM
    monkeypatch.delenv("IGAP4H_DEBUG_ALLOC_LIMIT")
    app = App(pb)
    # the Newton work space: refused, then built on the next call
    monkeypatch.setenv("IGAP4H_DEBUG_ALLOC_LIMIT", str(100 << 10))
    Uh = torch.tensor(P.affine_guess(pb), device=d)
    with pytest.raises(L.IgapError) as e:
        app.newton_solve_dev(Uh.data_ptr(), rtol=1e-10, atol=0.0, dirichlet_lift=1)
    assert e.value.code == L.IGAP4H_ENOMEM
    monkeypatch.delenv("IGAP4H_DEBUG_ALLOC_LIMIT")
    info = app.newton_solve_dev(Uh.data_ptr(), rtol=1e-10, atol=0.0, dirichlet_lift=1)
    torch.cuda.synchronize()
    assert info["converged"] == 1
    app.close()


def test_update_neumann_reevaluates_hook(oracle):
    state = {"t": 0.1}
    pb = P.stretch_problem((3, 2, 4), a=0.01, noise=0.1)
    pb.neumann = lambda face, order, idof, xq: state["t"] * (1 + idof) * (1 + xq[0] - xq[2])
    U = P.smooth_state(pb, amp=0.05, noise=1e-3)
    app = App(pb)
    R1 = app.residual(U)
    assert np.linalg.norm(R1 - oracle.residual(pb, U)) <= 1e-12 * np.linalg.norm(R1)
    state["t"] = 0.35
    assert np.array_equal(app.residual(U), R1)       # the context keeps the load made at create
    app.update_neumann()
    R2 = app.residual(U)
    R0 = oracle.residual(pb, U)
    assert not np.array_equal(R2, R1)
    assert np.linalg.norm(R2 - R0) <= 1e-12 * np.linalg.norm(R0)
