"""Implicit getime time stepping on the GPU (igap4standard_amd/transient.py):
every step's solution satisfies the oracle's getime residual with the same
history to the reference's absolute tolerance 1e-10 (src/snes_init.c:18), the
history shift follows example/ex1/main.c:55, and the FP32-MFMA-tangent form
follows the FP64 trajectory (only the Jacobian differs)."""
import sys

import numpy as np
import pytest

from igap4standard_amd import problem as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_time_steps_match_oracle_residual(oracle):
    from igap4standard_amd.transient import TimeStepper
    say("transient 6^3")
    pb = P.transient_problem(6, dt=1e-3)
    v0 = P.initial_velocity(pb, std=1e-3, corr_len=0.3)
    n = pb.ndof
    traj = {}
    for mixed in (False, True):
        ts = TimeStepper(pb, mixed=mixed)
        ts.set_state(np.zeros(n), -1e-3 * v0)
        us = []
        for k in range(3):
            before = ts.Uh.cpu().numpy()
            info = ts.step()
            after = ts.Uh.cpu().numpy()
            assert info["converged"] == 1, info
            U_new = after[:n]
            assert np.array_equal(after[n:2 * n], U_new) and np.array_equal(after[2 * n:], before[n:2 * n])
            R = oracle.residual(pb, np.concatenate([U_new, before[n:2 * n], before[2 * n:]]))
            say("  mixed", mixed, "step", k, "newton", info["newton_its"], "|R| oracle %.2e" % np.linalg.norm(R))
            assert np.linalg.norm(R) < 1e-10
            us.append(U_new)
        traj[mixed] = np.array(us)
    rel = np.linalg.norm(traj[True] - traj[False]) / np.linalg.norm(traj[False])
    assert rel < 1e-8, rel
    assert np.linalg.norm(traj[False][-1]) > 0.0


def test_lagged_tangent_steps(oracle):
    """Modified Newton (tangent of the first step reused for all steps,
    lag_jacobian with lag_persist): every step still reaches the absolute
    tolerance and the oracle's residual; one tangent assembly in total."""
    from igap4standard_amd.transient import TimeStepper
    pb = P.transient_problem(6, dt=1e-3)
    v0 = P.initial_velocity(pb, std=1e-3, corr_len=0.3)
    n = pb.ndof
    ts = TimeStepper(pb, mixed=True, lag_jacobian=1000, lag_persist=1)
    ts.set_state(np.zeros(n), -1e-3 * v0)
    tangents = 0
    for k in range(3):
        before = ts.Uh.cpu().numpy()
        info = ts.step()
        assert info["converged"] == 1, info
        tangents += info["tangents"]
        U_new = ts.Uh.cpu().numpy()[:n]
        R = oracle.residual(pb, np.concatenate([U_new, before[n:2 * n], before[2 * n:]]))
        assert np.linalg.norm(R) < 1e-10
    assert tangents == 1



def test_getime_energy_matches_oracle(oracle):
    """igap4h_energy_dev for getime: W at (u0 + u1) / 2 plus the kinetic
    energy rho |u1 - u0|^2 / (2 dt^2) integrated over the volume
    (phys/getime/density.c via iga_intdensity), against the oracle."""
    import torch
    from igap4standard_amd.snesiga import App
    pb = P.reference_example(3, physics=P.GETIME)
    rng = np.random.default_rng(4)
    Uh = rng.uniform(-1, 1, 3 * pb.ndof) * 0.02
    app = App(pb)
    d = torch.tensor(Uh, device="cuda:0")
    e = app.energy_dev(d.data_ptr())
    e0 = oracle.energy(pb, Uh)
    assert abs(e - e0) <= 1e-12 * abs(e0), (e, e0)
