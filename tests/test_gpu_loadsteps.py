"""Incremental loading on the GPU (igap4standard_amd/loadsteps.py, with
igap4h_update_bc between steps): every load step converges to the reference's
absolute tolerance and satisfies the oracle's residual for that step's
Dirichlet data; the Dirichlet rows follow the data; energy grows with load."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_load_steps_match_oracle(oracle):
    from igap4standard_amd.loadsteps import run
    checks = []

    def cb(pb, U, rec):
        Un = U.cpu().numpy()
        R = oracle.residual(pb, Un)
        rows = oracle.dirichlet_rows(pb)
        checks.append((rec, float(np.linalg.norm(R)), float(np.abs(R[rows]).max())))
        print("step", rec["step"], rec["newton_its"], "|R| oracle %.2e" % checks[-1][1], flush=True)

    steps = run(n=4, steps=3, nquad=3, callback=cb)
    assert len(steps) == 3
    for rec, rn, rd in checks:
        assert rec["converged"] == 1 and rec["fnorm"] < 1e-10
        assert rn < 1e-10 and rd < 1e-12
    e = [r["energy"] for r in steps]
    assert e[0] < e[1] < e[2]
    # quadratic in the load for a small stretch: E(k) ~ k^2
    assert abs(e[2] / e[0] - 9.0) < 0.1
