"""Incremental loading of the clamp/stretch problem (phys/ge): u_x = a k / K
at x = 1 for k = 1..K, one GPU Newton solve per step from the previous step
(the first from 1e-4 sin(i), as example/ex1/main.c:49), the Dirichlet data
pushed with igap4h_update_bc between steps.

    python -m igap4standard_amd.loadsteps --n 8 --steps 5 --nquad 3   (one JSON line)
"""
import json
import sys
import time

import numpy as np

from . import _lib as L
from . import problem as P
from .snesiga import App


def run(n=8, steps=5, a=0.01, nquad=3, lam=1.0, mu=0.5, ell=0.1, mixed=False, device="cuda:0", callback=None):
    import torch
    dev = torch.device(device)
    pb = P.stretch_problem(n, a=a / steps, lam=lam, mu=mu, ell=ell, nquad=nquad)
    app = App(pb)
    kw = {}
    if mixed:
        kw = dict(tangent_prec=L.IGAP4H_FP32, vals_type=L.IGAP4H_VALS_F32)
    o = app.newton_opts(dirichlet_lift=1, **kw)   # reference tolerance: ||R|| < 1e-10
    U = torch.tensor(1e-4 * np.sin(np.arange(pb.ndof)), device=dev)
    out = []
    for k in range(1, steps + 1):
        pb.par_dirichlet[0] = a * k / steps
        app.update_bc()
        t0 = time.perf_counter()
        info = app.newton_solve_dev(U.data_ptr(), o)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rec = {"step": k, "u_x": a * k / steps, "newton_its": info["newton_its"], "ksp_its": info["ksp_its"],
               "fnorm": info["fnorm"], "converged": info["converged"], "s": dt,
               "energy": app.energy_dev(U.data_ptr())}
        out.append(rec)
        if callback:
            callback(pb, U, rec)
    app.close()
    return out


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=8)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--nquad", type=int, default=3)
    ap.add_argument("--mixed", action="store_true")
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    steps = run(args.n, args.steps, nquad=args.nquad, mixed=args.mixed,
                callback=lambda pb, U, r: print("loadsteps:", json.dumps(r), file=sys.stderr, flush=True))
    print(json.dumps({"n": args.n, "nquad": args.nquad, "form": "mixed" if args.mixed else "fp64",
                      "total_s": time.perf_counter() - t0, "steps": steps}), flush=True)


if __name__ == "__main__":
    main()
