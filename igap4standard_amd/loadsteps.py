"""Incremental loading of the clamp/stretch problem (phys/ge): u_x = a k / K
at x = 1 for k = 1..K, one GPU Newton solve per step from the previous step
(the first from 1e-4 sin(i), as example/ex1/main.c:49), the Dirichlet data
pushed with igap4h_update_bc between steps.

By default the Newton solves follow the reference's semantics: the Dirichlet
rows start from the guess's values (R = U - g there) and the tangent's
Dirichlet columns are zeroed without a right-hand-side correction
(src/snesiga_Tangent.c:270).  lift=True sets U = g on those rows first.

The guess of each step is the application's choice (the reference's
`initguess` hook, include/snes.h:24): the previous step's solution (the
first step from `init`: 1e-4 sin(i) as example/ex1/main.c:49, or zero),
plus, with extend=True, the increment of the imposed affine field
u_i = a_iJ X_J carried into the interior (its control values at every node),
so that the guess meets the new Dirichlet data with no boundary layer.

    python -m igap4standard_amd.loadsteps --n 8 --steps 5 --nquad 3   (one JSON line)
"""
import json
import sys
import time

import numpy as np

from . import _lib as L
from . import problem as P
from .snesiga import App

FORMS = {
    "fp64": {},
    "mixed": dict(tangent_prec=L.IGAP4H_FP32, vals_type=L.IGAP4H_VALS_F32),
    "mixed_f16s": dict(tangent_prec=L.IGAP4H_FP32, vals_type=L.IGAP4H_VALS_F32,
                       mg_smooth_vals=L.IGAP4H_SMOOTH_F16),
}


def problem(n=8, steps=5, a=0.01, nquad=3, lam=1.0, mu=0.5, ell=0.1, noise=0.0, seed=0x5EED0002):
    """The clamp/stretch problem at its first load step u_x = a / steps."""
    return P.stretch_problem(n, a=a / steps, lam=lam, mu=mu, ell=ell, noise=noise, seed=seed, nquad=nquad)


def affine_increment(pb, da):
    """Control values at every node of u_x = da X (the increment of the
    imposed field between two load steps)."""
    inc = P.Problem(nelem=pb.nelem, knots=pb.knots)
    inc.par_dirichlet = np.zeros(9)
    inc.par_dirichlet[0] = da
    return P.affine_guess(inc, nhist=1)


def run(n=8, steps=5, a=0.01, nquad=3, lam=1.0, mu=0.5, ell=0.1, noise=0.0, seed=0x5EED0002, form="fp64",
        lift=False, init="sin", extend=False, device="cuda:0", callback=None, pb=None, kmax=None, **newton_kw):
    """Load steps k = 1..steps (only the first kmax if given); returns one
    record per step.  callback(pb, U, rec) runs after each step with the
    device state U (slice 0)."""
    import torch
    dev = torch.device(device)
    if pb is None:
        pb = problem(n, steps, a, nquad, lam, mu, ell, noise, seed)
    app = App(pb)
    kw = dict(FORMS[form])
    kw.update(newton_kw)
    o = app.newton_opts(dirichlet_lift=int(bool(lift)), **kw)   # reference tolerance: ||R|| < 1e-10
    U0 = 1e-4 * np.sin(np.arange(pb.ndof)) if init == "sin" else np.zeros(pb.ndof)
    U = torch.tensor(U0, device=dev)
    dU = torch.tensor(affine_increment(pb, a / steps), device=dev) if extend else None
    out = []
    for k in range(1, (kmax or steps) + 1):
        pb.par_dirichlet[0] = a * k / steps
        app.update_bc()
        if extend:
            U += dU
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = app.newton_solve_dev(U.data_ptr(), o)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        its = info["newton_its"]
        rec = {"step": k, "u_x": a * k / steps, "newton_its": its, "ksp_its": info["ksp_its"],
               "ksp_its_per_newton": info["ksp_its_hist"][:its], "fnorm0": info["fnorm0"],
               "fnorm": info["fnorm"], "converged": info["converged"], "s": dt,
               "ms_per_newton_step": 1000.0 * dt / max(its, 1), "t_tangent_s": info["t_tangent"],
               "t_solve_s": info["t_solve"], "energy": app.energy_dev(U.data_ptr())}
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
    ap.add_argument("--noise", type=float, default=0.0, help="relative Lame noise per element")
    ap.add_argument("--form", default="fp64", choices=sorted(FORMS))
    ap.add_argument("--lift", action="store_true", help="set U = g on the Dirichlet rows before each solve")
    ap.add_argument("--init", default="sin", choices=("sin", "zero"), help="first guess: 1e-4 sin(i) or zero")
    ap.add_argument("--extend", action="store_true", help="add the affine load increment to each guess")
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    steps = run(args.n, args.steps, nquad=args.nquad, noise=args.noise, form=args.form, lift=args.lift,
                init=args.init, extend=args.extend,
                callback=lambda pb, U, r: print("loadsteps:", json.dumps(r), file=sys.stderr, flush=True))
    print(json.dumps({"n": args.n, "nquad": args.nquad, "form": args.form, "lift": args.lift, "init": args.init,
                      "extend": args.extend, "total_s": time.perf_counter() - t0, "steps": steps}), flush=True)


if __name__ == "__main__":
    main()
