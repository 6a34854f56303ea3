"""Implicit time stepping of phys/getime on the GPU.

The reference's driver (example/ex1/main.c:46-60) solves one SNES problem per
step and then shifts the history, U_hist[h] = U_hist[h-1] for h = torder..1
(:55), starting each solve from the previous step (initguess, :116-124).
Here every step is one igap4h_newton_solve_dev on the device history
U_hist = [U_{n+1}, U_n, U_{n-1}] and the shift is two device copies.

    python -m igap4standard_amd.transient --n 48 --steps 100   (one JSON line)
"""
import json
import sys
import time

import numpy as np

from . import _lib as L
from . import problem as P
from .snesiga import App


class TimeStepper:
    def __init__(self, pb, mixed=True, device="cuda:0", **newton_kw):
        import torch
        if pb.physics != P.GETIME:
            raise ValueError("TimeStepper needs a phys/getime problem")
        self.torch = torch
        self.dev = torch.device(device)
        self.pb = pb
        self.app = App(pb)
        n = pb.ndof
        self.n = n
        self.Uh = torch.zeros(3 * n, device=self.dev, dtype=torch.float64)
        kw = dict(atol=1e-10, rtol=0.0)   # the reference's absolute tolerance (src/snes_init.c:18)
        # the mass term rho/dt^2 makes these systems well conditioned (~20 GMRES
        # iterations per Newton step): a lighter smoother than the static default
        # (48^3: 147 ms per step against 163 with degree 5 over [lmax/200, lmax])
        kw.update(mg_pre=2, mg_post=2, mg_cheb_ratio=20.0)
        if mixed:
            kw.update(tangent_prec=L.IGAP4H_FP32, vals_type=L.IGAP4H_VALS_F32)
        kw.update(newton_kw)
        self.opts = self.app.newton_opts(**kw)

    def set_state(self, U_n, U_nm1):
        """History: slice 1 = U_n, slice 2 = U_{n-1}; slice 0 (the guess) = U_n."""
        t = self.torch
        self.Uh[self.n:2 * self.n] = t.as_tensor(U_n, device=self.dev)
        self.Uh[2 * self.n:] = t.as_tensor(U_nm1, device=self.dev)
        self.Uh[:self.n] = self.Uh[self.n:2 * self.n]

    def step(self):
        info = self.app.newton_solve_dev(self.Uh.data_ptr(), self.opts)
        n = self.n
        self.Uh[2 * n:].copy_(self.Uh[n:2 * n])   # U_{n-1} <- U_n
        self.Uh[n:2 * n].copy_(self.Uh[:n])       # U_n <- U_{n+1}; slice 0 stays as the next guess
        return info


def run(n=48, steps=100, dt=1e-3, fp64=False, lag=1, persist=False, say=None, tangent_form=L.IGAP4H_TANGENT_AUTO):
    """Config-3 run: n^3 getime, all faces clamped, v0 a Gaussian random field
    (std 1e-3, correlation 0.2), `steps` steps of dt.  Returns (summary dict,
    the TimeStepper, whose context stays open)."""
    import torch
    say = say or (lambda *a: None)
    pb = P.transient_problem(n, dt=dt)
    v0 = P.initial_velocity(pb)
    ts = TimeStepper(pb, mixed=not fp64, lag_jacobian=lag, lag_persist=int(persist))
    ts.app.set_tangent_form(tangent_form)
    ts.set_state(np.zeros(pb.ndof), -dt * v0)
    say("%d^3 getime, %d dofs, dt %g, %s tangent" % (n, pb.ndof, dt, "FP64" if fp64 else "FP32"))
    infos, norms = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        info = ts.step()
        infos.append(info)
        if info["converged"] != 1:
            say("step %d did not converge: %s" % (k, info))
            break
        if k % 10 == 0 or k == steps - 1:
            u = float(torch.linalg.vector_norm(ts.Uh[:ts.n]))
            norms.append((k + 1, u))
            say("step %d: %d Newton, %d Krylov, |R| %.2e, |U| %.4e, %.1f s" % (
                k + 1, info["newton_its"], info["ksp_its"], info["fnorm"], u, time.perf_counter() - t0))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    done = len(infos)
    out = {"n": n, "ndof": pb.ndof, "nelem": pb.n_elem, "dt": dt, "steps": done,
           "tangent": "fp64" if fp64 else "fp32-mfma (fp32 values)", "total_s": wall,
           "ms_per_step": 1000.0 * wall / max(done, 1),
           "newton_its_mean": float(np.mean([i["newton_its"] for i in infos])),
           "ksp_its_mean": float(np.mean([i["ksp_its"] for i in infos])),
           "tangents": int(sum(i["tangents"] for i in infos)), "lag_jacobian": lag, "lag_persist": persist,
           "max_final_fnorm": float(max(i["fnorm"] for i in infos)),
           "all_converged": all(i["converged"] == 1 for i in infos),
           "elements_per_s_per_newton_step": pb.n_elem * sum(i["newton_its"] for i in infos) / wall,
           "u_norm_history": norms}
    return out, ts


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=48)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--fp64", action="store_true", help="FP64 tangent (default: FP32-MFMA tangent, FP32 values)")
    ap.add_argument("--lag", type=int, default=1, help="rebuild the tangent every LAG Newton steps")
    ap.add_argument("--persist", action="store_true", help="count the lag across time steps")
    ap.add_argument("--tangent-form", choices=("auto", "tstrip", "sv"), default="auto",
                    help="element tangent kernel (igap4h_set_tangent_form)")
    args = ap.parse_args(argv)
    form = {"auto": L.IGAP4H_TANGENT_AUTO, "tstrip": L.IGAP4H_TANGENT_TSTRIP, "sv": L.IGAP4H_TANGENT_SV}[args.tangent_form]
    out, ts = run(args.n, args.steps, args.dt, args.fp64, args.lag, args.persist,
                  say=lambda *a: print("transient:", *a, file=sys.stderr, flush=True), tangent_form=form)
    ts.app.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
