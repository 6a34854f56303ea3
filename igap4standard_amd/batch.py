"""Batched assembly of many independent static phys/ge instances on one GPU
(igap4h_batch_* in include/igap4h.h): the contexts of all instances, their
device arrays, the autotuner, and a runner for the 256-instance workload.

    python -m igap4standard_amd.batch [--count 256] [--reps 5]   (one JSON line)
"""
import ctypes
import json
import sys
import time

import numpy as np

from . import _lib as L
from . import problem as P
from .snesiga import App

SIZES = (8, 12, 16, 20, 24)


def batch_problems(count=256, seed=0x5EED0004, a=0.005, nquad=3):
    """count clamp/stretch instances: n uniform from SIZES, lambda and mu
    log-uniform in [0.1, 10], l uniform in [0.02, 0.2] (splitmix64 draws from
    seed, four per instance), u_x = a at x = 1, nquad^3 Gauss points."""
    u = P.uniform01(seed, 4 * count).reshape(count, 4)
    pbs = []
    for k in range(count):
        n = SIZES[min(int(u[k, 0] * len(SIZES)), len(SIZES) - 1)]
        lam = 10.0 ** (2.0 * u[k, 1] - 1.0)
        mu = 10.0 ** (2.0 * u[k, 2] - 1.0)
        ell = 0.02 + 0.18 * u[k, 3]
        pbs.append(P.stretch_problem(n, a=a, lam=lam, mu=mu, ell=ell, nquad=nquad))
    return pbs


class Batch:
    """Contexts, device arrays and the library batch for a list of problems.
    U_hist of instance i starts as problem.smooth_state (non-equilibrium)."""

    def __init__(self, pbs, prec=L.IGAP4H_FP32, vals_type=L.IGAP4H_VALS_F32, device="cuda:0"):
        import torch
        self.torch = torch
        dev = torch.device(device)
        self.pbs = pbs
        self.apps = [App(pb) for pb in pbs]
        vdt = torch.float32 if vals_type == L.IGAP4H_VALS_F32 else torch.float64
        self.U = [torch.tensor(P.smooth_state(pb, amp=1e-3, noise=1e-6, seed=0x5EED0012 + i), device=dev)
                  for i, pb in enumerate(pbs)]
        self.R = [torch.empty(a.ndof, device=dev, dtype=torch.float64) for a in self.apps]
        self.vals = [torch.empty(a.nnz, device=dev, dtype=vdt) for a in self.apps]
        n = len(pbs)
        vp = ctypes.c_void_p
        ctxs = (vp * n)(*[a.ctx for a in self.apps])
        us = (vp * n)(*[t.data_ptr() for t in self.U])
        rs = (vp * n)(*[t.data_ptr() for t in self.R])
        vs = (vp * n)(*[t.data_ptr() for t in self.vals])
        b = vp()
        L.check(L.lib().igap4h_batch_create(ctxs, n, us, rs, vs, prec, vals_type, ctypes.byref(b)))
        self.b = b
        self.prec, self.vals_type = prec, vals_type
        ne, q = ctypes.c_int64(), ctypes.c_int()
        L.check(L.lib().igap4h_batch_sizes(b, ctypes.byref(ne), ctypes.byref(q)))
        self.nelem = ne.value

    def variant(self):
        ne, q = ctypes.c_int64(), ctypes.c_int()
        L.check(L.lib().igap4h_batch_sizes(self.b, ctypes.byref(ne), ctypes.byref(q)))
        return q.value

    def set_variant(self, qc):
        L.check(L.lib().igap4h_batch_set_variant(self.b, qc))

    def tune(self, reps=3):
        best = ctypes.c_int()
        nv = L.IGAP4H_BATCH_MAX_VARIANTS
        ms = (ctypes.c_double * nv)()
        qs = (ctypes.c_int * nv)()
        L.check(L.lib().igap4h_batch_tune(self.b, reps, ctypes.byref(best), ms, qs))
        return best.value, {int(qs[k]): float(ms[k]) for k in range(nv) if qs[k]}

    def assemble(self, stream=None):
        L.check(L.lib().igap4h_batch_assemble(self.b, ctypes.c_void_p(stream or 0)))

    def close(self):
        if getattr(self, "b", None):
            L.lib().igap4h_batch_destroy(self.b)
            self.b = None
        for a in getattr(self, "apps", []):
            a.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(count=256, reps=5, fp64=False, seed=0x5EED0004, sequential=True, say=None, variant=None):
    """Config-4 run: count instances, autotuner (or the given variant), timed
    graph replays of the whole assembly (and, optionally, the same work
    instance by instance)."""
    import torch
    say = say or (lambda *a: None)
    pbs = batch_problems(count, seed=seed)
    prec, vt = (L.IGAP4H_FP64, L.IGAP4H_VALS_F64) if fp64 else (L.IGAP4H_FP32, L.IGAP4H_VALS_F32)
    t0 = time.perf_counter()
    B = Batch(pbs, prec, vt)
    t_setup = time.perf_counter() - t0
    say("%d instances, %d elements, set-up %.1f s" % (len(pbs), B.nelem, t_setup))
    if variant is None:
        best, table = B.tune(reps)
        say("autotuner (variant = 10 form + Gauss points per MFMA chunk -> tangent ms):", table, "-> %d" % best)
    else:
        B.set_variant(variant)
        best, table = variant, {}
    stream = torch.cuda.current_stream()
    B.assemble(stream.cuda_stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        B.assemble(stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    ms_seq = None
    if sequential:  # the same work instance by instance (residual_dev + tangent_dev per context)
        def seq():
            for a, U, R, V in zip(B.apps, B.U, B.R, B.vals):
                a.residual_dev(U.data_ptr(), R.data_ptr(), stream.cuda_stream)
                a.tangent_dev(U.data_ptr(), V.data_ptr(), prec, vt, stream.cuda_stream)
        seq()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            seq()
        e1.record(stream)
        torch.cuda.synchronize()
        ms_seq = e0.elapsed_time(e1) / reps
    sizes = {n: sum(1 for pb in pbs if pb.nelem[0] == n) for n in SIZES}
    out = {"instances": len(pbs), "nelem": B.nelem, "sizes": sizes, "nquad": 3, "setup_s": t_setup,
           "precision": "fp64" if fp64 else "fp32-mfma tangent (fp64 chunk accumulation), fp32 values, fp64 residual",
           "autotuner_ms": table, "variant": best,
           "variant_key": "10 x form (0 T strip, 1 strain variation) + Gauss points per MFMA chunk",
           "graph_ms_per_assembly": ms,
           "elements_per_s": B.nelem / (ms * 1e-3), "sequential_ms_per_assembly": ms_seq,
           "graph_speedup_vs_sequential": (ms_seq / ms) if ms_seq else None}
    B.close()
    return out


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--variant", type=int, default=None, help="skip the autotuner, use this variant")
    ap.add_argument("--no-sequential", action="store_true", help="graph replays only")
    args = ap.parse_args(argv)
    out = run(args.count, args.reps, args.fp64, sequential=not args.no_sequential,
              say=lambda *a: print("batch:", *a, file=sys.stderr, flush=True), variant=args.variant)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
