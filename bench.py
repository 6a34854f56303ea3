"""Newton-step benchmark on one MI355X: phys/ge at 64^3 quadratic C1 B-spline
elements (nq = 4, 16.8 M Gauss points), clamped at x = 0 and stretched
u_x = 0.01 X at x = 1, per-element Lame fields with seeded +-10 % noise
(lambda = 1, mu = 0.5, l = 0.1).

A step is one full Newton solve on the GPU from the affine guess u = 0.01 X
(Dirichlet rows set to their values) to ||R|| / ||R0|| < 1e-10: per Newton
iteration the FP64 residual, the tangent on the AIJ pattern, GMRES(300) with a
multigrid V-cycle and the backtracking line search, all with inputs resident
in HBM.  Forms: plain FP64 (the primary line); mixed (FP32-MFMA tangent
with FP64 accumulation over Gauss-point chunks, FP32 matrix values in the
Krylov solve and the multigrid hierarchy's fine level); mixed_f16s (mixed,
the fine-level smoothing passes on a row-scaled binary16 copy of the
tangent); fp64_f16s (FP64 tangent, residual and Krylov solve, only the
smoothing passes on the binary16 copy).

Also timed: the assembly alone (residual + tangent at a non-equilibrium state
U = 1e-3 sin(pi X) sin(pi Y) sin(pi Z) (1, .5, .25) + 1e-6 noise) and the
fine-level row pass (SpMV) that the Krylov and multigrid passes share.

Extra keys ("configs"), each timed on the GPU in the same run:
  config1  8^3, 5 load steps u_x = 0.01 k / 5, 3x3x3 Gauss, first guess
           1e-4 sin(i), reference Dirichlet semantics (no lift);
  config2  64^3, 5 load steps (Lame +-10 %), each from the previous solution
           plus the affine load increment, in the three forms; also the
           reference Dirichlet semantics and the 1e-4 sin(i) first guess;
  config3  48^3 getime, 100 steps of dt = 1e-3, FP32-MFMA tangent, plus the
           getime tangent kernel's MFMA roofline;
  config4  256 batched instances {8..24}^3 (one graph replay per assembly),
           autotuner table and FP32 MFMA fraction;
  config5  32^3 adjoint: solves, gradient, backward kernel vs central
           differences on 64 elements, dL/dE vs differences of the loss.

    python bench.py [--steps K] [--warmup W] [--n 64] [--no-cpu] [--no-extras]

Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic flops of K_e per element and Gauss point: 9 dof blocks x
# (2*10*10*27 + 2*27*10*27) = 179,820  (i.e. 11,508,480 per element at nq = 4)
FLOP_KE_PER_QP = 9 * (2 * 10 * 10 * 27 + 2 * 27 * 10 * 27)
# algorithmic flops of R_e per Gauss point: u-slot interpolation 1,620 +
# projection 1,620 + basis products 540 + the pointwise stress 291 (ge,
# counted on igap4h_physics.hpp's residual_grad: 121 adds, 170 multiplies)
FLOP_RE_PER_QP = 1620 + 1620 + 540 + 291
# flops the sum-factorised residual kernel executes per element (interpolation
# 18.6 k, pointwise stress 18.6 k, projection ~18.6 k, basis products and
# weights): ~71 k of the 260,544 algorithmic
FLOP_RE_EXEC_PER_ELEM = 71000
PEAK = {"f64": 78.6, "f32": 157.3}   # MI355X dense MFMA TFLOP/s (the vector FP64 / FP32 peaks are the same)
# MFMA flops the strain-variation tangent executes per Gauss point, in 16 x 16 x 4
# tiles: ge waves 0-2 5 material slots x 3 k-steps (12 strain components per
# dof), wave 3 the 3 Gamma tiles x 3 (54 MFMAs); getime waves 0-2 4 material
# slots x 6 k-steps (24 components) + 1 Gamma slot x 3, wave 3 3 material slots
# x 6 (99 MFMAs)
SV_MFMA_FLOP_PER_QP = {"ge": (3 * 5 * 3 + 3 * 3) * 2 * 16 * 16 * 4, "getime": (3 * (4 * 6 + 3) + 3 * 6) * 2 * 16 * 16 * 4}
PEAK_HBM = 8000.0                    # GB/s


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summaries (FETCH_SIZE and WRITE_SIZE runs, scaled by a plain copy)."""
    for name in ("k_tangent_ge_64cube_traffic.txt", "k_spmv_line_64cube_traffic.txt"):
        try:
            for line in open(os.path.join(ROOT, "profiles", name)):
                if line.startswith("traffic_bytes_per_launch") and kernel in line:
                    return float(line.split("=")[1].split()[0])
        except OSError:
            pass
    return None


METRIC = "Newton-step s & elements/s, 64\u00b3 3D ge; FP32-MFMA tangent vs FP64 @ \u2016R\u2016/\u2016R0\u2016<1e-10"


_CPU = {}


def _cpu_worker(w):
    """One worker of the CPU baseline: its own disjoint range of elements."""
    import numpy as np

    from oracle import oracle as O
    d = _CPU
    pb, U, n_k, n_r = d["pb"], d["U"], d["n_k"], d["n_r"]
    e0 = 1 + w * n_k
    t0 = time.perf_counter()
    O.tangent_elements_timed(pb, U, d["rowptr"], d["col"], d["vals"], e0, e0 + n_k)
    t_k = time.perf_counter() - t0
    R = np.zeros(pb.ndof)
    r0 = (w * n_r) % max(1, pb.n_elem - n_r)
    t0 = time.perf_counter()
    O.residual_elements_timed(pb, U, R, r0, r0 + n_r)
    t_r = time.perf_counter() - t0
    return t_k, t_r


def host_description():
    """CPU model, clock and core counts of this host (lscpu, else /proc/cpuinfo)."""
    d = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU max MHz", "CPU MHz", "CPU(s)", "Socket(s)", "Core(s) per socket",
                             "Thread(s) per core"):
                d[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            if k.strip() in ("model name", "cpu MHz") and k.strip() not in d:
                d[k.strip()] = v.strip()
    except OSError:
        pass
    d["cpus_usable"] = len(os.sched_getaffinity(0))
    return d


def cpu_baseline(pb, U, budget_s=10.0, procs=None):
    """The reference algorithm restated in C (oracle/): per-point basis, naive
    81x81x10x10 contraction, AIJ insertion, one process per core (as the
    reference runs under mpiexec), each on its own range of elements.  Runs
    before the GPU is touched, so the workers are plain forks sharing the
    pattern arrays copy-on-write."""
    import multiprocessing as mp

    import numpy as np

    from oracle import oracle as O
    rowptr, col = O.pattern(pb)
    vals = np.zeros(len(col))
    t0 = time.perf_counter()
    O.tangent_elements_timed(pb, U, rowptr, col, vals, 0, 1)
    t_one = time.perf_counter() - t0
    nproc = procs or min(16, len(os.sched_getaffinity(0)))
    n_k = max(1, min((pb.n_elem - 1) // nproc, int(0.7 * budget_s / max(t_one, 1e-6))))
    n_r = min(pb.n_elem, 64 * n_k)
    _CPU.update(pb=pb, U=U, rowptr=rowptr, col=col, vals=vals, n_k=n_k, n_r=n_r)
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(nproc) as pool:
        res = pool.map(_cpu_worker, range(nproc))
    wall = time.perf_counter() - t0
    _CPU.clear()
    rate = sum(1.0 / (t_k / n_k + t_r / n_r) for t_k, t_r in res)  # all workers run at once
    t_k1 = sum(t_k for t_k, _ in res) / (nproc * n_k)
    t_r1 = sum(t_r for _, t_r in res) / (nproc * n_r)
    return {"value": rate, "unit": "elements/s", "cores": nproc, "kind": "port",
            "sample": "oracle/ C restatement of the snesiga_Tangent + snesiga_Residual loops, %d processes "
                      "(one per core, as mpiexec -n %d), each %d tangent + %d residual elements of the same "
                      "%dx%dx%d workload (%.1f s wall). Assembly only: an upper bound on a CPU Newton step's "
                      "elements/s (the CPU linear solve is not timed)" % ((nproc, nproc, n_k, n_r) + tuple(pb.nelem) + (wall,)),
            "tangent_s_per_element_per_core": t_k1, "residual_s_per_element_per_core": t_r1,
            "single_core_elements_per_s": 1.0 / (t_k1 + t_r1),
            "assembly_s_per_newton_step": pb.n_elem / rate,
            "newton_step_note": "a CPU Newton step costs at least one residual + one tangent assembly of all "
                                "elements (assembly_s_per_newton_step, on these cores); the linear solve and "
                                "line-search residuals come on top and are not timed",
            "host": host_description()}


def _events(torch, stream, fn, reps):
    """Mean ms of fn() over reps launches, HIP events on `stream`."""
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _steps_summary(steps):
    nt = sum(r["newton_its"] for r in steps)
    return {"total_s": sum(r["s"] for r in steps), "newton_its": [r["newton_its"] for r in steps],
            "ksp_its": [r["ksp_its_per_newton"] for r in steps],
            "ms_per_newton_step": 1000.0 * sum(r["s"] for r in steps) / max(nt, 1),
            "step_s": [r["s"] for r in steps], "converged": [r["converged"] for r in steps],
            "fnorm": ["%.2e" % r["fnorm"] for r in steps],
            "all_converged": all(r["converged"] == 1 for r in steps)}


def config_legs(say):
    """The other workloads, each on its own contexts (closed before the next)."""
    import torch

    from igap4standard_amd import _lib as L
    from igap4standard_amd import adjoint, batch, loadsteps, transient
    out = {}

    say("config1: 8^3, 5 load steps")
    st = loadsteps.run(n=8, steps=5, a=0.01, nquad=3, form="fp64", lift=False, init="sin")
    out["config1_8cube_5_load_steps"] = dict(_steps_summary(st), workload=(
        "8^3 clamp/stretch, lambda 1, mu 0.5, l 0.1, 3x3x3 Gauss, u_x = 0.01 k/5, first guess 1e-4 sin(i) "
        "(example/ex1/main.c:49), reference Dirichlet semantics (no lift), ||R|| < 1e-10, FP64"))

    say("config2: 64^3, 5 load steps")
    c2 = {"workload": "64^3 clamp/stretch, Lame +-10 % (seed 0x5EED0002), nq 4, u_x = 0.01 k/5; each step from "
                      "the previous solution plus the affine load increment (the app's initguess), Dirichlet rows "
                      "lifted, ||R|| < 1e-10 (reference absolute tolerance)"}
    for form in ("fp64", "mixed", "mixed_f16s"):
        st = loadsteps.run(n=64, steps=5, a=0.01, nquad=4, noise=0.1, form=form, lift=True, init="zero",
                           extend=True)
        c2[form] = _steps_summary(st)
        say("  ", form, json.dumps(c2[form]))
    for form in ("mixed", "mixed_f16s"):
        c2[form + "_speedup_vs_fp64"] = c2["fp64"]["total_s"] / c2[form]["total_s"]
    st = loadsteps.run(n=64, steps=5, a=0.01, nquad=4, noise=0.1, form="fp64", lift=False, init="zero",
                       extend=True, max_newton=10)
    c2["fp64_reference_dirichlet"] = dict(_steps_summary(st), note=(
        "no lift: the tangent's zeroed Dirichlet columns (snesiga_Tangent.c:270) drop the coupling of the "
        "Dirichlet rows' correction, the interior step is wrong and the backtracking line search fails "
        "(converged -3)"))
    st = loadsteps.run(n=64, steps=5, a=0.01, nquad=4, noise=0.1, form="fp64", lift=False, init="sin",
                       extend=True, kmax=1, max_newton=3, ksp_max_it=2000)
    c2["fp64_sin_guess_first_step"] = dict(_steps_summary(st), note=(
        "first load step from 1e-4 sin(i) (example/ex1/main.c:49), capped at 3 Newton steps of at most "
        "2000 GMRES iterations: the guess oscillates at the grid scale, its strain gradients (~1e-4/h^2) "
        "make the tangent indefinite at 64^3"))
    out["config2_64cube_5_load_steps"] = c2

    say("config3: getime 48^3, 100 steps")
    c3, ts = transient.run(48, 100, 1e-3, fp64=False, say=lambda *a: None)
    stream = torch.cuda.current_stream()
    app, pb = ts.app, ts.pb
    dR = torch.empty(app.ndof, device="cuda:0", dtype=torch.float64)
    c3["residual_ms"] = _events(torch, stream, lambda: app.residual_dev(ts.Uh.data_ptr(), dR.data_ptr(),
                                                                        stream.cuda_stream), 5)
    for prec, vt, dt, nm, peak in ((L.IGAP4H_FP32, L.IGAP4H_VALS_F32, torch.float32, "k_tangent_sv<float,2,1>",
                                    PEAK["f32"]),
                                   (L.IGAP4H_FP64, L.IGAP4H_VALS_F64, torch.float64, "k_tangent_sv<double,2,1>",
                                    PEAK["f64"])):
        V = torch.empty(app.nnz, device="cuda:0", dtype=dt)
        ms = _events(torch, stream, lambda: app.tangent_dev(ts.Uh.data_ptr(), V.data_ptr(), prec, vt,
                                                            stream.cuda_stream), 3)
        tf = pb.n_elem * FLOP_KE_PER_QP * pb.nquad ** 3 / (ms * 1e-3) / 1e12
        c3["roofline_tangent_" + ("f32" if prec == L.IGAP4H_FP32 else "f64")] = {
            "bound": "mfma", "kernel": nm + " (27 colour launches per tangent)", "tangent_ms": ms,
            "achieved": tf, "peak": peak, "unit": "TFLOP/s", "frac": tf / peak,
            "algorithmic": "11,508,480 flop per element"}
        ex = pb.n_elem * SV_MFMA_FLOP_PER_QP["getime"] * pb.nquad ** 3 / (ms * 1e-3) / 1e12
        c3["roofline_tangent_" + ("f32" if prec == L.IGAP4H_FP32 else "f64")].update(
            {"mfma_executed_tflops": ex, "mfma_executed_frac": ex / peak})
        del V
    ts.app.close()
    c3["workload"] = ("48^3 getime (torder 2), all faces clamped, rho 1, lambda 1, mu 0.5, l 0.05, dt 1e-3, v0 a "
                      "Gaussian random field (std 1e-3, correlation 0.2), FP64 residual, FP32-MFMA tangent, "
                      "||R|| < 1e-10 per step")
    out["config3_getime_48cube_100_steps"] = c3
    say("  ", json.dumps({k: v for k, v in c3.items() if k != "u_norm_history"}))

    say("config4: 256 batched instances")
    c4 = batch.run(256, reps=3, fp64=False, say=lambda *a: say("  ", *a))
    best = min(c4["autotuner_ms"].values())
    flop = c4["nelem"] * FLOP_KE_PER_QP * 27
    c4["roofline_tangent_f32"] = {"bound": "mfma", "kernel": "k_tangent_ge_batch<float,QC> / k_tangent_sv_batch<float,QC> (tuned variant)",
                                  "achieved": flop / (best * 1e-3) / 1e12, "peak": PEAK["f32"], "unit": "TFLOP/s",
                                  "frac": flop / (best * 1e-3) / 1e12 / PEAK["f32"]}
    c4["assembly_tflops_f32_tangent_share"] = flop / (c4["graph_ms_per_assembly"] * 1e-3) / 1e12
    out["config4_batch_256"] = c4
    say("  ", json.dumps(c4))

    say("config5: 32^3 adjoint")
    c5 = adjoint.run(32, nfd=64, h=1e-6, loss_fd=4, loss_h=1e-3)
    out["config5_adjoint_32cube"] = c5
    say("  ", json.dumps(c5))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed Newton solves per form")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--asm-steps", type=int, default=5, help="timed assembly steps per form")
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of work per CPU baseline process")
    ap.add_argument("--cpu-procs", type=int, default=0, help="CPU baseline processes (default: cores, at most 16)")
    ap.add_argument("--forms", default="fp64,mixed,mixed_f16s,fp64_f16s",
                    help="comma list of fp64, mixed, mixed_f16s / mixed_bf16s (the mixed form with the "
                         "fine-level multigrid smoothing passes on a row-scaled binary16 / bfloat16 copy "
                         "of the FP32 tangent), fp64_f16s (FP64 tangent and Krylov, smoothing passes on the "
                         "binary16 copy of the FP64 tangent)")
    ap.add_argument("--no-newton", action="store_true", help="assembly and SpMV only")
    ap.add_argument("--no-extras", action="store_true", help="only the 64^3 headline workload")
    ap.add_argument("--tangent-form", choices=("auto", "tstrip", "sv"), default="auto",
                    help="element tangent kernel (igap4h_set_tangent_form); auto: the faster per precision")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": a.numel() * 8}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the assembly path is single-GPU in this build "
              "(multi-GPU sharding is a design note, not implemented); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from igap4standard_amd import _lib as L
    from igap4standard_amd import problem as P
    from igap4standard_amd.snesiga import App

    def say(*a):
        print("bench:", *a, file=sys.stderr, flush=True)

    n = args.n
    pb = P.stretch_problem(n, a=0.01, lam=1.0, mu=0.5, ell=0.1, noise=0.1, seed=0x5EED0002, nquad=4)
    U = P.smooth_state(pb, amp=1e-3, noise=1e-6, seed=0x5EED0012)
    cpu = None
    if not args.no_cpu:  # before the GPU is initialised (its workers are forks)
        say("CPU baseline (%.0f s per process)" % args.cpu_budget)
        cpu = cpu_baseline(pb, U, args.cpu_budget, args.cpu_procs or None)
        say("CPU baseline", json.dumps(cpu))
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    say("building context for %d^3 (%d dofs)" % (n, pb.ndof))
    app = App(pb)
    app.set_tangent_form({"auto": L.IGAP4H_TANGENT_AUTO, "tstrip": L.IGAP4H_TANGENT_TSTRIP,
                          "sv": L.IGAP4H_TANGENT_SV}[args.tangent_form])
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream
    dUh = torch.tensor(U, device=dev, dtype=torch.float64)
    dR = torch.empty(app.ndof, device=dev, dtype=torch.float64)
    dK64 = torch.empty(app.nnz, device=dev, dtype=torch.float64)
    dK32 = torch.empty(app.nnz, device=dev, dtype=torch.float32)
    forms = {"fp64": (L.IGAP4H_FP64, L.IGAP4H_VALS_F64, dK64), "mixed": (L.IGAP4H_FP32, L.IGAP4H_VALS_F32, dK32),
             "mixed_f16s": (L.IGAP4H_FP32, L.IGAP4H_VALS_F32, dK32),
             "mixed_bf16s": (L.IGAP4H_FP32, L.IGAP4H_VALS_F32, dK32),
             "fp64_f16s": (L.IGAP4H_FP64, L.IGAP4H_VALS_F64, dK64)}
    smooth = {"mixed_f16s": L.IGAP4H_SMOOTH_F16, "mixed_bf16s": L.IGAP4H_SMOOTH_BF16, "fp64_f16s": L.IGAP4H_SMOOTH_F16}
    wanted = [f for f in args.forms.split(",") if f in forms]
    wanted_asm = [f for f in wanted if f not in smooth] or ["mixed"]  # the 16-bit forms assemble as mixed

    # ---- assembly alone (residual + tangent), HIP events on the launch stream
    asm = {}
    for name in wanted_asm:
        prec, vt, dK = forms[name]
        for _ in range(2):
            app.residual_dev(dUh.data_ptr(), dR.data_ptr(), sp)
            app.tangent_dev(dUh.data_ptr(), dK.data_ptr(), prec, vt, sp)
        torch.cuda.synchronize()
        ev = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(args.asm_steps)]
        t0 = time.perf_counter()
        for e0, e1, e2 in ev:
            e0.record(stream)
            app.residual_dev(dUh.data_ptr(), dR.data_ptr(), sp)
            e1.record(stream)
            app.tangent_dev(dUh.data_ptr(), dK.data_ptr(), prec, vt, sp)
            e2.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        r_ms = sum(a.elapsed_time(b) for a, b, _ in ev) / args.asm_steps
        k_ms = sum(b.elapsed_time(c) for _, b, c in ev) / args.asm_steps
        ms = wall * 1000.0 / args.asm_steps
        asm[name] = {"ms_per_assembly": ms, "elements_per_s": pb.n_elem / (ms / 1000.0), "residual_ms": r_ms,
                     "tangent_ms": k_ms,
                     "tangent_tflops": pb.n_elem * FLOP_KE_PER_QP * pb.nquad ** 3 / (k_ms * 1e-3) / 1e12,
                     "residual_tflops": pb.n_elem * FLOP_RE_PER_QP * pb.nquad ** 3 / (r_ms * 1e-3) / 1e12}
        say("assembly", name, json.dumps(asm[name]))

    # ---- fine-level row pass (y = K x), the traversal every Krylov/multigrid pass makes
    spmv = {}
    x = torch.tensor(np.random.default_rng(1).standard_normal(app.ndof), device=dev)
    y = torch.empty_like(x)
    for name in wanted_asm:
        prec, vt, dK = forms[name]
        app.tangent_dev(dUh.data_ptr(), dK.data_ptr(), prec, vt, sp)
        for _ in range(3):
            app.spmv_dev(dK.data_ptr(), vt, x.data_ptr(), y.data_ptr(), sp)
        nrep = 20
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(nrep):
            app.spmv_dev(dK.data_ptr(), vt, x.data_ptr(), y.data_ptr(), sp)
        e1.record(stream)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1000.0 / nrep
        vb = 8 if vt == L.IGAP4H_VALS_F64 else 4
        alg = app.nnz * vb + 3 * app.ndof * 8 + (app.ndof + 1) * 8   # values, x, y, rowptr
        spmv[name] = {"us": us, "algorithmic_bytes": alg, "GBps": alg / (us * 1e-6) / 1e9}
        say("spmv", name, json.dumps(spmv[name]))

    # ---- Newton solves (the timed step)
    newton = {}
    if not args.no_newton:
        U0 = torch.tensor(P.affine_guess(pb), device=dev, dtype=torch.float64)
        Uw = torch.empty_like(U0)
        for name in wanted:
            prec, vt, _ = forms[name]
            o = app.newton_opts(rtol=1e-10, atol=0.0, dirichlet_lift=1, tangent_prec=prec, vals_type=vt,
                                mg_smooth_vals=smooth.get(name, L.IGAP4H_SMOOTH_SAME))
            for _ in range(args.warmup):
                Uw.copy_(U0)
                app.newton_solve_dev(Uw.data_ptr(), o, sp)
            torch.cuda.synchronize()
            runs = []
            t0 = time.perf_counter()
            for _ in range(args.steps):
                Uw.copy_(U0)
                runs.append(app.newton_solve_dev(Uw.data_ptr(), o, sp))
            torch.cuda.synchronize()
            wall = (time.perf_counter() - t0) / args.steps
            info = runs[-1]
            its = info["newton_its"]
            newton[name] = {"solve_s": wall, "newton_its": its, "ms_per_newton_step": wall * 1000.0 / its,
                            "elements_per_s": pb.n_elem * its / wall, "ksp_its": info["ksp_its"],
                            "ksp_its_per_step": info["ksp_its_hist"], "converged": info["converged"],
                            "rel_fnorm": info["fnorm"] / info["fnorm0"], "fnorm": info["fnorm"],
                            "abs_1e-10_reached": bool(info["fnorm"] < 1e-10),
                            "t_tangent_s": info["t_tangent"], "t_solve_s": info["t_solve"],
                            "t_residual_s": info["t_residual"],
                            "fnorm_hist": ["%.3e" % f for f in info["fnorm_hist"]]}
            say("newton", name, json.dumps(newton[name]))

    main_asm = asm.get("fp64") or asm[wanted_asm[0]]
    main_sp = spmv.get("fp64") or spmv[wanted_asm[0]]
    main_nt = newton.get("fp64") or (newton[wanted[0]] if newton else None)
    # the tangent kernels of the chosen form (auto: the strain variation for phys/ge in both precisions)
    tstrip = args.tangent_form == "tstrip"
    k64 = "k_tangent_ge<double,2>" if tstrip else "k_tangent_sv<double,2,0>"
    k32 = "k_tangent_ge<float,4>" if tstrip else "k_tangent_sv<float,2,0>"
    tangent_kernel = k64 if "fp64" in asm else k32
    spmv_kernel = "k_spmv_line<double,0,0>" if "fp64" in spmv else "k_spmv_line<float,0,0>"
    if main_nt:
        value, ms_step = main_nt["elements_per_s"], main_nt["ms_per_newton_step"]
    else:
        value, ms_step = main_asm["elements_per_s"], main_asm["ms_per_assembly"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "phys/ge Newton solve to ||R||/||R0||<1e-10 (residual, tangent, GMRES(300) + "
                               "multigrid V-cycle with block-Jacobi Chebyshev smoothing (degree 5 on the finest "
                               "level, 1 below, interval [lmax/200, lmax]), line search), %d^3 quadratic C1 "
                               "B-spline elements, nq=4, clamp/stretch "
                               "a=0.01, Lame +-10%% per element; value = elements x Newton steps / solve time; "
                               "ms_per_step = per Newton step; steps = timed solves" % n,
                   "nelem": pb.n_elem, "ndof": app.ndof, "nnz": app.nnz, "nquad": 4, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": main_sp["GBps"], "peak": PEAK_HBM, "unit": "GB/s",
                     "frac": main_sp["GBps"] / PEAK_HBM, "traffic": profiled_traffic(spmv_kernel),
                     "kernel": spmv_kernel + " (fine-level row pass; the Chebyshev smoothing and Krylov "
                               "passes that dominate a Newton step traverse the matrix the same way)",
                     "algorithmic": "nnz x value bytes + x, y (FP64) + rowptr per launch"},
        "roofline_tangent": {"bound": "mfma", "achieved": main_asm["tangent_tflops"], "peak": PEAK["f64"],
                             "unit": "TFLOP/s", "frac": main_asm["tangent_tflops"] / PEAK["f64"],
                             "traffic": profiled_traffic(tangent_kernel), "kernel": tangent_kernel + " (27 colour launches per tangent)",
                             "algorithmic": "11,508,480 flop per element (9 x (2*10*10*27 + 2*27*10*27) per Gauss point)"},
        "roofline_residual": {"bound": "fp64 valu (latency: one wave per element)",
                              "achieved": pb.n_elem * FLOP_RE_EXEC_PER_ELEM / (main_asm["residual_ms"] * 1e-3) / 1e12,
                              "peak": PEAK["f64"], "unit": "TFLOP/s",
                              "frac": pb.n_elem * FLOP_RE_EXEC_PER_ELEM / (main_asm["residual_ms"] * 1e-3) / 1e12 / PEAK["f64"],
                              "kernel": "k_residual_elem<0> (sum-factorised)",
                              "executed": "~%d flop per element" % FLOP_RE_EXEC_PER_ELEM,
                              "algorithmic": "%d flop per element (%d per Gauss point)" % (FLOP_RE_PER_QP * pb.nquad ** 3,
                                                                                     FLOP_RE_PER_QP),
                              "algorithmic_tflops": main_asm["residual_tflops"]},
        "newton": newton,
        "assembly": asm,
        "spmv": spmv,
    }
    if "mixed" in asm:
        ex = pb.n_elem * SV_MFMA_FLOP_PER_QP["ge"] * pb.nquad ** 3 / (asm["mixed"]["tangent_ms"] * 1e-3) / 1e12
        if tstrip:
            ex = None
        line["roofline_tangent_f32"] = {"bound": "mfma", "achieved": asm["mixed"]["tangent_tflops"], "peak": PEAK["f32"],
                                        "unit": "TFLOP/s", "frac": asm["mixed"]["tangent_tflops"] / PEAK["f32"],
                                        "traffic": profiled_traffic(k32),
                                        "kernel": k32 + " (27 colour launches per tangent)",
                                        "algorithmic": "11,508,480 flop per element",
                                        "mfma_executed_tflops": ex, "mfma_executed_frac": ex / PEAK["f32"] if ex else None}
    if "fp64" in asm and not tstrip:
        ex = pb.n_elem * SV_MFMA_FLOP_PER_QP["ge"] * pb.nquad ** 3 / (asm["fp64"]["tangent_ms"] * 1e-3) / 1e12
        line["roofline_tangent"].update({"mfma_executed_tflops": ex, "mfma_executed_frac": ex / PEAK["f64"]})
    if "mixed" in asm and "fp64" in asm:
        line["assembly"]["mixed"]["roofline_frac_f32"] = asm["mixed"]["tangent_tflops"] / PEAK["f32"]
        line["assembly"]["mixed"]["tangent_speedup_vs_fp64"] = asm["fp64"]["tangent_ms"] / asm["mixed"]["tangent_ms"]
    for name in ("mixed", "mixed_f16s", "mixed_bf16s", "fp64_f16s"):
        if name in newton and "fp64" in newton:
            line["newton"][name + "_speedup_vs_fp64"] = newton["fp64"]["solve_s"] / newton[name]["solve_s"]
    line["cpu_baseline"] = cpu
    app.close()
    del dK64, dK32, dUh, dR
    if not args.no_extras:
        line["configs"] = config_legs(say)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
