#!/usr/bin/env python3
"""Benchmark of the radiative-transfer operator and its PCG on one MI355X.

Workload (the ball64 configuration): cubes of a 64^3 grid on [-1,1]^3 with
centroid r <= 1 (824,256 tets, 147,253 vertices), kappa = 50 exp(-r/0.2),
I_p = (1 - 0.8 r)^4, no scattering, harmonic order K = 3 (28 harmonics,
the 110-point Lebedev order), FP64.

A step is one application of the operator A (computeLhs: volume + vacuum
boundary kernels) to a device-resident vector.  value = DOF*ordinate/s =
nP * slm * steps / wall seconds of the timed loop (device-synchronised on
both sides).  Also reported: the roofline of the operator kernel (HIP
events on its stream), the FP32 operator of the mixed solver, the time to
a 1e-8 relative FP64 residual of every solver (the plain FP64 PCG is the
reference algorithm, main.cpp:125-162, run on the device; the learned
initial guess feeds the mixed solver) and the oracle timed on the host
cores.  --batch N runs the batching workload, --c3 the K = 5 scattering
workload, each as its own JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-1e-8 residual & DOF·ordinate/s, learned+mixed vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0
CHEB_DEGREE, CHEB_RATIO, CHEB_SPLIT = 6, 300.0, 2


def operator_bytes(nT, nP, nF, slm, p):
    """Algorithmic bytes of one operator application (the tet-centric
    minimum): per tet 4 u32 indices + 1/(kappa V) + 3 stored s rows
    (16 + 10p), f read and r written once (2 nP slm p), per boundary face
    3 u32 indices + its area vector's axis weight (12 + p)."""
    return nT * (16 + 10 * p) + 2 * nP * slm * p + nF * (12 + p)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, each scaled by the
    factor the same counter reads on a plain copy of known size); None when
    no summary covers this kernel."""
    fn = os.path.join(ROOT, "profiles", "traffic_%s.txt" % workload)
    if not os.path.exists(fn):
        return None
    for line in open(fn):
        parts = line.split("|")
        if len(parts) >= 3 and parts[0].strip() == kernel:
            try:
                return float(parts[-1])
            except ValueError:
                return None
    return None


def cpu_baseline(P, seconds=12.0):
    import numpy as np
    import oracle
    a = P.ang
    op = oracle.OracleProblem(P.mesh.arrays(), dict(omega=a.omega, omega_pos=a.omega_pos, Ox=a.Ox, Oy=a.Oy,
                                                    Oz=a.Oz, massfac=a.massfac, slm=a.slm))
    f = op.rhs()
    op.lhs(f)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        op.lhs(f)
        n += 1
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))
    return dict(value=op.nP * op.slm * n / dt, unit="DOF*ordinate/s", cores=cores, kind="port",
                sample="%d applications of the FP64 oracle operator (kernels.cu restated, OpenMP over "
                       "vertices) on the full ball64 mesh, %.1f s" % (n, dt))


def bench_batch(count, steps, warmup):
    """The batching workload: `count` independent slabs (n ~ U{8..32}),
    log-normal opacity, per-problem albedo, K = 2, solved together to a 1e-8
    relative FP64 residual per problem."""
    import transfer_amd as ta
    t0 = time.perf_counter()
    def progress(p):
        if (p + 1) % 32 == 0 or p + 1 == count:
            print("[bench-batch] %d/%d meshes, %.1f s" % (p + 1, count, time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    B = ta.problems.batch(count, progress=progress)
    t1 = time.perf_counter()
    ctx = ta.Batch(B.meshes, B.ang, B.albedo)
    t2 = time.perf_counter()
    nP, slm = ctx.nP, ctx.slm
    nT = sum(m.nT for m in B.meshes)
    print("[bench-batch] %d problems, %d vertices, %d tets; meshes %.1f s, assembly+upload %.1f s"
          % (count, nP, nT, t1 - t0, t2 - t1), file=sys.stderr, flush=True)
    ctx.autotune(0)
    ctx.autotune(1)
    out = dict(metric="batched time-to-1e-8 residual & DOF*ordinate/s", unit="DOF*ordinate/s", n_gpus=1,
               steps=steps, warmup=warmup, higher_is_better=True, data="synthetic", dtype="f32+f64",
               config=dict(workload="batch%d: slabs n x n x 4n, n~U{8..32}, ln kappa~N(0,1), albedo~U[0,0.95], K=2"
                                    % count, nP=nP, nT=nT, slm=slm, problems=count))
    for prec, key in ((0, "operator_fp64"), (1, "operator_fp32")):
        w, k, _, _ = ctx.bench_operator(warmup, steps, precision=prec)
        out[key] = dict(kernel=ctx.kernel_name(prec), kernel_ms=k, dof_ordinate_per_s=nP * slm * steps / w)
        print("[bench-batch] %s %s %.3f ms" % (key, ctx.kernel_name(prec), k), file=sys.stderr, flush=True)
    out["value"] = out["operator_fp32"]["dof_ordinate_per_s"]
    out["ms_per_step"] = out["operator_fp32"]["kernel_ms"]
    solves = {}
    ctx.set_chebyshev(CHEB_DEGREE, CHEB_RATIO, CHEB_SPLIT)
    for name, solver, prec in (("batched_mixed_fp32_ir_jacobi", ta.TR_SOLVER_MIXED, 2),
                               ("batched_mixed_fp32_ir_chebyshev", ta.TR_SOLVER_MIXED, 4),
                               ("batched_pcg_fp64_jacobi", ta.TR_SOLVER_PCG_FP64, 2),
                               ("batched_pcg_fp64_chebyshev", ta.TR_SOLVER_PCG_FP64, 4)):
        print("[bench-batch] solving: %s" % name, file=sys.stderr, flush=True)
        d_x = ctx.allocVector()
        st, per = ctx.solve_batch(d_x, solver=solver, use_prec=prec,
                                  maxit=20000 if solver != ta.TR_SOLVER_MIXED else 30)
        solves[name] = dict(seconds=st.seconds, max_iters=st.iters, inner_iters_total=st.preck,
                            worst_true_rel=st.true_rel, converged=st.status == 0,
                            problems_converged=sum(p["status"] == 0 for p in per))
        print("[bench-batch] %s: %.3f s, worst true rel %.2e, %d/%d converged"
              % (name, st.seconds, st.true_rel, solves[name]["problems_converged"], count),
              file=sys.stderr, flush=True)
        ctx.deviceFree(d_x)
    out["time_to_1e-8"] = solves
    print(json.dumps(out), flush=True)


def bench_sphere(n, steps, warmup):
    """The curved-boundary workload: the two-region sphere of the soltype-1
    check (problems.sphere, every boundary face off the axes), the FP64
    operator with the reference's axis rule and with general normals, at
    K = 3 and 5."""
    import transfer_amd as ta
    out = dict(metric="FP64 operator time on a curved sphere mesh, axis rule vs general normals", unit="us",
               n_gpus=1, steps=steps, warmup=warmup, higher_is_better=False, data="synthetic", dtype="f64",
               config=dict(workload="sphere%d: %d^3 Kuhn cubes mapped onto the unit ball, core r < 0.3" % (n, n)))
    for K in (3, 5):
        P = ta.problems.sphere(n, maxk=K)
        for gen in (False, True):
            ctx = ta.HipContext(P.mesh, P.ang, general_normals=gen)
            ctx.autotune(0)
            w, k, _, _ = ctx.bench_operator(warmup, steps, precision=0)
            key = "K%d_%s" % (K, "general_normals" if gen else "axis_rule")
            out[key] = dict(operator_us_per_step=w / steps * 1e6, kernel=ctx.kernel_name(0), nP=ctx.nP,
                            nF=P.mesh.nF)
            print("[bench-sphere] %s: %.1f us per application (%s)" % (key, w / steps * 1e6, ctx.kernel_name(0)),
                  file=sys.stderr, flush=True)
            ctx.close()
    print(json.dumps(out), flush=True)


def bench_c3(steps, warmup, solve=True, copy_calibration=False):
    """The high-order scattering workload: the 96^3-grid staircase ball
    (2,780,400 tets, 485,385 vertices), K = 5 (66 harmonics, the 194-point
    Lebedev order), Henyey-Greenstein g = 0.6 truncated to P_7 with albedo
    0.9, kappa = 1 + 9 U[0,1) per tet, I_p = 1.  Reports the FP64 operator,
    the FP32 operator with the BF16 matrix-core projection (the autotuned
    default) and the exact FP32 operator, each FP32 operator's error against
    the FP64 one on a random vector, and the solves to a 1e-8 relative FP64
    residual."""
    import numpy as np
    import transfer_amd as ta
    t0 = time.perf_counter()
    P = ta.problems.ball(96, maxk=5, kind="hg")
    ctx = ta.HipContext(P.mesh, P.ang)
    nP, nT, slm = P.mesh.nP, P.mesh.nT, P.ang.slm
    print("[bench-c3] %s nP=%d nT=%d slm=%d, setup %.1f s" % (P.name, nP, nT, slm, time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    ctx.autotune(0)
    if copy_calibration:
        cms = ctx.bench_copy(256 << 20, 10)
        print("[bench-c3] copy 256 MiB: %.3f ms" % cms, file=sys.stderr, flush=True)
    ctx.set_angular_precision(0)
    ctx.autotune(1)
    exact32 = ctx.schedule(1)
    ctx.set_angular_precision(3)
    ctx.autotune(1)
    out = dict(metric="C3 operator DOF*ordinate/s and time-to-1e-8", unit="DOF*ordinate/s", n_gpus=1, steps=steps,
               warmup=warmup, higher_is_better=True, data="synthetic", dtype="f64",
               config=dict(workload="ball96-hg: ball of 96^3-grid cubes, K=5 (slm 66), HG g=0.6 P7, albedo 0.9, "
                                    "kappa=1+9U", nP=nP, nT=nT, nF=P.mesh.nF, slm=slm))
    for prec, key, p in ((0, "operator_fp64", 8), (1, "operator_fp32", 4), (1, "operator_fp32_exact", 4)):
        if key == "operator_fp32_exact":
            ctx.set_angular_precision(0)
            ctx.set_schedule(1, J=exact32["J"], WPB=exact32["WPB"], variant=exact32["variant"])
        w, k, _, _ = ctx.bench_operator(warmup, steps, precision=prec)
        b = operator_bytes(nT, nP, P.mesh.nF, slm, p)
        out[key] = dict(kernel=ctx.kernel_name(prec), kernel_ms=k, dof_ordinate_per_s=nP * slm * steps / w,
                        algorithmic_bytes=b, achieved_gbs=b / (k * 1e-3) / 1e9, schedule=ctx.schedule(prec))
        print("[bench-c3] %s %s %.1f us" % (key, ctx.kernel_name(prec), k * 1e3), file=sys.stderr, flush=True)
    out["value"] = out["operator_fp64"]["dof_ordinate_per_s"]
    out["ms_per_step"] = out["operator_fp64"]["kernel_ms"]
    f, a64, a32 = ctx.allocVector(), ctx.allocVector(), ctx.allocVector()
    ctx.copyToDev(f, np.random.default_rng(1).standard_normal(ctx.N()))
    ctx.computeLhs(f, a64)
    A64 = ctx.copyToHost(a64)
    for key, aprec in (("operator_fp32_exact", 0), ("operator_fp32", 3)):
        ctx.set_angular_precision(aprec)
        if aprec == 0:
            ctx.set_schedule(1, J=exact32["J"], WPB=exact32["WPB"], variant=exact32["variant"])
        else:
            ctx.autotune(1)
        ctx.computeLhsPrec(1, f, a32)
        A32 = ctx.copyToHost(a32)
        out[key]["rel_error_vs_fp64"] = float(np.linalg.norm(A32 - A64) / np.linalg.norm(A64))
    for v in (f, a64, a32):
        ctx.deviceFree(v)
    if not solve:
        print(json.dumps(out), flush=True)
        return
    solves = {}
    # the mixed solver with the exact FP32 operator (VALU schedule) and with
    # the BF16 matrix-core projection; no Chebyshev polynomial here: it
    # amplifies the BF16 operator's rounding past what CG tolerates
    for name, solver, prec, kw in (("pcg_fp64_device", ta.TR_SOLVER_PCG_FP64, 0, {}),
                                   ("pcg_fp64_device_jacobi", ta.TR_SOLVER_PCG_FP64, 2, {}),
                                   ("mixed_fp32_ir_jacobi", ta.TR_SOLVER_MIXED, 2, dict(inner_tol=1e-5, maxit=30)),
                                   ("mixed_bf16mfma_ir_jacobi", ta.TR_SOLVER_MIXED, 2,
                                    dict(inner_tol=1e-5, maxit=30))):
        if name == "mixed_fp32_ir_jacobi":
            ctx.set_angular_precision(0)
            ctx.set_schedule(1, J=exact32["J"], WPB=exact32["WPB"], variant=exact32["variant"])
        elif name == "mixed_bf16mfma_ir_jacobi":
            ctx.set_angular_precision(3)
            ctx.autotune(1)
        d = ctx.allocVector()
        p = dict(maxit=5000, tol_abs=0.0, tol_rel=1e-8, inner_maxit=5000)
        p.update(kw)
        st, _ = ctx.solve(d, solver=solver, use_prec=prec, **p)
        solves[name] = dict(seconds=st.seconds, iters=st.iters, inner_iters=st.preck, true_rel=st.true_rel,
                            converged=st.status == 0)
        print("[bench-c3] %s: %.3f s, %d iters (%d inner), true rel %.2e" % (name, st.seconds, st.iters, st.preck,
                                                                           st.true_rel), file=sys.stderr, flush=True)
        ctx.deviceFree(d)
    out["time_to_1e-8"] = solves
    best = min(v["seconds"] for k, v in solves.items() if "mixed" in k and v["converged"])
    out["speedup_mixed_vs_plain_fp64"] = solves["pcg_fp64_device"]["seconds"] / best
    print(json.dumps(out), flush=True)


def learned_mixed(ctx, P, seeds=48, n=12, use_prec=4, inner_tol=1e-4):
    """The learned initial guess in front of the mixed solver on the bench
    mesh.  The network (8 -> 64 -> 64 -> 1, BF16 weights) is trained first on
    solver-generated cube problems (that time is reported, not counted);
    the timed solve includes building the guess (features, network and
    guess vector on the device), after one untimed warm-up call.  The solver falls back to a zero start when
    the guess's residual is not below ||b||."""
    import transfer_amd as ta
    from transfer_amd import learned
    t0 = time.perf_counter()
    feat, targ, _ = learned.generate(list(range(seeds)), n=n, maxk=P.ang.maxk, chunk=24)
    g = learned.LearnedGuess(seed=3)
    hist = g.fit(feat, targ, steps=800, batch=8192, lr=3e-3)
    train_s = time.perf_counter() - t0
    # one untimed call (first-launch code loading and the context's feature
    # scratch, as the operator's warm-up steps); the timed call below
    # recomputes features, prediction and guess from scratch
    ctx.deviceFree(g.write_guess(ctx, P.mesh, P.albedo, Ip=None))
    ctx.synchronize()
    t1 = time.perf_counter()
    d_f = g.write_guess(ctx, P.mesh, P.albedo, Ip=None)
    ctx.synchronize()
    guess_s = time.perf_counter() - t1
    st, _ = ctx.solve(d_f, solver=ta.TR_SOLVER_MIXED, use_prec=use_prec, x0=True, maxit=30, tol_abs=0.0,
                      tol_rel=1e-8, inner_maxit=20000, inner_tol=inner_tol)
    ctx.deviceFree(d_f)
    return dict(seconds=guess_s + st.seconds, guess_seconds=guess_s, solve_seconds=st.seconds, iters=st.iters,
                inner_iters=st.preck, true_rel=st.true_rel, converged=st.status == 0, fallback=st.fallback,
                training_seconds=train_s, training_samples=int(len(targ)), training_loss=float(hist[-10:].mean()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also time a plain 256 MiB device copy (counter calibration)")
    ap.add_argument("--c3", action="store_true", help="run the K = 5 scattering workload (ball96-hg) instead")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--only", default="", help="run only the named time_to_1e-8 solve (for short profiles)")
    ap.add_argument("--sphere", type=int, default=0,
                    help="run the curved sphere workload with this many cubes per axis instead (48: the measured one)")
    ap.add_argument("--batch", type=int, default=0,
                    help="run the batching workload with this many problems instead (256 = the full one)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(the multi-GPU sharding is a design note, not an implementation)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.batch:
        bench_batch(args.batch, args.steps, args.warmup)
        return
    if args.sphere:
        bench_sphere(args.sphere, min(args.steps, 50), min(args.warmup, 5))
        return
    if args.c3:
        bench_c3(min(args.steps, 50), min(args.warmup, 5), solve=not args.no_solve,
                 copy_calibration=args.copy_calibration)
        return

    import transfer_amd as ta
    P = ta.problems.ball(args.n)
    ctx = ta.HipContext(P.mesh, P.ang)
    nP, nT, slm = P.mesh.nP, P.mesh.nT, P.ang.slm
    print("[bench] %s nP=%d nT=%d nF=%d slm=%d" % (P.name, nP, nT, P.mesh.nF, slm), file=sys.stderr, flush=True)
    tune = ctx.autotune(0)
    sched = ctx.schedule(0)
    print("[bench] autotuned FP64 schedule %s" % sched, file=sys.stderr, flush=True)
    if args.copy_calibration:
        cms = ctx.bench_copy(256 << 20, 10)
        print("[bench] copy 256 MiB: %.3f ms, %.0f GB/s (read+write)" % (cms, 2 * (256 << 20) / cms / 1e6),
              file=sys.stderr, flush=True)
    wall, k_ms, k_min, k_max = ctx.bench_operator(args.warmup, args.steps, precision=0)
    value = nP * slm * args.steps / wall
    ms_per_step = wall / args.steps * 1e3
    vbytes = operator_bytes(nT, nP, P.mesh.nF, slm, 8)
    achieved = vbytes / (k_ms * 1e-3) / 1e9
    kname = ctx.kernel_name(0)
    print("[bench] FP64 operator %.1f us/step wall, kernel %s %.1f us (min %.1f, max %.1f)"
          % (ms_per_step * 1e3, kname, k_ms * 1e3, k_min * 1e3, k_max * 1e3), file=sys.stderr, flush=True)
    out = dict(metric=METRIC, value=value, unit="DOF*ordinate/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=ms_per_step, higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="%s: ball of 64^3-grid cubes, K=3 (slm 28), kappa=50exp(-r/0.2), "
                                    "I_p=(1-0.8r)^4, FP64 operator apply" % P.name,
                           nP=nP, nT=nT, nF=P.mesh.nF, slm=slm))
    out["roofline"] = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                           frac=achieved / HBM_PEAK_GBS, traffic=measured_traffic(kname, P.name),
                           kernel=kname, kernel_ms=k_ms, algorithmic_bytes=vbytes)
    out["schedule_fp64"] = sched
    # the FP32 operator of the mixed solver (BF16 matrix-core projection)
    tune32 = ctx.autotune(1)
    w32, k32, _, _ = ctx.bench_operator(args.warmup, args.steps, precision=1)
    b32 = operator_bytes(nT, nP, P.mesh.nF, slm, 4)
    out["operator_fp32"] = dict(kernel=ctx.kernel_name(1), kernel_ms=k32, dof_ordinate_per_s=nP * slm * args.steps / w32,
                                achieved_gbs=b32 / (k32 * 1e-3) / 1e9, algorithmic_bytes=b32,
                                schedule=ctx.schedule(1))
    print("[bench] FP32 operator kernel %s %.1f us" % (ctx.kernel_name(1), k32 * 1e3), file=sys.stderr, flush=True)
    if not args.no_solve:
        solves = {}
        runs = [("reference_pcg_fp64_host", ta.TR_SOLVER_REFERENCE, 0, {}),
                ("pcg_fp64_device", ta.TR_SOLVER_PCG_FP64, 0, {}),
                ("pcg_fp64_device_jacobi", ta.TR_SOLVER_PCG_FP64, 2, {}),
                ("pcg_fp64_device_multilevel", ta.TR_SOLVER_PCG_FP64, 3, {}),
                ("pcg_fp64_device_chebyshev", ta.TR_SOLVER_PCG_FP64, 4, {}),
                ("mixed_fp32_ir_jacobi", ta.TR_SOLVER_MIXED, 2, dict(inner_tol=1e-5, maxit=30)),
                ("mixed_fp32_ir_multilevel", ta.TR_SOLVER_MIXED, 3, dict(inner_tol=1e-4, maxit=30)),
                ("mixed_fp32_ir_chebyshev", ta.TR_SOLVER_MIXED, 4, dict(inner_tol=1e-4, maxit=30))]
        if args.only:
            runs = [r for r in runs if r[0] == args.only]
        # Chebyshev preconditioner: degree 6 over [lambda_max / 300, lambda_max],
        # steps with the gathered values split into BF16 hi + lo
        ctx.set_chebyshev(CHEB_DEGREE, CHEB_RATIO, CHEB_SPLIT)
        for name, solver, prec, kw in runs:
            if solver == ta.TR_SOLVER_MIXED:
                out["schedule_fp32"] = ctx.schedule(1)
            d_f = ctx.allocVector()
            p = dict(maxit=20000, tol_abs=0.0, tol_rel=1e-8, inner_maxit=20000)
            p.update(kw)
            st, _ = ctx.solve(d_f, solver=solver, use_prec=prec, **p)
            solves[name] = dict(seconds=st.seconds, iters=st.iters, inner_iters=st.preck, true_rel=st.true_rel,
                                converged=st.status == 0)
            ctx.deviceFree(d_f)
            print("[bench] %s: %.4f s, %d iters (%d inner), true rel %.2e" % (name, st.seconds, st.iters, st.preck,
                                                                            st.true_rel), file=sys.stderr, flush=True)
        if not args.no_learned:
            lm = learned_mixed(ctx, P)
            solves["learned_mixed_fp32_ir_chebyshev"] = lm
            print("[bench] learned_mixed_fp32_ir_chebyshev: %.4f s (guess %.4f s, solve %.4f s), %d inner, "
                  "fallback %d, true rel %.2e; training %.1f s on %d samples"
                  % (lm["seconds"], lm["guess_seconds"], lm["solve_seconds"], lm["inner_iters"], lm["fallback"],
                     lm["true_rel"], lm["training_seconds"], lm["training_samples"]), file=sys.stderr, flush=True)
        out["time_to_1e-8"] = solves
        if args.only:
            print(json.dumps(out), flush=True)
            return
        # the plain FP64 GPU solver: the reference's method (main.cpp's PCG
        # as configured) run on the device
        plain = solves["pcg_fp64_device"]["seconds"]
        best_plain = min(v["seconds"] for k, v in solves.items() if k.startswith(("pcg_fp64", "reference"))
                         and v["converged"])
        mixed = min((v["seconds"], k) for k, v in solves.items() if "mixed" in k and v["converged"])
        out["speedup_vs_plain_fp64"] = plain / mixed[0]
        out["speedup_vs_best_plain_fp64"] = best_plain / mixed[0]
        out["best_mixed"] = mixed[1]
        if "learned_mixed_fp32_ir_chebyshev" in solves:
            out["speedup_learned_mixed_vs_plain_fp64"] = plain / solves["learned_mixed_fp32_ir_chebyshev"]["seconds"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(P)
        print("[bench] cpu baseline %.3g DOF*ordinate/s" % out["cpu_baseline"]["value"], file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
