"""Benchmark: EFIE assemble + solve on one MI355X.

Workload (default ``--config C3``): EFIE (Maxwell3D.singlelayer, k = 4.0) on a
unit-sphere mesh with 13,334 Fibonacci vertices (26,664 triangles, 39,996 RT
dofs), plane wave d = x, p = z.  One step = FP64 assembly of the resident
system matrix (default quadstrat) + LU of its FP32 copy on MFMA (trailing updates
from split-bf16 products by default, ``--lu fp32`` for FP32 products) +
GMRES-IR with FP64 residuals until ||b - A x|| / ||b|| <= 1e-11.  The plain FP64
GPU path (FP64 assembly + FP64 LU + getrs) is timed the same way, over warm steps
outside the mixed path's timed region; the mixed solution's relative error against
it is reported.  A timed step allocates no device memory: the resident matrix
buffers of the previous step are reused from the context's pool.

After the timed steps, a batched sample (64 requests at t = 0, N in [500, 8000])
and a Poisson serving sample (200 arrivals at 20 req/s, N up to 30,000) run
through the continuous-batching service (batched assembly, two worker contexts)
and report throughput and p50/p99 latency.

Prints one JSON line (see the keys below).  ``--gpus`` above 1 is refused:
this build runs on one GPU.
"""
import argparse
import json
import subprocess
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

METRIC = "EFIE assemble+solve sec at 1e-8 rel err (mixed vs FP64); p50/p99 serve latency"
PEAK_FP32_MFMA = 157.3e12
PEAK_FP64_MFMA = 78.6e12
PEAK_BF16_MFMA = 2.5e15

CONFIGS = {
    # name: (mesh vertices, jitter, seed, wavenumber)
    "C1": (236, 0.0, None, 2.0),
    "C2": (5804, 0.02, 2, 2 * np.pi),
    "C3": (13334, 0.0, None, 4.0),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(bj, name):
    nv, jit, seed, k = CONFIGS[name]
    mesh = bj.meshsphere(1.0, nv=nv, jitter=jit, seed=seed)
    X = bj.raviartthomas(mesh)
    t = bj.Maxwell3D.singlelayer(wavenumber=k)
    E = bj.Maxwell3D.planewave(direction=bj.x_hat, polarization=bj.z_hat, wavenumber=k)
    e = bj.cross(bj.cross(bj.n, E), bj.n)
    return mesh, X, t, e, k


def _sampled_cpu_estimate(bj, mesh, X, t, ncells, nlu):
    """Oracle assembly of ``ncells`` sampled test cells against all trial cells, scaled by
    F / ncells, plus LAPACK getrf of an nlu^2 complex matrix, scaled by (N / nlu)^3."""
    from oracle import oracle
    import scipy.linalg as sla
    F, N = mesh.numcells, X.numfunctions()
    rng = np.random.default_rng(0)
    cells = np.sort(rng.choice(F, size=min(ncells, F), replace=False))
    sub = mesh.subset(cells)
    S = len(cells)
    # all three local shapes of every sampled cell as separate test functions
    T = bj.RTBasis(sub, np.arange(0, 3 * S + 1, dtype=np.int32), np.repeat(np.arange(S), 3).astype(np.int32),
                   np.tile(np.arange(3), S).astype(np.int32), np.ones(3 * S))
    t0 = time.perf_counter()
    oracle.assemble(0, t.gamma, t.alpha, t.beta, 0, (2, 3, 6, 7, 5, 5, 4, 3), T, X)
    ta = time.perf_counter() - t0
    nlu = min(nlu, N)
    A = rng.standard_normal((nlu, nlu)) + 1j * rng.standard_normal((nlu, nlu))
    t0 = time.perf_counter()
    lu = sla.lu_factor(A, check_finite=False, overwrite_a=True)
    tl = time.perf_counter() - t0
    del lu, A
    return ta * F / S + tl * (N / nlu) ** 3, ta, tl, S, nlu


def _full_cpu_run(bj, name):
    """The whole reference path on the CPU: oracle assembly, LAPACK getrf + getrs."""
    from oracle import oracle
    import scipy.linalg as sla
    mesh, X, t, e, k = build_problem(bj, name)
    b = oracle.planewave(X, [1.0, 0.0, 0.0], [0.0, 0.0, 1.0], 1j * k)
    t0 = time.perf_counter()
    Z = oracle.assemble(0, t.gamma, t.alpha, t.beta, 0, (2, 3, 6, 7, 5, 5, 4, 3), X, X)
    ta = time.perf_counter() - t0
    t0 = time.perf_counter()
    lu = sla.lu_factor(Z, check_finite=False, overwrite_a=True)
    sla.lu_solve(lu, b, check_finite=False)
    tl = time.perf_counter() - t0
    del lu, Z
    return {"assemble_s": ta, "lu_solve_s": tl, "total_s": ta + tl, "dofs": X.numfunctions()}


def cpu_baseline(bj, mesh, X, t, k, ncells=128, nlu=6000, farfield=None, anchors=True):
    """Oracle (C/OpenMP port of the reference algorithm) + LAPACK, on a bounded sample of the
    C3 workload, extrapolated; anchored by full runs of C1 (median of 3) and C2 (once), where
    the same extrapolation is compared with the measured time.
    farfield = (directions, coefficients): also the oracle's far-field rate on a sample of
    48 of those directions, reported under "farfield"."""
    from oracle import oracle
    F, N = mesh.numcells, X.numfunctions()
    est, ta, tl, S, nl = _sampled_cpu_estimate(bj, mesh, X, t, ncells, nlu)
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    out = {"value": est, "unit": "s", "cores": threads, "kind": "port",
           "sample": (f"extrapolated, not timed in full: oracle assembly of {S} test cells x {F} trial cells "
                      f"({ta:.2f} s) scaled by {F}/{S}; LAPACK getrf of a {nl}^2 complex matrix ({tl:.2f} s) "
                      f"scaled by ({N}/{nl})^3; CPU restatement of the reference algorithm, not BEAST.jl")}
    if anchors:
        log("bench: CPU anchors (C1 x3, C2 once)")
        c1 = [_full_cpu_run(bj, "C1") for _ in range(3)]
        c1m = {key: float(np.median([r[key] for r in c1])) for key in ("assemble_s", "lu_solve_s", "total_s")}
        m2, X2, t2, _, _ = build_problem(bj, "C2")
        est2 = _sampled_cpu_estimate(bj, m2, X2, t2, ncells, nlu)[0]
        del m2, X2
        c2 = _full_cpu_run(bj, "C2")
        out["anchors"] = {
            "C1": dict(c1m, dofs=c1[0]["dofs"], runs=3, statistic="median"),
            "C2": dict(c2, runs=1, extrapolated_s=est2, extrapolation_over_measured=est2 / c2["total_s"]),
        }
    if farfield is not None:
        Y, x = farfield
        sub = Y[:: max(1, len(Y) // 48)][:48]
        t0 = time.perf_counter()
        oracle.farfield(0, 1j * k, 1.0, sub, x, X)
        tf = time.perf_counter() - t0
        out["farfield"] = {"terms_per_s": len(sub) * F * len(bj.trgauss(10)[1]) / tf,
                           "sample": f"{len(sub)} directions", "cores": threads}
    return out


def farfield_directions(nth=181, nph=361):
    th, ph = np.meshgrid(np.linspace(0, np.pi, nth), np.linspace(0, 2 * np.pi, nph), indexing="ij")
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1).reshape(-1, 3)


def farfield_sample(bj, ctx, X, x, k, nth=181, nph=361):
    """Far field of the solved current on a full (theta, phi) grid of directions
    (potential(MWFarField3D(wavenumber=k), points, x, X), default SingleNumQStrat(10)),
    outside the timed region.  Work unit: one (direction, rule point) term."""
    Y = farfield_directions(nth, nph)
    op = bj.MWFarField3D(wavenumber=k)
    bj.potential(op, Y[:256], x, X)
    ctx.profile(True)
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        ff = bj.potential(op, Y, x, X)
        walls.append(time.perf_counter() - t0)
    kern = {}
    for name in ("k_ff_current", "k_ff_partial_cells", "k_ff_partial", "k_ff_finish"):
        n_, s_, w_ = ctx.kernel_stats(name)
        if n_:
            kern[name] = {"launches": n_, "avg_ms": round(1e3 * s_ / n_, 3)}
    ctx.profile(False)
    terms = float(len(Y)) * X.geo.numcells * len(bj.trgauss(10)[1])
    wall = float(np.median(walls))
    out = {"workload": f"MWFarField3D(wavenumber={k}) of the solution at {nth}x{nph} directions, "
                       f"{X.geo.numcells} cells x 25 rule points", "directions": len(Y), "terms": terms,
           "wall_s": wall, "terms_per_s": terms / wall, "kernels": kern}
    return out


def gemm_library_comparator(shapes=((16384, 16384, 1024), (38912, 38912, 1024))):
    """The trailing update's split-bf16 products as one real GEMM through the vendor library
    (torch.addmm on ROCm dispatches to hipBLASLt): C'[M x 2N] (FP32, C's interleaved storage)
    -= A'[M x 6K] (bf16) B'[6K x 2N] (bf16), the 12 bf16 products per complex entry and k of
    k_cgemm_bf16x3 concatenated along k.  Same executed work (24 M N K) and the same FP32
    read-modify-write of C; a measured ceiling for this formulation, reported beside the
    kernel's rate.  Inputs are random (timing only)."""
    import torch
    out = []
    for (M, N, K) in shapes:
        A = torch.randn(M, 6 * K, device="cuda:0").to(torch.bfloat16)
        Bt = torch.randn(2 * N, 6 * K, device="cuda:0").to(torch.bfloat16)
        C = torch.randn(M, 2 * N, device="cuda:0", dtype=torch.float32)
        run = lambda: torch.addmm(C, A, Bt.t(), beta=1.0, alpha=-1.0, out_dtype=torch.float32)
        run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        # the same products with a bf16 result and no read of C: the cost of the FP32 C update
        torch.mm(A, Bt.t())
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            torch.mm(A, Bt.t())
        e1.record()
        torch.cuda.synchronize()
        ms_nc = e0.elapsed_time(e1) / reps
        out.append({"M": M, "N": N, "K": K, "ms": ms, "executed_tflops": 24.0 * M * N * K / (ms * 1e-3) / 1e12,
                    "algorithmic_tflops": 8.0 * M * N * K / (ms * 1e-3) / 1e12,
                    "bf16_result_no_c_read_executed_tflops": 24.0 * M * N * K / (ms_nc * 1e-3) / 1e12})
        del A, Bt, C
        torch.cuda.empty_cache()
    return {"library": "hipBLASLt through torch.addmm(out_dtype=float32)", "shapes": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C3", choices=sorted(CONFIGS))
    ap.add_argument("--tol", type=float, default=1e-11)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-anchors", action="store_true", help="skip the full C1/C2 CPU runs")
    ap.add_argument("--fp64-steps", type=int, default=3, help="timed steps of the plain FP64 GPU path")
    ap.add_argument("--lu", default="fp32split", choices=["fp32", "fp32split"],
                    help="factorisation of the mixed path")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the Poisson serving sample (0 skips serving)")
    ap.add_argument("--batch-requests", type=int, default=64, help="requests of the batched sample (0 skips)")
    ap.add_argument("--no-gemm-comparator", action="store_true",
                    help="skip timing the vendor-library GEMM of the same split-bf16 update")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three plain 256 MiB device copies (calibrates HBM byte counters) and exit")
    ap.add_argument("--assembly-only", type=int, default=0, metavar="K",
                    help="only assemble the config's resident matrix K times (for kernel counters) and exit")
    ap.add_argument("--gemm-comparator-only", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gemm_comparator_only:
        print(json.dumps(gemm_library_comparator()), flush=True)
        return
    if args.copy_calibration:
        import torch
        n = 64 << 20   # float32 elements: 256 MiB
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n,
                                               "copies": 3, "launch_order": "one fill of src, then the copies"}}),
              flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: the assembly and the LU "
              "are measured on one GPU's share of the work", file=sys.stderr)
        sys.exit(2)
    bj = __graft_entry__.load_package()
    ctx = bj.default_context()
    mesh, X, t, e, k = build_problem(bj, args.config)
    N, F = X.numfunctions(), mesh.numcells
    log(f"bench: {args.config} F={F} N={N} k={k}")
    if args.assembly_only > 0:
        times = []
        for _ in range(args.assembly_only):
            ctx.synchronize()
            ts = time.perf_counter()
            M = bj.DeviceMatrix.assemble(t, X)
            ctx.synchronize()
            times.append(time.perf_counter() - ts)
            del M
        print(json.dumps({"assembly_only": {"config": args.config, "dofs": N, "assemble_s": times}}), flush=True)
        return
    b = bj.assemble(e, X)

    # plain FP64 GPU path (also the reference solution for the error), timed like the
    # mixed path: warm steps, each assembly + FP64 LU + getrs
    def fp64_step():
        ts = time.perf_counter()
        M64 = bj.DeviceMatrix.assemble(t, X)
        ctx.synchronize()
        ta = time.perf_counter()
        M64.factor("fp64")
        xx = M64.solve(b)
        te = time.perf_counter()
        del M64
        return xx, {"assemble_s": ta - ts, "factor_solve_s": te - ta, "total_s": te - ts}

    x64, ph = fp64_step()
    log(f"bench: FP64 path warmup {ph}")
    fp64_times = []
    for i in range(args.fp64_steps):
        x64, ph = fp64_step()
        fp64_times.append(ph["total_s"])
        log(f"bench: FP64 path step {i} {ph}")
    t_fp64 = float(np.mean(fp64_times))

    def step(profile=False):
        ts = time.perf_counter()
        M = bj.DeviceMatrix.assemble(t, X)
        ctx.synchronize()
        ta = time.perf_counter()
        M.factor(args.lu)
        tf = time.perf_counter()
        x = M.solve(b, tol=args.tol, maxit=10)
        te = time.perf_counter()
        st = dict(M.stats)
        del M
        return x, {"assemble_s": ta - ts, "factor_s": tf - ta, "solve_s": te - tf, "total_s": te - ts,
                   "ir_iterations": st["iterations"], "gmres_iterations": st["inner"], "residual": st["residual"]}

    for w in range(args.warmup):
        x, ph = step()
        log(f"bench: warmup {w} {ph}")
    dom = "k_cgemm_bf16x3" if args.lu == "fp32split" else "k_cgemm"
    ctx.synchronize()
    ctx.profile(True, only=dom)   # events around the dominant kernel's launches only
    times, phases = [], []
    t_all = time.perf_counter()
    for s in range(args.steps):
        x, ph = step()
        ctx.synchronize()
        times.append(ph["total_s"])
        phases.append(ph)
        log(f"bench: step {s} {ph}")
    t_all = time.perf_counter() - t_all
    g_launch, g_sec, g_work = ctx.kernel_stats(dom)
    ctx.profile(False)
    # per-kernel breakdown from one further step, outside the timed region
    ctx.profile(True)
    step()
    ctx.synchronize()
    breakdown = {}
    for name in ("k_far_tiles", "k_classify_count", "k_sauter_schwab", "k_wilton", dom, "k_split", "k_tslu_panel",
                 "k_blockrow_trsm", "k_trinv", "k_tslu_apply", "k_laswp", "k_cgemm", "k_trsv_inv", "k_matvec",
                 "k_residual"):
        n_, s_, w_ = ctx.kernel_stats(name)
        if n_:
            breakdown[name] = {"launches": n_, "ms": round(1e3 * s_, 3)}
    f_launch, f_sec, f_work = ctx.kernel_stats("k_far_tiles")
    ctx.profile(False)
    farfield = farfield_sample(bj, ctx, X, x, k)
    err = float(np.linalg.norm(x - x64) / np.linalg.norm(x64))
    value = t_all / args.steps
    achieved = g_work / g_sec if g_sec > 0 else 0.0
    executed = None
    if dom == "k_cgemm_bf16x3":
        # the profiler counts executed bf16 MFMA flops (24 M N K); the algorithmic work of the
        # complex update is 8 M N K
        executed = achieved
        achieved = executed / 3.0
        peak = PEAK_BF16_MFMA
        work_note = ("algorithmic flops per launch = 8 M N K (complex C -= A B, 4 real products); the kernel "
                     "executes 24 M N K bf16 MFMA flops (each real product as 3 split-bf16 products), reported "
                     "as executed_tflops")
    else:
        peak = PEAK_FP32_MFMA
        work_note = "8 M N K real flops per launch (complex update, counted as 4 real products)"
    # HBM bytes per launch of the dominant kernel, from the committed counter passes
    # (FETCH_SIZE / WRITE_SIZE scaled by a copy kernel) for this workload only
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{dom}_traffic_{args.config.lower()}.json")
    if os.path.exists(tpath) and args.lu == "fp32split":
        with open(tpath) as fh:
            traffic = json.load(fh).get("traffic_bytes_per_launch")
    out = {
        "metric": METRIC, "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * value, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "c64 LU, c128 assembly/residual", "data": "synthetic",
        "config": {"workload": f"EFIE assemble+solve, unit sphere {F} triangles / {N} RT dofs, k={k}, "
                               f"mixed precision ({args.lu} LU + FP64 GMRES-IR to {args.tol:g})",
                   "config": args.config, "dofs": N, "triangles": F, "wavenumber": k},
        "rel_err_vs_fp64": err,
        "fp64_path_s": t_fp64,
        "fp64_path": {"steps": len(fp64_times), "warmup": 1, "mean_s": t_fp64,
                      "median_s": float(np.median(fp64_times)), "step_s": fp64_times},
        "speedup_vs_fp64_path": t_fp64 / value,
        "speedup_vs_fp64_path_median": float(np.median(fp64_times)) / float(np.median(times)),
        "step_s": {"mean": value, "median": float(np.median(times)), "min": float(np.min(times)),
                   "max": float(np.max(times))},
        "phases": {key: float(np.median([p[key] for p in phases])) for key in phases[0]},
        "roofline": {"kernel": dom + " (LU trailing updates)", "bound": "mfma",
                     "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic, "traffic_unit": "bytes per launch",
                     "executed_tflops": None if executed is None else executed / 1e12,
                     "executed_frac": None if executed is None else executed / peak,
                     "work": work_note,
                     "launches": g_launch // max(1, args.steps), "avg_launch_ms": 1e3 * g_sec / max(1, g_launch)},
        "kernels": dict(breakdown, k_far_tiles={"launches": f_launch, "ms": round(1e3 * f_sec, 3),
                                                "gflops_ref_formulation": f_work / max(f_sec, 1e-30) / 1e9}),
        "farfield": farfield,
    }
    if dom == "k_cgemm_bf16x3" and not args.no_gemm_comparator:
        # in a child process: torch brings its own HIP runtime, which does not initialise in a
        # process that has loaded this library's
        log("bench: vendor-library GEMM of the same update (child process)")
        try:
            cp = subprocess.run([sys.executable, os.path.abspath(__file__), "--gemm-comparator-only"],
                                capture_output=True, text=True, timeout=300)
            line = [ln for ln in cp.stdout.splitlines() if ln.startswith("{")]
            out["roofline"]["library_comparator"] = (json.loads(line[-1]) if cp.returncode == 0 and line
                                                     else {"error": (cp.stderr or "")[-200:]})
        except Exception as ex:   # the comparator is reported, not required
            out["roofline"]["library_comparator"] = {"error": str(ex)[:200]}
    if args.serve_requests > 0 or args.batch_requests > 0:
        from importlib import import_module
        sv = import_module("beast_jl_amd.serve")
        # two worker contexts (this one and one more): four HIP streams, the GPU's hardware queues
        ctx.pool(0)
        ctxs = [ctx, bj.Context(0)]
        rule = {"footprint": "24 N^2 bytes per request (FP64 matrix 16 N^2 + FP32 factors 8 N^2)",
                "budget_gb": 250.0, "hbm_gb": 288.0, "workers": len(ctxs), "streams": 2 * len(ctxs)}
        if args.batch_requests > 0:
            log(f"bench: batched sample ({args.batch_requests} requests at t=0)")
            srv = sv.Server(tol=1e-10, contexts=ctxs)
            st = srv.run(sv.batched_requests(seed=4, count=args.batch_requests), realtime=False)
            srv.close()
            out["batched"] = dict(st, admission=rule,
                                  workload="seeded mix of spheres / 2x1x1 ellipsoids / cubes, N log-uniform in "
                                           "[500, 8000], k ~ U[0.5, 8], 1-4 plane waves each, EFIE, mixed precision")
            log(f"bench: batched {st}")
        if args.serve_requests > 0:
            log(f"bench: serving sample ({args.serve_requests} Poisson arrivals at 20/s)")
            srv = sv.Server(tol=1e-10, contexts=ctxs)
            reqs = sv.poisson_requests(seed=5, rate=20.0, count=args.serve_requests, nmax=30000)
            st = srv.run(reqs)
            srv.close()
            lat = sorted((r.arrival, r.latency) for r in reqs)
            half = len(lat) // 2
            out["serving"] = dict(st, admission=rule,
                                  p50_second_half_s=float(np.percentile([l for _, l in lat[half:]], 50)),
                                  p99_second_half_s=float(np.percentile([l for _, l in lat[half:]], 99)),
                                  arrival_rate_rps=20.0,
                                  workload=f"{args.serve_requests} requests, Poisson 20/s (seed 5), geometry mix "
                                           "with N log-uniform in [500, 30000], 30% repeats on resident factors, "
                                           "EFIE or MFIE (n x K - Id/2, RT-tested) with p = 0.5, mixed precision")
            log(f"bench: serving {st}")
    if not args.no_cpu_baseline:
        log("bench: CPU baseline sample")
        out["cpu_baseline"] = cpu_baseline(bj, mesh, X, t, k, farfield=(farfield_directions(), x),
                                           anchors=not args.no_cpu_anchors)
        out["farfield"]["cpu_oracle"] = out["cpu_baseline"]["farfield"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
