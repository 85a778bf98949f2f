"""Benchmark: 256^3 Helmholtz solves/s, mixed precision vs FP64, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]

Workload: single-domain 256^3 box [0,1]^3, tanh stretch 2.0 in y, uniform x
and z, homogeneous Dirichlet walls; RHS = 8 seeded random-phase cos modes
(wavenumbers 1-16) + 1e-3 white noise.  sigma is the prediction step's
Helmholtz coefficient -Re*fac(1) (BDF2, Re=3000, dt=1e-3).  A step is one
complete solve g -> u with inputs resident in HBM.  The mixed solves (FP32
MFMA, three-part bf16 or scaled two-part fp16 MFMA transforms, FP64-residual
refinement) iterate to
a relative residual <= 1e-12; the FP64 solve is the reference algorithm in
FP64 on MFMA.

Prints one JSON line (value = the faster mixed variant's solves/s) with the
roofline of its dominant kernel (the FP64 residual, two launches per solve
and the largest single-kernel share of the solve; its average launch time
from HIP event pairs around each launch in a second pass of the same
solves), the roofline of its transform kernels (transform_roofline: the
launches of an inner solve's transforms back to back between two HIP
events) and the CPU baseline (the reference's own FP64 solve_3d, built from its sources
into oracle/_ref, on the host's cores; the numpy oracle where that build is absent; a
bounded sample of the same workload).  Extra keys: poisson_sigma0 (the same problem at sigma = 0,
FP64 and the fastest mixed variant), influence (the 4x4x2-subdomain
influence-matrix assembly, responses/s of device time) and ns_steps_per_s
(100 timed steps of the 192x256x192 channel, FP64 and fp16x2 velocity solves).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mflops3d-md_amd"))

METRIC = "Helmholtz solves/s @256³ mixed vs FP64 at ≤1e-12 rel residual; NS steps/s"
PEAK = {"fp64_mfma_tflops": 78.6, "fp32_mfma_tflops": 157.3, "bf16_mfma_tflops": 2500.0,
        "fp16_mfma_tflops": 2500.0, "hbm_gbs": 8000.0}
COPY_PROBE_BYTES = 256 << 20
# HBM bytes per launch of the transform kernels, from the FETCH_SIZE / WRITE_SIZE
# PMC runs (scaled by the copy probe); written by the profiling step
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "helmholtz256_pmc_traffic.md")


def measured_traffic(kernel_prefix):
    """HBM bytes per launch of the kernels whose name starts with
    kernel_prefix, from the JSON block of TRAFFIC_FILE; None if absent."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    start, end = text.find("```json"), text.rfind("```")
    if start < 0 or end <= start:
        return None
    data = json.loads(text[start + 7:end])
    return data.get("bytes_per_launch", {}).get(kernel_prefix)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-solves", type=int, default=0,
                    help="CPU baseline calls (0: 20 of the reference build, 2 of the numpy port)")
    ap.add_argument("--workload", choices=("helmholtz", "influence", "md", "sgs"),
                    default="helmholtz")
    ap.add_argument("--sgs-fine", type=int, nargs=3, default=(256, 256, 256))
    ap.add_argument("--sgs-coarse", type=int, nargs=3, default=(64, 64, 64))
    ap.add_argument("--sgs-steps", type=int, default=200)
    ap.add_argument("--sgs-snap", type=int, default=50)
    ap.add_argument("--sgs-train", type=int, default=300)
    ap.add_argument("--ns-steps", type=int, default=100, help="timed NS steps (0: skip)")
    ap.add_argument("--no-poisson", action="store_true", help="skip the sigma=0 solves")
    ap.add_argument("--no-influence", action="store_true",
                    help="skip the influence-matrix assembly key")
    ap.add_argument("--ns-warmup", type=int, default=3)
    ap.add_argument("--ns-shape", type=int, nargs=3, default=(192, 256, 192))
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy a 256 MiB buffer 3 times first (the known-size kernel that "
                         "scales PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "influence":
        return run_influence(args)
    if args.workload == "md":
        return run_md(args)
    if args.workload == "sgs":
        return run_sgs(args)

    import numpy as np
    import torch

    import mflops as mf
    from mflops import synthetic

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_probe:
        src = torch.ones(COPY_PROBE_BYTES // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 3 x %d bytes read and written" % COPY_PROBE_BYTES)
        del src, dst
    n = args.n
    re, dt = 3000.0, 1e-3
    sigma = -re * 1.5 / dt
    log("building %d^3 problem" % n)
    gx, gy, gz, g = synthetic.helmholtz_256_problem(n)
    sc = mf.SolverCoeffs3D(gx, gy, gz, (1,) * 6, 6)
    gd = mf.to_device(g)
    bc = mf.BoundaryCondition.zeros(n, n, n)
    u64 = mf.fortran_zeros((n, n, n))

    def run(prec, u, k):
        rep = None
        for _ in range(k):
            rep = sc.solve(gd, u, bc, sigma, prec, tol=args.tol)
        return rep

    variants = (("fp32", mf.MF_MIXED), ("bf16x3", mf.MF_MIXED_BF16X3),
                ("fp16x2", mf.MF_MIXED_FP16X2))
    umx = {name: mf.fortran_zeros((n, n, n)) for name, _ in variants}
    log("warmup")
    run(mf.MF_FP64, u64, args.warmup)
    for name, prec in variants:
        run(prec, umx[name], args.warmup)
    torch.cuda.synchronize()

    def timed(prec, u):
        """Wall time per solve over args.steps solves with no events on the
        stream, then the same solves again with an event pair around every
        kernel class (the per-kernel durations behind the roofline and the
        breakdown; the event records themselves add a few microseconds
        between kernels, so they stay out of the timed pass)."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = run(prec, u, args.steps)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        sc.profile(True)
        sc.profile_read(reset=True)
        run(prec, u, args.steps)
        torch.cuda.synchronize()
        prof = sc.profile_read(reset=True)
        sc.profile(False)
        return t, rep, prof

    t64, _, prof64 = timed(mf.MF_FP64, u64)
    log("fp64: %.3f ms/solve" % (t64 * 1e3))

    # mixed timed regions, one per inner-solve arithmetic
    res = {}
    for name, prec in variants:
        tmx, rep, prof = timed(prec, umx[name])
        res[name] = (tmx, rep, prof)
        log("mixed-%s: %.3f ms/solve, %d inner solves, residual %.2e"
            % (name, tmx * 1e3, rep.iterations, rep.residual))
    best = min(res, key=lambda k: res[k][0])
    tmx, rep, profmx = res[best]

    a64 = mf.to_host(u64)
    amx = {name: mf.to_host(u) for name, u in umx.items()}
    inner = (slice(1, -1),) * 3
    scale = float(np.abs(a64[inner]).max())
    err_mx_64 = {name: float(np.abs(a[inner] - a64[inner]).max() / scale) for name, a in amx.items()}

    # the transform kernels' own durations: the first five transforms of a
    # solve launched back to back, or for the fp16x2 plane pairs an inner
    # solve's three launches (per-launch event pairs would add the host's
    # enqueue gaps)
    kt = {name: sc.time_transforms(prec, reps=10) for name, prec in
          (("fp64", mf.MF_FP64),) + variants}
    npts = (n - 2) ** 3
    # the plane pairs ran if the fp16x2 timing covered six transforms in
    # three launches: x+y forward (8 B/point), z pair (8), y+x backward into
    # FP64 u (12)
    f_planes = 2.0 * npts * 6 * (n - 2) / 3
    planes = abs(kt["fp16x2"][1] - f_planes) < 1e-6 * f_planes

    def roof(name, peak, ops=1.0, traf=None, bytes_per_point=8.0):
        """ops: hardware MFMA flops per algorithmic flop (6 for bf16x3, 3 for
        fp16x2); traf: measured HBM bytes per launch (PMC), or None.  The
        bound is the slower of the two roofs: MFMA at peak, or HBM for one
        read and one write of the field (bytes_per_point in all)."""
        ms, flops = kt[name]
        avg = ms * 1e-3
        t_mfma = ops * flops / (peak * 1e12)
        t_hbm = bytes_per_point * npts / (PEAK["hbm_gbs"] * 1e9)
        common = {"traffic": traf, "avg_launch_ms": round(ms, 4), "flops_per_launch": flops,
                  "bytes_per_launch": bytes_per_point * npts,
                  "mfma_frac": round(t_mfma / avg, 4), "hbm_frac": round(t_hbm / avg, 4)}
        if t_mfma >= t_hbm:
            ach = ops * flops / avg / 1e12
            return dict({"bound": "mfma", "achieved": round(ach, 2), "peak": peak,
                         "unit": "TFLOP/s", "frac": round(ach / peak, 4)}, **common)
        ach = bytes_per_point * npts / avg / 1e9
        return dict({"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK["hbm_gbs"],
                     "unit": "GB/s", "frac": round(ach / PEAK["hbm_gbs"], 4)}, **common)

    if planes:
        tr = [measured_traffic(k) for k in ("h2p_fwd_kernel", "h2c_kernel", "h2p_bwd_kernel<store>")]
        traf_h2 = sum(tr) / 3 if all(t is not None for t in tr) else None
        h2 = roof("fp16x2", PEAK["fp16_mfma_tflops"], 3.0, traf=traf_h2, bytes_per_point=28.0 / 3)
        h2["kernels"] = "h2p_fwd_kernel, h2c_kernel (z pair), h2p_bwd_kernel (store), one launch each"
    else:
        h2 = roof("fp16x2", PEAK["fp16_mfma_tflops"], 3.0, traf=measured_traffic("gemm_h2_kernel"))
    roofs = {"fp32": roof("fp32", PEAK["fp32_mfma_tflops"],
                          traf=measured_traffic("gemm_mode_kernel<float")),
             "bf16x3": roof("bf16x3", PEAK["bf16_mfma_tflops"], 6.0,
                            traf=measured_traffic("gemm_bf3_kernel")),
             "fp16x2": h2}
    r64 = roof("fp64", PEAK["fp64_mfma_tflops"], traf=measured_traffic("gemm_mode_kernel<double"),
               bytes_per_point=16.0)
    # the fastest mixed solve's dominant kernel: its FP64 residual (two
    # launches per solve, the largest single-kernel share of the solve time),
    # timed by the event pairs of the profiled pass
    ms_r, cnt_r, work_r = profmx["residual_fp64"]
    res_gbs = (work_r / cnt_r) / (ms_r / cnt_r * 1e-3) / 1e9 if cnt_r else None
    res_roof = None
    if cnt_r:
        res_roof = {"kernel": "residual_zs_kernel<true>", "bound": "hbm",
                    "achieved": round(res_gbs, 1), "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                    "frac": round(res_gbs / PEAK["hbm_gbs"], 4),
                    "traffic": measured_traffic("residual_zs_kernel<true>"),
                    "avg_launch_ms": round(ms_r / cnt_r, 4), "bytes_per_launch": work_r / cnt_r,
                    "launches_per_solve": cnt_r / args.steps}

    def breakdown(p):
        return {k: {"ms_per_solve": round(v[0] / args.steps, 4), "launches_per_solve": v[1] / args.steps}
                for k, v in p.items() if v[1]}

    dtypes = {"fp32": "f32 MFMA transforms + f64 residual",
              "bf16x3": "3-part bf16 MFMA transforms (f32 accumulate) + f64 residual",
              "fp16x2": "scaled 2-part fp16 MFMA transforms (f32 accumulate) + f64 residual"}
    out = {
        "metric": METRIC,
        "value": round(1.0 / tmx, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tmx * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtypes[best],
        "data": "synthetic",
        "config": {"workload": "helmholtz-%d^3-tanh2.0-dirichlet" % n, "n": n,
                   "interior": n - 2, "sigma": sigma, "tol": args.tol, "so": 6,
                   "inner_solve": best,
                   "rhs": "8 seeded cos modes k=1..16 + 1e-3 noise (seed 2)"},
        "roofline": res_roof if res_roof else roofs[best],
        "transform_roofline": roofs[best],
        "fp64": {"value": round(1.0 / t64, 3), "unit": "solves/s", "ms_per_step": round(t64 * 1e3, 4),
                 "roofline": r64},
        "mixed_over_fp64": round(t64 / tmx, 3),
        "mixed": {name: {"value": round(1.0 / r[0], 3), "ms_per_step": round(r[0] * 1e3, 4),
                         "inner_solves": r[1].iterations, "residual": r[1].residual,
                         "residual_after_first": r[1].residual0,
                         "err_vs_fp64_gpu": err_mx_64[name], "roofline": roofs[name],
                         "breakdown": breakdown(r[2])}
                  for name, r in res.items()},
        "refinement": {"inner_solves": rep.iterations, "residual": rep.residual,
                       "residual_after_first": rep.residual0},
        "residual_kernel_gbs": round(res_gbs, 1) if res_gbs else None,
        "ns_steps_per_s": None,
    }

    if not args.no_cpu:
        out["cpu_baseline"], errs = cpu_baseline(args, gx, gy, gz, g, sigma, a64, amx[best])
        out.update(errs)
    if not args.no_poisson:
        out["poisson_sigma0"] = poisson_segment(args, sc, gd, bc, best, dict(variants)[best], inner)
    del sc, gd, u64, umx
    torch.cuda.empty_cache()
    if not args.no_influence:
        out["influence"] = influence_key(args)
    if args.ns_steps > 0:
        out["ns_steps_per_s"] = ns_segment(args)
    print(json.dumps(out), flush=True)


def poisson_segment(args, sc, gd, bc, name, prec, inner):
    """The same 256^3 problem at sigma = 0 (the Poisson operator, the worst
    conditioned case of the solver): FP64 and the fastest mixed variant, each
    args.steps timed solves after one untimed one; inner solves, residual and
    the mixed result's distance to the FP64 one."""
    import numpy as np
    import torch

    import mflops as mf

    n = gd.shape[0]
    res = {"sigma": 0.0, "inner_solve": name}
    outs = {}
    for key, p in (("fp64", mf.MF_FP64), (name, prec)):
        u = mf.fortran_zeros((n, n, n))
        sc.solve(gd, u, bc, 0.0, p, tol=args.tol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rep = sc.solve(gd, u, bc, 0.0, p, tol=args.tol)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        res[key] = {"value": round(1.0 / t, 3), "ms_per_step": round(t * 1e3, 4)}
        if p != mf.MF_FP64:
            res[key].update(inner_solves=rep.iterations, residual=rep.residual,
                            residual_after_first=rep.residual0)
        outs[key] = mf.to_host(u)
        del u
        log("poisson sigma=0 %s: %.3f ms/solve" % (key, t * 1e3))
    a, b = outs["fp64"][inner], outs[name][inner]
    res["mixed_over_fp64"] = round(res[name]["value"] / res["fp64"]["value"], 3)
    res["err_vs_fp64_gpu"] = float(np.abs(a - b).max() / np.abs(a).max())
    return res


def ns_segment(args, cpu_steps=0):
    """NS prediction-projection steps/s on the 192x256x192 channel (Re=3000,
    dt=1e-3, tanh stretch 2.5 in y): args.ns_warmup untimed then
    args.ns_steps timed steps, for FP64 velocity solves and for the scaled
    two-part fp16 mixed velocity solves (tol 1e-12); the pressure solve is FP64
    in both."""
    import numpy as np
    import torch

    import mflops as mf
    from mflops import synthetic

    nx, ny, nz = args.ns_shape
    re, dt = 3000.0, 1e-3
    log("ns: building %dx%dx%d channel" % (nx, ny, nz))
    gx, gy, gz, u, v, w, bc = synthetic.channel_problem(nx, ny, nz, re=re)
    res = {"unit": "steps/s", "config": {"workload": "channel-%dx%dx%d-tanh2.5" % (nx, ny, nz),
                                         "re": re, "dt": dt, "tou": 2, "top": 2, "nlt": 2,
                                         "pt": 1, "warmup": args.ns_warmup,
                                         "steps": args.ns_steps,
                                         "pressure_solve": "fp64"}}
    fields = [mf.to_device(f) for f in (u, v, w)]
    bcs = [mf.BoundaryCondition(*[mf.to_device(b) for b in bc[c]]) for c in range(3)]
    finals = {}
    for name, prec in (("fp64", mf.MF_FP64), ("mixed_fp16x2", mf.MF_MIXED_FP16X2)):
        s = mf.NavierStokes(gx, gy, gz, re, dt)
        for c in range(3):
            s.set_bc(c, bcs[c])
        s.set_state(*fields)
        s.set_flow_rate(True, 2.0 / re)
        s.step(args.ns_warmup, prec, 1e-12)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.step(args.ns_steps, prec, 1e-12)
        torch.cuda.synchronize()
        ts = (time.perf_counter() - t0) / args.ns_steps
        res[name] = round(1.0 / ts, 3)
        res[name + "_ms_per_step"] = round(ts * 1e3, 4)
        finals[name] = mf.to_host(s.get()[0])
        log("ns %s: %.3f ms/step" % (name, ts * 1e3))
        del s
    a, b = finals["fp64"], finals["mixed_fp16x2"]
    res["u_mixed_vs_fp64"] = float(np.abs(a - b).max() / np.abs(a).max())
    res["value"] = res["mixed_fp16x2"]
    return res


def cpu_baseline(args, gx, gy, gz, g, sigma, a64, amx):
    """The reference's own FP64 solver on the host: solve_3d
    (class_solver_3d.f90:83) compiled from the reference's sources by
    oracle/build_ref.sh (amdflang, MKL's LAPACK/BLAS on GNU OpenMP threads)
    into oracle/_ref/, timed over a bounded sample of calls on the same
    problem (the coefficient set-up is timed apart).  Without that build, the
    FP64 oracle (numpy + OpenBLAS dgemm, the reference's algorithm) instead.
    Also states the GPU results' distance to the baseline's solution."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref_solver

    n = g.shape[0]
    if ref_solver.available():
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        calls = args.cpu_solves or 20
        log("cpu baseline: reference solve_3d, %d calls on %d threads" % (calls, threads))
        u, t, t_setup = ref_solver.solve(gx, gy, gz, g, sigma, nrep=calls, threads=threads)
        log("cpu baseline: %.3f s/solve (set-up %.2f s)" % (t, t_setup))
        base = {"value": round(1.0 / t, 4), "unit": "solves/s", "cores": threads,
                "kind": "reference",
                "sample": "%d solve_3d calls of the reference on the %d^3 problem (amdflang -O3, "
                          "MKL dgemm on %d threads; solver_coeffs_init_3d %.2f s, not counted)"
                          % (calls, n, threads, t_setup)}
        key = "err_vs_reference"
    else:
        import mf_oracle as o

        try:
            from threadpoolctl import threadpool_info

            cores = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
        except Exception:  # noqa: BLE001
            cores = os.cpu_count() or 1
        calls = args.cpu_solves or 2
        log("cpu baseline: oracle setup")
        sc = o.SolverCoeffs3D(gx, gy, gz, (1,) * 6, 6)
        z = [np.zeros((n - 2, n - 2, 2), order="F") for _ in range(3)]
        u = np.zeros_like(g)
        t0 = time.perf_counter()
        for _ in range(calls):
            o.solver_3d(sc, g, u, *z, sigma)
        t = (time.perf_counter() - t0) / calls
        log("cpu baseline: %.3f s/solve on %d threads" % (t, cores))
        base = {"value": round(1.0 / t, 4), "unit": "solves/s", "cores": cores, "kind": "port",
                "sample": "%d full %d^3 FP64 solves (numpy/OpenBLAS dgemm)" % (calls, n)}
        key = "err_vs_oracle64"
    inner = (slice(1, -1),) * 3
    scale = float(np.abs(u[inner]).max())
    errs = {key: {"fp64_gpu": float(np.abs(a64[inner] - u[inner]).max() / scale),
                  "mixed_gpu": float(np.abs(amx[inner] - u[inner]).max() / scale)}}
    return base, errs


def run_sgs(args):
    """Config 5, the learned SGS closure: a fine channel run (default 256^3,
    Re 5000, tanh stretch 2.5 in y, perturbation 0.3), snapshots filtered by
    the LSQ spline onto the coarse grid (default 64^3; nodes = output points)
    give the training set (features: filtered gradients and width; targets:
    tau_ij), the MLP 10-64-64-6 is trained with full-batch Adam, and coarse
    runs with no model, Smagorinsky (C_s 0.165) and the learned closure start
    from the filtered initial field.  value = the learned closure's
    a-posteriori error (relative L2 of the velocity against the filtered fine
    run, mean over u, v, w) after the same number of steps."""
    import torch

    import mflops as mf
    from mflops import sgs

    torch.cuda.set_device(0)
    res = sgs.run(mf, fine_shape=tuple(args.sgs_fine), coarse_shape=tuple(args.sgs_coarse),
                  re=5000.0, steps=args.sgs_steps, snap_every=args.sgs_snap,
                  train_steps=args.sgs_train, log=lambda m: log("sgs: " + m))
    mean = {k: sum(v) / 3.0 for k, v in res.errors.items()}
    out = {
        "metric": "learned SGS a-posteriori error vs the filtered fine run (config 5)",
        "value": round(mean["learned"], 6),
        "unit": "relative L2 error",
        "n_gpus": 1,
        "steps": args.sgs_steps,
        "higher_is_better": False,
        "dtype": "f64 (closure f32, bf16 weights)",
        "data": "synthetic",
        "config": {"workload": "sgs-channel-%dx%dx%d-to-%dx%dx%d-re5000"
                   % (tuple(args.sgs_fine) + tuple(args.sgs_coarse)),
                   "snapshot_every": args.sgs_snap, "adam_steps": args.sgs_train,
                   "samples": res.samples, "mlp": "10-64-64-6 tanh"},
        "errors": {k: [round(e, 6) for e in v] for k, v in res.errors.items()},
        "error_mean": {k: round(v, 6) for k, v in mean.items()},
        "training_loss": [round(res.loss[0], 5), round(res.loss[1], 5)],
        "timings_s": {k: round(v, 3) for k, v in res.timings.items()},
    }
    print(json.dumps(out), flush=True)


def run_md(args):
    """Config 3 end to end: the 4x4x2 multidomain layout of run_influence
    (slab sizes from {48,64,96}, seed 1, asin grids, Dirichlet outer faces),
    sigma of the NS velocity solve.  Times the device assembly of the
    influence matrix (every unit-impulse response, scattered into one dense
    block per subdomain) and then multidomain solves of a manufactured
    problem (zero-interface solves, GMRES on the interface system to 1e-10,
    final solves); reports solves/s, GMRES iterations, the block-matvec HBM
    rate and the error against the manufactured solution."""
    import numpy as np
    import torch

    import mflops as mf
    from mflops import synthetic

    torch.cuda.set_device(0)
    sigma = -3000.0 * 1.5 / 1e-3
    slab, subs = synthetic.subdomain_layout()
    nd = tuple(len(v) for v in slab)
    grids = [[None] * nd[ax] for ax in range(3)]
    for sd in subs:
        for ax in range(3):
            grids[ax][sd["index"][ax]] = sd["grids"][ax]
    log("md: slab sizes x=%s y=%s z=%s" % tuple(slab))
    t0 = time.perf_counter()
    h = mf.Multidomain(grids, sigma)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    log("md: influence matrix assembled (%d unknowns, %.1f GB of blocks) in %.2f s"
        % (h.ninf, h.matrix_bytes / 1e9, t_asm))

    def f(X, Y, Z):
        return np.sin(np.pi * X) * np.sin(2 * np.pi * Y) * np.cos(np.pi * Z)

    lapc = -(np.pi ** 2) * (1 + 4 + 1)
    order = [(a, b, c) for c in range(nd[2]) for b in range(nd[1]) for a in range(nd[0])]
    gd, bd, exact = [], [], []
    for s in order:
        gx, gy, gz = (grids[ax][s[ax]] for ax in range(3))
        X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
        U = f(X, Y, Z)
        gd.append(mf.to_device(np.asfortranarray((lapc + sigma) * U)))
        bc = [np.stack([U[0, 1:-1, 1:-1], U[-1, 1:-1, 1:-1]], -1),
              np.stack([U[1:-1, 0, 1:-1], U[1:-1, -1, 1:-1]], -1),
              np.stack([U[1:-1, 1:-1, 0], U[1:-1, 1:-1, -1]], -1)]
        bd.append(torch.tensor(np.concatenate([b.ravel(order="F") for b in bc]), device="cuda"))
        exact.append(U)
    u = [mf.fortran_zeros(g.shape) for g in gd]
    # matvec rate (the GMRES inner product is one pass over all blocks)
    x = torch.ones(h.ninf, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    h.matvec(x, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        h.matvec(x, y)
    torch.cuda.synchronize()
    t_mv = (time.perf_counter() - t0) / 3
    log("md: matvec %.2f ms (%.0f GB/s over the blocks)" % (t_mv * 1e3, h.matrix_bytes / t_mv / 1e9))
    _, rep = h.solve(gd, bd, u=u, tol=1e-10)  # warm
    torch.cuda.synchronize()
    steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(steps):
        _, rep = h.solve(gd, bd, u=u, tol=1e-10)
    torch.cuda.synchronize()
    t_solve = (time.perf_counter() - t0) / steps
    err = max(float(np.abs(mf.to_host(ui) - e)[1:-1, 1:-1, 1:-1].max()) for ui, e in zip(u, exact))
    log("md: solve %.3f s, GMRES %d iterations, residual %.2e, max error %.2e"
        % (t_solve, rep.iterations, rep.residual, err))
    out = {
        "metric": "multidomain (influence-matrix) solves/s (config 3)",
        "value": round(1.0 / t_solve, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "md-4x4x2-{48,64,96}-seed1", "slab_sizes": slab,
                   "subdomains": len(order), "sigma": sigma, "tol": 1e-10,
                   "rhs": "manufactured sin(pi x) sin(2 pi y) cos(pi z)"},
        "ms_per_step": round(t_solve * 1e3, 2),
        "interface_unknowns": h.ninf,
        "influence_blocks_gb": round(h.matrix_bytes / 1e9, 2),
        "assembly_s": round(t_asm, 2),
        "gmres_iterations": rep.iterations,
        "gmres_residual": rep.residual,
        "matvec_ms": round(t_mv * 1e3, 3),
        "matvec_gbs": round(h.matrix_bytes / t_mv / 1e9, 1),
        "max_error_vs_manufactured": err,
    }
    print(json.dumps(out), flush=True)


def influence_assembly(reps):
    """The influence matrix of the 4x4x2 = 32-subdomain layout (per-slab
    sizes drawn from {48,64,96}, seed 1, asin-stretched FP64 grids, Dirichlet
    interfaces, sigma of the NS velocity solve), formed reps times; returns
    the layout, the response count and the fastest run's device time, flops,
    block bytes and wall time of the whole create."""
    import torch

    import mflops as mf
    from mflops import synthetic

    sigma = -3000.0 * 1.5 / 1e-3
    slab, subs = synthetic.subdomain_layout()
    nd = tuple(len(v) for v in slab)
    grids = [[None] * nd[ax] for ax in range(3)]
    for sd in subs:
        for ax in range(3):
            grids[ax][sd["index"][ax]] = sd["grids"][ax]
    log("influence: slab sizes x=%s y=%s z=%s" % tuple(slab))
    responses = 0
    for sd in subs:
        shape = tuple(len(g) for g in sd["grids"])
        responses += sum(mf.face_points(shape, f) for f in sd["interfaces"])
    times, walls = [], []
    h = None
    for _ in range(reps):
        del h
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h = mf.Multidomain(grids, sigma)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
        ms, flop, nbytes = h.assembly_stats()
        times.append(ms)
    del h
    torch.cuda.empty_cache()
    ms = min(times)
    log("influence: %d impulse responses, %.1f GB of blocks; device assembly %.1f ms (runs %s), "
        "create incl. host set-up %.2f s" % (responses, nbytes / 1e9, ms, ["%.1f" % t for t in times],
                                           min(walls)))
    return dict(sigma=sigma, slab=slab, subs=subs, responses=responses, ms=ms, flop=flop,
                nbytes=nbytes, wall=min(walls))


def influence_key(args):
    """The influence assembly's figures for the headline line (two runs)."""
    a = influence_assembly(2)
    tflops = a["flop"] / (a["ms"] * 1e-3) / 1e12
    return {"workload": "influence-4x4x2-{48,64,96}-seed1", "value": round(a["responses"] / (a["ms"] * 1e-3), 1),
            "unit": "responses/s", "responses": a["responses"], "device_ms": round(a["ms"], 2),
            "fp64_tflops": round(tflops, 2), "frac_fp64_peak": round(tflops / PEAK["fp64_mfma_tflops"], 4),
            "create_s_incl_host_setup": round(a["wall"], 3)}


def run_influence(args):
    """Influence-matrix decomposition on one GPU: 4x4x2 = 32 subdomains with
    per-slab sizes drawn from {48,64,96} (seed 1), asin-stretched FP64 grids,
    Dirichlet interfaces, sigma of the NS velocity solve.  The influence
    matrix (md_influence_matrix_set, class_field.f90:1396-1510: one unit-
    impulse response per interface unknown of every subdomain) is formed by
    spectral contraction in one HIP-graph-captured sequence whose block
    products run as persistent grouped FP64-MFMA launches.  Reports impulse
    responses per second of device time (each block column is one response),
    the grouped GEMM's FP64 rate, and, on a bounded sample (the faces of the
    largest subdomain, 4096 impulses), the rate of the per-impulse solve path
    that forms the same responses with one solve each."""
    import numpy as np
    import torch

    import mflops as mf

    torch.cuda.set_device(0)
    a = influence_assembly(max(1, min(args.steps, 3)))
    sigma, slab, subs, responses = a["sigma"], a["slab"], a["subs"], a["responses"]
    ms, flop, nbytes = a["ms"], a["flop"], a["nbytes"]
    walls = [a["wall"]]
    tflops = flop / (ms * 1e-3) / 1e12
    # the per-impulse solve path on a bounded sample: the largest subdomain
    big = max(subs, key=lambda sd: np.prod([len(g) for g in sd["grids"]]))
    gx, gy, gz = big["grids"]
    shape = (len(gx), len(gy), len(gz))
    n1, n2, n3 = (q - 2 for q in shape)
    sc = mf.SolverCoeffs3D(gx, gy, gz, (1,) * 6, 6)
    dcs = [mf.DerivativeCoefficients(g, solver=True, so=6) for g in (gx, gy, gz)]
    stride = 2 * (n2 * n3 + n1 * n3 + n1 * n2)
    cnt = 4096
    out = torch.empty((cnt, stride), device="cuda", dtype=torch.float64)
    face = big["interfaces"][0]
    mf.influence_impulses(sc, dcs, sigma, face, 0, 1, out=out[:1])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mf.influence_impulses(sc, dcs, sigma, face, 0, cnt, out=out)
    torch.cuda.synchronize()
    t_imp = time.perf_counter() - t0
    log("influence: per-impulse solve path on %s: %d impulses in %.3f s (%.0f responses/s)"
        % (shape, cnt, t_imp, cnt / t_imp))
    del sc, dcs, out
    out = {
        "metric": "influence-matrix unit-impulse responses/s (config 3)",
        "value": round(responses / (ms * 1e-3), 1),
        "unit": "responses/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "influence-4x4x2-{48,64,96}-seed1", "slab_sizes": slab,
                   "subdomains": len(subs), "sigma": sigma},
        "responses": responses,
        "device_ms": round(ms, 2),
        "create_s_incl_host_setup": round(min(walls), 3),
        "block_gb": round(nbytes / 1e9, 2),
        "grouped_gemm_flop": flop,
        "fp64_tflops_over_device_time": round(tflops, 2),
        "frac_fp64_peak": round(tflops / PEAK["fp64_mfma_tflops"], 4),
        "block_write_gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
        "per_impulse_path_responses_per_s": round(cnt / t_imp, 1),
        "per_impulse_path_sample": "%d impulses on one face of the %dx%dx%d subdomain" % ((cnt,) + shape),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
