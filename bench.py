#!/usr/bin/env python3
"""Benchmark of the Kore solve path on one MI355X.

Workload (default): the shift-and-invert eigenproblem of Kore's hydrodynamic
operator at m=1, E=1e-6, ricb=0.35, no-slip, symmetric, N=200, lmax=300
(60 000 unknowns, 300 degree blocks of 200), target sigma = -1e-3+0.5i,
nev=10, tol=1e-12.  A and B are assembled in-process (the reference's
assembly restated by kore_amd.assembly) and uploaded before timing.

One step = one eigensolve from the resident matrices: factor A - sigma B by
block cyclic reduction, run block Krylov-Schur, download the eigenpairs.
Both paths are timed in the same run: the plain FP64 factor (the headline
value: this workload is the FP64 path at scale) and the complex64 factor with
FP64 iterative refinement ("mixed", reported beside it).  Every
eigenpair must have ||(A - lambda B)x|| / (||A||_F ||x||) <= 1e-10.

Prints ONE JSON line (bench contract); extra keys: per-path throughput and
p50/p99 latency, the roofline of the dominant kernel class (HIP events on
the library stream) and the CPU baseline (the oracle, timed on this host).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s + p50/p99 latency, mixed-prec+IR vs FP64 block LU, res≤1e-10"
PEAK_HBM_GBS = 8000.0
PEAK_FP64 = 78.6
PEAK_FP32 = 157.3

WORKLOADS = {
    "hydro_E1e-6_N200": dict(par=dict(m=1, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-6, N=200, lmax=300),
                    sigma=-1e-3 + 0.5j, nev=10, tol=1e-12,
                    name="eigen m=1 E=1e-6 N=200 lmax=300 nev=10 sigma=-1e-3+0.5i"),
    "hydro_E1e-4_N24": dict(par=dict(m=1, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-4, N=24, lmax=40),
                    sigma=1.0j, nev=5, tol=1e-12,
                    name="eigen m=1 E=1e-4 N=24 lmax=40 nev=5 sigma=1i"),
    "mac_E1e-6_N150": dict(par=dict(m=2, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-6, N=150, lmax=201, magnetic=1,
                             thermal=1, Le2=1e-8, Pm=1e-3, Prandtl=1.0, Ra_gap=6e5),
                    sigma=0.1j, nev=8, tol=1e-12,
                    name="MAC eigen m=2 E=1e-6 Le=1e-4 Pm=1e-3 Pr=1 N=150 lmax=201 nev=8 sigma=0.1i"),
    "hydro_E1e-7_N400": dict(par=dict(m=1, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-7, N=400, lmax=500),
                    sigma=-1e-3 + 0.5j, nev=10, tol=1e-12,
                    name="eigen m=1 E=1e-7 N=400 lmax=500 nev=10 sigma=-1e-3+0.5i"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(x, q):
    x = np.sort(np.asarray(x))
    if len(x) == 0:
        return float("nan")
    k = (len(x) - 1) * q / 100.0
    f, c = math.floor(k), math.ceil(k)
    return float(x[f] + (x[c] - x[f]) * (k - f))


def run_path(S, ctx, wl, precision, steps, warmup, block, ncv, timing_steps=2):
    """Warm-up, then ``steps`` timed eigensolves (no instrumentation), then
    ``timing_steps`` more with HIP-event timers around every kernel (the
    per-class device times of the roofline; kept out of the timed region)."""
    from kore_amd import backend, eigs
    sigma, nev, tol = wl["sigma"], wl["nev"], wl["tol"]
    res = None
    for _ in range(warmup):
        S.factor(sigma, precision)
        res = eigs.eigs(S, sigma, nev=nev, ncv=ncv, tol=tol, block=block)
    ctx.sync()
    lat = []
    finfo = None
    host = {}
    ctx.sync()
    t_all = time.perf_counter()
    for k in range(steps):
        t0 = time.perf_counter()
        finfo = S.factor(sigma, precision)
        t1 = time.perf_counter()
        res = eigs.eigs(S, sigma, nev=nev, ncv=ncv, tol=tol, block=block)
        ctx.sync()
        lat.append(time.perf_counter() - t0)
        host["factor_call"] = host.get("factor_call", 0.0) + (t1 - t0)
        for kk, v in res.host_seconds.items():
            host[kk] = host.get(kk, 0.0) + v
        log("  step %d (%s): %.2f ms, nconv %d, sweeps %d, restarts %d" % (
            k, "c64+IR" if precision else "fp64", 1e3 * lat[-1], res.nconv, res.op_calls, res.restarts))
    ctx.sync()
    total = time.perf_counter() - t_all
    ctx.timing_reset()
    ctx.set_timing(True)
    for _ in range(timing_steps):
        S.factor(sigma, precision)
        eigs.eigs(S, sigma, nev=nev, ncv=ncv, tol=tol, block=block)
    ctx.sync()
    ctx.set_timing(False)
    timing = ctx.timing()
    return dict(lat=lat, total=total, res=res, finfo=finfo, timing=timing, timing_steps=timing_steps,
                host_ms_per_step={k: 1e3 * v / steps for k, v in host.items()})


# HBM bytes per solve-sweep dispatch measured with rocprofv3 FETCH_SIZE /
# WRITE_SIZE (scaled by a plain 1 GiB copy) on the default workload (FP64,
# Krylov block 1); see profiles/hbm_traffic_eigensolve_m1_E1e-6_N200_fp64.md.
MEASURED_SWEEP_TRAFFIC = {("hydro_E1e-6_N200", 0, 1): 40.02e6}


def roofline(run, precision, workload=None, block=None):
    t = run["timing"]
    steps = run["timing_steps"]
    f = run["finfo"]
    cand = {}
    # solve sweeps: factor bytes streamed per sweep
    cnt, ms = t["skinny"]
    if cnt and ms > 0:
        sweeps = cnt / max(1, f.solve_launches)
        gbs = sweeps * f.solve_bytes / (ms * 1e-3) / 1e9
        cand["skinny"] = (ms, dict(bound="hbm", achieved=gbs, peak=PEAK_HBM_GBS, unit="GB/s",
                                   kernel="skinny_kernel (cyclic-reduction solve sweep)",
                                   per_launch_bytes=f.solve_bytes / max(1, f.solve_launches),
                                   launches_per_sweep=f.solve_launches,
                                   avg_launch_ms=ms / cnt))
    peak = PEAK_FP32 if precision else PEAK_FP64
    for key, flops in (("inv", f.flops_inv), ("gemm", f.flops_gemm)):
        cnt, ms = t[key]
        if cnt and ms > 0:
            tf = steps * flops / (ms * 1e-3) / 1e12
            cand[key] = (ms, dict(bound="mfma", achieved=tf, peak=peak, unit="TFLOP/s",
                                  kernel="gj_inverse (panel + MFMA update kernels)" if key == "inv" else "gemm_batched_kernel",
                                  avg_launch_ms=ms / cnt))
    if not cand:
        return None, {}
    dom = max(cand, key=lambda k: cand[k][0])
    r = dict(cand[dom][1])
    r["frac"] = r["achieved"] / r["peak"]
    r["traffic"] = MEASURED_SWEEP_TRAFFIC.get((workload, precision, block)) if dom == "skinny" else None
    classes = {k: dict(launches_per_step=v[0] / steps, device_ms_per_step=v[1] / steps) for k, v in t.items()}
    return r, classes


def cpu_baseline(A, B, wl, budget_s=45.0):
    """The oracle on this host: SuperLU factor (scipy) of A - sigma B and the
    oracle's single-vector Krylov-Schur to convergence, one BLAS thread."""
    import scipy.sparse.linalg as sla
    try:
        from threadpoolctl import threadpool_limits
        lim = threadpool_limits(limits=1)
    except Exception:
        lim = None
    from oracle import krylov_schur
    sigma, nev, tol = wl["sigma"], wl["nev"], wl["tol"]
    t0 = time.perf_counter()
    lu = sla.splu((A - sigma * B).tocsc())
    t_fac = time.perf_counter() - t0
    nops = [0]
    t_op = [0.0]

    class Budget(Exception):
        pass

    t1 = time.perf_counter()

    def op(x):
        if time.perf_counter() - t1 > budget_s:
            raise Budget()
        nops[0] += 1
        return lu.solve(B @ x)

    complete = True
    try:
        krylov_schur.krylov_schur(op, A.shape[0], sigma, nev, tol=tol, maxit=100)
    except Budget:
        complete = False
    t_ks = time.perf_counter() - t1
    if lim is not None:
        lim.unregister() if hasattr(lim, "unregister") else None
    if complete:
        t_eig = t_fac + t_ks
        sample = "full eigensolve: SuperLU factor %.1f s + oracle Krylov-Schur %.1f s (%d OP applications)" % (
            t_fac, t_ks, nops[0])
    else:
        per_op = t_ks / max(1, nops[0])
        est_ops = 100
        t_eig = t_fac + per_op * est_ops
        sample = ("SuperLU factor %.1f s + %d OP applications in %.1f s, extrapolated to %d applications"
                  % (t_fac, nops[0], t_ks, est_ops))
    return dict(value=1.0 / t_eig, unit="solves/s", cores=1, kind="port", sample=sample,
                seconds_per_solve=t_eig, stand_in="scipy SuperLU (not MUMPS)")


def run_serving(args):
    """Serving workload: a seeded trace of forced (70 %) and eigen (30 %, nev 6)
    requests of uneven (m, E, N, lmax), Poisson arrivals at 80 % of the
    capacity measured on the FP64 path; both paths serve the same trace."""
    from kore_amd import backend, serving
    e_lo, e_hi = (float(v) for v in args.e_range.split(","))
    trace = serving.serving_trace(nreq=args.requests, seed=args.seed, e_range=(e_lo, e_hi), n_cap=args.n_cap)
    log("serving: %d requests, N in [%d, %d], %d forced" % (
        len(trace), min(r.N for r in trace), max(r.N for r in trace), sum(r.kind == "forced" for r in trace)))
    # calibration: mean service time of the FP64 path, one slot, a few requests
    cal = trace[:min(6, len(trace))]
    srv = serving.Server(slots=1, precision=backend.PREC_FP64, block=args.block)
    srv.run(cal[:1])                       # warm-up
    t0 = time.perf_counter()
    res = srv.run([dataclasses_replace(r, arrival=0.0) for r in cal])
    srv.close()
    mean_service = float(np.mean([r.service for r in res]))
    rate = 0.8 / mean_service
    log("calibration: mean FP64 service %.3f s -> arrival rate %.2f /s" % (mean_service, rate))
    rng = np.random.default_rng(args.seed + 1)
    t = 0.0
    timed = []
    for r in trace:
        t += float(rng.exponential(1.0 / rate))
        timed.append(dataclasses_replace(r, arrival=t))
    paths = {}
    for name in args.paths.split(","):
        prec = backend.PREC_C64 if name == "mixed" else backend.PREC_FP64
        srv = serving.Server(slots=args.slots, precision=prec, block=args.block, log=None)
        t0 = time.perf_counter()
        out = srv.run(timed)
        wall = time.perf_counter() - t0
        summ = serving.summarize(out, wall)
        summ.update(max_inflight=srv.max_inflight, peak_predicted_bytes=srv.mem_peak,
                    failed=[r.rid for r in out if not r.ok][:10])
        srv.close()
        paths[name] = summ
        log("  %s: %.3f solves/s, p50 %.3f s, p99 %.3f s, ok %d/%d, worst residual %.2e, max in flight %d" % (
            name, summ["solves_per_s"], summ["p50_s"], summ["p99_s"], summ["ok"], summ["requests"],
            summ["worst_residual"], srv.max_inflight))
    head = paths.get("mixed") or paths[next(iter(paths))]
    ok = all(p["ok"] == p["requests"] for p in paths.values())
    out = {
        "metric": METRIC, "value": head["solves_per_s"] if ok else 0.0, "unit": "solves/s", "n_gpus": 1,
        "steps": len(trace), "warmup": 1, "ms_per_step": 1e3 / max(head["solves_per_s"], 1e-12),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64+f64-IR" if "mixed" in paths else "f64",
        "data": "synthetic: seeded request trace, Kore hydrodynamic operators assembled in-process",
        "config": {"workload": "serving %d requests E in [%g, %g] N cap %d, 70%% forced / 30%% eigen nev 6, "
                               "Poisson at 80%% of FP64 capacity" % (len(trace), e_lo, e_hi, args.n_cap),
                   "slots": args.slots},
        "p50_ms": 1e3 * head["p50_s"], "p99_ms": 1e3 * head["p99_s"], "residual_ok": ok,
        "calibration": {"mean_fp64_service_s": mean_service, "arrival_rate_per_s": rate},
        "paths": paths, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def calibrate_copy(nbytes=1 << 30, reps=3):
    """A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE counter
    readings (KiB per dispatch) against nbytes give the scale that turns the
    counters of the solver kernels into HBM bytes."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read": nbytes, "bytes_written": nbytes, "reps": reps}))


def dataclasses_replace(obj, **kw):
    import dataclasses
    return dataclasses.replace(obj, **kw)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hydro_E1e-6_N200", choices=sorted(WORKLOADS) + ["serving"])
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--slots", type=int, default=2)
    ap.add_argument("--e-range", default="1e-5,1e-4")
    ap.add_argument("--n-cap", type=int, default=200)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--block", type=int, default=1)
    ap.add_argument("--ncv", type=int, default=None)
    ap.add_argument("--paths", default="fp64,mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy, three times (scale for HBM counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "serving":
        run_serving(args)
        return
    from kore_amd import assembly, backend, eigs
    from kore_amd.params import Params
    wl = WORKLOADS[args.workload]
    par = Params(**wl["par"])
    t0 = time.perf_counter()
    A, B, _ = assembly.assemble(par)
    log("assembled sizmat %d nnz(A) %d in %.1f s" % (A.shape[0], A.nnz, time.perf_counter() - t0))
    ctx = backend.Context(0)
    S = backend.System(ctx, backend.layout_from_params(par), A, B)
    paths = {}
    for name in args.paths.split(","):
        prec = backend.PREC_C64 if name == "mixed" else backend.PREC_FP64
        log("path %s" % name)
        r = run_path(S, ctx, wl, prec, args.steps, args.warmup, args.block, args.ncv)
        res = r["res"]
        resid = eigs.residuals(A, B, res.lam[:wl["nev"]], res.X[:, :wl["nev"]])
        rl, classes = roofline(r, prec, args.workload, args.block)
        paths[name] = dict(
            solves_per_s=args.steps / r["total"], ms_per_step=1e3 * r["total"] / args.steps,
            p50_ms=1e3 * pct(r["lat"], 50), p99_ms=1e3 * pct(r["lat"], 99),
            nconv=int(res.nconv), sweeps=int(res.op_calls), restarts=int(res.restarts),
            residual_max=float(resid.max()), factor_ms=float(r["finfo"].ms),
            factor_gflop=float(r["finfo"].flops / 1e9), factor_gbytes=float(r["finfo"].bytes / 1e9),
            roofline=rl, kernel_classes=classes, host_ms_per_step=r["host_ms_per_step"], lam=[[float(l.real), float(l.imag)] for l in res.lam[:wl["nev"]]])
        log("  %s: %.3f solves/s, p50 %.1f ms, p99 %.1f ms, residual %.2e" % (
            name, paths[name]["solves_per_s"], paths[name]["p50_ms"], paths[name]["p99_ms"], resid.max()))
    # parity between the two paths (relative eigenvalue difference)
    if "mixed" in paths and "fp64" in paths:
        lm = np.array([complex(*v) for v in paths["mixed"]["lam"]])
        lf = np.array([complex(*v) for v in paths["fp64"]["lam"]])
        paths["mixed"]["max_rel_diff_vs_fp64"] = float(max(np.min(np.abs(lf - l)) / abs(l) for l in lm))
    head_name = "fp64" if "fp64" in paths else next(iter(paths))
    head = paths[head_name]
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(A, B, wl)
        log("  cpu: %.4f solves/s (%s)" % (cpu["value"], cpu["sample"]))
    ok = all(p["residual_max"] <= 1e-10 and p["nconv"] >= wl["nev"] for p in paths.values())
    out = {
        "metric": METRIC,
        "value": head["solves_per_s"] if ok else 0.0,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head_name == "fp64" else "c64+f64-IR",
        "data": "synthetic: Kore hydrodynamic A, B assembled in-process from the stated parameters",
        "config": {"workload": wl["name"], "sizmat": int(A.shape[0]), "block": args.block,
                   "ncv": args.ncv, "paths": list(paths), "headline_path": head_name},
        "p50_ms": head["p50_ms"],
        "p99_ms": head["p99_ms"],
        "residual_ok": ok,
        "paths": paths,
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    S.close()
    ctx.close()


if __name__ == "__main__":
    main()
