#!/usr/bin/env python3
"""Benchmark of the Kore solve path on one MI355X.

Default workload (the metric's configuration): serving a stream of Kore
solves.  1024 requests with Poisson arrivals, 70 % forced solves and 30 %
eigensolves (nev 6), m uniform in {0..4}, log10 E uniform in [-7, -5],
N = min(ceil(8 E^-1/3), 500) (so N in [372, 500], 0.17-0.30 M unknowns per
request), lmax = ceil(1.2 N) rounded to parity, forcing frequency w uniform
in [-2, 2] (eigen target sigma = i w), seeded unit-norm complex right-hand
sides.  Matrices are assembled in-process on host threads (the reference's
assembly restated by kore_amd.assembly) as requests arrive.

One "step" is one request served (arrival to result on the host).  Both
paths serve the same trace at the same arrival rate, 80 % of the FP64 path's
capacity (measured by a saturated calibration pass of the first requests):
  fp64   FP64 factor for every request;
  mixed  complex64 factor on FP32 MFMA + FP64 GMRES for forced solves (FP64
         refactor if it stalls); FP64 factor for eigensolves (their ~60
         operator applications would each need several complex64 sweeps).
Every result must meet residual <= 1e-10 (forced: normwise backward error;
eigen: ||(A - lambda B)x|| / (||A||_F ||x||)).

Other workloads (--workload): single eigensolves of one configuration,
timed back to back (hydro_E1e-6_N200, hydro_E1e-4_N24, hydro_E1e-7_N400,
mac_E1e-6_N150).

Prints ONE JSON line (bench contract).  Extra keys: per-path throughput,
capacity and latency, the roofline of the dominant kernel class (HIP events
on the library stream in a separate instrumented pass) and the CPU baseline
(the oracle's block-Thomas on this host's cores, and SuperLU for the eigen
workloads).
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
        label = "fp64" if not precision else ("fp64 after c64 fallback" if res.fallback else "c64+IR")
        log("  step %d (%s): %.2f ms, nconv %d, sweeps %d, restarts %d" % (
            k, label, 1e3 * lat[-1], res.nconv, res.op_calls, res.restarts))
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


def roofline(run, precision, bsz, n):
    """Roofline of the dominant kernel class of an eigensolve step.  Solve
    sweeps are priced at the block-Thomas LU's bytes (the algorithmic bytes
    of one single-column solve); the bytes the cyclic
    reduction's sweep actually streams are reported beside them
    (traffic_ratio).  Factor classes at the flops actually run."""
    t = run["timing"]
    steps = run["timing_steps"]
    f = run["finfo"]
    s = 8 if precision else 16
    alg = solve_bytes_8d(bsz, s)
    cand = {}
    cnt, ms = t["skinny"]
    if cnt and ms > 0:
        sweeps = cnt / max(1, f.solve_launches)
        gbs = sweeps * alg / (ms * 1e-3) / 1e9
        cand["skinny"] = (ms, dict(bound="hbm", achieved=gbs, peak=PEAK_HBM_GBS, unit="GB/s",
                                   kernel="skinny + offsweep (cyclic-reduction solve sweep)",
                                   algorithmic_bytes_per_sweep=alg, sweep_bytes_streamed=f.solve_bytes,
                                   traffic_ratio=f.solve_bytes / alg, launches_per_sweep=f.solve_launches,
                                   sweeps=sweeps / steps, avg_launch_ms=ms / cnt))
    peak = PEAK_FP32 if precision else PEAK_FP64
    for key, flops in (("inv", f.flops_inv), ("gemm", f.flops_gemm), ("sparse", f.flops_sparse)):
        cnt, ms = t[key]
        if cnt and ms > 0:
            tf = steps * flops / (ms * 1e-3) / 1e12
            cand[key] = (ms, dict(bound="mfma", achieved=tf, peak=peak, unit="TFLOP/s",
                                  kernel={"inv": "gj_inverse (panel, update, outer-update, permutation kernels)",
                                          "gemm": "gemm_batched_kernel",
                                          "sparse": "spcol_kernel + sprow_kernel (level 0)"}[key],
                                  avg_launch_ms=ms / cnt))
    if not cand:
        return None, {}
    dom = max(cand, key=lambda k: cand[k][0])
    r = dict(cand[dom][1])
    r["frac"] = r["achieved"] / r["peak"]
    r["traffic"] = None      # measured by rocprofv3 --pmc passes, committed under profiles/
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
    lam_cpu = None
    try:
        lam_cpu = krylov_schur.krylov_schur(op, A.shape[0], sigma, nev, tol=tol, maxit=100)[0][:nev]
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
                seconds_per_solve=t_eig, factor_s=t_fac, eigensolve_s=t_eig, host_cores=cpu_threads(),
                stand_in="scipy SuperLU + oracle Krylov-Schur, one thread (not MUMPS)", lam=lam_cpu)


def solve_bytes_8d(bsz, s=16):
    """Bytes of one single-column solve by the block-Thomas LU (the
    algorithmic bytes of the path): sum_k s (n_k^2 + n_k n_{k+1}) + 3 s sizmat."""
    bsz = [int(b) for b in bsz]
    t = sum(s * (bsz[k] ** 2 + (bsz[k] * bsz[k + 1] if k + 1 < len(bsz) else 0)) for k in range(len(bsz)))
    return float(t + 3 * s * sum(bsz))


def cpu_threads():
    return len(os.sched_getaffinity(0))


def cpu_block_thomas(par, sigma, nops=None, max_blocks=None):
    """CPU baseline (a stand-in for the sparse direct solver): the oracle's block-Thomas LU
    (numpy/LAPACK zgetrf/zgetrs/zgemm through OpenBLAS, threads = cores) on the
    request's own matrices: factor and one solve timed; an eigensolve is
    estimated as factor + nops solves.  With max_blocks, only the first
    max_blocks degree blocks of the chain are factored and solved (a bounded
    sample) and the times are scaled to all blocks: every step of the chain
    costs the same (one LU, one solve and one product of blocks)."""
    from kore_amd import assembly, backend
    from oracle import blocktri, layout
    try:
        from threadpoolctl import threadpool_limits
        lim = threadpool_limits(limits=cpu_threads())
    except Exception:
        lim = None
    A0, B0 = assembly.assemble(par, which="unscaled")
    lay = backend.layout_from_params(par)
    n, sizes, perm = backend.layout_blocks(lay)
    t0 = time.perf_counter()
    M = layout.permute(A0 - sigma * B0, perm)
    D, L, U, off = layout.block_tridiag(M, sizes)
    t1 = time.perf_counter()
    nb = len(sizes)
    kb = nb if not max_blocks else min(nb, int(max_blocks))
    bt = blocktri.BlockThomas(D[:kb], L[:kb], U[:kb])
    t2 = time.perf_counter()
    rng = np.random.default_rng(0)
    nk = int(off[kb])
    b = rng.standard_normal(nk) + 1j * rng.standard_normal(nk)
    x = bt.solve(b, off[:kb + 1])
    t3 = time.perf_counter()
    Mk = M[:nk, :nk]
    res = float(np.linalg.norm(Mk @ x - b) / (np.sqrt(np.sum(np.abs(Mk.data) ** 2)) * np.linalg.norm(x)
                                              + np.linalg.norm(b)))
    if lim is not None and hasattr(lim, "unregister"):
        lim.unregister()
    scale = nb / kb
    out = dict(factor_s=(t2 - t1) * scale, solve_s=(t3 - t2) * scale, extract_s=t1 - t0, backward_error=res,
               n=int(n), blocks=int(nb), sample_blocks=int(kb), sample_factor_s=t2 - t1, sample_solve_s=t3 - t2)
    out["forced_s"] = out["factor_s"] + out["solve_s"]
    if nops:
        out["eigen_s_est"] = out["factor_s"] + nops * out["solve_s"]
    return out


def classes_from(results):
    """Device ms and algorithmic work per kernel class, summed over
    instrumented results."""
    ms = {}
    flops = dict(inv=0.0, gemm=0.0, sparse=0.0)
    prec = set()
    for r in results:
        for k, v in r.kernel_ms.items():
            ms[k] = ms.get(k, 0.0) + v
        for key, f in r.factor_flops.items():
            prec.add(key)
            for c in flops:
                flops[c] += f[c]
    return ms, flops, prec


def serving_roofline(inst, head_name):
    """Roofline of the dominant kernel of an instrumented serving sample.
    The factorisation's batched GEMM (one kernel, gemm_batched_kernel) is
    priced per precision at the flops of the GEMMs run, over its launches
    (HIP events around each launch); the solve sweeps at the block-Thomas
    bytes.  The kernel reported is the one with the most device time among
    the single-kernel classes (GEMM per precision, solve sweep); every class's
    share is reported beside it."""
    ms, flops, precs = classes_from(inst)
    tot = sum(ms.values())
    g = {}
    for r in inst:
        keys = list(r.factor_flops)
        if len(keys) != 1:             # a fallback mixed two precisions: left out of the per-precision sums
            continue
        e = g.setdefault(keys[0], [0.0, 0, 0.0])
        e[0] += r.kernel_ms.get("gemm", 0.0)
        e[1] += r.kernel_launches.get("gemm", 0)
        e[2] += r.factor_flops[keys[0]]["gemm"]
    cands = {}
    for key, (gms, cnt, fl) in g.items():
        if gms > 0 and cnt:
            peak = PEAK_FP32 if key == "c64" else PEAK_FP64
            ach = fl / (gms * 1e-3) / 1e12
            cands["gemm_" + key] = (gms, dict(bound="mfma", achieved=ach, peak=peak, unit="TFLOP/s",
                                              kernel="gemm_batched_kernel (%s factorisations)" % key,
                                              launches=cnt, avg_launch_ms=gms / cnt,
                                              flops_per_launch=fl / cnt))
    sk = ms.get("skinny", 0.0)
    nsk = sum(r.kernel_launches.get("skinny", 0) for r in inst)
    if sk > 0:
        sb = sum(r.sweep_bytes_8d for r in inst)
        cands["skinny"] = (sk, dict(bound="hbm", achieved=sb / (sk * 1e-3) / 1e9, peak=PEAK_HBM_GBS, unit="GB/s",
                                    kernel="skinny + offsweep (solve sweep)", launches=nsk,
                                    avg_launch_ms=sk / max(1, nsk)))
    dom = max(cands, key=lambda k: cands[k][0])
    rl = dict(cands[dom][1])
    rl["frac"] = rl["achieved"] / rl["peak"]
    rl["traffic"] = None           # HBM bytes per launch: rocprofv3 --pmc passes, committed under profiles/
    rl["traffic_profile"] = {"skinny": "profiles/serving_sweep_roofline.md"}.get(
        dom, "profiles/serving_gemm_c64_roofline.md")
    rl["other_kernels"] = {k: dict(achieved=round(v[1]["achieved"], 3), unit=v[1]["unit"],
                                   frac=round(v[1]["achieved"] / v[1]["peak"], 4), launches=v[1]["launches"],
                                   avg_launch_ms=round(v[1]["avg_launch_ms"], 4))
                           for k, v in cands.items() if k != dom}
    rl.update(dominant=dom, class_device_ms={k: round(v, 3) for k, v in ms.items()},
              class_share={k: round(v / tot, 3) for k, v in ms.items()} if tot > 0 else {},
              class_flops={k: v for k, v in flops.items()},
              sample="%d requests of the trace on the %s path, one slot, HIP events around every launch" % (
                  len(inst), head_name),
              factor_precisions=sorted(precs))
    return rl


def summarize_subset(results, first):
    """Latency and throughput of the requests rid < first (a prefix of the trace)."""
    sub = [r for r in results if r.rid < first]
    lat = np.array([r.latency for r in sub])
    return dict(requests=len(sub), p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_service_s=float(np.mean([r.service for r in sub])), ok=int(sum(r.ok for r in sub)))


def roofline_sample(args, trace, head_name):
    """The instrumented pass behind the serving line's roofline: the first
    --roofline-sample requests of the trace, all queued at once, one slot,
    HIP events around every launch (outside the timed region).
    `bench.py --roofline-only` runs this pass alone (the command profiled
    with rocprofv3 for the committed kernel summaries)."""
    import dataclasses
    from kore_amd import backend, serving
    hp = backend.PREC_C64 if head_name == "mixed" else backend.PREC_FP64
    sample = [dataclasses.replace(r, arrival=0.0) for r in trace[:args.roofline_sample]]
    # an untimed pass first: the tile tuner measures at the first
    # factorisation of each block-size class, which is not the sample's work
    srv = serving.Server(slots=1, precision=hp, block=args.serving_block)
    srv.run(sample)
    srv.close()
    srv = serving.Server(slots=1, precision=hp, block=args.serving_block, instrument=True)
    inst = srv.run(sample)
    srv.close()
    by_id = {q.rid: q for q in sample}
    for r in inst:
        # single-column solve sweeps at the block-Thomas bytes (one per OP column / GMRES step)
        q = by_id[r.rid]
        _, sizes, _ = backend.layout_blocks(backend.layout_from_params(q.params()))
        r.sweep_bytes_8d = r.sweeps * solve_bytes_8d(sizes, 8 if r.precision else 16)
    rl = serving_roofline(inst, head_name)
    log("roofline: dominant %s, %.2f %s of %.1f (%.3f), %d launches of %.3f ms" % (
        rl["dominant"], rl["achieved"], rl["unit"], rl["peak"], rl["frac"], rl["launches"], rl["avg_launch_ms"]))
    return rl


def run_serving(args):
    """Serving workload (module docstring).  The trace of --requests requests
    (1024, the serving configuration) is the timed region of the headline
    (mixed) path; a "step" is one window of its arrivals, --steps windows of
    about requests/steps requests each.  The FP64 path serves the first
    --compare-requests requests of the same trace at the same arrival times
    (both full traces would not fit a ten-minute run), and the mixed path's
    figures over the same prefix are reported beside it."""
    import dataclasses
    from kore_amd import backend, serving
    e_lo, e_hi = (float(v) for v in args.e_range.split(","))
    trace = serving.serving_trace(nreq=args.requests, seed=args.seed, e_range=(e_lo, e_hi), n_cap=args.n_cap)
    log("serving: %d requests, N in [%d, %d], %d forced, %d eigen" % (
        len(trace), min(r.N for r in trace), max(r.N for r in trace), sum(r.kind == "forced" for r in trace),
        sum(r.kind == "eigen" for r in trace)))
    at0 = lambda rs: [dataclasses.replace(r, arrival=0.0) for r in rs]
    paths_req = args.paths.split(",")
    # warm-up, untimed: W requests per path (kernels loaded, memory pool
    # filled), then one request of each kind and block-size class of the
    # trace per path, so that the tile tuner has measured every (precision,
    # class) before the timed region
    reps = {}
    for r in trace:
        _, sizes, _ = backend.layout_blocks(backend.layout_from_params(r.params()))
        reps.setdefault(((int(max(sizes)) + 63) // 64, r.kind), r)
    log("warm-up: %d requests, then %d of distinct kind and block-size class, per path" % (
        max(1, args.warmup), len(reps)))
    # (the server runs the first request of each precision and class alone,
    # so that the tuner's timing sees no other request's kernels)
    for prec in (backend.PREC_FP64, backend.PREC_C64):
        srv = serving.Server(slots=args.slots, precision=prec, block=args.serving_block)
        srv.run(at0(trace[:max(1, args.warmup)]))
        srv.run(at0(list(reps.values())))
        srv.close()
    # capacity: a saturated pass (all requests queued at once) over the first
    # requests of the trace, per path
    cal = at0(trace[:min(args.calibrate, len(trace))])
    # the operator cache (repeated structures: one axpy per request after two
    # assemblies) is warmed on the calibration requests first, untimed, so
    # that both paths' capacity passes see the same cache
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(6) as ex:
        list(ex.map(serving.OPERATOR_CACHE.assemble, cal))
    capacity = {}
    for name, prec in (("fp64", backend.PREC_FP64), ("mixed", backend.PREC_C64)):
        if name not in paths_req:
            continue
        srv = serving.Server(slots=args.slots, precision=prec, block=args.serving_block)
        t0 = time.perf_counter()
        out = srv.run(cal)
        wall = time.perf_counter() - t0
        srv.close()
        capacity[name] = dict(solves_per_s=len(cal) / wall, requests=len(cal), wall_s=wall,
                              ok=int(sum(r.ok for r in out)))
        log("capacity %s: %.2f solves/s over %d requests (saturated)" % (name, len(cal) / wall, len(cal)))
    base_cap = capacity.get("fp64", next(iter(capacity.values())))["solves_per_s"]
    rate = args.load * base_cap
    rng = np.random.default_rng(args.seed + 1)
    t = 0.0
    timed = []
    for r in trace:
        t += float(rng.exponential(1.0 / rate))
        timed.append(dataclasses.replace(r, arrival=t))
    log("arrivals: Poisson at %.2f /s (%.0f %% of FP64 capacity), trace spans %.1f s" % (rate, 100 * args.load, t))
    head_name = "mixed" if "mixed" in paths_req else paths_req[0]
    order = [head_name] + [p for p in paths_req if p != head_name]
    paths, outs = {}, {}
    prefix = min(args.compare_requests, len(timed))
    for name in order:
        prec = backend.PREC_C64 if name == "mixed" else backend.PREC_FP64
        tr = timed if name == head_name else timed[:prefix]
        done = [0]

        def progress(msg, _done=done, _name=name, _n=len(tr)):
            _done[0] += 1
            if _done[0] % 64 == 0:
                log("  %s: %d / %d served" % (_name, _done[0], _n))
        srv = serving.Server(slots=args.slots, precision=prec, block=args.serving_block, log=progress)
        h0, m0 = serving.OPERATOR_CACHE.hits, serving.OPERATOR_CACHE.misses
        t0 = time.perf_counter()
        out = srv.run(tr)
        wall = time.perf_counter() - t0
        summ = serving.summarize(out, wall)
        summ.update(operator_cache=dict(hits=serving.OPERATOR_CACHE.hits - h0,
                                        full_assemblies=serving.OPERATOR_CACHE.misses - m0))
        summ.update(max_inflight=srv.max_inflight, peak_predicted_bytes=srv.mem_peak,
                    failed=[r.rid for r in out if not r.ok][:10],
                    capacity_solves_per_s=capacity.get(name, {}).get("solves_per_s"),
                    trace_requests=len(tr))
        srv.close()
        paths[name] = summ
        outs[name] = out
        for r in out:
            if not r.ok:
                log("  FAILED request %d (%s): residual %.3e, precision %d, fallback %s, error %s" % (
                    r.rid, r.kind, r.residual, r.precision, r.fallback, r.error or "-"))
        log("  %s: %d requests, %.3f solves/s, p50 %.3f s, p99 %.3f s, ok %d/%d, worst residual %.2e, "
            "max in flight %d, fallbacks %d" % (name, len(tr), summ["solves_per_s"], summ["p50_s"], summ["p99_s"],
                                                summ["ok"], summ["requests"], summ["worst_residual"],
                                                srv.max_inflight, summ["fallbacks"]))
    # like for like: every path over the common prefix of the trace
    common = {name: summarize_subset(outs[name], prefix) for name in outs}
    # roofline: an instrumented pass (HIP events around every launch) over a
    # sample of the trace on the headline path; the dominant class by device time
    rl = roofline_sample(args, trace, head_name)
    cpu = None
    if not args.no_cpu_baseline:
        req = next(r for r in trace if r.kind == "forced")
        mean_ops = paths[head_name].get("mean_eigen_op_columns") or 60.0
        log("cpu baseline: request %d (N %d, m %d) ..." % (req.rid, req.N, req.m))
        c = cpu_block_thomas(req.params(), req.sigma, nops=int(round(mean_ops)), max_blocks=args.cpu_blocks)
        nf = sum(r.kind == "forced" for r in trace)
        per_req = (nf * c["forced_s"] + (len(trace) - nf) * c["eigen_s_est"]) / len(trace)
        cpu = dict(value=1.0 / per_req, unit="solves/s", cores=cpu_threads(), kind="port",
                   sample="block-Thomas LU (oracle, OpenBLAS zgetrf/zgetrs/zgemm, %d threads) of request %d "
                          "(N %d, %d unknowns in %d degree blocks): the first %d blocks factored (%.2f s) and "
                          "solved (%.3f s), scaled to all blocks (the chain costs the same per block): factor "
                          "%.1f s, one solve %.2f s; the trace's 70 %% forced (factor + solve) and 30 %% eigen "
                          "(factor + %d solves, estimated) averaged"
                          % (cpu_threads(), req.rid, req.N, c["n"], c["blocks"], c["sample_blocks"],
                             c["sample_factor_s"], c["sample_solve_s"], c["factor_s"], c["solve_s"],
                             int(round(mean_ops))),
                   stand_in="oracle block-Thomas (not MUMPS)", detail=c)
        log("  cpu: %.4f solves/s (%s)" % (cpu["value"], cpu["sample"]))
    head = paths[head_name]
    ok = all(p["ok"] == p["requests"] for p in paths.values())
    wall = head["wall_s"]
    out = {
        "metric": METRIC, "value": head["solves_per_s"] if ok else 0.0, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / max(1, args.steps),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 factor + f64 GMRES (forced), f64 (eigen)" if head_name == "mixed" else "f64",
        "data": "synthetic: seeded request trace, Kore hydrodynamic operators assembled in-process",
        "config": {"workload": "serving %d requests, E log-uniform in [%g, %g], N = min(ceil(8 E^-1/3), %d), "
                               "lmax = ceil(1.2 N), m in 0..4, 70%% forced / 30%% eigen nev 6, Poisson arrivals "
                               "at %.0f%% of FP64 capacity" % (len(trace), e_lo, e_hi, args.n_cap, 100 * args.load),
                   "slots": args.slots, "headline_path": head_name, "arrival_rate_per_s": rate,
                   "step": "one window of the arrival trace (%d requests / %d steps); warm-up steps are single "
                           "requests per path" % (len(trace), args.steps),
                   "fp64_path_requests": prefix},
        "p50_ms": 1e3 * head["p50_s"], "p99_ms": 1e3 * head["p99_s"], "residual_ok": ok,
        "capacity": capacity, "paths": paths, "common_prefix": common, "roofline": rl, "cpu_baseline": cpu,
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="serving", choices=sorted(WORKLOADS) + ["serving"])
    ap.add_argument("--requests", type=int, default=1024)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--e-range", default="1e-7,1e-5")
    ap.add_argument("--n-cap", type=int, default=500)
    ap.add_argument("--load", type=float, default=0.8, help="arrival rate as a fraction of the FP64 capacity")
    ap.add_argument("--calibrate", type=int, default=24, help="requests of the saturated capacity passes")
    ap.add_argument("--roofline-sample", type=int, default=8)
    ap.add_argument("--compare-requests", type=int, default=256,
                    help="requests of the trace (from its start) the FP64 path serves")
    ap.add_argument("--cpu-blocks", type=int, default=96, help="degree blocks of the bounded CPU baseline sample")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--block", type=int, default=1, help="Krylov block size of the eigen workloads")
    # Krylov block size of the serving trace's eigen requests: 2 measured 5-10 %
    # faster than 1 at N = 420-500 (a sweep's cost hardly grows with a second column)
    ap.add_argument("--serving-block", type=int, default=2)
    ap.add_argument("--ncv", type=int, default=None)
    ap.add_argument("--paths", default="fp64,mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--factor-only", action="store_true",
                    help="eigen workloads: only factor the workload's system --steps times per path")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the instrumented roofline pass of the serving line (the command profiled with rocprofv3)")
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
        if args.roofline_only:
            from kore_amd import serving
            e_lo, e_hi = (float(v) for v in args.e_range.split(","))
            trace = serving.serving_trace(nreq=args.requests, seed=args.seed, e_range=(e_lo, e_hi),
                                          n_cap=args.n_cap)
            rl = roofline_sample(args, trace, "mixed" if "mixed" in args.paths.split(",") else "fp64")
            print(json.dumps({"roofline": rl}), flush=True)
            return
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
    _, sizes, _ = backend.layout_blocks(backend.layout_from_params(par))
    if args.factor_only:
        # the workload's factorisation alone, --steps times per path (the
        # command profiled for the factorisation's counters)
        out = {}
        for name in args.paths.split(","):
            prec = backend.PREC_C64 if name == "mixed" else backend.PREC_FP64
            ms = []
            for i in range(args.steps):
                fi = S.factor(wl["sigma"], prec)
                ms.append(fi.ms)
                log("  %s factorisation %d: %.2f ms" % (name, i, fi.ms))
            out[name] = dict(factor_ms=ms, flops=fi.flops, flops_inv=fi.flops_inv, flops_gemm=fi.flops_gemm,
                             flops_sparse=fi.flops_sparse, form=int(fi.form))
        print(json.dumps({"workload": args.workload, "factor_only": out}), flush=True)
        S.close()
        ctx.close()
        return
    paths = {}
    for name in args.paths.split(","):
        prec = backend.PREC_C64 if name == "mixed" else backend.PREC_FP64
        log("path %s" % name)
        r = run_path(S, ctx, wl, prec, args.steps, args.warmup, args.block, args.ncv)
        res = r["res"]
        resid = eigs.residuals(A, B, res.lam[:wl["nev"]], res.X[:, :wl["nev"]])
        rl, classes = roofline(r, prec, sizes, A.shape[0])
        paths[name] = dict(
            solves_per_s=args.steps / r["total"], ms_per_step=1e3 * r["total"] / args.steps,
            p50_ms=1e3 * pct(r["lat"], 50), p99_ms=1e3 * pct(r["lat"], 99),
            nconv=int(res.nconv), sweeps=int(res.op_calls), restarts=int(res.restarts),
            residual_max=float(resid.max()), factor_ms=float(r["finfo"].ms),
            factor_gflop=float(r["finfo"].flops / 1e9), factor_gbytes=float(r["finfo"].bytes / 1e9),
            roofline=rl, kernel_classes=classes, host_ms_per_step=r["host_ms_per_step"],
            lam=[[float(l.real), float(l.imag)] for l in res.lam[:wl["nev"]]],
            factor_precision_used="c64" if res.precision == backend.PREC_C64 else "fp64",
            fallback_to_fp64=bool(res.fallback), c64_op_columns=int(res.c64_op_columns),
            c64_ir_max_sweeps=int(res.c64_ir_max), op_columns=int(res.op_columns))
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
        lam_cpu = cpu.pop("lam")
        if lam_cpu is not None and head_name in paths:
            lg = np.array([complex(*v) for v in paths[head_name]["lam"]])
            cpu["max_rel_diff_gpu_vs_cpu"] = float(max(np.min(np.abs(lg - l)) / abs(l) for l in lam_cpu))
        log("  cpu: %.4f solves/s (%s)" % (cpu["value"], cpu["sample"]))
        if A.shape[0] <= 100000:
            bt = cpu_block_thomas(par, wl["sigma"], nops=int(paths[head_name]["op_columns"]))
            cpu["block_thomas_threads"] = dict(threads=cpu_threads(), **bt)
            log("  cpu block-Thomas (%d threads): factor %.2f s, solve %.3f s" % (cpu_threads(), bt["factor_s"],
                                                                                 bt["solve_s"]))
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
