"""Request loop for a stream of Kore solves on one MI355X (continuous batching).

A request is one Kore problem (m, E, N, lmax, ...) plus either a forced solve
at frequency w (the system A(w) = A0 - i w B0, bin/assemble.py:623 vs 445,
with a seeded unit-norm complex right-hand side) or an eigensolve with target
sigma (nev pairs nearest sigma).  Both factor A0 - sigma B0 (sigma = i w for
forced solves), so they share one kernel set.

Precision policy of the mixed path: forced solves factor in complex64 and
solve by FP64 GMRES preconditioned with that factor (kore_solve_gmres), with
an FP64 refactorisation if GMRES does not reach the tolerance; eigensolves,
which are bound by their many solve sweeps, factor in FP64 from the start.

Pipeline:

* at its arrival, a request's host work runs on the CPU assembly threads:
  its operators (A0, B0), from the operator cache when its structure repeats,
  and the host half of its system's set-up (kore_sys_prepare), so that a slot
  only uploads and computes; at most ``lookahead`` requests are prepared and
  not yet dispatched, and when a slot frees the first of them (in the order
  their host work completed, or eigen requests first with eigen_first) that
  fits the memory left is dispatched (a full assembly holds up no later
  request);
* admission: a request enters a device slot when a slot is free and its
  predicted device footprint (library estimate of the factor bytes plus the
  system and Krylov workspace) fits in the HBM budget;
* each slot owns a device context (its own HIP stream): requests in flight
  run concurrently and their kernels overlap on the device; a finished
  request frees its slot and memory at once and the next admissible request
  starts (continuous batching, not a FIFO of whole batches);
* the first request of each (factor precision, block-size class) in the
  process runs alone: the library's tile tuner times candidate tiles during
  that factorisation, and other requests' kernels would skew the timing;
* results are checked on the device from the resident matrices (forced:
  ||A x - b|| / (||A||_F ||x|| + ||b||); eigen: ||(A - lam B) x|| /
  (||A||_F ||x||)), and latency is measured from arrival to result on the
  host.
"""
from __future__ import annotations

import collections
import contextlib
import dataclasses
import math
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from typing import List, Optional

import numpy as np
import scipy.sparse as ss

from . import assembly, backend, eigs
from .params import Params

try:
    from threadpoolctl import threadpool_limits
except ImportError:                      # pragma: no cover
    threadpool_limits = None


@dataclasses.dataclass
class Request:
    rid: int
    kind: str                  # "forced" or "eigen"
    m: int
    Ek: float
    N: int
    lmax: int
    omega: float               # forcing frequency; eigen target is sigma = i*omega
    nev: int = 6
    seed: int = 0
    arrival: float = 0.0       # seconds after the trace start

    @property
    def sigma(self) -> complex:
        return 1j * self.omega

    def params(self) -> Params:
        return Params(m=self.m, symm=1, ricb=0.35, bci=1, bco=1, Ek=self.Ek, N=self.N, lmax=self.lmax)


@dataclasses.dataclass
class Result:
    rid: int
    kind: str
    latency: float
    service: float
    residual: float
    ok: bool
    lam: Optional[np.ndarray] = None
    x: Optional[np.ndarray] = None
    precision: int = 0
    ir_sweeps: int = 0
    error: str = ""
    stages: dict = dataclasses.field(default_factory=dict)   # seconds: setup, factor, solve, check
    fallback: bool = False     # the complex64 factor was replaced by an FP64 one
    n: int = 0                 # unknowns
    factor_flops: dict = dataclasses.field(default_factory=dict)   # algorithmic flops of the factorisations run
    sweeps: int = 0            # single-column solve sweeps (OP applications, GMRES steps)
    solve_bytes: float = 0.0   # bytes one sweep reads (factor blocks + level-0 couplings)
    kernel_ms: dict = dataclasses.field(default_factory=dict)      # device ms per kernel class (instrumented runs)
    kernel_launches: dict = dataclasses.field(default_factory=dict)  # launches per kernel class (instrumented runs)


def serving_trace(nreq=1024, seed=1234, e_range=(1e-7, 1e-5), n_cap=500, forced_frac=0.7, rate=None, nev=6,
                  m_values=(0, 1, 2, 3, 4)) -> List[Request]:
    """The serving workload: m ~ U{0..4}, log10 E ~ U, N = min(ceil(8 E^-1/3), cap),
    lmax = ceil(1.2 N) rounded up so that lmax - m + 1 is even, w ~ U[-2, 2],
    70 % forced / 30 % eigen, Poisson arrivals at ``rate`` (None: all at t=0)."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for i in range(nreq):
        m = int(rng.choice(m_values))
        E = float(10 ** rng.uniform(math.log10(e_range[0]), math.log10(e_range[1])))
        N = int(min(math.ceil(8 * E ** (-1.0 / 3.0)), n_cap))
        N += N % 2
        lmax = int(math.ceil(1.2 * N))
        lmax += (lmax - m + 1) % 2
        w = float(rng.uniform(-2, 2))
        kind = "forced" if rng.uniform() < forced_frac else "eigen"
        if rate:
            t += float(rng.exponential(1.0 / rate))
        out.append(Request(i, kind, m, E, N, lmax, w, nev=nev, seed=int(rng.integers(1 << 31)), arrival=t))
    return out


def rhs_for(req: Request, n: int) -> np.ndarray:
    rng = np.random.default_rng(req.seed)
    b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2.0)
    return b / np.linalg.norm(b)


def footprint(par: Params, precision: int, nev: int = 6, block: int = 1) -> int:
    """Predicted device bytes of one request (factor + system + workspace)."""
    lay = backend.layout_from_params(par)
    fb = backend.factor_bytes(lay, precision)
    n = par.sizmat
    nnz_est = 48 * n
    sysb = nnz_est * (4 + 32) + n * (8 + 4 + 8) + 4 * n * 16 * 16
    ncv = max(2 * nev, nev + 15, 96)
    wsb = n * (2 * ncv + 4 * block) * 16
    return int(fb + sysb + wsb)


def assemble_request(req: Request):
    par = req.params()
    A0, B0 = assembly.assemble(par, which="unscaled")
    return par, A0, B0


_AXPY_THREADS = 4
_AXPY_POOL = None
_AXPY_LOCK = threading.Lock()


def _affine_values(a0, a1, e):
    """a0 + e a1 over the value arrays of one pattern, split over a few
    threads for large arrays (numpy releases the interpreter lock in these
    loops): at N = 500 (10.7 million values) one thread takes 70-100 ms of
    a request's latency.  Each entry is computed as before, e a1 then + a0."""
    global _AXPY_POOL
    out = np.empty_like(a0)
    n = a0.size
    if n < (1 << 20):
        np.multiply(a1, e, out=out)
        np.add(out, a0, out=out)
        return out
    with _AXPY_LOCK:
        if _AXPY_POOL is None:
            _AXPY_POOL = ThreadPoolExecutor(max_workers=_AXPY_THREADS)
    step = (n + _AXPY_THREADS - 1) // _AXPY_THREADS

    def part(lo):
        hi = min(n, lo + step)
        np.multiply(a1[lo:hi], e, out=out[lo:hi])
        np.add(out[lo:hi], a0[lo:hi], out=out[lo:hi])

    list(_AXPY_POOL.map(part, range(0, n, step)))
    return out


@dataclasses.dataclass
class AffineA:
    """A = a0 + e a1 on the CSR pattern (indptr, indices), not formed."""
    indptr: np.ndarray
    indices: np.ndarray
    a0: np.ndarray
    a1: np.ndarray
    e: float
    shape: tuple

    def materialize(self):
        return ss.csr_matrix((_affine_values(self.a0, self.a1, self.e), self.indices, self.indptr), shape=self.shape,
                             copy=False)


class OperatorCache:
    """Operators of repeated structures: in the hydrodynamic problem of the
    serving trace the Ekman number multiplies only the viscous terms (the
    boundary rows and the Coriolis terms do not depend on it), so on one
    structure (N, m, lmax) A(E) = A0 + E A1 on a fixed pattern, and B does not
    depend on E.  The first request of a structure is assembled in full and
    kept; the second is assembled in full too, and the two give A1 = (A(E2) -
    A(E1)) / (E2 - E1) and A0 = A(E2) - E2 A1 (only if both have the same
    pattern and the same B); every later request of the structure costs one
    axpy over A's values instead of an assembly.  Least recently used
    structures are dropped beyond ``max_entries`` (each holds A0, A1, the
    pattern and B: about 0.4 GB at N = 500)."""

    def __init__(self, max_entries=12):
        self.max_entries = max_entries
        self._lock = threading.Lock()
        self._first = collections.OrderedDict()
        self._affine = collections.OrderedDict()
        self.hits = 0
        self.misses = 0

    @staticmethod
    def key(req: Request):
        return (req.N, req.m, req.lmax)

    def _trim(self, d):
        while len(d) > self.max_entries:
            d.popitem(last=False)

    def assemble(self, req: Request):
        """(par, A, B) of the request, as assemble_request gives them."""
        par, A, B = self.operators(req)
        return par, (A.materialize() if isinstance(A, AffineA) else A), B

    def operators(self, req: Request):
        """(par, A, B) with A, for a structure whose A0 and A1 are known, as
        an AffineA (A0, A1 and E on the pattern, A itself not formed)."""
        par = req.params()
        k = self.key(req)
        with self._lock:
            aff = self._affine.get(k)
            if aff is not None:
                self._affine.move_to_end(k)
            first = self._first.get(k) if aff is None else None
        if aff is not None:
            indptr, indices, a0, a1, B, shape = aff
            with self._lock:
                self.hits += 1
            return par, AffineA(indptr, indices, a0, a1, req.Ek, shape), B
        A, B = assembly.assemble(par, which="unscaled")
        A = ss.csr_matrix(A)
        B = ss.csr_matrix(B)
        with self._lock:
            self.misses += 1
            if first is None:
                self._first[k] = (req.Ek, A, B)
                self._trim(self._first)
            elif first[0] != req.Ek and np.array_equal(first[1].indptr, A.indptr) \
                    and np.array_equal(first[1].indices, A.indices) and (first[2] != B).nnz == 0:
                e1, A1 = first[0], first[1]
                a1 = (A.data - A1.data) / (req.Ek - e1)
                a0 = A.data - req.Ek * a1
                self._affine[k] = (A.indptr, A.indices, a0, a1, B, A.shape)
                self._trim(self._affine)
                self._first.pop(k, None)
        return par, A, B


OPERATOR_CACHE = OperatorCache()


def residual(A0, B0, req: Request, x=None, lam=None, X=None, b=None):
    """Host check (reference formula; used by tests):
    forced ||A x - b|| / (||A||_F ||x|| + ||b||) with A = A0 - sigma B0;
    eigen max_c ||(A0 - lam_c B0) x_c|| / (||A0||_F ||x_c||)."""
    nA = np.sqrt(np.sum(np.abs((A0 - req.sigma * B0).data) ** 2)) if req.kind == "forced" else \
        np.sqrt(np.sum(np.abs(A0.data) ** 2))
    if req.kind == "forced":
        M = A0 - req.sigma * B0
        r = M @ x - b
        return float(np.linalg.norm(r) / (nA * np.linalg.norm(x) + np.linalg.norm(b)))
    return float(eigs.residuals(A0, B0, lam, X).max()) if len(lam) else float("inf")


def residual_device(S, req: Request, x=None, lam=None, X=None, b=None):
    """The same check as ``residual`` from the matrices resident on the device
    (kore_residual, kore_sys_norm): the host only forms vector norms."""
    if req.kind == "forced":
        nA = S.norm(req.sigma)
        r = S.residual_norms(x, b=b, shift=req.sigma)[0]
        return float(r / (nA * np.linalg.norm(x) + np.linalg.norm(b)))
    if len(lam) == 0:
        return float("inf")
    nA = S.norm(0.0)
    r = S.residual_norms(X, lam=lam)
    return float(np.max(r / (nA * np.linalg.norm(X, axis=0))))


# (factor precision, block-size class of 64) pairs whose level-0 column tile
# the library has measured in this process: its tuner times candidate tiles
# at the first factorisation of each pair, so that factorisation runs with no
# other request on the device (another slot's kernels would skew the timing)
_TUNED = set()


def tuner_key(par, precision):
    _, sizes, _ = backend.layout_blocks(backend.layout_from_params(par))
    return (int(precision), (int(max(sizes)) + 63) // 64)


MAX_SLOTS = 4


class Server:
    """Continuous-batching request loop over ``slots`` device contexts."""

    def __init__(self, device=0, slots=2, mem_budget=200 << 30, precision=backend.PREC_C64, block=1,
                 assembly_workers=6, tol=1e-12, ir_tol=1e-13, keep_results=False, log=None,
                 gmres_tol=1e-15, gmres_maxit=60, instrument=False, operator_cache=OPERATOR_CACHE,
                 prepare_systems=True, lookahead=None, host_threads=None, blas_threads=1, eigen_first=False):
        # one HIP stream per slot; the process gets four hardware queues, and a
        # run with a fifth slot under the serving trace stopped making progress
        # (cause not found), so the loop takes at most MAX_SLOTS
        if not 1 <= slots <= MAX_SLOTS:
            raise ValueError("slots must be 1..%d (one stream per slot, four hardware queues)" % MAX_SLOTS)
        self.device = device
        self.slots = slots
        self.mem_budget = int(mem_budget)
        self.precision = precision
        self.block = block
        self.tol = tol
        self.ir_tol = ir_tol
        self.gmres_tol = gmres_tol
        self.gmres_maxit = gmres_maxit
        self.keep_results = keep_results
        self.log = log
        self.instrument = instrument
        self.ctxs = [backend.Context(device) for _ in range(slots)]
        if instrument:               # HIP events around every launch: per-class device times
            for c in self.ctxs:
                c.set_timing(True)
        self.free_slots = queue.Queue()
        for i in range(slots):
            self.free_slots.put(i)
        self.mem_in_use = 0
        self.mem_peak = 0
        self.max_inflight = 0
        self._inflight = 0
        self._alone = False          # a request whose tile the tuner measures runs by itself
        self._cv = threading.Condition()
        self.asm_pool = ThreadPoolExecutor(max_workers=assembly_workers)
        # operators of repeated structures from the process-wide cache (None:
        # every request assembled in full); the host half of each system's
        # set-up on the assembly threads too (prepare_systems)
        self.operator_cache = operator_cache
        self.prepare_systems = prepare_systems
        self.lookahead = lookahead if lookahead is not None else assembly_workers + slots
        # Host CPU is shared by the assembly threads, the set-ups and the slot
        # threads that launch the kernels, within the process's CPU quota.
        # While serving: `blas_threads` BLAS threads (an OpenBLAS pool keeps
        # every worker spinning for about 0.1 s after each call, even a vector
        # norm, which costs a 16-thread pool 1.5 CPU-seconds per call and
        # throttled the launching threads), and `host_threads` threads per
        # system set-up (None: the library's default, the CPU share up to 16;
        # measured at N = 500 with BLAS on one thread: 16 threads set a system
        # up in 32 ms against 80 ms with 2 under a Poisson trace, at the same
        # saturated capacity); None leaves either setting alone.
        self.host_threads = host_threads
        self.blas_threads = blas_threads
        # dispatch order among prepared requests when a slot frees: the order
        # their host work completed in, or eigen requests first
        self.eigen_first = eigen_first

    def close(self):
        self.asm_pool.shutdown(wait=True)
        for c in self.ctxs:
            c.close()

    # -- one request on one slot ------------------------------------------------
    def _prepare(self, req: Request):
        """Host work of a request, on the assembly threads ahead of a free
        slot: its operators (from the cache when the structure repeats) and
        the host half of its system's set-up (kore_sys_prepare)."""
        t0 = time.perf_counter()
        par, A0, B0 = (self.operator_cache.operators(req) if self.operator_cache is not None
                       else assemble_request(req))
        t1 = time.perf_counter()
        lay = backend.layout_from_params(par)
        plan = None
        if self.prepare_systems and isinstance(A0, AffineA):
            # a cached structure: the plan forms A0 + E A1 as it stages the values
            plan = backend.SystemPlan.affine(lay, A0.indptr, A0.indices, A0.a0, A0.a1, A0.e, B0, A0.shape)
        elif self.prepare_systems:
            plan = backend.SystemPlan(lay, A0, B0)
        elif isinstance(A0, AffineA):
            A0 = A0.materialize()
        return par, A0, B0, plan, (t1 - t0, time.perf_counter() - t1)

    def _execute(self, slot, req: Request, par, A0, B0, t0, plan=None) -> Result:
        ctx = self.ctxs[slot]
        t_start = time.perf_counter()
        st = {}
        flops = {}
        lay = backend.layout_from_params(par)
        S = backend.System(ctx, lay, A0, B0, plan=plan)
        if self.instrument:
            ctx.timing_reset()

        def factor(prec):
            fi = S.factor(req.sigma, prec)
            key = "c64" if prec == backend.PREC_C64 else "fp64"
            f = flops.setdefault(key, dict(inv=0.0, gemm=0.0, sparse=0.0, n=0))
            f["inv"] += fi.flops_inv
            f["gemm"] += fi.flops_gemm
            f["sparse"] += fi.flops_sparse
            f["n"] += 1
            return fi
        try:
            t1 = time.perf_counter()
            st["setup"] = t1 - t_start
            # shift-and-invert eigensolves are bound by the solve sweeps, which a
            # complex64 factor cannot shorten (each operator application then
            # needs refinement sweeps); they take the FP64 factor directly
            prec = self.precision if req.kind == "forced" else backend.PREC_FP64
            fi = factor(prec)
            t2 = time.perf_counter()
            st["factor"] = t2 - t1
            fallback = False
            if req.kind == "forced":
                b = rhs_for(req, S.n)
                # FP64 GMRES preconditioned by the factor: with an FP64 factor it
                # ends after the first solve; with a complex64 factor it wins back
                # FP64 accuracy.  Either way it reports the normwise backward
                # error ||b - (A - sigma B) x|| / (||A - sigma B||_F ||x|| + ||b||)
                # from the resident matrices, the check of every forced result.
                x, info = S.solve_gmres(b, tol=self.gmres_tol, maxit=self.gmres_maxit)
                sweeps = info.iterations
                if not info.converged and S.precision != backend.PREC_FP64:
                    # stalled (low Ekman number): FP64 refactorisation
                    fallback = True
                    t3 = time.perf_counter()
                    S.release_factor()
                    fi = factor(backend.PREC_FP64)
                    st["factor"] += time.perf_counter() - t3
                    x, info = S.solve_gmres(b, tol=self.gmres_tol, maxit=self.gmres_maxit)
                    sweeps += info.iterations
                t4 = time.perf_counter()
                st["solve"] = t4 - t2 - (st["factor"] - (t2 - t1))
                res = float(info.rel_correction)
                st["check"] = 0.0
                out = Result(req.rid, req.kind, 0.0, 0.0, res, bool(info.converged) and res <= 1e-10,
                             x=x if self.keep_results else None, precision=S.precision, ir_sweeps=info.iterations,
                             fallback=fallback, sweeps=sweeps)
            else:
                r = eigs.eigs(S, req.sigma, nev=req.nev, tol=self.tol, block=self.block, ir_tol=self.ir_tol)
                t4 = time.perf_counter()
                st["solve"] = t4 - t2
                k = min(req.nev, r.nconv)
                res = residual_device(S, req, lam=r.lam[:k], X=r.X[:, :k])
                st["check"] = time.perf_counter() - t4
                out = Result(req.rid, req.kind, 0.0, 0.0, res, (res <= 1e-10) and r.nconv >= req.nev,
                             lam=r.lam[:k], x=r.X[:, :k] if self.keep_results else None,
                             precision=r.precision, ir_sweeps=r.ir_max_iterations, fallback=r.fallback,
                             sweeps=r.op_columns)
            out.solve_bytes = float(S.last_factor.solve_bytes)
            if self.instrument:
                tm = ctx.timing()
                out.kernel_ms = {k: v[1] for k, v in tm.items()}
                out.kernel_launches = {k: v[0] for k, v in tm.items()}
        finally:
            S.close()
        now = time.perf_counter()
        out.service = now - t_start
        out.latency = now - (t0 + req.arrival)
        out.stages = st
        out.n = S.n
        out.factor_flops = flops
        return out

    def run(self, trace: List[Request]) -> List[Result]:
        """Serve the trace (arrival times relative to now); returns results by id."""
        blas = (threadpool_limits(limits=self.blas_threads, user_api="blas")
                if threadpool_limits is not None and self.blas_threads else contextlib.nullcontext())
        prev = backend.host_threads()
        if self.host_threads:
            backend.set_host_threads(self.host_threads)
        try:
            with blas:
                return self._run(trace)
        finally:
            backend.set_host_threads(prev)

    def _run(self, trace: List[Request]) -> List[Result]:
        t0 = time.perf_counter()
        results = {}
        lock = threading.Lock()
        order = sorted(trace, key=lambda r: r.arrival)
        # a request's host work (operators, system plan) starts at its arrival,
        # on the assembly threads; at most `lookahead` requests are prepared and
        # not yet dispatched (each holds its operators and staged system).
        # Prepared requests are dispatched in the order their host work
        # completed (eigen requests first with eigen_first), so that a full
        # assembly (a structure not yet cached: 0.5-1.6 s at N = 400-500)
        # holds up no request that arrived after it.
        ready_q = queue.Queue()
        window = threading.Semaphore(self.lookahead)

        def submitter():
            for req in order:
                dt = t0 + req.arrival - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                window.acquire()
                fut = self.asm_pool.submit(self._prepare, req)
                fut.add_done_callback(lambda f, req=req: ready_q.put((req, f)))

        sub = threading.Thread(target=submitter, daemon=True)
        sub.start()
        threads = []

        def worker(slot, req, par, A0, B0, need, plan, key, host):
            try:
                res = self._execute(slot, req, par, A0, B0, t0, plan)
            except Exception as e:  # report, keep serving
                res = Result(req.rid, req.kind, time.perf_counter() - (t0 + req.arrival), 0.0, float("inf"), False,
                             error=repr(e))
            # before the slot: the operators and the host set-up, then the
            # wait for a host worker and for a slot
            res.stages["operators"], res.stages["host_setup"] = host
            res.stages["wait"] = max(0.0, res.latency - res.service - host[0] - host[1])
            with lock:
                results[req.rid] = res
            with self._cv:
                self.mem_in_use -= need
                self._inflight -= 1
                if key is not None:
                    _TUNED.add(key)
                    self._alone = False
                self.free_slots.put(slot)
                self._cv.notify_all()
            if self.log:
                self.log("req %d %s done: latency %.3f s, service %.3f s, residual %.1e" % (
                    req.rid, req.kind, res.latency, res.service, res.residual))

        pending = []    # host work done, not dispatched: (req, par, A0, B0, plan, host, need, key)
        done = 0        # requests dispatched or reported

        def take_ready(block):
            """Move requests whose host work has finished into pending; those
            whose host work failed, or that can never fit the budget, are
            reported at once."""
            nonlocal done
            items = [ready_q.get()] if block else []
            while True:
                try:
                    items.append(ready_q.get_nowait())
                except queue.Empty:
                    break
            for req, fut in items:
                try:
                    par, A0, B0, plan, host = fut.result()
                except Exception as e:  # its host work failed: report it, keep serving
                    with lock:
                        results[req.rid] = Result(req.rid, req.kind, time.perf_counter() - (t0 + req.arrival), 0.0,
                                                  float("inf"), False, error=repr(e))
                    window.release()
                    done += 1
                    continue
                # FP64 bytes even on the mixed path: eigensolves and refinement
                # fallbacks hold an FP64 factor
                need = footprint(par, backend.PREC_FP64, req.nev, self.block)
                if need > self.mem_budget:
                    with lock:
                        results[req.rid] = Result(req.rid, req.kind, 0.0, 0.0, float("inf"), False,
                                                  error="request needs %d bytes, budget %d" % (need, self.mem_budget))
                    if plan is not None:
                        plan.close()
                    window.release()
                    done += 1
                    continue
                # the factor precision of the request's first factorisation
                key = tuner_key(par, self.precision if req.kind == "forced" else backend.PREC_FP64)
                pending.append((req, par, A0, B0, plan, host, need, key))

        def rank(item):
            # completion order of the host work, or eigen requests (the
            # longest in a slot) before forced ones
            return (0 if (self.eigen_first and item[0].kind == "eigen") else 1)

        def admissible(item):
            need, key = item[6], item[7]
            return (not self.free_slots.empty() and self.mem_in_use + need <= self.mem_budget and not self._alone
                    and (key in _TUNED or self._inflight == 0))

        while done < len(order):
            if not pending:
                take_ready(block=True)
                continue
            pick = None
            with self._cv:
                # a request of a class the tuner has not measured goes first,
                # alone; otherwise the best-ranked request that fits (stable:
                # ties keep the order the host work completed in)
                untuned = [it for it in pending if it[7] not in _TUNED]
                cands = untuned[:1] if untuned else sorted(pending, key=rank)
                pick = next((it for it in cands if admissible(it)), None)
                if pick is None:
                    self._cv.wait(timeout=0.005)    # a slot frees, or more host work completes
                else:
                    req, par, A0, B0, plan, host, need, key = pick
                    if key in _TUNED:
                        key = None
                    else:
                        self._alone = True
                    slot = self.free_slots.get()
                    self.mem_in_use += need
                    self.mem_peak = max(self.mem_peak, self.mem_in_use)
                    self._inflight += 1
                    self.max_inflight = max(self.max_inflight, self._inflight)
            if pick is None:
                take_ready(block=False)
                continue
            pending.remove(pick)
            window.release()
            done += 1
            th = threading.Thread(target=worker, args=(slot, req, par, A0, B0, need, plan, key, host), daemon=True)
            th.start()
            threads.append(th)
        sub.join()
        for th in threads:
            th.join()
        return [results[r.rid] for r in trace]


def summarize(results: List[Result], wall: float):
    lat = np.array([r.latency for r in results])
    ok = np.array([r.ok for r in results])
    stages = {}
    for r in results:
        for k, v in r.stages.items():
            stages[k] = stages.get(k, 0.0) + v / len(results)
    forced = [r for r in results if r.kind == "forced"]
    eigen = [r for r in results if r.kind == "eigen"]
    return dict(requests=len(results), ok=int(ok.sum()), solves_per_s=len(results) / wall, wall_s=wall,
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_service_s=float(np.mean([r.service for r in results])),
                mean_stage_s=stages,
                fp64_results=int(sum(r.precision == backend.PREC_FP64 for r in results)),
                c64_results=int(sum(r.precision == backend.PREC_C64 for r in results)),
                fallbacks=int(sum(r.fallback for r in results)),
                mean_forced_sweeps=float(np.mean([r.sweeps for r in forced])) if forced else 0.0,
                mean_eigen_op_columns=float(np.mean([r.sweeps for r in eigen])) if eigen else 0.0,
                forced_p50_s=float(np.percentile([r.latency for r in forced], 50)) if forced else 0.0,
                eigen_p50_s=float(np.percentile([r.latency for r in eigen], 50)) if eigen else 0.0,
                worst_residual=float(max(r.residual for r in results)))
