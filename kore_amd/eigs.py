"""Shift-and-invert block Krylov-Schur eigensolver on the device workspace.

Replaces what ``E.solve()`` does inside SLEPc for bin/solve.py:91-123
(EPS GNHEP, Krylov-Schur, ST sinvert, TARGET_MAGNITUDE by default): it finds
the eigenpairs of A x = lambda B x nearest the target sigma from the
operator OP = (A - sigma B)^{-1} B, whose eigenvalues are
theta = 1/(lambda - sigma).

Design (MI355X): one solve of the cyclic-reduction factor costs the same HBM
traffic for 1 to 16 right-hand sides, so the Krylov space is grown ``bs``
vectors at a time (block Krylov-Schur, Zhou & Saad 2008; Stewart 2001 for
bs = 1).  Everything of length sizmat stays on the device: the OP
applications, the block classical Gram-Schmidt with one reorthogonalisation
(two passes of V^H W then W -= V h), the Cholesky-QR of each new block and
the restart combination V <- V Q.  A whole expansion (OP, Gram-Schmidt and
Cholesky-QR of every new block) is one library call (kore_ws_expand) whose
coefficients stay on the device until the block is done.  The host holds
only the small projected matrix and computes its ordered Schur form in FP64
(scipy/LAPACK).

Krylov decomposition kept throughout:  OP V[:, :j] = V[:, :j+bs] H[:j+bs, :j].
Convergence: a Ritz pair (theta, y) is converged when ||Bres y|| <= tol |theta|
(relative criterion, SLEPc's default EPS_CONV_REL); pairs are counted in
wanted order, consecutively, as SLEPc does.
"""
from __future__ import annotations

import concurrent.futures
import dataclasses
import threading
import time

import numpy as np
import scipy.linalg as sla

from .backend import KORE_WARN_BREAKDOWN, KORE_WARN_IR_NOT_CONVERGED, PREC_FP64, Workspace


def sort_key(which: str, lam, tau):
    """Ascending key, wanted eigenvalues first (bin/solve.py:99-117)."""
    which = which.upper()
    lam = np.asarray(lam)
    table = {
        "TM": lambda: np.abs(lam - tau),
        "TR": lambda: np.abs(lam.real - tau.real),
        "TI": lambda: np.abs(lam.imag - tau.imag),
        "LM": lambda: -np.abs(lam),
        "SM": lambda: np.abs(lam),
        "LR": lambda: -lam.real,
        "SR": lambda: lam.real,
        "LI": lambda: -lam.imag,
        "SI": lambda: lam.imag,
    }
    if which not in table:
        raise ValueError("unknown which_eigenpairs %r" % which)
    return table[which]()


@dataclasses.dataclass
class EigResult:
    lam: np.ndarray            # converged eigenvalues, wanted order
    X: np.ndarray              # eigenvectors (Kore order), unit 2-norm columns
    errest: np.ndarray         # ||OP x - theta x|| / |theta| estimates
    nconv: int
    restarts: int
    op_columns: int            # OP applications (vectors)
    op_calls: int              # block applications
    seconds: float
    ir_max_iterations: int = 0
    host_seconds: dict = dataclasses.field(default_factory=dict)   # wall time per workspace primitive
    precision: int = 0         # factor precision the converged pairs were computed with
    fallback: bool = False     # True if a complex64 factor was replaced by an FP64 one
    breakdowns: int = 0        # blocks that hit an invariant subspace (replaced by random blocks)
    c64_op_columns: int = 0    # OP applications made with a complex64 factor (before a fallback, if any)
    c64_ir_max: int = 0        # most refinement sweeps one complex64 OP application needed


@dataclasses.dataclass
class _Fallback:
    """An eigensolve abandoned on its complex64 factor: the work done so far."""
    op_columns: int
    ir_max: int


class _Timed:
    """Wraps a workspace so that the wall time of each primitive is summed."""

    def __init__(self, ws, acc):
        self._ws = ws
        self._acc = acc

    def __getattr__(self, name):
        f = getattr(self._ws, name)
        if not callable(f):
            return f
        acc = self._acc

        def g(*a, **k):
            t = time.perf_counter()
            try:
                return f(*a, **k)
            finally:
                acc[name] = acc.get(name, 0.0) + time.perf_counter() - t
        return g


def _cholqr(ws: Workspace, src: int, dst: int, nc: int):
    """Orthonormalise V[:, src:src+nc] into V[:, dst:dst+nc]; returns R with
    V_src = V_dst R.  Two Cholesky-QR passes (the second in place via src);
    a single column is normalised once, as kore_ws_expand does."""
    G = ws.dot(src, nc, src, nc)
    G = 0.5 * (G + G.conj().T)
    R1 = sla.cholesky(G, lower=False)
    ws.gemm(dst, nc, src, nc, sla.solve_triangular(R1, np.eye(nc), lower=False), alpha=1.0, beta=0.0)
    if nc == 1:
        return R1
    G2 = ws.dot(dst, nc, dst, nc)
    G2 = 0.5 * (G2 + G2.conj().T)
    R2 = sla.cholesky(G2, lower=False)
    ws.gemm(src, nc, dst, nc, sla.solve_triangular(R2, np.eye(nc), lower=False), alpha=1.0, beta=0.0)
    ws.copy(src, dst, nc)
    return R2 @ R1


def expand_host(ws, j0, j1, bs, cw, H, ir_tol=1e-14, ir_maxit=0, ir_stop=0):
    """Block Arnoldi expansion driven from the host through the workspace
    primitives; the same contract and arithmetic as kore_ws_expand (OP, two
    classical Gram-Schmidt passes, two Cholesky-QR passes per block, one for
    a single column).
    Returns (status, j_done, info)."""
    iters, conv = 0, True
    for j in range(j0, j1, bs):
        inf = ws.apply_op(j, cw, bs, ir_tol=ir_tol, ir_maxit=ir_maxit)
        iters = max(iters, inf.iterations)
        ok = bool(getattr(inf, "converged", 1))
        conv = conv and ok
        if ir_stop > 0 and (inf.iterations > ir_stop or not ok):
            return KORE_WARN_IR_NOT_CONVERGED, j, _Info(iters, conv)
        nb = j + bs
        h = ws.dot(0, nb, cw, bs)
        ws.gemm(cw, bs, 0, nb, h, alpha=-1.0, beta=1.0)
        h2 = ws.dot(0, nb, cw, bs)
        ws.gemm(cw, bs, 0, nb, h2, alpha=-1.0, beta=1.0)
        H[:nb, j:j + bs] = h + h2
        try:
            H[nb:nb + bs, j:j + bs] = _cholqr(ws, cw, nb, bs)
        except np.linalg.LinAlgError:
            return KORE_WARN_BREAKDOWN, j, _Info(iters, conv)
    return 0, j1, _Info(iters, conv)


@dataclasses.dataclass
class _Info:
    iterations: int
    converged: bool


def _triu_eigvecs(S):
    """Unit-norm eigenvectors of an upper-triangular matrix (columns)."""
    # Z is upper triangular with a unit diagonal: row k of (S - d_j I) z_j = 0
    # gives Z[k, j] = -S[k, k+1:] Z[k+1:, j] / (d_k - d_j) for j > k, so all
    # eigenvectors advance together, one row at a time from the bottom.
    # Near-equal diagonal entries are separated by at least eps |d_j| (as
    # LAPACK's ztrevc perturbs them), and a column that grows past 1e100 is
    # rescaled on the spot so that no entry overflows.
    n = S.shape[0]
    d = np.diag(S)
    Z = np.eye(n, dtype=complex)
    smin = np.maximum(np.finfo(float).eps * np.abs(d), np.finfo(float).tiny / np.finfo(float).eps)
    for k in range(n - 2, -1, -1):
        num = S[k, k + 1:] @ Z[k + 1:, k + 1:]
        den = d[k] - d[k + 1:]
        small = np.abs(den) < smin[k + 1:]
        den[small] = smin[k + 1:][small]
        Z[k, k + 1:] = -num / den
        big = np.abs(Z[k, k + 1:]) > 1e100
        if big.any():
            cols = np.nonzero(big)[0] + k + 1
            Z[:, cols] /= np.abs(Z[k, cols])
    return Z / np.linalg.norm(Z, axis=0)


@dataclasses.dataclass
class _Ritz:
    """Rayleigh-Ritz data of the decomposition of size j (wanted order)."""
    j: int
    S: np.ndarray
    Q: np.ndarray
    Bres: np.ndarray
    order: np.ndarray
    theta: np.ndarray
    Y: np.ndarray
    res: np.ndarray
    nconv: int


def _rayleigh_ritz(H, j, bs, sigma, which, tol, host):
    """Ordered Schur form of T = H[:j, :j] (host, FP64, LAPACK), Ritz vectors
    and residual estimates ||Bres y|| with Bres = H[j:j+bs, :j]."""
    T = H[:j, :j]
    Bres = H[j:j + bs, :j]
    tq = time.perf_counter()
    S, Q = sla.schur(T, output="complex")
    host["schur"] = host.get("schur", 0.0) + time.perf_counter() - tq
    theta_s = np.diag(S).copy()
    Ys = Q @ _triu_eigvecs(S)
    res_s = np.linalg.norm(Bres @ Ys, axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):     # theta = 0 (null space of B) sorts last
        order = np.argsort(sort_key(which, sigma + 1.0 / theta_s, sigma), kind="stable")
    theta = theta_s[order]
    res = res_s[order]
    conv = res <= tol * np.abs(theta)
    nconv = 0
    while nconv < j and conv[nconv]:
        nconv += 1
    return _Ritz(j, S, Q, Bres, order, theta, Ys[:, order], res, nconv)


_POOLS = []                        # threadpoolctl controller, found once per process


def eigs(*args, driver="auto", **kwargs) -> EigResult:
    """Eigenpairs nearest sigma of the pencil held by a factored System.

    driver "native": one kore_eigs call (the library's own Krylov-Schur: the
    projected problem on the host in C++, beside the device expansion);
    "python": the same algorithm driven from here through the workspace
    primitives (``_eigs``, LAPACK for the projected problem); "auto" takes
    the native driver unless the call needs what only the Python one offers
    (a stand-in workspace, a start block of several columns, a per-restart
    log).  Both give the same eigenvalues; the native one measured faster on
    MI355X (26 against 30 ms per eigensolve of the 60 000-unknown inertial
    mode problem, 6.9 against 10 ms at 960 unknowns).

    With the Python driver the projected problem is tiny (at most a few
    hundred columns): multithreaded BLAS only adds start-up cost to each
    call, so the host side runs with one BLAS thread.  The BLAS libraries are
    located once (about 10 ms) and only their thread count is switched per
    call."""
    if driver == "auto":
        v0 = kwargs.get("v0")
        one_col = v0 is None or np.asarray(v0).ndim == 1 or np.asarray(v0).shape[1] == 1
        native_ok = (kwargs.get("workspace") is None and kwargs.get("log") is None and one_col
                     and hasattr(args[0] if args else kwargs.get("sysm"), "eigs_native"))
        driver = "native" if native_ok else "python"
    if driver == "native":
        return _eigs_native(*args, **kwargs)
    if driver != "python":
        raise ValueError("unknown eigensolver driver %r" % driver)
    if not _POOLS:
        try:
            from threadpoolctl import ThreadpoolController
            _POOLS.append(ThreadpoolController())
        except ImportError:        # pragma: no cover
            _POOLS.append(None)
    if _POOLS[0] is None:          # pragma: no cover
        return _eigs(*args, **kwargs)
    with _POOLS[0].limit(limits=1):
        return _eigs(*args, **kwargs)


def _eigs_native(sysm, sigma, nev, ncv=None, tol=1e-12, maxit=100, which="TM", block=1, seed=0,
                 ir_tol=1e-13, ir_maxit=12, v0=None, log=None, workspace=None, fallback_sweeps=4,
                 check_every=None) -> EigResult:
    """kore_eigs with the arguments of ``_eigs`` (the system's current
    factor is used when it is at sigma)."""
    if workspace is not None or log is not None:
        raise ValueError("the native driver takes no stand-in workspace or log")
    n = sysm.n
    bs = max(1, min(int(block), 16))
    # every converged pair is returned, as the Python driver does: room for
    # the whole basis (untouched columns cost no memory)
    m = ncv if ncv else max(2 * nev, nev + 15, 96)
    m = max(m, nev + 2 * bs)
    kmax = int(np.ceil(m / bs) * bs)
    prec = sysm.precision if sysm.precision is not None else PREC_FP64
    v = None
    if v0 is not None:
        v = np.asarray(v0, dtype=complex).reshape(n, -1)[:, 0]
    lam, X, err, info, _ = sysm.eigs_native(sigma, nev=nev, ncv=ncv or 0, tol=tol, maxit=maxit, which=which,
                                            block=bs, precision=prec, seed=seed, v0=v, max_pairs=kmax,
                                            check_every=-1 if check_every is None else check_every,
                                            fallback_sweeps=fallback_sweeps if fallback_sweeps is not None else 0,
                                            ir_tol=ir_tol, ir_maxit=ir_maxit)
    return EigResult(lam=lam, X=X, errest=err, nconv=int(info.nconv), restarts=int(info.restarts),
                     op_columns=int(info.op_columns), op_calls=int(info.op_calls), seconds=float(info.seconds),
                     ir_max_iterations=int(info.ir_max_iterations),
                     host_seconds={"projected": float(info.host_seconds)}, precision=int(info.precision),
                     fallback=bool(info.fallback), breakdowns=int(info.breakdowns),
                     c64_op_columns=int(info.c64_op_columns), c64_ir_max=int(info.c64_ir_max))


def _eigs(sysm, sigma, nev, ncv=None, tol=1e-12, maxit=100, which="TM", block=1, seed=0,
          ir_tol=1e-13, ir_maxit=12, v0=None, log=None, workspace=None, fallback_sweeps=4,
          check_every=None) -> EigResult:
    """fallback_sweeps: with a complex64 factor, if one OP application needs
    more than this many refinement sweeps (or does not converge), the system
    is refactored in FP64 at the same shift and the iteration restarts: every
    OP application costs ~one FP64 sweep instead of many complex64 ones.

    check_every: columns between the early convergence checks made while the
    basis grows (0 disables them).  Each check is a Rayleigh-Ritz step on the
    decomposition reached so far, computed on the host while the device
    expands the next columns; when the wanted pairs have converged the
    iteration ends there, without expanding to ncv.  By default the base
    spacing is max(4, 2 bs) columns; a chunk grows (up to 4x) so that its
    device time covers the predicted Schur time, or so that it ends a base
    step short of where the trend of the worst wanted residual reaches tol.
    An explicit value is used as a fixed spacing."""
    c64 = sysm.precision != PREC_FP64
    res = _eigs_once(sysm, sigma, nev, ncv, tol, maxit, which, block, seed, ir_tol, ir_maxit, v0, log, workspace,
                     fallback_sweeps if c64 else None, check_every)
    if isinstance(res, _Fallback):
        if log:
            log("complex64 refinement too slow after %d OP columns: refactoring in FP64" % res.op_columns)
        fb = res
        sysm.factor(sigma, PREC_FP64)
        res = _eigs_once(sysm, sigma, nev, ncv, tol, maxit, which, block, seed, ir_tol, ir_maxit, v0, log,
                         workspace, None, check_every)
        res.fallback = True
        res.c64_op_columns = fb.op_columns
        res.c64_ir_max = fb.ir_max
    elif c64:
        res.c64_op_columns = res.op_columns
        res.c64_ir_max = res.ir_max_iterations
    res.precision = sysm.precision
    return res


def _eigs_once(sysm, sigma, nev, ncv, tol, maxit, which, block, seed, ir_tol, ir_maxit, v0, log, workspace,
               fallback_sweeps, check_every):
    """Eigenpairs nearest ``sigma`` of the pencil held by ``sysm``.

    ``sysm`` must have been factored at ``sigma`` (System.factor).
    ``workspace`` constructs the vector block; by default the system's cached
    device Workspace is used (tests of the host logic pass a stand-in class
    with the same methods).
    """
    t0 = time.perf_counter()
    sigma = complex(sigma)
    n = sysm.n
    bs = max(1, min(int(block), 16))
    if ncv is None:
        # with the early convergence checks a large space costs nothing until
        # it is needed and saves restarts (each restart is a serial host Schur
        # step).  Measured on the 60 000-unknown m=1, E=1e-6 inertial-mode
        # problem (nev=10): bs=1 converges after ~80 OP columns, bs=2 after
        # ~120 in ~60 sweeps, bs=4 after ~230 in ~58 sweeps plus 3 restarts;
        # with the sweep cost nearly flat in bs, bs=1 with 96 columns was the
        # fastest (18 solves/s against 16 for bs=2 and 11 for bs=4)
        ncv = max(2 * nev, nev + 15, 96)
    ncv = max(ncv, nev + 2 * bs)
    ncv = int(np.ceil(ncv / bs) * bs)
    ncv = min(ncv, max(bs, (n // bs) * bs - bs))
    m = ncv
    # columns: [0, m+bs) basis + residual block ; [m+bs, m+2bs) new block ; [m+2bs, 2m+2bs) restart temp
    cW = m + bs
    cT = m + 2 * bs
    if workspace is None:
        ws_raw = sysm.get_workspace(2 * m + 2 * bs)
        own_ws = False
    else:
        ws_raw = workspace(sysm, 2 * m + 2 * bs)
        own_ws = True
    host = {}
    ws = _Timed(ws_raw, host)
    rng = np.random.default_rng(seed)
    if v0 is None and hasattr(ws_raw, "random"):
        ws.random(cW, bs, seed)                  # generated on the device: no host pass, no upload
    else:
        if v0 is None:
            X0 = rng.standard_normal((n, bs)) + 1j * rng.standard_normal((n, bs))
        else:
            X0 = np.asarray(v0, dtype=complex).reshape(n, -1)
            if X0.shape[1] < bs:
                X0 = np.hstack([X0, rng.standard_normal((n, bs - X0.shape[1])) + 0j])
            X0 = X0[:, :bs]
        ws.set(cW, X0)
    _cholqr(ws, cW, 0, bs)
    H = np.zeros((m + bs, m), dtype=complex, order="F")
    j = 0
    restarts = 0
    op_cols = 0
    op_calls = 0
    ir_iters = 0
    nconv = 0
    breakdowns = 0
    mixed = sysm.precision != PREC_FP64
    expand = getattr(ws, "expand", None) if hasattr(ws_raw, "expand") else None
    stop_expand = getattr(ws_raw, "stop_expand", None) if expand is not None else None
    if expand is None:
        def expand(j0, j1, bs_, cw, Hm, ir_tol=1e-14, ir_maxit=0, ir_stop=0):
            return expand_host(ws, j0, j1, bs_, cw, Hm, ir_tol, ir_maxit, ir_stop)
    adaptive = check_every is None                 # explicit spacing: fixed chunks (deterministic)
    if check_every is None:
        check_every = max(4, 2 * bs)
    check_every = int(np.ceil(check_every / bs) * bs) if check_every > 0 else 0
    j_first = int(np.ceil((nev + bs) / bs) * bs)    # fewest columns worth a check
    j_base = j                                      # columns of the last restart (no check there)
    rr = None
    t_col = None                                    # device wall time per OP column (last chunk)
    t_rr, j_rr = None, None                         # last check's host time and size
    hidden = {}                                     # host time spent beside the device
    worker = None
    trend = []                                      # (columns, worst wanted relative residual) per check
    try:
        while True:
            # ---- expand the decomposition to m columns; from j_first on, each
            # chunk of check_every columns is expanded on the device (a persistent
            # worker thread: the library call releases the GIL) while the host runs
            # the Rayleigh-Ritz step of the decomposition reached before the chunk
            rr = None
            while j < m:
                j_stop, early = m, None
                can_check = bool(check_every) and j >= j_first and j > j_base
                threaded = can_check and j + check_every < m
                if threaded:
                    chunk = check_every
                    if adaptive:
                        # a chunk long enough to cover the check (Schur cost ~ j^3) ...
                        if t_col and t_rr:
                            chunk = max(chunk, int(np.ceil(t_rr * (j / j_rr) ** 3 / t_col / bs)) * bs)
                        # ... unless the worst wanted residual, seen falling at
                        # two checks, reaches tol on its log-linear trend before
                        # that: then the chunk ends at the trend's column, where
                        # the next check is the one expected to succeed (the
                        # expansion beside a successful check is stopped).  The
                        # trend is pessimistic as convergence speeds up.  A
                        # trend column beyond the cover: a base step short of it.
                        if len(trend) >= 2 and trend[-1][1] < trend[-2][1]:
                            (ja, ra), (jb, rb) = trend[-2], trend[-1]
                            slope = (np.log(rb) - np.log(ra)) / (jb - ja)
                            j_conv = jb + (np.log(tol) - np.log(rb)) / slope
                            to_conv = int(np.ceil((j_conv - j) / bs)) * bs
                            chunk = max(bs, to_conv) if to_conv <= chunk else max(chunk, to_conv - check_every)
                        chunk = min(chunk, 4 * check_every)
                    j_stop = min(m, j + chunk)
                elif check_every and not can_check:
                    j_stop = min(m, max(j_first, j + check_every))   # grow to where a check is worth it
                if stop_expand is not None:
                    stop_expand(None)
                if threaded:
                    Hs = np.asfortranarray(H[:j + bs, :j])
                    started = threading.Event()
                    def work(j0=j, j1=j_stop):
                        started.set()
                        t = time.perf_counter()
                        r = expand(j0, j1, bs, cW, H, ir_tol=ir_tol, ir_maxit=ir_maxit if mixed else 0,
                                   ir_stop=fallback_sweeps or 0)
                        return r, time.perf_counter() - t
                    if worker is None:                      # one thread per eigensolve, reused by every chunk
                        worker = concurrent.futures.ThreadPoolExecutor(max_workers=1)
                    fut = worker.submit(work)
                    # the worker enters the library call (which releases the GIL)
                    # before the check starts: LAPACK holds the GIL, and a worker
                    # still waiting for it would leave the device idle
                    started.wait()
                    try:
                        t = time.perf_counter()
                        early = _rayleigh_ritz(Hs, j, bs, sigma, which, tol, hidden)
                        t_rr, j_rr = time.perf_counter() - t, j
                        if stop_expand is not None:
                            if early.nconv >= nev:
                                stop_expand(0)              # the rest of the chunk is not needed
                            elif adaptive and trend:
                                # this check's residual extends the trend: when it
                                # now reaches tol before the chunk's end, the chunk
                                # ends there instead (the next check is made there)
                                wr = float(np.max(early.res[:nev] / np.abs(early.theta[:nev])))
                                (ja, ra) = trend[-1]
                                if np.isfinite(wr) and 0 < wr < ra:
                                    slope = (np.log(wr) - np.log(ra)) / (j - ja)
                                    j_end = j + max(1, int(np.ceil((np.log(tol) - np.log(wr)) / slope / bs))) * bs
                                    if j_end < j_stop:
                                        stop_expand(j_end)
                    finally:
                        (st, jd, ws_info), t_exp = fut.result()    # re-raises the worker's exception
                    if jd > j:
                        t_col = t_exp / (jd - j)
                else:
                    t = time.perf_counter()
                    st, jd, ws_info = expand(j, j_stop, bs, cW, H, ir_tol=ir_tol, ir_maxit=ir_maxit if mixed else 0,
                                             ir_stop=fallback_sweeps or 0)
                    if jd > j:
                        t_col = (time.perf_counter() - t) / (jd - j)
                ir_iters = max(ir_iters, ws_info.iterations)
                done_blocks = (jd - j) // bs
                if st == KORE_WARN_IR_NOT_CONVERGED and fallback_sweeps is not None:
                    if own_ws:
                        ws.close()
                    return _Fallback(op_columns=op_cols + bs * done_blocks, ir_max=ir_iters)
                if st == KORE_WARN_BREAKDOWN:
                    done_blocks += 1
                op_calls += done_blocks
                op_cols += bs * done_blocks
                if early is not None:
                    if early.nconv >= nev:
                        rr = early                          # converged before the chunk
                        break
                    worst = float(np.max(early.res[:nev] / np.abs(early.theta[:nev])))
                    if np.isfinite(worst) and worst > 0:
                        trend.append((early.j, worst))
                if st == KORE_WARN_BREAKDOWN:
                    # invariant subspace: continue with a fresh random block
                    breakdowns += 1
                    nb_all = jd + bs
                    Xr = rng.standard_normal((n, bs)) + 1j * rng.standard_normal((n, bs))
                    ws.set(cW, Xr)
                    for _ in range(2):
                        hh = ws.dot(0, nb_all, cW, bs)
                        ws.gemm(cW, bs, 0, nb_all, hh, alpha=-1.0, beta=1.0)
                    _cholqr(ws, cW, nb_all, bs)
                    H[nb_all:nb_all + bs, jd:jd + bs] = 0.0
                    jd = nb_all
                j = jd
            # ---- Rayleigh-Ritz on the projected matrix (host, FP64, LAPACK)
            if rr is None:
                rr = _rayleigh_ritz(H, m, bs, sigma, which, tol, host)
            nconv = rr.nconv
            if log:
                log("restart %d: size %d, nconv %d, worst wanted residual %.3e" % (
                    restarts, rr.j, nconv, float(np.max(rr.res[:nev] / np.abs(rr.theta[:nev])))))
            if nconv >= nev or restarts >= maxit:
                break
            # ---- thick restart: keep the 'keep' best Ritz values' invariant subspace
            keep = nconv + max(1, (m - nconv) // 2)
            keep = max(keep, nev)
            keep = int(np.ceil(keep / bs) * bs)       # the expansion then ends exactly at m
            keep = max(bs, min(keep, m - bs))
            select = np.zeros(m, dtype=np.int32)
            select[rr.order[:keep]] = 1
            tq = time.perf_counter()
            St, Qt, _, msel, _, _, info = sla.lapack.ztrsen(select, rr.S, rr.Q, job="N")
            host["reorder"] = host.get("reorder", 0.0) + time.perf_counter() - tq
            if info != 0 or msel != keep:
                raise RuntimeError("Schur reordering failed (info %d, %d of %d)" % (info, msel, keep))
            ws.gemm(cT, keep, 0, m, Qt[:, :keep], alpha=1.0, beta=0.0)
            ws.copy(cT, 0, keep)
            ws.copy(m, keep, bs)
            Hn = np.zeros_like(H)
            Hn[:keep, :keep] = np.triu(St[:keep, :keep])
            Hn[keep:keep + bs, :keep] = rr.Bres @ Qt[:, :keep]
            H = Hn
            j = keep
            j_base = keep
            trend = []
            restarts += 1
    finally:
        if worker is not None:
            worker.shutdown(wait=True)
    k = nconv
    X = np.zeros((n, 0), dtype=complex)
    if k > 0:
        # the basis is orthonormal, so ||V y|| = ||y||: normalising y gives
        # unit eigenvectors without a host pass over the downloaded vectors
        Yk = rr.Y[:, :k] / np.linalg.norm(rr.Y[:, :k], axis=0)
        ws.gemm(cT, k, 0, rr.j, Yk, alpha=1.0, beta=0.0)
        X = ws.get(cT, k)
    lam = sigma + 1.0 / rr.theta[:k]
    if own_ws:
        ws.close()
    total = time.perf_counter() - t0
    host["projected"] = total - sum(host.values())
    if hidden:
        host["checks_beside_device"] = sum(hidden.values())
    return EigResult(lam=lam, X=X, errest=rr.res[:k] / np.abs(rr.theta[:k]), nconv=k, restarts=restarts,
                     op_columns=op_cols, op_calls=op_calls, seconds=total,
                     ir_max_iterations=ir_iters, host_seconds=host, breakdowns=breakdowns)


def residuals(A, B, lam, X):
    """||(A - lam B) x|| / (||A||_F ||x||) per pair (host FP64 check)."""
    import scipy.sparse as ss
    A = ss.csr_matrix(A)
    B = ss.csr_matrix(B)
    nA = np.sqrt(np.sum(np.abs(A.data) ** 2))
    out = []
    for i in range(len(lam)):
        x = X[:, i]
        r = A @ x - lam[i] * (B @ x)
        out.append(np.linalg.norm(r) / (nA * np.linalg.norm(x)))
    return np.array(out)
