"""ctypes binding of the HIP library (include/kore_amd.h).

There is no fallback: if ``libkore_amd.so`` is missing or cannot be loaded,
``lib()`` raises ``KoreLibraryError``.  Build it with
``python -m kore_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import threading
import weakref

import numpy as np
import scipy.sparse as ss

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("KORE_AMD_LIB", os.path.join(HERE, "libkore_amd.so"))

KORE_OK = 0
KORE_WARN_IR_NOT_CONVERGED = 1
KORE_WARN_BREAKDOWN = 2
KORE_WARN_NOT_CONVERGED = 3
WHICH = {w: i for i, w in enumerate(["TM", "TR", "TI", "LM", "SM", "LR", "SR", "LI", "SI"])}
PREC_FP64 = 0
PREC_C64 = 1
FORM_AUTO, FORM_SWEEP, FORM_LEAN = range(3)
TIMER_INV, TIMER_GEMM, TIMER_SKINNY, TIMER_SPMV, TIMER_ORTHO, TIMER_OTHER, TIMER_SPARSE = range(7)
TIMER_NAMES = ["inv", "gemm", "skinny", "spmv", "ortho", "other", "sparse"]


class KoreLibraryError(RuntimeError):
    pass


class KoreError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("kore_amd error %d: %s" % (code, msg))
        self.code = code


class Layout(C.Structure):
    _fields_ = [(n, C.c_int32) for n in
                ("m", "lmax", "symm", "N1", "hydro", "magnetic", "thermal", "compositional", "symmB0")]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("rel_correction", C.c_double)]


class FactorInfo(C.Structure):
    _fields_ = [("levels", C.c_int32), ("nblk", C.c_int32), ("flops", C.c_double),
                ("flops_inv", C.c_double), ("flops_gemm", C.c_double), ("flops_sparse", C.c_double),
                ("solve_bytes", C.c_double),
                ("solve_launches", C.c_int32), ("bytes", C.c_int64), ("ms", C.c_double), ("form", C.c_int32)]


class SysState(C.Structure):
    _fields_ = [("sizmat", C.c_int64), ("nblk", C.c_int32), ("has_b", C.c_int32), ("factored", C.c_int32),
                ("precision", C.c_int32), ("sre", C.c_double), ("sim", C.c_double)]


class EigsOpts(C.Structure):
    _fields_ = [("nev", C.c_int32), ("ncv", C.c_int32), ("tol", C.c_double), ("maxit", C.c_int32),
                ("which", C.c_int32), ("block", C.c_int32), ("precision", C.c_int32),
                ("fallback_sweeps", C.c_int32), ("ir_tol", C.c_double), ("ir_maxit", C.c_int32),
                ("check_every", C.c_int32), ("seed", C.c_uint64)]


class EigsInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nconv", "returned", "restarts", "op_columns", "op_calls",
                                         "ir_max_iterations", "precision", "fallback", "breakdowns",
                                         "c64_op_columns", "c64_ir_max", "basis_columns")] + \
               [("seconds", C.c_double), ("host_seconds", C.c_double)]


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_D = C.c_double
_PI32 = C.POINTER(C.c_int32)
_PI64 = C.POINTER(C.c_int64)
_PD = C.POINTER(C.c_double)

# (name, restype, argtypes) for every symbol of include/kore_amd.h
SIGNATURES = [
    ("kore_version", _I32, []),
    ("kore_last_error", C.c_char_p, []),
    ("kore_ctx_create", _I32, [_I32, _I64, C.POINTER(_P)]),
    ("kore_ctx_destroy", None, [_P]),
    ("kore_ctx_sync", _I32, [_P]),
    ("kore_ctx_bytes_in_use", _I64, [_P]),
    ("kore_mem_trim", _I32, [_I32]),
    ("kore_mem_stats", _I32, [_I32, _PI64, _PI64]),
    ("kore_ctx_set_graphs", _I32, [_P, _I32]),
    ("kore_ctx_set_timing", _I32, [_P, _I32]),
    ("kore_ctx_timing", _I32, [_P, _I32, _PI64, _PD]),
    ("kore_ctx_timing_reset", _I32, [_P]),
    ("kore_layout_blocks", _I32, [C.POINTER(Layout), _PI64, _PI32, _PI32, _I32, _PI64]),
    ("kore_factor_bytes", _I64, [C.POINTER(Layout), _I32]),
    ("kore_sys_create", _I32, [_P, C.POINTER(Layout), _I64, _PI32, _PI32, _PD, _PI32, _PI32, _PD, _I32,
                               C.POINTER(_P)]),
    ("kore_sys_destroy", None, [_P]),
    ("kore_sys_prepare", _I32, [C.POINTER(Layout), _I64, _PI32, _PI32, _PD, _PI32, _PI32, _PD, _I32,
                                C.POINTER(_P)]),
    ("kore_sys_prepare_affine", _I32, [C.POINTER(Layout), _I64, _PI32, _PI32, _PD, _PD, _D, _PI32, _PI32, _PD,
                                       _I32, C.POINTER(_P)]),
    ("kore_sys_create_prepared", _I32, [_P, _P, C.POINTER(_P)]),
    ("kore_sys_plan_destroy", None, [_P]),
    ("kore_sys_update_a", _I32, [_P, _PD]),
    ("kore_sys_norm", _I32, [_P, _D, _D, _PD]),
    ("kore_sys_set_factor_form", _I32, [_P, _I32]),
    ("kore_set_host_threads", _I32, [_I32]),
    ("kore_host_threads", _I32, []),
    ("kore_cpu_share", _I32, []),
    ("kore_sys_query", _I32, [_P, C.POINTER(SysState)]),
    ("kore_factor_async", _I32, [_P, _D, _D, _I32, _P]),
    ("kore_factor_check", _I32, [_P]),
    ("kore_solve_async", _I32, [_P, _I32, _P, _I64, _P, _I64, _I32, _P]),
    ("kore_eigs_default_opts", None, [C.POINTER(EigsOpts)]),
    ("kore_eigs", _I32, [_P, _D, _D, C.POINTER(EigsOpts), _PD, _I32, _PD, _PD, _I64, _PD, C.POINTER(EigsInfo)]),
    ("kore_schur", _I32, [_I32, _PD, _I64, _PD, _I64]),
    ("kore_schur_reorder", _I32, [_I32, _PD, _I64, _PD, _I64, _PI32, _PI32]),
    ("kore_factor", _I32, [_P, _D, _D, _I32, C.POINTER(FactorInfo)]),
    ("kore_factor_release", _I32, [_P]),
    ("kore_solve", _I32, [_P, _I32, _PD, _I64, _PD, _I64, _D, _I32, C.POINTER(SolveInfo)]),
    ("kore_residual", _I32, [_P, _I32, _PD, _I64, _PD, _PD, _I64, _D, _D, _PD]),
    ("kore_residual_noalloc", _I32, [_P, _I32, _PD, _I64, _PD, _PD, _I64, _D, _D, _PD]),
    ("kore_solve_gmres", _I32, [_P, _PD, _PD, _D, _I32, _I32, C.POINTER(SolveInfo)]),
    ("kore_ws_create", _I32, [_P, _I32, C.POINTER(_P)]),
    ("kore_ws_destroy", None, [_P]),
    ("kore_ws_set", _I32, [_P, _I32, _I32, _PD, _I64]),
    ("kore_ws_get", _I32, [_P, _I32, _I32, _PD, _I64]),
    ("kore_ws_apply_op", _I32, [_P, _I32, _I32, _I32, _D, _I32, C.POINTER(SolveInfo)]),
    ("kore_ws_spmv", _I32, [_P, _I32, _I32, _I32, _D, _D, _D, _D]),
    ("kore_ws_dot", _I32, [_P, _I32, _I32, _I32, _I32, _PD]),
    ("kore_ws_gemm", _I32, [_P, _I32, _I32, _I32, _I32, _PD, _D, _D]),
    ("kore_ws_copy", _I32, [_P, _I32, _I32, _I32]),
    ("kore_ws_random", _I32, [_P, _I32, _I32, C.c_uint64]),
    ("kore_ws_stop_expand", _I32, [_P, _I32]),
    ("kore_ws_expand", _I32, [_P, _I32, _I32, _I32, _I32, _D, _I32, _I32, _PD, _I64, C.POINTER(C.c_int32),
                              C.POINTER(SolveInfo)]),
]

_lib = None
_lock = threading.Lock()


def lib():
    """Load the HIP library (once).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIBPATH):
            raise KoreLibraryError("HIP library %s not found; build it with `python -m kore_amd.build`" % LIBPATH)
        try:
            L = C.CDLL(LIBPATH)
        except OSError as e:
            raise KoreLibraryError("cannot load %s: %s" % (LIBPATH, e))
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        return L


def _check(rc, allow_warn=True):
    if rc < 0 or (rc > 0 and not allow_warn):
        msg = lib().kore_last_error()
        raise KoreError(rc, msg.decode() if msg else "")
    return rc


def _dptr(a):
    return a.ctypes.data_as(_PD)


def make_layout(m, lmax, symm, N1, hydro=1, magnetic=0, thermal=0, compositional=0, symmB0=-1) -> Layout:
    return Layout(int(m), int(lmax), int(symm), int(N1), int(hydro), int(magnetic), int(thermal),
                  int(compositional), int(symmB0))


def layout_from_params(par) -> Layout:
    return make_layout(par.m, par.lmax, par.symm, par.N1, par.hydro, par.magnetic, par.thermal,
                       par.compositional, par.symmB0 if par.magnetic else -1)


def layout_blocks(lay: Layout):
    """(sizmat, block sizes, permutation) computed by the library (host only)."""
    L = lib()
    n = C.c_int64(0)
    nb = C.c_int32(0)
    _check(L.kore_layout_blocks(C.byref(lay), C.byref(n), C.byref(nb), None, 0, None))
    sizes = np.zeros(nb.value, dtype=np.int32)
    perm = np.zeros(n.value, dtype=np.int64)
    _check(L.kore_layout_blocks(C.byref(lay), C.byref(n), C.byref(nb), sizes.ctypes.data_as(_PI32), nb.value,
                                perm.ctypes.data_as(_PI64)))
    return n.value, sizes, perm


def factor_bytes(lay: Layout, precision: int) -> int:
    return int(lib().kore_factor_bytes(C.byref(lay), int(precision)))


def mem_stats(device: int = 0):
    """(cached device bytes, device allocations so far) of the library's pool."""
    cached = C.c_int64(0)
    allocs = C.c_int64(0)
    _check(lib().kore_mem_stats(int(device), C.byref(cached), C.byref(allocs)))
    return cached.value, allocs.value


def set_host_threads(n: int):
    """Threads per call of the library's parallel host loops (system set-up,
    staging copies); n <= 0 restores the default (kore_set_host_threads)."""
    _check(lib().kore_set_host_threads(int(n)))


def host_threads() -> int:
    return int(lib().kore_host_threads())


def cpu_share() -> int:
    """The process's CPU share as the library sees it (cgroup CPU quota, else
    the affinity mask)."""
    return int(lib().kore_cpu_share())


def mem_trim(device: int = 0):
    """Returns the library's cached device and pinned blocks to the runtime."""
    _check(lib().kore_mem_trim(int(device)))


_LIVE_CONTEXTS = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    """Contexts still open at interpreter exit are closed (with their systems
    and workspaces) while the HIP runtime is intact, not left to object
    finalisation at an unspecified point of the shutdown."""
    for ctx in list(_LIVE_CONTEXTS):
        try:
            ctx.close()
        except Exception:
            pass


class Context:
    """A device context: one HIP stream, a memory budget, kernel timers."""

    def __init__(self, device: int = 0, mem_budget: int = 0):
        L = lib()
        p = _P()
        _check(L.kore_ctx_create(int(device), int(mem_budget), C.byref(p)))
        self._p = p
        self._children = weakref.WeakSet()
        _LIVE_CONTEXTS.add(self)

    @property
    def handle(self):
        return self._p

    def sync(self):
        _check(lib().kore_ctx_sync(self._p))

    def bytes_in_use(self) -> int:
        return int(lib().kore_ctx_bytes_in_use(self._p))

    def set_graphs(self, on: bool):
        """Replay solve sweeps from captured HIP graphs, or launch them kernel by kernel (default)."""
        _check(lib().kore_ctx_set_graphs(self._p, 1 if on else 0))

    def set_timing(self, on: bool):
        _check(lib().kore_ctx_set_timing(self._p, 1 if on else 0))

    def timing(self):
        out = {}
        for k, name in enumerate(TIMER_NAMES):
            cnt = C.c_int64(0)
            ms = C.c_double(0)
            _check(lib().kore_ctx_timing(self._p, k, C.byref(cnt), C.byref(ms)))
            out[name] = (cnt.value, ms.value)
        return out

    def timing_reset(self):
        _check(lib().kore_ctx_timing_reset(self._p))

    def close(self):
        if getattr(self, "_p", None):
            for ch in list(self._children):
                ch.close()
            lib().kore_ctx_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def schur(T):
    """(S, Q) with T = Q S Q^H, S upper triangular (kore_schur, host only)."""
    S = np.asfortranarray(np.array(T, dtype=np.complex128))
    n = S.shape[0]
    Q = np.zeros((n, n), dtype=np.complex128, order="F")
    _check(lib().kore_schur(n, _dptr(S), max(n, 1), _dptr(Q), max(n, 1)))
    return S, Q


def schur_reorder(S, Q, select):
    """(S', Q', m): the selected diagonal entries moved to the top (kore_schur_reorder, host only)."""
    S = np.asfortranarray(np.array(S, dtype=np.complex128))
    Q = np.asfortranarray(np.array(Q, dtype=np.complex128))
    n = S.shape[0]
    sel = np.ascontiguousarray(np.asarray(select, dtype=np.int32))
    m = C.c_int32(0)
    _check(lib().kore_schur_reorder(n, _dptr(S), max(n, 1), _dptr(Q), max(n, 1), sel.ctypes.data_as(_PI32),
                                    C.byref(m)))
    return S, Q, m.value


def _csr_args(A, B):
    """(n, nnz_a, keep-alive arrays, argument list) of the CSR pair for kore_sys_create / _prepare."""
    A = ss.csr_matrix(A)
    a = (A.indptr.astype(np.int32, copy=False), A.indices.astype(np.int32, copy=False),
         np.ascontiguousarray(A.data.astype(np.complex128, copy=False)))
    args = [a[0].ctypes.data_as(_PI32), a[1].ctypes.data_as(_PI32), _dptr(a[2])]
    keep = [a]
    if B is not None:
        B = ss.csr_matrix(B)
        bc = np.iscomplexobj(B.data)
        b = (B.indptr.astype(np.int32, copy=False), B.indices.astype(np.int32, copy=False),
             np.ascontiguousarray(B.data.astype(np.complex128 if bc else np.float64, copy=False)))
        keep.append(b)
        args += [b[0].ctypes.data_as(_PI32), b[1].ctypes.data_as(_PI32), b[2].ctypes.data_as(_PD), 1 if bc else 0]
    else:
        args += [None, None, None, 0]
    return A.shape[0], int(A.nnz), keep, args


class SystemPlan:
    """The host half of a system's set-up (kore_sys_prepare): needs no
    device, so it can run ahead of a free slot; System.from_plan uploads it."""

    def __init__(self, lay: Layout, A, B=None):
        n, nnz_a, keep, args = _csr_args(A, B)
        p = _P()
        _check(lib().kore_sys_prepare(C.byref(lay), n, *args, C.byref(p)))
        self._p = p
        self.n = n
        self.nnz_a = nnz_a
        self.lay = lay

    @classmethod
    def affine(cls, lay: Layout, indptr, indices, a0, a1, e, B=None, shape=None):
        """The plan of A = a0 + e a1 on the CSR pattern (indptr, indices),
        without building A (kore_sys_prepare_affine)."""
        n = int(shape[0]) if shape is not None else len(indptr) - 1
        ip = np.ascontiguousarray(indptr, dtype=np.int32)
        ix = np.ascontiguousarray(indices, dtype=np.int32)
        v0 = np.ascontiguousarray(a0, dtype=np.complex128)
        v1 = np.ascontiguousarray(a1, dtype=np.complex128)
        if len(ip) != n + 1 or v0.shape != v1.shape or v0.size != ix.size:
            raise ValueError("A0 and A1 must hold one value per entry of the pattern")
        _, _, keep, args = _csr_args(ss.csr_matrix((n, n), dtype=np.complex128), B)
        p = _P()
        _check(lib().kore_sys_prepare_affine(C.byref(lay), n, ip.ctypes.data_as(_PI32), ix.ctypes.data_as(_PI32),
                                             _dptr(v0), _dptr(v1), float(e), *args[3:], C.byref(p)))
        self = cls.__new__(cls)
        self._p = p
        self.n = n
        self.nnz_a = int(ix.size)
        self.lay = lay
        return self

    def close(self):
        if getattr(self, "_p", None):
            lib().kore_sys_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class System:
    """A + sigma B system resident on the device (Kore ordering on the host)."""

    def __init__(self, ctx: Context, lay: Layout, A, B=None, plan: SystemPlan = None):
        L = lib()
        self.ctx = ctx
        self.lay = lay
        p = _P()
        if plan is not None:
            if plan._p is None:
                raise ValueError("the system plan was already used")
            rc = L.kore_sys_create_prepared(ctx.handle, plan._p, C.byref(p))
            plan.close()
            _check(rc)
            n, nnz_a = plan.n, plan.nnz_a
        else:
            n, nnz_a, keep, args = _csr_args(A, B)
            _check(L.kore_sys_create(ctx.handle, C.byref(lay), n, *args, C.byref(p)))
            del keep                           # the library keeps its own copies
        self.n = n
        self._p = p
        self.nnz_a = nnz_a
        self._children = weakref.WeakSet()
        ctx._children.add(self)
        self.sigma = None
        self.precision = None
        self.last_factor = None

    @property
    def handle(self):
        return self._p

    def update_a(self, data):
        """New values of A on the pattern given at creation (kore_sys_update_a);
        the system must be factored again."""
        data = np.ascontiguousarray(np.asarray(data, dtype=np.complex128))
        if data.shape != (self.nnz_a,):
            raise ValueError("expected %d values of A, got %s" % (self.nnz_a, data.shape))
        _check(lib().kore_sys_update_a(self._p, _dptr(data)))
        self.sigma = None

    def set_factor_form(self, form):
        """Form of the later factorisations (FORM_AUTO, FORM_SWEEP, FORM_LEAN;
        kore_sys_set_factor_form)."""
        _check(lib().kore_sys_set_factor_form(self._p, int(form)))

    def norm(self, sigma=0j) -> float:
        out = C.c_double(0)
        _check(lib().kore_sys_norm(self._p, complex(sigma).real, complex(sigma).imag, C.byref(out)))
        return out.value

    def factor(self, sigma, precision=PREC_FP64) -> FactorInfo:
        info = FactorInfo()
        sigma = complex(sigma)
        _check(lib().kore_factor(self._p, sigma.real, sigma.imag, int(precision), C.byref(info)))
        self.sigma = sigma
        self.precision = precision
        self.last_factor = info
        return info

    def release_factor(self):
        _check(lib().kore_factor_release(self._p))

    def query(self) -> SysState:
        out = SysState()
        _check(lib().kore_sys_query(self._p, C.byref(out)))
        return out

    def factor_async(self, sigma, precision=PREC_FP64, stream=0):
        """Queues the factorisation behind the work on the HIP stream
        ``stream`` (an integer hipStream_t of the HIP runtime the library is
        linked against; 0: the default stream); ``stream`` waits for it
        (kore_factor_async)."""
        sigma = complex(sigma)
        _check(lib().kore_factor_async(self._p, sigma.real, sigma.imag, int(precision), _P(int(stream) or None)))
        self.sigma = sigma
        self.precision = precision

    def factor_check(self):
        """Waits for the context's stream; raises on a zero pivot of a queued factorisation."""
        _check(lib().kore_factor_check(self._p))

    def solve_async(self, d_b: int, d_x: int, nrhs=1, ldb=None, ldx=None, ir_sweeps=0, stream=0):
        """Queues x = (A - sigma B)^{-1} b on device arrays (addresses d_b,
        d_x of complex128 columns in Kore order) behind ``stream`` (kore_solve_async)."""
        _check(lib().kore_solve_async(self._p, int(nrhs), _P(int(d_b)), int(ldb or self.n), _P(int(d_x)),
                                      int(ldx or self.n), int(ir_sweeps), _P(int(stream) or None)))

    def eigs_native(self, tau, nev, ncv=0, tol=1e-12, maxit=100, which="TM", block=1, precision=PREC_FP64,
                    seed=0, v0=None, max_pairs=None, vectors=True, check_every=-1, fallback_sweeps=4,
                    ir_tol=1e-13, ir_maxit=12):
        """Eigenpairs nearest tau through the library's own Krylov-Schur
        (kore_eigs): returns (lam, X, errest, EigsInfo, status)."""
        L = lib()
        o = EigsOpts()
        L.kore_eigs_default_opts(C.byref(o))
        o.nev, o.ncv, o.tol, o.maxit = int(nev), int(ncv or 0), float(tol), int(maxit)
        o.which, o.block, o.precision, o.seed = WHICH[which.upper()], int(block), int(precision), int(seed)
        o.check_every, o.fallback_sweeps, o.ir_tol, o.ir_maxit = (int(check_every), int(fallback_sweeps),
                                                                  float(ir_tol), int(ir_maxit))
        kmax = int(max_pairs if max_pairs is not None else max(nev, 1) + 16)
        lam = np.zeros(kmax, dtype=np.complex128)
        err = np.zeros(kmax)
        X = np.zeros((self.n, kmax), dtype=np.complex128, order="F") if vectors else None
        v = None
        if v0 is not None:
            v = np.ascontiguousarray(np.asarray(v0, dtype=np.complex128).reshape(self.n))
        info = EigsInfo()
        tau = complex(tau)
        rc = _check(L.kore_eigs(self._p, tau.real, tau.imag, C.byref(o), _dptr(v) if v is not None else None,
                                kmax, _dptr(lam), _dptr(X) if X is not None else None, self.n, _dptr(err),
                                C.byref(info)))
        st = self.query()
        self.sigma = complex(st.sre, st.sim)
        self.precision = st.precision
        k = info.returned
        return lam[:k], (X[:, :k] if X is not None else None), err[:k], info, rc

    def get_workspace(self, ncol: int) -> "Workspace":
        """A device workspace of at least ncol columns, kept with the system
        so that repeated eigensolves do not re-allocate it."""
        ws = getattr(self, "_ws_cached", None)
        if ws is None or ws._p is None or ws.ncol < ncol:
            if ws is not None:
                ws.close()
            ws = Workspace(self, int(ncol))
            self._ws_cached = ws
        return ws

    def solve(self, b, ir_tol=1e-14, ir_maxit=0):
        """x = (A - sigma B)^{-1} b for b of shape (n,) or (n, nrhs) (Kore order)."""
        b = np.asarray(b, dtype=np.complex128)
        one = b.ndim == 1
        bb = np.asfortranarray(b.reshape(self.n, -1))
        x = np.zeros_like(bb, order="F")
        info = SolveInfo()
        rc = lib().kore_solve(self._p, bb.shape[1], _dptr(bb), self.n,
                              _dptr(x), self.n, float(ir_tol), int(ir_maxit), C.byref(info))
        _check(rc)      # 1 (refinement not converged) is reported through info.converged
        return (x[:, 0] if one else x), info

    def solve_gmres(self, b, tol=1e-14, maxit=60, restart=20):
        """x = (A - sigma B)^{-1} b (one column, Kore order) by FP64 GMRES
        right-preconditioned with the factor (kore_solve_gmres)."""
        bb = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(self.n))
        x = np.zeros(self.n, dtype=np.complex128)
        info = SolveInfo()
        rc = lib().kore_solve_gmres(self._p, _dptr(bb), _dptr(x), float(tol), int(maxit), int(restart),
                                    C.byref(info))
        _check(rc)      # 1: not converged, reported through info.converged
        return x, info

    def residual_norms(self, X, lam=None, b=None, shift=0j):
        """||(A - (shift + lam_c) B) x_c - b_c||_2 per column, on the device
        from the stored matrices (X, b in Kore order; lam, b optional)."""
        X = np.asfortranarray(np.asarray(X, dtype=np.complex128).reshape(self.n, -1))
        k = X.shape[1]
        lp = None
        if lam is not None:
            lam = np.ascontiguousarray(np.asarray(lam, dtype=np.complex128).reshape(k))
            lp = _dptr(lam)
        bp = None
        if b is not None:
            b = np.asfortranarray(np.asarray(b, dtype=np.complex128).reshape(self.n, k))
            bp = _dptr(b)
        out = np.zeros(k)
        shift = complex(shift)
        _check(lib().kore_residual(self._p, k, _dptr(X), self.n, lp, bp, self.n, shift.real, shift.imag, _dptr(out)))
        return out

    def close(self):
        if getattr(self, "_p", None):
            for ch in list(self._children):
                ch.close()
            lib().kore_sys_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Workspace:
    """ncol device vectors (degree-major order) for the Krylov iteration."""

    def __init__(self, sysm: System, ncol: int):
        self.sys = sysm
        self.n = sysm.n
        self.ncol = ncol
        p = _P()
        _check(lib().kore_ws_create(sysm.handle, int(ncol), C.byref(p)))
        self._p = p
        sysm._children.add(self)

    def set(self, c0, X):
        X = np.asfortranarray(np.asarray(X, dtype=np.complex128).reshape(self.n, -1))
        _check(lib().kore_ws_set(self._p, int(c0), X.shape[1], _dptr(X), self.n))

    def random(self, c0, nc, seed):
        """Pseudo-random columns generated on the device (Krylov start block)."""
        _check(lib().kore_ws_random(self._p, int(c0), int(nc), int(seed) & 0xFFFFFFFFFFFFFFFF))

    def get(self, c0, nc):
        X = np.zeros((self.n, nc), dtype=np.complex128, order="F")
        _check(lib().kore_ws_get(self._p, int(c0), int(nc), _dptr(X), self.n))
        return X

    def apply_op(self, src, dst, nc, ir_tol=1e-14, ir_maxit=0):
        info = SolveInfo()
        rc = lib().kore_ws_apply_op(self._p, int(src), int(dst), int(nc), float(ir_tol), int(ir_maxit),
                                    C.byref(info))
        _check(rc)
        return info

    def spmv(self, src, dst, nc, ca, cb):
        ca, cb = complex(ca), complex(cb)
        _check(lib().kore_ws_spmv(self._p, int(src), int(dst), int(nc), ca.real, ca.imag, cb.real, cb.imag))

    def dot(self, a0, na, b0, nb):
        H = np.zeros((na, nb), dtype=np.complex128, order="F")
        _check(lib().kore_ws_dot(self._p, int(a0), int(na), int(b0), int(nb), _dptr(H)))
        return H

    def gemm(self, d0, nd, s0, ns, Q, alpha=1.0, beta=0.0):
        Q = np.asfortranarray(np.asarray(Q, dtype=np.complex128).reshape(ns, nd))
        _check(lib().kore_ws_gemm(self._p, int(d0), int(nd), int(s0), int(ns), _dptr(Q),
                                  float(alpha), float(beta)))

    def copy(self, src, dst, nc):
        _check(lib().kore_ws_copy(self._p, int(src), int(dst), int(nc)))

    def expand(self, j0, j1, bs, cw, H, ir_tol=1e-14, ir_maxit=0, ir_stop=0):
        """Block Arnoldi expansion on the device (kore_ws_expand); H is a
        Fortran-ordered complex128 array whose columns j0..j1 are filled.
        Returns (status, j_done, SolveInfo)."""
        if not (isinstance(H, np.ndarray) and H.dtype == np.complex128 and H.flags.f_contiguous):
            raise ValueError("H must be a Fortran-ordered complex128 array")
        info = SolveInfo()
        jd = C.c_int32(0)
        rc = lib().kore_ws_expand(self._p, int(j0), int(j1), int(bs), int(cw), float(ir_tol), int(ir_maxit),
                                  int(ir_stop), _dptr(H), H.shape[0], C.byref(jd), C.byref(info))
        _check(rc)
        return rc, int(jd.value), info

    def stop_expand(self, j_end=0):
        """Ask an expand() running on another thread to start no block at a
        column >= j_end (0: end after the block in flight); None clears the
        request (kore_ws_stop_expand)."""
        _check(lib().kore_ws_stop_expand(self._p, -1 if j_end is None else int(j_end)))

    def close(self):
        if getattr(self, "_p", None):
            # a system and its cached workspace form a reference cycle, so the
            # collector may finalise the system first: kore_sys_destroy then
            # destroyed this workspace already
            if getattr(self.sys, "_p", None):
                lib().kore_ws_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
