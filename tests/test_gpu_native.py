"""GPU tests of the one-call eigensolver (kore_eigs) and of the
stream-ordered entry points (kore_factor_async, kore_solve_async), through
the C ABI.  kore_eigs is compared with the oracle (SuperLU + Krylov-Schur on
the host) and with the reference's spin-over golden; the stream-ordered
solves with the blocking ones."""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as sla

from conftest import SPINOVER_LAMBDA

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from kore_amd import backend
    c = backend.Context(0)
    yield c
    c.close()


def _system(ctx, par, A, B):
    from kore_amd import backend
    return backend.System(ctx, backend.layout_from_params(par), A, B)


def _oracle_pairs(A, B, sigma, nev):
    from oracle import krylov_schur
    lu = sla.splu((A - sigma * B).tocsc())
    lam, _, _, _ = krylov_schur.krylov_schur(lambda x: lu.solve(B @ x), A.shape[0], sigma, nev, tol=1e-13,
                                             maxit=200)
    return lam[:nev]


@pytest.mark.parametrize("precision,block", [(0, 1), (0, 3), (1, 1)])
def test_kore_eigs_hydro_small(ctx, hydro_small, precision, block):
    """The eigenpairs from one kore_eigs call (the system unfactored before
    the call: kore_eigs factors it at tau) match the oracle to 1e-10 and have
    true residuals <= 1e-10; with a complex64 factor they come from the
    complex64 path (no fallback)."""
    from kore_amd import eigs
    par, A, B = hydro_small
    say("case: kore_eigs hydro_small precision", precision, "block", block)
    S = _system(ctx, par, A, B)
    lam, X, err, info, rc = S.eigs_native(par.tau, nev=5, tol=1e-12, block=block, precision=precision)
    say("  nconv", info.nconv, "restarts", info.restarts, "ops", info.op_columns, "fallback", info.fallback,
        "host s %.4f of %.4f" % (info.host_seconds, info.seconds))
    assert rc == 0 and info.nconv >= 5
    assert S.precision == precision and S.sigma == par.tau
    if precision == 1:
        assert info.fallback == 0 and info.precision == 1 and info.c64_op_columns == info.op_columns
    for lo in _oracle_pairs(A, B, par.tau, 5):
        d = np.min(np.abs(lam - lo)) / abs(lo)
        say("  lambda", lo, "rel diff", d)
        assert d <= 1e-10
    assert eigs.residuals(A, B, lam[:5], X[:, :5]).max() <= 1e-10
    assert np.all(err[:5] <= 1e-12)
    np.testing.assert_allclose(np.linalg.norm(X, axis=0), 1.0, rtol=1e-12)
    S.close()


def test_kore_eigs_spinover_golden(ctx, spinover):
    """The reference's spin-over test (tests/test_spinover.py:21-29, rtol
    1e-8) through one C call."""
    par, A, B = spinover
    say("case: kore_eigs spin-over")
    S = _system(ctx, par, A, B)
    lam, X, err, info, rc = S.eigs_native(par.tau, nev=par.nev, tol=1e-13, block=4)
    assert rc == 0
    got = lam[np.argmax(lam.real)]
    say("  spin-over lambda", got)
    np.testing.assert_allclose([got.real, got.imag], [SPINOVER_LAMBDA.real, SPINOVER_LAMBDA.imag], rtol=1e-8,
                               atol=1e-20)
    S.close()


def test_kore_eigs_agrees_with_python_driver(ctx, hydro_small):
    """kore_eigs and the Python driver (eigs.eigs, LAPACK for the projected
    problem) find the same eigenvalues with fixed check spacing; a warm start
    from a converged eigenvector needs fewer operator applications."""
    from kore_amd import eigs
    par, A, B = hydro_small
    say("case: kore_eigs vs eigs.eigs")
    S = _system(ctx, par, A, B)
    S.factor(par.tau, 0)
    r = eigs.eigs(S, par.tau, nev=5, tol=1e-12, check_every=8, driver="python")
    lam, X, _, info, _ = S.eigs_native(par.tau, nev=5, tol=1e-12, check_every=8)
    for l in r.lam[:5]:
        assert np.min(np.abs(lam - l)) / abs(l) <= 1e-10
    lam2, _, _, info2, _ = S.eigs_native(par.tau, nev=1, tol=1e-12, v0=X[:, 0], check_every=4)
    say("  cold ops", info.op_columns, "warm ops (nev 1)", info2.op_columns)
    assert abs(lam2[0] - lam[0]) / abs(lam[0]) <= 1e-10
    assert info2.op_columns <= 8
    S.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_kore_eigs_mac(ctx, precision):
    """Magneto-thermal pencil (blocks alternating 72 / 48) through kore_eigs,
    against the oracle to 1e-10 (complex64: fallback after 8 sweeps)."""
    from kore_amd import assembly, eigs
    from kore_amd.params import Params
    par = Params(m=2, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-4, N=24, lmax=29, magnetic=1, thermal=1,
                 Lambda=0.1, Pm=1e-2, Prandtl=1.0, Ra_gap=6e5)
    A, B, _ = assembly.assemble(par)
    sigma = 0.1j
    say("case: kore_eigs MAC precision", precision)
    S = _system(ctx, par, A, B)
    lam, X, _, info, rc = S.eigs_native(sigma, nev=6, tol=1e-12, maxit=200, precision=precision, fallback_sweeps=8)
    say("  fallback", info.fallback, "c64 ops", info.c64_op_columns, "ir max", info.c64_ir_max)
    assert rc == 0 and info.nconv >= 6
    for l in _oracle_pairs(A, B, sigma, 6):
        assert np.min(np.abs(lam - l)) / abs(l) <= 1e-10
    assert eigs.residuals(A, B, lam[:6], X[:, :6]).max() <= 1e-10
    S.close()


def test_kore_eigs_breakdown_recovers(ctx, hydro_small):
    """Start vector in the null space of B': the first operator application
    gives zero, the expansion reports the breakdown, kore_eigs continues
    with a fresh random direction, and the eigenvalues match the oracle on
    the same pencil."""
    from conftest import pencil_with_null_block
    par, A, B = hydro_small
    A2, Bn, E = pencil_with_null_block(A, B)
    say("case: kore_eigs breakdown")
    S = _system(ctx, par, A2, Bn)
    lam, X, _, info, rc = S.eigs_native(par.tau, nev=4, tol=1e-12, v0=E[:, 0])
    say("  breakdowns", info.breakdowns, "nconv", info.nconv)
    assert rc == 0 and info.breakdowns >= 1 and info.nconv >= 4
    for l in _oracle_pairs(A2, Bn, par.tau, 4):
        assert np.min(np.abs(lam - l)) / abs(l) <= 1e-10
    S.close()


class _Hip:
    """The HIP runtime the library is linked against (the system ROCm
    libamdhip64, already loaded with the library): its streams and device
    memory are the ones the stream-ordered entry points take.  (torch wheels
    bundle a HIP runtime of their own, whose handles are not valid here.)"""

    def __init__(self):
        import ctypes as C
        from kore_amd import backend
        backend.lib()
        self.C = C
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]

    def ok(self, rc):
        assert rc == 0, "HIP error %d" % rc

    def malloc(self, nbytes):
        p = self.C.c_void_p()
        self.ok(self.h.hipMalloc(self.C.byref(p), nbytes))
        return p.value

    def stream(self):
        s = self.C.c_void_p()
        self.ok(self.h.hipStreamCreate(self.C.byref(s)))
        return s.value


@pytest.mark.parametrize("precision", [0, 1])
def test_solve_async_matches_blocking(ctx, hydro_small, precision):
    """kore_factor_async + kore_solve_async on device arrays, queued behind a
    caller's HIP stream: the FP64 solve equals the blocking kore_solve bit for
    bit; the complex64 solve with 3 refinement sweeps matches the FP64
    solution to 1e-10.  On every repetition the right-hand sides are copied
    into place on the caller's stream just before the call (the library must
    wait for that copy) and the solution is copied out on it right after (the
    stream must wait for the library); the solution array is cleared between
    repetitions."""
    par, A, B = hydro_small
    say("case: async solve precision", precision)
    hip = _Hip()
    S = _system(ctx, par, A, B)
    n = S.n
    rng = np.random.default_rng(9)
    b = np.asfortranarray(rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3)))
    sigma = 0.41 + 0.8j
    S.factor(sigma, 0)
    x_ref, _ = S.solve(b)
    nbytes = b.nbytes
    d_src, d_b, d_x, d_out = (hip.malloc(nbytes) for _ in range(4))
    st = hip.stream()
    H2D, D2H, D2D = 1, 2, 3
    try:
        hip.ok(hip.h.hipMemcpy(d_src, b.ctypes.data, nbytes, H2D))
        for rep in range(3):
            hip.ok(hip.h.hipMemsetAsync(d_b, 0, nbytes, st))
            S.factor_async(sigma, precision, stream=st)
            hip.ok(hip.h.hipMemcpyAsync(d_b, d_src, nbytes, D2D, st))       # b, just before the call
            S.solve_async(d_b, d_x, nrhs=3, ir_sweeps=0 if precision == 0 else 3, stream=st)
            hip.ok(hip.h.hipMemcpyAsync(d_out, d_x, nbytes, D2D, st))      # x, right after
            hip.ok(hip.h.hipMemsetAsync(d_x, 0, nbytes, st))
            hip.ok(hip.h.hipStreamSynchronize(st))
            S.factor_check()
            out = np.zeros_like(b)
            hip.ok(hip.h.hipMemcpy(out.ctypes.data, d_out, nbytes, D2H))
            if precision == 0:
                assert np.array_equal(out, x_ref), rep
            else:
                assert np.linalg.norm(out - x_ref) / np.linalg.norm(x_ref) <= 1e-10, rep
    finally:
        hip.h.hipStreamSynchronize(st)
        hip.h.hipStreamDestroy(st)
        for p in (d_src, d_b, d_x, d_out):
            hip.h.hipFree(p)
        S.close()


def test_prepared_system_equals_created(ctx, hydro_small):
    """A system set up in two steps (kore_sys_prepare on the host, then
    kore_sys_create_prepared) factors and solves bit for bit as one from
    kore_sys_create; a used plan is refused."""
    from kore_amd import backend
    par, A, B = hydro_small
    lay = backend.layout_from_params(par)
    plan = backend.SystemPlan(lay, A, B)
    S1 = backend.System(ctx, lay, None, None, plan=plan)
    S2 = _system(ctx, par, A, B)
    rng = np.random.default_rng(4)
    b = rng.standard_normal((A.shape[0], 2)) + 1j * rng.standard_normal((A.shape[0], 2))
    for S in (S1, S2):
        S.factor(0.2 + 0.9j, 0)
    x1, _ = S1.solve(b)
    x2, _ = S2.solve(b)
    assert np.array_equal(x1, x2)
    with pytest.raises(ValueError):
        backend.System(ctx, lay, None, None, plan=plan)
    S1.close()
    S2.close()


def test_prepare_threads_do_not_change_the_system(ctx):
    """The host set-up splits rows over kore_set_host_threads threads (one per
    4096 rows at most); systems prepared with one thread and with six factor
    and solve bit for bit alike (25 600 unknowns, so six ranges)."""
    from kore_amd.params import Params
    from kore_amd import assembly, backend
    par = Params(m=1, symm=1, Ek=1e-4, N=64, lmax=400)
    A, B, _ = assembly.assemble(par)
    lay = backend.layout_from_params(par)
    rng = np.random.default_rng(8)
    b = rng.standard_normal((A.shape[0], 2)) + 1j * rng.standard_normal((A.shape[0], 2))
    out = []
    try:
        for nt in (1, 6):
            backend.set_host_threads(nt)
            plan = backend.SystemPlan(lay, A, B)
            S = backend.System(ctx, lay, None, None, plan=plan)
            S.factor(0.3 + 0.8j, 0)
            out.append(S.solve(b)[0])
            S.close()
    finally:
        backend.set_host_threads(0)
    assert np.array_equal(out[0], out[1])
    # the shift lies close to an eigenvalue (|x| ~ 1e14 |b|): the solve is
    # checked by its normwise backward error, which the conditioning does not
    # inflate
    M = (A - (0.3 + 0.8j) * B).tocsr()
    nM = np.sqrt(np.sum(np.abs(M.data) ** 2))
    for j in range(2):
        x = out[0][:, j]
        berr = np.linalg.norm(M @ x - b[:, j]) / (nM * np.linalg.norm(x) + np.linalg.norm(b[:, j]))
        assert berr <= 1e-12, (j, berr)


def test_eigenvectors_match_oracle_up_to_phase(ctx, hydro_small):
    """The eigenvectors themselves, not only their residuals: each device
    eigenvector spans the same line as the oracle's (SuperLU + Krylov-Schur)
    for the same eigenvalue, |<x, y>| / (|x| |y|) >= 1 - 1e-9 (the pairs of
    this problem are well separated: the nearest other eigenvalue is more
    than 1e-3 |lambda| away)."""
    from oracle import krylov_schur
    par, A, B = hydro_small
    S = _system(ctx, par, A, B)
    lam, X, _, info, rc = S.eigs_native(par.tau, nev=5, tol=1e-13)
    lu = sla.splu((A - par.tau * B).tocsc())
    lo, Xo, _, _ = krylov_schur.krylov_schur(lambda x: lu.solve(B @ x), A.shape[0], par.tau, 5, tol=1e-13,
                                             maxit=200)
    for i in range(5):
        j = int(np.argmin(np.abs(lam - lo[i])))
        assert abs(lam[j] - lo[i]) <= 1e-10 * abs(lo[i])
        others = np.delete(lo, i)
        assert np.min(np.abs(others - lo[i])) > 1e-3 * abs(lo[i])
        x, y = X[:, j], Xo[:, i]
        c = abs(np.vdot(x, y)) / (np.linalg.norm(x) * np.linalg.norm(y))
        say("  lambda", lo[i], "|cos| - 1 = %.2e" % (c - 1))
        assert c >= 1 - 1e-9
    S.close()
