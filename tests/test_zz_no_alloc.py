"""kore_residual_noalloc (the allocation-free variant of kore_residual) gives
the existing entry point's residual norms, called through the C ABI, on two
calls in a row, the second with a smaller input (fewer unknowns and columns),
so that the static scratch array's leftover contents from the first call
must not leak into the second."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _norms(lib, S, fn, X, lam, b, shift):
    n, k = X.shape
    X = np.asfortranarray(X)
    out = np.zeros(k)
    lp = lam.ctypes.data_as(C.POINTER(C.c_double)) if lam is not None else None
    bp = np.asfortranarray(b).ctypes.data_as(C.POINTER(C.c_double)) if b is not None else None
    rc = getattr(lib, fn)(S.handle, k, X.ctypes.data_as(C.POINTER(C.c_double)), n, lp, bp, n,
                          shift.real, shift.imag, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, lib.kore_last_error()
    return out


def test_residual_noalloc_matches_residual():
    from kore_amd import assembly, backend
    from kore_amd.params import Params
    lib = backend.lib()
    ctx = backend.Context(0)
    rng = np.random.default_rng(11)
    cases = [(Params(m=1, symm=1, Ek=1e-4, N=48, lmax=60), 5, True),     # larger first: 2880 unknowns, 5 columns
             (Params(m=2, symm=1, Ek=1e-3, N=16, lmax=21), 2, False)]    # then smaller: 320 unknowns, 2 columns
    for par, k, with_b in cases:
        A, B, _ = assembly.assemble(par)
        S = backend.System(ctx, backend.layout_from_params(par), A, B)
        n = S.n
        print("case: residual_noalloc n %d columns %d with b %s" % (n, k, with_b), flush=True)
        X = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
        lam = np.ascontiguousarray(rng.standard_normal(k) + 1j * rng.standard_normal(k))
        b = (rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))) if with_b else None
        shift = 0.3 - 0.2j
        ref = _norms(lib, S, "kore_residual", X, lam, b, shift)
        got = _norms(lib, S, "kore_residual_noalloc", X, lam, b, shift)
        print("  max rel diff", np.max(np.abs(got - ref) / ref), flush=True)
        assert np.all(ref > 0)
        assert np.max(np.abs(got - ref) / ref) <= 1e-14
        S.close()
    ctx.close()
