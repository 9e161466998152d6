"""Host-only dense routines of the projected eigenproblem (kore_schur,
kore_schur_reorder), checked against LAPACK (scipy's zgees / ztrsen) and
against what the mathematics fixes: T = Q S Q^H with Q unitary and S upper
triangular, the eigenvalues of T on S's diagonal, and after a reordering the
selected eigenvalues leading with the same invariant subspace as LAPACK's.
These run on the CPU (the library loads without a GPU; no device call)."""
import numpy as np
import pytest
import scipy.linalg as sla


def _lib_or_skip():
    from kore_amd import backend
    try:
        backend.lib()
    except backend.KoreLibraryError as e:       # pragma: no cover
        pytest.skip(str(e))
    return backend


def _match(a, b):
    """max over a of the distance to the nearest unused entry of b."""
    b = list(b)
    worst = 0.0
    for x in a:
        d = [abs(x - y) for y in b]
        k = int(np.argmin(d))
        worst = max(worst, d[k])
        b.pop(k)
    return worst


def _cases():
    rng = np.random.default_rng(42)
    out = []
    for n in (0, 1, 2, 3, 8, 31, 64, 97):
        out.append(("random%d" % n, rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))))
    # an Arnoldi-shaped matrix: upper Hessenberg with one full trailing row block
    H = np.triu(rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40)), -1)
    out.append(("hessenberg40", H))
    out.append(("zero5", np.zeros((5, 5), dtype=complex)))
    out.append(("diag6", np.diag(np.arange(6) + 1j)))
    # a Jordan block (defective, one repeated eigenvalue) plus a distinct one
    J = np.eye(5, dtype=complex) * (2 + 1j) + np.eye(5, k=1)
    J[4, 4] = -1.0
    out.append(("jordan5", J))
    # the thick-restart shape: triangular block plus a coupling row
    K = np.triu(rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20)))
    K[12:, :12] = 0
    K[12, :12] = rng.standard_normal(12)
    out.append(("restart20", K))
    return out


@pytest.mark.parametrize("name,T", _cases(), ids=[c[0] for c in _cases()])
def test_schur_matches_lapack(name, T):
    backend = _lib_or_skip()
    n = T.shape[0]
    S, Q = backend.schur(T)
    if n == 0:
        return
    nt = max(np.linalg.norm(T), 1.0)
    assert np.linalg.norm(Q @ S @ Q.conj().T - T) <= 1e-13 * nt * max(1, n / 10)
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(n)) <= 1e-13 * max(1, n / 10)
    assert np.all(np.tril(S, -1) == 0)
    S2, _ = sla.schur(T, output="complex")
    tol = 1e-12 * nt if name != "jordan5" else 1e-2       # a defective eigenvalue moves by eps^(1/4)
    assert _match(np.diag(S), np.diag(S2)) <= tol


@pytest.mark.parametrize("name,T", [c for c in _cases() if c[1].shape[0] >= 2], ids=[c[0] for c in _cases()
                                                                                  if c[1].shape[0] >= 2])
def test_reorder_matches_lapack(name, T):
    backend = _lib_or_skip()
    n = T.shape[0]
    rng = np.random.default_rng(n)
    S, Q = backend.schur(T)
    for sel in (rng.random(n) < 0.5, np.arange(n) >= n // 2, np.zeros(n, bool), np.ones(n, bool)):
        sel = sel.astype(np.int32)
        S3, Q3, m = backend.schur_reorder(S, Q, sel)
        assert m == int(sel.sum())
        nt = max(np.linalg.norm(T), 1.0)
        assert np.linalg.norm(Q3 @ S3 @ Q3.conj().T - T) <= 1e-13 * nt * max(1, n / 10)
        assert np.linalg.norm(Q3.conj().T @ Q3 - np.eye(n)) <= 1e-13 * max(1, n / 10)
        assert np.all(np.tril(S3, -1) == 0)
        # the selected eigenvalues lead, in their original relative order (exact copies)
        assert np.array_equal(np.diag(S3)[:m], np.diag(S)[sel.astype(bool)])
        if name == "jordan5" or m in (0, n):
            continue
        St, Qt, _, msel, _, _, info = sla.lapack.ztrsen(sel, S, Q, job="N")
        assert info == 0 and msel == m
        # the same invariant subspace as LAPACK's
        P1 = Q3[:, :m] @ Q3[:, :m].conj().T
        P2 = Qt[:, :m] @ Qt[:, :m].conj().T
        assert np.linalg.norm(P1 - P2) <= 1e-10


def test_schur_planted_error_is_caught():
    """The reconstruction check above fails for a wrong Schur vector: swap two
    columns of Q (a plausible indexing mistake) and the identity breaks."""
    backend = _lib_or_skip()
    rng = np.random.default_rng(3)
    T = rng.standard_normal((12, 12)) + 1j * rng.standard_normal((12, 12))
    S, Q = backend.schur(T)
    Qb = Q.copy()
    Qb[:, [3, 4]] = Qb[:, [4, 3]]
    assert np.linalg.norm(Qb @ S @ Qb.conj().T - T) > 1e-3 * np.linalg.norm(T)
