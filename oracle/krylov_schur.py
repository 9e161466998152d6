"""Shift-and-invert Krylov-Schur eigensolver.  TEST INFRASTRUCTURE ONLY.

What it restates: the eigensolver bin/solve.py:91-123 configures -- SLEPc EPS
of type GNHEP with the default Krylov-Schur method (G. W. Stewart, "A
Krylov-Schur algorithm for large eigenproblems", SIMAX 23, 2001) and the
spectral transformation ``-st_type sinvert``: it iterates on
OP = (A - sigma B)^{-1} B, whose eigenvalues theta = 1/(lambda - sigma) are
largest for lambda nearest the target sigma; lambda = sigma + 1/theta.
SLEPc itself is not vendored (un-pinned PETSc/SLEPc 3.12-3.19,
README.md:31-70, .github/workflows/main.yml:36-47); its published algorithm is
followed:

* Arnoldi expansion with classical Gram-Schmidt and one reorthogonalisation;
* Schur form of the projected matrix sorted by the wanted criterion, residual
  estimate ||beta * y_last|| of each Ritz pair, relative convergence test
  ``res / |theta| <= tol`` (SLEPc's EPS_CONV_REL), converged pairs counted
  consecutively in sorted order;
* thick restart keeping half of the non-converged Schur vectors.

Pinned by: dense QZ (scipy.linalg.eig, LAPACK zggev) on the assembled
plain-method case, ARPACK on the same operator, and the reference's spin-over
golden eigenvalue (tests/spinover/reference.eig:1), in tests/test_oracle.py.
"""
import numpy as np
import scipy.linalg as sla


def sort_key(which, lam, tau):
    """Ascending key: the wanted eigenvalues first (bin/solve.py:99-117)."""
    which = which.upper()
    if which == "TM":
        return np.abs(lam - tau)
    if which == "TR":
        return np.abs(lam.real - tau.real)
    if which == "TI":
        return np.abs(lam.imag - tau.imag)
    if which == "LM":
        return -np.abs(lam)
    if which == "SM":
        return np.abs(lam)
    if which == "LR":
        return -lam.real
    if which == "SR":
        return lam.real
    if which == "LI":
        return -lam.imag
    if which == "SI":
        return lam.imag
    raise ValueError(which)


def krylov_schur(apply_op, n, sigma, nev, ncv=None, tol=1e-12, maxit=100, which="TM",
                 v0=None, seed=0):
    """Eigenpairs of the pencil via OP = (A - sigma B)^{-1} B.

    apply_op(x) -> OP x for a vector x of length n.
    Returns (lam[nconv], X[n, nconv], resest[nconv], iterations).
    """
    if ncv is None:
        ncv = max(2 * nev, nev + 15)
    ncv = min(ncv, n)
    rng = np.random.default_rng(seed)
    V = np.zeros((n, ncv + 1), dtype=complex)
    H = np.zeros((ncv + 1, ncv), dtype=complex)
    v = v0 if v0 is not None else rng.standard_normal(n) + 1j * rng.standard_normal(n)
    V[:, 0] = v / np.linalg.norm(v)
    j0 = 0
    nconv = 0
    it = 0
    while True:
        it += 1
        for j in range(j0, ncv):
            w = apply_op(V[:, j])
            h = V[:, :j + 1].conj().T @ w
            w = w - V[:, :j + 1] @ h
            h2 = V[:, :j + 1].conj().T @ w            # one reorthogonalisation (CGS2)
            w = w - V[:, :j + 1] @ h2
            h = h + h2
            beta = np.linalg.norm(w)
            H[:j + 1, j] = h
            H[j + 1, j] = beta
            V[:, j + 1] = w / beta
        T = H[:ncv, :ncv]
        brow = H[ncv, :ncv]
        theta_all = np.linalg.eigvals(T)
        lam_all = sigma + 1.0 / theta_all
        key_all = sort_key(which, lam_all, sigma)
        order = np.argsort(key_all, kind="stable")
        # Schur form with the wanted values leading, in the wanted order
        S, Q = sla.schur(T, output="complex")
        S, Q = _reorder(S, Q, lambda th: sort_key(which, sigma + 1.0 / th, sigma))
        theta = np.diag(S).copy()
        # residual estimates |b^T y_i| for the Ritz vectors y_i = Q z_i
        Z = _triu_eigvecs(S)
        Y = Q @ Z
        res = np.abs(brow @ Y)
        conv = res / np.abs(theta) <= tol
        nconv = 0
        while nconv < ncv and conv[nconv]:
            nconv += 1
        if nconv >= nev or it >= maxit:
            break
        keep = nconv + max(1, (ncv - nconv) // 2)
        keep = min(keep, ncv - 1)
        Vk = V[:, :ncv] @ Q[:, :keep]
        V[:, :keep] = Vk
        V[:, keep] = V[:, ncv]
        Hn = np.zeros_like(H)
        Hn[:keep, :keep] = S[:keep, :keep]
        Hn[keep, :keep] = brow @ Q[:, :keep]
        H = Hn
        j0 = keep
    k = nconv
    X = V[:, :ncv] @ Y[:, :k]
    X /= np.linalg.norm(X, axis=0)
    lam = sigma + 1.0 / theta[:k]
    return lam, X, res[:k] / np.abs(theta[:k]), it


def _reorder(S, Q, key):
    """Reorder a complex Schur form so that diagonal keys ascend (bubble of
    adjacent Givens swaps, the standard ztrexc approach)."""
    S = S.copy(); Q = Q.copy()
    n = S.shape[0]
    keys = key(np.diag(S))
    order = np.argsort(keys, kind="stable")
    # move entries one by one to their slot
    cur = list(range(n))
    for target, src in enumerate(order):
        pos = cur.index(src)
        while pos > target:
            _swap(S, Q, pos - 1)
            cur[pos - 1], cur[pos] = cur[pos], cur[pos - 1]
            pos -= 1
    return S, Q


def _swap(S, Q, i):
    """Swap adjacent diagonal entries i, i+1 of upper-triangular S."""
    a, b, c = S[i, i], S[i + 1, i + 1], S[i, i + 1]
    x = np.array([c, b - a])
    nrm = np.linalg.norm(x)
    if nrm == 0:
        return
    cs, sn = x[0] / nrm, x[1] / nrm
    G = np.array([[cs.conjugate(), sn.conjugate()], [-sn, cs]])
    S[[i, i + 1], :] = G @ S[[i, i + 1], :]
    S[:, [i, i + 1]] = S[:, [i, i + 1]] @ G.conj().T
    Q[:, [i, i + 1]] = Q[:, [i, i + 1]] @ G.conj().T
    S[i + 1, i] = 0.0


def _triu_eigvecs(S):
    """Eigenvectors of an upper-triangular matrix (columns, unit norm)."""
    n = S.shape[0]
    Z = np.zeros((n, n), dtype=complex)
    d = np.diag(S)
    for i in range(n):
        z = np.zeros(n, dtype=complex)
        z[i] = 1.0
        for r in range(i - 1, -1, -1):
            den = d[r] - d[i]
            if den == 0:
                den = 1e-300
            z[r] = -(S[r, r + 1:i + 1] @ z[r + 1:i + 1]) / den
        Z[:, i] = z / np.linalg.norm(z)
    return Z
