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
the restart combination V <- V Q.  The host holds only the small projected
matrix and computes its ordered Schur form in FP64 (scipy/LAPACK).

Krylov decomposition kept throughout:  OP V[:, :j] = V[:, :j+bs] H[:j+bs, :j].
Convergence: a Ritz pair (theta, y) is converged when ||Bres y|| <= tol |theta|
(relative criterion, SLEPc's default EPS_CONV_REL); pairs are counted in
wanted order, consecutively, as SLEPc does.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np
import scipy.linalg as sla

from .backend import PREC_FP64, Workspace


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


def _cholqr(ws: Workspace, src: int, dst: int, nc: int):
    """Orthonormalise V[:, src:src+nc] into V[:, dst:dst+nc]; returns R with
    V_src = V_dst R.  Two Cholesky-QR passes (the second in place via src)."""
    G = ws.dot(src, nc, src, nc)
    G = 0.5 * (G + G.conj().T)
    R1 = sla.cholesky(G, lower=False)
    ws.gemm(dst, nc, src, nc, sla.solve_triangular(R1, np.eye(nc), lower=False), alpha=1.0, beta=0.0)
    G2 = ws.dot(dst, nc, dst, nc)
    G2 = 0.5 * (G2 + G2.conj().T)
    R2 = sla.cholesky(G2, lower=False)
    ws.gemm(src, nc, dst, nc, sla.solve_triangular(R2, np.eye(nc), lower=False), alpha=1.0, beta=0.0)
    ws.copy(src, dst, nc)
    return R2 @ R1


def _triu_eigvecs(S):
    n = S.shape[0]
    d = np.diag(S)
    Z = np.zeros((n, n), dtype=complex)
    for i in range(n):
        z = np.zeros(n, dtype=complex)
        z[i] = 1.0
        for r in range(i - 1, -1, -1):
            den = d[r] - d[i]
            if den == 0:
                den = np.finfo(float).tiny
            z[r] = -(S[r, r + 1:i + 1] @ z[r + 1:i + 1]) / den
        Z[:, i] = z / np.linalg.norm(z)
    return Z


def eigs(sysm, sigma, nev, ncv=None, tol=1e-12, maxit=100, which="TM", block=8, seed=0,
         ir_tol=1e-15, ir_maxit=12, v0=None, log=None, workspace=Workspace) -> EigResult:
    """Eigenpairs nearest ``sigma`` of the pencil held by ``sysm``.

    ``sysm`` must have been factored at ``sigma`` (System.factor).
    ``workspace`` constructs the vector block (the device Workspace; tests of
    the host logic pass a stand-in with the same methods).
    """
    t0 = time.perf_counter()
    sigma = complex(sigma)
    n = sysm.n
    bs = max(1, min(int(block), 16))
    if ncv is None:
        ncv = max(2 * nev, nev + 15)
    ncv = max(ncv, nev + 2 * bs)
    ncv = int(np.ceil(ncv / bs) * bs)
    ncv = min(ncv, max(bs, (n // bs) * bs - bs))
    m = ncv
    # columns: [0, m+bs) basis + residual block ; [m+bs, m+2bs) new block ; [m+2bs, 2m+2bs) restart temp
    cW = m + bs
    cT = m + 2 * bs
    ws = workspace(sysm, 2 * m + 2 * bs)
    rng = np.random.default_rng(seed)
    if v0 is None:
        X0 = rng.standard_normal((n, bs)) + 1j * rng.standard_normal((n, bs))
    else:
        X0 = np.asarray(v0, dtype=complex).reshape(n, -1)
        if X0.shape[1] < bs:
            X0 = np.hstack([X0, rng.standard_normal((n, bs - X0.shape[1])) + 0j])
        X0 = X0[:, :bs]
    ws.set(cW, X0)
    _cholqr(ws, cW, 0, bs)
    H = np.zeros((m + bs, m), dtype=complex)
    j = 0
    restarts = 0
    op_cols = 0
    op_calls = 0
    ir_iters = 0
    nconv = 0
    mixed = sysm.precision != PREC_FP64
    while True:
        # ---- expand the decomposition to m columns
        while j < m:
            ws_info = ws.apply_op(j, cW, bs, ir_tol=ir_tol, ir_maxit=ir_maxit if mixed else 0)
            ir_iters = max(ir_iters, ws_info.iterations)
            op_cols += bs
            op_calls += 1
            nb_all = j + bs
            h = ws.dot(0, nb_all, cW, bs)
            ws.gemm(cW, bs, 0, nb_all, h, alpha=-1.0, beta=1.0)
            h2 = ws.dot(0, nb_all, cW, bs)
            ws.gemm(cW, bs, 0, nb_all, h2, alpha=-1.0, beta=1.0)
            h = h + h2
            try:
                R = _cholqr(ws, cW, j + bs, bs)
            except np.linalg.LinAlgError:
                # invariant subspace: continue with a fresh random block
                Xr = rng.standard_normal((n, bs)) + 1j * rng.standard_normal((n, bs))
                ws.set(cW, Xr)
                for _ in range(2):
                    hh = ws.dot(0, nb_all, cW, bs)
                    ws.gemm(cW, bs, 0, nb_all, hh, alpha=-1.0, beta=1.0)
                _cholqr(ws, cW, j + bs, bs)
                R = np.zeros((bs, bs), dtype=complex)
            H[:nb_all, j:j + bs] = h
            H[j + bs:j + 2 * bs, j:j + bs] = R
            j += bs
        # ---- Rayleigh-Ritz on the projected matrix (host, FP64)
        T = H[:m, :m]
        Bres = H[m:m + bs, :m]
        # Schur form ordered by the wanted criterion
        S, Q = sla.schur(T, output="complex")
        d = np.diag(S)
        kd = sort_key(which, sigma + 1.0 / d, sigma)
        perm = np.argsort(kd, kind="stable")
        S, Q = _reorder(S, Q, np.argsort(perm))
        theta = np.diag(S).copy()
        Z = _triu_eigvecs(S)
        Y = Q @ Z
        res = np.linalg.norm(Bres @ Y, axis=0)
        conv = res <= tol * np.abs(theta)
        nconv = 0
        while nconv < m and conv[nconv]:
            nconv += 1
        if log:
            log("restart %d: nconv %d, worst wanted residual %.3e" % (restarts, nconv,
                float(np.max(res[:nev] / np.abs(theta[:nev])))))
        if nconv >= nev or restarts >= maxit:
            break
        # ---- thick restart
        keep = nconv + max(1, (m - nconv) // 2)
        keep = max(keep, nev)
        keep = int(np.ceil(keep / bs) * bs)       # the expansion then ends exactly at m
        keep = max(bs, min(keep, m - bs))
        ws.gemm(cT, keep, 0, m, Q[:, :keep], alpha=1.0, beta=0.0)
        ws.copy(cT, 0, keep)
        ws.copy(m, keep, bs)
        Hn = np.zeros_like(H)
        Hn[:keep, :keep] = S[:keep, :keep]
        Hn[keep:keep + bs, :keep] = Bres @ Q[:, :keep]
        H = Hn
        j = keep
        restarts += 1
    k = nconv
    X = np.zeros((n, 0), dtype=complex)
    if k > 0:
        ws.gemm(cT, k, 0, m, Y[:, :k], alpha=1.0, beta=0.0)
        X = ws.get(cT, k)
        X = X / np.linalg.norm(X, axis=0)
    lam = sigma + 1.0 / theta[:k]
    ws.close()
    return EigResult(lam=lam, X=X, errest=res[:k] / np.abs(theta[:k]), nconv=k, restarts=restarts,
                     op_columns=op_cols, op_calls=op_calls, seconds=time.perf_counter() - t0,
                     ir_max_iterations=ir_iters)


def _reorder(S, Q, target_pos):
    """Reorder the complex Schur form so that diagonal entry i moves to
    position target_pos[i] (adjacent Givens swaps, the ztrexc method)."""
    S = S.copy()
    Q = Q.copy()
    n = S.shape[0]
    order = np.argsort(target_pos)          # order[p] = current index to place at p
    cur = list(range(n))
    for p, src in enumerate(order):
        pos = cur.index(src)
        while pos > p:
            i = pos - 1
            a, b, c = S[i, i], S[i + 1, i + 1], S[i, i + 1]
            x = np.array([c, b - a])
            nr = np.linalg.norm(x)
            if nr != 0:
                cs, sn = x[0] / nr, x[1] / nr
                G = np.array([[np.conj(cs), np.conj(sn)], [-sn, cs]])
                S[[i, i + 1], :] = G @ S[[i, i + 1], :]
                S[:, [i, i + 1]] = S[:, [i, i + 1]] @ G.conj().T
                Q[:, [i, i + 1]] = Q[:, [i, i + 1]] @ G.conj().T
                S[i + 1, i] = 0.0
            cur[pos - 1], cur[pos] = cur[pos], cur[pos - 1]
            pos -= 1
    return S, Q


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
