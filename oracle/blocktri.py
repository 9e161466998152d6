"""FP64 block-tridiagonal LU (block Thomas).  TEST INFRASTRUCTURE ONLY.

What it restates: the sparse direct LU that bin/solve.py obtains from
PETSc/MUMPS for the shift-and-invert operator (solve.py:123, ST sinvert) and
for the forced problem (solve.py:220-227, KSP preonly + PC lu).  Any exact
direct factorisation of A - sigma*B gives the same solution up to rounding;
this one eliminates degree by degree in l-major order, with LAPACK partial
pivoting inside each diagonal block:

    S_0 = D_0,  X_k = S_k^{-1} U_k,  S_{k+1} = D_{k+1} - L_{k+1} X_k
    y_k = S_k^{-1} (b_k - L_k y_{k-1}),   x_k = y_k - X_k x_{k+1}

Pinned by: scipy.sparse.linalg.splu (SuperLU) and numpy.linalg.solve on the
same assembled systems, and the residual ||Mx - b|| / (||M||_F ||x|| + ||b||)
at the 1e-13 level (tests/test_oracle.py).
"""
import numpy as np
import scipy.linalg as sla


class BlockThomas:
    def __init__(self, D, L, U):
        self.n = len(D)
        self.L = L
        self.lu, self.X = [], []
        S = D[0].astype(complex)
        for k in range(self.n):
            lu = sla.lu_factor(S, check_finite=False)
            self.lu.append(lu)
            if k < self.n - 1:
                X = sla.lu_solve(lu, U[k], check_finite=False)
                self.X.append(X)
                S = D[k + 1] - L[k + 1] @ X

    def solve(self, b, off):
        """b: (sizmat,) or (sizmat, nrhs) in l-major order."""
        b = np.asarray(b, dtype=complex)
        y = [None] * self.n
        for k in range(self.n):
            r = b[off[k]:off[k + 1]]
            if k > 0:
                r = r - self.L[k] @ y[k - 1]
            y[k] = sla.lu_solve(self.lu[k], r, check_finite=False)
        x = [None] * self.n
        x[-1] = y[-1]
        for k in range(self.n - 2, -1, -1):
            x[k] = y[k] - self.X[k] @ x[k + 1]
        return np.concatenate(x, axis=0)


def relres(M, x, b):
    """||Mx - b|| / (||M||_F ||x|| + ||b||), the forced-solve residual."""
    import scipy.sparse as ss
    M = ss.csr_matrix(M)
    r = M @ x - b
    nf = np.sqrt(np.sum(np.abs(M.data) ** 2))
    return np.linalg.norm(r) / (nf * np.linalg.norm(x) + np.linalg.norm(b))
