"""FP64 block cyclic reduction with explicit block inverses.  TEST INFRASTRUCTURE ONLY.

This is the algorithm the HIP factorisation runs, written plainly so that the
device buffers can be compared level by level.  At a level with blocks
t = 0..T-1 (odd t are eliminated, even t kept):

    Dinv_t = D_t^{-1}                                   (t odd)
    E_t = L_t Dinv_{t-1},  F_t = U_t Dinv_{t+1}          (t even)
    D_t <- D_t - E_t U_{t-1} - F_t L_{t+1}
    L_t <- -E_t L_{t-1}   (couples to t-2),   U_t <- -F_t U_{t+1}

until one block remains (Dinv of it is stored).  Solve:

    forward,  per level:   b_t <- b_t - E_t b_{t-1} - F_t b_{t+1}   (t even)
    root:                  x_0 = Dinv_0 b_0
    backward, per level:   x_t = Dinv_t (b_t - L_t x_{t-1} - U_t x_{t+1})  (t odd)

With tail=True (what the device runs) the last level (two blocks, t = 0, 1)
and the root are one step: with G_1 = Dinv_1 L_1, P = Dinv_0 F_0,
R = G_1 Dinv_0 and Q = Dinv_1 + G_1 P,

    x_0 = Dinv_0 b_0 - P b_1,    x_1 = Q b_1 - R b_0

(substituting the last forward step into the root and the root into the
last backward step).

It solves the same linear system as the reference's direct solve
(bin/solve.py:123, :227), so it is pinned by agreement with
``blocktri.BlockThomas`` and scipy's splu on assembled systems, and by the
residual bound (tests/test_oracle.py).  Row equilibration by powers of two
(exact in binary floating point) is applied first, as the device path does.
"""
import numpy as np


def pow2_row_scale(D, L, U):
    """Per-row scale 2^-e with 2^e the nearest power of two to max|row|."""
    s = []
    for k in range(len(D)):
        mx = np.abs(D[k]).max(axis=1)
        if L[k] is not None:
            mx = np.maximum(mx, np.abs(L[k]).max(axis=1))
        if U[k] is not None:
            mx = np.maximum(mx, np.abs(U[k]).max(axis=1))
        e = np.where(mx > 0, np.round(np.log2(np.where(mx > 0, mx, 1.0))), 0.0)
        s.append(2.0 ** (-e))
    return s


class BCR:
    def __init__(self, D, L, U, equilibrate=True, dtype=complex, tail=True):
        n = len(D)
        self.dtype = dtype
        if equilibrate:
            self.rs = pow2_row_scale(D, L, U)
        else:
            self.rs = [np.ones(d.shape[0]) for d in D]
        rs = self.rs
        D = [(rs[k][:, None] * D[k]).astype(dtype) for k in range(n)]
        L = [None if L[k] is None else (rs[k][:, None] * L[k]).astype(dtype) for k in range(n)]
        U = [None if U[k] is None else (rs[k][:, None] * U[k]).astype(dtype) for k in range(n)]
        self.levels = []
        ids = list(range(n))
        while len(ids) > 1:
            T = len(ids)
            lev = dict(ids=list(ids), Dinv={}, E={}, F={}, L={}, U={})
            for t in range(1, T, 2):
                lev["Dinv"][t] = np.linalg.inv(D[t])
                lev["L"][t] = L[t]
                lev["U"][t] = U[t]
            nD, nL, nU = [], [], []
            for t in range(0, T, 2):
                d = D[t].copy()
                lt = ut = None
                if t - 1 >= 0:
                    E = L[t] @ lev["Dinv"][t - 1]
                    lev["E"][t] = E
                    d -= E @ U[t - 1]
                    if t - 2 >= 0:
                        lt = -(E @ L[t - 1])
                if t + 1 <= T - 1:
                    F = U[t] @ lev["Dinv"][t + 1]
                    lev["F"][t] = F
                    d -= F @ L[t + 1]
                    if t + 2 <= T - 1:
                        ut = -(F @ U[t + 1])
                nD.append(d); nL.append(lt); nU.append(ut)
            self.levels.append(lev)
            D, L, U = nD, nL, nU
            ids = ids[0::2]
        self.root = np.linalg.inv(D[0])
        self.tail = None
        if tail and self.levels:
            last = self.levels[-1]
            G1 = last["Dinv"][1] @ last["L"][1]
            P = self.root @ last["F"][0]
            self.tail = dict(P=P, R=G1 @ self.root, Q=last["Dinv"][1] + G1 @ P)

    def solve(self, b, off):
        b = np.asarray(b, dtype=complex)
        n = len(off) - 1
        bb = [self.rs[k].reshape((-1,) + (1,) * (b.ndim - 1)) * b[off[k]:off[k + 1]]
              for k in range(n)]
        bb = [v.astype(self.dtype) for v in bb]
        stack = []
        cur = bb
        levels = self.levels[:-1] if self.tail else self.levels
        for lev in levels:
            T = len(cur)
            new = []
            for t in range(0, T, 2):
                v = cur[t].copy()
                if t in lev["E"]:
                    v -= lev["E"][t] @ cur[t - 1]
                if t in lev["F"]:
                    v -= lev["F"][t] @ cur[t + 1]
                new.append(v)
            stack.append(cur)
            cur = new
        if self.tail:
            tl = self.tail
            x = [self.root @ cur[0] - tl["P"] @ cur[1], tl["Q"] @ cur[1] - tl["R"] @ cur[0]]
        else:
            x = [self.root @ cur[0]]
        for lev, bl in zip(reversed(levels), reversed(stack)):
            T = len(bl)
            out = [None] * T
            for t in range(0, T, 2):
                out[t] = x[t // 2]
            for t in range(1, T, 2):
                r = bl[t].copy()
                if lev["L"][t] is not None:
                    r -= lev["L"][t] @ out[t - 1]
                if lev["U"][t] is not None:
                    r -= lev["U"][t] @ out[t + 1]
                out[t] = lev["Dinv"][t] @ r
            x = out
        return np.concatenate(x, axis=0)
