"""l-major block layout of Kore's matrices.  TEST INFRASTRUCTURE ONLY.

Kore stores unknowns section-major: every poloidal-velocity degree, then
every toroidal one, then the magnetic, thermal and compositional sections
(bin/assemble.py:613-614, 730-731, 826-827, 925-926, 1022-1023).  Each
section holds N1 radial coefficients for each degree l of its list
(``ell``, bin/utils.py:174-183).  Couplings reach only l +- 1
(bin/operators.py:65-113; bin/assemble.py:642-650, 739-747), so grouping
the unknowns by degree l makes A - sigma B block-tridiagonal.

Pinned by: structural invariants (a permutation; no nonzero outside the
three block diagonals, checked on every assembled fixture), brute force on a
tiny system, and the block sizes the reference's own layout implies.
"""
import numpy as np
import scipy.sparse as ss


def ell(m, lmax, vsymm):
    """Poloidal/toroidal degree lists (bin/utils.py:174-183)."""
    lm1 = lmax - m + 1
    s = int(vsymm * 0.5 + 0.5)
    sg = int(np.sign(m))
    ll = np.arange(m + 1 - sg, lmax + 2 - sg)
    idp = np.arange((sg + s) % 2, lm1, 2)
    idt = np.arange((sg + s + 1) % 2, lm1, 2)
    return ll[idp], ll[idt], ll


def sections(m, lmax, symm, N1, hydro=1, magnetic=0, thermal=0, compositional=0, symmB0=-1):
    """[(name, row_offset, degree_list)] in Kore's section-major order."""
    top, bot, _ = ell(m, lmax, symm)
    nb = (lmax - m + 1) // 2
    n = N1 * nb
    out, off = [], 0
    if hydro:
        out += [("upol", off, top), ("utor", off + n, bot)]
        off += 2 * n
    if magnetic:
        # bpol (f) and btor (g) take the FLOW's degree lists: loc_mag_f = loc_bot,
        # loc_mag_g = loc_top for an antisymmetric B0 (bin/assemble.py:69-74)
        f_l = top if symmB0 == 1 else bot
        g_l = bot if symmB0 == 1 else top
        out += [("bpol", off, f_l), ("btor", off + n, g_l)]
        off += 2 * n
    if thermal:
        out += [("theta", off, top)]
        off += n
    if compositional:
        out += [("xi", off, top)]
        off += n
    return out, off


def lmajor_permutation(m, lmax, symm, N1, **kw):
    """perm, block_sizes: new index i holds old unknown perm[i]."""
    secs, size = sections(m, lmax, symm, N1, **kw)
    _, _, ll = ell(m, lmax, symm)
    perm, sizes = [], []
    for l in ll:
        nb = 0
        for name, off, lst in secs:
            hit = np.where(lst == l)[0]
            if len(hit):
                perm.append(off + hit[0] * N1 + np.arange(N1))
                nb += N1
        if nb:
            sizes.append(nb)
    perm = np.concatenate(perm)
    assert len(perm) == size and len(np.unique(perm)) == size
    return perm, np.array(sizes, dtype=np.int64)


def permute(M, perm):
    M = ss.csr_matrix(M)
    return M[perm][:, perm].tocsr()


def block_tridiag(M, sizes):
    """Dense diagonal D[k], sub-diagonal L[k] (couples k -> k-1) and
    super-diagonal U[k] (couples k -> k+1) blocks of an l-ordered matrix.
    Raises if any nonzero lies outside the three block diagonals."""
    M = ss.csr_matrix(M)
    off = np.concatenate([[0], np.cumsum(sizes)])
    nblk = len(sizes)
    coo = M.tocoo()
    bi = np.searchsorted(off, coo.row, side="right") - 1
    bj = np.searchsorted(off, coo.col, side="right") - 1
    if np.any(np.abs(bi - bj) > 1):
        raise ValueError("matrix is not block-tridiagonal in this ordering")
    D, L, U = [], [], []
    dense = M.toarray() if M.shape[0] <= 20000 else None
    for k in range(nblk):
        r = slice(off[k], off[k + 1])
        get = (lambda rr, cc: dense[rr, cc]) if dense is not None else (lambda rr, cc: M[rr, cc].toarray())
        D.append(get(r, slice(off[k], off[k + 1])))
        L.append(get(r, slice(off[k - 1], off[k])) if k > 0 else None)
        U.append(get(r, slice(off[k + 1], off[k + 2])) if k < nblk - 1 else None)
    return D, L, U, off
