"""CPU tests of the host side: the C-ABI library (load, exports, host-only
layout entry points) and the block Krylov-Schur driver's logic, run against
a numpy stand-in of the device workspace (no GPU needed)."""
import os
import re

import numpy as np
import pytest
import scipy.linalg as la
import scipy.sparse.linalg as sla

from conftest import ROOT

from kore_amd import assembly, backend, eigs
from kore_amd.params import Params
from oracle import layout


def header_functions():
    src = open(os.path.join(ROOT, "include", "kore_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kore_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = backend.lib()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    bound = {s[0] for s in backend.SIGNATURES}
    assert set(names) == bound


@pytest.mark.parametrize("m,lmax,symm,N1,kw", [
    (1, 40, 1, 24, {}), (1, 64, -1, 68, {}), (0, 9, 1, 10, {}), (3, 10, -1, 6, {}),
    (2, 17, 1, 16, dict(magnetic=1, thermal=1)), (2, 17, 1, 16, dict(magnetic=1, thermal=1, compositional=1)),
])
def test_library_layout_matches_oracle(m, lmax, symm, N1, kw):
    lay = backend.make_layout(m, lmax, symm, N1, **kw)
    n, sizes, perm = backend.layout_blocks(lay)
    po, so = layout.lmajor_permutation(m, lmax, symm, N1, **kw)
    np.testing.assert_array_equal(perm, po)
    np.testing.assert_array_equal(sizes, so)


def test_library_rejects_bad_layout():
    lay = backend.make_layout(2, 200, 1, 150)        # lmax - m + 1 odd
    with pytest.raises(backend.KoreError):
        backend.layout_blocks(lay)


def test_factor_bytes_prediction():
    small = backend.factor_bytes(backend.make_layout(1, 40, 1, 24), 0)
    big = backend.factor_bytes(backend.make_layout(1, 300, 1, 200), 0)
    c64 = backend.factor_bytes(backend.make_layout(1, 300, 1, 200), 1)
    assert 0 < small < big
    assert 0.45 < c64 / big < 0.55


def test_affine_plan_argument_checks():
    """kore_sys_prepare_affine refuses a non-finite coefficient and A0, A1
    arrays that do not match the pattern, before any host or device work."""
    import scipy.sparse as ss
    par = Params(m=1, symm=1, Ek=1e-3, N=12, lmax=8)
    A, B = assembly.assemble(par, which="unscaled")
    A = ss.csr_matrix(A)
    lay = backend.layout_from_params(par)
    with pytest.raises(backend.KoreError, match="not finite"):
        backend.SystemPlan.affine(lay, A.indptr, A.indices, A.data, A.data, float("nan"), B, A.shape)
    with pytest.raises(ValueError):
        backend.SystemPlan.affine(lay, A.indptr, A.indices, A.data, A.data[:-1], 1.0, B, A.shape)


def test_host_thread_setting():
    """kore_set_host_threads: a positive count is used as given, 0 restores
    the default (the process's CPU share, at most 16, unless
    KORE_AMD_HOST_THREADS sets it), more than 64 is refused."""
    import os
    share = backend.cpu_share()
    assert 1 <= share <= (os.cpu_count() or share)
    default = backend.host_threads()
    try:
        backend.set_host_threads(3)
        assert backend.host_threads() == 3
        with pytest.raises(backend.KoreError):
            backend.set_host_threads(65)
        assert backend.host_threads() == 3
        backend.set_host_threads(0)
        assert backend.host_threads() == default
        if "KORE_AMD_HOST_THREADS" not in os.environ:
            assert default == min(share, 16)
    finally:
        backend.set_host_threads(0)


class NumpyWorkspace:
    """Stand-in for backend.Workspace: same methods, numpy arrays, OP by SuperLU."""

    def __init__(self, sysm, ncol):
        self.sys = sysm
        self.n = sysm.n
        self.V = np.zeros((sysm.n, ncol), dtype=complex)

    def set(self, c0, X):
        X = np.asarray(X).reshape(self.n, -1)
        self.V[:, c0:c0 + X.shape[1]] = X

    def get(self, c0, nc):
        return self.V[:, c0:c0 + nc].copy()

    def apply_op(self, src, dst, nc, ir_tol=0, ir_maxit=0):
        assert not (src < dst + nc and dst < src + nc)
        self.V[:, dst:dst + nc] = self.sys.lu.solve(self.sys.B @ self.V[:, src:src + nc])

        class I:
            iterations = 1
        return I()

    def dot(self, a0, na, b0, nb):
        return self.V[:, a0:a0 + na].conj().T @ self.V[:, b0:b0 + nb]

    def gemm(self, d0, nd, s0, ns, Q, alpha=1.0, beta=0.0):
        assert not (s0 < d0 + nd and d0 < s0 + ns)
        self.V[:, d0:d0 + nd] = beta * self.V[:, d0:d0 + nd] + alpha * self.V[:, s0:s0 + ns] @ Q

    def copy(self, src, dst, nc):
        self.V[:, dst:dst + nc] = self.V[:, src:src + nc]

    def close(self):
        pass


class StoppableNumpyWorkspace(NumpyWorkspace):
    """NumpyWorkspace with the library's expansion contract: expand() block by
    block (host-driven), honouring stop_expand(j_end) requests made from
    another thread between blocks, as kore_ws_expand / kore_ws_stop_expand."""

    def __init__(self, sysm, ncol):
        super().__init__(sysm, ncol)
        self.stop_at = None
        self.requests = []
        self.stopped = []

    def stop_expand(self, j_end=0):
        self.stop_at = j_end
        if j_end is not None:
            self.requests.append(j_end)

    def expand(self, j0, j1, bs, cw, H, ir_tol=1e-14, ir_maxit=0, ir_stop=0):
        j = j0
        while j < j1:
            if self.stop_at is not None and j >= self.stop_at:
                self.stopped.append((j0, j1, j))
                break
            st, jd, info = eigs.expand_host(self, j, j + bs, bs, cw, H, ir_tol, ir_maxit, ir_stop)
            if st != 0:
                return st, jd, info
            j = jd
        return 0, j, eigs._Info(1, True)


class NumpySystem:
    def __init__(self, A, B, sigma):
        self.n = A.shape[0]
        self.B = B
        self.lu = sla.splu((A - sigma * B).tocsc())
        self.precision = backend.PREC_FP64


@pytest.mark.parametrize("block", [1, 3, 4, 8])
def test_block_krylov_schur_logic_vs_qz(block):
    par = Params(m=1, symm=1, Ek=1e-3, N=16, lmax=20)
    A, B, _ = assembly.assemble(par)
    sigma = 0.9j
    w = la.eig(A.toarray(), B.toarray(), right=False)
    w = w[np.isfinite(w)]
    w = w[np.argsort(np.abs(w - sigma))]
    S = NumpySystem(A, B, sigma)
    r = eigs.eigs(S, sigma, nev=6, tol=1e-13, block=block, workspace=NumpyWorkspace)
    assert r.nconv >= 6
    for i in range(6):
        assert abs(r.lam[i] - w[i]) / abs(w[i]) <= 1e-10
    res = eigs.residuals(A, B, r.lam[:6], r.X[:, :6])
    assert res.max() <= 1e-13


@pytest.mark.parametrize("which", ["TR", "TI"])
def test_block_krylov_schur_other_targets(which):
    par = Params(m=1, symm=1, Ek=1e-3, N=16, lmax=20)
    A, B, _ = assembly.assemble(par)
    sigma = -0.05 + 0.5j
    w = la.eig(A.toarray(), B.toarray(), right=False)
    w = w[np.isfinite(w)]
    key = eigs.sort_key(which, w, sigma)
    S = NumpySystem(A, B, sigma)
    r = eigs.eigs(S, sigma, nev=3, tol=1e-12, block=4, which=which, workspace=NumpyWorkspace)
    assert r.nconv >= 3
    # every returned eigenvalue is a true eigenvalue of the pencil
    for lam in r.lam[:3]:
        assert np.min(np.abs(w - lam)) / abs(lam) <= 1e-9


def test_restart_path_is_exercised():
    """Small subspace forces thick restarts; the result must not change."""
    par = Params(m=1, symm=1, Ek=1e-3, N=16, lmax=20)
    A, B, _ = assembly.assemble(par)
    sigma = 0.9j
    S = NumpySystem(A, B, sigma)
    r1 = eigs.eigs(S, sigma, nev=4, ncv=12, tol=1e-12, block=2, workspace=NumpyWorkspace)
    r2 = eigs.eigs(S, sigma, nev=4, ncv=40, tol=1e-12, block=4, workspace=NumpyWorkspace)
    assert r1.restarts >= 1
    for i in range(4):
        assert abs(r1.lam[i] - r2.lam[i]) / abs(r2.lam[i]) <= 1e-10


def test_early_convergence_checks():
    """The Rayleigh-Ritz checks made while the basis grows (host step beside
    the device expansion) stop the iteration as soon as the wanted pairs have
    converged: fewer OP applications than a full expansion to ncv, the same
    eigenvalues as dense QZ, and residuals at the tolerance."""
    par = Params(m=1, symm=1, Ek=1e-3, N=16, lmax=20)
    A, B, _ = assembly.assemble(par)
    sigma = 0.9j
    w = la.eig(A.toarray(), B.toarray(), right=False)
    w = w[np.isfinite(w)]
    w = w[np.argsort(np.abs(w - sigma))]
    S = NumpySystem(A, B, sigma)
    full = eigs.eigs(S, sigma, nev=4, ncv=120, tol=1e-12, block=2, check_every=0, workspace=NumpyWorkspace)
    early = eigs.eigs(S, sigma, nev=4, ncv=120, tol=1e-12, block=2, check_every=8, workspace=NumpyWorkspace)
    assert early.op_columns % 8 == 0 or early.op_columns == 120
    assert full.op_columns == 120 and full.restarts == 0
    assert early.op_columns < full.op_columns and early.restarts == 0
    for r in (full, early):
        assert r.nconv >= 4
        for i in range(4):
            assert abs(r.lam[i] - w[i]) / abs(w[i]) <= 1e-10
        assert eigs.residuals(A, B, r.lam[:4], r.X[:, :4]).max() <= 1e-12


def test_expansion_stop_requests():
    """With a workspace that honours stop requests (as the device library
    does), the adaptive checks end an expansion chunk once a check beside it
    has converged, or at the column where the residual trend reaches tol:
    requests are made, the eigenpairs still match dense QZ with residuals at
    the tolerance, and fewer OP columns are applied than a full expansion."""
    par = Params(m=1, symm=1, Ek=1e-3, N=16, lmax=20)
    A, B, _ = assembly.assemble(par)
    sigma = 0.9j
    w = la.eig(A.toarray(), B.toarray(), right=False)
    w = w[np.isfinite(w)]
    w = w[np.argsort(np.abs(w - sigma))]
    S = NumpySystem(A, B, sigma)
    made = []

    def ws_factory(sysm, ncol):
        ws = StoppableNumpyWorkspace(sysm, ncol)
        made.append(ws)
        return ws

    full = eigs.eigs(S, sigma, nev=4, ncv=120, tol=1e-12, block=1, check_every=0, workspace=NumpyWorkspace)
    r = eigs.eigs(S, sigma, nev=4, ncv=120, tol=1e-12, block=1, workspace=ws_factory)
    ws = made[-1]
    assert ws.requests and ws.requests[-1] == 0          # the converged check stops its chunk
    assert r.nconv >= 4 and r.restarts == 0 and r.op_columns < full.op_columns
    for i in range(4):
        assert abs(r.lam[i] - w[i]) / abs(w[i]) <= 1e-10
    assert eigs.residuals(A, B, r.lam[:4], r.X[:, :4]).max() <= 1e-12
    # an end column set before the call ends the expansion exactly there
    ws2 = StoppableNumpyWorkspace(S, 12)
    ws2.set(0, np.ones((S.n, 1)) / np.sqrt(S.n))
    H = np.zeros((11, 10), dtype=complex, order="F")
    ws2.stop_expand(6)
    st, jd, _ = ws2.expand(0, 10, 1, 11, H)
    assert st == 0 and jd == 6 and ws2.stopped == [(0, 10, 6)]


def test_triangular_eigenvectors():
    """S z_j = d_j z_j for every column, unit norm, and the same directions as
    LAPACK's eigenvectors of the triangular matrix (equal up to a phase)."""
    rng = np.random.default_rng(5)
    n = 40
    S = np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    Z = eigs._triu_eigvecs(S)
    d = np.diag(S)
    assert np.abs(S @ Z - Z * d).max() <= 1e-10 * np.abs(S).max() * np.abs(Z).max() * n
    assert np.allclose(np.linalg.norm(Z, axis=0), 1.0)
    w, V = np.linalg.eig(S)
    for j in range(n):
        i = int(np.argmin(np.abs(w - d[j])))
        assert abs(abs(np.vdot(V[:, i], Z[:, j])) - 1.0) <= 1e-8
    # a repeated diagonal entry must not produce inf / nan
    S2 = S.copy()
    S2[3, 3] = S2[7, 7]
    assert np.all(np.isfinite(eigs._triu_eigvecs(S2)))


def test_breakdown_path_recovers(hydro_small):
    """A start block in the null space of B' breaks down at the first
    expansion (W = 0 exactly); the random replacement block must lead to the
    pencil's true eigenvalues nearest sigma (dense QZ)."""
    from conftest import pencil_with_null_block
    par, A, B = hydro_small
    A2, Bn, E = pencil_with_null_block(A, B)
    sigma = 1.0j
    S = NumpySystem(A2, Bn, sigma)
    r = eigs.eigs(S, sigma, nev=4, tol=1e-12, block=2, v0=E, workspace=NumpyWorkspace)
    assert r.breakdowns >= 1 and r.nconv >= 4
    w = la.eigvals(A2.toarray(), Bn.toarray())
    w = w[np.isfinite(w)]
    near = w[np.argsort(np.abs(w - sigma))[:4]]
    for lam in r.lam[:4]:
        assert np.min(np.abs(near - lam)) / abs(lam) <= 1e-9
