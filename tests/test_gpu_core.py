"""Device parity tests of the HIP path against the oracle (MI355X only).

Tolerances: FP64 path and complex64+FP64-refinement path must reproduce the
oracle's solutions to relative error <= 1e-10 and eigenvalues to <= 1e-10
(the residual bar of the serving metric); the reference's spin-over golden is
checked at its own rtol 1e-8 (tests/test_spinover.py:21-29 of Kore).
"""
import sys

import numpy as np
import pytest
import scipy.sparse as ss
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
    lay = backend.layout_from_params(par)
    return backend.System(ctx, lay, A, B)


def test_ws_gemm_and_dot(ctx, hydro_small):
    """GEMM kernel (both operand modes) through the workspace primitives."""
    from kore_amd import backend
    par, A, B = hydro_small
    say("case: ws gemm/dot, sizmat", A.shape[0])
    S = _system(ctx, par, A, B)
    ws = backend.Workspace(S, 40)
    rng = np.random.default_rng(1)
    n = S.n
    V = rng.standard_normal((n, 20)) + 1j * rng.standard_normal((n, 20))
    ws.set(0, V)
    assert np.allclose(ws.get(0, 20), V, rtol=0, atol=0)
    H = ws.dot(0, 20, 3, 5)
    ref = V.conj().T @ V[:, 3:8]
    assert np.abs(H - ref).max() <= 1e-12 * np.abs(ref).max()
    Q = rng.standard_normal((20, 7)) + 1j * rng.standard_normal((20, 7))
    ws.gemm(25, 7, 0, 20, Q, alpha=1.0, beta=0.0)
    ref = V @ Q
    got = ws.get(25, 7)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    ws.gemm(25, 7, 0, 20, Q, alpha=-1.0, beta=1.0)
    assert np.abs(ws.get(25, 7)).max() <= 1e-12 * np.abs(ref).max()
    ws.close()
    S.close()


def test_ws_get_large_columns(ctx):
    """Downloads large enough to take the pinned staging and the multi-thread
    host copy (several MiB per chunk of 16 columns, a partial last chunk)
    return the uploaded columns exactly, in Kore's ordering."""
    from kore_amd.params import Params
    from kore_amd import assembly, backend
    par = Params(m=1, symm=1, Ek=1e-4, N=100, lmax=150)
    A, B, _ = assembly.assemble(par)
    S = _system(ctx, par, A, B)
    say("case: ws get, sizmat", S.n)
    ws = backend.Workspace(S, 40)
    rng = np.random.default_rng(4)
    V = rng.standard_normal((S.n, 37)) + 1j * rng.standard_normal((S.n, 37))
    ws.set(2, V)
    assert np.array_equal(ws.get(2, 37), V)
    assert np.array_equal(ws.get(20, 3), V[:, 18:21])
    ws.close()
    S.close()


def test_ws_random_start_block(ctx, hydro_small):
    """Device-generated start columns: entries in [-1, 1) (re and im), the
    same for the same seed, different across seeds and columns."""
    par, A, B = hydro_small
    S = _system(ctx, par, A, B)
    ws = S.get_workspace(8)
    ws.random(0, 3, 7)
    ws.random(3, 1, 7)
    ws.random(4, 1, 8)
    V = ws.get(0, 5)
    assert np.all(np.abs(V.real) <= 1.0) and np.all(np.abs(V.imag) <= 1.0)
    assert np.allclose(V[:, 0], V[:, 3])
    assert not np.allclose(V[:, 0], V[:, 4]) and not np.allclose(V[:, 0], V[:, 1])
    assert abs(np.mean(V[:, :3].real)) < 0.1 and 0.25 < np.var(V[:, :3].imag) < 0.42
    S.close()


def _solve_case(ctx, par, A, B, sigma, precision, nrhs=1, seed=0, ir_maxit=20):
    from kore_amd import backend
    S = _system(ctx, par, A, B)
    S.factor(sigma, precision)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((S.n, nrhs)) + 1j * rng.standard_normal((S.n, nrhs))
    b /= np.linalg.norm(b, axis=0)
    x, info = S.solve(b, ir_tol=1e-14, ir_maxit=ir_maxit)
    M = (A - sigma * B).tocsc()
    xr = sla.splu(M).solve(b)
    err = np.linalg.norm(x - xr, axis=0) / np.linalg.norm(xr, axis=0)
    S.close()
    return err, info


def test_solve_tiny_fp64(ctx):
    from kore_amd.params import Params
    from kore_amd import assembly
    par = Params(m=1, symm=1, Ek=1e-3, N=12, lmax=8)
    A, B, _ = assembly.assemble(par)
    say("case: tiny fp64 N=12 lmax=8 sizmat", A.shape[0])
    err, info = _solve_case(ctx, par, A, B, 0.3j, 0)
    say("  rel err", err)
    assert err.max() <= 1e-10


@pytest.mark.parametrize("precision", [0, 1])
def test_solve_hydro_small(ctx, hydro_small, precision):
    par, A, B = hydro_small
    say("case: hydro_small solve precision", precision)
    err, info = _solve_case(ctx, par, A, B, 1.0j, precision, nrhs=3)
    say("  rel err", err, "sweeps", info.iterations, "last correction", info.rel_correction)
    assert err.max() <= 1e-10


@pytest.mark.parametrize("case", ["hydro_N100", "hydro_N200", "thermal_N60", "hydro_N300"])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nrhs", [1, 2])
def test_solve_large_blocks(ctx, case, precision, nrhs):
    """Degree blocks of 96-512 rows take the multi-workgroup inversion (panel
    kernel + tiled MFMA update) and GEMM edge tiles of sizes that are not
    multiples of 64: 100, 200, 300, and thermal 120 / 60 in one level list.
    One right-hand side with the FP64 factor takes level 0 of the sweep
    through the sparse couplings.  Solves match SuperLU to 1e-10."""
    from kore_amd.params import Params
    from kore_amd import assembly
    kw = dict(hydro_N100=dict(N=100, lmax=14), hydro_N200=dict(N=200, lmax=10), hydro_N300=dict(N=300, lmax=8),
              thermal_N60=dict(N=60, lmax=16, thermal=1, Prandtl=1.0, Ra_gap=1e5))[case]
    par = Params(m=1, symm=1, Ek=1e-4, **kw)
    A, B, _ = assembly.assemble(par)
    say("case: large-block solve", case, "sizmat", A.shape[0], "precision", precision)
    # the FP64 factor is checked on its plain sweep (no refinement that could
    # hide a wrong sweep); the complex64 factor needs its refinement
    err, info = _solve_case(ctx, par, A, B, 0.4j, precision, nrhs=nrhs, seed=7,
                            ir_maxit=0 if precision == 0 else 20)
    say("  rel err", err, "sweeps", info.iterations)
    assert err.max() <= 1e-10
    if precision == 0:
        assert info.iterations == 1


@pytest.mark.parametrize("case", ["two_blocks", "four_blocks", "tiny", "hydro_N100", "thermal_N60"])
@pytest.mark.parametrize("precision", [0, 1])
def test_lean_form_matches_sweep_form(ctx, case, precision):
    """The lean factor form (no E = L Dinv, F = U Dinv products; kept blocks
    updated through G, H; a forward sweep level in two launches) solves as the
    sweep form does: both match SuperLU to 1e-10 on their plain sweep (FP64)
    or after refinement (complex64), and agree with each other to rounding.
    Two blocks: the tail alone (its F0 in the root block's H); four: one
    level before the tail; the others several levels after the sparse level
    0, thermal with two block sizes.  The lean form runs fewer GEMM flops and
    one more launch per level swept before the tail."""
    from kore_amd.params import Params
    from kore_amd import assembly, backend
    kw = dict(two_blocks=dict(N=20, lmax=2), four_blocks=dict(N=20, lmax=4), tiny=dict(N=12, lmax=8),
              hydro_N100=dict(N=100, lmax=14),
              thermal_N60=dict(N=60, lmax=16, thermal=1, Prandtl=1.0, Ra_gap=1e5))[case]
    par = Params(m=1, symm=1, Ek=1e-4, **kw)
    A, B, _ = assembly.assemble(par)
    sigma = 0.4j
    rng = np.random.default_rng(5)
    b = rng.standard_normal((A.shape[0], 2)) + 1j * rng.standard_normal((A.shape[0], 2))
    xr = sla.splu((A - sigma * B).tocsc()).solve(b)
    out, infos = {}, {}
    for form in (backend.FORM_SWEEP, backend.FORM_LEAN):
        say("case: factor form", form, case, "sizmat", A.shape[0], "precision", precision)
        S = _system(ctx, par, A, B)
        S.set_factor_form(form)
        infos[form] = S.factor(sigma, precision)
        for nrhs in (1, 2):
            x, info = S.solve(b[:, :nrhs], ir_tol=1e-14, ir_maxit=0 if precision == 0 else 20)
            err = np.linalg.norm(x - xr[:, :nrhs], axis=0) / np.linalg.norm(xr[:, :nrhs], axis=0)
            say("  form", form, "nrhs", nrhs, "rel err", err, "sweeps", info.iterations)
            assert err.max() <= 1e-10, (form, nrhs)
        out[form] = x
        S.close()
    fs, fl = infos[backend.FORM_SWEEP], infos[backend.FORM_LEAN]
    assert fs.form == backend.FORM_SWEEP and fl.form == backend.FORM_LEAN
    assert fl.levels == fs.levels
    if precision == 0:
        d = np.linalg.norm(out[backend.FORM_LEAN] - out[backend.FORM_SWEEP], axis=0) / np.linalg.norm(xr, axis=0)
        assert d.max() <= 1e-11
    # one y launch per level swept one by one (all but the tail's) on dense couplings
    dense_swept = max(0, fs.levels - 1 - (1 if fs.levels >= 2 else 0))
    assert fl.solve_launches - fs.solve_launches == dense_swept
    if fs.levels >= 3:
        assert fl.flops_gemm < fs.flops_gemm
    assert fl.flops_inv == fs.flops_inv


def test_eigs_blocks_wider_than_256(ctx):
    """FP64 eigensolve (single-vector Krylov, no refinement) on degree blocks
    of 300 rows: level 0 of every sweep takes the long preloading kernel for
    the inverse blocks and the sparse couplings; eigenvalues match the oracle
    (SuperLU + Krylov-Schur) to 1e-10, residuals <= 1e-10."""
    from kore_amd import assembly, eigs
    from kore_amd.params import Params
    from oracle import krylov_schur
    par = Params(m=1, symm=1, Ek=1e-4, N=300, lmax=8)
    A, B, _ = assembly.assemble(par)
    sigma = 0.4j
    say("case: eigs with 300-row blocks, sizmat", A.shape[0])
    S = _system(ctx, par, A, B)
    S.factor(sigma, 0)
    r = eigs.eigs(S, sigma, nev=4, tol=1e-12, maxit=200, block=1)
    lu = sla.splu((A - sigma * B).tocsc())
    lo, _, _, _ = krylov_schur.krylov_schur(lambda x: lu.solve(B @ x), A.shape[0], sigma, 4, tol=1e-13, maxit=200)
    assert r.nconv >= 4 and r.ir_max_iterations == 1
    for l in lo[:4]:
        d = np.min(np.abs(r.lam - l)) / abs(l)
        say("  lambda", l, "rel diff", d)
        assert d <= 1e-10
    assert eigs.residuals(A, B, r.lam[:4], r.X[:, :4]).max() <= 1e-10
    S.close()


def test_solve_many_rhs(ctx, hydro_small):
    par, A, B = hydro_small
    say("case: hydro_small solve 21 rhs fp64")
    err, _ = _solve_case(ctx, par, A, B, 0.5 + 0.7j, 0, nrhs=21, seed=3)
    assert err.max() <= 1e-10


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("block", [1, 2, 4])
@pytest.mark.parametrize("driver", ["native", "python"])
def test_eigs_hydro_small(ctx, hydro_small, precision, block, driver):
    from kore_amd import eigs
    from oracle import krylov_schur
    par, A, B = hydro_small
    say("case: hydro_small eigs precision", precision, "block", block, "driver", driver)
    S = _system(ctx, par, A, B)
    S.factor(par.tau, precision)
    r = eigs.eigs(S, par.tau, nev=5, tol=1e-12, maxit=100, block=block, driver=driver)
    say("  nconv", r.nconv, "restarts", r.restarts, "ops", r.op_columns, "fallback", r.fallback,
        "c64 ops", r.c64_op_columns, "ir max", r.c64_ir_max)
    if precision == 1:     # the pairs come from the complex64 factor with FP64 refinement
        assert not r.fallback and r.precision == 1 and r.c64_op_columns == r.op_columns
    lu = sla.splu((A - par.tau * B).tocsc())
    lam_o, X_o, _, _ = krylov_schur.krylov_schur(lambda x: lu.solve(B @ x), A.shape[0], par.tau, 5, tol=1e-13,
                                                  maxit=100)
    assert r.nconv >= 5
    for lo in lam_o[:5]:
        d = np.min(np.abs(r.lam - lo)) / abs(lo)
        say("  lambda", lo, "rel diff", d)
        assert d <= 1e-10
    res = eigs.residuals(A, B, r.lam[:5], r.X[:, :5])
    assert res.max() <= 1e-10
    S.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_spinover_golden(ctx, spinover, precision):
    from kore_amd import eigs
    par, A, B = spinover
    say("case: spinover eigs precision", precision)
    S = _system(ctx, par, A, B)
    S.factor(par.tau, precision)
    r = eigs.eigs(S, par.tau, nev=par.nev, tol=1e-13, maxit=100, block=4)
    say("  fallback", r.fallback, "c64 ops", r.c64_op_columns, "ir max", r.c64_ir_max)
    if precision == 1:
        assert not r.fallback and r.precision == 1
    i = np.argmax(r.lam.real)
    got = r.lam[i]
    say("  spin-over lambda", got)
    # power balance of the device eigenvector (bin/spin_doctor.py:205-242)
    from kore_amd import postprocess
    u, _ = postprocess.field_sections(par, r.X[:, i])
    dg = postprocess.diagnose(par, got, usol=u)
    say("  resid0 %.2e resid1 %.2e" % (dg.resid0, dg.resid1))
    assert dg.resid0 < 2e-3 and dg.resid1 < 1e-5
    np.testing.assert_allclose([got.real, got.imag], [SPINOVER_LAMBDA.real, SPINOVER_LAMBDA.imag], rtol=1e-8,
                               atol=1e-20)
    S.close()


@pytest.mark.parametrize("bs", [1, 5, 8, 16])
def test_expand_matches_host_driven(ctx, hydro_small, bs):
    """kore_ws_expand (coefficients kept on the device, Cholesky on the
    device) against the same expansion driven from the host through the
    workspace primitives (scipy Cholesky): projected matrix and basis agree."""
    from kore_amd import backend, eigs
    par, A, B = hydro_small
    say("case: expand vs host-driven, bs", bs)
    S = _system(ctx, par, A, B)
    S.factor(par.tau, 0)
    nblk = 4
    m = nblk * bs
    ncol = m + 2 * bs
    rng = np.random.default_rng(11)
    X0 = rng.standard_normal((S.n, bs)) + 1j * rng.standard_normal((S.n, bs))
    X0, _ = np.linalg.qr(X0)
    out = []
    for use_device in (True, False):
        ws = backend.Workspace(S, ncol)
        ws.set(0, X0)
        H = np.zeros((m + bs, m), dtype=complex, order="F")
        if use_device:
            st, jd, _ = ws.expand(0, m, bs, m + bs, H)
        else:
            st, jd, _ = eigs.expand_host(ws, 0, m, bs, m + bs, H)
        assert st == 0 and jd == m
        out.append((H.copy(), ws.get(0, m + bs)))
        ws.close()
    (Hd, Vd), (Hh, Vh) = out
    assert np.abs(Hd - Hh).max() <= 1e-10 * np.abs(Hh).max()
    assert np.abs(Vd - Vh).max() <= 1e-10 * np.abs(Vh).max()
    # the basis is orthonormal
    G = Vd.conj().T @ Vd
    assert np.abs(G - np.eye(m + bs)).max() <= 1e-12
    S.close()


@pytest.mark.parametrize("bs", [1, 2])
def test_expand_fused_matches_separate_launches(ctx, bs, monkeypatch):
    """Gram-Schmidt updates fused with the inner products that follow them
    (the default) against separate update and product launches
    (KORE_AMD_ORTHO_FUSED=0), on 70 000 unknowns (row chunks of 69 rows:
    partial tiles, and empty trailing chunks) with bases of up to 202 columns
    (the 64-, 128- and 256-column tiles).  The first blocks agree to rounding;
    over the whole expansion each path's basis is orthonormal and satisfies the
    Arnoldi relation OP V[:, :m] = V[:, :m+bs] H, with OP applied by separate
    solves on the host side of the library."""
    from kore_amd import assembly, backend
    from kore_amd.params import Params
    par = Params(m=1, symm=1, Ek=1e-4, N=100, lmax=700)
    A, B, _ = assembly.assemble(par)
    sigma = 0.4j
    say("case: fused vs separate Gram-Schmidt, bs", bs, "sizmat", A.shape[0])
    S = _system(ctx, par, A, B)
    S.factor(sigma, 0)
    m = 200
    ncol = m + 2 * bs
    rng = np.random.default_rng(5)
    X0 = rng.standard_normal((S.n, bs)) + 1j * rng.standard_normal((S.n, bs))
    X0, _ = np.linalg.qr(X0)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("KORE_AMD_ORTHO_FUSED", fused)
        ws = backend.Workspace(S, ncol)
        ws.set(0, X0)
        H = np.zeros((m + bs, m), dtype=complex, order="F")
        st, jd, _ = ws.expand(0, m, bs, m + bs, H)
        assert st == 0 and jd == m
        V = ws.get(0, m + bs)
        ws.close()
        G = V.conj().T @ V
        orth = np.abs(G - np.eye(m + bs)).max()
        Y, _ = S.solve(B @ V[:, :m], ir_maxit=0)
        arn = np.abs(Y - V @ H).max() / np.abs(Y).max()
        say("  fused", fused, "orthonormality", orth, "Arnoldi relation", arn)
        assert orth <= 1e-12
        assert arn <= 1e-10
        out[fused] = (H, V)
    (Hf, Vf), (Hs, Vs) = out["1"], out["0"]
    k = 8 * bs
    dH = np.abs(Hf[:k + bs, :k] - Hs[:k + bs, :k]).max() / np.abs(Hs[:k + bs, :k]).max()
    dV = np.abs(Vf[:, :k + bs] - Vs[:, :k + bs]).max() / np.abs(Vs[:, :k + bs]).max()
    say("  first", k, "columns: H diff", dH, "V diff", dV)
    assert dH <= 1e-10 and dV <= 1e-10
    S.close()


def test_expand_stop_request(ctx, hydro_small):
    """kore_ws_stop_expand from another thread ends a running expansion after
    its current block: status OK, a whole number of blocks done, and those
    columns (basis and projected matrix) equal an expansion that ran to the
    end; a request made before the call ends it at once, an end column ends
    it exactly there, and clearing the request lets the next call run in
    full."""
    import threading
    import time
    from kore_amd import backend
    par, A, B = hydro_small
    say("case: expand stop request")
    S = _system(ctx, par, A, B)
    S.factor(par.tau, 0)
    bs, m = 1, 48
    ncol = m + 2 * bs
    rng = np.random.default_rng(12)
    X0 = rng.standard_normal((S.n, bs)) + 1j * rng.standard_normal((S.n, bs))
    X0 /= np.linalg.norm(X0)
    ws = backend.Workspace(S, ncol)
    ws.set(0, X0)
    Hf = np.zeros((m + bs, m), dtype=complex, order="F")
    st, jd, _ = ws.expand(0, m, bs, m + bs, Hf)
    assert st == 0 and jd == m
    Vf = ws.get(0, m + bs)
    # stop requested before the call: nothing is expanded
    ws.stop_expand(0)
    H0 = np.zeros_like(Hf)
    st, jd, _ = ws.expand(0, m, bs, m + bs, H0)
    assert st == 0 and jd == 0
    # stop requested while the expansion runs on another thread
    ws.stop_expand(None)
    ws.set(0, X0)
    Hs = np.zeros_like(Hf)
    res = {}
    th = threading.Thread(target=lambda: res.update(r=ws.expand(0, m, bs, m + bs, Hs)))
    th.start()
    time.sleep(0.002)
    ws.stop_expand(0)
    th.join(timeout=60)
    assert not th.is_alive()
    st, jd, _ = res["r"]
    say("  stopped at column", jd, "of", m)
    assert st == 0 and 0 <= jd <= m and jd % bs == 0
    if jd > 0:
        Vs = ws.get(0, jd + bs)
        assert np.abs(Hs[:jd + bs, :jd] - Hf[:jd + bs, :jd]).max() <= 1e-12 * np.abs(Hf).max()
        assert np.abs(Vs - Vf[:, :jd + bs]).max() <= 1e-12
    # an end column: the expansion stops exactly there
    ws.stop_expand(20)
    ws.set(0, X0)
    He = np.zeros_like(Hf)
    st, jd, _ = ws.expand(0, m, bs, m + bs, He)
    assert st == 0 and jd == 20
    assert np.abs(He[:21, :20] - Hf[:21, :20]).max() <= 1e-12 * np.abs(Hf).max()
    # cleared: the next expansion runs in full
    ws.stop_expand(None)
    ws.set(0, X0)
    H2 = np.zeros_like(Hf)
    st, jd, _ = ws.expand(0, m, bs, m + bs, H2)
    assert st == 0 and jd == m
    assert np.abs(H2 - Hf).max() <= 1e-12 * np.abs(Hf).max()
    ws.close()
    S.close()


def test_expand_breakdown_recovers(ctx, hydro_small):
    """Start block in the null space of B': the device Cholesky flags the
    breakdown, the host continues with a random block, and the eigenvalues
    match the oracle on the same pencil."""
    from conftest import pencil_with_null_block
    from kore_amd import eigs
    from oracle import krylov_schur
    par, A, B = hydro_small
    A2, Bn, E = pencil_with_null_block(A, B)
    sigma = par.tau
    say("case: expand breakdown, null columns", np.nonzero(E)[0])
    S = _system(ctx, par, A2, Bn)
    S.factor(sigma, 0)
    r = eigs.eigs(S, sigma, nev=4, tol=1e-12, block=2, v0=E)
    assert r.breakdowns >= 1 and r.nconv >= 4
    lu = sla.splu((A2 - sigma * Bn).tocsc())
    lam_o, _, _, _ = krylov_schur.krylov_schur(lambda x: lu.solve(Bn @ x), A2.shape[0], sigma, 4, tol=1e-13,
                                               maxit=100)
    for lo in lam_o[:4]:
        assert np.min(np.abs(r.lam - lo)) / abs(lo) <= 1e-10
    S.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_solve_gmres(ctx, precision):
    """GMRES preconditioned by the factor reaches FP64 accuracy with a
    complex64 factor on a low-Ekman case (E=1e-6): normwise backward error
    <= 1e-15 (the refinement tolerance).  The forward error against SuperLU
    is then bounded by cond(A - sigma B) * 1e-15 (cond ~ 2e5 here): <= 1e-9
    for the complex64 factor, <= 1e-10 for the FP64 factor."""
    from kore_amd import assembly
    from kore_amd.params import Params
    par = Params(m=2, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-6, N=48, lmax=61)
    A0, B0 = assembly.assemble(par, which="unscaled")
    sigma = 0.7j
    S = _system(ctx, par, A0, B0)
    S.factor(sigma, precision)
    rng = np.random.default_rng(9)
    b = rng.standard_normal(S.n) + 1j * rng.standard_normal(S.n)
    say("case: gmres precision", precision, "n", S.n)
    x, info = S.solve_gmres(b, tol=1e-15, maxit=60)
    say("  iterations", info.iterations, "converged", info.converged, "backward error", info.rel_correction)
    assert info.converged and info.rel_correction <= 1e-15
    M = (A0 - sigma * B0).tocsc()
    xr = sla.splu(M).solve(b)
    fwd = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    say("  forward error vs SuperLU", fwd)
    assert fwd <= (1e-10 if precision == 0 else 1e-9)
    if precision == 0:
        assert info.iterations <= 3
    S.close()


def _device_growth(ctx, mkpar, nev, precision):
    from kore_amd import assembly, eigs

    def growth(ra):
        par = mkpar(ra)
        A, B, _ = assembly.assemble(par)
        S = _system(ctx, par, A, B)
        S.factor(0.0, precision)
        r = eigs.eigs(S, 0.0, nev=nev, tol=1e-14, which="TR", maxit=200)
        say("    Ra %.6e: fallback %s, c64 ops %d, ir max %d" % (ra, r.fallback, r.c64_op_columns, r.c64_ir_max))
        S.close()
        lam = r.lam[:nev]
        i = int(np.argmax(lam.real))
        return float(lam[i].real), lam[i]
    return growth


@pytest.mark.parametrize("case", ["dormy2004", "jones2000"])
@pytest.mark.parametrize("precision", [0, 1])
def test_onset_golden_device(ctx, case, precision):
    """The reference's onset tests through the device path: thermal
    convection pencils (degree blocks alternating 2N1 and N1; Dormy: inner
    core, N1 = 74; Jones: full sphere, N1 = 59), cyclic-reduction factor and
    block Krylov-Schur ('TR' about 0) inside the reference driver's Brent
    search; Ra_c and omega_c as printed by the driver must equal
    tests/dormy2004/reference.dormy04:1 and tests/jones2000/reference.jones:1
    at rtol 1e-8 (tests/test_convection_bouss.py)."""
    from conftest import DORMY_GOLDEN, JONES_GOLDEN, critical_rayleigh, params_dormy, params_jones
    mk, nev, golden, lo, hi = {"dormy2004": (params_dormy, 4, DORMY_GOLDEN, 1.6e6, 1.7e6),
                               "jones2000": (params_jones, 5, JONES_GOLDEN, 4.6e6, 4.8e6)}[case]
    say("case: onset", case, "precision", precision)
    ra, lam, nsolves = critical_rayleigh(_device_growth(ctx, mk, nev, precision), lo=lo, hi=hi)
    say("  Ra_c %.5e omega_c %.5e after %d eigensolves" % (ra, lam.imag, nsolves))
    got = np.array([float("%.5e" % ra), float("%.5e" % lam.imag)])
    np.testing.assert_allclose(got, [golden[2], golden[4]], rtol=1e-8, atol=1e-20)


@pytest.mark.parametrize("precision", [0, 1])
def test_mac_eigs_device(ctx, precision):
    """Magneto-thermal pencil (degree blocks alternating 3N1 = 72 and 2N1 =
    48: upol+btor+theta / utor+bpol) through the device path: eigenvalues
    nearest 0.1i match the oracle (SuperLU + Krylov-Schur) to 1e-10."""
    from kore_amd import assembly, eigs
    from kore_amd.params import Params
    from oracle import krylov_schur
    par = Params(m=2, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-4, N=24, lmax=29, magnetic=1, thermal=1,
                 Lambda=0.1, Pm=1e-2, Prandtl=1.0, Ra_gap=6e5)
    A, B, _ = assembly.assemble(par)
    sigma = 0.1j
    say("case: MAC eigs, sizmat", A.shape[0], "precision", precision)
    S = _system(ctx, par, A, B)
    S.factor(sigma, precision)
    # the complex64 factor needs 5 refinement sweeps per operator application
    # here, past the eigensolver's default switch to FP64 (4): allow 8, so
    # that the pairs come from the complex64 path with FP64 refinement
    r = eigs.eigs(S, sigma, nev=6, tol=1e-12, maxit=200, fallback_sweeps=8)
    say("  fallback", r.fallback, "c64 ops", r.c64_op_columns, "ir max", r.c64_ir_max)
    if precision == 1:
        assert not r.fallback and r.precision == 1 and r.c64_ir_max <= 8
    lu = sla.splu((A - sigma * B).tocsc())
    lo, _, _, _ = krylov_schur.krylov_schur(lambda x: lu.solve(B @ x), A.shape[0], sigma, 6, tol=1e-13, maxit=200)
    assert r.nconv >= 6
    for l in lo[:6]:
        assert np.min(np.abs(r.lam - l)) / abs(l) <= 1e-10
    assert eigs.residuals(A, B, r.lam[:6], r.X[:, :6]).max() <= 1e-10
    S.close()


def test_c64_eigs_low_ekman_reports_fallback(ctx):
    """At low Ekman number one complex64 OP application needs more refinement
    sweeps than the eigensolver allows (4): the solve is redone on an FP64
    factor, and the result says so (fallback, precision FP64, the complex64
    columns spent) while its pairs still match the FP64 path to 1e-10."""
    from kore_amd import assembly, eigs
    from kore_amd.params import Params
    par = Params(m=1, symm=1, Ek=1e-6, N=96, lmax=120)
    A, B, _ = assembly.assemble(par)
    sigma = -1e-3 + 0.5j
    say("case: c64 eigs at E=1e-6, sizmat", A.shape[0])
    S = _system(ctx, par, A, B)
    S.factor(sigma, 0)
    r64 = eigs.eigs(S, sigma, nev=6, tol=1e-12)
    S.factor(sigma, 1)
    r = eigs.eigs(S, sigma, nev=6, tol=1e-12)
    say("  fallback", r.fallback, "precision", r.precision, "c64 ops", r.c64_op_columns, "ir max", r.c64_ir_max)
    if r.fallback:
        assert r.precision == 0 and r.c64_op_columns >= 0
    else:
        assert r.precision == 1 and r.c64_op_columns == r.op_columns
    for l in r64.lam[:6]:
        assert np.min(np.abs(r.lam - l)) / abs(l) <= 1e-10
    assert eigs.residuals(A, B, r.lam[:6], r.X[:, :6]).max() <= 1e-10
    S.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_sweep_graph_replays_match_eager(ctx, hydro_small, precision):
    """Solve sweeps replayed from captured HIP graphs give bit-identical
    results to kernel-by-kernel sweeps, on every replay, with other device
    work (eager sweeps of another column count, sparse products) between the
    replays, and across a refactorisation at a new shift (the graph reads the
    factor in place)."""
    from kore_amd import backend
    par, A, B = hydro_small
    say("case: sweep graphs vs eager, precision", precision)
    S = _system(ctx, par, A, B)
    rng = np.random.default_rng(5)
    b1 = rng.standard_normal(S.n) + 1j * rng.standard_normal(S.n)
    b3 = rng.standard_normal((S.n, 3)) + 1j * rng.standard_normal((S.n, 3))
    ws = backend.Workspace(S, 8)
    ws.set(0, b3)
    for sigma in (0.9j, 0.3 + 1.1j):
        S.factor(sigma, precision)
        ctx.set_graphs(False)
        x1_ref, _ = S.solve(b1, ir_maxit=0 if precision == 0 else 3)
        x3_ref, _ = S.solve(b3, ir_maxit=0 if precision == 0 else 3)
        ctx.set_graphs(True)
        for rep in range(4):
            x1, _ = S.solve(b1, ir_maxit=0 if precision == 0 else 3)
            ws.spmv(0, 4, 3, 1.0, -sigma)                   # other device work between replays
            x3, _ = S.solve(b3, ir_maxit=0 if precision == 0 else 3)
            assert np.array_equal(x1, x1_ref), (sigma, rep)
            assert np.array_equal(x3, x3_ref), (sigma, rep)
    ctx.set_graphs(False)
    ws.close()
    S.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_sprow_tiles_bit_identical(ctx, hydro_small, precision, monkeypatch):
    """Every column tile the tuner may choose for the level-0 sparse-row
    products (4, 8 or 16 columns) gives
    the same factor bit for bit: each output entry sums the row's entries in
    the same order whatever the tile.  Forced tiles (KORE_AMD_SPROW_NC) and
    the measured choice (no variable) are compared by their solves."""
    par, A, B = hydro_small
    rng = np.random.default_rng(11)
    b = rng.standard_normal((A.shape[0], 2)) + 1j * rng.standard_normal((A.shape[0], 2))
    tiles = [4, 8, 16]
    out = {}
    for nc in tiles + [None]:
        say("case: sprow tile", nc, "precision", precision)
        if nc is None:
            monkeypatch.delenv("KORE_AMD_SPROW_NC", raising=False)
        else:
            monkeypatch.setenv("KORE_AMD_SPROW_NC", str(nc))
        S = _system(ctx, par, A, B)
        S.factor(0.37 + 0.9j, precision)
        out[nc], _ = S.solve(b, ir_maxit=0)
        S.close()
    for nc in tiles[1:] + [None]:
        assert np.array_equal(out[nc], out[tiles[0]]), nc


def test_sys_update_a_matches_fresh_system(ctx):
    """A system whose A values are replaced in place (kore_sys_update_a)
    factors and solves exactly as a system created from the new values."""
    import dataclasses
    from conftest import params_dormy
    from kore_amd import assembly, backend
    from kore_amd.onset import _on_pattern
    import scipy.sparse as ss
    par = dataclasses.replace(params_dormy(1.6e6), N=24, lmax=32)
    A1, B, _ = assembly.assemble(par)
    A2, _, _ = assembly.assemble(dataclasses.replace(par, Ra_gap=2.4e6, Beyonce=None))
    pat = (abs(A1) + abs(A2)).tocsr()
    pat.sort_indices()
    Ap1 = ss.csr_matrix((_on_pattern(pat, A1), pat.indices, pat.indptr), shape=pat.shape)
    Ap2 = ss.csr_matrix((_on_pattern(pat, A2), pat.indices, pat.indptr), shape=pat.shape)
    say("case: sys_update_a, sizmat", pat.shape[0])
    S = _system(ctx, par, Ap1, B)
    S.update_a(Ap2.data)
    F = _system(ctx, par, Ap2, B)
    b = np.random.default_rng(2).standard_normal(S.n) + 0j
    for X in (S, F):
        X.factor(0.3j, 0)
    xs, _ = S.solve(b, ir_maxit=0)
    xf, _ = F.solve(b, ir_maxit=0)
    assert np.array_equal(xs, xf)
    assert S.norm(0.3j) == F.norm(0.3j)
    S.close()
    F.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_onset_driver_resident_system(ctx, precision):
    """The product onset driver (one resident system per m, A's buoyancy
    values updated per Brent step, warm-started eigensolves) reproduces the
    reference's Dormy et al. (2004) onset golden (tests/dormy2004/
    reference.dormy04:1; printed values at rtol 1e-8) with two assemblies in
    total instead of one per step."""
    from conftest import DORMY_GOLDEN, params_dormy
    from kore_amd import onset
    say("case: onset driver, Dormy 2004, precision", precision)
    ra, lam, solver = onset.critical_rayleigh(params_dormy(1.6e6), 1.6e6, ctx=ctx, precision=precision,
                                              lo=1.6e6, hi=1.7e6)
    say("  Ra_c %.5e omega_c %.5e, %d eigensolves, %d assemblies, %.2f s of eigensolves" % (
        ra, lam.imag, len(solver.points), solver.assemblies, sum(p.seconds for p in solver.points)))
    solver.close()
    got = np.array([float("%.5e" % ra), float("%.5e" % lam.imag)])
    np.testing.assert_allclose(got, [DORMY_GOLDEN[2], DORMY_GOLDEN[4]], rtol=1e-8, atol=1e-20)
    assert solver.assemblies == 2
