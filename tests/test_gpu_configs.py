"""Full-size parity of the device path against the oracle (MI355X only).

The three full-size workloads -- m=1, E=1e-6, N=200, lmax=300 (60 000
unknowns); m=1, E=1e-7, N=400, lmax=500 (200 000); MAC m=2, E=1e-6, N=150,
lmax=201 (75 000) -- are assembled here by the in-repo assembler and solved
on the device.  The oracle's answers on the same matrices (SuperLU +
oracle Krylov-Schur, written by tests/golden/make_config_goldens.py) are
stored as fixtures:

* residuals ||(A - lambda B) x|| / (||A||_F ||x||) <= 1e-10 per pair;
* eigenvalues: relative difference <= 1e-10 (the bar of bin/solve.py's
  FP64 path) wherever FP64 determines the eigenvalue that well.  These
  pencils are far from normal at low Ekman number: the relative condition
  number kappa_rel = ||x|| ||y|| ||A||_F / (|y^H B x| |lambda|) (x, y right
  and left eigenvectors; stored with the fixture) is 1e13 - 1e18 for the
  wanted eigenvalues of the three workloads, so the normwise backward error
  both solvers reach (1e-17) bounds an eigenvalue only to kappa_rel * 1e-17,
  1e-4 and worse.  Two backward-stable FP64 answers (the oracle's and the
  device's) agree far better than that worst case: measured over the 28
  eigenvalues, their relative difference is about 1e-24 kappa_rel (9e-10 at
  kappa_rel 3.7e15, 2.4e-8 at 9.1e15, 1.75e-7 near 1e17).  The bar is ten
  times that, max(1e-10, min(1e-6, 1e-23 kappa_rel)).  Agreement to 1e-10 for
  these would need residuals in extended precision on both sides; neither
  the reference (SLEPc + MUMPS in FP64) nor this path uses them;
* eigenvectors, phase-invariantly: the ratios (w_k^H x) / (w_0^H x) for
  seeded probe vectors w_k agree with the oracle's to max(1e-7, 1e-21
  kappa_rel) (measured about 1e-22 kappa_rel, up to 3e-5; a wrong vector
  misses by orders of magnitude);
* the 200 000-unknown forced solve x = (A0 - 0.5i B0)^{-1} b (seeded
  unit-norm complex b): the backward error is <= 1e-15 on both paths, and
  ||x|| and the probe projections w_k^H x agree with SuperLU's to 1e-8, or
  to ten times cond_1(A0 - 0.5i B0) times the two backward errors where that
  is larger (the complex64 + GMRES path stops at a backward error near
  1e-15, the FP64 factor reaches ~1e-19).
"""
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(name):
    with open(os.path.join(GOLDEN, "config_%s.json" % name)) as f:
        return json.load(f)


def cz(v):
    a = np.asarray(v, dtype=float).reshape(-1, 2)
    return a[:, 0] + 1j * a[:, 1]


def probes(n, seed, k):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))) / np.sqrt(2.0)


@pytest.fixture(scope="module")
def ctx():
    from kore_amd import backend
    c = backend.Context(0)
    yield c
    c.close()


def _eigen_case(ctx, name, precision):
    from kore_amd import assembly, backend, eigs
    from kore_amd.params import Params
    g = load(name)
    par = Params(**g["par"])
    A, B, _ = assembly.assemble(par)
    assert A.shape[0] == g["sizmat"]
    sigma = complex(*g["sigma"])
    nev = g["nev"]
    say("case: %s eigen precision %d sizmat %d" % (name, precision, A.shape[0]))
    S = backend.System(ctx, backend.layout_from_params(par), A, B)
    S.factor(sigma, precision)
    r = eigs.eigs(S, sigma, nev=nev, tol=g["tol"])
    say("  nconv %d ops %d fallback %s precision %d c64 ops %d" % (r.nconv, r.op_columns, r.fallback, r.precision,
                                                                   r.c64_op_columns))
    assert r.nconv >= nev
    lam_o = cz(g["lam"])
    W = probes(A.shape[0], g["probe_seed"], g["nprobe"])
    res_all = eigs.residuals(A, B, r.lam, r.X)
    say("  max residual %.2e" % res_all[:nev].max())
    assert res_all[:nev].max() <= 1e-10
    for i, lo in enumerate(lam_o):
        j = int(np.argmin(np.abs(r.lam - lo)))
        d = abs(r.lam[j] - lo) / abs(lo)
        x = r.X[:, j]
        P = W.conj().T @ x
        got = P[1:] / P[0]
        want = cz(g["vec_ratios"][i])
        dv = np.max(np.abs(got - want) / np.abs(want))
        # bar scaled by the eigenvalue's relative condition number (module docstring)
        kappa_rel = g["kappa_abs"][i] * g["normA_fro"] / abs(lo)
        tol_l = max(1e-10, min(1e-6, 1e-23 * kappa_rel))
        tol_v = max(1e-7, 1e-21 * kappa_rel)
        say("  lambda %s rel diff %.2e (tol %.1e, kappa_rel %.1e), vector ratios rel diff %.2e (tol %.1e)" % (
            lo, d, tol_l, kappa_rel, dv, tol_v))
        assert d <= tol_l
        assert dv <= tol_v
    S.close()
    return r


@pytest.mark.parametrize("precision", [0, 1])
def test_hydro_E1e6_N200_eigen(ctx, precision):
    _eigen_case(ctx, "hydro_E1e-6_N200", precision)


@pytest.mark.parametrize("precision", [0, 1])
def test_mac_E1e6_N150_eigen(ctx, precision):
    _eigen_case(ctx, "mac_E1e-6_N150", precision)


@pytest.mark.parametrize("precision", [0, 1])
def test_hydro_E1e7_N400_eigen(ctx, precision):
    _eigen_case(ctx, "hydro_E1e-7_N400", precision)


@pytest.mark.parametrize("path", ["fp64", "c64_gmres"])
def test_hydro_E1e7_N400_forced(ctx, path):
    from kore_amd import assembly, backend
    from kore_amd.params import Params
    g = load("hydro_E1e-7_N400")
    f = g["forced"]
    par = Params(**g["par"])
    A0, B0 = assembly.assemble(par, which="unscaled")
    n = A0.shape[0]
    rng = np.random.default_rng(f["rhs_seed"])
    b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2.0)
    b /= np.linalg.norm(b)
    sigma = 1j * f["omega"]
    say("case: forced %s sizmat %d omega %g" % (path, n, f["omega"]))
    S = backend.System(ctx, backend.layout_from_params(par), A0, B0)
    if path == "fp64":
        S.factor(sigma, backend.PREC_FP64)
        x, info = S.solve(b, ir_maxit=0)
    else:
        S.factor(sigma, backend.PREC_C64)
        x, info = S.solve_gmres(b, tol=1e-15, maxit=60)
        say("  gmres iterations %d converged %d backward error %.2e" % (info.iterations, info.converged,
                                                                       info.rel_correction))
        assert info.converged
    nM = S.norm(sigma)
    berr = S.residual_norms(x, b=b, shift=sigma)[0] / (nM * np.linalg.norm(x) + 1.0)
    W = probes(n, g["probe_seed"], g["nprobe"])
    P = W.conj().T @ x
    want = cz(f["proj"])
    dp = np.max(np.abs(P - want) / np.abs(want))
    dn = abs(np.linalg.norm(x) - f["xnorm"]) / f["xnorm"]
    tol = max(1e-8, 10 * f["cond1"] * (berr + f["backward_error"]))
    say("  backward error %.2e, |x| rel diff %.2e, projections rel diff %.2e (tol %.1e, cond_1 %.1e)" % (
        berr, dn, dp, tol, f["cond1"]))
    assert berr <= 1e-15
    assert dn <= tol and dp <= tol
    S.close()
