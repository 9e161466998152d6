#!/usr/bin/env python3
"""Write the oracle's answers for the full-size workloads as small fixtures.

The matrices of these workloads (60 000 to 200 000 unknowns) come from the
in-repo assembler (kore_amd.assembly, itself pinned operator by operator
against the reference's own modules by tests/golden/ops_*.npz), so only the
oracle's results need to be stored, not the matrices:

* eigenvalues: SuperLU (scipy.sparse.linalg.splu) of A - sigma B and the
  oracle's shift-and-invert Krylov-Schur (oracle/krylov_schur.py), the nev
  pairs nearest sigma, as the reference's bin/solve.py:91-158 computes them
  with SLEPc + MUMPS;
* eigenvectors, phase-invariantly: for each unit eigenvector x the ratios
  (w_k^H x) / (w_0^H x), k = 1..3, with seeded Gaussian probe vectors w_k
  (complex, Kore ordering; numpy default_rng(PROBE_SEED));
* the forced solve of the 200 000-unknown workload: x = (A0 - i w B0)^{-1} b
  (bin/solve.py:207-233 with B_forced = b) for a seeded unit-norm complex
  Gaussian b (default_rng(RHS_SEED)), stored as ||x|| and w_k^H x, k = 0..3.

Every quantity is a few complex numbers, so the fixture stays tiny while the
device result is compared entry-free but complete (a random projection of a
wrong vector disagrees with probability one).

Conditioning (``--conditioning``, added to an existing fixture): for each
eigenvalue its absolute condition number kappa = ||x|| ||y|| / |y^H B x|
(x, y the right and left eigenvectors, each one step of inverse iteration
with SuperLU's factor of A - lambda B at the computed lambda), so that a
test can bound the difference of two backward-stable answers by
kappa ||A||_F eps / |lambda| where that exceeds 1e-10; and for the forced
solve the 1-norm condition number of A0 - i w B0 (scipy's onenormest on
SuperLU solves), which bounds the forward error of a solution with a given
backward error.

Run once in the build container (a few minutes, one BLAS thread):
    python tests/golden/make_config_goldens.py [name ...]
    python tests/golden/make_config_goldens.py --conditioning [name ...]
writes (or augments) tests/golden/config_<name>.json.
"""
import json
import os
import sys
import time

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

PROBE_SEED = 20261203
RHS_SEED = 4242
NPROBE = 4

CASES = {
    # m=1, E=1e-6, N=200, lmax=300: 60 000 unknowns, 300 degree blocks of 200
    "hydro_E1e-6_N200": dict(par=dict(m=1, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-6, N=200, lmax=300),
                             sigma=-1e-3 + 0.5j, nev=10, tol=1e-12, forced=None),
    # m=1, E=1e-7, N=400, lmax=500: 200 000 unknowns; plus a forced solve at w = 0.5
    "hydro_E1e-7_N400": dict(par=dict(m=1, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-7, N=400, lmax=500),
                             sigma=-1e-3 + 0.5j, nev=10, tol=1e-12, forced=0.5),
    # MAC: m=2, E=1e-6, Le=1e-4, Pm=1e-3, Pr=1, N=150, lmax=201: 75 000 unknowns
    "mac_E1e-6_N150": dict(par=dict(m=2, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-6, N=150, lmax=201, magnetic=1,
                                    thermal=1, Le2=1e-8, Pm=1e-3, Prandtl=1.0, Ra_gap=6e5),
                           sigma=0.1j, nev=8, tol=1e-12, forced=None),
}


def probes(n):
    rng = np.random.default_rng(PROBE_SEED)
    return (rng.standard_normal((n, NPROBE)) + 1j * rng.standard_normal((n, NPROBE))) / np.sqrt(2.0)


def rhs(n):
    rng = np.random.default_rng(RHS_SEED)
    b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2.0)
    return b / np.linalg.norm(b)


def vec_ratios(W, X):
    """Phase-invariant signature of unit columns X: (w_k^H x) / (w_0^H x)."""
    P = W.conj().T @ X
    return P[1:] / P[0]


def cpx(a):
    return [[float(np.real(v)), float(np.imag(v))] for v in np.ravel(a)]


def make(name):
    from kore_amd import assembly
    from kore_amd.params import Params
    from oracle import krylov_schur
    case = CASES[name]
    par = Params(**case["par"])
    t0 = time.perf_counter()
    A, B, _ = assembly.assemble(par)
    n = A.shape[0]
    sigma, nev, tol = case["sigma"], case["nev"], case["tol"]
    lu = spla.splu((A - sigma * B).tocsc())
    lam, X, res, it = krylov_schur.krylov_schur(lambda x: lu.solve(B @ x), n, sigma, nev, tol=tol, maxit=200)
    lam, X = lam[:nev], X[:, :nev]
    nA = np.sqrt(np.sum(np.abs(A.data) ** 2))
    resid = [float(np.linalg.norm(A @ X[:, i] - lam[i] * (B @ X[:, i])) / (nA * np.linalg.norm(X[:, i])))
             for i in range(nev)]
    W = probes(n)
    out = dict(case=name, par=case["par"], sigma=cpx(sigma)[0], nev=nev, tol=tol, sizmat=int(n),
               lam=cpx(lam), residual=resid, vec_ratios=[cpx(vec_ratios(W, X[:, i])) for i in range(nev)],
               probe_seed=PROBE_SEED, nprobe=NPROBE, restarts=int(it))
    del lu
    if case["forced"] is not None:
        w = case["forced"]
        A0, B0 = assembly.assemble(par, which="unscaled")
        M = (A0 - 1j * w * B0).tocsc()
        b = rhs(n)
        x = spla.splu(M).solve(b)
        nM = np.sqrt(np.sum(np.abs(M.data) ** 2))
        berr = float(np.linalg.norm(M @ x - b) / (nM * np.linalg.norm(x) + 1.0))
        out["forced"] = dict(omega=w, rhs_seed=RHS_SEED, xnorm=float(np.linalg.norm(x)),
                             proj=cpx(W.conj().T @ x), backward_error=berr)
    out["seconds"] = time.perf_counter() - t0
    path = os.path.join(HERE, "config_%s.json" % name)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("%s: n %d, %d pairs, max residual %.2e, %.1f s -> %s" % (name, n, nev, max(resid), out["seconds"], path))


def conditioning(name):
    from kore_amd import assembly
    from kore_amd.params import Params
    path = os.path.join(HERE, "config_%s.json" % name)
    with open(path) as f:
        out = json.load(f)
    case = CASES[name]
    par = Params(**case["par"])
    t0 = time.perf_counter()
    A, B, _ = assembly.assemble(par)
    n = A.shape[0]
    nA = float(np.sqrt(np.sum(np.abs(A.data) ** 2)))
    lam = np.array([complex(*v) for v in out["lam"]])
    rng = np.random.default_rng(99)
    kappa = []
    for l in lam:
        lu = spla.splu((A - l * B).tocsc())
        r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        x = lu.solve(r)
        y = lu.solve(r, trans="H")
        x /= np.linalg.norm(x)
        y /= np.linalg.norm(y)
        kappa.append(float(1.0 / abs(np.vdot(y, B @ x))))
        del lu
    out["kappa_abs"] = kappa
    out["normA_fro"] = nA
    if case["forced"] is not None:
        w = case["forced"]
        A0, B0 = assembly.assemble(par, which="unscaled")
        M = (A0 - 1j * w * B0).tocsc()
        lu = spla.splu(M)
        inv = spla.LinearOperator(M.shape, matvec=lu.solve, rmatvec=lambda v: lu.solve(v, trans="H"),
                                  dtype=complex)
        out["forced"]["cond1"] = float(spla.onenormest(inv) * spla.norm(M, 1))
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("%s: kappa_abs in [%.2e, %.2e], %.1f s" % (name, min(kappa), max(kappa), time.perf_counter() - t0))


if __name__ == "__main__":
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(limits=1)
    except ImportError:
        pass
    args = sys.argv[1:]
    cond = "--conditioning" in args
    args = [a for a in args if a != "--conditioning"]
    for nm in (args or sorted(CASES)):
        (conditioning if cond else make)(nm)
