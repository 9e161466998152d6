"""Critical Rayleigh number of thermal convection (onset search) on the device.

Restates the reference's onset driver (tests/jones2000/find_Rac.py:36-113,
bin/find_Rac_pbs.py): for each azimuthal wavenumber m, the growth rate
sigma(Ra) = max Re(lambda) of the eigenpairs nearest the target (which
'TR') is bracketed from Ra_min in steps of 0.01 in log10 Ra
(bracket_brentq, find_Rac.py:55-78) and its root found by Brent's method on
log10 Ra with xtol = rtol = 1e-6 (find_Rac.py:77); the critical
(Ek, ricb, Ra_c, m, omega_c) is appended to ``critical_params.dat`` with the
reference's format (find_Rac.py:104-108).

Where the reference re-runs assemble.py and solve.py (a new PETSc problem)
for every Ra, here one system stays resident on the device per m:

* A depends on Ra only through the buoyancy force, linearly (Beyonce =
  -Ra Ek^2 / Pr, bin/parameters.py:186, 274): A(Ra) = A(Ra_1) + (Ra - Ra_1) dA
  with dA = (A(Ra_2) - A(Ra_1)) / (Ra_2 - Ra_1) from two assemblies on one
  pattern; B does not depend on Ra.  Each Brent step uploads the new values
  of A into the resident system (kore_sys_update_a) and refactors, reusing
  the system's pattern, layout, factor plan and device memory;
* the Krylov start vector of each eigensolve is the previous step's leading
  eigenvector (warm start): the leading mode changes little between
  neighbouring Ra.
"""
from __future__ import annotations

import dataclasses
import os
import sys
import time

import numpy as np
import scipy.sparse as ss
from scipy.optimize import brentq

from . import assembly, backend, eigs
from .params import Params


def _on_pattern(P, M):
    """Values of the CSR matrix M at the entries of the CSR pattern P (zeros where M has none)."""
    P = ss.csr_matrix(P)
    M = ss.csr_matrix(M)
    rows = np.repeat(np.arange(P.shape[0]), np.diff(P.indptr))
    return np.asarray(M[rows, P.indices]).ravel().astype(np.complex128)


@dataclasses.dataclass
class OnsetPoint:
    Ra: float
    growth: float
    lam: complex
    seconds: float


class OnsetSolver:
    """Growth rate sigma(Ra) of one (m, Ek, ...) thermal problem with the
    system resident on the device; ``par`` gives every knob but Ra_gap."""

    def __init__(self, par: Params, ctx=None, precision=backend.PREC_FP64, ra_ref=1e6, log=None):
        if not par.thermal:
            raise ValueError("the onset search needs a thermal problem (thermal = 1)")
        self.par = par
        self.precision = precision
        self.log = log
        self.ctx = ctx if ctx is not None else backend.Context(0)
        self._own_ctx = ctx is None
        ra1, ra2 = float(ra_ref), 2.0 * float(ra_ref)
        A1, B, _ = assembly.assemble(dataclasses.replace(par, Ra_gap=ra1, Beyonce=None))
        A2, _, _ = assembly.assemble(dataclasses.replace(par, Ra_gap=ra2, Beyonce=None))
        # one pattern for every Ra: the union of the two assemblies
        pat = (abs(A1) + abs(A2)).tocsr()
        pat.sort_indices()
        a1 = _on_pattern(pat, A1)
        a2 = _on_pattern(pat, A2)
        self.ra1 = ra1
        self.da = (a2 - a1) / (ra2 - ra1)
        self.a1 = a1
        A = ss.csr_matrix((a1, pat.indices, pat.indptr), shape=pat.shape)
        self.S = backend.System(self.ctx, backend.layout_from_params(par), A, B)
        self.v0 = None
        self.points = []
        self.assemblies = 2

    def a_values(self, ra: float):
        return self.a1 + (ra - self.ra1) * self.da

    def growth(self, ra: float):
        """(max Re lambda, lambda) of the nev eigenpairs nearest the target at Ra_gap = ra."""
        t = time.perf_counter()
        self.S.update_a(self.a_values(ra))
        tau = complex(self.par.tau)
        self.S.factor(tau, self.precision)
        r = eigs.eigs(self.S, tau, nev=self.par.nev, tol=self.par.tol, which=self.par.which_eigenpairs,
                      maxit=max(200, self.par.maxit), v0=self.v0)
        lam = r.lam[:max(1, min(self.par.nev, r.nconv))]
        i = int(np.argmax(lam.real))
        self.v0 = r.X[:, i].copy()              # warm start of the next step
        p = OnsetPoint(ra, float(lam[i].real), complex(lam[i]), time.perf_counter() - t)
        self.points.append(p)
        if self.log:
            self.log("Ra = %e: growth %.6e, lambda %s, %.3f s" % (ra, p.growth, p.lam, p.seconds))
        return p.growth, p.lam

    def close(self):
        self.S.close()
        if self._own_ctx:
            self.ctx.close()


def bracket_brentq(f, x1, dx=0.01, tol=1e-6, maxiter=200):
    """Root of f on log10 Ra: march from x1 in steps of dx (down if f(x1) > 0)
    until the sign changes, then Brent's method with xtol = rtol = tol
    (find_Rac.py:55-78)."""
    y1 = f(x1)
    dx = abs(dx)
    if y1 > 0:
        dx = -dx
    x2 = x1
    while True:
        x2 += dx
        y2 = f(x2)
        if y2 * y1 < 0:
            break
        x1, y1 = x2, y2
    lo, hi = min(x1, x2), max(x1, x2)
    return brentq(f, lo, hi, maxiter=maxiter, xtol=tol, rtol=tol)


def critical_rayleigh(par: Params, ra_min: float, ctx=None, precision=backend.PREC_FP64, log=None, lo=None,
                      hi=None):
    """(Ra_c, lambda_c, solver) for the problem ``par`` (Ra_gap ignored).
    With lo and hi the bracket is given (Brent directly on [log10 lo,
    log10 hi]); otherwise it is searched from ra_min as the reference does."""
    solver = OnsetSolver(par, ctx=ctx, precision=precision, ra_ref=ra_min, log=log)
    cache = {}

    def f(x):
        ra = 10 ** x
        if ra not in cache:
            cache[ra] = solver.growth(ra)
        return cache[ra][0]
    if lo is not None and hi is not None:
        x0 = brentq(f, np.log10(lo), np.log10(hi), xtol=1e-6, rtol=1e-6, maxiter=200)
    else:
        x0 = bracket_brentq(f, np.log10(ra_min))
    ra = 10 ** x0
    lam = solver.growth(ra)[1]
    return ra, lam, solver


def write_critical(path, par: Params, ra_c: float, m: int, omega_c: float):
    """Appends one line in the reference's format ('%.3e %.2f %.5e %d %.5e',
    find_Rac.py:104-108)."""
    X = np.array([par.Ek, par.ricb, ra_c, int(m), omega_c])
    with open(path, "a") as f:
        np.savetxt(f, X.reshape(1, -1), fmt=["%.3e", "%.2f", "%.5e", "%d", "%.5e"])


def main(argv=None):
    """python -m kore_amd.onset [parameters.py] [Ra_min] [m_min] [m_max]
    (the knobs of find_Rac.py:21-25 as arguments)."""
    from .params import load_parameters
    argv = sys.argv[1:] if argv is None else argv
    ppath = argv[0] if argv else "parameters.py"
    base = load_parameters(ppath) if os.path.exists(ppath) else Params()
    ra_min = float(argv[1]) if len(argv) > 1 else 4.6e6
    m0 = int(argv[2]) if len(argv) > 2 else base.m
    m1 = int(argv[3]) if len(argv) > 3 else m0
    ctx = backend.Context(0)
    for m in range(m0, m1 + 1):
        par = dataclasses.replace(base, m=m, Beyonce=None)
        t = time.perf_counter()
        ra, lam, solver = critical_rayleigh(par, ra_min, ctx=ctx, log=lambda s: print(s, flush=True))
        print("Rac for m=%d: %.5e (omega %.5e) in %.1f s, %d eigensolves, %d assemblies" % (
            m, ra, lam.imag, time.perf_counter() - t, len(solver.points), solver.assemblies), flush=True)
        write_critical("critical_params.dat", par, ra, m, lam.imag)
        solver.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
