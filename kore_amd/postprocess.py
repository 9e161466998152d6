"""Power-balance diagnostics of Kore solutions (the step after the solve).

Restates what bin/spin_doctor.py:22-242 computes from the solve stage's
output files with bin/utils4pp.py's per-degree integrals: the kinetic energy
KE (poloidal KP, toroidal KT), the viscous dissipation Dkin = E Dkin0
(integral of u . lap u), the internal dissipation Dint = E Dint0, the rate
of working of thermal buoyancy Wthm, the thermal "energy" TE, diffusion Dthm
and advection Wadv_thm, and the relative residuals of the power balance of a
solution with eigenvalue lambda = sigma + i w:

    resid0 = |Dint0 + Dkin0 - pss| / max(|Dint0|, |Dkin0|, |pss|)
    resid1 = |2 sigma KE - Dkin - Wlor + Wthm| / max(|2 sigma KE|, |Dkin|, |Wlor|, |Wthm|)
    resid3 = |2 sigma TE - Dthm - Wadv_thm| / max(|2 sigma TE|, |Dthm|, |Wadv_thm|)

(bin/spin_doctor.py:227-242).  They vanish for an exact eigenvector: a
solver-independent physics check of every returned solution.

Units: the reference's rotation time scale (Gaspard = 1, bin/parameters.py:
273-278), so its OmgTau factors (spin_doctor.py:158-162) are 1 and
Dkin = Ek Dkin0, Dint = Ek Dint0, Wthm = BV2 Wthm0, Dthm = Etherm Dthm0.

Design: the reference evaluates each degree l in its own process
(utils4pp.diagnose with a multiprocessing pool); here all degrees are one
set of matrix products: Chebyshev coefficients (rows: degrees) times the
Chebyshev values at the N Gauss-Chebyshev quadrature radii, derivative
coefficients by the three-term recursion of bin/utils.py:273-294 applied to
all rows at once.  Forced solutions: the reference's balance needs the power
of the forcing (``repow``), which bin/spin_doctor.py:216-225 references but
never defines; for them only the integrals are reported (resid0/resid1 with
pss = pvf = 0 are not meaningful and are returned as NaN).

Magnetic diagnostics (Lorentz power, ohmic dissipation, induction;
utils4pp.py:299-360, 565-776) are not restated.
"""
from __future__ import annotations

import dataclasses
import os
import sys

import numpy as np

from .params import Params, ell


def _dcheb_rows(C, ricb, rcmb=1.0):
    """Chebyshev coefficients of d/dr of each row of C (bin/utils.py:273-294):
    out[s-2] = 2 (s-1) c[s-1], out[k] = out[k+2] + 2 (k+1) c[k+1], out[0] /= 2,
    times the domain factor 1/rcmb (full sphere) or 2/(rcmb - ricb)."""
    s = C.shape[1]
    out = np.zeros_like(C)
    if s >= 2:
        out[:, s - 2] = 2.0 * (s - 1) * C[:, s - 1]
    for k in range(s - 3, -1, -1):
        out[:, k] = out[:, k + 2] + 2.0 * (k + 1) * C[:, k + 1]
    out[:, 0] = out[:, 0] / 2.0
    return out / rcmb if ricb == 0 else 2.0 * out / (rcmb - ricb)


def _cheb_values(x, N):
    """V[i, k] = T_k(x_i)."""
    V = np.zeros((len(x), N))
    V[:, 0] = 1.0
    if N > 1:
        V[:, 1] = x
    for k in range(2, N):
        V[:, k] = 2 * x * V[:, k - 1] - V[:, k - 2]
    return V


def expand_rows(sec, N1, N, ricb, m, vsymm, parity_offset):
    """Coefficient rows (degrees x N) of one section of the solution vector.
    With ricb = 0 the N1 stored coefficients per degree are the ones of one
    parity, spread to N with zeros (bin/utils4pp.py:67-160): poloidal (and
    scalar) rows start at (m + 1 - s) % 2, toroidal rows at (m + s) % 2, s =
    (vsymm + 1) / 2."""
    rows = np.asarray(sec).reshape(-1, N1)
    if ricb != 0:
        return rows
    s = int((vsymm + 1) // 2)
    first = (m + 1 - s) % 2 if parity_offset == "pol" else (m + s) % 2
    out = np.zeros((rows.shape[0], N), dtype=complex)
    out[:, first::2] = rows
    return out


@dataclasses.dataclass
class Diagnostics:
    KE: float
    KP: float
    KT: float
    Dkin: float
    Dint: float
    Wthm: float
    TE: float
    Dthm: float
    Wadv_thm: float
    resid0: float
    resid1: float
    resid3: float
    Dkin0: float = 0.0
    Dint0: float = 0.0
    per_l: dict = dataclasses.field(default_factory=dict)


class Quadrature:
    """Gauss-Chebyshev radii of the reference's diagnose() (utils4pp.py:794-830)."""

    def __init__(self, par: Params):
        N = par.N
        self.N = N
        i = np.arange(N)
        xk = np.cos((i + 0.5) * np.pi / N)
        self.Ra, self.Rb = par.ricb, 1.0
        self.rk = 0.5 * (self.Rb - self.Ra) * (xk + 1) + self.Ra
        r0 = par.ricb + (np.sign(par.ricb) - 1) * 1.0          # xcheb, utils4pp.py:15-25
        self.x0 = 2 * (self.rk - r0) / (1.0 - r0) - 1
        self.sqx = np.sqrt(1 - xk ** 2)
        self.V = _cheb_values(self.x0, N)

    def integrate(self, F):
        """Chebyshev-Gauss quadrature of each row of F (utils4pp.py:54-64)."""
        return (np.pi / self.N) * (F * self.sqx).sum(axis=-1) * (self.Rb - self.Ra) / 2

    def values(self, C, nder, ricb):
        """[f, f', f'', ...] at the quadrature radii for coefficient rows C."""
        out = [C @ self.V.T]
        D = C
        for _ in range(nder):
            D = _dcheb_rows(D, ricb)
            out.append(D @ self.V.T)
        return out


def flow_integrals(par: Params, usol, q: Quadrature = None):
    """Per-degree integrals of the flow section (2n coefficients: poloidal then
    toroidal, Kore order), utils4pp.py:221-298 and 426-488: returns
    dict(l=..., KE=..., Dkin0=..., Dint0=..., pol=bool array)."""
    q = q or Quadrature(par)
    lp, lt, _ = ell(par.m, par.lmax, par.symm)
    n = par.n
    P = expand_rows(usol[:n], par.N1, par.N, par.ricb, par.m, par.symm, "pol")
    T = expand_rows(usol[n:2 * n], par.N1, par.N, par.ricb, par.m, par.symm, "tor")
    rk, r2 = q.rk, q.rk ** 2
    # poloidal: plm and derivatives up to 3; qlm, slm up to 2 (cheb2space_pol)
    p0, p1, p2, p3 = q.values(P, 3, par.ricb)
    L1 = (lp * (lp + 1.0))[:, None]
    q0 = L1 * p0 / rk
    s0 = p1 + p0 / rk
    q1 = (L1 * p1 - q0) / rk
    s1 = p2 + q1 / L1
    q2 = (L1 * p2 - 2 * q1) / rk
    s2 = p3 + q2 / L1
    lpc = lp[:, None].astype(float)
    f0 = 4 * np.pi / (2 * lpc + 1)
    kin_p = f0 * (r2 * np.abs(q0) ** 2 + r2 * L1 * np.abs(s0) ** 2)
    dif_p = 2 * np.real(f0 * (L1 * r2 * np.conj(s0) * s2 + 2 * rk * L1 * np.conj(s0) * s1
                              - L1 ** 2 * np.conj(s0) * s0 - (lpc ** 2 + lpc + 2) * np.conj(q0) * q0
                              + 2 * rk * np.conj(q0) * q1 + r2 * np.conj(q0) * q2
                              + 2 * L1 * (np.conj(q0) * s0 + q0 * np.conj(s0))))
    int_p = 2 * f0 * (L1 * np.abs(q0 + rk * s1 - s0) ** 2 + 3 * np.abs(rk * q1) ** 2
                      + L1 * (lpc - 1) * (lpc + 2) * np.abs(s0) ** 2)
    # toroidal: tlm and derivatives up to 2 (cheb2space_tor)
    t0, t1, t2 = q.values(T, 2, par.ricb)
    L2 = (lt * (lt + 1.0))[:, None]
    ltc = lt[:, None].astype(float)
    g0 = 4 * np.pi / (2 * ltc + 1)
    kin_t = g0 * r2 * L2 * np.abs(t0) ** 2
    dif_t = 2 * np.real(g0 * (L2 * r2 * np.conj(t0) * t2 + 2 * rk * L2 * np.conj(t0) * t1 - L2 ** 2 * np.conj(t0) * t0))
    int_t = 2 * g0 * (L2 * np.abs(rk * t1 - t0) ** 2 + L2 * (ltc - 1) * (ltc + 2) * np.abs(t0) ** 2)
    l_all = np.concatenate([lp, lt])
    return dict(l=l_all, pol=np.concatenate([np.ones(len(lp), bool), np.zeros(len(lt), bool)]),
                KE=np.concatenate([q.integrate(kin_p), q.integrate(kin_t)]),
                Dkin0=np.concatenate([q.integrate(dif_p), q.integrate(dif_t)]),
                Dint0=np.concatenate([q.integrate(int_p), q.integrate(int_t)]),
                q0=q0)


def thermal_integrals(par: Params, tsol, usol=None, q: Quadrature = None):
    """Per-degree thermal integrals (utils4pp.py:363-424, 536-562, 777-792):
    TE, Dthm0 (integral of theta lap theta), Wadv (advection power) on the
    poloidal degree list, and the buoyancy power Wthm0 of the flow."""
    q = q or Quadrature(par)
    lp, _, _ = ell(par.m, par.lmax, par.symm)
    # a scalar section takes the poloidal parity (bin/utils4pp.py:140-150)
    H = expand_rows(tsol, par.N1, par.N, par.ricb, par.m, par.symm, "pol")
    h0, h1, h2 = q.values(H, 2, par.ricb)
    rk, r2 = q.rk, q.rk ** 2
    lpc = lp[:, None].astype(float)
    L1 = lpc * (lpc + 1)
    f0 = 4 * np.pi / (2 * lpc + 1)
    te = f0 * r2 * np.abs(h0) ** 2
    td = f0 * (2 * rk * 2 * np.real(h0 * np.conj(h1)) + r2 * 2 * np.real(h0 * np.conj(h2)) - 2 * L1 * np.abs(h0) ** 2)
    out = dict(l=lp, TE=q.integrate(te), Dthm0=q.integrate(td), Wadv=np.zeros(len(lp)), Wthm0=np.zeros(len(lp)))
    if usol is not None and par.hydro:
        n = par.n
        P = expand_rows(usol[:n], par.N1, par.N, par.ricb, par.m, par.symm, "pol")
        p0 = q.values(P, 0, par.ricb)[0]
        q0 = L1 * p0 / rk
        plm0 = q0 * rk / L1
        if par.heating == "internal":
            fr = r2
        elif par.heating == "differential":
            fr = 1 / rk
        else:
            raise ValueError("heating %r is not supported by the diagnostics" % par.heating)
        out["Wadv"] = q.integrate(f0 * fr * L1 * 2 * np.real(np.conj(plm0) * h0))
        out["Wthm0"] = q.integrate(f0 * r2 * L1 * 2 * np.real(np.conj(plm0) * h0))
    return out


def _rel(terms):
    terms = [abs(t) for t in terms]
    den = max(terms[1:] + [abs(terms[0])])
    return terms, den


def diagnose(par: Params, lam: complex = None, usol=None, tsol=None) -> Diagnostics:
    """Diagnostics of one solution (Kore order): usol the flow section (2n),
    tsol the temperature section (n) if thermal; lam its eigenvalue (None:
    a forced solution, whose balance needs the forcing power)."""
    q = Quadrature(par)
    KE = KP = KT = Dkin0 = Dint0 = Wthm0 = TE = Dthm0 = Wadv = 0.0
    per_l = {}
    if par.hydro and usol is not None:
        f = flow_integrals(par, np.asarray(usol), q)
        KE = float(f["KE"].sum())
        KP = float(f["KE"][f["pol"]].sum())
        KT = float(f["KE"][~f["pol"]].sum())
        Dkin0 = float(f["Dkin0"].sum())
        Dint0 = float(f["Dint0"].sum())
        per_l["flow"] = f
    if par.thermal and tsol is not None:
        t = thermal_integrals(par, np.asarray(tsol), usol, q)
        TE = float(t["TE"].sum())
        Dthm0 = float(t["Dthm0"].sum())
        Wadv = float(t["Wadv"].sum())
        Wthm0 = float(t["Wthm0"].sum())
        per_l["thermal"] = t
    Dkin = par.Ek * Dkin0
    Dint = par.Ek * Dint0
    Wthm = par.Beyonce * Wthm0          # BV2 (bin/parameters.py:274)
    Dthm = par.ThermaD * Dthm0          # Etherm (bin/parameters.py:277)
    resid0 = resid1 = resid3 = float("nan")
    if lam is not None:
        sigma = float(np.real(lam))
        pss = 0.0
        if par.Ek != 0 and par.hydro:
            resid0 = abs(Dint0 + Dkin0 - pss) / max(abs(Dint0), abs(Dkin0), abs(pss))
        if par.hydro:
            resid1 = abs(2 * sigma * KE - Dkin + Wthm) / max(abs(2 * sigma * KE), abs(Dkin), abs(Wthm))
        if par.thermal:
            resid3 = abs(2 * sigma * TE - Dthm - Wadv) / max(abs(2 * sigma * TE), abs(Dthm), abs(Wadv))
    return Diagnostics(KE, KP, KT, Dkin, Dint, Wthm, TE, Dthm, Wadv, resid0, resid1, resid3, Dkin0, Dint0, per_l)


def field_sections(par: Params, X):
    """(flow, temperature) sections of solution columns X (Kore order)."""
    n = par.n
    flow = X[:2 * n] if par.hydro else None
    off = 2 * n * par.hydro + 2 * n * par.magnetic
    temp = X[off:off + n] if par.thermal else None
    return flow, temp


def main(argv=None):
    """Reads eigenvalues0.dat and the .field files of the current directory
    (bin/spin_doctor.py:27-66) and prints the balance table."""
    from .params import load_parameters
    argv = sys.argv[1:] if argv is None else argv
    ppath = argv[0] if argv else "parameters.py"
    par = load_parameters(ppath) if os.path.exists(ppath) else Params()
    n = par.n
    lam = None
    if os.path.isfile("eigenvalues0.dat"):
        ev = np.loadtxt("eigenvalues0.dat").reshape((-1, 2))
        lam = ev[:, 0] + 1j * ev[:, 1]
    u = t = None
    if os.path.isfile("real_flow.field"):
        u = (np.loadtxt("real_flow.field").reshape((2 * n, -1)) + 1j * np.loadtxt("imag_flow.field").reshape((2 * n, -1)))
    if os.path.isfile("real_temperature.field"):
        t = (np.loadtxt("real_temperature.field").reshape((n, -1))
             + 1j * np.loadtxt("imag_temperature.field").reshape((n, -1)))
    ns = u.shape[1] if u is not None else (t.shape[1] if t is not None else 0)
    print("   i     sigma            w             resid0     resid1     resid3     Tor/Pol")
    for i in range(ns):
        li = lam[i] if lam is not None and par.forcing == 0 else None
        d = diagnose(par, li, None if u is None else u[:, i], None if t is None else t[:, i])
        s, w = (li.real, li.imag) if li is not None else (0.0, par.wf)
        print(" %3d  % 12.9f  % 12.9f   %8.2e   %8.2e   %8.2e   %8.2e" % (
            i, s, w, d.resid0, d.resid1, d.resid3, d.KT / d.KP if d.KP else float("nan")))
    return 0


if __name__ == "__main__":
    sys.exit(main())
