"""Right-hand sides of Kore's forced problems (B_forced.npz).

Restates the forcing branch of bin/assemble.py:77-427 for the forcings the
reference can build:

* 2  eccentricity tide as a boundary flow (l = 2 poloidal, boundary rows of
     the ICB and CMB; m = 0 or 2, symmetric, no-slip), assemble.py:106-130,
     values from utils.eccen_tide (utils.py:1119-1191);
* 3  Rovira-Navarro et al. (2018) tidal body force (l = 1, 3 toroidal and
     l = 2 poloidal interior rows; m < 3), assemble.py:133-178, values from
     utils.marc_tide (utils.py:1079-1116);
* 6  Buffett (2010) ICB radial velocity (l = 2 poloidal ICB rows; m = 1,
     antisymmetric), assemble.py:238-270;
* 7  longitudinal libration as a boundary flow (m = 0: l = 1 toroidal; m = 2:
     l = 2 poloidal), assemble.py:272-310;
* 9  m = 2 radial velocity at the boundaries (l = 2 poloidal), assemble.py:355-383;
* 10 order-m radial velocity at the CMB (l = m poloidal), assemble.py:386-414.

Forcings 2, 3 and 6 scale with the knob ``forcing_amplitude`` (which
bin/parameters.py does not define: a user's parameters.py must), the others
with forcing_amplitude_cmb / _icb.  Forcings 1, 4, 5 and 8 multiply by
operator objects (op.Ig, Ib, r2Ib, op.Iv) that the reference's own
submatrices.py never writes, so the reference cannot build them; they are
refused here with that reason.

Rows: the poloidal section starts at 0 and the toroidal one at n, with the
N1 coefficients of the k-th degree of each section's list at k N1
(ell, utils.py:174-183); boundary conditions occupy a degree's first rows.
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sft
import scipy.sparse as ss

from .params import Params, ell


def chebco_pow(powr: float, N: int, tol: float, ricb: float, rcmb: float = 1.0) -> np.ndarray:
    """First N Chebyshev coefficients of r**powr on the solution's radial
    domain ([ricb, rcmb], or [-rcmb, rcmb] without an inner core): a DCT-II of
    the values at the N Chebyshev points, coefficients below tol set to zero
    (utils.chebco, utils.py:251-269)."""
    x = np.cos(np.pi * (np.arange(N) + 0.5) / N)
    r = rcmb * x if ricb == 0 else ricb + (rcmb - ricb) * (x + 1) / 2.0
    c = sft.dct(r ** powr) / N
    c[0] /= 2.0
    c[np.abs(c) <= tol] = 0.0
    return c


# Eccentricity tide at the boundaries: rational fits in the inner-core radius
# eta, (a0 + a1 y + a2 y^2 + a3 y^3) / (b0 + b1 y + b2 y^2 + b3 y^3) with
# y = eta - 37/42, for the poloidal scalar P and its radial derivative dP, per
# (boundary, m) -- the fitted constants of utils.eccen_tide (utils.py:1119-1191).
_ECCEN = {
    ('cmb', -2): ((3.00927601455473e-07, 9.923006738722803e-07, 2.0022902696912132e-06, 2.078150971132207e-06), (0.9796806966104893, 10.138063592395676, 22.793605369253733, 49.52238501116611),
                   (3.6923409583928235e-08, 5.525597059314317e-07, 1.491072486300637e-06, 1.7131360325476716e-06), (0.990016066975477, 10.903688391808103, 28.973467794986213, 54.35051022571307)),
    ('cmb', 0): ((-1.0530272472650882e-07, -3.4723290320109935e-07, -7.006556396692487e-07, -7.272013553918712e-07), (0.9796806966105374, 10.138063592396957, 22.793605369262107, 49.52238501117491),
                  (-1.2920501863486681e-08, -1.933556188506115e-07, -5.217666801343201e-07, -5.994727342457128e-07), (0.990016066975396, 10.903688391807878, 28.973467794991336, 54.35051022572941)),
    ('cmb', 2): ((-4.2989657350784934e-08, -1.4175723912463045e-07, -2.860414670987557e-07, -2.9687871016172726e-07), (0.9796806966105591, 10.1380635923966, 22.793605369256564, 49.522385011165795),
                  (-5.274772797703058e-09, -7.893710084733917e-08, -2.1301035518586083e-07, -2.4473371893559314e-07), (0.990016066975294, 10.903688391807277, 28.973467794994058, 54.35051022573984)),
    ('icb', -2): ((1.450645275222846e-09, 1.6226580438374377e-08, 3.8162533600883e-08, 2.1621482543044536e-08), (1.0040474189168584, 9.816014125318514, 12.467747302017761, -0.13326826070343542),
                   (5.690951704457902e-11, 6.365766130462465e-10, 1.4971346844873185e-09, 8.482212366654875e-10), (1.0040474189168667, 9.816014125318599, 12.467747302017914, -0.13326826070340814)),
    ('icb', 0): ((-5.076201031536069e-10, -5.678120334892032e-09, -1.3354104944854432e-08, -7.565942816136093e-09), (1.0040474189168602, 9.81601412531852, 12.467747302017683, -0.13326826070345546),
                  (-1.991418260963427e-11, -2.22755412021775e-10, -5.238880076023669e-10, -2.96815602688989e-10), (1.0040474189168587, 9.816014125318501, 12.46774730201766, -0.13326826070345907)),
    ('icb', 2): ((-2.072350393175482e-10, -2.3180829197677505e-09, -5.451790514411792e-09, -3.0887832204348696e-09), (1.0040474189168525, 9.816014125318436, 12.46774730201756, -0.13326826070347508),
                  (-8.129931006368303e-12, -9.093951614946213e-11, -2.138763834981828e-10, -1.2117446238077878e-10), (1.0040474189168571, 9.816014125318477, 12.467747302017562, -0.13326826070347858)),
}
_ECCEN_Y0 = 0.8809523809523809


def eccentricity_tide(m: int, eta: float, boundary: str) -> np.ndarray:
    """[P, dP] of the eccentricity tide at the ICB or CMB for m in {-2, 0, 2}."""
    nP, dPn, ndP, ddP = _ECCEN[(boundary, int(m))]
    y = eta - _ECCEN_Y0
    pw = np.array([1.0, y, y * y, y ** 3])
    return np.array([np.dot(nP, pw) / np.dot(dPn, pw), np.dot(ndP, pw) / np.dot(ddP, pw)])


def tidal_body_force(omega: float, l: int, m: int, loc: str, N1: int, ricb: float) -> np.ndarray:
    """Coefficients of the Rovira-Navarro et al. (2018) tidal body force for
    degree l (utils.marc_tide): poloidal l = 2 ('top', N1-4 interior rows,
    r^5 profile), toroidal l = 1 and 3 ('bot', N1-2 interior rows, r^4)."""
    phase = {0: 1j, 1: 1.0, 2: 1j}[m]
    if l == 2 and loc == "top":
        return -6j * phase * (omega + m) * chebco_pow(5, N1 - 4, 2e-16, ricb)
    r4 = chebco_pow(4, N1 - 2, 2e-16, ricb)
    amp = {(0, 1): 4 / 5, (0, 3): -24 / 5, (1, 1): (2 / 5) * np.sqrt(3), (1, 3): -(16 / 5) * np.sqrt(2),
           (2, 3): -8 / np.sqrt(5)}
    if (m, l) not in amp:
        raise ValueError("no tidal body force term for l = %d, m = %d" % (l, m))
    return amp[(m, l)] * phase * r4


def _pos(par: Params, section: str, l: int) -> int:
    top, bot, _ = ell(par.m, par.lmax, par.symm)
    lst, base = (top, 0) if section == "pol" else (bot, par.n)
    hit = np.where(lst == l)[0]
    if len(hit) == 0:
        raise ValueError("degree %d is not in the %s list for m = %d, symm = %d" % (l, section, par.m, par.symm))
    return base + par.N1 * int(hit[0])


def _amplitude(par: Params) -> float:
    if par.forcing_amplitude is None:
        raise ValueError("forcing %d scales with forcing_amplitude, which the parameters do not set" % par.forcing)
    return float(par.forcing_amplitude)


def forcing_vector(par: Params) -> ss.csr_matrix:
    """B_forced (sizmat x 1, complex) of ``par.forcing`` (module docstring)."""
    f = int(par.forcing)
    w = float(par.forcing_frequency)
    rows, data = [], []
    if f == 2:
        if not (par.symm == 1 and par.m in (0, 2) and par.bci + par.bco == 2):
            raise ValueError("the eccentricity tide needs symm = 1, m = 0 or 2 and no-slip boundaries")
        amp = _amplitude(par)
        ms = int(par.m * np.sign(w))
        fc = 1j * amp * eccentricity_tide(ms, par.ricb, "cmb")
        fi = 1j * amp * eccentricity_tide(ms, par.ricb, "icb")
        p = _pos(par, "pol", 2)
        rows, data = [p, p + 1, p + 2, p + 3], [fi[0], fi[1], fc[0], fc[1]]
    elif f == 3:
        if par.m >= 3:
            raise ValueError("the tidal body force needs m < 3")
        amp = _amplitude(par)
        wf = w
        terms = [(3, "tor", 2), (2, "pol", 4)] + ([(1, "tor", 2)] if par.m < 2 else [])
        for l, sec, skip in terms:
            p = _pos(par, sec, l)
            v = tidal_body_force(wf, l, par.m, "top" if sec == "pol" else "bot", par.N1, par.ricb) * amp
            rows += list(range(p + skip, p + par.N1))
            data += list(v)
    elif f == 6:
        if not (par.m == 1 and par.symm == -1):
            raise ValueError("the ICB radial velocity forcing needs m = 1, symm = -1")
        l, hft = 2, 0.0025                       # inner-core hydrostatic flattening
        C = 1j * (2 / 3) * np.sqrt(6) * hft * par.ricb
        P = (par.ricb / (l * (l + 1))) * C * _amplitude(par)
        second = -P / par.ricb if par.bci == 1 else (2 - l * (l + 1)) * P / par.ricb ** 2
        p = _pos(par, "pol", 2)
        rows, data = [p, p + 1], [P, second]
    elif f == 7:
        if not (par.symm == 1 and par.bci == 1 and par.bco == 1 and par.m in (0, 2)):
            raise ValueError("longitudinal libration requires m = 0 or 2, symm = 1 and no-slip boundaries")
        if par.m == 0:
            p = _pos(par, "tor", 1)
            data = [1j * w * par.forcing_amplitude_cmb / 2, 1j * w * par.forcing_amplitude_icb * par.ricb / 2]
        else:
            L = 6.0
            p = _pos(par, "pol", 2)
            data = [-w / L * par.forcing_amplitude_cmb, -w / L * par.forcing_amplitude_icb * par.ricb ** 2]
        rows = [p, p + 1]
    elif f == 9:
        if not (par.symm == 1 and par.m == 2 and par.bci == 1):
            raise ValueError("the m = 2 radial forcing needs symm = 1, m = 2 and bci = 1")
        L = 6.0
        p = _pos(par, "pol", 2)
        rows, data = [p, p + 1], [par.forcing_amplitude_cmb / L, par.forcing_amplitude_icb * par.ricb / L]
    elif f == 10:
        if par.symm != 1:
            raise ValueError("the order-m radial forcing needs symm = 1")
        L = par.m * (par.m + 1.0)
        p = _pos(par, "pol", par.m)
        rows, data = [p, p + 1], [1.0 / L, 0.0]
    elif f in (1, 4, 5, 8):
        raise NotImplementedError(
            "forcing %d multiplies by an operator object (op.Ig, Ib, r2Ib or op.Iv, bin/assemble.py:92, 187, 212, "
            "350) that the reference's submatrices.py never writes: the reference cannot build it" % f)
    else:
        raise ValueError("forcing %d is not a forced problem" % f)
    return ss.csr_matrix((np.asarray(data, dtype=complex), (np.asarray(rows), np.zeros(len(rows), dtype=int))),
                         shape=(par.sizmat, 1))
