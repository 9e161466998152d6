"""Solver and physics knobs, with the derived problem sizes.

Kore configures a run through a plain Python module ``parameters.py``
(bin/parameters.py) whose globals are read by every stage.  ``Params`` holds
the subset of those knobs that the solve path and the hydrodynamic /
magneto-thermal assembly read, under the reference's own names, and
``load_parameters`` reads a user's ``parameters.py`` into one.

Derived sizes follow bin/utils.py:26-44:

* ``N1 = N`` when an inner core is present (``ricb > 0``), else ``N/2``;
* ``n = N1*(lmax-m+1)/2`` unknowns per scalar section;
* ``sizmat = 2n*hydro + 2n*magnetic + n*thermal + n*compositional``.
"""
from __future__ import annotations

import dataclasses
import importlib.util
import math
import os
from typing import Optional

import numpy as np


def ncheb(ek: float) -> int:
    """Radial truncation rule of bin/parameters.py:6-16 (``Ncheb``)."""
    out = int(17 * ek ** -0.2) if ek != 0 else 48
    return max(48, out + out % 2)


@dataclasses.dataclass
class Params:
    # hydrodynamics (bin/parameters.py:26-51)
    hydro: int = 1
    m: int = 1
    symm: int = -1
    ricb: float = 0.35
    bci: int = 1
    bco: int = 1
    Ek: float = 1e-3
    anelastic: int = 0
    # forcing (bin/parameters.py:76-92)
    forcing: int = 0
    forcing_frequency: float = 1.0
    forcing_amplitude_cmb: float = 1.0
    forcing_amplitude_icb: float = 0.0
    forcing_amplitude: Optional[float] = None    # body / tide forcings 2, 3, 6 (not in bin/parameters.py)
    # magnetic field (bin/parameters.py:103-141)
    magnetic: int = 0
    B0: str = "axial"
    innercore: str = "insulator"
    mantle: str = "insulator"
    Lambda: float = 0.1
    Pm: float = 0.001
    Le2: Optional[float] = None
    Em: Optional[float] = None
    cnorm: object = "mag_energy"
    # thermal (bin/parameters.py:162-214)
    thermal: int = 0
    Prandtl: float = 1.0
    heating: str = "differential"
    Ra_gap: float = 6e5
    bci_thermal: int = 0
    bco_thermal: int = 0
    Beyonce: Optional[float] = None      # buoyancy factor; None -> -Ra Ek^2 / Prandtl
    ThermaD: Optional[float] = None      # thermal diffusion factor; None -> Ek / Prandtl
    compositional: int = 0
    # time-unit switches (bin/parameters.py:273-278); None -> reference default
    Gaspard: float = 1.0
    ViscosD: Optional[float] = None
    # resolution (bin/parameters.py:292-304)
    N: Optional[int] = None
    lmax: Optional[int] = None
    ncpus: int = 4
    # solver knobs (bin/parameters.py:316-347)
    tau: complex = 1.0j
    which_eigenpairs: str = "TM"
    nev: int = 5
    ncv: Optional[int] = None
    maxit: int = 50
    tol: float = 1e-15

    def __post_init__(self):
        if self.N is None:
            self.N = ncheb(self.Ek)
        if self.lmax is None:
            g = 1.0
            self.lmax = int(2 * self.ncpus * (np.floor_divide(g * self.N, 2 * self.ncpus)) + self.m - 1)
        if self.ViscosD is None:
            self.ViscosD = self.Ek
        if self.Em is None:
            self.Em = self.Ek / self.Pm
        if self.Le2 is None:
            self.Le2 = self.Lambda * self.Em
        if self.Beyonce is None:
            # Ra = Ra_gap / (rcmb - ricb)^3, BV2 = -Ra Ek^2 / Prandtl (bin/parameters.py:177-181, 274)
            ra = self.Ra_gap / (1.0 - self.ricb) ** 3
            self.Beyonce = -ra * self.Ek ** 2 / self.Prandtl
        if self.ThermaD is None:
            self.ThermaD = self.Ek / self.Prandtl       # Etherm (bin/parameters.py:167, 277)
        self.tau = complex(self.tau)
        self.validate()

    # ------------------------------------------------------------------ sizes
    def validate(self):
        if (self.lmax - self.m + 1) % 2:
            raise ValueError("lmax - m + 1 must be even (got lmax=%d, m=%d); Kore's l lists "
                             "split evenly into poloidal/toroidal sets" % (self.lmax, self.m))
        if self.symm not in (1, -1):
            raise ValueError("symm must be +1 or -1")
        if self.ricb == 0 and self.N % 2:
            raise ValueError("N must be even when ricb = 0")

    @property
    def N1(self) -> int:
        return self.N if self.ricb > 0 else self.N // 2

    @property
    def nb(self) -> int:
        """Blocks (degrees l) per section."""
        return (self.lmax - self.m + 1) // 2

    @property
    def n(self) -> int:
        return self.N1 * self.nb

    @property
    def sizmat(self) -> int:
        n = self.n
        return 2 * n * self.hydro + 2 * n * self.magnetic + n * self.thermal + n * self.compositional

    @property
    def s(self) -> int:
        return int((self.symm + 1) // 2)

    @property
    def m_top(self) -> int:
        v = self.m + 1 - self.s
        return 2 if v == 0 else v

    @property
    def m_bot(self) -> int:
        v = self.m + self.s
        return 2 if v == 0 else v

    @property
    def symm1(self) -> int:
        return (2 * int(np.sign(self.m)) - 1) * self.symm

    @property
    def vP(self) -> int:
        """Radial parity of the poloidal flow coefficients without an inner
        core (bin/submatrices.py:45-47); toroidal vT = -vP, temperature vP."""
        return int((1 - 2 * (self.m % 2)) * self.symm)

    @property
    def symmB0(self) -> int:
        if self.B0 in ("axial", "dipole", "G21 dipole", "Luo_S1"):
            return -1
        if self.B0 == "Luo_S2":
            return 1
        raise ValueError("unsupported B0 %r" % self.B0)

    @property
    def wf(self) -> float:
        return 0.0 if self.forcing == 0 else float(self.forcing_frequency)

    @property
    def fields(self):
        out = []
        if self.hydro:
            out += ["flow"]
        if self.magnetic:
            out += ["magnetic"]
        if self.thermal:
            out += ["temperature"]
        if self.compositional:
            out += ["composition"]
        return out


_KNOBS = [f.name for f in dataclasses.fields(Params)]


def load_parameters(path: str) -> Params:
    """Read a Kore ``parameters.py`` (same variable names) into ``Params``.

    The file is executed as a module, as Kore itself does with
    ``import parameters as par`` (bin/solve.py:24).  Knobs this path does not
    use are ignored.
    """
    path = os.path.abspath(path)
    spec = importlib.util.spec_from_file_location("_kore_user_parameters", path)
    mod = importlib.util.module_from_spec(spec)
    cwd = os.getcwd()
    try:
        spec.loader.exec_module(mod)
    finally:
        os.chdir(cwd)
    kw = {}
    for k in _KNOBS:
        if hasattr(mod, k):
            kw[k] = getattr(mod, k)
    # the reference derives these at module level; keep its values verbatim
    for k in ("Le2", "Em", "ViscosD", "Beyonce", "ThermaD"):
        if hasattr(mod, k):
            kw[k] = getattr(mod, k)
    if hasattr(mod, "MagnetD"):
        kw["Em"] = getattr(mod, "MagnetD")
    if hasattr(mod, "Hendrik"):
        kw["Le2"] = getattr(mod, "Hendrik")
    return Params(**kw)


def ell(m: int, lmax: int, vsymm: int):
    """Degrees of the poloidal and toroidal sections (bin/utils.py:174-183).

    Returns ``(l_pol, l_tor, l_all)``; for m = 0 the degrees start at 1.
    Which parity goes to the poloidal list is fixed by ``(sign(m) + s) % 2``
    with ``s = 1`` for equatorially symmetric flow.
    """
    lm1 = lmax - m + 1
    s = int(vsymm * 0.5 + 0.5)
    sg = int(np.sign(m))
    l_all = np.arange(m + 1 - sg, lmax + 2 - sg, dtype=np.int64)
    first_pol = (sg + s) % 2
    return l_all[first_pol:lm1:2], l_all[1 - first_pol:lm1:2], l_all


def suggested_ncv(nev: int) -> int:
    """Krylov subspace size used when the caller does not set one."""
    return max(2 * nev, nev + 15)


def frob(x) -> float:
    return math.sqrt(float(np.vdot(x, x).real))
