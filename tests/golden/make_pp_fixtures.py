#!/usr/bin/env python3
"""Write golden vectors for the post-processing integrals (run once, in a
container holding the Kore reference at /root/reference).

For each case a ``parameters.py`` with the case's knobs substituted into the
reference's own defaults file is written into a temporary directory (as
make_fixtures.py does); a child process imports the reference's numpy-only
modules ``utils`` and ``utils4pp`` there and records, for a seeded complex
coefficient vector (flow: poloidal then toroidal, and temperature for
thermal cases; Kore ordering, coefficients decaying like exp(-k/8)):

* ``utils4pp.diagnose`` (bin/utils4pp.py:794-858): per degree the kinetic
  energy, the integrals of u . lap u (Dkin0) and of the internal dissipation
  (Dint0), the buoyancy power (Wthm0), and per poloidal degree the thermal
  energy, theta lap theta (Dthm0) and the advection power.

bin/spin_doctor.py itself is not run: it reads par.OmgTau, which the
reference's current parameters.py (bin/parameters.py:268-271) leaves
commented out, and for forced problems an undefined ``repow``
(spin_doctor.py:216-225).  The per-degree integrals it sums are what is
pinned here.

Usage:  python tests/golden/make_pp_fixtures.py     (writes tests/golden/pp_*.npz)
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_fixtures import REF_BIN, make_params   # noqa: E402

CASES = {
    "hydro_m1_N24_L40": dict(m="1", symm="1", ricb="0.35", Ek="1e-4", N="24", lmax="40", ncpus="2"),
    "spinover": dict(ncpus="2"),
    "thermal_m9_N24_L32": dict(m="9", symm="1", ricb="0.35", Ek="2e-5", N="24", lmax="32", ncpus="2",
                               thermal="1", Prandtl="1", heating="'differential'", Ra_gap="1654042.168683",
                               bci_thermal="0", bco_thermal="0"),
    "jones_m9_N24_L32": dict(m="9", symm="1", ricb="0", Ek="np.sqrt(4/1e9)", N="24", lmax="32", ncpus="2",
                             thermal="1", Prandtl="1", heating="'internal'", Ra_gap="4.6e6",
                             bci_thermal="1", bco_thermal="1"),
}

CHILD = r'''
import sys, os, json
import numpy as np
sys.path.insert(0, os.getcwd())
sys.path.insert(1, REF_BIN)
import parameters as par
import utils as ut
import utils4pp as upp
rng = np.random.default_rng(SEED)
n = ut.n
decay = np.exp(-np.arange(ut.N1) / 8.0)
def vec(nsec):
    v = (rng.standard_normal((nsec * n,)) + 1j * rng.standard_normal((nsec * n,)))
    return (v.reshape(-1, ut.N1) * decay).ravel()
u = vec(2)
t = vec(1) if par.thermal else None
usol2 = upp.expand_reshape_sol(u, par.symm)
tsol2 = upp.expand_reshape_sol(t, par.symm) if par.thermal else 0
udgn, bdgn, tdgn, cdgn = upp.diagnose(usol2, 0, tsol2, 0, par.ricb, ut.rcmb, 2)
out = dict(u=u, udgn=np.asarray(udgn))
if par.thermal:
    out["t"] = t
    out["tdgn"] = np.asarray(tdgn)
meta = dict(m=int(par.m), symm=int(par.symm), ricb=float(par.ricb), Ek=float(par.Ek), N=int(par.N),
            lmax=int(par.lmax), thermal=int(par.thermal), heating=str(par.heating),
            Prandtl=float(par.Prandtl), N1=int(ut.N1), n=int(ut.n))
np.savez(OUT, meta=json.dumps(meta), **out)
'''


def make(name, over):
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "parameters.py"), "w") as f:
            f.write(make_params(over))
        out = os.path.join(HERE, "pp_%s.npz" % name)
        code = CHILD.replace("REF_BIN", repr(REF_BIN)).replace("SEED", "2026").replace("OUT", repr(out))
        r = subprocess.run([sys.executable, "-c", code], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("%s failed:\n%s" % (name, r.stderr[-4000:]))
        print("wrote", out)


if __name__ == "__main__":
    for nm, ov in CASES.items():
        if len(sys.argv) == 1 or nm in sys.argv[1:]:
            make(nm, ov)
