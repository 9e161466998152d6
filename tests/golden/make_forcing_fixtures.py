#!/usr/bin/env python3
"""Write golden B_forced vectors (run once, in a container holding the Kore
reference at /root/reference).

bin/assemble.py cannot be imported here (it needs mpi4py), so for each case
a child process imports the reference's numpy-only modules (parameters,
utils, operators) with a ``parameters.py`` holding the case's knobs, takes
the source text of assemble.py's forcing branch (the ``if par.forcing == 1:
... elif par.forcing == 10:`` chain of main(), bin/assemble.py:77-427) and
executes it with the names main() defines before it (rank 0, alltop /
allbot from ut.ell); the branch writes B_forced.npz, which is stored.

Two names the branch uses are missing from the reference itself and are
supplied: ``sign`` (forcing 2 calls an unimported sign(); numpy's is given)
and the knob ``forcing_amplitude`` (read by forcing 2, 3 and 6, absent from
bin/parameters.py; appended to the case's parameters.py).  Forcings 1, 4, 5
and 8 use operator objects (op.Ig, Ib, r2Ib, op.Iv) that the reference's
submatrices.py never writes, so those branches raise in the reference and
have no fixture.

Usage:  python tests/golden/make_forcing_fixtures.py   (writes tests/golden/forcing_*.npz)
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_fixtures import REF_BIN, make_params   # noqa: E402

CASES = {
    "f2_m2": dict(forcing="2", m="2", symm="1", ricb="0.35", Ek="1e-4", N="24", lmax="27", ncpus="2",
                  forcing_frequency="-1.3"),
    "f3_m0": dict(forcing="3", m="0", symm="1", ricb="0.35", Ek="1e-4", N="24", lmax="25", ncpus="2",
                  forcing_frequency="0.7"),
    "f3_m1": dict(forcing="3", m="1", symm="-1", ricb="0.35", Ek="1e-4", N="24", lmax="26", ncpus="2",
                  forcing_frequency="1.1"),
    "f6_m1": dict(forcing="6", m="1", symm="-1", ricb="0.35", Ek="1e-4", N="24", lmax="26", ncpus="2",
                  forcing_frequency="0.9"),
    "f6_m1_stressfree": dict(forcing="6", m="1", symm="-1", ricb="0.35", bci="0", Ek="1e-4", N="24", lmax="26",
                             ncpus="2", forcing_frequency="0.9"),
    "f7_m0": dict(forcing="7", m="0", symm="1", ricb="0.35", Ek="1e-4", N="24", lmax="25", ncpus="2",
                  forcing_frequency="1.2", forcing_amplitude_icb="0.3"),
    "f7_m2": dict(forcing="7", m="2", symm="1", ricb="0.35", Ek="1e-4", N="24", lmax="27", ncpus="2",
                  forcing_frequency="-0.8", forcing_amplitude_icb="0.4"),
    "f9_m2": dict(forcing="9", m="2", symm="1", ricb="0.35", Ek="1e-4", N="24", lmax="27", ncpus="2",
                  forcing_frequency="0.5", forcing_amplitude_icb="0.25"),
    "f10_m3": dict(forcing="10", m="3", symm="1", ricb="0.35", Ek="1e-4", N="24", lmax="28", ncpus="2",
                   forcing_frequency="0.5"),
}
EXTRA = "\nforcing_amplitude = 0.75\n"

CHILD = r'''
import sys, os, json, textwrap
import numpy as np
import scipy.sparse as ss
sys.path.insert(0, os.getcwd())
sys.path.insert(1, REF_BIN)
import parameters as par
import utils as ut
import operators as op
src = open(os.path.join(REF_BIN, "assemble.py")).read().splitlines()
i0 = next(i for i, l in enumerate(src) if l.strip() == "if par.forcing == 1: # -------------------------------------------------------------------------- Yufeng's forcing" or l.strip().startswith("if par.forcing == 1:"))
i1 = next(i for i, l in enumerate(src) if i > i0 and l.strip().startswith("elif par.forcing == 0:"))
code = textwrap.dedent("\n".join(src[i0:i1]))
alltop, allbot = ut.ell(par.m, par.lmax, par.symm)[:2]
ns = dict(np=np, ss=ss, ut=ut, par=par, op=op, rank=0, alltop=alltop, allbot=allbot, sign=np.sign)
exec(compile(code, "assemble.py forcing branch", "exec"), ns)
z = np.load("B_forced.npz")
meta = dict(m=int(par.m), symm=int(par.symm), ricb=float(par.ricb), Ek=float(par.Ek), N=int(par.N),
            lmax=int(par.lmax), bci=int(par.bci), bco=int(par.bco), forcing=int(par.forcing),
            forcing_frequency=float(par.forcing_frequency), forcing_amplitude=float(par.forcing_amplitude),
            forcing_amplitude_cmb=float(par.forcing_amplitude_cmb),
            forcing_amplitude_icb=float(par.forcing_amplitude_icb), sizmat=int(ut.sizmat))
np.savez(OUT, meta=json.dumps(meta), data=z["data"], indices=z["indices"], indptr=z["indptr"], shape=z["shape"])
'''


def make(name, over):
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "parameters.py"), "w") as f:
            f.write(make_params(over) + EXTRA)
        out = os.path.join(HERE, "forcing_%s.npz" % name)
        code = CHILD.replace("REF_BIN", repr(REF_BIN)).replace("OUT", repr(out))
        r = subprocess.run([sys.executable, "-c", code], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("%s failed:\n%s" % (name, r.stderr[-4000:]))
        print("wrote", out)


if __name__ == "__main__":
    for nm, ov in CASES.items():
        if len(sys.argv) == 1 or nm in sys.argv[1:]:
            make(nm, ov)
