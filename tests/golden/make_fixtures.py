#!/usr/bin/env python3
"""Write golden vectors for the operator-assembly stage that feeds the solve path.

This script is run ONCE, in a container that holds the Kore reference tree at
/root/reference (it never runs on a GPU machine and nothing at run time reads
the reference).  For each case it writes a ``parameters.py`` with the case's
knobs substituted into the reference's own defaults file (the way the
reference's tests do it with ``sed``, tests/jones2000/find_Rac.py:31,44), then
in a fresh subprocess imports the reference modules that depend only on
numpy/scipy and records their outputs:

* ``submatrices.main`` (bin/submatrices.py:28-597): every radial operator
  ``*.mtx`` it writes, stored as CSR arrays (for magnetic cases after
  ``compute_profiles`` has written the radial profiles, bin/compute_profiles.py,
  as the reference's workflow does);
* ``operators.u / coriolis / viscous_diffusion`` (bin/operators.py:22-194) and,
  for the magneto-thermal case, ``lorentz / b / induction /
  magnetic_diffusion / buoyancy / theta / thermal_advection /
  thermal_diffusion`` for every degree l and block offset the assembly uses;
* ``bc_variables.Ta / Tb`` (bin/bc_variables.py:54-55) and the
  ``utils`` globals ``N1, n, sizmat, m_top, m_bot`` plus ``utils.ell``
  (bin/utils.py:26-44, 174-183).

``bin/assemble.py`` is not imported: it needs ``mpi4py``, which is absent
here, and no stand-in module is written for it.  Its global placement logic is
pinned instead by the reference's spin-over golden eigenvalue
(tests/spinover/reference.eig) and by structural invariants.

Usage:  python tests/golden/make_fixtures.py            (writes tests/golden/*.npz)
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

REF_BIN = "/root/reference/bin"
HERE = os.path.dirname(os.path.abspath(__file__))

# Knob overrides per case.  Values are Python source text.
CASES = {
    # Plain-method case: hydrodynamic inertial modes, m=1, E=1e-4, ricb=0.35,
    # no-slip, equatorially symmetric, N=24, lmax=40.
    "hydro_m1_E1e-4_N24_L40": dict(
        m="1", symm="1", ricb="0.35", bci="1", bco="1", Ek="1e-4",
        N="24", lmax="40", ncpus="2", itau="1.0", nev="5", tol="1e-12"),
    # The reference's spin-over test (tests/spinover/params.spinover is
    # identical to bin/parameters.py): m=1, symm=-1, Ek=1e-3, N=68, lmax=64.
    "spinover": dict(),
    # Thermal convection (Boussinesq, differential heating, isothermal walls),
    # the setting of the reference's Dormy et al. 2004 onset test
    # (tests/dormy2004/params.dormy04) at a small resolution.
    "thermal_m9_N24_L32": dict(
        m="9", symm="1", ricb="0.35", bci="1", bco="1", Ek="2e-5",
        N="24", lmax="32", ncpus="2", thermal="1", Prandtl="1", heating="'differential'",
        Ra_gap="1654042.168683", bci_thermal="0", bco_thermal="0"),
    # Full sphere (no inner core): thermal convection with internal heating
    # and fixed-flux boundary, the setting of the reference's Jones et al.
    # 2000 onset test (tests/jones2000/params.jones) at a small resolution.
    "jones_m9_N24_L32": dict(
        m="9", symm="1", ricb="0", bci="1", bco="1", Ek="np.sqrt(4/1e9)",
        N="24", lmax="32", ncpus="2", thermal="1", Prandtl="1", heating="'internal'",
        Ra_gap="4.6e6", bci_thermal="1", bco_thermal="1"),
    # Small magneto-thermal (MAC) case: 5 fields, m=2, symmetric flow, axial B0.
    "mac_m2_N16_L17": dict(
        m="2", symm="1", ricb="0.35", bci="1", bco="1", Ek="1e-3",
        N="16", lmax="17", ncpus="2", magnetic="1", thermal="1",
        Lambda="0.1", Pm="1e-3", Prandtl="1", Ra_gap="6e5"),
}

# Degrees sampled for the per-l operator fixtures (all of them for small cases).
MAX_L_SAMPLES = 12


def make_params(overrides):
    src = open(os.path.join(REF_BIN, "parameters.py")).read()
    for name, val in overrides.items():
        pat = re.compile(r"^([ \t]*)%s\s*=.*$" % re.escape(name), re.M)
        if not pat.search(src):
            raise KeyError(name)
        src = pat.sub(lambda mo: "%s%s = %s" % (mo.group(1), name, val), src, count=1)
    return src


CHILD = r'''
import sys, os, json
import numpy as np
import scipy.sparse as ss
import scipy.io as sio
sys.path.insert(0, os.getcwd())
sys.path.insert(1, REF_BIN)
import parameters as par
import utils as ut
if par.magnetic:
    import compute_profiles      # writes radProfs.mat (the diffusivity profile submatrices reads)
import submatrices
submatrices.main(2)
import bc_variables as bv
import operators as op

out = {}
meta = dict(N1=int(ut.N1), n=int(ut.n), sizmat=int(ut.sizmat), m_top=int(ut.m_top),
            m_bot=int(ut.m_bot), N=int(par.N), lmax=int(par.lmax), m=int(par.m),
            symm=int(par.symm), ricb=float(par.ricb), Ek=float(par.Ek),
            hydro=int(par.hydro), magnetic=int(par.magnetic), thermal=int(par.thermal),
            compositional=int(par.compositional), symmB0=int(ut.symmB0),
            tau=[float(np.real(par.tau)), float(np.imag(par.tau))], nev=int(par.nev))
ll = ut.ell(par.m, par.lmax, par.symm)
out["ell_top"] = np.asarray(ll[0]); out["ell_bot"] = np.asarray(ll[1]); out["ell_all"] = np.asarray(ll[2])
llm = ut.ell(par.m, par.lmax, par.symm*ut.symmB0)
out["ellmag_top"] = np.asarray(llm[0]); out["ellmag_bot"] = np.asarray(llm[1])
out["Ta"] = np.asarray(bv.Ta); out["Tb"] = np.asarray(bv.Tb)

def put(key, mtx):
    if isinstance(mtx, (int, float)) and mtx == 0:
        return
    m = ss.csr_matrix(mtx)
    m.sort_indices()
    out[key + "__data"] = m.data.astype(complex)
    out[key + "__indices"] = m.indices.astype(np.int32)
    out[key + "__indptr"] = m.indptr.astype(np.int32)
    out[key + "__shape"] = np.asarray(m.shape)

import glob
for f in sorted(glob.glob("*.mtx")):
    put("mtx:" + f[:-4], ss.csr_matrix(sio.mmread(f)))

def sample(ls):
    ls = list(ls)
    if len(ls) <= MAX_L:
        return ls
    k = MAX_L // 2
    return ls[:k] + ls[-k:]

top, bot = ll[0], ll[1]
for l in sample(top):
    l = int(l)
    put("u:u:upol:%d" % l, op.u(l, 'u', 'upol', 0))
    put("visc:u:upol:%d" % l, op.viscous_diffusion(l, 'u', 'upol', 0))
    c = op.coriolis(l, 'u', 'upol', 0); put("cori:u:upol:0:%d" % l, c[0])
    for i in (-1, 1):
        c = op.coriolis(l, 'u', 'utor', i)
        put("cori:u:utor:%d:%d" % (i, l), c[0]); out["offd:cori:u:utor:%d:%d" % (i, l)] = np.asarray(c[1])
for l in sample(bot):
    l = int(l)
    put("u:v:utor:%d" % l, op.u(l, 'v', 'utor', 0))
    put("visc:v:utor:%d" % l, op.viscous_diffusion(l, 'v', 'utor', 0))
    c = op.coriolis(l, 'v', 'utor', 0); put("cori:v:utor:0:%d" % l, c[0])
    for i in (-1, 1):
        c = op.coriolis(l, 'v', 'upol', i)
        put("cori:v:upol:%d:%d" % (i, l), c[0]); out["offd:cori:v:upol:%d:%d" % (i, l)] = np.asarray(c[1])

if par.magnetic:
    # only the (degree, offset) pairs bin/assemble.py:653-684, 765-796, 837-861,
    # 936-960 actually request (l+i inside the target degree set)
    topset, botset = set(int(x) for x in ll[0]), set(int(x) for x in ll[1])
    bpolset, btorset = set(int(x) for x in llm[0]), set(int(x) for x in llm[1])
    for sec, comp, offs, src, tgt in (('u', 'bpol', (-2, -1, 0, 1, 2), top, bpolset),
                                      ('u', 'btor', (-1, 0, 1), top, btorset),
                                      ('v', 'bpol', (-1, 0, 1), bot, bpolset),
                                      ('v', 'btor', (-2, -1, 0, 1, 2), bot, btorset)):
        for l in sample(src):
            for i in offs:
                if int(l) + i not in tgt:
                    continue
                c = op.lorentz(int(l), sec, comp, i)
                put("lorentz:%s:%s:%d:%d" % (sec, comp, i, l), c[0])
                out["offd:lorentz:%s:%s:%d:%d" % (sec, comp, i, l)] = np.asarray(c[1])
    f_l = llm[0] if ut.symmB0 == 1 else llm[1]
    g_l = llm[1] if ut.symmB0 == 1 else llm[0]
    f_l, g_l = ll[0] if ut.symmB0 == 1 else ll[1], ll[1] if ut.symmB0 == 1 else ll[0]
    for l in sample(f_l):
        l = int(l)
        put("b:f:bpol:%d" % l, op.b(l, 'f', 'bpol', 0))
        put("mdiff:f:bpol:%d" % l, op.magnetic_diffusion(l, 'f', 'bpol', 0))
        for comp, offs, tgt in (('upol', (-2, -1, 0, 1, 2), topset), ('utor', (-1, 0, 1), botset)):
            for i in offs:
                if l + i not in tgt:
                    continue
                c = op.induction(l, 'f', comp, i)
                put("induction:f:%s:%d:%d" % (comp, i, l), c[0])
                out["offd:induction:f:%s:%d:%d" % (comp, i, l)] = np.asarray(c[1])
    for l in sample(g_l):
        l = int(l)
        put("b:g:btor:%d" % l, op.b(l, 'g', 'btor', 0))
        put("mdiff:g:btor:%d" % l, op.magnetic_diffusion(l, 'g', 'btor', 0))
        for comp, offs, tgt in (('upol', (-1, 0, 1), topset), ('utor', (-2, -1, 0, 1, 2), botset)):
            for i in offs:
                if l + i not in tgt:
                    continue
                c = op.induction(l, 'g', comp, i)
                put("induction:g:%s:%d:%d" % (comp, i, l), c[0])
                out["offd:induction:g:%s:%d:%d" % (comp, i, l)] = np.asarray(c[1])
    meta["cnorm"] = float(ut.B0_norm())
    meta["Le2"] = float(par.Le2); meta["Em"] = float(par.Em); meta["Hendrik"] = float(par.Hendrik)
    meta["MagnetD"] = float(par.MagnetD)

if par.thermal:
    meta["Beyonce"] = float(par.Beyonce); meta["ThermaD"] = float(par.ThermaD); meta["heating"] = str(par.heating)
    meta["Ra_gap"] = float(par.Ra_gap); meta["Prandtl"] = float(par.Prandtl)
    for l in sample(top):
        l = int(l)
        put("buoy:u:%d" % l, op.buoyancy(l, 'u', '', 0))
        put("theta:h:%d" % l, op.theta(l, 'h', '', 0))
        put("tadv:h:%d" % l, op.thermal_advection(l, 'h', 'upol', 0))
        put("tdiff:h:%d" % l, op.thermal_diffusion(l, 'h', '', 0))

np.savez_compressed(OUTFILE, meta=json.dumps(meta), **out)
print("wrote", OUTFILE, len(out), "arrays")
'''


def run_case(name, overrides):
    outfile = os.path.join(HERE, "ops_%s.npz" % name)
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "parameters.py"), "w") as f:
            f.write(make_params(overrides))
        code = (CHILD.replace("REF_BIN", repr(REF_BIN))
                .replace("OUTFILE", repr(outfile))
                .replace("MAX_L", str(MAX_L_SAMPLES)))
        subprocess.run([sys.executable, "-c", code], cwd=td, check=True)
    return outfile


def main():
    only = sys.argv[1:]
    for name, ov in CASES.items():
        if only and name not in only:
            continue
        run_case(name, ov)


if __name__ == "__main__":
    main()
