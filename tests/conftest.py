import json
import os
import sys

import numpy as np
import pytest
import scipy.sparse as ss

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_ops_fixture(name):
    """Golden operator fixture written by tests/golden/make_fixtures.py."""
    d = np.load(os.path.join(GOLDEN, "ops_%s.npz" % name))
    meta = json.loads(str(d["meta"]))

    def get(key):
        return ss.csr_matrix((d[key + "__data"], d[key + "__indices"], d[key + "__indptr"]),
                             shape=tuple(d[key + "__shape"]))

    keys = sorted(set(k.split("__")[0] for k in d.files if "__" in k))
    return d, meta, get, keys


# Reference golden (tests/spinover/reference.eig:1 of Kore): lambda = re + i im
SPINOVER_LAMBDA = complex(-1.019947097016500742e-01, 1.005733690103750355e+00)


def params_hydro_small():
    from kore_amd.params import Params
    # hydrodynamic inertial modes: m=1, E=1e-4, ricb=0.35, no-slip, symmetric, N=24, lmax=40
    return Params(m=1, symm=1, ricb=0.35, bci=1, bco=1, Ek=1e-4, N=24, lmax=40, tau=1.0j, nev=5, tol=1e-12)


def params_spinover():
    from kore_amd.params import Params
    return Params()     # bin/parameters.py defaults = tests/spinover/params.spinover


@pytest.fixture(scope="session")
def hydro_small():
    from kore_amd import assembly
    par = params_hydro_small()
    A, B, _ = assembly.assemble(par)
    return par, A, B


@pytest.fixture(scope="session")
def spinover():
    from kore_amd import assembly
    par = params_spinover()
    A, B, _ = assembly.assemble(par)
    return par, A, B


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pencil_with_null_block(A, B, k=2):
    """(A, B') with B' = B minus k of its nonzero columns, and the unit
    vectors on those columns: OP = (A - sigma B')^{-1} B' maps them to exactly
    zero, so a Krylov start block made of them breaks down at once (the
    invariant-subspace path of the expansion)."""
    B = ss.csc_matrix(B)
    nzcols = np.nonzero(np.diff(B.indptr))[0]
    cols = nzcols[len(nzcols) // 3: len(nzcols) // 3 + k]
    keep = np.ones(B.shape[1])
    keep[cols] = 0.0
    Bn = ss.csr_matrix(B @ ss.diags(keep))
    Bn.eliminate_zeros()
    E = np.zeros((A.shape[0], k), dtype=complex)
    E[cols, np.arange(k)] = 1.0
    return ss.csr_matrix(A), Bn, E


# Reference golden of the Dormy et al. (2004) onset test, tests/dormy2004/reference.dormy04:1
# (Ek, ricb, Ra_c, m, omega_c as printed with '%.3e %.2f %.5e %d %.5e').
DORMY_GOLDEN = (2.000e-05, 0.35, 1.65404e+06, 9, -1.10162e-02)


def params_dormy(Ra_gap):
    """tests/dormy2004/params.dormy04 of the reference: thermal convection,
    differential heating, isothermal no-slip walls, m=9, Ek=2e-5, Pr=1,
    N = Ncheb(Ek)/2 = 74, lmax = 156."""
    from kore_amd.params import Params
    return Params(m=9, symm=1, ricb=0.35, bci=1, bco=1, Ek=2e-5, N=74, lmax=156, thermal=1, Prandtl=1.0,
                  heating="differential", Ra_gap=Ra_gap, bci_thermal=0, bco_thermal=0,
                  which_eigenpairs="TR", tau=0.0, nev=4, tol=1e-15)


def critical_rayleigh(growth, lo=1.6e6, hi=1.7e6):
    """Brent's method on log10(Ra_gap) with the reference driver's tolerances
    (tests/jones2000/find_Rac.py:68-88: xtol = rtol = 1e-6); growth(Ra_gap)
    returns (max Re lambda, the corresponding lambda)."""
    from scipy.optimize import brentq
    cache = {}

    def f(x):
        ra = 10 ** x
        if ra not in cache:
            cache[ra] = growth(ra)
        return cache[ra][0]
    x0 = brentq(f, np.log10(lo), np.log10(hi), xtol=1e-6, rtol=1e-6, maxiter=200)
    ra = 10 ** x0
    lam = growth(ra)[1]
    return ra, lam, len(cache)


# Reference golden of the Jones et al. (2000) onset test, tests/jones2000/reference.jones:1.
JONES_GOLDEN = (6.325e-05, 0.00, 4.66986e+06, 9, -1.93444e-02)


def params_jones(Ra_gap):
    """tests/jones2000/params.jones of the reference: full sphere (ricb = 0),
    internal heating, fixed-flux CMB, no-slip, m=9, Ek = sqrt(4/Ta) with
    Ta = 1e9, Pr = 1, N = Ncheb(Ek) = 118, lmax = 124."""
    from kore_amd.params import Params
    return Params(m=9, symm=1, ricb=0.0, bci=1, bco=1, Ek=float(np.sqrt(4 / 1e9)), N=118, lmax=124, thermal=1,
                  Prandtl=1.0, heating="internal", Ra_gap=Ra_gap, bci_thermal=1, bco_thermal=1,
                  which_eigenpairs="TR", tau=0.0, nev=5, tol=1e-15)
