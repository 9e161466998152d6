"""Post-processing integrals (bin/utils4pp.py via bin/spin_doctor.py) against
golden vectors written by the reference's own modules
(tests/golden/make_pp_fixtures.py), and the power balance of computed
eigenvectors (the residuals vanish for exact eigenpairs)."""
import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as sla

from conftest import GOLDEN

CASES = ["hydro_m1_N24_L40", "spinover", "thermal_m9_N24_L32", "jones_m9_N24_L32"]


def load(name):
    d = np.load(os.path.join(GOLDEN, "pp_%s.npz" % name))
    return d, json.loads(str(d["meta"]))


def params_of(meta):
    from kore_amd.params import Params
    kw = dict(m=meta["m"], symm=meta["symm"], ricb=meta["ricb"], Ek=meta["Ek"], N=meta["N"], lmax=meta["lmax"],
              thermal=meta["thermal"], Prandtl=meta["Prandtl"], heating=meta["heating"])
    return Params(**kw)


@pytest.mark.parametrize("name", CASES)
def test_flow_integrals_match_reference(name):
    from kore_amd import postprocess as pp
    from kore_amd.params import ell
    d, meta = load(name)
    par = params_of(meta)
    assert par.N1 == meta["N1"] and par.n == meta["n"]
    f = pp.flow_integrals(par, d["u"])
    ref = d["udgn"]                       # rows: every degree in ascending order; KE, Dkin0, Dint0, Wlor0, Wthm0, Wcmp0
    _, _, ll = ell(par.m, par.lmax, par.symm)
    order = np.argsort(f["l"])
    assert np.array_equal(f["l"][order], ll)
    for col, key in ((0, "KE"), (1, "Dkin0"), (2, "Dint0")):
        got = f[key][order]
        scale = np.abs(ref[:, col]).max()
        assert np.abs(got - ref[:, col]).max() <= 1e-12 * scale, key


@pytest.mark.parametrize("name", ["thermal_m9_N24_L32", "jones_m9_N24_L32"])
def test_thermal_integrals_match_reference(name):
    from kore_amd import postprocess as pp
    d, meta = load(name)
    par = params_of(meta)
    t = pp.thermal_integrals(par, d["t"], d["u"])
    ref = d["tdgn"]                       # rows: poloidal degrees; TE, Dthm0, Wadv
    for col, key in ((0, "TE"), (1, "Dthm0"), (2, "Wadv")):
        scale = np.abs(ref[:, col]).max()
        assert np.abs(t[key] - ref[:, col]).max() <= 1e-12 * scale, key
    # buoyancy power of the flow equation (utils4pp.py:363-370 in flow_worker)
    f_ref = d["udgn"][:, 4]
    from kore_amd.params import ell
    lp, _, ll = ell(par.m, par.lmax, par.symm)
    idx = np.searchsorted(ll, lp)
    assert np.abs(t["Wthm0"] - f_ref[idx]).max() <= 1e-12 * np.abs(f_ref).max()
    assert np.abs(np.delete(f_ref, idx)).max() == 0.0


def test_power_balance_of_eigenvectors(spinover):
    """For eigenpairs of the spin-over pencil (oracle: SuperLU + Krylov-Schur)
    the balance residuals are small -- resid1 (energy) ~1e-6, resid0
    (dissipation) ~1e-3, the accuracy of the reference's N-point
    Gauss-Chebyshev quadrature of these boundary-layer integrands -- while a
    vector perturbed by 1e-4 of white noise breaks both (resid ~1): the
    diagnostics see the physics."""
    from kore_amd import postprocess as pp
    from oracle import krylov_schur
    par, A, B = spinover
    lu = sla.splu((A - par.tau * B).tocsc())
    lam, X, _, _ = krylov_schur.krylov_schur(lambda x: lu.solve(B @ x), A.shape[0], par.tau, 3, tol=1e-13)
    for i in range(3):
        u, _ = pp.field_sections(par, X[:, i])
        dg = pp.diagnose(par, lam[i], usol=u)
        assert dg.resid0 < 2e-3 and dg.resid1 < 1e-5, (lam[i], dg.resid0, dg.resid1)
    rng = np.random.default_rng(0)
    xb = X[:, 0] + 1e-4 * np.linalg.norm(X[:, 0]) * rng.standard_normal(A.shape[0]) / np.sqrt(A.shape[0])
    u, _ = pp.field_sections(par, xb)
    bad = pp.diagnose(par, lam[0], usol=u)
    assert bad.resid0 > 0.1 and bad.resid1 > 0.1
