"""End-to-end drop-in runs of kore_amd.solve on the device (MI355X only):
a Kore case directory in, the reference's files out, checked against the
oracle (eigenvalues rel. <= 1e-10; forced solution rel. <= 1e-10)."""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as sla

from kore_amd import assembly, solve
from kore_amd.params import Params

pytestmark = pytest.mark.gpu

PARAMS_TEMPLATE = """
import numpy as np
hydro = 1
m = {m}
symm = {symm}
ricb = 0.35
bci = 1
bco = 1
Ek = {Ek}
forcing = {forcing}
forcing_frequency = {w}
forcing_amplitude_cmb = 1.0
forcing_amplitude_icb = 0.0
magnetic = 0
thermal = 0
compositional = 0
N = {N}
lmax = {lmax}
tau = {tau}
which_eigenpairs = 'TM'
nev = {nev}
maxit = 100
tol = 1e-12
ncpus = 2
"""


def write_case(d, **kw):
    (d / "parameters.py").write_text(PARAMS_TEMPLATE.format(**kw))


@pytest.mark.parametrize("prec", ["fp64", "c64"])
def test_cli_eigen(tmp_path, monkeypatch, prec):
    from oracle import krylov_schur
    kw = dict(m=1, symm=1, Ek=1e-4, forcing=0, w=0.0, N=24, lmax=40, tau="1.0j", nev=5)
    write_case(tmp_path, **kw)
    par = Params(m=1, symm=1, Ek=1e-4, N=24, lmax=40)
    A, B, _ = assembly.assemble(par)
    assembly.save_npz_csr(str(tmp_path / "A.npz"), A)
    assembly.save_npz_csr(str(tmp_path / "B.npz"), B)
    monkeypatch.chdir(tmp_path)
    print("case: cli eigen", prec, flush=True)
    assert solve.main(["-st_type", "sinvert", "-eps_error_relative", "::ascii_info_detail",
                       "-kore_precision", prec]) == 0
    ev = np.loadtxt("eigenvalues0.dat")
    lam = ev[:, 0] + 1j * ev[:, 1]
    lu = sla.splu((A - 1j * B).tocsc())
    lo, _, _, _ = krylov_schur.krylov_schur(lambda x: lu.solve(B @ x), A.shape[0], 1j, 5, tol=1e-13)
    for l in lo[:5]:
        assert np.min(np.abs(lam - l)) / abs(l) <= 1e-10
    rf = np.loadtxt("real_flow.field")
    assert rf.shape == (2 * par.n, len(lam))
    assert os.path.exists("timing.dat") and not os.path.exists("no_conv_solution")


def test_cli_forced_libration(tmp_path, monkeypatch):
    kw = dict(m=2, symm=1, Ek=1e-4, forcing=7, w=0.8, N=24, lmax=41, tau="1.0j", nev=1)
    write_case(tmp_path, **kw)
    par = Params(m=2, symm=1, Ek=1e-4, N=24, lmax=41, forcing=7, forcing_frequency=0.8)
    A = assembly.assemble(par, which="forced")
    bf = assembly.forcing_vector(par)
    assembly.save_npz_csr(str(tmp_path / "A.npz"), A)
    assembly.save_npz_csr(str(tmp_path / "B_forced.npz"), bf)
    monkeypatch.chdir(tmp_path)
    print("case: cli forced libration", flush=True)
    assert solve.main([]) == 0
    x = np.loadtxt("real_flow.field") + 1j * np.loadtxt("imag_flow.field")
    b = np.asarray(bf.toarray()).ravel()
    xr = sla.splu(A.tocsc()).solve(b)
    assert np.linalg.norm(x - xr[:2 * par.n]) / np.linalg.norm(xr[:2 * par.n]) <= 1e-10
