"""CPU tests of the drop-in solve.py contract (argv parsing, field split,
writers, file formats).  The GPU end-to-end run is in tests/test_gpu_cli.py."""
import io
import os

import numpy as np
import pytest
import scipy.sparse as ss

from kore_amd import assembly, solve
from kore_amd.params import Params, load_parameters


def test_parse_petsc_options():
    o = solve.parse_petsc_options("-st_type sinvert -eps_error_relative ::ascii_info_detail -eps_nev 7 "
                                  "-eps_target -1e-3+0.5i -mat_mumps_icntl_14 1000 -eps_true_residual".split())
    assert o["st_type"] == "sinvert"
    assert o["eps_error_relative"] == "::ascii_info_detail"
    assert o["eps_nev"] == "7"
    assert solve.parse_scalar(o["eps_target"]) == complex(-1e-3, 0.5)
    assert o["mat_mumps_icntl_14"] == "1000"
    assert o["eps_true_residual"] is None


@pytest.mark.parametrize("txt,val", [("0.5", 0.5), ("1.0i", 1j), ("0+1.0i", 1j), ("-2e-3-0.7i", -2e-3 - 0.7j),
                                     ("3", 3.0)])
def test_parse_scalar(txt, val):
    assert solve.parse_scalar(txt) == val


def test_field_split_offsets():
    """solve.py:169-190: flow 2n, magnetic 2n after it, temperature n, composition n."""
    par = Params(m=2, symm=1, N=8, lmax=5, magnetic=1, thermal=1, compositional=1)
    n = par.n
    V = np.arange(par.sizmat, dtype=float).reshape(-1, 1) * (1 + 1j)
    f = solve.split_fields(V, par)
    assert f["flow"].shape[0] == 2 * n and f["flow"][0, 0] == 0
    assert f["magnetic"][0, 0] == 2 * n * (1 + 1j)
    assert f["temperature"][0, 0] == 4 * n * (1 + 1j)
    assert f["composition"][0, 0] == 5 * n * (1 + 1j)


def test_writers_match_numpy_savetxt(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    lam = np.array([1 + 2j, -0.5 + 0.25j])
    solve.write_eigenvalues(lam)
    ref = io.BytesIO()
    np.savetxt(ref, np.hstack([lam.real.reshape(-1, 1), lam.imag.reshape(-1, 1)]))
    assert open("eigenvalues0.dat", "rb").read() == ref.getvalue()
    V = np.array([[1 + 2j, 3 - 1j], [0.5j, -2.0]])
    solve.write_fields({"flow": V})
    assert np.array_equal(np.loadtxt("real_flow.field"), V.real)
    assert np.array_equal(np.loadtxt("imag_flow.field"), V.imag)
    # '%.18e' default format, space separated
    assert open("real_flow.field").readline().split()[0] == "%.18e" % 1.0


def test_npz_round_trip(tmp_path, hydro_small):
    par, A, B = hydro_small
    p = tmp_path / "A.npz"
    assembly.save_npz_csr(str(p), A)
    A2 = assembly.load_csr(str(p))
    assert (A2 != A).nnz == 0
    z = np.load(str(p))
    assert set(z.files) == {"data", "indices", "indptr", "shape"}


def test_load_parameters_reads_kore_file(tmp_path):
    src = """
import numpy as np
hydro = 1
m = 2
symm = 1
ricb = 0.35
bci = 1
bco = 1
Ek = 1e-4
forcing = 0
magnetic = 0
thermal = 0
compositional = 0
N = 24
lmax = 41
tau = 0.1 + 1.0j
which_eigenpairs = 'TM'
nev = 4
maxit = 30
tol = 1e-11
ncpus = 2
unrelated_knob = 'ignored'
"""
    f = tmp_path / "parameters.py"
    f.write_text(src)
    par = load_parameters(str(f))
    assert (par.m, par.lmax, par.N, par.nev, par.maxit) == (2, 41, 24, 4, 30)
    assert par.tau == 0.1 + 1.0j and par.tol == 1e-11
    assert par.sizmat == 2 * 24 * (41 - 2 + 1) // 2


def test_rank_detection(monkeypatch):
    monkeypatch.setenv("OMPI_COMM_WORLD_RANK", "3")
    assert solve.mpi_rank() == 3
    monkeypatch.delenv("OMPI_COMM_WORLD_RANK")
    monkeypatch.setenv("PMI_RANK", "0")
    assert solve.mpi_rank() == 0


def test_nonzero_rank_exits_immediately(monkeypatch, tmp_path):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("PMI_RANK", "1")
    assert solve.main([]) == 0
    assert not os.path.exists("timing.dat")


def test_relative_error_is_slepc_definition():
    """-eps_error_relative prints ||A x - lambda B x|| / (|lambda| ||x||)
    (SLEPc's EPS_ERROR_RELATIVE), not a norm of B x in the denominator."""
    import scipy.sparse as ss
    from kore_amd.solve import relative_errors
    A = ss.csr_matrix(np.diag([2.0 + 0j, 3.0]))
    B = ss.csr_matrix(np.diag([1.0, 0.5]))
    x = np.array([[3.0 + 0j], [4.0]])            # ||x|| = 5
    lam = np.array([2.0 + 0j])
    # A x - 2 B x = (6 - 6, 12 - 4) = (0, 8): 8 / (2 * 5)
    assert abs(relative_errors(A, B, lam, x)[0] - 0.8) < 1e-15
