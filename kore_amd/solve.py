#!/usr/bin/env python3
"""Drop-in for Kore's ``bin/solve.py`` on one MI355X.

Usage (in a Kore case directory holding parameters.py and A.npz, B.npz or
B_forced.npz, as written by Kore's assemble.py):

    python -m kore_amd.solve -st_type sinvert -eps_error_relative ::ascii_info_detail

Contract kept from the reference (bin/solve.py):

* argv is read as the PETSc/SLEPc options database and overrides
  parameters.py (solve.py:16, 120, 224): -eps_nev, -eps_ncv, -eps_tol,
  -eps_max_it, -eps_target, -eps_error_relative, -eps_true_residual,
  -st_type (sinvert only; cayley is rejected), -nbl; solver-package options
  (-st_pc_factor_mat_solver_type, -pc_factor_mat_solver_type,
  -mat_mumps_icntl_*, -mat_superlu_dist_*, -ksp_*, -pc_type, -eps_balance,
  -eps_monitor*, -eps_converged_reason, -eps_conv_rel) are accepted and
  noted as not applicable;
* eigenproblems (forcing = 0): eigenvalues0.dat (nconv x [Re, Im]), one
  column per eigenpair in real_/imag_{flow,magnetic,temperature,composition}.field
  (solve.py:160-199, 275-306); no_conv_solution when nothing converged;
* forced problems: the same field files with one column (solve.py:207-268),
  "Solver crashed, got nan's!" and no files on NaN/Inf;
* timing.dat is appended with the wall time from start to files written
  (solve.py:36-37, 308-311); the exit code is always 0 (solve.py:314).

Extra options of this backend: -kore_precision fp64|c64 (c64: complex64
factor + FP64 refinement; forced problems by FP64 GMRES preconditioned with
the complex64 factor), -kore_block <bs> (Krylov block size),
-kore_device <id>, -kore_params <path to parameters.py>.
Only rank 0 of an MPI launch works; other ranks exit at once.
"""
from __future__ import annotations

import os
import sys
from timeit import default_timer as timer

import numpy as np

NOT_APPLICABLE = ("st_pc_factor_mat_solver_type", "pc_factor_mat_solver_type", "mat_mumps_icntl_",
                  "mat_superlu_dist_", "ksp_", "pc_type", "eps_balance", "eps_monitor", "eps_converged_reason",
                  "eps_conv_rel", "st_ksp_type", "st_pc_type", "mat_type", "log_view", "eps_view")


def parse_petsc_options(argv):
    """-name [value] pairs; a token starting with '-' that is not a number is a new option."""
    def is_number(t):
        try:
            parse_scalar(t)
            return True
        except ValueError:
            return False

    out = {}
    i = 0
    toks = list(argv)
    while i < len(toks):
        t = toks[i]
        if t.startswith("-") and not is_number(t):
            name = t.lstrip("-")
            val = None
            if i + 1 < len(toks) and (not toks[i + 1].startswith("-") or is_number(toks[i + 1])):
                val = toks[i + 1]
                i += 1
            out[name] = val
        i += 1
    return out


def parse_scalar(s):
    """PETSc scalar text: '0.5', '-1e-3+0.5i', '0+1.0j', '1.0i'."""
    t = s.strip().replace("I", "i").replace("j", "i")
    if t.endswith("i"):
        return complex(t[:-1] + "j") if any(c.isdigit() for c in t[:-1]) and t[:-1] not in ("+", "-", "") \
            else complex(t[:-1] + "1j")
    return complex(float(t))


def mpi_rank():
    for k in ("OMPI_COMM_WORLD_RANK", "PMI_RANK", "PMIX_RANK", "MPI_RANK", "SLURM_PROCID"):
        if k in os.environ:
            try:
                return int(os.environ[k])
            except ValueError:
                pass
    return 0


def split_fields(V, par):
    """Field blocks of a solution matrix (rows = unknowns), solve.py:169-190, 243-261."""
    n = par.n
    out = {}
    if par.hydro == 1:
        out["flow"] = V[0:2 * n]
    if par.magnetic == 1:
        off = 2 * n * par.hydro
        out["magnetic"] = V[off:off + 2 * n]
    if par.thermal == 1:
        off = 2 * n + par.magnetic * 2 * n
        out["temperature"] = V[off:off + n]
    if par.compositional == 1:
        off = 2 * n + par.magnetic * 2 * n + par.thermal * n
        out["composition"] = V[off:off + n]
    return out


def write_fields(fields):
    for name, blk in fields.items():
        with open("real_%s.field" % name, "wb") as f:
            np.savetxt(f, np.real(blk))
        with open("imag_%s.field" % name, "wb") as f:
            np.savetxt(f, np.imag(blk))


def write_eigenvalues(lam):
    eigval = np.hstack([np.real(lam).reshape(-1, 1), np.imag(lam).reshape(-1, 1)])
    with open("eigenvalues0.dat", "wb") as f:
        np.savetxt(f, eigval)


def relative_errors(A, B, lam, X):
    """SLEPc's relative error of each pair, ||A x - lambda B x|| / (|lambda| ||x||),
    the quantity -eps_error_relative prints for bin/solve.py:125-129."""
    out = []
    for i in range(len(lam)):
        x = X[:, i]
        res = A @ x - lam[i] * (B @ x)
        out.append(float(np.linalg.norm(res) / (abs(lam[i]) * np.linalg.norm(x) + 1e-300)))
    return out


def error_table(lam, rel):
    lines = [" ---------------------- --------------------",
             "            k             ||Ax-kBx||/||kx||",
             " ---------------------- --------------------"]
    for l, r in zip(lam, rel):
        lines.append("  %10.6f%+10.6fi      %12g" % (l.real, l.imag, r))
    lines.append(" ---------------------- --------------------")
    return "\n".join(lines)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if mpi_rank() != 0:
        return 0
    from .assembly import load_csr
    from .params import load_parameters
    from . import backend, eigs as keigs

    tic = timer()
    opts = parse_petsc_options(argv)
    par = load_parameters(opts.get("kore_params") or os.path.join(os.getcwd(), "parameters.py"))
    for name in opts:
        if any(name.startswith(p) for p in NOT_APPLICABLE):
            print("kore_amd: option -%s does not apply to the MI355X backend (ignored)" % name)
    st = (opts.get("st_type") or "sinvert").lower()
    if st != "sinvert":
        raise SystemExit("kore_amd: -st_type %s is not implemented (only sinvert)" % st)
    prec = {"fp64": backend.PREC_FP64, "c64": backend.PREC_C64}[(opts.get("kore_precision") or "fp64").lower()]
    block = int(opts.get("kore_block") or 1)
    device = int(opts.get("kore_device") or 0)

    A = load_csr("A.npz")
    lay = backend.layout_from_params(par)
    ctx = backend.Context(device)
    success = 0
    if par.forcing == 0:
        B = load_csr("B.npz")
        if "nbl" in opts and opts["nbl"] is not None and int(opts["nbl"]) != B.shape[0]:
            raise SystemExit("kore_amd: -nbl %s differs from B's size %d" % (opts["nbl"], B.shape[0]))
        nev = int(opts.get("eps_nev") or par.nev)
        ncv = int(opts["eps_ncv"]) if opts.get("eps_ncv") else par.ncv
        tol = float(opts.get("eps_tol") or par.tol)
        maxit = int(opts.get("eps_max_it") or par.maxit)
        tau = parse_scalar(opts["eps_target"]) if opts.get("eps_target") else par.tau
        S = backend.System(ctx, lay, A, B)
        S.factor(tau, prec)
        r = keigs.eigs(S, tau, nev=nev, ncv=ncv, tol=tol, maxit=maxit, which=par.which_eigenpairs, block=block)
        if r.nconv > 0:
            write_eigenvalues(r.lam)
            write_fields(split_fields(r.X, par))
            if "eps_error_relative" in opts or "eps_true_residual" in opts:
                print(error_table(r.lam, relative_errors(A, B, r.lam, r.X)))
            success = r.nconv
        else:
            print("No converged solution found")
            np.savetxt("no_conv_solution", [0])
        S.close()
    else:
        b0 = load_csr("B_forced.npz")
        b = np.asarray(b0[:, 0].toarray()).ravel().astype(complex)
        S = backend.System(ctx, lay, A, None)
        S.factor(0.0, prec)
        if prec == backend.PREC_FP64:
            x, info = S.solve(b, ir_tol=1e-15, ir_maxit=1)
        else:
            # complex64 factor as the preconditioner of FP64 GMRES; FP64 refactor if it stalls
            x, info = S.solve_gmres(b, tol=1e-15, maxit=60)
            if not info.converged:
                S.factor(0.0, backend.PREC_FP64)
                x, info = S.solve(b, ir_tol=1e-15, ir_maxit=1)
        S.close()
        fields = split_fields(x.reshape(-1, 1), par)
        flow = fields.get("flow", x.reshape(-1, 1))
        if np.sum([np.isnan(flow.real), np.isnan(flow.imag), np.isinf(flow.real), np.isinf(flow.imag)]) > 0:
            success = 0
            print("Solver crashed, got nan's!")
        else:
            success = 1
            print("Solution(s) computed")
            write_fields(fields)
    ctx.close()
    toc2 = timer()
    print("Solve done in", toc2 - tic, "seconds")
    with open("timing.dat", "ab") as f:
        np.savetxt(f, np.array([toc2 - tic]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
