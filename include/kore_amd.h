/*
 * kore_amd -- MI355X (gfx950) solver backend for Kore's solve path.
 *
 * C ABI: plain pointers and sizes.  Complex numbers are interleaved
 * (re, im) doubles, so numpy complex128 arrays pass without copies.  Host
 * matrices and vectors are in Kore's own (section-major) unknown ordering;
 * the library permutes them to degree-major order internally.
 *
 * Status codes: 0 OK; > 0 numerical warning; < 0 error (message from
 * kore_last_error(), which is per host thread).
 *
 * Each entry point names the reference operation it replaces.  The
 * reference (Kore, bin/solve.py) drives PETSc/SLEPc/MUMPS through
 * petsc4py/slepc4py; the equivalent calls are cited as file:line.
 */
#ifndef KORE_AMD_H
#define KORE_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KORE_OK 0
#define KORE_WARN_IR_NOT_CONVERGED 1
#define KORE_WARN_BREAKDOWN 2        /* Krylov block is numerically rank deficient */
#define KORE_WARN_NOT_CONVERGED 3    /* fewer than nev eigenpairs converged within maxit restarts */
#define KORE_ERR_ARG (-1)
#define KORE_ERR_HIP (-2)
#define KORE_ERR_OOM (-3)
#define KORE_ERR_STATE (-4)
#define KORE_ERR_SINGULAR (-5)
#define KORE_ERR_UNSUPPORTED (-6)

#define KORE_PREC_FP64 0   /* complex128 factor */
#define KORE_PREC_C64 1    /* complex64 factor (FP32 MFMA); FP64 refinement in kore_solve */

/* Form of a factorisation (kore_sys_set_factor_form).  Both hold the same
 * inverse blocks and give the same solutions to rounding. */
#define KORE_FORM_AUTO 0    /* LEAN from degree blocks of 320 unknowns, SWEEP below */
#define KORE_FORM_SWEEP 1   /* keeps E = L Dinv, F = U Dinv: two launches per level in a solve sweep */
#define KORE_FORM_LEAN 2    /* no E, F: about a quarter fewer GEMM flops, three launches per level */

#define KORE_TIMER_INV 0      /* block inversions of the factorisation        */
#define KORE_TIMER_GEMM 1     /* Schur-complement GEMMs of the factorisation  */
#define KORE_TIMER_SKINNY 2   /* solve sweeps (one event pair per kernel)     */
#define KORE_TIMER_SPMV 3     /* sparse products (B x, residuals)             */
#define KORE_TIMER_ORTHO 4    /* Krylov workspace dot / gemm                  */
#define KORE_TIMER_OTHER 5    /* block scatter, gathers, conversions          */
#define KORE_TIMER_SPARSE 6   /* level 0's products with the sparse couplings */
#define KORE_TIMER_COUNT 7

/* Which eigenvalues are wanted, in the order of bin/solve.py:99-117
 * (which_eigenpairs of bin/parameters.py): nearest the target in modulus,
 * real part or imaginary part; largest / smallest modulus, real part,
 * imaginary part. */
#define KORE_WHICH_TM 0
#define KORE_WHICH_TR 1
#define KORE_WHICH_TI 2
#define KORE_WHICH_LM 3
#define KORE_WHICH_SM 4
#define KORE_WHICH_LR 5
#define KORE_WHICH_SR 6
#define KORE_WHICH_LI 7
#define KORE_WHICH_SI 8

typedef struct kore_ctx kore_ctx;
typedef struct kore_sys kore_sys;
typedef struct kore_ws kore_ws;
typedef struct kore_sys_plan kore_sys_plan;

/* Field content and truncation of a Kore problem: the globals of
 * bin/parameters.py (m, lmax, symm, hydro, magnetic, thermal,
 * compositional) and bin/utils.py:26-52 (N1, symmB0). */
typedef struct {
  int32_t m, lmax, symm, N1;
  int32_t hydro, magnetic, thermal, compositional;
  int32_t symmB0;
} kore_layout;

typedef struct {
  int32_t iterations;      /* solve sweeps performed (1 = plain solve, no refinement) */
  int32_t converged;       /* 1 if the refinement met ir_tol (or none was asked) */
  double rel_correction;   /* max over columns of the last ||dx|| / ||x||; -1 without refinement */
} kore_solve_info;

typedef struct {
  int32_t levels;          /* reduction levels */
  int32_t nblk;            /* degree blocks */
  double flops;            /* algorithmic real flops of the factorisation actually run */
  double flops_inv;        /*   of which block inversions (8 n^3 each)     */
  double flops_gemm;       /*   of which Schur-complement GEMMs (8 m n k)  */
  double flops_sparse;     /*   of which level 0's products with the sparse couplings (8 nnz n) */
  double solve_bytes;      /* bytes one single-column solve sweep reads: factor blocks, plus, with an
                              FP64 factor, level 0's sparse couplings read from the matrix */
  int32_t solve_launches;  /* kernel launches of one single-column solve sweep */
  int64_t bytes;           /* device bytes held by the factor */
  double ms;               /* device time, HIP events on the library stream */
  int32_t form;            /* KORE_FORM_SWEEP or KORE_FORM_LEAN */
} kore_factor_info;

/* State of a system (kore_sys_query). */
typedef struct {
  int64_t sizmat;          /* unknowns */
  int32_t nblk;            /* degree blocks */
  int32_t has_b;           /* 1 if B was given */
  int32_t factored;        /* 1 if a factorisation of A - sigma B is held */
  int32_t precision;       /* its precision (KORE_PREC_*), -1 if none */
  double sre, sim;         /* its shift sigma */
} kore_sys_state;

/* Options of kore_eigs; kore_eigs_default_opts fills the defaults given here. */
typedef struct {
  int32_t nev;             /* wanted eigenpairs (-eps_nev; default 6) */
  int32_t ncv;             /* basis columns; <= 0: max(2 nev, nev + 15, 96) (-eps_ncv) */
  double tol;              /* relative tolerance on ||OP y - theta y|| / |theta| (-eps_tol; 1e-12) */
  int32_t maxit;           /* restarts (-eps_max_it; 100) */
  int32_t which;           /* KORE_WHICH_* (default TM) */
  int32_t block;           /* Krylov block size 1..16 (default 1) */
  int32_t precision;       /* factor precision (default FP64); with KORE_PREC_C64 every OP application is
                              refined in FP64, and the system is refactored in FP64 (and the iteration
                              restarted) as soon as one application needs more than fallback_sweeps */
  int32_t fallback_sweeps; /* default 4 */
  double ir_tol;           /* refinement tolerance of complex64 OP applications (1e-13) */
  int32_t ir_maxit;        /* refinement sweeps at most (12) */
  int32_t check_every;     /* early convergence checks while the basis grows: < 0 adaptive (default),
                              0 none, > 0 fixed spacing in columns */
  uint64_t seed;           /* start block (device generator) when no v0 is given */
} kore_eigs_opts;

typedef struct {
  int32_t nconv;           /* converged pairs (may exceed nev) */
  int32_t returned;        /* pairs written: min(nconv, max_pairs) */
  int32_t restarts;
  int32_t op_columns;      /* OP applications (vectors) of the run that produced the pairs */
  int32_t op_calls;        /* block applications */
  int32_t ir_max_iterations;
  int32_t precision;       /* precision of the factor the pairs were computed with */
  int32_t fallback;        /* 1 if a complex64 factor was replaced by an FP64 one */
  int32_t breakdowns;      /* invariant subspaces met (replaced by random blocks) */
  int32_t c64_op_columns;  /* OP applications made with the complex64 factor */
  int32_t c64_ir_max;      /* most refinement sweeps of one complex64 application */
  int32_t basis_columns;   /* ncv used */
  double seconds;          /* wall time of the call */
  double host_seconds;     /* of which serial host work on the projected matrix */
} kore_eigs_info;

/* ---- library / context ---------------------------------------------------- */
int32_t kore_version(void);
const char* kore_last_error(void);

/* Creates the device context (one HIP stream at default priority).
 * Replaces: PETSc/SLEPc initialisation, bin/solve.py:14-18, 31-34. */
int32_t kore_ctx_create(int32_t device, int64_t mem_budget_bytes, kore_ctx** out);
void kore_ctx_destroy(kore_ctx* ctx);
int32_t kore_ctx_sync(kore_ctx* ctx);
int64_t kore_ctx_bytes_in_use(const kore_ctx* ctx);
/* Device and pinned host blocks released by systems, factors and
 * workspaces are cached by the library (per device, shared by every context
 * on it) and reused by later allocations, so that a stream of requests does
 * not call hipMalloc / hipFree (which wait for the whole device) once warm.
 * kore_mem_trim returns every cached block of the device (and every cached
 * pinned host block) to the runtime; kore_mem_stats reports the cached
 * device bytes and the number of device allocations made so far (pool
 * misses, all devices).  Destroy systems before their context.
 * No counterpart in the reference: PETSc allocates per object and frees at
 * MatDestroy / PetscFinalize (bin/solve.py:314-319). */
int32_t kore_mem_trim(int32_t device);
int32_t kore_mem_stats(int32_t device, int64_t* cached_bytes, int64_t* device_allocs);
/* enable = 1: solve sweeps are captured as HIP graphs at the first sweep of
 * each factorisation and column count, and replayed afterwards.  Default
 * off (or KORE_AMD_GRAPHS=1 at context creation): kernel-by-kernel launches
 * measured faster on MI355X.  Per-kernel timing
 * (below) always launches kernel by kernel.  No counterpart in the
 * reference (MUMPS solves inside KSPSolve / STApply, bin/solve.py:123, 227). */
int32_t kore_ctx_set_graphs(kore_ctx* ctx, int32_t enable);
/* Per-kernel-class device timing (HIP events around each launch). */
int32_t kore_ctx_set_timing(kore_ctx* ctx, int32_t enable);
int32_t kore_ctx_timing(kore_ctx* ctx, int32_t timer_class, int64_t* count, double* total_ms);
int32_t kore_ctx_timing_reset(kore_ctx* ctx);

/* ---- layout (host only) ---------------------------------------------------- */
/* Degree-major permutation and block sizes: new index i holds Kore unknown
 * perm[i].  Follows bin/utils.py:26-44 and :174-183 (ell) and the section
 * offsets of bin/assemble.py:613-614, 730-731, 826-827, 925-926, 1022-1023.
 * blk_sizes/perm may be NULL (query); *nblk and *sizmat are always set. */
int32_t kore_layout_blocks(const kore_layout* lay, int64_t* sizmat, int32_t* nblk, int32_t* blk_sizes,
                           int32_t blk_cap, int64_t* perm);
/* Predicted device bytes of a factorisation (admission control). */
int64_t kore_factor_bytes(const kore_layout* lay, int32_t precision);

/* ---- system ------------------------------------------------------------------ */
/* Uploads A (complex CSR) and B (real or complex CSR), both sizmat x sizmat
 * in Kore's ordering, as read from A.npz / B.npz by ut.load_csr.
 * Replaces: the MPIAIJ matrix set-up of bin/solve.py:40-60 (A) and 65-86 (B).
 * B may be NULL (forced problems: b_indptr = NULL). */
int32_t kore_sys_create(kore_ctx* ctx, const kore_layout* lay, int64_t sizmat,
                        const int32_t* a_indptr, const int32_t* a_indices, const double* a_data,
                        const int32_t* b_indptr, const int32_t* b_indices, const double* b_data,
                        int32_t b_is_complex, kore_sys** out);
/* kore_sys_create in two steps: kore_sys_prepare does the host work (checks,
 * the degree-order union pattern, the level-0 coupling views, values staged
 * in pinned memory) and needs no context or device, so a serving loop can
 * run it on its assembly threads ahead of a free device slot;
 * kore_sys_create_prepared allocates the device buffers on ctx and uploads
 * (the plan is used up; destroy it afterwards, or instead, with
 * kore_sys_plan_destroy).  Replaces: the same set-up as kore_sys_create. */
int32_t kore_sys_prepare(const kore_layout* lay, int64_t sizmat,
                         const int32_t* a_indptr, const int32_t* a_indices, const double* a_data,
                         const int32_t* b_indptr, const int32_t* b_indices, const double* b_data,
                         int32_t b_is_complex, kore_sys_plan** out);
int32_t kore_sys_create_prepared(kore_ctx* ctx, kore_sys_plan* plan, kore_sys** out);
/* kore_sys_prepare for A = A0 + e A1 on one pattern (a_indptr, a_indices):
 * each value a0 + e a1 (the product, then the sum, each rounded) is formed as
 * it is staged, so a caller holding A0 and A1 never builds A.  Kore's A is
 * affine in the Ekman number on a fixed pattern (the viscous force Ek nabla^2 u,
 * bin/operators.py:136; boundary rows and the Coriolis terms do not carry
 * Ek), so a serving loop keeps A0, A1 per structure.  Replaces: the same
 * set-up as kore_sys_prepare, for a request of a structure already seen. */
int32_t kore_sys_prepare_affine(const kore_layout* lay, int64_t sizmat,
                                const int32_t* a_indptr, const int32_t* a_indices,
                                const double* a0_data, const double* a1_data, double e,
                                const int32_t* b_indptr, const int32_t* b_indices, const double* b_data,
                                int32_t b_is_complex, kore_sys_plan** out);
void kore_sys_plan_destroy(kore_sys_plan* plan);
/* New values of A (a_data: interleaved complex, one per entry of the CSR
 * given to kore_sys_create, in that order; same pattern), uploaded into the
 * resident system; B, the layout and the device buffers stay.  The system
 * must be factored again.  An entry that was a zero outside the three block
 * diagonals at creation cannot take a nonzero value (KORE_ERR_UNSUPPORTED).
 * Replaces: re-running assemble.py and re-creating the PETSc matrices for
 * every Brent step of the critical-Rayleigh search
 * (tests/jones2000/find_Rac.py:36-52), where only A's buoyancy values change. */
int32_t kore_sys_update_a(kore_sys* sys, const double* a_data);
/* Also destroys the system's workspaces that are still alive. */
void kore_sys_destroy(kore_sys* sys);
/* Size and factorisation state of a system. */
int32_t kore_sys_query(const kore_sys* sys, kore_sys_state* out);
/* Form of the system's later factorisations (KORE_FORM_*; default AUTO); a
 * held factorisation of another form is rebuilt by the next kore_factor.
 * No counterpart in the reference (MUMPS chooses its own factor layout):
 * for callers that know whether a factor serves few sweeps (a forced
 * problem: LEAN) or many (an eigensolve of small blocks: SWEEP). */
int32_t kore_sys_set_factor_form(kore_sys* sys, int32_t form);
/* Threads one call of the library's parallel host loops may use
 * (kore_sys_prepare / kore_sys_create, pinned-staging copies); n <= 0
 * restores the default: KORE_AMD_HOST_THREADS, else the process's CPU share
 * (cgroup CPU quota, else affinity mask), at most 16.  Process-wide.  No
 * counterpart in the reference: for callers that run several set-ups at once
 * (a serving loop) and must keep the process within its CPU quota. */
int32_t kore_set_host_threads(int32_t n);
/* The current per-call host thread count, and the process's CPU share. */
int32_t kore_host_threads(void);
int32_t kore_cpu_share(void);
/* ||A - sigma B||_F of the stored (unscaled) matrices. */
int32_t kore_sys_norm(kore_sys* sys, double sre, double sim, double* out);

/* Factors A - sigma*B by block cyclic reduction over the degree blocks.
 * Replaces: the numeric LU of ST sinvert inside E.solve() (bin/solve.py:123)
 * and of PC lu inside K.solve() (bin/solve.py:220-227); MUMPS. */
int32_t kore_factor(kore_sys* sys, double sre, double sim, int32_t precision, kore_factor_info* info);
/* Drops the factorisations (both precisions: kore_factor keeps the plan of
 * the other precision when the precision changes); their device memory
 * returns to the library's cache. */
int32_t kore_factor_release(kore_sys* sys);

/* Solves (A - sigma B) x = b for nrhs columns (host arrays, Kore order,
 * column-major with leading dimensions ldb/ldx).  With a complex64 factor the
 * solution is refined in FP64 (residual b - (A - sigma B) x by FP64 SpMV,
 * correction by the factor) until the last correction ||dx|| <= ir_tol ||x||,
 * the corrections stop shrinking, or 1 + ir_maxit sweeps; with an FP64 factor
 * ir_maxit extra sweeps are optional (0 = plain solve).
 * Replaces: KSPSolve of bin/solve.py:227 (forced problems). */
int32_t kore_solve(kore_sys* sys, int32_t nrhs, const double* b, int64_t ldb, double* x, int64_t ldx,
                   double ir_tol, int32_t ir_maxit, kore_solve_info* info);

/* Solves (A - sigma B) x = b for one right-hand side (host arrays, Kore
 * order) by restarted FP64 GMRES right-preconditioned with the factor:
 * every Krylov step is one factor solve and one FP64 sparse product, so a
 * complex64 factor acts as a preconditioner and the answer carries FP64
 * accuracy where plain refinement would stall.  Stops when
 * ||b - (A - sigma B) x|| <= tol (||A - sigma B||_F ||x|| + ||b||), or after
 * maxit factor solves (KORE_WARN_IR_NOT_CONVERGED); restart <= 30.
 * info: iterations = factor solves, converged, rel_correction = final
 * normwise backward error.
 * Replaces: KSPSolve of bin/solve.py:227 (forced problems), with GMRES
 * (-ksp_type gmres, tools/ramp.sh) preconditioned by the factor. */
int32_t kore_solve_gmres(kore_sys* sys, const double* b, double* x, double tol, int32_t maxit, int32_t restart,
                         kore_solve_info* info);

/* Residual norms of nrhs candidate solutions from the stored (unscaled)
 * matrices, on the device: r_c = (A - (sigma + lam_c) B) x_c - b_c and
 * rnorm[c] = ||r_c||_2.  x and b are host arrays in Kore order (column-major,
 * ldx / ldb); lam (nrhs interleaved complex) and b may be NULL (zero).
 * Replaces: the host-side true residuals behind -eps_true_residual and the
 * error table of bin/solve.py:125-129, and the check of a forced solution. */
int32_t kore_residual(kore_sys* sys, int32_t nrhs, const double* x, int64_t ldx, const double* lam,
                      const double* b, int64_t ldb, double sre, double sim, double* rnorm);
/* The same residual norms, allocating nothing: the staging, gathered and
 * residual vectors and the norms' partial sums live in a static 64 MiB device
 * array of the library (each device holds its own copy until the process
 * ends) instead of the system's buffers or newly allocated memory; no
 * hipMalloc / hipFree or allocating library routine is called.  A call needs
 * 3 * sizmat * 16 bytes plus 8 KiB of that array per column (columns are taken
 * in as large chunks as fit); a larger system is rejected with KORE_ERR_ARG.
 * Calls that overlap on one device share the array and race.
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * device allocation). */
int32_t kore_residual_noalloc(kore_sys* sys, int32_t nrhs, const double* x, int64_t ldx, const double* lam,
                              const double* b, int64_t ldb, double sre, double sim, double* rnorm);

/* ---- stream-ordered (asynchronous) variants ------------------------------- */
/* For a serving loop that keeps its vectors in device memory and orders its
 * work on its own HIP streams.  stream is a hipStream_t (NULL: the legacy
 * default stream).  The work runs on the context's stream, ordered after
 * everything queued on `stream` before the call, and `stream` waits for it
 * (HIP events both ways): the call returns without waiting for the device.
 * Calls on one system must not overlap with calls on it from other threads.
 * kore_factor_async: as kore_factor; the first factorisation of a system in
 * a precision (plan set-up, and the tile tuner's measurement for a new
 * block-size class) blocks until it is built.  Zero pivots are reported by
 * kore_factor_check, which waits for the context's stream.
 * kore_solve_async: d_b, d_x are device arrays in Kore order (column-major,
 * ldb, ldx >= sizmat); nrhs columns; after the factor solve, ir_sweeps FP64
 * refinement sweeps with no convergence test (use 2-4 with a complex64 factor,
 * 0 with an FP64 one).
 * Replaces: the numeric factorisation and KSPSolve of bin/solve.py:220-227 as
 * queued work of a serving loop (the reference runs one blocking solve per
 * process). */
int32_t kore_factor_async(kore_sys* sys, double sre, double sim, int32_t precision, void* stream);
int32_t kore_factor_check(kore_sys* sys);
int32_t kore_solve_async(kore_sys* sys, int32_t nrhs, const double* d_b, int64_t ldb, double* d_x, int64_t ldx,
                         int32_t ir_sweeps, void* stream);

/* ---- eigensolver ------------------------------------------------------------- */
/* The nev eigenpairs of A x = lambda B x nearest the target tau (in the sense
 * of opts->which) by shift-and-invert block Krylov-Schur on the device.  The
 * system is factored at tau (opts->precision) unless it already holds that
 * factorisation.  Writes min(nconv, max_pairs) pairs in wanted order: lam
 * (interleaved complex), X (sizmat x max_pairs complex, column-major, ldx,
 * Kore order, unit 2-norm columns; may be NULL), errest (relative error
 * estimates; may be NULL).  v0: optional start vector (sizmat complex, Kore
 * order).  Returns KORE_WARN_NOT_CONVERGED if fewer than nev converged.
 * Replaces: E.solve() and the extraction loop of bin/solve.py:91-158 (EPS
 * GNHEP, Krylov-Schur, ST sinvert, EPSGetEigenpair). */
void kore_eigs_default_opts(kore_eigs_opts* opts);
int32_t kore_eigs(kore_sys* sys, double tau_re, double tau_im, const kore_eigs_opts* opts, const double* v0,
                  int32_t max_pairs, double* lam, double* X, int64_t ldx, double* errest, kore_eigs_info* info);
/* Host only: the projected-problem step of Krylov-Schur (SLEPc's DS object,
 * DSSolve / DSSort inside E.solve(), bin/solve.py:123).  kore_schur: Schur
 * form of the complex n x n matrix T in place (upper triangular) with its
 * unitary Schur vectors in Q, T_in = Q T Q^H.  kore_schur_reorder: moves the
 * diagonal entries with select[k] != 0 to the top of (S, Q), keeping their
 * relative order; *m = number moved. */
int32_t kore_schur(int32_t n, double* T, int64_t ldt, double* Q, int64_t ldq);
int32_t kore_schur_reorder(int32_t n, double* S, int64_t lds, double* Q, int64_t ldq, const int32_t* select,
                           int32_t* m);

/* ---- Krylov workspace (device, degree-major order) ------------------------- */
/* A block of ncol device vectors of length sizmat (complex128).
 * Replaces: SLEPc's BV basis object inside EPS Krylov-Schur. */
int32_t kore_ws_create(kore_sys* sys, int32_t ncol, kore_ws** out);
/* Must precede kore_sys_destroy of its system (which destroys it otherwise). */
void kore_ws_destroy(kore_ws* ws);
/* Host <-> device columns, host side in Kore order.
 * Replaces: VecSetValues / Scatter.toZero gathers (bin/solve.py:143-158). */
int32_t kore_ws_set(kore_ws* ws, int32_t c0, int32_t nc, const double* host, int64_t ld);
int32_t kore_ws_get(kore_ws* ws, int32_t c0, int32_t nc, double* host, int64_t ld);
/* V[:, dst+j] = (A - sigma B)^{-1} B V[:, src+j], j < nc  (the shift-and-invert
 * operator).  Replaces: STApply of ST sinvert (MatMult(B) + MUMPS solve). */
int32_t kore_ws_apply_op(kore_ws* ws, int32_t src, int32_t dst, int32_t nc, double ir_tol, int32_t ir_maxit,
                         kore_solve_info* info);
/* V[:, dst+j] = (ca A + cb B) V[:, src+j].  Replaces: MatMult (true residuals). */
int32_t kore_ws_spmv(kore_ws* ws, int32_t src, int32_t dst, int32_t nc, double ca_re, double ca_im,
                     double cb_re, double cb_im);
/* H (na x nb, column-major, host) = V[:, a0:a0+na]^H V[:, b0:b0+nb].
 * Replaces: BVDot / BVDotVec of the Gram-Schmidt orthogonalisation. */
int32_t kore_ws_dot(kore_ws* ws, int32_t a0, int32_t na, int32_t b0, int32_t nb, double* H);
/* V[:, d0:d0+nd] = beta V[:, d0:d0+nd] + alpha V[:, s0:s0+ns] Q  (Q host ns x nd,
 * column-major); destination and source ranges must not overlap.
 * Replaces: BVMult / BVMultInPlace (orthogonalisation update, restart). */
int32_t kore_ws_gemm(kore_ws* ws, int32_t d0, int32_t nd, int32_t s0, int32_t ns, const double* Q,
                     double alpha, double beta);
/* V[:, dst+j] = V[:, src+j]. */
int32_t kore_ws_copy(kore_ws* ws, int32_t src, int32_t dst, int32_t nc);
/* V[:, c0:c0+nc] = pseudo-random entries (re, im uniform in [-1, 1)), a
 * function of seed, column and row only, generated on the device.
 * Replaces: SLEPc's random initial vector of EPS (BVSetRandomColumn), which
 * Krylov-Schur uses when no initial space is given (bin/solve.py:123). */
int32_t kore_ws_random(kore_ws* ws, int32_t c0, int32_t nc, uint64_t seed);
/* Block Arnoldi expansion of the Krylov-Schur basis from column j0 to j1 in
 * blocks of bs: for j = j0, j0+bs, ...: W = OP V[:, j:j+bs] (in the scratch
 * block V[:, cw:cw+bs], cw >= j1+bs), two passes of classical Gram-Schmidt of
 * W against V[:, 0:j+bs], two passes of Cholesky QR giving V[:, j+bs:j+2bs] R
 * = W (one pass when bs = 1: the normalised column).  H (host, column-major, ldh >= j1+bs) receives the projected columns:
 * H[0:j+bs, j:j+bs] = Gram-Schmidt coefficients, H[j+bs:j+2bs, j:j+bs] = R.
 * ir_tol / ir_maxit as in kore_ws_apply_op; ir_stop > 0 ends the expansion
 * when one block needs more than ir_stop refinement sweeps or does not
 * converge (KORE_WARN_IR_NOT_CONVERGED).  KORE_WARN_BREAKDOWN: the Cholesky
 * factor of block j_done broke down (its Gram-Schmidt coefficients are in H,
 * the orthogonalised block in V[:, cw:cw+bs]).  *j_done = first column not
 * completed (j1 on success, less after a stop request, see below).
 * Replaces: the Arnoldi expansion inside SLEPc's Krylov-Schur (STApply +
 * BVOrthogonalize per step) under E.solve() (bin/solve.py:123). */
int32_t kore_ws_expand(kore_ws* ws, int32_t j0, int32_t j1, int32_t bs, int32_t cw, double ir_tol,
                       int32_t ir_maxit, int32_t ir_stop, double* H, int64_t ldh, int32_t* j_done,
                       kore_solve_info* info);
/* j_end >= 0: a kore_ws_expand running on another thread (or started later)
 * starts no block at a column >= j_end: it ends at j_end, or after the block in
 * flight if it is already past (KORE_OK, *j_done = first column not
 * completed); j_end < 0 clears the request, which stays set until then (clear
 * it before each expansion).  Lets the host end an expansion when a
 * convergence check made beside it has succeeded (j_end = 0), or at the column
 * where the residual trend of that check predicts convergence.  No
 * counterpart in the reference: SLEPc's Krylov-Schur checks convergence only
 * at the end of each expansion to ncv (bin/solve.py:123). */
int32_t kore_ws_stop_expand(kore_ws* ws, int32_t j_end);

#ifdef __cplusplus
}
#endif
#endif
