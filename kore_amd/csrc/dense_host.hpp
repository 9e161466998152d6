// Small dense complex kernels of the projected eigenproblem (host, FP64).
//
// The Krylov-Schur driver keeps the projected matrix (at most a few hundred
// columns) on the host: these routines give its Schur form, reorder it so
// that the wanted Ritz values lead, and return the Ritz vectors of the
// triangular factor.  Column-major storage with explicit leading dimensions.
#pragma once
#include <complex>
#include <cstdint>

namespace kore_dense {

using zc = std::complex<double>;

// Schur form in place: T (n x n) is overwritten by upper-triangular S and Q
// (n x n) receives the unitary Schur vectors, T_in = Q S Q^H.  Householder
// reduction to Hessenberg form, then implicitly shifted single-shift QR with
// Wilkinson shifts (exceptional shifts every 10 iterations without
// deflation) and the Ahues-Tisseur deflation test.  Returns 0, or the number
// of eigenvalues left unconverged after 30 n iterations.
int schur(int n, zc* T, int64_t ldt, zc* Q, int64_t ldq);

// Reorders the Schur form (S, Q) so that the entries with select[k] != 0
// lead, in their original relative order, by swaps of neighbouring diagonal
// entries (one plane rotation each).  *m = number selected.
void schur_reorder(int n, zc* S, int64_t lds, zc* Q, int64_t ldq, const int32_t* select, int* m);

// Unit-norm eigenvectors (columns of Z, n x n, upper triangular) of the
// upper-triangular S.
void triu_eigvecs(int n, const zc* S, int64_t lds, zc* Z, int64_t ldz);

// Cholesky factor G = R^H R of a Hermitian positive-definite n x n matrix:
// R (upper) overwrites G's upper triangle (the strict lower triangle is
// zeroed).  Returns 0, or k + 1 if the k-th pivot is not positive.
int cholesky_upper(int n, zc* G, int64_t ld);

// Inverse of an upper-triangular R in place.
void triu_inverse(int n, zc* R, int64_t ld);

}  // namespace kore_dense
