// Kernel descriptors and launchers (host-visible declarations).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kore {

template <class T> struct cx;

// C = beta*C + alpha*(op(A1)*B1 + op(A2)*B2); column-major; op = identity
// (opA = 0) or conjugate transpose of a K x m stored A (opA = 1).
template <class T> struct GemmDesc {
  const cx<T>* A[2];
  const cx<T>* B[2];
  cx<T>* C;
  int64_t lda[2], ldb[2], ldc;
  int m, n, k[2];
  int opA;
  int tiles_m, tiles_n;
  T alpha, beta;
};

// in-place inverse of an n x n column-major block
template <class T> struct InvDesc {
  cx<T>* M;
  int64_t ld;
  int n;
  int* ipiv;        // n ints of device scratch (pivot row of each column)
  int* info;        // 0 or (1-based column of the first zero pivot)
  cx<T>* scratch;   // n x n block used for the final permutation
};

// Y = beta*Y + sum_p alpha_p A_p X_p  (p < 3) for skinny X (nrhs <= 16
// columns); A_p column-major m x k_p; unused products have A_p = null.
template <class T> struct SkinnyDesc {
  const cx<T>* A[3];
  const cx<T>* X[3];
  cx<T>* Y;
  int64_t lda[3], ldx[3], ldy;
  int m, k[3];
  T alpha[3];
  T beta;
};

template <class T>
hipError_t launch_gemm_batched(const GemmDesc<T>* d_descs, int ndesc, int max_tiles, hipStream_t s);
template <class T>
// nmin / nmax: smallest / largest block size in the list; blocks of 96..512
// rows take the multi-workgroup form (panel kernel + tiled MFMA update).
hipError_t launch_gj_inverse(const InvDesc<T>* d_descs, int ndesc, int nmin, int nmax, hipStream_t s);
template <class T>
hipError_t launch_skinny(const SkinnyDesc<T>* d_descs, int ndesc, int max_m, int nrhs, hipStream_t s);
// the same for one right-hand side, descriptors of at most np (= 1)
// product each and blocks up to 256 wide (wider ones stream)
template <class T>
hipError_t launch_skinny_np(const SkinnyDesc<T>* d_descs, int ndesc, int max_m, int np, hipStream_t s);
// single product with the descriptor passed by value
template <class T>
hipError_t launch_skinny1(const SkinnyDesc<T>& d, int nrhs, hipStream_t s);

// ---- sparse / vector kernels (kernels_vec.hip) -----------------------------
// CSR on the union pattern of A and B, rows and columns already in l-major
// order; va, vb are the complex values of A and B on that pattern.
struct CsrDev {
  int64_t n;
  const int32_t* indptr;
  const int32_t* indices;
  const cx<double>* va;
  const cx<double>* vb;
};

// ---- level-0 sparse couplings (kernels_sparse.hip) ------------------------
// Compact CSR of the entries outside each row's diagonal block (degree
// order): row i holds ptr[i+1] - ptr[i] entries, the first nlo[i] in the
// previous block, the rest in the next one, sorted by (global) column; val
// holds r_i (a - sigma b) of the current shift.
struct OffCsr {
  const int32_t* ptr;
  const int32_t* nlo;
  const int32_t* col;
  cx<double>* val;
};
// Column view of the same entries per (block t, side): for column c of the
// neighbouring block, entries ptr[cbase + c] .. ptr[cbase + c + 1] give the
// compact-CSR entry q and the local row (in block t).
struct OffCsc {
  const int32_t* ptr;
  const int32_t* q;
  const int32_t* row;
};
// C (m x n, ldc) = beta C + alpha sum_{p<2} S_p X_p, S_p the side[p]
// entries (0: previous block, 1: next block; -1: unused) of the rows
// row0 .. row0+m-1; X_p column-major with its first row at global column xcol0[p].
template <class T> struct SpRowDesc {
  const cx<T>* X[2];
  int64_t ldx[2];
  int64_t xcol0[2];
  int side[2];
  int xrows[2];         // rows of X_p (the neighbouring block's size, <= 512)
  cx<T>* C;
  int64_t ldc;
  int64_t row0;
  int m, n;
  T alpha, beta;
};
// C (m x n, ldc) = beta C + alpha X S, S the entries of one side of a block
// (columns: the neighbouring block's n unknowns; column segment cbase).
template <class T> struct SpColDesc {
  const cx<T>* X;
  int64_t ldx;
  cx<T>* C;
  int64_t ldc;
  int64_t cbase;
  int m, n;
  T alpha, beta;
};
hipError_t launch_offvals(const CsrDev& m, double sre, double sim, const double* scale, const int32_t* split,
                          const OffCsr& o, hipStream_t s);
// nc: columns per workgroup tile (4, 8; 16 for complex64)
template <class T>
hipError_t launch_sprow(const OffCsr& o, const SpRowDesc<T>* d_descs, int ndesc, int max_m, int max_n, int nc,
                        hipStream_t s);
template <class T>
hipError_t launch_spcol(const OffCsc& cc, const cx<double>* oval, const SpColDesc<T>* d_descs, int ndesc, int max_n,
                        hipStream_t s);
// y[:, c] -= S v[:, c] on the rows of blocks of the given parity (S: all off-block entries)
template <class T>
hipError_t launch_offsweep(const OffCsr& o, const int32_t* blk_of_row, int64_t n, int parity, const cx<T>* v,
                           int64_t ldv, cx<T>* y, int64_t ldy, int nrhs, hipStream_t s);

// Y[:, c] = beta*Y[:, c] + (ca*A + cb*B) X[:, c],  c < nrhs  (complex128)
// rows of blocks of the given parity: y -= scale * (off-diagonal blocks of A - (sre + i sim) B) v
hipError_t launch_offdiag(const CsrDev& m, double sre, double sim, const double* scale, const int32_t* split,
                          int parity, const cx<double>* v, cx<double>* y, hipStream_t s);
hipError_t launch_spmv(const CsrDev& m, double ca_re, double ca_im, double cb_re, double cb_im,
                       const cx<double>* X, int64_t ldx, cx<double>* Y, int64_t ldy, int nrhs,
                       double beta, hipStream_t s, const double* rscale = nullptr);
// s[i] = 2^-round(log2 max_j |a_ij - sigma b_ij|)
hipError_t launch_row_scale(const CsrDev& m, double sre, double sim, double* scale, hipStream_t s);
// Scatter R (A - sigma B) into the dense level-0 blocks.
template <class T>
hipError_t launch_build_blocks(const CsrDev& m, double sre, double sim, const double* scale,
                               const int32_t* blk_of_row, const int64_t* blk_off,
                               cx<T>* const* Dp, cx<T>* const* Lp, cx<T>* const* Up,
                               hipStream_t s);
// y[i] = scale[i] * x[perm[i]]   (perm may be null: identity), with type conversion
template <class Tin, class Tout>
hipError_t launch_gather_scale(const cx<Tin>* x, int64_t ldx, cx<Tout>* y, int64_t ldy, int64_t n,
                               int nrhs, const int64_t* perm, const double* scale, double mult,
                               hipStream_t s);
// y[perm[i]] = x[i]
hipError_t launch_scatter(const cx<double>* x, int64_t ldx, cx<double>* y, int64_t ldy, int64_t n,
                          int nrhs, const int64_t* perm, hipStream_t s);
// x[:, c] += mult * d[:, c]   (d complex64 or complex128)
template <class Td>
hipError_t launch_axpy(cx<double>* x, int64_t ldx, const cx<Td>* d, int64_t ldd, int64_t n, int nrhs,
                       double mult, hipStream_t s);
// partial[b*nrhs + c] = sum over rows in chunk b of |x[:, c]|^2 (b < nblocks)
template <class T>
hipError_t launch_colnorm2(const cx<T>* x, int64_t ldx, int64_t n, int nrhs, double* partial,
                           int nblocks, hipStream_t s);
// out[c] = max over chunks of per-chunk max |x| ... (max abs per column)
hipError_t launch_colmaxabs(const cx<double>* x, int64_t ldx, int64_t n, int nrhs, double* partial,
                            int nblocks, hipStream_t s);
// partial[c] (na x nb, column-major) = V[chunk c rows, :na]^H W[chunk c rows, :nb], nb <= 16
hipError_t launch_tsdot(const cx<double>* V, int64_t ldv, int na, const cx<double>* W, int64_t ldw, int nb,
                        int64_t n, int nchunk, int64_t rows_per_chunk, cx<double>* partial, hipStream_t s);
// W (n x nbw, nbw <= 2) -= V[:, :na] H (H: na x nbw, ld ldh), then per row chunk c
// partial[c] = V[chunk rows, :na]^H W[chunk rows] (gram = 0, na x nbw) or
// W[chunk rows]^H W[chunk rows] (gram = 1, nbw x nbw), column-major, as launch_tsdot
// writes them; na <= ORTHO_STEP_MAXA
constexpr int ORTHO_STEP_MAXA = 256;
hipError_t launch_ortho_step(const cx<double>* V, int64_t ldv, int na, cx<double>* W, int64_t ldw, int nbw,
                             const cx<double>* H, int ldh, int64_t n, int nchunk, int64_t rows_per_chunk, int gram,
                             cx<double>* partial, hipStream_t s);
// sum nparts consecutive arrays of len doubles: out[i] = sum_p in[p*len + i]
hipError_t launch_sum_parts(const double* in, int nparts, int len, double* out, hipStream_t s);
// partial[b] (b < nblocks) = slice sums of |va - sigma vb|^2 over nnz entries
hipError_t launch_csr_fro2(const cx<double>* va, const cx<double>* vb, int64_t nnz, double sre, double sim,
                           double* partial, int nblocks, hipStream_t s);
hipError_t launch_normalize1(const double* part, int nparts, const cx<double>* w, cx<double>* v, int64_t n,
                             cx<double>* R, cx<double>* Rinv, int* flag, hipStream_t s);
hipError_t launch_chol_inv(const cx<double>* G, int ldg, int n, cx<double>* R, cx<double>* Rinv, int* flag,
                           hipStream_t s);
// fill complex buffer with zeros (graph-capture-safe clear)
template <class T>
hipError_t launch_zero(cx<T>* p, int64_t n, hipStream_t s);
hipError_t launch_rand_cols(cx<double>* V, int64_t ldv, int64_t n, int nc, uint64_t seed, hipStream_t s);
// copy columns (complex128)
hipError_t launch_copy_cols(const cx<double>* x, int64_t ldx, cx<double>* y, int64_t ldy, int64_t n,
                            int nrhs, hipStream_t s);

}  // namespace kore
