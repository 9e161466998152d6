// Dense block kernels of the block-cyclic-reduction factorisation (gfx950).
//
//   gemm_batched   C = beta*C + alpha*(A1*B1 + A2*B2) for a list of complex
//                  blocks, on MFMA (f64 16x16x4 or f32 16x16x4).
//   gj_inverse     in-place inverse of a list of complex blocks by blocked
//                  Gauss-Jordan elimination with partial (row) pivoting,
//                  one workgroup per block, panel staged in LDS, the rank-w
//                  updates on MFMA.
//
// All blocks are column-major with interleaved complex storage.
#include "kore_device.hpp"
#include "kore_kernels.hpp"

namespace kore {

// ---------------------------------------------------------------------------
// batched complex GEMM
// ---------------------------------------------------------------------------
// Workgroup = 4 waves, output tile 64x64, each wave a 32x32 quadrant made of
// 2x2 MFMA 16x16 tiles; K staged through LDS in slabs of 16.
constexpr int GT_M = 64, GT_N = 64, GT_K = 16;
constexpr int GT_PADB = 1;

template <class T>
__global__ __launch_bounds__(256) void gemm_batched_kernel(const GemmDesc<T>* __restrict__ descs) {
  using A4 = typename acc4<T>::type;
  const GemmDesc<T> d = descs[blockIdx.y];
  const int ntiles = d.tiles_m * d.tiles_n;
  if ((int)blockIdx.x >= ntiles) return;

  __shared__ cx<T> sA[GT_K][GT_M];
  __shared__ cx<T> sB[GT_N][GT_K + GT_PADB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int l16 = lane & 15, lq = lane >> 4;

  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int i0 = (tile % d.tiles_m) * GT_M;
    const int j0 = (tile / d.tiles_m) * GT_N;
    A4 cre[2][2], cim[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        cre[a][b] = A4{0, 0, 0, 0};
        cim[a][b] = A4{0, 0, 0, 0};
      }

    for (int p = 0; p < 2; ++p) {
      const cx<T>* __restrict__ A = d.A[p];
      const cx<T>* __restrict__ B = d.B[p];
      const int K = d.k[p];
      if (A == nullptr || K <= 0) continue;
      const int64_t lda = d.lda[p], ldb = d.ldb[p];
      for (int k0 = 0; k0 < K; k0 += GT_K) {
        __syncthreads();
        // A slab -> sA[k][i]
        if (d.opA == 0) {
          const int kk = tid >> 4, ii = (tid & 15) * 4;
          const int gk = k0 + kk;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int gi = i0 + ii + q;
            cx<T> v = czero<T>();
            if (gk < K && gi < d.m) v = A[(int64_t)gk * lda + gi];
            sA[kk][ii + q] = v;
          }
        } else {  // A_eff(i,k) = conj(A_store(k,i)), A_store column-major K x m
          const int ii = tid >> 2, kk = (tid & 3) * 4;
          const int gi = i0 + ii;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int gk = k0 + kk + q;
            cx<T> v = czero<T>();
            if (gk < K && gi < d.m) {
              v = A[(int64_t)gi * lda + gk];
              v.im = -v.im;
            }
            sA[kk + q][ii] = v;
          }
        }
        // B slab -> sB[j][k]
        {
          const int jj = tid >> 2, kk = (tid & 3) * 4;
          const int gj = j0 + jj;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int gk = k0 + kk + q;
            cx<T> v = czero<T>();
            if (gk < K && gj < d.n) v = B[(int64_t)gj * ldb + gk];
            sB[jj][kk + q] = v;
          }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GT_K; kk += 4) {
          cx<T> a[2], b[2];
#pragma unroll
          for (int s = 0; s < 2; ++s) {
            a[s] = sA[kk + lq][wr + 16 * s + l16];
            b[s] = sB[wc + 16 * s + l16][kk + lq];
          }
#pragma unroll
          for (int rs = 0; rs < 2; ++rs)
#pragma unroll
            for (int cs = 0; cs < 2; ++cs) cmma<T>(cre[rs][cs], cim[rs][cs], a[rs], b[cs]);
        }
      }
    }
    // epilogue: lane holds C[i0+wr+16rs+l16][j0+wc+16cs+mfma_mrow(lq,r)]
#pragma unroll
    for (int rs = 0; rs < 2; ++rs)
#pragma unroll
      for (int cs = 0; cs < 2; ++cs)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = i0 + wr + 16 * rs + l16;
          const int gj = j0 + wc + 16 * cs + mfma_mrow<T>(lq, r);
          if (gi < d.m && gj < d.n) {
            cx<T>* c = d.C + (int64_t)gj * d.ldc + gi;
            cx<T> v{d.alpha * cre[rs][cs][r], d.alpha * cim[rs][cs][r]};
            if (d.beta != T(0)) {
              cx<T> o = *c;
              v.re += d.beta * o.re;
              v.im += d.beta * o.im;
            }
            *c = v;
          }
        }
  }
}

template <class T> __device__ __forceinline__ T tiny_pivot() { return sizeof(T) == 8 ? T(1e-150) : T(1e-18); }

// ---------------------------------------------------------------------------
// batched Gauss-Jordan inverse with partial pivoting
// ---------------------------------------------------------------------------
// For a panel J = [p0, p0+w):
//   1. copy the not-yet-pivoted rows of the panel to LDS and factor them
//      with partial pivoting (pivot = max |re|+|im|, first index on ties);
//   2. apply the w row swaps to the whole matrix, record ipiv;
//   3. X = (pivot block)^{-1} from its LU;
//   4. T = X * M[J, :] with T[:, J] = X, staged in LDS (w x n);
//   5. every row i not in J:  M[i, j] = [j not in J] M[i, j] - sum_k M[i, p0+k] T[k, j]
//      rows in J:             M[J, :] = T.
// Each wave owns whole 32-row tiles in step 5, so the reads of M[i, J]
// (the multipliers) precede all writes of row i.  At the end, the column
// interchanges are undone in reverse order:  A^{-1} = (P A)^{-1} P.
template <class T, int WP, int NMAX>
__global__ __launch_bounds__(256) void gj_inverse_kernel(const InvDesc<T>* __restrict__ descs) {
  using A4 = typename acc4<T>::type;
  constexpr int LDP = NMAX + 4;          // panel: column c at sW[c*LDP + r]
  constexpr int LDT = NMAX + 4;          // T:     row   k at sW[k*LDT + j]
  __shared__ cx<T> sW[WP * LDP];
  __shared__ cx<T> sX[WP * WP];          // column-major X
  __shared__ T redv[4];
  __shared__ int redi[4];
  __shared__ int sPiv[WP];
  __shared__ int sSing;

  const InvDesc<T> d = descs[blockIdx.x];
  const int n = d.n;
  const int64_t ld = d.ld;
  cx<T>* __restrict__ M = d.M;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  if (tid == 0) sSing = 0;

  for (int p0 = 0; p0 < n; p0 += WP) {
    const int w = min(WP, n - p0);
    const int rows = n - p0;
    __syncthreads();
    // 1. panel -> LDS (zero the unused columns)
    for (int idx = tid; idx < WP * rows; idx += 256) {
      const int c = idx / rows, r = idx - c * rows;
      cx<T> v = czero<T>();
      if (c < w) v = M[(int64_t)(p0 + c) * ld + p0 + r];
      sW[c * LDP + r] = v;
    }
    __syncthreads();
    for (int c = 0; c < w; ++c) {
      // pivot search over rows c..rows-1 of column c
      T best = T(-1);
      int bidx = 0x7fffffff;
      for (int r = c + tid; r < rows; r += 256) {
        const T v = cabs1(sW[c * LDP + r]);
        if (v > best) { best = v; bidx = r; }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const T ov = shfl_xor_t(best, off);
        const int oi = shfl_xor_t(bidx, off);
        if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
      }
      if (lane == 0) { redv[wave] = best; redi[wave] = bidx; }
      __syncthreads();
      if (tid == 0) {
        T bv = redv[0];
        int bi = redi[0];
        for (int q = 1; q < 4; ++q)
          if (redv[q] > bv || (redv[q] == bv && redi[q] < bi)) { bv = redv[q]; bi = redi[q]; }
        if (bi < c || bi >= rows) bi = c;
        sPiv[c] = bi;
        if (!(bv > T(0)) && sSing == 0) sSing = p0 + c + 1;
      }
      __syncthreads();
      const int piv = sPiv[c];
      if (piv != c && tid < WP) {
        cx<T> t0 = sW[tid * LDP + c];
        sW[tid * LDP + c] = sW[tid * LDP + piv];
        sW[tid * LDP + piv] = t0;
      }
      __syncthreads();
      cx<T> pv = sW[c * LDP + c];
      T den = pv.re * pv.re + pv.im * pv.im;
      if (!(den > T(0))) { pv = cx<T>{tiny_pivot<T>(), T(0)}; den = pv.re * pv.re; }
      const cx<T> ipv{pv.re / den, -pv.im / den};
      for (int r = c + 1 + tid; r < rows; r += 256) {
        const cx<T> l = cmul(sW[c * LDP + r], ipv);
        sW[c * LDP + r] = l;
        for (int j = c + 1; j < w; ++j) {
          const cx<T> u = sW[j * LDP + c];
          cx<T> x = sW[j * LDP + r];
          x.re -= l.re * u.re - l.im * u.im;
          x.im -= l.re * u.im + l.im * u.re;
          sW[j * LDP + r] = x;
        }
      }
      __syncthreads();
    }
    // 2. row interchanges on the whole matrix
    for (int col = tid; col < n; col += 256) {
      cx<T>* mc = M + (int64_t)col * ld + p0;
      for (int c = 0; c < w; ++c) {
        const int piv = sPiv[c];
        if (piv != c) {
          cx<T> t0 = mc[c];
          mc[c] = mc[piv];
          mc[piv] = t0;
        }
      }
    }
    if (tid < w) d.ipiv[p0 + tid] = p0 + sPiv[tid];
    // 3. X = (L U)^{-1}, column tid
    if (tid < w) {
      cx<T> y[WP];
#pragma unroll
      for (int i = 0; i < WP; ++i) {
        if (i >= w) { y[i] = czero<T>(); continue; }
        cx<T> s = (i == tid) ? cx<T>{T(1), T(0)} : czero<T>();
#pragma unroll
        for (int k = 0; k < i; ++k) s = csub(s, cmul(sW[k * LDP + i], y[k]));
        y[i] = s;
      }
#pragma unroll
      for (int ii = WP - 1; ii >= 0; --ii) {
        if (ii >= w) continue;
        cx<T> s = y[ii];
#pragma unroll
        for (int k = ii + 1; k < WP; ++k)
          if (k < w) s = csub(s, cmul(sW[k * LDP + ii], y[k]));
        cx<T> u = sW[ii * LDP + ii];
        T den = u.re * u.re + u.im * u.im;
        if (!(den > T(0))) { u = cx<T>{tiny_pivot<T>(), T(0)}; den = u.re * u.re; }
        y[ii] = cmul(s, cx<T>{u.re / den, -u.im / den});
      }
#pragma unroll
      for (int i = 0; i < WP; ++i) sX[tid * WP + i] = y[i];
    }
    if (tid >= w && tid < WP) {
      for (int i = 0; i < WP; ++i) sX[tid * WP + i] = czero<T>();
    }
    __syncthreads();
    // 4. T = X * M[J, :], T[:, J] = X ; rows k >= w of T are zero
    for (int j = tid; j < n; j += 256) {
      if (j >= p0 && j < p0 + w) {
        for (int k = 0; k < WP; ++k) sW[k * LDT + j] = (k < w) ? sX[(j - p0) * WP + k] : czero<T>();
      } else {
        cx<T> v[WP];
        const cx<T>* mc = M + (int64_t)j * ld + p0;
#pragma unroll
        for (int k = 0; k < WP; ++k) v[k] = (k < w) ? mc[k] : czero<T>();
#pragma unroll
        for (int k = 0; k < WP; ++k) {
          cx<T> s = czero<T>();
#pragma unroll
          for (int q = 0; q < WP; ++q) s = cfma(s, sX[q * WP + k], v[q]);
          sW[k * LDT + j] = s;
        }
      }
    }
    __syncthreads();
    // 5. rank-w update, one wave per 32-row tile
    const int ntr = (n + 31) / 32, ntc = (n + 31) / 32;
    for (int tr = wave; tr < ntr; tr += 4) {
      const int i0 = tr * 32;
      cx<T> g[2][WP / 4];
#pragma unroll
      for (int rs = 0; rs < 2; ++rs)
#pragma unroll
        for (int kk = 0; kk < WP / 4; ++kk) {
          const int gi = i0 + 16 * rs + l16, gk = 4 * kk + lq;
          g[rs][kk] = (gi < n && gk < w) ? M[(int64_t)(p0 + gk) * ld + gi] : czero<T>();
        }
      for (int tc = 0; tc < ntc; ++tc) {
        const int j0 = tc * 32;
        A4 cre[2][2], cim[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            cre[a][b] = A4{0, 0, 0, 0};
            cim[a][b] = A4{0, 0, 0, 0};
          }
#pragma unroll
        for (int kk = 0; kk < WP / 4; ++kk) {
          cx<T> b[2];
#pragma unroll
          for (int cs = 0; cs < 2; ++cs) {
            const int gj = j0 + 16 * cs + l16;
            b[cs] = (gj < n) ? sW[(4 * kk + lq) * LDT + gj] : czero<T>();
          }
#pragma unroll
          for (int rs = 0; rs < 2; ++rs)
#pragma unroll
            for (int cs = 0; cs < 2; ++cs) cmma<T>(cre[rs][cs], cim[rs][cs], g[rs][kk], b[cs]);
        }
#pragma unroll
        for (int rs = 0; rs < 2; ++rs)
#pragma unroll
          for (int cs = 0; cs < 2; ++cs)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int gi = i0 + 16 * rs + l16;
              const int gj = j0 + 16 * cs + mfma_mrow<T>(lq, r);
              if (gi < n && gj < n) {
                cx<T>* c = M + (int64_t)gj * ld + gi;
                cx<T> v;
                if (gi >= p0 && gi < p0 + w) {
                  v = sW[(gi - p0) * LDT + gj];
                } else if (gj >= p0 && gj < p0 + w) {
                  v = cx<T>{-cre[rs][cs][r], -cim[rs][cs][r]};
                } else {
                  v = *c;
                  v.re -= cre[rs][cs][r];
                  v.im -= cim[rs][cs][r];
                }
                *c = v;
              }
            }
      }
    }
  }
  __syncthreads();
  // undo the column interchanges in reverse order
  for (int p = n - 1; p >= 0; --p) {
    const int q = d.ipiv[p];
    if (q != p) {
      cx<T>* a = M + (int64_t)p * ld;
      cx<T>* b = M + (int64_t)q * ld;
      for (int r = tid; r < n; r += 256) {
        cx<T> t0 = a[r];
        a[r] = b[r];
        b[r] = t0;
      }
    }
    __syncthreads();
  }
  if (tid == 0 && d.info) *d.info = sSing;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class T>
hipError_t launch_gemm_batched(const GemmDesc<T>* d_descs, int ndesc, int max_tiles, hipStream_t s) {
  if (ndesc <= 0 || max_tiles <= 0) return hipSuccess;
  dim3 grid(max_tiles, ndesc);
  hipLaunchKernelGGL(gemm_batched_kernel<T>, grid, dim3(256), 0, s, d_descs);
  return hipGetLastError();
}

template <class T>
hipError_t launch_gj_inverse(const InvDesc<T>* d_descs, int ndesc, int nmax, hipStream_t s) {
  if (ndesc <= 0) return hipSuccess;
  if constexpr (sizeof(T) == 8) {
    if (nmax <= 256)
      hipLaunchKernelGGL((gj_inverse_kernel<T, 32, 256>), dim3(ndesc), dim3(256), 0, s, d_descs);
    else if (nmax <= 512)
      hipLaunchKernelGGL((gj_inverse_kernel<T, 16, 512>), dim3(ndesc), dim3(256), 0, s, d_descs);
    else
      return hipErrorInvalidValue;
  } else {
    if (nmax <= 512)
      hipLaunchKernelGGL((gj_inverse_kernel<T, 32, 512>), dim3(ndesc), dim3(256), 0, s, d_descs);
    else
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template hipError_t launch_gemm_batched<double>(const GemmDesc<double>*, int, int, hipStream_t);
template hipError_t launch_gemm_batched<float>(const GemmDesc<float>*, int, int, hipStream_t);
template hipError_t launch_gj_inverse<double>(const InvDesc<double>*, int, int, hipStream_t);
template hipError_t launch_gj_inverse<float>(const InvDesc<float>*, int, int, hipStream_t);

}  // namespace kore
