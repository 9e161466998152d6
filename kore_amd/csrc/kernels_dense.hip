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
// Rows are never moved while the block is reduced.  Column j is eliminated
// with pivot row r_j (largest |re|+|im| among the rows not used yet, first
// index on ties); the in-place Gauss-Jordan update then leaves in column j of
// M the column r_j of the transformed identity, so that at the end
//     A^{-1}[k, l] = M[r_k, r^{-1}(l)],
// which is applied once, through the caller's scratch block.
// For a panel J = [p0, p0+w) with pivot rows P = (r_p0, ..., r_{p0+w-1}):
//   1. the w panel columns (all rows) are copied to LDS and factored there
//      with partial pivoting over the unused rows;
//   2. X = M[P, J]^{-1} from that LU;
//   3. T = X * M[P, :], with T[:, J] = X, staged in LDS (w x n);
//   4. rows i not in P:  M[i, j] = [j not in J] M[i, j] - sum_k M[i, p0+k] T[k, j];
//      rows P:           M[P, :] = T.
// Each wave owns whole 32-row tiles in step 4, so the reads of M[i, J] (the
// multipliers) precede all writes of row i.
template <class T, int WP, int NMAX>
__global__ __launch_bounds__(256) void gj_inverse_kernel(const InvDesc<T>* __restrict__ descs) {
  using A4 = typename acc4<T>::type;
  constexpr int LDP = NMAX + 4;          // panel: column c at sW[c*LDP + r]
  constexpr int LDT = NMAX + 4;          // T:     row   k at sW[k*LDT + j]
  __shared__ cx<T> sW[WP * LDP];
  __shared__ cx<T> sX[WP * WP];          // column-major X
  __shared__ T redv[4];
  __shared__ int redi[4];
  __shared__ int sPiv[WP];               // pivot rows of the current panel
  __shared__ short sRole[NMAX];          // -1 unused, -2 used earlier, c >= 0: pivot c of this panel
  __shared__ int sSing;

  const InvDesc<T> d = descs[blockIdx.x];
  const int n = d.n;
  const int64_t ld = d.ld;
  cx<T>* __restrict__ M = d.M;
  int* __restrict__ rowof = d.ipiv;      // r_j: pivot row of column j
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  if (tid == 0) sSing = 0;
  for (int r = tid; r < n; r += 256) sRole[r] = -1;

  for (int p0 = 0; p0 < n; p0 += WP) {
    const int w = min(WP, n - p0);
    __syncthreads();
    // 1. panel (all rows) -> LDS, zero the unused columns
    for (int idx = tid; idx < WP * n; idx += 256) {
      const int c = idx / n, r = idx - c * n;
      sW[c * LDP + r] = (c < w) ? M[(int64_t)(p0 + c) * ld + r] : czero<T>();
    }
    __syncthreads();
    for (int c = 0; c < w; ++c) {
      T best = T(-1);
      int bidx = 0x7fffffff;
      for (int r = tid; r < n; r += 256) {
        if (sRole[r] != -1) continue;
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
      // every thread reduces the four wave results itself (no second barrier)
      T bv = redv[0];
      int p = redi[0];
#pragma unroll
      for (int q = 1; q < 4; ++q)
        if (redv[q] > bv || (redv[q] == bv && redi[q] < p)) { bv = redv[q]; p = redi[q]; }
      if (p < 0 || p >= n) {   // no finite candidate left: take the first unused row
        p = 0;
        while (p < n && sRole[p] != -1) ++p;
        if (p >= n) p = 0;
      }
      if (tid == 0) {
        sPiv[c] = p;
        if (!(bv > T(0)) && sSing == 0) sSing = p0 + c + 1;
      }
      cx<T> pv = sW[c * LDP + p];
      T den = pv.re * pv.re + pv.im * pv.im;
      if (!(den > T(0))) { pv = cx<T>{tiny_pivot<T>(), T(0)}; den = pv.re * pv.re; }
      const cx<T> ipv{pv.re / den, -pv.im / den};
      for (int r = tid; r < n; r += 256) {
        if (r == p || sRole[r] != -1) continue;
        const cx<T> l = cmul(sW[c * LDP + r], ipv);
        sW[c * LDP + r] = l;
        for (int j = c + 1; j < w; ++j) {
          const cx<T> u = sW[j * LDP + p];
          cx<T> x = sW[j * LDP + r];
          x.re -= l.re * u.re - l.im * u.im;
          x.im -= l.re * u.im + l.im * u.re;
          sW[j * LDP + r] = x;
        }
      }
      __syncthreads();                 // all reads of sRole for this column done
      if (tid == 0) sRole[p] = (short)c;
      __syncthreads();
    }
    if (tid < w) rowof[p0 + tid] = sPiv[tid];
    // 2. X = (pivot block)^{-1}; the LU of M[P, J] is in the panel rows P:
    //    U(i, k) = sW[k][P_i] (k >= i),  L(i, k) = sW[k][P_i] (k < i, unit diagonal)
    if (tid < w) {
      cx<T> y[WP];
#pragma unroll
      for (int i = 0; i < WP; ++i) {
        if (i >= w) { y[i] = czero<T>(); continue; }
        cx<T> sacc = (i == tid) ? cx<T>{T(1), T(0)} : czero<T>();
        const int pi = sPiv[i];
#pragma unroll
        for (int k = 0; k < i; ++k) sacc = csub(sacc, cmul(sW[k * LDP + pi], y[k]));
        y[i] = sacc;
      }
#pragma unroll
      for (int ii = WP - 1; ii >= 0; --ii) {
        if (ii >= w) continue;
        const int pi = sPiv[ii];
        cx<T> sacc = y[ii];
#pragma unroll
        for (int k = ii + 1; k < WP; ++k)
          if (k < w) sacc = csub(sacc, cmul(sW[k * LDP + pi], y[k]));
        cx<T> u = sW[ii * LDP + pi];
        T den = u.re * u.re + u.im * u.im;
        if (!(den > T(0))) { u = cx<T>{tiny_pivot<T>(), T(0)}; den = u.re * u.re; }
        y[ii] = cmul(sacc, cx<T>{u.re / den, -u.im / den});
      }
#pragma unroll
      for (int i = 0; i < WP; ++i) sX[tid * WP + i] = y[i];
    }
    if (tid >= w && tid < WP) {
      for (int i = 0; i < WP; ++i) sX[tid * WP + i] = czero<T>();
    }
    __syncthreads();
    // 3. T = X * M[P, :], T[:, J] = X ; rows k >= w of T are zero
    for (int j = tid; j < n; j += 256) {
      if (j >= p0 && j < p0 + w) {
        for (int k = 0; k < WP; ++k) sW[k * LDT + j] = (k < w) ? sX[(j - p0) * WP + k] : czero<T>();
      } else {
        cx<T> v[WP];
        const cx<T>* mc = M + (int64_t)j * ld;
#pragma unroll
        for (int k = 0; k < WP; ++k) v[k] = (k < w) ? mc[sPiv[k]] : czero<T>();
#pragma unroll 2
        for (int k = 0; k < WP; ++k) {
          cx<T> sacc = czero<T>();
#pragma unroll
          for (int q = 0; q < WP; ++q) sacc = cfma(sacc, sX[q * WP + k], v[q]);
          sW[k * LDT + j] = sacc;
        }
      }
    }
    __syncthreads();
    // 4. rank-w update, one wave per 32-row tile
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
                cx<T>* cp = M + (int64_t)gj * ld + gi;
                const int role = sRole[gi];
                cx<T> v;
                if (role >= 0 && role < WP) {
                  v = sW[role * LDT + gj];
                } else if (gj >= p0 && gj < p0 + w) {
                  v = cx<T>{-cre[rs][cs][r], -cim[rs][cs][r]};
                } else {
                  v = *cp;
                  v.re -= cre[rs][cs][r];
                  v.im -= cim[rs][cs][r];
                }
                *cp = v;
              }
            }
      }
    }
    __syncthreads();
    // this panel's pivot rows become "used earlier"
    for (int c = tid; c < w; c += 256) sRole[sPiv[c]] = -2;
  }
  __syncthreads();
  // A^{-1}[k, l] = M[r_k, r^{-1}(l)]: permute into the scratch block, copy back
  int* rinv = reinterpret_cast<int*>(sX);          // n ints fit in sX (WP*WP complex >= NMAX ints)
  for (int j = tid; j < n; j += 256) rinv[rowof[j]] = j;
  __syncthreads();
  cx<T>* __restrict__ S = d.scratch;
  for (int l = 0; l < n; ++l) {
    const cx<T>* src = M + (int64_t)rinv[l] * ld;
    for (int k = tid; k < n; k += 256) S[(int64_t)l * n + k] = src[rowof[k]];
  }
  __syncthreads();
  for (int l = 0; l < n; ++l)
    for (int k = tid; k < n; k += 256) M[(int64_t)l * ld + k] = S[(int64_t)l * n + k];
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
