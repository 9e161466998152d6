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
// 2x2 MFMA 16x16 tiles; K staged through LDS in slabs of GT_K = 16 (32 measured
// slower: the larger LDS and register footprint halves the workgroups per CU).
constexpr int GT_M = 64, GT_N = 64, GT_K = 16;
constexpr int GT_EPT = GT_K * GT_M / 256;   // slab elements of A (and of B) per thread
constexpr int GT_PADB = 1;

// One 64 x 64 tile of C = alpha sum_p A_p B_p + beta C for the workgroup
// (256 threads); sA / sB are the workgroup's LDS staging.
template <class T>
__device__ __forceinline__ void gemm_tile_body(const GemmDesc<T>& d, const int tile, cx<T> (&sA)[GT_K][GT_M],
                                               cx<T> (&sB)[GT_N][GT_K + GT_PADB]) {
  using A4 = typename acc4<T>::type;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int l16 = lane & 15, lq = lane >> 4;
  {
    const int i0 = (tile % d.tiles_m) * GT_M;
    const int j0 = (tile / d.tiles_m) * GT_N;
    // 16 x 16 sub-tiles wholly outside C (edge tiles of sizes that are not
    // multiples of 64, e.g. 200 = 3*64 + 8) are skipped, wave-uniformly
    const bool rv0 = i0 + wr < d.m, rv1 = i0 + wr + 16 < d.m;
    const bool cv0 = j0 + wc < d.n, cv1 = j0 + wc + 16 < d.n;
    // complex64: three real products per complex step (Karatsuba's 3M
    // form): cre = sum ar br, cim = sum ai bi, csum = sum (ar + ai)(br + bi),
    // and at the end re = cre - cim, im = csum - cre - cim: 3/4 of the
    // matrix-core work of the four-product form, with a normwise error bound
    // of the same order (Higham, Accuracy and Stability of Numerical
    // Algorithms, section 23.2.4); the complex64 factor is a preconditioner
    // whose answers FP64 refinement or GMRES corrects.  complex128 keeps the
    // four products (cre, cim = real and imaginary parts): measured, the 3M
    // form makes the FP64 factor 10 % faster at N = 500 but moves the most
    // ill-conditioned MAC eigenvalues (relative condition 1e16) further from
    // the oracle's than the four-product form does.
    constexpr bool k3m = sizeof(T) == 4;
    A4 cre[2][2], cim[2][2], csum[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        cre[a][b] = A4{0, 0, 0, 0};
        cim[a][b] = A4{0, 0, 0, 0};
        csum[a][b] = A4{0, 0, 0, 0};
      }

    // Slabs of GT_K k-columns over the (up to two) products, register
    // double-buffered: slab s+1 is requested from memory before slab s is
    // multiplied, so the load latency hides behind the matrix-core work.
    cx<T> ra[GT_EPT], rb[GT_EPT];
    auto load = [&](int p, int k0) {      // selects, not indexing: d stays in registers
      const cx<T>* __restrict__ A = p ? d.A[1] : d.A[0];
      const cx<T>* __restrict__ B = p ? d.B[1] : d.B[0];
      const int K = p ? d.k[1] : d.k[0];
      const int64_t lda = p ? d.lda[1] : d.lda[0], ldb = p ? d.ldb[1] : d.ldb[0];
      if (d.opA == 0) {        // A slab element (k = tid / (64/EPT), i = EPT (tid % (64/EPT)) + q)
        const int gk = k0 + tid / (GT_M / GT_EPT), ii = (tid % (GT_M / GT_EPT)) * GT_EPT;
#pragma unroll
        for (int q = 0; q < GT_EPT; ++q) {
          const int gi = i0 + ii + q;
          ra[q] = czero<T>();
          if (gk < K && gi < d.m) ra[q] = gld(A + (int64_t)gk * lda + gi);
        }
      } else {                 // A_eff(i,k) = conj(A_store(k,i)), A_store column-major K x m
        const int gi = i0 + (tid >> 2), kk = (tid & 3) * GT_EPT;
#pragma unroll
        for (int q = 0; q < GT_EPT; ++q) {
          const int gk = k0 + kk + q;
          ra[q] = czero<T>();
          if (gk < K && gi < d.m) {
            ra[q] = gld(A + (int64_t)gi * lda + gk);
            ra[q].im = -ra[q].im;
          }
        }
      }
      const int gj = j0 + (tid >> 2), kk = (tid & 3) * GT_EPT;   // B slab element (j, k)
#pragma unroll
      for (int q = 0; q < GT_EPT; ++q) {
        const int gk = k0 + kk + q;
        rb[q] = czero<T>();
        if (gk < K && gj < d.n) rb[q] = gld(B + (int64_t)gj * ldb + gk);
      }
    };
    auto valid = [&](int p) { return d.A[p] != nullptr && d.k[p] > 0; };
    int p = valid(0) ? 0 : (valid(1) ? 1 : 2), k0 = 0;
    if (p < 2) load(p, 0);
    while (p < 2) {
      __syncthreads();                      // the previous slab has been consumed
      if (d.opA == 0) {
        const int kk = tid / (GT_M / GT_EPT), ii = (tid % (GT_M / GT_EPT)) * GT_EPT;
#pragma unroll
        for (int q = 0; q < GT_EPT; ++q) sA[kk][ii + q] = ra[q];
      } else {
        const int ii = tid >> 2, kk = (tid & 3) * GT_EPT;
#pragma unroll
        for (int q = 0; q < GT_EPT; ++q) sA[kk + q][ii] = ra[q];
      }
      {
        const int jj = tid >> 2, kk = (tid & 3) * GT_EPT;
#pragma unroll
        for (int q = 0; q < GT_EPT; ++q) sB[jj][kk + q] = rb[q];
      }
      __syncthreads();
      int pn = p, kn = k0 + GT_K;           // next slab
      if (kn >= (p ? d.k[1] : d.k[0])) {
        pn = (p == 0 && valid(1)) ? 1 : 2;
        kn = 0;
      }
      if (pn < 2) load(pn, kn);
#pragma unroll
      for (int kk = 0; kk < GT_K; kk += 4) {
        cx<T> a[2], b[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          a[s] = sA[kk + lq][wr + 16 * s + l16];
          b[s] = sB[wc + 16 * s + l16][kk + lq];
        }
        if constexpr (k3m) {
          T as[2], bs[2];
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            as[q] = a[q].re + a[q].im;
            bs[q] = b[q].re + b[q].im;
          }
#pragma unroll
          for (int rs = 0; rs < 2; ++rs)
#pragma unroll
            for (int cs = 0; cs < 2; ++cs)
              if ((rs ? rv1 : rv0) && (cs ? cv1 : cv0)) {
                cre[rs][cs] = mfma16(b[cs].re, a[rs].re, cre[rs][cs]);
                cim[rs][cs] = mfma16(b[cs].im, a[rs].im, cim[rs][cs]);
                csum[rs][cs] = mfma16(bs[cs], as[rs], csum[rs][cs]);
              }
        } else {
#pragma unroll
          for (int rs = 0; rs < 2; ++rs)
#pragma unroll
            for (int cs = 0; cs < 2; ++cs)
              if ((rs ? rv1 : rv0) && (cs ? cv1 : cv0)) cmma<T>(cre[rs][cs], cim[rs][cs], a[rs], b[cs]);
        }
      }
      p = pn;
      k0 = kn;
    }
    __syncthreads();                        // LDS is reused by the next tile
    // epilogue: lane holds C[i0+wr+16rs+l16][j0+wc+16cs+mfma_mrow(lq,r)].
    // Per row sub-tile, the eight old values of C are loaded together before
    // any store, so that their latencies overlap (a load after a store to a
    // possibly aliasing address would wait for the previous one).
    const bool acc = d.beta != T(0);
#pragma unroll
    for (int rs = 0; rs < 2; ++rs) {
      const int gi = i0 + wr + 16 * rs + l16;
      cx<T> o[2][4];
#pragma unroll
      for (int cs = 0; cs < 2; ++cs)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gj = j0 + wc + 16 * cs + mfma_mrow<T>(lq, r);
          o[cs][r] = czero<T>();
          if (acc && gi < d.m && gj < d.n) o[cs][r] = gld(d.C + (int64_t)gj * d.ldc + gi);
        }
#pragma unroll
      for (int cs = 0; cs < 2; ++cs)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gj = j0 + wc + 16 * cs + mfma_mrow<T>(lq, r);
          if (gi < d.m && gj < d.n) {
            T vre = cre[rs][cs][r], vim = cim[rs][cs][r];
            if constexpr (k3m) {
              vre = cre[rs][cs][r] - cim[rs][cs][r];
              vim = csum[rs][cs][r] - cre[rs][cs][r] - cim[rs][cs][r];
            }
            cx<T> v{d.alpha * vre, d.alpha * vim};
            if (acc) {
              v.re += d.beta * o[cs][r].re;
              v.im += d.beta * o[cs][r].im;
            }
            gst(d.C + (int64_t)gj * d.ldc + gi, v);
          }
        }
    }
  }
}

template <class T>
__global__ __launch_bounds__(256, 3) void gemm_batched_kernel(const GemmDesc<T>* __restrict__ descs) {
  const GemmDesc<T> d = descs[blockIdx.y];
  const int ntiles = d.tiles_m * d.tiles_n;
  if ((int)blockIdx.x >= ntiles) return;
  __shared__ cx<T> sA[GT_K][GT_M];
  __shared__ cx<T> sB[GT_N][GT_K + GT_PADB];
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) gemm_tile_body<T>(d, tile, sA, sB);
}

template <class T> __device__ __forceinline__ T tiny_pivot() { return sizeof(T) == 8 ? T(1e-150) : T(1e-18); }

// ---------------------------------------------------------------------------
// batched Gauss-Jordan inverse with partial pivoting
// ---------------------------------------------------------------------------
// One workgroup of 512 threads per block.  Rows are never moved: column j is
// eliminated with pivot row r_j (within a factor 1 - 2^-11 of the largest
// |re|+|im| among the rows not used yet, first index on ties), so that at the end
//     A^{-1}[k, l] = M[r_k, r^{-1}(l)],
// applied once through the caller's scratch block.
//
// Panel J = [p0, p0+w) with pivot rows P:
//   1. Gauss-Jordan on the panel columns over all rows, in registers (each
//      row owned by TPR adjacent lanes, CPT columns each).  One column costs
//      two barriers: the pivot is a packed 32-bit magnitude|row key reduced
//      within each wave by DPP and across the waves through LDS; then the
//      pivot row's lanes store it for all to read.
//      Afterwards the panel holds C: rows P hold X = M[P, J]^{-1} (permuted),
//      the other rows hold -M[i, J] X -- the final columns J of M.
//   2. R = M[P, :] (old pivot rows) is staged in LDS.
//   3. Outside J:  M[i, j] <- [i not in P] M[i, j] + sum_k C[i, k] R[k, j]
//      (matrix cores; C read back from the block, R from LDS).
// Cross-lane moves within rows of 16 lanes (DPP) and lane reads.
template <int CTRL> __device__ __forceinline__ unsigned dpp_u32(unsigned x) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ double dpp_t(double x) {
  return __hiloint2double((int)dpp_u32<CTRL>((unsigned)__double2hiint(x)), (int)dpp_u32<CTRL>((unsigned)__double2loint(x)));
}
template <int CTRL> __device__ __forceinline__ float dpp_t(float x) {
  return __uint_as_float(dpp_u32<CTRL>(__float_as_uint(x)));
}

// 32-bit pivot key of the Gauss-Jordan kernels: valid flag (bit 31), the
// magnitude's exponent and top 11 mantissa bits (bits 30-9), 511 - row.  The
// chosen pivot is within a factor 1 - 2^-11 of the column's largest entry
// (threshold partial pivoting; ties go to the lower row); one max per
// reduction stage instead of a 64-bit compare and two selects.
__device__ __forceinline__ unsigned pivot_key32(double mag, int r) {
  if (!(mag <= 1.7976931348623157e308)) mag = 0.0;   // NaN / inf never win
  const unsigned long long b = (unsigned long long)__double_as_longlong(mag);
  return 0x80000000u | (((unsigned)(b >> 41) & 0x3FFFFFu) << 9) | (unsigned)(511 - r);
}
// Maximum over the wavefront, returned to every lane: quad swaps and row
// mirrors leave each row of 16 lanes with its maximum, then the four rows
// are combined through lanes 15, 31, 47 and 63.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  v = max(v, dpp_u32<0xB1>(v));        // quad_perm [1,0,3,2]
  v = max(v, dpp_u32<0x4E>(v));        // quad_perm [2,3,0,1]
  v = max(v, dpp_u32<0x141>(v));       // row_half_mirror
  v = max(v, dpp_u32<0x140>(v));       // row_mirror
  const unsigned a = max((unsigned)__builtin_amdgcn_readlane((int)v, 15), (unsigned)__builtin_amdgcn_readlane((int)v, 31));
  const unsigned b = max((unsigned)__builtin_amdgcn_readlane((int)v, 47), (unsigned)__builtin_amdgcn_readlane((int)v, 63));
  return max(a, b);
}


template <class T, int WP, int NMAX>
__global__ __launch_bounds__(512) void gj_inverse_kernel(const InvDesc<T>* __restrict__ descs) {
  using A4 = typename acc4<T>::type;
  constexpr int NT = 512;
  constexpr int NW = NT / 64;
  constexpr int TPR = NT / NMAX;         // threads per row
  constexpr int CPT = WP / TPR;          // panel columns per thread
  constexpr int LDR = NMAX + 4;          // R: row k at sR[k*LDR + j]
  static_assert(TPR >= 1 && CPT * TPR == WP, "panel split");
  __shared__ cx<T> sR[WP * LDR];
  __shared__ cx<T> sRow[2][WP];           // the pivot row (double-buffered)
  __shared__ unsigned sKey[2][NW];
  __shared__ short sRole[NMAX];          // -1 unused, -2 used earlier, c >= 0: pivot c of this panel
  __shared__ int sPiv[WP];
  __shared__ int sSing;

  const InvDesc<T> d = descs[blockIdx.x];
  const int n = d.n;
  const int64_t ld = d.ld;
  cx<T>* __restrict__ M = d.M;
  int* __restrict__ rowof = d.ipiv;      // r_j: pivot row of column j
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  const int r = tid / TPR, h = tid % TPR;   // owned row and column slice
  const bool live = r < n;
  int role = -1;                          // register copy of sRole[r]
  if (tid == 0) sSing = 0;
  for (int q = tid; q < NMAX; q += NT) sRole[q] = -1;

  for (int p0 = 0; p0 < n; p0 += WP) {
    const int w = min(WP, n - p0);
    // 1. panel -> registers
    cx<T> v[CPT];
#pragma unroll
    for (int jj = 0; jj < CPT; ++jj) {
      const int j = h * CPT + jj;
      v[jj] = (live && j < w) ? M[(int64_t)(p0 + j) * ld + r] : czero<T>();
    }
    // Two barriers per column: every wave publishes its best 32-bit key
    // (pivot_key32: threshold pivoting within 1 - 2^-11 of the largest
    // entry); after the first barrier all threads pick the best of the
    // eight, the pivot row's lanes store it, and after the second barrier
    // all read it (one row of stores per column instead of one per wave).
#pragma unroll
    for (int c = 0; c < WP; ++c) {
      if (c < w) {                        // uniform
        const int par = c & 1;
        unsigned key = 0;
        if (live && h == c / CPT && role == -1) key = pivot_key32((double)cabs1(v[c % CPT]), r);
        key = wave_max_u32(key);
        if (lane == 0) sKey[par][wave] = key;
        __syncthreads();
        unsigned kk[NW];
#pragma unroll
        for (int qq = 0; qq < NW; ++qq) kk[qq] = sKey[par][qq];
#pragma unroll
        for (int st = 1; st < NW; st *= 2)
#pragma unroll
          for (int i = 0; i + st < NW; i += 2 * st) kk[i] = max(kk[i], kk[i + st]);
        const unsigned best = kk[0];
        // every column has an unused row left, so a valid key always exists
        const int p = best ? 511 - (int)(best & 0x1FFu) : 0;
        if (tid == 0) {
          sRole[p] = (short)c;
          sPiv[c] = p;
        }
        if (r == p) role = c;
        if (best != 0 && r == p) {
#pragma unroll
          for (int jj = 0; jj < CPT; ++jj) sRow[par][h * CPT + jj] = v[jj];
        }
        __syncthreads();
        const cx<T>* __restrict__ prow = sRow[par];
        cx<T> pv = prow[c];
        T den = pv.re * pv.re + pv.im * pv.im;
        if (!(den > T(0))) {
          if (tid == 0 && sSing == 0) sSing = p0 + c + 1;
          pv = cx<T>{tiny_pivot<T>(), T(0)};
          den = pv.re * pv.re;
        }
        const T iden = T(1) / den;
        const cx<T> ipv{pv.re * iden, -pv.im * iden};
        // this row's entry in column c, held by lane slice c / CPT of the row
        cx<T> vc = v[c % CPT];
        if constexpr (TPR == 2) {
          if (c / CPT == 0) {
            vc.re = dpp_t<0xA0>(vc.re);   // quad_perm [0,0,2,2]
            vc.im = dpp_t<0xA0>(vc.im);
          } else {
            vc.re = dpp_t<0xF5>(vc.re);   // quad_perm [1,1,3,3]
            vc.im = dpp_t<0xF5>(vc.im);
          }
        }
        if (live) {
          if (r == p) {
#pragma unroll
            for (int jj = 0; jj < CPT; ++jj) {
              const int j = h * CPT + jj;
              v[jj] = (j == c) ? ipv : cmul(v[jj], ipv);
            }
          } else {
            const cx<T> l = cmul(vc, ipv);
#pragma unroll
            for (int jj = 0; jj < CPT; ++jj) {
              const int j = h * CPT + jj;
              if (j == c) {
                v[jj] = cx<T>{-l.re, -l.im};
              } else if (j < w) {
                const cx<T> u = prow[j];
                v[jj].re -= l.re * u.re - l.im * u.im;
                v[jj].im -= l.re * u.im + l.im * u.re;
              }
            }
          }
        }
      }
    }
    // C -> columns J of the block (final values)
    if (live) {
#pragma unroll
      for (int jj = 0; jj < CPT; ++jj) {
        const int j = h * CPT + jj;
        if (j < w) M[(int64_t)(p0 + j) * ld + r] = v[jj];
      }
    }
    if (tid < w) rowof[p0 + tid] = sPiv[tid];
    // 2. R = M[P, :] outside J (rows k >= w are zero)
    for (int idx = tid; idx < WP * n; idx += NT) {
      const int k = idx % WP, j = idx / WP;
      cx<T> x = czero<T>();
      if (k < w && (j < p0 || j >= p0 + w)) x = M[(int64_t)j * ld + sPiv[k]];
      sR[k * LDR + j] = x;
    }
    __syncthreads();
    // 3. rank-w update outside J, one wave per 32-row tile
    const int ntr = (n + 31) / 32, ntc = (n + 31) / 32;
    for (int tr = wave; tr < ntr; tr += NW) {
      const int i0 = tr * 32;
      cx<T> g[2][WP / 4];
#pragma unroll
      for (int rs = 0; rs < 2; ++rs)
#pragma unroll
        for (int kk = 0; kk < WP / 4; ++kk) {
          const int gi = i0 + 16 * rs + l16, gk = 4 * kk + lq;
          g[rs][kk] = (gi < n && gk < w) ? M[(int64_t)(p0 + gk) * ld + gi] : czero<T>();
        }
      int prole[2];
#pragma unroll
      for (int rs = 0; rs < 2; ++rs) {
        const int gi = i0 + 16 * rs + l16;
        prole[rs] = gi < n ? (int)sRole[gi] : -1;
      }
      for (int tc = 0; tc < ntc; ++tc) {
        const int j0 = tc * 32;
        if (j0 >= p0 && j0 + 32 <= p0 + w) continue;   // tile inside J: already final
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
            b[cs] = (gj < n) ? sR[(4 * kk + lq) * LDR + gj] : czero<T>();
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
            for (int rr = 0; rr < 4; ++rr) {
              const int gi = i0 + 16 * rs + l16;
              const int gj = j0 + 16 * cs + mfma_mrow<T>(lq, rr);
              if (gi < n && gj < n && (gj < p0 || gj >= p0 + w)) {
                cx<T>* cp = M + (int64_t)gj * ld + gi;
                cx<T> x{cre[rs][cs][rr], cim[rs][cs][rr]};
                if (!(prole[rs] >= 0)) {
                  const cx<T> o = *cp;
                  x.re += o.re;
                  x.im += o.im;
                }
                *cp = x;
              }
            }
      }
    }
    __syncthreads();
    // this panel's pivot rows become "used earlier"
    if (role >= 0) role = -2;
    for (int c = tid; c < w; c += NT) sRole[sPiv[c]] = -2;
  }
  __syncthreads();
  // A^{-1}[k, l] = M[r_k, r^{-1}(l)]: permute into the scratch block, copy back
  int* rinv = reinterpret_cast<int*>(sR);
  for (int j = tid; j < n; j += NT) rinv[rowof[j]] = j;
  __syncthreads();
  cx<T>* __restrict__ S = d.scratch;
  for (int l = 0; l < n; ++l) {
    const cx<T>* src = M + (int64_t)rinv[l] * ld;
    for (int k = tid; k < n; k += NT) S[(int64_t)l * n + k] = src[rowof[k]];
  }
  __syncthreads();
  for (int l = 0; l < n; ++l)
    for (int k = tid; k < n; k += NT) M[(int64_t)l * ld + k] = S[(int64_t)l * n + k];
  if (tid == 0 && d.info) *d.info = sSing;
}

// ---------------------------------------------------------------------------
// batched Gauss-Jordan inverse, multi-workgroup form (large blocks)
// ---------------------------------------------------------------------------
// The same elimination as gj_inverse_kernel, split so that the rank-w update
// of a block spreads over many workgroups (one 64 x 64 tile each) instead of
// one compute unit: per panel J = [p0, p0+w)
//   gj_panel_kernel   one workgroup per block: Gauss-Jordan on the panel
//                     columns in registers (pivot rows P), C -> M[:, J],
//                     R = M[P, :] (rows before the update, zero on J) -> scratch;
//   gj_update_kernel  M[i, j] <- [i not in P] M[i, j] + sum_k C[i, k] R[k, j]
//                     for j outside J, on matrix cores;
// then gj_perm_kernel twice: A^{-1}[k, l] = M[r_k, r^{-1}(l)] through the
// scratch block and back.  Scratch layout (per block, n x n complex): R
// (32 x n, R[k, j] at R[j 32 + k]), then n ints of row roles (-1 unused,
// c >= 0: the pivot row of column c).  The
// multi-panel composition: with T the product of the panel's elementary
// transforms, T differs from the identity only in the pivot columns P, so for
// a column j outside J,  T M[:, j] = [i not in P] M[i, j] + T[:, P] M[P, j],
// and T[:, P] is exactly what in-place elimination leaves in M[:, J].
// v[idx] by selects, so a register array never becomes indexable memory
template <int N, class V>
__device__ __forceinline__ V pick(const V (&v)[N], int idx) {
  V x = v[0];
#pragma unroll
  for (int q = 1; q < N; ++q)
    if (q == idx) x = v[q];
  return x;
}

constexpr int GJM_WP = 32;
constexpr int GJ_MULTI_MIN = 96;          // blocks at least this large take the multi-workgroup form
// Two-level panels for blocks of at least GJ_OUTER_MIN rows: the panels of
// 32 columns are grouped into outer panels O of GJ_WO columns.  A panel's
// rank-32 update is applied to the columns of its outer panel only; once the
// outer panel is complete, the columns outside it receive the composite
// transform of its GJ_WO eliminations in one rank-GJ_WO product (the same
// composition argument as for a single panel, with P = the outer panel's
// pivot rows).  That halves the passes over the block and turns most of the
// update into a product with K = GJ_WO on the matrix cores.  Scratch layout
// then: R (32 x n), n ints of row roles, R_O (GJ_WO x n): 161 n entries, so
// the block needs n >= 161.  Below 256 rows the two extra launches per outer
// panel cost more than the passes they save (measured at 200 rows).
constexpr int GJ_WO = 128;
constexpr int GJ_OUTER_MIN = 256;

// Outer panel [q0, q0 + wo) holding column p0 of an n x n block.
__device__ __forceinline__ void gj_outer_range(int n, int p0, int& q0, int& wo) {
  if (n >= GJ_OUTER_MIN) {
    q0 = (p0 / GJ_WO) * GJ_WO;
    wo = min(GJ_WO, n - q0);
  } else {
    q0 = 0;
    wo = n;
  }
}

// NT threads, NT / NMAX per row.  A column step is a dependent chain (pivot
// key, wave and workgroup reductions, pivot-row read, reciprocal) plus the
// row updates.  Blocks of up to 256 rows: complex128 takes 512 threads, two
// per row, so that a second wave per SIMD hides the chain's latencies;
// complex64, whose updates cost half as much, takes 256 threads, one per row
// (measured at 200-row blocks: 256 threads make the complex64 factorisation
// 1.6 % faster and the complex128 one 4 % slower).  Larger blocks: 512.
template <class T>
constexpr int GJ_PANEL_NT256 = sizeof(T) == 4 ? 256 : 512;

template <class T, int NMAX, int NT>
__global__ __launch_bounds__(NT) void gj_panel_kernel(const InvDesc<T>* __restrict__ descs, int p0) {
  constexpr int NW = NT / 64, WP = GJM_WP;
  constexpr int TPR = NT / NMAX, CPT = WP / TPR;
  static_assert(TPR >= 1 && CPT * TPR == WP, "panel split");
  // the pivot row; its two column slices are one complex apart beyond their
  // natural offset, so that the two slices of a wave's reads fall in
  // different LDS banks
  constexpr int SL = CPT + (TPR > 1 ? 1 : 0);
  __shared__ cx<T> sRow[2][TPR * SL];
  __shared__ unsigned sKey[2][NW];
  __shared__ int sPiv[WP];
  __shared__ int sSing;
  const InvDesc<T> d = descs[blockIdx.x];
  const int n = d.n;
  if (p0 >= n) return;                    // uniform: this block is already done
  const int64_t ld = d.ld;
  cx<T>* __restrict__ M = d.M;
  int* __restrict__ rowof = d.ipiv;
  cx<T>* __restrict__ R = d.scratch;
  int* __restrict__ roles = reinterpret_cast<int*>(d.scratch + (size_t)WP * n);
  const int w = min(WP, n - p0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = tid / TPR, h = tid % TPR;
  const bool live = r < n;
  if (tid == 0) sSing = 0;
  for (int q = tid; q < WP; q += NT) sPiv[q] = 0;
  __syncthreads();
  // every row is unused at the first panel (the scratch holds an earlier call's
  // roles); afterwards a row's role is the column it pivots, or -1
  int role = -2;
  if (live) role = p0 == 0 ? -1 : roles[r];
  int q0, wo;
  gj_outer_range(n, p0, q0, wo);
  cx<T> myipv = czero<T>();               // this row's inverse pivot, if it becomes a pivot row
  cx<T> v[CPT];
#pragma unroll
  for (int jj = 0; jj < CPT; ++jj) {
    const int j = h * CPT + jj;
    v[jj] = (live && j < w) ? M[(int64_t)(p0 + j) * ld + r] : czero<T>();
  }
  // one elimination step per panel column; the two half loops keep each
  // unrolled body small enough that the compiler unrolls it fully, so that v
  // is indexed by constants only and stays in registers
  auto step = [&](const int c) {
    if (c < w) {                          // uniform
      const int par = c & 1;
      unsigned key = 0;
      if (live && h == c / CPT && role == -1) key = pivot_key32((double)cabs1(pick<CPT>(v, c % CPT)), r);
      key = wave_max_u32(key);
      if (lane == 0) sKey[par][wave] = key;
      __syncthreads();
      // the workgroup's best key by a tree over the waves' keys
      unsigned kk[NW];
#pragma unroll
      for (int qq = 0; qq < NW; ++qq) kk[qq] = sKey[par][qq];
#pragma unroll
      for (int st = 1; st < NW; st *= 2)
#pragma unroll
        for (int i = 0; i + st < NW; i += 2 * st) kk[i] = max(kk[i], kk[i + st]);
      const unsigned best = kk[0];
      const int p = best ? 511 - (int)(best & 0x1FFu) : 0;
      if (tid == 0) sPiv[c] = p;
      if (r == p) role = p0 + c;
      // only the pivot row's lanes write it, behind a second barrier: one
      // row of stores per column instead of every wave storing its best
      // candidate before the keys are compared (measured: a 500-row FP64
      // panel 102.5 -> 74.7 us for one block, 258 -> 204 us for 300)
      if (best != 0 && r == p) {
#pragma unroll
        for (int jj = 0; jj < CPT; ++jj) sRow[par][h * SL + jj] = v[jj];
      }
      __syncthreads();
      const cx<T>* __restrict__ prow = sRow[par];
      cx<T> pv = prow[(c / CPT) * SL + c % CPT];
      T den = pv.re * pv.re + pv.im * pv.im;
      if (!(den > T(0))) {
        if (tid == 0 && sSing == 0) sSing = p0 + c + 1;
        pv = cx<T>{tiny_pivot<T>(), T(0)};
        den = pv.re * pv.re;
      }
      const T iden = T(1) / den;
      const cx<T> ipv{pv.re * iden, -pv.im * iden};
      cx<T> vc = pick<CPT>(v, c % CPT);
      if constexpr (TPR == 2) {
        if (c / CPT == 0) {
          vc.re = dpp_t<0xA0>(vc.re);
          vc.im = dpp_t<0xA0>(vc.im);
        } else {
          vc.re = dpp_t<0xF5>(vc.re);
          vc.im = dpp_t<0xF5>(vc.im);
        }
      }
      // Rows past n and columns past w hold zeros (and the pivot row's copy
      // too), so the update needs no row or column guard: zeros stay zeros.
      // Column c is held by slice c / CPT at register c % CPT.
      // The pivot row is not scaled here: its column c becomes 1 and the row
      // is multiplied by its inverse pivot once, after the panel (a later
      // step updates it like any other row, which commutes with that scale).
      // So every lane runs the same code, the pivot row with multiplier zero.
      const bool isp = r == p;
      if (isp) myipv = ipv;
      cx<T> l = cmul(vc, ipv);
      if (isp) l = czero<T>();
      const cx<T> nl{-l.re, -l.im}, sc = isp ? cx<T>{T(1), T(0)} : nl;
#pragma unroll
      for (int jj = 0; jj < CPT; ++jj) {
        const cx<T> u = prow[h * SL + jj];
        v[jj] = (jj == c % CPT && h == c / CPT) ? sc : cfma(v[jj], nl, u);   // four FMAs
      }
    }
  };
#pragma unroll
  for (int c = 0; c < WP / 2; ++c) step(c);
#pragma unroll
  for (int c = WP / 2; c < WP; ++c) step(c);
  if (role >= p0) {                       // pivot rows of this panel: the deferred scale
#pragma unroll
    for (int jj = 0; jj < CPT; ++jj) v[jj] = cmul(v[jj], myipv);
  }
  __syncthreads();
  // C -> M[:, J]; R = M[P, :] on the outer panel's columns outside J, and
  // those entries of M cleared (disjoint columns, so the order of the two
  // does not matter; the update kernel has not run yet): the update adds
  // C R to every row, which replaces the pivot rows and accumulates into the
  // others
  if (live) {
#pragma unroll
    for (int jj = 0; jj < CPT; ++jj) {
      const int j = h * CPT + jj;
      if (j < w) M[(int64_t)(p0 + j) * ld + r] = v[jj];
    }
    if (h == 0) roles[r] = role;
  }
  if (tid < w) rowof[p0 + tid] = sPiv[tid];
  constexpr int RB = 8;                   // loads in flight per thread
  for (int base = 0; base < WP * wo; base += RB * NT) {
    cx<T> x[RB];
#pragma unroll
    for (int t = 0; t < RB; ++t) {
      const int idx = base + t * NT + tid;
      const int k = idx % WP, j = q0 + idx / WP;
      x[t] = czero<T>();
      if (idx < WP * wo && k < w && (j < p0 || j >= p0 + w)) {
        cx<T>* p = M + (int64_t)j * ld + sPiv[k];
        x[t] = *p;
        *p = czero<T>();                  // the update then replaces the pivot rows
      }
    }
#pragma unroll
    for (int t = 0; t < RB; ++t) {
      const int idx = base + t * NT + tid;
      if (idx < WP * wo) R[(int64_t)q0 * WP + idx] = x[t];
    }
  }
  // info: reset by the first panel, then the first singular column wins
  if (tid == 0 && d.info && (p0 == 0 || (sSing && *d.info == 0))) *d.info = sSing;
}

// The panel's update: M[:, j] += M[:, J] R[:, j] for the columns j of its
// outer panel outside J (pivot rows cleared by the panel kernel), on the
// batched GEMM's tile with descriptors made here (columns [q0, p0), then
// [p0 + w, q0 + wo)); K = w <= 32.
template <class T>
__global__ __launch_bounds__(256, 3) void gj_update_kernel(const InvDesc<T>* __restrict__ descs, int p0) {
  __shared__ cx<T> sA[GT_K][GT_M];
  __shared__ cx<T> sB[GT_N][GT_K + GT_PADB];
  const InvDesc<T> v = descs[blockIdx.y];
  const int n = v.n;
  if (p0 >= n) return;
  int q0, wo;
  gj_outer_range(n, p0, q0, wo);
  const int w = min(GJM_WP, n - p0);
  const int n0 = p0 - q0, n1 = q0 + wo - (p0 + w);
  const int tm = (n + 63) / 64, tn0 = (n0 + 63) / 64, tn1 = (n1 + 63) / 64;
  int t = blockIdx.x;
  if (t >= tm * (tn0 + tn1)) return;      // uniform over the workgroup
  int jstart = q0, ncols = n0;
  if (t >= tm * tn0) {
    t -= tm * tn0;
    jstart = p0 + w;
    ncols = n1;
  }
  GemmDesc<T> d{};
  d.A[0] = v.M + (int64_t)p0 * v.ld;
  d.lda[0] = v.ld;
  d.k[0] = w;
  d.B[0] = v.scratch + (size_t)jstart * GJM_WP;
  d.ldb[0] = GJM_WP;
  d.A[1] = nullptr;
  d.B[1] = nullptr;
  d.k[1] = 0;
  d.C = v.M + (int64_t)jstart * v.ld;
  d.ldc = v.ld;
  d.m = n;
  d.n = ncols;
  d.opA = 0;
  d.tiles_m = tm;
  d.tiles_n = (ncols + 63) / 64;
  d.alpha = T(1);
  d.beta = T(1);
  gemm_tile_body<T>(d, t, sA, sB);
}

// phase 0: S[l n + k] = M[r_k, r^{-1}(l)];  phase 1: M <- S.   Grid: (column chunks, blocks).
// Phase 0 reads each source column whole (coalesced) into LDS and gathers
// its rows from there.
template <class T>
__global__ __launch_bounds__(256) void gj_perm_kernel(const InvDesc<T>* __restrict__ descs, int phase) {
  __shared__ int sRinv[512];
  __shared__ int sRow[512];
  __shared__ cx<T> sCol[512];
  const InvDesc<T> d = descs[blockIdx.y];
  const int n = d.n;
  constexpr int CC = 16;                                   // columns per workgroup
  const int l0 = blockIdx.x * CC;
  if (l0 >= n) return;
  const int64_t ld = d.ld;
  cx<T>* __restrict__ M = d.M;
  cx<T>* __restrict__ S = d.scratch;
  if (phase == 0) {
    for (int j = threadIdx.x; j < n; j += 256) {
      const int rj = d.ipiv[j];
      sRow[j] = rj;
      sRinv[rj] = j;
    }
    __syncthreads();
    for (int l = l0; l < min(n, l0 + CC); ++l) {
      const cx<T>* src = M + (int64_t)sRinv[l] * ld;
      for (int k = threadIdx.x; k < n; k += 256) sCol[k] = src[k];
      __syncthreads();
      for (int k = threadIdx.x; k < n; k += 256) S[(int64_t)l * n + k] = sCol[sRow[k]];
      __syncthreads();                                     // sCol is refilled for the next column
    }
  } else {
    for (int l = l0; l < min(n, l0 + CC); ++l)
      for (int k = threadIdx.x; k < n; k += 256) M[(int64_t)l * ld + k] = S[(int64_t)l * n + k];
  }
}

// End of outer panel O = [q0, q0 + wo): R_O[k, j] = M[P_O[k], j] for the
// columns j outside O (at R_O[j GJ_WO + k]), and those entries of M are
// cleared (the outer product then adds T_O[:, P_O] R_O to every row, which
// replaces the pivot rows and accumulates into the others).
template <class T>
__global__ __launch_bounds__(256) void gj_outer_gather_kernel(const InvDesc<T>* __restrict__ descs, int q0) {
  const InvDesc<T> d = descs[blockIdx.y];
  const int n = d.n;
  if (n < GJ_OUTER_MIN || q0 >= n) return;
  const int wo = min(GJ_WO, n - q0);
  const int64_t ld = d.ld;
  cx<T>* __restrict__ M = d.M;
  cx<T>* __restrict__ RO = d.scratch + (size_t)(GJM_WP + 1) * n;
  const int* __restrict__ piv = d.ipiv + q0;
  const int total = (n - wo) * GJ_WO;
  for (int idx = blockIdx.x * 256 + threadIdx.x; idx < total; idx += gridDim.x * 256) {
    const int k = idx % GJ_WO, jc = idx / GJ_WO;
    const int j = jc < q0 ? jc : jc + wo;
    cx<T> x = czero<T>();
    if (k < wo) {
      cx<T>* p = M + (int64_t)j * ld + piv[k];
      x = *p;
      *p = czero<T>();
    }
    RO[(int64_t)j * GJ_WO + k] = x;
  }
}

// M[:, j] += M[:, O] R_O[:, j] for the columns outside O: the batched GEMM's
// tile on descriptors made here (columns [0, q0), then [q0 + wo, n)).
template <class T>
__global__ __launch_bounds__(256, 3) void gj_outer_update_kernel(const InvDesc<T>* __restrict__ descs, int q0) {
  __shared__ cx<T> sA[GT_K][GT_M];
  __shared__ cx<T> sB[GT_N][GT_K + GT_PADB];
  const InvDesc<T> v = descs[blockIdx.y];
  const int n = v.n;
  if (n < GJ_OUTER_MIN || q0 >= n) return;
  const int wo = min(GJ_WO, n - q0);
  const int tm = (n + 63) / 64, tn0 = (q0 + 63) / 64, n1 = n - q0 - wo, tn1 = (n1 + 63) / 64;
  int t = blockIdx.x;
  if (t >= tm * (tn0 + tn1)) return;      // uniform over the workgroup
  int jstart = 0, ncols = q0;
  if (t >= tm * tn0) {
    t -= tm * tn0;
    jstart = q0 + wo;
    ncols = n1;
  }
  GemmDesc<T> d{};
  d.A[0] = v.M + (int64_t)q0 * v.ld;
  d.lda[0] = v.ld;
  d.k[0] = wo;
  d.B[0] = v.scratch + (size_t)(GJM_WP + 1) * n + (size_t)jstart * GJ_WO;
  d.ldb[0] = GJ_WO;
  d.A[1] = nullptr;
  d.B[1] = nullptr;
  d.k[1] = 0;
  d.C = v.M + (int64_t)jstart * v.ld;
  d.ldc = v.ld;
  d.m = n;
  d.n = ncols;
  d.opA = 0;
  d.tiles_m = tm;
  d.tiles_n = (ncols + 63) / 64;
  d.alpha = T(1);
  d.beta = T(1);
  gemm_tile_body<T>(d, t, sA, sB);
}

template <class T>
static hipError_t launch_gj_inverse_multi(const InvDesc<T>* d_descs, int ndesc, int nmax, hipStream_t s) {
  const int nt = (nmax + 63) / 64;
  for (int p0 = 0; p0 < nmax; p0 += GJM_WP) {
    if (nmax <= 256)
      hipLaunchKernelGGL((gj_panel_kernel<T, 256, GJ_PANEL_NT256<T>>), dim3(ndesc), dim3(GJ_PANEL_NT256<T>), 0, s,
                         d_descs, p0);
    else
      hipLaunchKernelGGL((gj_panel_kernel<T, 512, 512>), dim3(ndesc), dim3(512), 0, s, d_descs, p0);
    // column tiles of the two ranges an update covers: at most 3 in an outer
    // panel of 128 columns, at most nt + 1 in a block below GJ_OUTER_MIN rows
    // (at most 5 when such a block shares a launch with larger ones)
    const int ntc = nmax >= GJ_OUTER_MIN ? 5 : nt + 1;
    hipLaunchKernelGGL((gj_update_kernel<T>), dim3(nt * ntc, ndesc), dim3(256), 0, s, d_descs, p0);
    // the end of an outer panel (of every block whose outer panel holds p0)
    const int pe = p0 + GJM_WP;
    if (nmax >= GJ_OUTER_MIN && (pe % GJ_WO == 0 || pe >= nmax)) {
      const int q0 = (p0 / GJ_WO) * GJ_WO;
      const int chunks = (nmax * GJ_WO + 255) / 256 < 256 ? (nmax * GJ_WO + 255) / 256 : 256;
      hipLaunchKernelGGL((gj_outer_gather_kernel<T>), dim3(chunks, ndesc), dim3(256), 0, s, d_descs, q0);
      hipLaunchKernelGGL((gj_outer_update_kernel<T>), dim3(nt * nt, ndesc), dim3(256), 0, s, d_descs, q0);
    }
  }
  const int nc = (nmax + 15) / 16;
  hipLaunchKernelGGL((gj_perm_kernel<T>), dim3(nc, ndesc), dim3(256), 0, s, d_descs, 0);
  hipLaunchKernelGGL((gj_perm_kernel<T>), dim3(nc, ndesc), dim3(256), 0, s, d_descs, 1);
  return hipGetLastError();
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
hipError_t launch_gj_inverse(const InvDesc<T>* d_descs, int ndesc, int nmin, int nmax, hipStream_t s) {
  if (ndesc <= 0) return hipSuccess;
  // the multi-workgroup form keeps R (32 x n) and the row roles in each block's
  // n x n scratch, which needs n > 32
  if (nmax >= GJ_MULTI_MIN && nmax <= 512 && nmin > GJM_WP)
    return launch_gj_inverse_multi<T>(d_descs, ndesc, nmax, s);
  if constexpr (sizeof(T) == 8) {
    if (nmax <= 256)
      hipLaunchKernelGGL((gj_inverse_kernel<T, 32, 256>), dim3(ndesc), dim3(512), 0, s, d_descs);
    else if (nmax <= 512)
      hipLaunchKernelGGL((gj_inverse_kernel<T, 16, 512>), dim3(ndesc), dim3(512), 0, s, d_descs);
    else
      return hipErrorInvalidValue;
  } else {
    if (nmax <= 256)
      hipLaunchKernelGGL((gj_inverse_kernel<T, 32, 256>), dim3(ndesc), dim3(512), 0, s, d_descs);
    else if (nmax <= 512)
      hipLaunchKernelGGL((gj_inverse_kernel<T, 16, 512>), dim3(ndesc), dim3(512), 0, s, d_descs);
    else
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template hipError_t launch_gemm_batched<double>(const GemmDesc<double>*, int, int, hipStream_t);
template hipError_t launch_gemm_batched<float>(const GemmDesc<float>*, int, int, hipStream_t);
template hipError_t launch_gj_inverse<double>(const InvDesc<double>*, int, int, int, hipStream_t);
template hipError_t launch_gj_inverse<float>(const InvDesc<float>*, int, int, int, hipStream_t);

}  // namespace kore
