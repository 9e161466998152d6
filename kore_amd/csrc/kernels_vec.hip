// Bandwidth-bound kernels of the Kore solve path (gfx950).
//
//   spmv          (ca A + cb B) X on the union CSR pattern, 16 lanes per row
//   row_scale     power-of-two row equilibration of A - sigma B
//   build_blocks  scatter R(A - sigma B) into the dense level-0 blocks
//   skinny        batched Y = beta Y + alpha (A1 X1 + A2 X2), nrhs <= 16,
//                 the block sweeps of the cyclic-reduction solve
//   gather/scatter/axpy/norms/zero/copy   vector plumbing
#include <type_traits>
#include "kore_device.hpp"
#include "kore_kernels.hpp"

namespace kore {

// ---------------------------------------------------------------------------
// CSR SpMV, 16 lanes per row
// ---------------------------------------------------------------------------
template <int NR>
__global__ __launch_bounds__(256) void spmv_kernel(CsrDev m, cx<double> ca, cx<double> cb,
                                                   const cx<double>* __restrict__ X, int64_t ldx,
                                                   cx<double>* __restrict__ Y, int64_t ldy, int nrhs,
                                                   double beta, const double* __restrict__ rscale) {
  const int g = threadIdx.x >> 4, gl = threadIdx.x & 15;
  const int64_t row = (int64_t)blockIdx.x * 16 + g;
  if (row >= m.n) return;
  const int s = m.indptr[row], e = m.indptr[row + 1];
  cx<double> acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = czero<double>();
  // a zero coefficient skips its value array (B x and A x read one of the two)
  const bool use_a = ca.re != 0.0 || ca.im != 0.0, use_b = cb.re != 0.0 || cb.im != 0.0;
  for (int k = s + gl; k < e; k += 16) {
    const int j = m.indices[k];
    cx<double> v = czero<double>();
    if (use_a) v = cmul(ca, m.va[k]);
    if (use_b) v = cadd(v, cmul(cb, m.vb[k]));
#pragma unroll
    for (int c = 0; c < NR; ++c)
      if (c < nrhs) acc[c] = cfma(acc[c], v, X[(int64_t)c * ldx + j]);
  }
#pragma unroll
  for (int c = 0; c < NR; ++c) {
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) {
      acc[c].re += __shfl_xor(acc[c].re, off, 64);
      acc[c].im += __shfl_xor(acc[c].im, off, 64);
    }
  }
  if (gl == 0) {
    const double sc = rscale ? rscale[row] : 1.0;   // optional row scale of the product
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (c < nrhs) {
        cx<double>* y = Y + (int64_t)c * ldy + row;
        cx<double> v{sc * acc[c].re, sc * acc[c].im};
        if (beta != 0.0) {
          const cx<double> o = *y;
          v.re += beta * o.re;
          v.im += beta * o.im;
        }
        *y = v;
      }
    }
  }
}

hipError_t launch_spmv(const CsrDev& m, double ca_re, double ca_im, double cb_re, double cb_im,
                       const cx<double>* X, int64_t ldx, cx<double>* Y, int64_t ldy, int nrhs,
                       double beta, hipStream_t s, const double* rscale) {
  if (m.n == 0) return hipSuccess;
  const cx<double> ca{ca_re, ca_im}, cb{cb_re, cb_im};
  const dim3 grid((unsigned)((m.n + 15) / 16));
  for (int c0 = 0; c0 < nrhs; c0 += 16) {
    const int nr = nrhs - c0 < 16 ? nrhs - c0 : 16;
    const cx<double>* x = X + (int64_t)c0 * ldx;
    cx<double>* y = Y + (int64_t)c0 * ldy;
    if (nr == 1)
      hipLaunchKernelGGL(spmv_kernel<1>, grid, dim3(256), 0, s, m, ca, cb, x, ldx, y, ldy, nr, beta, rscale);
    else if (nr <= 4)
      hipLaunchKernelGGL(spmv_kernel<4>, grid, dim3(256), 0, s, m, ca, cb, x, ldx, y, ldy, nr, beta, rscale);
    else
      hipLaunchKernelGGL(spmv_kernel<16>, grid, dim3(256), 0, s, m, ca, cb, x, ldx, y, ldy, nr, beta, rscale);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// level-0 couplings of the solve sweep, from the sparse matrix
// ---------------------------------------------------------------------------
// For the rows of the blocks of one parity:
//   y[i] -= scale[i] * sum over entries k of row i outside its diagonal block of
//           (va[k] - cs vb[k]) v[col k],
// i.e. y_t -= L_t v_{t-1} + U_t v_{t+1} with L, U the off-diagonal blocks of
// the row-equilibrated level-0 system R (A - cs B).  16 lanes per row.
__global__ __launch_bounds__(256) void offdiag_kernel(CsrDev m, cx<double> cs, const double* __restrict__ scale,
                                                      const int32_t* __restrict__ split, int parity,
                                                      const cx<double>* __restrict__ v, cx<double>* __restrict__ y) {
  const int g = threadIdx.x >> 4, gl = threadIdx.x & 15;
  const int64_t row = (int64_t)blockIdx.x * 16 + g;
  if (row >= m.n) return;
  const int32_t sp = split[row];
  if (((sp >> 30) & 1) != parity) return;          // uniform over the row's 16 lanes
  const int nlo = sp & 0x7FFF, nhi = (sp >> 15) & 0x7FFF;
  const int s = m.indptr[row], e = m.indptr[row + 1];
  // the row's entries are sorted by column: those of the neighbouring blocks
  // are its first nlo and last nhi
  cx<double> acc = czero<double>();
  for (int q = gl; q < nlo + nhi; q += 16) {
    const int k = q < nlo ? s + q : e - nhi + (q - nlo);
    acc = cfma(acc, csub(m.va[k], cmul(cs, m.vb[k])), v[m.indices[k]]);
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) {
    acc.re += __shfl_xor(acc.re, off, 64);
    acc.im += __shfl_xor(acc.im, off, 64);
  }
  if (gl == 0) {
    const double sc = scale[row];
    const cx<double> o = y[row];
    y[row] = cx<double>{o.re - sc * acc.re, o.im - sc * acc.im};
  }
}

hipError_t launch_offdiag(const CsrDev& m, double sre, double sim, const double* scale, const int32_t* split,
                          int parity, const cx<double>* v, cx<double>* y, hipStream_t s) {
  if (m.n == 0) return hipSuccess;
  hipLaunchKernelGGL(offdiag_kernel, dim3((unsigned)((m.n + 15) / 16)), dim3(256), 0, s, m, cx<double>{sre, sim},
                     scale, split, parity, v, y);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// row equilibration
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void row_scale_kernel(CsrDev m, cx<double> cs, double* __restrict__ scale) {
  const int g = threadIdx.x >> 4, gl = threadIdx.x & 15;
  const int64_t row = (int64_t)blockIdx.x * 16 + g;
  if (row >= m.n) return;
  const int s = m.indptr[row], e = m.indptr[row + 1];
  double mx = 0.0;
  for (int k = s + gl; k < e; k += 16) {
    const cx<double> v = csub(m.va[k], cmul(cs, m.vb[k]));
    mx = fmax(mx, sqrt(v.re * v.re + v.im * v.im));
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if (gl == 0) {
    double sc = 1.0;
    if (mx > 0.0 && isfinite(mx)) sc = ldexp(1.0, -(int)rint(log2(mx)));
    scale[row] = sc;
  }
}

hipError_t launch_row_scale(const CsrDev& m, double sre, double sim, double* scale, hipStream_t s) {
  if (m.n == 0) return hipSuccess;
  hipLaunchKernelGGL(row_scale_kernel, dim3((unsigned)((m.n + 15) / 16)), dim3(256), 0, s, m,
                     cx<double>{sre, sim}, scale);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// level-0 block scatter
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void build_blocks_kernel(CsrDev m, cx<double> cs, const double* __restrict__ scale,
                                                           const int32_t* __restrict__ blk_of_row,
                                                           const int64_t* __restrict__ blk_off,
                                                           cx<T>* const* __restrict__ Dp,
                                                           cx<T>* const* __restrict__ Lp,
                                                           cx<T>* const* __restrict__ Up) {
  const int g = threadIdx.x >> 4, gl = threadIdx.x & 15;
  const int64_t row = (int64_t)blockIdx.x * 16 + g;
  if (row >= m.n) return;
  const int kb = blk_of_row[row];
  const int64_t ro = blk_off[kb];
  const int64_t nk = blk_off[kb + 1] - ro;
  const int64_t li = row - ro;
  const double sc = scale[row];
  const int s = m.indptr[row], e = m.indptr[row + 1];
  for (int k = s + gl; k < e; k += 16) {
    const int j = m.indices[k];
    const int kc = blk_of_row[j];
    const int64_t lj = j - blk_off[kc];
    const cx<double> v = csub(m.va[k], cmul(cs, m.vb[k]));
    const cx<T> w{(T)(sc * v.re), (T)(sc * v.im)};
    cx<T>* dst = nullptr;
    if (kc == kb) dst = Dp[kb];
    else if (kc == kb - 1) dst = Lp[kb];
    else if (kc == kb + 1) dst = Up[kb];
    if (dst) dst[lj * nk + li] = w;
  }
}

template <class T>
hipError_t launch_build_blocks(const CsrDev& m, double sre, double sim, const double* scale,
                               const int32_t* blk_of_row, const int64_t* blk_off, cx<T>* const* Dp,
                               cx<T>* const* Lp, cx<T>* const* Up, hipStream_t s) {
  if (m.n == 0) return hipSuccess;
  hipLaunchKernelGGL(build_blocks_kernel<T>, dim3((unsigned)((m.n + 15) / 16)), dim3(256), 0, s, m,
                     cx<double>{sre, sim}, scale, blk_of_row, blk_off, Dp, Lp, Up);
  return hipGetLastError();
}
template hipError_t launch_build_blocks<double>(const CsrDev&, double, double, const double*, const int32_t*,
                                                const int64_t*, cx<double>* const*, cx<double>* const*,
                                                cx<double>* const*, hipStream_t);
template hipError_t launch_build_blocks<float>(const CsrDev&, double, double, const double*, const int32_t*,
                                               const int64_t*, cx<float>* const*, cx<float>* const*,
                                               cx<float>* const*, hipStream_t);

// ---------------------------------------------------------------------------
// batched skinny products for the solve sweeps
// ---------------------------------------------------------------------------
// Workgroup: 16 rows x 16 k-slices.  A wave covers 16 consecutive rows of 4
// k-columns, so its A loads are four 256-byte segments; the X entries are LDS
// broadcasts (the 16 lanes of a k-slice share k).  Short k-ranges per thread
// keep the dependent load chains short, which matters for the deep reduction
// levels where a launch holds only a handful of blocks.
constexpr int SK_ROWS = 16, SK_KS = 16;
template <int NR> struct SkChunk { static constexpr int value = NR >= 8 ? 128 : (NR >= 4 ? 256 : 512); };
template <int NR> struct SkLds {
  static constexpr int x = SkChunk<NR>::value * NR;           // staged X chunk
  static constexpr int p = (SK_KS - 1) * SK_ROWS * NR;         // partial sums of k-slices 1..15
  static constexpr int value = x > p ? x : p;
};

// PRE (the batched sweep kernel; the single-product kernel streams): a product
// with K <= min(SkPreK, X chunk) -- 512 for one or two right-hand sides, 256
// otherwise -- has all of this thread's A entries and its share of the X chunk
// requested before the first wait, so the product costs one memory round trip
// instead of one per unrolled group.
constexpr int SK_PRE_K = 256;
template <int NR> struct SkPreK { static constexpr int value = NR <= 2 ? 512 : SK_PRE_K; };   // k entries preloaded

// PRE with one or two right-hand sides and every product's K <= 256 (the
// sweeps of blocks up to 256 wide): the A entries and X chunks of all (up to
// three) products are requested before the first wait, so a launch costs one
// memory round trip instead of one per product.  Products are accumulated in
// the same order, with the same alpha-scaled X values, as the per-product form.
// NP: products held in registers (3; 1 for launches whose descriptors hold one).
template <class T, int NR, int NP>
__device__ __forceinline__ void skinny_fused_pre(const SkinnyDesc<T>& d, int nrhs, int row, int ks, cx<T> (&acc)[NR]) {
  constexpr int NA = SK_PRE_K / SK_KS;               // A entries per thread and product
  constexpr int NXL = (SK_PRE_K * NR + 255) / 256;   // X entries per thread and product
  __shared__ cx<T> sXf[NP][SK_PRE_K][NR];
  cx<T> xr[NP][NXL];
  cx<T> a[NP][NA];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const cx<T>* __restrict__ X = d.X[p];
    const int K = d.A[p] != nullptr ? d.k[p] : 0;
    const int64_t ldx = d.ldx[p];
#pragma unroll
    for (int t = 0; t < NXL; ++t) {
      const int idx = threadIdx.x + 256 * t;
      const int kk = idx / NR, c = idx - kk * NR;
      xr[p][t] = czero<T>();
      if (idx < SK_PRE_K * NR && kk < K && c < nrhs) xr[p][t] = gld(X + (int64_t)c * ldx + kk);
    }
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int K = d.A[p] != nullptr ? d.k[p] : 0;
    const cx<T>* __restrict__ Ac = d.A[p] + row;
    const int64_t lda = d.lda[p];
#pragma unroll
    for (int t = 0; t < NA; ++t) {
      const int kk = ks + SK_KS * t;
      a[p][t] = czero<T>();
      if (row < d.m && kk < K) a[p][t] = gld(Ac + (int64_t)kk * lda);
    }
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const T al = d.alpha[p];
#pragma unroll
    for (int t = 0; t < NXL; ++t) {
      const int idx = threadIdx.x + 256 * t;
      if (idx < SK_PRE_K * NR) {
        const int kk = idx / NR, c = idx - kk * NR;
        sXf[p][kk][c] = cx<T>{xr[p][t].re * al, xr[p][t].im * al};
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    if (d.A[p] == nullptr || d.k[p] <= 0) continue;   // uniform
#pragma unroll
    for (int t = 0; t < NA; ++t) {
      const int kk = ks + SK_KS * t;                  // entries past K are zero on both sides
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = cfma(acc[c], a[p][t], sXf[p][kk][c]);
    }
  }
}

template <class T, int NR, bool PRE, int NP = 3>
__device__ __forceinline__ void skinny_tile(const SkinnyDesc<T>& d, int nrhs) {
  constexpr int KC = SkChunk<NR>::value;
  const int r0 = blockIdx.x * SK_ROWS;
  if (r0 >= d.m) return;
  const int rl = threadIdx.x & (SK_ROWS - 1), ks = threadIdx.x / SK_ROWS;
  const int row = r0 + rl;
  // one LDS region: the X chunk during the sweep, then the slice partial sums
  __shared__ cx<T> lds[SkLds<NR>::value];
  cx<T>(*sX)[NR] = reinterpret_cast<cx<T>(*)[NR]>(lds);
  cx<T>(*sP)[SK_ROWS][NR] = reinterpret_cast<cx<T>(*)[SK_ROWS][NR]>(lds);
  cx<T> acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = czero<T>();
  bool fused = false;
  if constexpr (PRE && NR <= 2) {
    fused = true;                                       // uniform over the workgroup
    for (int p = 0; p < 3; ++p)
      if (d.A[p] != nullptr && (d.k[p] > SK_PRE_K || p >= NP)) fused = false;
    if (fused) skinny_fused_pre<T, NR, NP>(d, nrhs, row, ks, acc);
  }
  for (int p = 0; p < 3 && !fused; ++p) {
    const cx<T>* __restrict__ A = d.A[p];
    const cx<T>* __restrict__ X = d.X[p];
    const int K = d.k[p];
    if (A == nullptr || K <= 0) continue;
    const int64_t lda = d.lda[p], ldx = d.ldx[p];
    const T al = d.alpha[p];
    if constexpr (PRE) {
      // k entries staged at once (launches of fewer than three products keep
      // to the short form, so that their register count stays low; longer
      // products stream)
      constexpr int KPP = NP < 3 ? SK_PRE_K : SkPreK<NR>::value;
      constexpr int KP = KPP < KC ? KPP : KC;
      // the product, with NA A entries per thread (NA * 16 >= K)
      auto product = [&](auto na) {
        constexpr int NA = decltype(na)::value;
        constexpr int NXL = (KC * NR + 255) / 256;       // X entries per thread
        cx<T> xr[NXL];
#pragma unroll
        for (int t = 0; t < NXL; ++t) {
          const int idx = threadIdx.x + 256 * t;
          const int kk = idx / NR, c = idx - kk * NR;
          xr[t] = czero<T>();
          if (idx < KC * NR && kk < K && c < nrhs) xr[t] = gld(X + (int64_t)c * ldx + kk);
        }
        cx<T> a[NA];
        const cx<T>* __restrict__ Ac = A + row;
#pragma unroll
        for (int t = 0; t < NA; ++t) {
          const int kk = ks + SK_KS * t;
          a[t] = czero<T>();
          if (row < d.m && kk < K) a[t] = gld(Ac + (int64_t)kk * lda);
        }
        __syncthreads();                  // earlier readers of the X chunk are done
#pragma unroll
        for (int t = 0; t < NXL; ++t) {
          const int idx = threadIdx.x + 256 * t;
          if (idx < KC * NR) {
            const int kk = idx / NR, c = idx - kk * NR;
            sX[kk][c] = cx<T>{xr[t].re * al, xr[t].im * al};
          }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NA; ++t) {
          const int kk = ks + SK_KS * t;  // < KP <= KC; entries past K are zero on both sides
#pragma unroll
          for (int c = 0; c < NR; ++c) acc[c] = cfma(acc[c], a[t], sX[kk][c]);
        }
      };
      if (K <= SK_PRE_K && SK_PRE_K <= KP) {          // uniform: the short form when it suffices
        product(std::integral_constant<int, SK_PRE_K / SK_KS>{});
        continue;
      }
      if (K <= KP) {
        product(std::integral_constant<int, KP / SK_KS>{});
        continue;
      }
    }
    for (int kc0 = 0; kc0 < K; kc0 += KC) {
      __syncthreads();
      for (int idx = threadIdx.x; idx < KC * NR; idx += 256) {
        const int kk = idx / NR, c = idx - kk * NR;
        cx<T> v = czero<T>();
        if (kc0 + kk < K && c < nrhs) {
          v = gld(X + (int64_t)c * ldx + kc0 + kk);
          v.re *= al;
          v.im *= al;
        }
        sX[kk][c] = v;
      }
      __syncthreads();
      const int kend = (K - kc0) < KC ? (K - kc0) : KC;
      if (row < d.m) {
        const cx<T>* __restrict__ Ac = A + (int64_t)kc0 * lda + row;
        int kk = ks;
        for (; kk + 3 * SK_KS < kend; kk += 4 * SK_KS) {
          const cx<T> a0 = gld(Ac + (int64_t)kk * lda);
          const cx<T> a1 = gld(Ac + (int64_t)(kk + SK_KS) * lda);
          const cx<T> a2 = gld(Ac + (int64_t)(kk + 2 * SK_KS) * lda);
          const cx<T> a3 = gld(Ac + (int64_t)(kk + 3 * SK_KS) * lda);
#pragma unroll
          for (int c = 0; c < NR; ++c) {
            acc[c] = cfma(acc[c], a0, sX[kk][c]);
            acc[c] = cfma(acc[c], a1, sX[kk + SK_KS][c]);
            acc[c] = cfma(acc[c], a2, sX[kk + 2 * SK_KS][c]);
            acc[c] = cfma(acc[c], a3, sX[kk + 3 * SK_KS][c]);
          }
        }
        for (; kk < kend; kk += SK_KS) {
          const cx<T> a = gld(Ac + (int64_t)kk * lda);
#pragma unroll
          for (int c = 0; c < NR; ++c) acc[c] = cfma(acc[c], a, sX[kk][c]);
        }
      }
    }
  }
  __syncthreads();
  if (ks > 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) sP[ks - 1][rl][c] = acc[c];
  }
  __syncthreads();
  if (ks == 0 && row < d.m) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (c < nrhs) {
        cx<T> v = acc[c];
#pragma unroll
        for (int q = 0; q < SK_KS - 1; ++q) v = cadd(v, sP[q][rl][c]);
        cx<T>* y = d.Y + (int64_t)c * d.ldy + row;
        cx<T> o = v;
        if (d.beta != T(0)) {
          const cx<T> old = gld(y);
          o.re += d.beta * old.re;
          o.im += d.beta * old.im;
        }
        gst(y, o);
      }
    }
  }
}

template <class T, int NR>
__global__ __launch_bounds__(256) void skinny_kernel(const SkinnyDesc<T>* __restrict__ descs, int nrhs) {
  skinny_tile<T, NR, true>(descs[blockIdx.y], nrhs);   // descriptor read in place (uniform loads, no private copy)
}

// one right-hand side and at most NP products per descriptor, blocks up to
// 256 wide: registers for NP products only, so more workgroups per CU keep
// loads in flight (one product: a third of the three-product form; only
// NP = 1 is launched)
template <class T, int NP>
__global__ __launch_bounds__(256) void skinny_np_kernel(const SkinnyDesc<T>* __restrict__ descs) {
  skinny_tile<T, 1, true, NP>(descs[blockIdx.y], 1);
}

template <class T>
hipError_t launch_skinny_np(const SkinnyDesc<T>* d_descs, int ndesc, int max_m, int np, hipStream_t s) {
  if (ndesc <= 0 || max_m <= 0) return hipSuccess;
  const dim3 grid((unsigned)((max_m + SK_ROWS - 1) / SK_ROWS), (unsigned)ndesc);
  if (np == 1)
    hipLaunchKernelGGL((skinny_np_kernel<T, 1>), grid, dim3(256), 0, s, d_descs);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
template hipError_t launch_skinny_np<double>(const SkinnyDesc<double>*, int, int, int, hipStream_t);

// one product, descriptor passed by value (no descriptor upload)
template <class T, int NR>
__global__ __launch_bounds__(256) void skinny1_kernel(const SkinnyDesc<T> d, int nrhs) {
  skinny_tile<T, NR, false>(d, nrhs);
}

template <class T>
hipError_t launch_skinny1(const SkinnyDesc<T>& d, int nrhs, hipStream_t s) {
  if (d.m <= 0) return hipSuccess;
  if (nrhs < 1 || nrhs > 16) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((d.m + SK_ROWS - 1) / SK_ROWS), 1);
  if (nrhs == 1)
    hipLaunchKernelGGL((skinny1_kernel<T, 1>), grid, dim3(256), 0, s, d, nrhs);
  else if (nrhs <= 2)
    hipLaunchKernelGGL((skinny1_kernel<T, 2>), grid, dim3(256), 0, s, d, nrhs);
  else if (nrhs <= 4)
    hipLaunchKernelGGL((skinny1_kernel<T, 4>), grid, dim3(256), 0, s, d, nrhs);
  else if (nrhs <= 8)
    hipLaunchKernelGGL((skinny1_kernel<T, 8>), grid, dim3(256), 0, s, d, nrhs);
  else
    hipLaunchKernelGGL((skinny1_kernel<T, 16>), grid, dim3(256), 0, s, d, nrhs);
  return hipGetLastError();
}
template hipError_t launch_skinny1<double>(const SkinnyDesc<double>&, int, hipStream_t);

template <class T>
hipError_t launch_skinny(const SkinnyDesc<T>* d_descs, int ndesc, int max_m, int nrhs, hipStream_t s) {
  if (ndesc <= 0 || max_m <= 0) return hipSuccess;
  if (nrhs < 1 || nrhs > 16) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((max_m + SK_ROWS - 1) / SK_ROWS), (unsigned)ndesc);
  // the preloading form at every size: measured no slower than the streaming
  // loop on the largest sweep levels (thousands of workgroups) and faster on
  // the deep levels, where a launch holds a handful of blocks
  if (nrhs == 1)
    hipLaunchKernelGGL((skinny_kernel<T, 1>), grid, dim3(256), 0, s, d_descs, nrhs);
  else if (nrhs <= 2)
    hipLaunchKernelGGL((skinny_kernel<T, 2>), grid, dim3(256), 0, s, d_descs, nrhs);
  else if (nrhs <= 4)
    hipLaunchKernelGGL((skinny_kernel<T, 4>), grid, dim3(256), 0, s, d_descs, nrhs);
  else if (nrhs <= 8)
    hipLaunchKernelGGL((skinny_kernel<T, 8>), grid, dim3(256), 0, s, d_descs, nrhs);
  else
    hipLaunchKernelGGL((skinny_kernel<T, 16>), grid, dim3(256), 0, s, d_descs, nrhs);
  return hipGetLastError();
}
template hipError_t launch_skinny<double>(const SkinnyDesc<double>*, int, int, int, hipStream_t);
template hipError_t launch_skinny<float>(const SkinnyDesc<float>*, int, int, int, hipStream_t);

// ---------------------------------------------------------------------------
// vector plumbing
// ---------------------------------------------------------------------------
template <class Tin, class Tout>
__global__ void gather_scale_kernel(const cx<Tin>* __restrict__ x, int64_t ldx, cx<Tout>* __restrict__ y,
                                    int64_t ldy, int64_t n, const int64_t* __restrict__ perm,
                                    const double* __restrict__ scale, double mult) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n) return;
  const int64_t src = perm ? perm[i] : i;
  const cx<Tin> v = x[(int64_t)c * ldx + src];
  const double f = mult * (scale ? scale[i] : 1.0);
  y[(int64_t)c * ldy + i] = cx<Tout>{(Tout)(f * (double)v.re), (Tout)(f * (double)v.im)};
}

template <class Tin, class Tout>
hipError_t launch_gather_scale(const cx<Tin>* x, int64_t ldx, cx<Tout>* y, int64_t ldy, int64_t n, int nrhs,
                               const int64_t* perm, const double* scale, double mult, hipStream_t s) {
  if (n == 0 || nrhs == 0) return hipSuccess;
  hipLaunchKernelGGL((gather_scale_kernel<Tin, Tout>), dim3((unsigned)((n + 255) / 256), nrhs), dim3(256), 0, s,
                     x, ldx, y, ldy, n, perm, scale, mult);
  return hipGetLastError();
}
template hipError_t launch_gather_scale<double, double>(const cx<double>*, int64_t, cx<double>*, int64_t, int64_t,
                                                        int, const int64_t*, const double*, double, hipStream_t);
template hipError_t launch_gather_scale<double, float>(const cx<double>*, int64_t, cx<float>*, int64_t, int64_t,
                                                       int, const int64_t*, const double*, double, hipStream_t);
template hipError_t launch_gather_scale<float, double>(const cx<float>*, int64_t, cx<double>*, int64_t, int64_t,
                                                       int, const int64_t*, const double*, double, hipStream_t);

__global__ void scatter_kernel(const cx<double>* __restrict__ x, int64_t ldx, cx<double>* __restrict__ y,
                               int64_t ldy, int64_t n, const int64_t* __restrict__ perm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n) return;
  y[(int64_t)c * ldy + perm[i]] = x[(int64_t)c * ldx + i];
}

hipError_t launch_scatter(const cx<double>* x, int64_t ldx, cx<double>* y, int64_t ldy, int64_t n, int nrhs,
                          const int64_t* perm, hipStream_t s) {
  if (n == 0 || nrhs == 0) return hipSuccess;
  hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((n + 255) / 256), nrhs), dim3(256), 0, s, x, ldx, y, ldy, n,
                     perm);
  return hipGetLastError();
}

template <class Td>
__global__ void axpy_kernel(cx<double>* __restrict__ x, int64_t ldx, const cx<Td>* __restrict__ d, int64_t ldd,
                            int64_t n, double mult) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n) return;
  const cx<Td> v = d[(int64_t)c * ldd + i];
  cx<double>& o = x[(int64_t)c * ldx + i];
  o.re += mult * (double)v.re;
  o.im += mult * (double)v.im;
}

template <class Td>
hipError_t launch_axpy(cx<double>* x, int64_t ldx, const cx<Td>* d, int64_t ldd, int64_t n, int nrhs, double mult,
                       hipStream_t s) {
  if (n == 0 || nrhs == 0) return hipSuccess;
  hipLaunchKernelGGL((axpy_kernel<Td>), dim3((unsigned)((n + 255) / 256), nrhs), dim3(256), 0, s, x, ldx, d, ldd,
                     n, mult);
  return hipGetLastError();
}
template hipError_t launch_axpy<double>(cx<double>*, int64_t, const cx<double>*, int64_t, int64_t, int, double,
                                        hipStream_t);
template hipError_t launch_axpy<float>(cx<double>*, int64_t, const cx<float>*, int64_t, int64_t, int, double,
                                       hipStream_t);

template <class T, bool MAXABS>
__global__ __launch_bounds__(256) void colred_kernel(const cx<T>* __restrict__ x, int64_t ldx, int64_t n,
                                                     double* __restrict__ partial, int nrhs) {
  __shared__ double red[4];
  const int c = blockIdx.y;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    const cx<T> v = x[(int64_t)c * ldx + i];
    const double re = (double)v.re, im = (double)v.im;
    if (MAXABS)
      acc = fmax(acc, sqrt(re * re + im * im));
    else
      acc += re * re + im * im;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(acc, off, 64);
    acc = MAXABS ? fmax(acc, o) : acc + o;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int q = 1; q < 4; ++q) r = MAXABS ? fmax(r, red[q]) : r + red[q];
    partial[(int64_t)blockIdx.x * nrhs + c] = r;
  }
}

template <class T>
hipError_t launch_colnorm2(const cx<T>* x, int64_t ldx, int64_t n, int nrhs, double* partial, int nblocks,
                           hipStream_t s) {
  if (nrhs == 0) return hipSuccess;
  hipLaunchKernelGGL((colred_kernel<T, false>), dim3(nblocks, nrhs), dim3(256), 0, s, x, ldx, n, partial, nrhs);
  return hipGetLastError();
}
template hipError_t launch_colnorm2<double>(const cx<double>*, int64_t, int64_t, int, double*, int, hipStream_t);
template hipError_t launch_colnorm2<float>(const cx<float>*, int64_t, int64_t, int, double*, int, hipStream_t);

hipError_t launch_colmaxabs(const cx<double>* x, int64_t ldx, int64_t n, int nrhs, double* partial, int nblocks,
                            hipStream_t s) {
  if (nrhs == 0) return hipSuccess;
  hipLaunchKernelGGL((colred_kernel<double, true>), dim3(nblocks, nrhs), dim3(256), 0, s, x, ldx, n, partial,
                     nrhs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// tall-skinny inner products  P[chunk] = V[rows, i0:i0+64]^H W[rows, 0:nb]
// ---------------------------------------------------------------------------
// Workgroup = 64 V columns x 4 row slices; 32-row tiles of V are staged
// through LDS with coalesced loads (a column's rows are contiguous), W tiles
// likewise; each thread keeps nb <= NB complex accumulators.  The chunk
// partials are summed by sum_parts_kernel.
template <int NB>
__global__ __launch_bounds__(256) void tsdot_kernel(const cx<double>* __restrict__ V, int64_t ldv, int na,
                                                    const cx<double>* __restrict__ W, int64_t ldw, int nb,
                                                    int64_t n, int64_t rows_per_chunk,
                                                    cx<double>* __restrict__ partial) {
  constexpr int TR = 32;
  constexpr int NV = TR * (64 + 1), NR3 = 3 * 64 * NB;
  // the V tile and, after the sweep, the slice partial sums share one region
  __shared__ cx<double> lds[NV > NR3 ? NV : NR3];
  __shared__ cx<double> sWt[TR][NB];
  cx<double>(*sV)[64 + 1] = reinterpret_cast<cx<double>(*)[64 + 1]>(lds);
  cx<double>(*sRed)[64][NB] = reinterpret_cast<cx<double>(*)[64][NB]>(lds);
  const int tid = threadIdx.x;
  const int ci = tid & 63, rs = tid >> 6;
  const int i0 = blockIdx.y * 64;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
  const int64_t r1 = r0 + rows_per_chunk < n ? r0 + rows_per_chunk : n;
  cx<double> acc[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) acc[j] = czero<double>();
  for (int64_t rt = r0; rt < r1; rt += TR) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < (TR * 64) / 256; ++q) {
      const int idx = tid + 256 * q;
      const int col = idx / TR, row = idx - col * TR;
      cx<double> v = czero<double>();
      if (rt + row < r1 && i0 + col < na) v = V[(int64_t)(i0 + col) * ldv + rt + row];
      sV[row][col] = v;
    }
    for (int idx = tid; idx < TR * NB; idx += 256) {
      const int col = idx / TR, row = idx - col * TR;
      cx<double> v = czero<double>();
      if (rt + row < r1 && col < nb) v = W[(int64_t)col * ldw + rt + row];
      sWt[row][col] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = rs; r < TR; r += 4) {
      const cx<double> v = sV[r][ci];
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[j] = cfma_conj(acc[j], v, sWt[r][j]);
    }
  }
  __syncthreads();
  if (rs > 0) {
#pragma unroll
    for (int j = 0; j < NB; ++j) sRed[rs - 1][ci][j] = acc[j];
  }
  __syncthreads();
  if (rs == 0 && i0 + ci < na) {
    cx<double>* out = partial + (int64_t)blockIdx.x * na * nb;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < nb) {
        cx<double> t = acc[j];
        t = cadd(t, sRed[0][ci][j]);
        t = cadd(t, sRed[1][ci][j]);
        t = cadd(t, sRed[2][ci][j]);
        out[(int64_t)j * na + i0 + ci] = t;   // column-major na x nb
      }
    }
  }
}

hipError_t launch_tsdot(const cx<double>* V, int64_t ldv, int na, const cx<double>* W, int64_t ldw, int nb,
                        int64_t n, int nchunk, int64_t rows_per_chunk, cx<double>* partial, hipStream_t s) {
  if (na <= 0 || nb <= 0) return hipSuccess;
  if (nb > 16) return hipErrorInvalidValue;
  const dim3 grid((unsigned)nchunk, (unsigned)((na + 63) / 64));
  if (nb == 1)        // one column (single-vector Krylov): a quarter of the accumulators, twice the occupancy
    hipLaunchKernelGGL(tsdot_kernel<1>, grid, dim3(256), 0, s, V, ldv, na, W, ldw, nb, n, rows_per_chunk, partial);
  else if (nb <= 4)
    hipLaunchKernelGGL(tsdot_kernel<4>, grid, dim3(256), 0, s, V, ldv, na, W, ldw, nb, n, rows_per_chunk, partial);
  else if (nb <= 8)
    hipLaunchKernelGGL(tsdot_kernel<8>, grid, dim3(256), 0, s, V, ldv, na, W, ldw, nb, n, rows_per_chunk, partial);
  else
    hipLaunchKernelGGL(tsdot_kernel<16>, grid, dim3(256), 0, s, V, ldv, na, W, ldw, nb, n, rows_per_chunk, partial);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// one Gram-Schmidt update fused with the inner products that follow it
//   W <- W - V H   (W: n x NBW, V: n x na, H: na x NBW, column-major, ld ldh)
// then per row chunk the partials of  V^H W  (MODE 0: the next pass's
// coefficients, na x NBW) or  W^H W  (MODE 1: the block's Gram matrix,
// NBW x NBW), in tsdot's layout, for sum_parts_kernel / normalize1_kernel.
// ---------------------------------------------------------------------------
// One workgroup per row chunk and all na columns: each tile of TR rows of V is
// staged through LDS once and serves both the update (TPR threads per row, a
// shuffle reduction over them) and the products (one thread per column, four
// row slices), so a pass reads the basis once where an update launch and a
// tsdot launch read it twice.
template <int NBW, int NAMAX, int MODE>
__global__ __launch_bounds__(256) void ortho_step_kernel(const cx<double>* __restrict__ V, int64_t ldv, int na,
                                                         cx<double>* __restrict__ W, int64_t ldw,
                                                         const cx<double>* __restrict__ H, int ldh, int64_t n,
                                                         int64_t rows_per_chunk, cx<double>* __restrict__ partial) {
  constexpr int TR = NAMAX <= 128 ? 32 : 16, TPR = 256 / TR, LDV = NAMAX + 1, NG = NAMAX / 64;
  constexpr int RB = 4;                   // tile loads in flight per thread
  static_assert(3 * 64 * NG * NBW <= TR * LDV && TR * NBW * NBW <= TR * LDV, "reduction fits the tile");
  __shared__ cx<double> sV[TR * LDV];     // the V tile; at the end, the slice partials
  __shared__ cx<double> sH[NAMAX * NBW];
  __shared__ cx<double> sW[TR][NBW];
  const int tid = threadIdx.x;
  const int ur = tid / TPR, us = tid % TPR;        // update role: row ur, columns us + k TPR
  const int ci = tid & 63, rs = tid >> 6;          // product role: columns ci + 64 g, rows rs + 4 k
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
  const int64_t r1 = r0 + rows_per_chunk < n ? r0 + rows_per_chunk : n;
  for (int idx = tid; idx < na * NBW; idx += 256) {
    const int c = idx % na, b = idx / na;
    sH[c * NBW + b] = H[(int64_t)b * ldh + c];
  }
  cx<double> acc[NG][NBW];
  cx<double> gacc[NBW][NBW];
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int b = 0; b < NBW; ++b) acc[g][b] = czero<double>();
#pragma unroll
  for (int a = 0; a < NBW; ++a)
#pragma unroll
    for (int b = 0; b < NBW; ++b) gacc[a][b] = czero<double>();
  const int tile = TR * na;
  for (int64_t rt = r0; rt < r1; rt += TR) {
    __syncthreads();                      // the previous tile's readers are done; sH is visible
    for (int base = 0; base < tile; base += RB * 256) {
      cx<double> x[RB];
#pragma unroll
      for (int t = 0; t < RB; ++t) {
        const int idx = base + t * 256 + tid;
        const int col = idx / TR, row = idx - col * TR;
        x[t] = czero<double>();
        if (idx < tile && rt + row < r1) x[t] = V[(int64_t)col * ldv + rt + row];
      }
#pragma unroll
      for (int t = 0; t < RB; ++t) {
        const int idx = base + t * 256 + tid;
        const int col = idx / TR, row = idx - col * TR;
        if (idx < tile) sV[row * LDV + col] = x[t];
      }
    }
    __syncthreads();
    cx<double> s[NBW];
#pragma unroll
    for (int b = 0; b < NBW; ++b) s[b] = czero<double>();
    for (int c = us; c < na; c += TPR) {
      const cx<double> v = sV[ur * LDV + c];
#pragma unroll
      for (int b = 0; b < NBW; ++b) s[b] = cfma(s[b], v, sH[c * NBW + b]);
    }
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1)
#pragma unroll
      for (int b = 0; b < NBW; ++b) {
        s[b].re += __shfl_xor(s[b].re, off);
        s[b].im += __shfl_xor(s[b].im, off);
      }
    if (us == 0) {
      const int64_t row = rt + ur;
      cx<double> x[NBW];
#pragma unroll
      for (int b = 0; b < NBW; ++b) {
        x[b] = czero<double>();
        if (row < r1) {
          cx<double>* p = W + (int64_t)b * ldw + row;
          x[b] = csub(*p, s[b]);
          *p = x[b];
        }
        sW[ur][b] = x[b];
      }
      if constexpr (MODE == 1) {
#pragma unroll
        for (int a = 0; a < NBW; ++a)
#pragma unroll
          for (int b = 0; b < NBW; ++b) gacc[a][b] = cfma_conj(gacc[a][b], x[a], x[b]);
      }
    }
    if constexpr (MODE == 0) {
      __syncthreads();
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const int col = ci + 64 * g;
        if (col < na) {
          for (int r = rs; r < TR; r += 4) {
            const cx<double> v = sV[r * LDV + col];
#pragma unroll
            for (int b = 0; b < NBW; ++b) acc[g][b] = cfma_conj(acc[g][b], v, sW[r][b]);
          }
        }
      }
    }
  }
  __syncthreads();
  if constexpr (MODE == 0) {
    if (rs > 0) {
#pragma unroll
      for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int b = 0; b < NBW; ++b) sV[(((rs - 1) * 64 + ci) * NG + g) * NBW + b] = acc[g][b];
    }
    __syncthreads();
    if (rs == 0) {
      cx<double>* out = partial + (int64_t)blockIdx.x * na * NBW;
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const int col = ci + 64 * g;
        if (col < na) {
#pragma unroll
          for (int b = 0; b < NBW; ++b) {
            cx<double> t = acc[g][b];
#pragma unroll
            for (int q = 0; q < 3; ++q) t = cadd(t, sV[((q * 64 + ci) * NG + g) * NBW + b]);
            out[(int64_t)b * na + col] = t;
          }
        }
      }
    }
  } else {
    if (us == 0) {
#pragma unroll
      for (int a = 0; a < NBW; ++a)
#pragma unroll
        for (int b = 0; b < NBW; ++b) sV[(ur * NBW + a) * NBW + b] = gacc[a][b];
    }
    __syncthreads();
    if (tid < NBW * NBW) {
      const int a = tid % NBW, b = tid / NBW;
      cx<double> t = czero<double>();
      for (int q = 0; q < TR; ++q) t = cadd(t, sV[(q * NBW + a) * NBW + b]);
      partial[(int64_t)blockIdx.x * NBW * NBW + b * NBW + a] = t;   // column-major NBW x NBW
    }
  }
}

template <int NBW, int MODE>
static hipError_t launch_ortho_step_nbw(const cx<double>* V, int64_t ldv, int na, cx<double>* W, int64_t ldw,
                                        const cx<double>* H, int ldh, int64_t n, int nchunk, int64_t rows_per_chunk,
                                        cx<double>* partial, hipStream_t s) {
  const dim3 grid((unsigned)nchunk);
  if (na <= 64)
    hipLaunchKernelGGL((ortho_step_kernel<NBW, 64, MODE>), grid, dim3(256), 0, s, V, ldv, na, W, ldw, H, ldh, n,
                       rows_per_chunk, partial);
  else if (na <= 128)
    hipLaunchKernelGGL((ortho_step_kernel<NBW, 128, MODE>), grid, dim3(256), 0, s, V, ldv, na, W, ldw, H, ldh, n,
                       rows_per_chunk, partial);
  else
    hipLaunchKernelGGL((ortho_step_kernel<NBW, 256, MODE>), grid, dim3(256), 0, s, V, ldv, na, W, ldw, H, ldh, n,
                       rows_per_chunk, partial);
  return hipGetLastError();
}

hipError_t launch_ortho_step(const cx<double>* V, int64_t ldv, int na, cx<double>* W, int64_t ldw, int nbw,
                             const cx<double>* H, int ldh, int64_t n, int nchunk, int64_t rows_per_chunk, int gram,
                             cx<double>* partial, hipStream_t s) {
  if (na < 1 || na > ORTHO_STEP_MAXA || nbw < 1 || nbw > 2 || ldh < na || nchunk < 1 || rows_per_chunk < 1 ||
      (int64_t)nchunk * rows_per_chunk < n)
    return hipErrorInvalidValue;
  if (nbw == 1)
    return gram ? launch_ortho_step_nbw<1, 1>(V, ldv, na, W, ldw, H, ldh, n, nchunk, rows_per_chunk, partial, s)
                : launch_ortho_step_nbw<1, 0>(V, ldv, na, W, ldw, H, ldh, n, nchunk, rows_per_chunk, partial, s);
  return gram ? launch_ortho_step_nbw<2, 1>(V, ldv, na, W, ldw, H, ldh, n, nchunk, rows_per_chunk, partial, s)
              : launch_ortho_step_nbw<2, 0>(V, ldv, na, W, ldw, H, ldh, n, nchunk, rows_per_chunk, partial, s);
}

// out[i] = sum_p in[p*len + i]: 8 outputs x 32 part-slices per workgroup, so
// that each thread adds at most nparts/32 values (short load chains even for
// a handful of outputs); the slice sums are combined through LDS in a fixed
// order (deterministic).
__global__ __launch_bounds__(256) void sum_parts_kernel(const double* __restrict__ in, int nparts, int len,
                                                        double* __restrict__ out) {
  __shared__ double red[32][9];
  const int o = threadIdx.x & 7, sl = threadIdx.x >> 3;
  const int i = blockIdx.x * 8 + o;
  double s = 0.0;
  if (i < len) {
    int p = sl;
    for (; p + 96 < nparts; p += 128) {
      const double a = in[(int64_t)p * len + i], b = in[(int64_t)(p + 32) * len + i];
      const double c = in[(int64_t)(p + 64) * len + i], e = in[(int64_t)(p + 96) * len + i];
      s += (a + b) + (c + e);
    }
    for (; p < nparts; p += 32) s += in[(int64_t)p * len + i];
  }
  red[sl][o] = s;
  __syncthreads();
  if (sl < 8) {
    double t = red[sl][o] + red[sl + 8][o] + red[sl + 16][o] + red[sl + 24][o];
    red[sl][o] = t;
  }
  __syncthreads();
  if (sl == 0 && i < len) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) t += red[q][o];
    out[i] = t;
  }
}

hipError_t launch_sum_parts(const double* in, int nparts, int len, double* out, hipStream_t s) {
  if (len <= 0) return hipSuccess;
  hipLaunchKernelGGL(sum_parts_kernel, dim3((len + 7) / 8), dim3(256), 0, s, in, nparts, len, out);
  return hipGetLastError();
}

// Cholesky factor of a small Hermitian positive definite matrix and the
// inverse of the factor, one wavefront (n <= 16).  G (n x n, column-major, ld
// ldg) enters symmetrised as (G + G^H)/2.  On return R (upper triangular,
// G = R^H R) and Rinv = R^{-1}, both n x n column-major with ld n and zeros
// below the diagonal; *flag = 1 when a pivot is not positive (the block is
// numerically rank deficient), and then R = Rinv = I.  Used by the Cholesky
// QR of the Krylov expansion.
__global__ __launch_bounds__(64) void chol_inv_kernel(const cx<double>* __restrict__ G, int ldg, int n,
                                                      cx<double>* __restrict__ R, cx<double>* __restrict__ Rinv,
                                                      int* __restrict__ flag) {
  __shared__ cx<double> A[16][17];
  __shared__ cx<double> X[16][17];
  __shared__ int bad;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < n * n; idx += 64) {
    const int i = idx % n, j = idx / n;
    const cx<double> gij = G[(int64_t)j * ldg + i], gji = G[(int64_t)i * ldg + j];
    A[i][j] = {0.5 * (gij.re + gji.re), 0.5 * (gij.im - gji.im)};
  }
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      const double p = A[k][k].re;
      if (!(p > 0.0)) bad = 1;
      else A[k][k] = {sqrt(p), 0.0};
    }
    __syncthreads();
    if (bad) break;
    const double rkk = A[k][k].re;
    for (int j = k + 1 + tid; j < n; j += 64) {
      A[k][j].re /= rkk;
      A[k][j].im /= rkk;
    }
    __syncthreads();
    const int mk = n - k - 1;
    for (int idx = tid; idx < mk * mk; idx += 64) {
      const int i = k + 1 + idx % mk, j = k + 1 + idx / mk;
      if (i <= j) {
        // A[i][j] -= conj(R[k][i]) R[k][j]
        const cx<double> a = A[k][i], b = A[k][j];
        A[i][j].re -= a.re * b.re + a.im * b.im;
        A[i][j].im -= a.re * b.im - a.im * b.re;
      }
    }
    __syncthreads();
  }
  const bool fail = bad != 0;
  // column j of R^{-1} by back substitution, one thread per column
  if (!fail && tid < n) {
    const int j = tid;
    X[j][j] = {1.0 / A[j][j].re, 0.0};
    for (int i = j - 1; i >= 0; --i) {
      cx<double> s = czero<double>();
      for (int p = i + 1; p <= j; ++p) s = cfma(s, A[i][p], X[p][j]);
      X[i][j] = {-s.re / A[i][i].re, -s.im / A[i][i].re};
    }
  }
  __syncthreads();
  for (int idx = tid; idx < n * n; idx += 64) {
    const int i = idx % n, j = idx / n;
    cx<double> r = czero<double>(), ri = czero<double>();
    if (fail) {
      if (i == j) r = ri = {1.0, 0.0};
    } else if (i <= j) {
      r = A[i][j];
      ri = X[i][j];
    }
    R[(int64_t)j * n + i] = r;
    Rinv[(int64_t)j * n + i] = ri;
  }
  if (tid == 0) *flag = fail ? 1 : 0;
}

// Normalisation of one column from the split-K partials of its Gram value
// (tsdot partials, nparts x 2 doubles): every workgroup sums the partials in
// the same fixed order, then scales its rows, v = w / ||w||; workgroup 0
// writes R = ||w||, Rinv = 1/||w|| and *flag (1 when ||w||^2 is not positive,
// with R = Rinv = 1, as chol_inv_kernel does for a 1 x 1 block).
__global__ __launch_bounds__(256) void normalize1_kernel(const double* __restrict__ part, int nparts,
                                                         const cx<double>* __restrict__ w, cx<double>* __restrict__ v,
                                                         int64_t n, cx<double>* __restrict__ R,
                                                         cx<double>* __restrict__ Rinv, int* __restrict__ flag) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int p = threadIdx.x; p < nparts; p += 256) acc += part[2 * (int64_t)p];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  const double g = red[0];
  const bool fail = !(g > 0.0);
  const double r = fail ? 1.0 : sqrt(g), ri = fail ? 1.0 : 1.0 / r;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const cx<double> x = w[i];
    v[i] = cx<double>{x.re * ri, x.im * ri};
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    R[0] = cx<double>{r, 0.0};
    Rinv[0] = cx<double>{ri, 0.0};
    *flag = fail ? 1 : 0;
  }
}

hipError_t launch_normalize1(const double* part, int nparts, const cx<double>* w, cx<double>* v, int64_t n,
                             cx<double>* R, cx<double>* Rinv, int* flag, hipStream_t s) {
  if (n <= 0 || nparts <= 0) return hipErrorInvalidValue;
  const int64_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(normalize1_kernel, dim3((unsigned)(nb < 1024 ? nb : 1024)), dim3(256), 0, s, part, nparts, w, v,
                     n, R, Rinv, flag);
  return hipGetLastError();
}

hipError_t launch_chol_inv(const cx<double>* G, int ldg, int n, cx<double>* R, cx<double>* Rinv, int* flag,
                           hipStream_t s) {
  if (n < 1 || n > 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(64), 0, s, G, ldg, n, R, Rinv, flag);
  return hipGetLastError();
}

// partial[b] = sum over a grid-stride slice of |va[k] - sigma vb[k]|^2
// (Frobenius norm of A - sigma B on the union pattern; one value per block).
__global__ __launch_bounds__(256) void csr_fro2_kernel(const cx<double>* __restrict__ va,
                                                       const cx<double>* __restrict__ vb, int64_t nnz, double sre,
                                                       double sim, double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * 256) {
    const cx<double> a = va[k], b = vb[k];
    const double re = a.re - (sre * b.re - sim * b.im), im = a.im - (sre * b.im + sim * b.re);
    acc += re * re + im * im;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

hipError_t launch_csr_fro2(const cx<double>* va, const cx<double>* vb, int64_t nnz, double sre, double sim,
                           double* partial, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(csr_fro2_kernel, dim3(nblocks), dim3(256), 0, s, va, vb, nnz, sre, sim, partial);
  return hipGetLastError();
}

template <class T>
__global__ void zero_kernel(cx<T>* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = czero<T>();
}

template <class T>
hipError_t launch_zero(cx<T>* p, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(zero_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
  return hipGetLastError();
}
template hipError_t launch_zero<double>(cx<double>*, int64_t, hipStream_t);

// Pseudo-random columns: entry i of column c is (u0, u1), u uniform in
// [-1, 1), from the SplitMix64 finaliser (Steele, Lea & Flood 2014) of
// seed, c and i -- counter based, so the values do not depend on the launch
// shape.  For Krylov start vectors, which need only be generic.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__global__ __launch_bounds__(256) void rand_cols_kernel(cx<double>* __restrict__ V, int64_t ldv, int64_t n,
                                                        uint64_t seed) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n) return;
  const uint64_t k = splitmix64(seed ^ splitmix64(((uint64_t)c << 40) ^ (uint64_t)i));
  const double s = 1.0 / 2147483648.0;   // 2^-31
  const double re = (double)(int32_t)(uint32_t)k * s, im = (double)(int32_t)(uint32_t)(k >> 32) * s;
  V[(int64_t)c * ldv + i] = cx<double>{re, im};
}

hipError_t launch_rand_cols(cx<double>* V, int64_t ldv, int64_t n, int nc, uint64_t seed, hipStream_t s) {
  if (n <= 0 || nc <= 0) return hipSuccess;
  hipLaunchKernelGGL(rand_cols_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)nc), dim3(256), 0, s, V, ldv, n,
                     seed);
  return hipGetLastError();
}
template hipError_t launch_zero<float>(cx<float>*, int64_t, hipStream_t);

__global__ void copy_cols_kernel(const cx<double>* __restrict__ x, int64_t ldx, cx<double>* __restrict__ y,
                                 int64_t ldy, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n) return;
  y[(int64_t)c * ldy + i] = x[(int64_t)c * ldx + i];
}

hipError_t launch_copy_cols(const cx<double>* x, int64_t ldx, cx<double>* y, int64_t ldy, int64_t n, int nrhs,
                            hipStream_t s) {
  if (n == 0 || nrhs == 0) return hipSuccess;
  hipLaunchKernelGGL(copy_cols_kernel, dim3((unsigned)((n + 255) / 256), nrhs), dim3(256), 0, s, x, ldx, y, ldy, n);
  return hipGetLastError();
}

}  // namespace kore
