// Level 0 of the cyclic reduction through the sparse couplings (gfx950).
//
// At level 0 the couplings L_t, U_t of degree block t are the off-diagonal
// blocks of the row-equilibrated matrix R (A - sigma B) itself: radial
// operators of a few dozen entries per row, not dense blocks.  Multiplying
// them as dense n x n blocks (the rest of the reduction's GEMMs) would cost
// 8 n^3 real flops per product where 8 nnz(S) n do; so level 0 keeps them in
// a compact off-block CSR (values rebuilt per shift) plus a per-block CSC
// view, and forms the level's products with two memory-streaming kernels:
//
//   sprow  C = beta C + alpha sum_p S_p X_p   (sparse rows x dense; for the
//          kept blocks t: D_t -= L_t H_{t-1} + U_t G_{t+1},
//          L_t <- -L_t G_{t-1}, U_t <- -U_t H_{t+1})
//   spcol  C = beta C + alpha X S             (dense x sparse columns; for the
//          eliminated blocks t: G_t = Dinv_t L_t, H_t = Dinv_t U_t)
//   offsweep  y -= S v on the rows of one block parity (the level-0 step of
//          every solve sweep: b_even -= L y_odd forward, b_odd -= L x_even
//          backward), for complex128 or complex64 vectors and up to 16
//          right-hand sides
//   offvals   oval[q] = r_i (a_k - sigma b_k) for the off-block entries
#include "kore_device.hpp"
#include "kore_kernels.hpp"

namespace kore {

template <class T> __device__ __forceinline__ cx<double> up(cx<T> a) { return cx<double>{(double)a.re, (double)a.im}; }

// ---------------------------------------------------------------------------
// values of the off-block entries for a shift: 16 lanes per row
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void offvals_kernel(CsrDev m, cx<double> cs, const double* __restrict__ scale,
                                                      const int32_t* __restrict__ split, OffCsr o) {
  const int g = threadIdx.x >> 4, gl = threadIdx.x & 15;
  const int64_t row = (int64_t)blockIdx.x * 16 + g;
  if (row >= m.n) return;
  const int32_t sp = split[row];
  const int nlo = sp & 0x7FFF, nhi = (sp >> 15) & 0x7FFF;
  const int s = m.indptr[row], e = m.indptr[row + 1];
  const int q0 = o.ptr[row];
  const double sc = scale[row];
  // the union row's entries are sorted by column: the previous block's are
  // its first nlo, the next block's its last nhi, in the order of the
  // compact row
  for (int q = gl; q < nlo + nhi; q += 16) {
    const int k = q < nlo ? s + q : e - nhi + (q - nlo);
    const cx<double> v = csub(m.va[k], cmul(cs, m.vb[k]));
    o.val[q0 + q] = cx<double>{sc * v.re, sc * v.im};
  }
}

hipError_t launch_offvals(const CsrDev& m, double sre, double sim, const double* scale, const int32_t* split,
                          const OffCsr& o, hipStream_t s) {
  if (m.n == 0) return hipSuccess;
  hipLaunchKernelGGL(offvals_kernel, dim3((unsigned)((m.n + 15) / 16)), dim3(256), 0, s, m, cx<double>{sre, sim},
                     scale, split, o);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// sparse rows x dense: one workgroup per (descriptor, tile of SR_NC<T>
// columns).  The tile of each X_p (all its rows, SR_NC columns) is staged in
// LDS once, so every sparse entry reads its X values from LDS: an entry of
// column c is shared by the ~w rows of the band around c, which would
// otherwise re-read X[c, :] from the caches w times.  Each thread owns up to
// two rows of C (blocks of at most 512 rows) and SR_NC columns.
// ---------------------------------------------------------------------------
static constexpr int SR_MAXN = 512;   // rows of a degree block (and of an X tile)
static constexpr int SR_U = 8;        // sparse entries whose loads are in flight together

template <class T, int NC>
__global__ __launch_bounds__(256) void sprow_kernel(OffCsr o, const SpRowDesc<T>* __restrict__ descs) {
  __shared__ cx<T> sX[NC * SR_MAXN];      // X tile, column-major: sX[jj * SR_MAXN + c]
  const SpRowDesc<T>& d = descs[blockIdx.y];
  const int j0 = blockIdx.x * NC;
  if (j0 >= d.n) return;                  // uniform over the workgroup
  const int nc = d.n - j0 < NC ? d.n - j0 : NC;
  const int tid = threadIdx.x;
  cx<double> acc[2][NC];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[h][c] = czero<double>();
  for (int p = 0; p < 2; ++p) {
    if (d.side[p] < 0) continue;          // uniform
    const int nbk = d.xrows[p];           // rows of X_p (the neighbouring block's unknowns)
    const cx<T>* X = d.X[p] + (int64_t)j0 * d.ldx[p];
    __syncthreads();                      // the previous product's reads are done
    for (int idx = tid; idx < NC * nbk; idx += 256) {
      const int jj = idx / nbk, c = idx - jj * nbk;
      sX[jj * SR_MAXN + c] = jj < nc ? gld(X + (int64_t)jj * d.ldx[p] + c) : czero<T>();
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = tid + 256 * h;
      if (r >= d.m) continue;
      const int64_t row = d.row0 + r;
      const int s = o.ptr[row], nl = o.nlo[row];
      const int q0 = d.side[p] == 0 ? s : s + nl, q1 = d.side[p] == 0 ? s + nl : o.ptr[row + 1];
      const int xc0 = (int)d.xcol0[p];
      // entries in groups of SR_U: the group's loads are all issued before
      // the first is used, so a row pays the memory latency once per group
      // (the entries are still added in order, as one at a time would)
      int q = q0;
      for (; q + SR_U <= q1; q += SR_U) {
        cx<double> vv[SR_U];
        int cc[SR_U];
#pragma unroll
        for (int u = 0; u < SR_U; ++u) {
          cc[u] = o.col[q + u];
          vv[u] = o.val[q + u];
        }
#pragma unroll
        for (int u = 0; u < SR_U; ++u) {
          const int c = cc[u] - xc0;
#pragma unroll
          for (int jj = 0; jj < NC; ++jj) acc[h][jj] = cfma(acc[h][jj], vv[u], up(sX[jj * SR_MAXN + c]));
        }
      }
      for (; q < q1; ++q) {
        const cx<double> v = o.val[q];
        const int c = o.col[q] - xc0;
#pragma unroll
        for (int jj = 0; jj < NC; ++jj) acc[h][jj] = cfma(acc[h][jj], v, up(sX[jj * SR_MAXN + c]));
      }
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int r = tid + 256 * h;
    if (r >= d.m) continue;
#pragma unroll
    for (int jj = 0; jj < NC; ++jj) {
      if (jj < nc) {
        cx<T>* y = d.C + (int64_t)(j0 + jj) * d.ldc + r;
        cx<double> out{(double)d.alpha * acc[h][jj].re, (double)d.alpha * acc[h][jj].im};
        if (d.beta != T(0)) {
          const cx<double> old = up(gld(y));
          out.re += (double)d.beta * old.re;
          out.im += (double)d.beta * old.im;
        }
        gst(y, cx<T>{(T)out.re, (T)out.im});
      }
    }
  }
}

// nc: columns per workgroup tile, 4, 8 or 16 (LDS: 16 B x 512 rows x nc for
// complex128, so 128 KiB of gfx950's 160 KiB at 16; half that for complex64);
// chosen per block-size class by the library's tuner.
template <class T>
hipError_t launch_sprow(const OffCsr& o, const SpRowDesc<T>* d_descs, int ndesc, int max_m, int max_n, int nc,
                        hipStream_t s) {
  if (ndesc <= 0 || max_m <= 0 || max_n <= 0) return hipSuccess;
  if (max_m > SR_MAXN) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((max_n + nc - 1) / nc), (unsigned)ndesc);
  if (nc == 4)
    hipLaunchKernelGGL((sprow_kernel<T, 4>), grid, dim3(256), 0, s, o, d_descs);
  else if (nc == 8)
    hipLaunchKernelGGL((sprow_kernel<T, 8>), grid, dim3(256), 0, s, o, d_descs);
  else if (nc == 16)
    hipLaunchKernelGGL((sprow_kernel<T, 16>), grid, dim3(256), 0, s, o, d_descs);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
template hipError_t launch_sprow<double>(const OffCsr&, const SpRowDesc<double>*, int, int, int, int, hipStream_t);
template hipError_t launch_sprow<float>(const OffCsr&, const SpRowDesc<float>*, int, int, int, int, hipStream_t);

// ---------------------------------------------------------------------------
// dense x sparse columns: one wave per output column, lanes over the rows;
// the column's entry list is uniform over the wave (scalar loads) and each
// entry is one coalesced column segment of X
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void spcol_kernel(OffCsc cc, const cx<double>* __restrict__ oval,
                                                    const SpColDesc<T>* __restrict__ descs) {
  const SpColDesc<T>& d = descs[blockIdx.y];
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= d.n) return;
  const int k0 = cc.ptr[d.cbase + c], k1 = cc.ptr[d.cbase + c + 1];
  for (int r0 = 0; r0 < d.m; r0 += 64) {
    const int r = r0 + lane;
    cx<double> acc = czero<double>();
    if (r < d.m) {
      // entries in groups of SR_U, loads first (as in sprow_kernel)
      int k = k0;
      for (; k + SR_U <= k1; k += SR_U) {
        cx<double> vv[SR_U];
        cx<T> xx[SR_U];
#pragma unroll
        for (int u = 0; u < SR_U; ++u) {
          vv[u] = oval[cc.q[k + u]];
          xx[u] = gld(d.X + (int64_t)cc.row[k + u] * d.ldx + r);
        }
#pragma unroll
        for (int u = 0; u < SR_U; ++u) acc = cfma(acc, up(xx[u]), vv[u]);
      }
      for (; k < k1; ++k) {
        const cx<double> v = oval[cc.q[k]];
        acc = cfma(acc, up(gld(d.X + (int64_t)cc.row[k] * d.ldx + r)), v);
      }
      cx<T>* y = d.C + (int64_t)c * d.ldc + r;
      cx<double> out{(double)d.alpha * acc.re, (double)d.alpha * acc.im};
      if (d.beta != T(0)) {
        const cx<double> old = up(gld(y));
        out.re += (double)d.beta * old.re;
        out.im += (double)d.beta * old.im;
      }
      gst(y, cx<T>{(T)out.re, (T)out.im});
    }
  }
}

template <class T>
hipError_t launch_spcol(const OffCsc& cc, const cx<double>* oval, const SpColDesc<T>* d_descs, int ndesc, int max_n,
                        hipStream_t s) {
  if (ndesc <= 0 || max_n <= 0) return hipSuccess;
  const dim3 grid((unsigned)((max_n + 3) / 4), (unsigned)ndesc);
  hipLaunchKernelGGL(spcol_kernel<T>, grid, dim3(256), 0, s, cc, oval, d_descs);
  return hipGetLastError();
}
template hipError_t launch_spcol<double>(const OffCsc&, const cx<double>*, const SpColDesc<double>*, int, int,
                                         hipStream_t);
template hipError_t launch_spcol<float>(const OffCsc&, const cx<double>*, const SpColDesc<float>*, int, int,
                                        hipStream_t);

// ---------------------------------------------------------------------------
// level-0 step of a solve sweep: y[i] -= sum_q oval[q] v[col q] for the rows
// i of blocks of one parity, nrhs <= NR columns; 16 lanes per row
// ---------------------------------------------------------------------------
template <class T, int NR>
__global__ __launch_bounds__(256) void offsweep_kernel(OffCsr o, const int32_t* __restrict__ blk_of_row, int64_t n,
                                                       int parity, const cx<T>* __restrict__ v, int64_t ldv,
                                                       cx<T>* __restrict__ y, int64_t ldy, int nrhs) {
  const int g = threadIdx.x >> 4, gl = threadIdx.x & 15;
  const int64_t row = (int64_t)blockIdx.x * 16 + g;
  if (row >= n) return;
  if ((blk_of_row[row] & 1) != parity) return;      // uniform over the row's 16 lanes
  const int s = o.ptr[row], e = o.ptr[row + 1];
  cx<double> acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = czero<double>();
  for (int q = s + gl; q < e; q += 16) {
    const cx<double> a = o.val[q];
    const int64_t col = o.col[q];
#pragma unroll
    for (int c = 0; c < NR; ++c)
      if (c < nrhs) acc[c] = cfma(acc[c], a, up(v[c * ldv + col]));
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
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      if (c < nrhs) {
        cx<T>* p = y + c * ldy + row;
        const cx<T> old = *p;
        *p = cx<T>{(T)((double)old.re - acc[c].re), (T)((double)old.im - acc[c].im)};
      }
    }
  }
}

template <class T>
hipError_t launch_offsweep(const OffCsr& o, const int32_t* blk_of_row, int64_t n, int parity, const cx<T>* v,
                           int64_t ldv, cx<T>* y, int64_t ldy, int nrhs, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (nrhs < 1 || nrhs > 16) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((n + 15) / 16));
  if (nrhs == 1)
    hipLaunchKernelGGL((offsweep_kernel<T, 1>), grid, dim3(256), 0, s, o, blk_of_row, n, parity, v, ldv, y, ldy, nrhs);
  else if (nrhs <= 2)
    hipLaunchKernelGGL((offsweep_kernel<T, 2>), grid, dim3(256), 0, s, o, blk_of_row, n, parity, v, ldv, y, ldy, nrhs);
  else if (nrhs <= 4)
    hipLaunchKernelGGL((offsweep_kernel<T, 4>), grid, dim3(256), 0, s, o, blk_of_row, n, parity, v, ldv, y, ldy, nrhs);
  else if (nrhs <= 8)
    hipLaunchKernelGGL((offsweep_kernel<T, 8>), grid, dim3(256), 0, s, o, blk_of_row, n, parity, v, ldv, y, ldy, nrhs);
  else
    hipLaunchKernelGGL((offsweep_kernel<T, 16>), grid, dim3(256), 0, s, o, blk_of_row, n, parity, v, ldv, y, ldy,
                       nrhs);
  return hipGetLastError();
}
template hipError_t launch_offsweep<double>(const OffCsr&, const int32_t*, int64_t, int, const cx<double>*, int64_t,
                                            cx<double>*, int64_t, int, hipStream_t);
template hipError_t launch_offsweep<float>(const OffCsr&, const int32_t*, int64_t, int, const cx<float>*, int64_t,
                                           cx<float>*, int64_t, int, hipStream_t);

}  // namespace kore
