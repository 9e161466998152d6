// Device-side building blocks shared by the Kore solver kernels (gfx950).
//
// Storage conventions used by every kernel:
//   * complex numbers are interleaved (re, im), 16 B for complex128 and
//     8 B for complex64, so numpy complex arrays are used without copies;
//   * dense blocks are column-major: element (i, j) of a block with leading
//     dimension ld sits at j*ld + i.
//
// Matrix-core use: complex products are split into four real MFMA products
// on the re/im parts.  The MFMA operands are swapped (C^T = B^T A^T) so that
// the 16 lanes that share an output register walk down a column of C, which
// makes the column-major loads and stores of C coalesced.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kore {

template <class T> struct alignas(2 * sizeof(T)) cx { T re, im; };

template <class T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> __device__ __forceinline__ cx<T> cfma(cx<T> acc, cx<T> a, cx<T> b) {
  acc.re = fma(a.re, b.re, acc.re); acc.re = fma(-a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im); acc.im = fma(a.im, b.re, acc.im);
  return acc;
}
// acc += conj(a) * b
template <class T> __device__ __forceinline__ cx<T> cfma_conj(cx<T> acc, cx<T> a, cx<T> b) {
  acc.re = fma(a.re, b.re, acc.re); acc.re = fma(a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im); acc.im = fma(-a.im, b.re, acc.im);
  return acc;
}
template <class T> __device__ __forceinline__ cx<T> czero() { return {T(0), T(0)}; }

// Global-memory access through pointers that the compiler cannot prove global
// (those read from descriptor arrays in device memory).  Through a generic
// pointer it emits flat loads, which count against the LDS/scalar wait counter
// as well as the vector-memory one, so every LDS read or descriptor-field wait
// would also wait for all outstanding loads and serialise their latencies.
template <class T> struct cx_vec;
template <> struct cx_vec<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct cx_vec<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <class T> __device__ __forceinline__ cx<T> gld(const cx<T>* p) {
  typedef typename cx_vec<T>::type V;
  const V v = *(const __attribute__((address_space(1))) V*)p;
  return cx<T>{v.x, v.y};
}
template <class T> __device__ __forceinline__ void gst(cx<T>* p, cx<T> x) {
  typedef typename cx_vec<T>::type V;
  *(__attribute__((address_space(1))) V*)p = V{x.re, x.im};
}
template <class T> __device__ __forceinline__ T cabs1(cx<T> a) { return fabs(a.re) + fabs(a.im); }

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <class T> struct acc4;
template <> struct acc4<double> { typedef d4v type; };
template <> struct acc4<float> { typedef f4v type; };

// D += A(16x4) * B(4x16).  Lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Lane l receives D[mrow(l/16, i)][l%16] in register i = 0..3, where the row
// index differs between the two instructions (measured on gfx950):
//   f64 16x16x4:  mrow = 4*i + l/16      (rows interleaved across lane groups)
//   f32 16x16x4:  mrow = 4*(l/16) + i    (four consecutive rows per lane group)
__device__ __forceinline__ d4v mfma16(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma16(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class T> __device__ __forceinline__ int mfma_mrow(int lq, int i);
template <> __device__ __forceinline__ int mfma_mrow<double>(int lq, int i) { return 4 * i + lq; }
template <> __device__ __forceinline__ int mfma_mrow<float>(int lq, int i) { return 4 * lq + i; }

// One 16x16 complex tile step over k-slab of 4:
//   C[i][j] += sum_k A[i][k] B[k][j]
// with the caller supplying, for lane l,  a = A[i0 + l%16][k0 + l/16]  and
// b = B[k0 + l/16][j0 + l%16].  Accumulators hold C^T in MFMA layout, i.e.
// lane l, register r  ->  C[i0 + l%16][j0 + mfma_mrow<T>(l/16, r)].
template <class T>
__device__ __forceinline__ void cmma(typename acc4<T>::type& cre, typename acc4<T>::type& cim,
                                     cx<T> a, cx<T> b) {
  cre = mfma16(b.re, a.re, cre);
  cre = mfma16(b.im, -a.im, cre);
  cim = mfma16(b.im, a.re, cim);
  cim = mfma16(b.re, a.im, cim);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

template <class T> __device__ __forceinline__ T shfl_xor_t(T v, int m) { return __shfl_xor(v, m, 64); }

}  // namespace kore
