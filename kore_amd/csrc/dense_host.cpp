// Small dense complex kernels of the projected eigenproblem (host, FP64).
// Compiled with limited-range complex multiplication (no NaN/Inf recovery
// calls in the inner loops); the divisions that can see extreme magnitudes
// go through a scaled division.
#include "dense_host.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace kore_dense {

namespace {

inline double abs1(zc z) { return std::fabs(z.real()) + std::fabs(z.imag()); }

// a / b without overflow or underflow of |b|^2 (Smith's algorithm)
inline zc cdiv(zc a, zc b) {
  const double br = b.real(), bi = b.imag();
  if (std::fabs(br) >= std::fabs(bi)) {
    const double r = bi / br, d = br + bi * r;
    return zc((a.real() + a.imag() * r) / d, (a.imag() - a.real() * r) / d);
  }
  const double r = br / bi, d = bi + br * r;
  return zc((a.real() * r + a.imag()) / d, (a.imag() * r - a.real()) / d);
}

// Plane rotation G = [c s; -conj(s) c] (c real, >= 0) with G [x; y] = [r; 0].
inline void rotation(zc x, zc y, double& c, zc& s, zc& r) {
  const double ax = std::abs(x), ay = std::abs(y);
  if (ay == 0.0) {
    c = 1.0;
    s = 0.0;
    r = x;
    return;
  }
  if (ax == 0.0) {
    c = 0.0;
    s = 1.0;
    r = y;
    return;
  }
  const double nrm = std::hypot(ax, ay);
  const zc ph = x / ax;
  c = ax / nrm;
  s = ph * std::conj(y) / nrm;
  r = ph * nrm;
}

// rows p, p+1 <- G rows p, p+1 over columns [j0, j1)
inline void rot_rows(zc* M, int64_t ld, int p, int j0, int j1, double c, zc s) {
  const zc sc = std::conj(s);
  for (int j = j0; j < j1; ++j) {
    zc* col = M + (int64_t)j * ld;
    const zc a = col[p], b = col[p + 1];
    col[p] = c * a + s * b;
    col[p + 1] = c * b - sc * a;
  }
}

// columns p, p+1 <- columns p, p+1 G^H over rows [i0, i1)
inline void rot_cols(zc* M, int64_t ld, int p, int i0, int i1, double c, zc s) {
  const zc sc = std::conj(s);
  zc* u = M + (int64_t)p * ld;
  zc* v = u + ld;
  for (int i = i0; i < i1; ++i) {
    const zc a = u[i], b = v[i];
    u[i] = c * a + sc * b;
    v[i] = c * b - s * a;
  }
}

// Householder reduction to upper Hessenberg form, T <- P^H T P, Q <- P.
void hessenberg(int n, zc* T, int64_t ldt, zc* Q, int64_t ldq) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) Q[(int64_t)j * ldq + i] = (i == j) ? 1.0 : 0.0;
  std::vector<zc> v(n), w(n);
  for (int k = 0; k + 2 < n; ++k) {
    zc* colk = T + (int64_t)k * ldt;
    double tail = 0.0;
    for (int i = k + 2; i < n; ++i) tail += std::norm(colk[i]);
    if (tail == 0.0) continue;
    const zc alpha = colk[k + 1];
    const double aa = std::abs(alpha);
    const double bnorm = std::sqrt(aa * aa + tail);
    const zc ph = aa > 0.0 ? alpha / aa : zc(1.0);
    const zc beta = -ph * bnorm;                 // the new subdiagonal entry
    const int len = n - k - 1;
    v[0] = alpha - beta;
    for (int i = 1; i < len; ++i) v[i] = colk[k + 1 + i];
    const double tau = 2.0 / (std::norm(v[0]) + tail);
    // T[k+1:, k+1:] <- (I - tau v v^H) T[k+1:, k+1:]
    for (int j = k + 1; j < n; ++j) {
      zc* col = T + (int64_t)j * ldt + k + 1;
      zc acc = 0.0;
      for (int i = 0; i < len; ++i) acc += std::conj(v[i]) * col[i];
      acc *= tau;
      for (int i = 0; i < len; ++i) col[i] -= v[i] * acc;
    }
    // T[:, k+1:] <- T[:, k+1:] (I - tau v v^H), and the same for Q
    for (int pass = 0; pass < 2; ++pass) {
      zc* M = pass == 0 ? T : Q;
      const int64_t ld = pass == 0 ? ldt : ldq;
      std::fill(w.begin(), w.end(), zc(0.0));
      for (int j = 0; j < len; ++j) {
        const zc* col = M + (int64_t)(k + 1 + j) * ld;
        const zc vj = v[j];
        for (int i = 0; i < n; ++i) w[i] += col[i] * vj;
      }
      for (int j = 0; j < len; ++j) {
        zc* col = M + (int64_t)(k + 1 + j) * ld;
        const zc f = tau * std::conj(v[j]);
        for (int i = 0; i < n; ++i) col[i] -= w[i] * f;
      }
    }
    colk[k + 1] = beta;
    for (int i = k + 2; i < n; ++i) colk[i] = 0.0;
  }
}

}  // namespace

int schur(int n, zc* T, int64_t ldt, zc* Q, int64_t ldq) {
  if (n <= 0) return 0;
  hessenberg(n, T, ldt, Q, ldq);
  auto t = [&](int i, int j) -> zc& { return T[(int64_t)j * ldt + i]; };
  const double ulp = DBL_EPSILON;
  const double smlnum = DBL_MIN * ((double)n / ulp);
  int hi = n - 1;
  int its = 0, total = 0;
  const int maxits = 30 * std::max(10, n);
  while (hi > 0) {
    // deflation: the lowest negligible subdiagonal entry of the active block
    int l = hi;
    for (; l > 0; --l) {
      const double h = abs1(t(l, l - 1));
      if (h <= smlnum) break;
      double tst = abs1(t(l - 1, l - 1)) + abs1(t(l, l));
      if (tst == 0.0) {
        if (l >= 2) tst += abs1(t(l - 1, l - 2));
        if (l + 1 <= hi) tst += abs1(t(l + 1, l));
      }
      if (h <= ulp * tst) {
        // Ahues & Tisseur (LAWN 122): a finer test for small subdiagonals
        const double ab = std::max(h, abs1(t(l - 1, l)));
        const double ba = std::min(h, abs1(t(l - 1, l)));
        const double aa = std::max(abs1(t(l, l)), abs1(t(l - 1, l - 1) - t(l, l)));
        const double bb = std::min(abs1(t(l, l)), abs1(t(l - 1, l - 1) - t(l, l)));
        const double sc = aa + ab;
        if (ba * (ab / sc) <= std::max(smlnum, ulp * (bb * (aa / sc)))) break;
      }
    }
    if (l > 0) t(l, l - 1) = 0.0;
    if (l == hi) {     // one eigenvalue converged
      --hi;
      its = 0;
      continue;
    }
    if (++total > maxits) return hi + 1;
    ++its;
    zc mu;
    if (its % 10 == 0) {
      // exceptional shift: breaks cycles of the Wilkinson shift
      mu = t(hi, hi) + 0.75 * std::fabs(t(hi, hi - 1).real());
    } else {
      // Wilkinson shift: the eigenvalue of the trailing 2 x 2 block nearer t(hi, hi)
      const zc a = t(hi - 1, hi - 1), b = t(hi - 1, hi), c = t(hi, hi - 1), d = t(hi, hi);
      const zc half = 0.5 * (a - d);
      const zc disc = std::sqrt(half * half + b * c);
      const zc m1 = d + half + disc, m2 = d + half - disc;   // (a + d)/2 +- disc
      mu = abs1(m1 - d) <= abs1(m2 - d) ? m1 : m2;
    }
    // implicit single-shift QR step on rows/columns [l, hi], chasing the bulge
    zc x = t(l, l) - mu, y = t(l + 1, l);
    for (int k = l; k < hi; ++k) {
      if (k > l) {
        x = t(k, k - 1);
        y = t(k + 1, k - 1);
      }
      double c;
      zc s, r;
      rotation(x, y, c, s, r);
      if (k > l) {
        t(k, k - 1) = r;
        t(k + 1, k - 1) = 0.0;
      }
      rot_rows(T, ldt, k, k, n, c, s);
      rot_cols(T, ldt, k, 0, std::min(k + 3, hi + 1), c, s);
      rot_cols(Q, ldq, k, 0, n, c, s);
    }
  }
  // exact zeros below the diagonal
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) t(i, j) = 0.0;
  return 0;
}

void schur_reorder(int n, zc* S, int64_t lds, zc* Q, int64_t ldq, const int32_t* select, int* m) {
  auto s_ = [&](int i, int j) -> zc& { return S[(int64_t)j * lds + i]; };
  int ks = 0;
  for (int k = 0; k < n; ++k) {
    if (!select[k]) continue;
    // move entry k up to position ks by swaps (p, p+1), p = k-1 .. ks
    for (int p = k - 1; p >= ks; --p) {
      const zc t11 = s_(p, p), t22 = s_(p + 1, p + 1);
      // the rotation whose conjugate transpose has first column along the
      // eigenvector (s_(p, p+1), t22 - t11) of t22 in the 2 x 2 block
      double c;
      zc s, r;
      rotation(s_(p, p + 1), t22 - t11, c, s, r);
      rot_rows(S, lds, p, p, n, c, s);
      rot_cols(S, lds, p, 0, p + 2, c, s);
      rot_cols(Q, ldq, p, 0, n, c, s);
      s_(p, p) = t22;
      s_(p + 1, p + 1) = t11;
      s_(p + 1, p) = 0.0;
    }
    ++ks;
  }
  *m = ks;
}

void triu_eigvecs(int n, const zc* S, int64_t lds, zc* Z, int64_t ldz) {
  // Z is upper triangular with a unit diagonal: row k of (S - d_j I) z_j = 0
  // gives Z[k, j] = -S[k, k+1:j+1] Z[k+1:j+1, j] / (d_k - d_j), column by
  // column from the bottom.  Near-equal diagonal entries are separated by at
  // least eps |d_j|, and a column that grows past 1e100 is rescaled on the
  // spot so that nothing overflows.
  auto s_ = [&](int i, int j) { return S[(int64_t)j * lds + i]; };
  const double eps = DBL_EPSILON, tiny = DBL_MIN / DBL_EPSILON;
  for (int j = 0; j < n; ++j) {
    zc* z = Z + (int64_t)j * ldz;
    for (int i = 0; i < n; ++i) z[i] = 0.0;
    z[j] = 1.0;
    const zc dj = s_(j, j);
    const double smin = std::max(eps * std::abs(dj), tiny);
    for (int k = j - 1; k >= 0; --k) {
      zc num = 0.0;
      for (int q = k + 1; q <= j; ++q) num += s_(k, q) * z[q];
      zc den = s_(k, k) - dj;
      if (std::abs(den) < smin) den = smin;
      z[k] = -cdiv(num, den);
      const double a = std::abs(z[k]);
      if (a > 1e100)
        for (int q = k; q <= j; ++q) z[q] /= a;
    }
    double nrm = 0.0;
    for (int i = 0; i <= j; ++i) nrm += std::norm(z[i]);
    nrm = std::sqrt(nrm);
    for (int i = 0; i <= j; ++i) z[i] /= nrm;
  }
}

int cholesky_upper(int n, zc* G, int64_t ld) {
  auto g = [&](int i, int j) -> zc& { return G[(int64_t)j * ld + i]; };
  for (int j = 0; j < n; ++j) {
    double d = g(j, j).real();
    for (int k = 0; k < j; ++k) d -= std::norm(g(k, j));
    if (!(d > 0.0) || !std::isfinite(d)) return j + 1;
    const double r = std::sqrt(d);
    g(j, j) = r;
    for (int i = j + 1; i < n; ++i) {
      zc acc = g(j, i);
      for (int k = 0; k < j; ++k) acc -= std::conj(g(k, j)) * g(k, i);
      g(j, i) = acc / r;
    }
    for (int i = j + 1; i < n; ++i) g(i, j) = 0.0;
  }
  return 0;
}

void triu_inverse(int n, zc* R, int64_t ld) {
  auto r = [&](int i, int j) -> zc& { return R[(int64_t)j * ld + i]; };
  // column j of the inverse: X[j, j] = 1 / R[j, j]; X[i, j] = -(sum_{q=i+1..j}
  // R[i, q] X[q, j]) / R[i, i], rows from the bottom; computed in place from
  // the last column backwards since column j reads R's columns q <= j only
  // through rows i < q, which later columns no longer need.
  std::vector<zc> x(n);
  for (int j = n - 1; j >= 0; --j) {
    x[j] = 1.0 / r(j, j);
    for (int i = j - 1; i >= 0; --i) {
      zc acc = 0.0;
      for (int q = i + 1; q <= j; ++q) acc += r(i, q) * x[q];
      x[i] = -acc / r(i, i);
    }
    for (int i = 0; i <= j; ++i) r(i, j) = x[i];
  }
}

}  // namespace kore_dense
