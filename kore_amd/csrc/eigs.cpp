// Shift-and-invert block Krylov-Schur behind one C call (kore_eigs).
//
// Replaces E.solve() of bin/solve.py:91-123 (EPS GNHEP, Krylov-Schur, ST
// sinvert) and the extraction of bin/solve.py:131-158: the eigenpairs of
// A x = lambda B x nearest the target tau come from the operator
// OP = (A - tau B)^{-1} B, whose eigenvalues are theta = 1/(lambda - tau).
//
// Built on the library's own entry points: the Krylov basis is a device
// workspace (kore_ws_*), a whole block Arnoldi expansion is one call
// (kore_ws_expand: OP, two classical Gram-Schmidt passes and Cholesky-QR per
// block, on the device), and the host keeps only the projected matrix H,
// whose ordered Schur form comes from the dense routines of dense_host.
//
// Krylov decomposition kept throughout: OP V[:, :j] = V[:, :j+bs] H[:j+bs, :j].
// A Ritz pair (theta, y) has converged when ||H[j:j+bs, :j] y|| <= tol |theta|
// (SLEPc's default relative criterion); pairs count in wanted order,
// consecutively.  Thick restart keeps nconv + (m - nconv)/2 Ritz vectors.
//
// Early convergence checks: once the basis can hold the wanted pairs, the
// expansion proceeds in chunks; beside each chunk (a host thread, while the
// device expands) the Rayleigh-Ritz step of the decomposition reached
// before the chunk runs, and when it finds the wanted pairs converged the
// expansion is stopped (kore_ws_stop_expand).  Chunks grow so that their
// device time covers the check, or end where the trend of the worst wanted
// residual reaches tol.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "dense_host.hpp"
#include "kore_amd.h"
#include "kore_internal.hpp"

using kore_dense::zc;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int32_t err(int32_t code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int32_t err(int32_t code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return kore_internal_set_error(code, buf);
}

#define KTRY(x)                  \
  do {                           \
    const int32_t rc_ = (x);     \
    if (rc_ < 0) return rc_;     \
  } while (0)

// Ascending key, wanted eigenvalues first (bin/solve.py:99-117).  A Ritz
// value theta = 0 (the null space of B) maps to lambda = infinity and sorts last.
double sort_key(int which, zc lam, zc tau) {
  double k;
  switch (which) {
    case KORE_WHICH_TM: k = std::abs(lam - tau); break;
    case KORE_WHICH_TR: k = std::fabs(lam.real() - tau.real()); break;
    case KORE_WHICH_TI: k = std::fabs(lam.imag() - tau.imag()); break;
    case KORE_WHICH_LM: k = -std::abs(lam); break;
    case KORE_WHICH_SM: k = std::abs(lam); break;
    case KORE_WHICH_LR: k = -lam.real(); break;
    case KORE_WHICH_SR: k = lam.real(); break;
    case KORE_WHICH_LI: k = -lam.imag(); break;
    default: k = lam.imag(); break;   // KORE_WHICH_SI
  }
  return std::isnan(k) ? HUGE_VAL : k;
}

// Rayleigh-Ritz data of the decomposition of size j.
struct Ritz {
  int j = 0, bs = 1;
  std::vector<zc> S, Q, Z;          // Schur form, Schur vectors, eigenvectors of S (j x j each)
  std::vector<zc> Bres;             // H[j:j+bs, :j] (bs x j)
  std::vector<int> order;           // wanted order of the Schur diagonal
  std::vector<zc> theta;            // Ritz values in wanted order
  std::vector<double> res;          // ||Bres y|| in wanted order
  int nconv = 0;
  double worst(int nev) const {     // worst relative residual of the nev wanted pairs
    double w = 0;
    for (int i = 0; i < std::min(nev, j); ++i) w = std::max(w, res[i] / std::abs(theta[i]));
    return w;
  }
};

int32_t rayleigh_ritz(const zc* H, int64_t ldh, int j, int bs, zc tau, int which, double tol, Ritz& r) {
  r.j = j;
  r.bs = bs;
  r.S.assign((size_t)j * j, 0.0);
  r.Q.assign((size_t)j * j, 0.0);
  r.Z.assign((size_t)j * j, 0.0);
  r.Bres.assign((size_t)bs * j, 0.0);
  for (int c = 0; c < j; ++c) {
    for (int i = 0; i < j; ++i) r.S[(size_t)c * j + i] = H[(int64_t)c * ldh + i];
    for (int i = 0; i < bs; ++i) r.Bres[(size_t)c * bs + i] = H[(int64_t)c * ldh + j + i];
  }
  if (kore_dense::schur(j, r.S.data(), j, r.Q.data(), j) != 0)
    return err(KORE_ERR_SINGULAR, "QR iteration of the projected matrix (order %d) did not converge", j);
  kore_dense::triu_eigvecs(j, r.S.data(), j, r.Z.data(), j);
  // residual estimates ||Bres Q z|| per Ritz vector
  std::vector<zc> BQ((size_t)bs * j, 0.0);
  for (int c = 0; c < j; ++c)
    for (int q = 0; q < j; ++q) {
      const zc v = r.Q[(size_t)c * j + q];
      for (int i = 0; i < bs; ++i) BQ[(size_t)c * bs + i] += r.Bres[(size_t)q * bs + i] * v;
    }
  std::vector<double> res_s(j);
  std::vector<zc> th_s(j);
  std::vector<double> key(j);
  for (int c = 0; c < j; ++c) {
    double acc = 0;
    for (int i = 0; i < bs; ++i) {
      zc y = 0.0;
      for (int q = 0; q <= c; ++q) y += BQ[(size_t)q * bs + i] * r.Z[(size_t)c * j + q];
      acc += std::norm(y);
    }
    res_s[c] = std::sqrt(acc);
    th_s[c] = r.S[(size_t)c * j + c];
    key[c] = th_s[c] == 0.0 ? HUGE_VAL : sort_key(which, tau + 1.0 / th_s[c], tau);
  }
  r.order.resize(j);
  std::iota(r.order.begin(), r.order.end(), 0);
  std::stable_sort(r.order.begin(), r.order.end(), [&](int a, int b) { return key[a] < key[b]; });
  r.theta.resize(j);
  r.res.resize(j);
  for (int i = 0; i < j; ++i) {
    r.theta[i] = th_s[r.order[i]];
    r.res[i] = res_s[r.order[i]];
  }
  r.nconv = 0;
  while (r.nconv < j && r.res[r.nconv] <= tol * std::abs(r.theta[r.nconv])) ++r.nconv;
  return KORE_OK;
}

// Orthonormalises V[:, src:src+nc] into V[:, dst:dst+nc] (two Cholesky-QR
// passes; one for a single column); R (nc x nc) with V_src = V_dst R.
// KORE_WARN_BREAKDOWN when the block is numerically rank deficient.
int32_t cholqr(kore_ws* w, int src, int dst, int nc, std::vector<zc>& R) {
  std::vector<zc> G((size_t)nc * nc), G2((size_t)nc * nc);
  auto gram = [&](int c0, std::vector<zc>& M) -> int32_t {
    KTRY(kore_ws_dot(w, c0, nc, c0, nc, reinterpret_cast<double*>(M.data())));
    for (int a = 0; a < nc; ++a)
      for (int b = 0; b <= a; ++b) {
        const zc h = 0.5 * (M[(size_t)b * nc + a] + std::conj(M[(size_t)a * nc + b]));
        M[(size_t)b * nc + a] = h;
        M[(size_t)a * nc + b] = std::conj(h);
      }
    return KORE_OK;
  };
  KTRY(gram(src, G));
  if (kore_dense::cholesky_upper(nc, G.data(), nc) != 0) return KORE_WARN_BREAKDOWN;
  std::vector<zc> Ri = G;
  kore_dense::triu_inverse(nc, Ri.data(), nc);
  KTRY(kore_ws_gemm(w, dst, nc, src, nc, reinterpret_cast<const double*>(Ri.data()), 1.0, 0.0));
  if (nc == 1) {
    R = G;
    return KORE_OK;
  }
  KTRY(gram(dst, G2));
  if (kore_dense::cholesky_upper(nc, G2.data(), nc) != 0) return KORE_WARN_BREAKDOWN;
  std::vector<zc> Ri2 = G2;
  kore_dense::triu_inverse(nc, Ri2.data(), nc);
  KTRY(kore_ws_gemm(w, src, nc, dst, nc, reinterpret_cast<const double*>(Ri2.data()), 1.0, 0.0));
  KTRY(kore_ws_copy(w, src, dst, nc));
  R.assign((size_t)nc * nc, 0.0);   // R = R2 R1
  for (int c = 0; c < nc; ++c)
    for (int i = 0; i <= c; ++i) {
      zc acc = 0.0;
      for (int p = i; p <= c; ++p) acc += G2[(size_t)p * nc + i] * G[(size_t)c * nc + p];
      R[(size_t)c * nc + i] = acc;
    }
  return KORE_OK;
}

struct Outcome {
  bool fallback = false;   // a complex64 OP application needed too many refinement sweeps
  int fb_columns = 0, fb_ir = 0;
  Ritz rr;
  int nconv = 0, restarts = 0, op_columns = 0, op_calls = 0, ir_max = 0, breakdowns = 0;
  double host_s = 0;
};

// One Krylov-Schur run on the system's current factor; the final Ritz data
// stays in out.rr with the basis in the workspace w.
int32_t krylov_schur(kore_sys* sys, kore_ws* w, int64_t n, zc tau, const kore_eigs_opts& o, int bs, int m,
                     const double* v0, bool mixed, int fallback_sweeps, Outcome& out) {
  const int cW = m + bs, cT = m + 2 * bs;
  const int64_t ldh = m + bs;
  std::vector<zc> H((size_t)ldh * m, 0.0);
  const int nev = o.nev;
  std::vector<zc> R;
  if (v0) {
    KTRY(kore_ws_set(w, cW, 1, v0, n));
    if (bs > 1) KTRY(kore_ws_random(w, cW + 1, bs - 1, o.seed));
  } else {
    KTRY(kore_ws_random(w, cW, bs, o.seed));
  }
  {
    const int32_t rc = cholqr(w, cW, 0, bs, R);
    KTRY(rc);
    if (rc == KORE_WARN_BREAKDOWN) return err(KORE_ERR_ARG, "the start vector is zero or rank deficient");
  }
  int j = 0;
  const bool adaptive = o.check_every < 0;
  int check_every = adaptive ? std::max(4, 2 * bs) : o.check_every;
  if (check_every > 0) check_every = (check_every + bs - 1) / bs * bs;
  const int j_first = (nev + bs + bs - 1) / bs * bs;   // fewest columns worth a check
  int j_base = 0;                                       // columns at the last restart (no check there)
  double t_col = 0, t_rr = 0;
  int j_rr = 0;
  std::vector<std::pair<int, double>> trend;            // (columns, worst wanted residual) per check
  const double tol = o.tol;
  while (true) {
    bool have_rr = false;
    while (j < m) {
      int j_stop = m;
      const bool can_check = check_every > 0 && j >= j_first && j > j_base;
      const bool threaded = can_check && j + check_every < m;
      if (threaded) {
        int chunk = check_every;
        if (adaptive) {
          // a chunk long enough to cover the check (its cost grows as j^3) ...
          if (t_col > 0 && t_rr > 0)
            chunk = std::max(chunk, (int)std::ceil(t_rr * std::pow((double)j / j_rr, 3) / t_col / bs) * bs);
          // ... unless the worst wanted residual, falling at the last two
          // checks, reaches tol on its log-linear trend before that
          if (trend.size() >= 2 && trend.back().second < trend[trend.size() - 2].second) {
            const auto [ja, ra] = trend[trend.size() - 2];
            const auto [jb, rb] = trend.back();
            const double slope = (std::log(rb) - std::log(ra)) / (jb - ja);
            const double j_conv = jb + (std::log(tol) - std::log(rb)) / slope;
            const int to_conv = (int)std::ceil((j_conv - j) / bs) * bs;
            chunk = to_conv <= chunk ? std::max(bs, to_conv) : std::max(chunk, to_conv - check_every);
          }
          chunk = std::min(chunk, 4 * check_every);
        }
        j_stop = std::min(m, j + chunk);
      } else if (check_every > 0 && !can_check) {
        j_stop = std::min(m, std::max(j_first, j + check_every));   // grow to where a check is worth it
      }
      KTRY(kore_ws_stop_expand(w, -1));
      Ritz early;
      int32_t early_rc = KORE_ERR_STATE;
      double t_check = 0;
      std::thread th;
      std::vector<zc> Hs;
      if (threaded) {
        const int ld = j + bs;
        Hs.resize((size_t)ld * j);
        for (int c = 0; c < j; ++c)
          std::copy(H.begin() + (size_t)c * ldh, H.begin() + (size_t)c * ldh + ld, Hs.begin() + (size_t)c * ld);
        th = std::thread([&, j0 = j, jstop = j_stop, ld] {
          const double t = now_s();
          early_rc = rayleigh_ritz(Hs.data(), ld, j0, bs, tau, o.which, tol, early);
          t_check = now_s() - t;
          if (early_rc < 0) return;
          if (early.nconv >= nev) {
            kore_ws_stop_expand(w, 0);          // the rest of the chunk is not needed
          } else if (adaptive && !trend.empty()) {
            // this check's residual extends the trend: when it now reaches
            // tol before the chunk's end, the chunk ends there instead
            const double wr = early.worst(nev);
            const auto [ja, ra] = trend.back();
            if (std::isfinite(wr) && wr > 0 && wr < ra) {
              const double slope = (std::log(wr) - std::log(ra)) / (j0 - ja);
              const int j_end =
                  j0 + std::max(1, (int)std::ceil((std::log(tol) - std::log(wr)) / slope / bs)) * bs;
              if (j_end < jstop) kore_ws_stop_expand(w, j_end);
            }
          }
        });
      }
      int32_t jd = j;
      kore_solve_info inf{};
      const double t0 = now_s();
      const int32_t st = kore_ws_expand(w, j, j_stop, bs, cW, o.ir_tol, mixed ? o.ir_maxit : 0,
                                        fallback_sweeps > 0 ? fallback_sweeps : 0,
                                        reinterpret_cast<double*>(H.data()), ldh, &jd, &inf);
      const double t_exp = now_s() - t0;
      if (threaded) {
        th.join();
        if (early_rc < 0) return early_rc;
        t_rr = t_check;
        j_rr = j;
      }
      if (st < 0) return st;
      if (jd > j) t_col = t_exp / (jd - j);
      out.ir_max = std::max(out.ir_max, (int)inf.iterations);
      int done_blocks = (jd - j) / bs;
      if (st == KORE_WARN_IR_NOT_CONVERGED && fallback_sweeps > 0) {
        out.fallback = true;
        out.fb_columns = out.op_columns + bs * done_blocks;
        out.fb_ir = out.ir_max;
        return KORE_OK;
      }
      if (st == KORE_WARN_BREAKDOWN) done_blocks += 1;
      out.op_calls += done_blocks;
      out.op_columns += bs * done_blocks;
      if (threaded) {
        if (early.nconv >= nev) {   // converged before the chunk
          out.rr = std::move(early);
          have_rr = true;
          break;
        }
        const double worst = early.worst(nev);
        if (std::isfinite(worst) && worst > 0) trend.emplace_back(early.j, worst);
      }
      if (st == KORE_WARN_BREAKDOWN) {
        // invariant subspace: continue with a fresh random block, orthogonal
        // to the basis, coupled to it by a zero subdiagonal block
        ++out.breakdowns;
        const int nb_all = jd + bs;
        KTRY(kore_ws_random(w, cW, bs, o.seed + 7919u * (uint64_t)out.breakdowns));
        std::vector<zc> hh((size_t)nb_all * bs);
        for (int pass = 0; pass < 2; ++pass) {
          KTRY(kore_ws_dot(w, 0, nb_all, cW, bs, reinterpret_cast<double*>(hh.data())));
          KTRY(kore_ws_gemm(w, cW, bs, 0, nb_all, reinterpret_cast<const double*>(hh.data()), -1.0, 1.0));
        }
        const int32_t rc = cholqr(w, cW, nb_all, bs, R);
        KTRY(rc);
        if (rc == KORE_WARN_BREAKDOWN) return err(KORE_ERR_SINGULAR, "no direction left after a Krylov breakdown");
        for (int c = jd; c < jd + bs; ++c)
          for (int i = nb_all; i < nb_all + bs; ++i) H[(size_t)c * ldh + i] = 0.0;
        jd = nb_all;
      }
      j = jd;
    }
    if (!have_rr) {
      const double t = now_s();
      KTRY(rayleigh_ritz(H.data(), ldh, m, bs, tau, o.which, tol, out.rr));
      out.host_s += now_s() - t;
    }
    out.nconv = out.rr.nconv;
    if (out.nconv >= nev || out.restarts >= o.maxit) break;
    // thick restart: the invariant subspace of the 'keep' best Ritz values
    int keep = out.nconv + std::max(1, (m - out.nconv) / 2);
    keep = std::max(keep, nev);
    keep = (keep + bs - 1) / bs * bs;                // the expansion then ends exactly at m
    keep = std::max(bs, std::min(keep, m - bs));
    const double t = now_s();
    Ritz& r = out.rr;
    std::vector<int32_t> sel(m, 0);
    for (int i = 0; i < keep; ++i) sel[r.order[i]] = 1;
    int msel = 0;
    kore_dense::schur_reorder(m, r.S.data(), m, r.Q.data(), m, sel.data(), &msel);
    if (msel != keep) return err(KORE_ERR_STATE, "Schur reordering selected %d of %d", msel, keep);
    KTRY(kore_ws_gemm(w, cT, keep, 0, m, reinterpret_cast<const double*>(r.Q.data()), 1.0, 0.0));
    KTRY(kore_ws_copy(w, cT, 0, keep));
    KTRY(kore_ws_copy(w, m, keep, bs));
    std::vector<zc> Hn((size_t)ldh * m, 0.0);
    for (int c = 0; c < keep; ++c) {
      for (int i = 0; i <= c; ++i) Hn[(size_t)c * ldh + i] = r.S[(size_t)c * m + i];
      for (int i = 0; i < bs; ++i) {
        zc acc = 0.0;
        for (int q = 0; q < m; ++q) acc += r.Bres[(size_t)q * bs + i] * r.Q[(size_t)c * m + q];
        Hn[(size_t)c * ldh + keep + i] = acc;
      }
    }
    H.swap(Hn);
    out.host_s += now_s() - t;
    j = keep;
    j_base = keep;
    trend.clear();
    ++out.restarts;
  }
  return KORE_OK;
}

}  // namespace

extern "C" {

void kore_eigs_default_opts(kore_eigs_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->nev = 6;
  o->ncv = 0;
  o->tol = 1e-12;
  o->maxit = 100;
  o->which = KORE_WHICH_TM;
  o->block = 1;
  o->precision = KORE_PREC_FP64;
  o->seed = 0;
  o->check_every = -1;
  o->fallback_sweeps = 4;
  o->ir_tol = 1e-13;
  o->ir_maxit = 12;
}

int32_t kore_eigs(kore_sys* sys, double tau_re, double tau_im, const kore_eigs_opts* opts, const double* v0,
                  int32_t max_pairs, double* lam, double* X, int64_t ldx, double* errest, kore_eigs_info* info) {
  const double t_start = now_s();
  if (!sys || !opts || !lam || max_pairs < 0) return err(KORE_ERR_ARG, "bad kore_eigs arguments");
  kore_eigs_opts o = *opts;
  if (o.nev < 1) return err(KORE_ERR_ARG, "nev must be >= 1");
  if (o.which < KORE_WHICH_TM || o.which > KORE_WHICH_SI) return err(KORE_ERR_ARG, "unknown which %d", o.which);
  if (o.precision != KORE_PREC_FP64 && o.precision != KORE_PREC_C64) return err(KORE_ERR_ARG, "bad precision");
  kore_sys_state ss{};
  KTRY(kore_sys_query(sys, &ss));
  if (!ss.has_b) return err(KORE_ERR_STATE, "system has no B matrix");
  const int64_t n = ss.sizmat;
  if (X && ldx < n) return err(KORE_ERR_ARG, "ldx < sizmat");
  const zc tau(tau_re, tau_im);
  const int bs = std::max(1, std::min(o.block, 16));
  // basis size: a large space costs nothing until it is needed (the early
  // checks stop the expansion), and saves restarts
  int m = o.ncv > 0 ? o.ncv : std::max(std::max(2 * o.nev, o.nev + 15), 96);
  m = std::max(m, o.nev + 2 * bs);
  m = (m + bs - 1) / bs * bs;
  m = (int)std::min<int64_t>(m, std::max<int64_t>(bs, (n / bs) * bs - bs));
  if (m < o.nev + bs) return err(KORE_ERR_ARG, "system of %lld unknowns is too small for nev = %d",
                                 (long long)n, o.nev);
  // factor at tau unless the system already holds that factorisation
  if (!(ss.factored && ss.precision == o.precision && ss.sre == tau_re && ss.sim == tau_im))
    KTRY(kore_factor(sys, tau_re, tau_im, o.precision, nullptr));
  kore_ws* w = nullptr;
  KTRY(kore_ws_create(sys, 2 * m + 2 * bs, &w));
  struct WsGuard {
    kore_ws* w;
    ~WsGuard() { kore_ws_destroy(w); }
  } guard{w};
  const bool c64 = o.precision == KORE_PREC_C64;
  Outcome out;
  KTRY(krylov_schur(sys, w, n, tau, o, bs, m, v0, c64, c64 ? o.fallback_sweeps : 0, out));
  int used_prec = o.precision;
  int fb = 0, c64_cols = 0, c64_ir = 0;
  if (out.fallback) {
    // complex64 refinement too slow: refactor in FP64 at the same shift
    fb = 1;
    c64_cols = out.fb_columns;
    c64_ir = out.fb_ir;
    KTRY(kore_factor(sys, tau_re, tau_im, KORE_PREC_FP64, nullptr));
    used_prec = KORE_PREC_FP64;
    out = Outcome{};
    KTRY(krylov_schur(sys, w, n, tau, o, bs, m, v0, false, 0, out));
  } else if (c64) {
    c64_cols = out.op_columns;
    c64_ir = out.ir_max;
  }
  const Ritz& r = out.rr;
  const int k = std::min(out.nconv, max_pairs);
  zc* lz = reinterpret_cast<zc*>(lam);
  for (int i = 0; i < k; ++i) {
    lz[i] = tau + 1.0 / r.theta[i];
    if (errest) errest[i] = r.res[i] / std::abs(r.theta[i]);
  }
  if (X && k > 0) {
    // Ritz vectors y = Q z (unit norm: the basis is orthonormal, so ||V y|| = ||y||)
    const int j = r.j;
    std::vector<zc> Y((size_t)j * k, 0.0);
    for (int c = 0; c < k; ++c) {
      const int src = r.order[c];
      double nrm = 0;
      for (int i = 0; i < j; ++i) {
        zc acc = 0.0;
        for (int q = 0; q <= src; ++q) acc += r.Q[(size_t)q * j + i] * r.Z[(size_t)src * j + q];
        Y[(size_t)c * j + i] = acc;
        nrm += std::norm(acc);
      }
      nrm = std::sqrt(nrm);
      for (int i = 0; i < j; ++i) Y[(size_t)c * j + i] /= nrm;
    }
    const int cT = m + 2 * bs;
    KTRY(kore_ws_gemm(w, cT, k, 0, j, reinterpret_cast<const double*>(Y.data()), 1.0, 0.0));
    KTRY(kore_ws_get(w, cT, k, X, ldx));
  }
  if (info) {
    info->nconv = out.nconv;
    info->returned = k;
    info->restarts = out.restarts;
    info->op_columns = out.op_columns;
    info->op_calls = out.op_calls;
    info->ir_max_iterations = out.ir_max;
    info->precision = used_prec;
    info->fallback = fb;
    info->breakdowns = out.breakdowns;
    info->c64_op_columns = c64_cols;
    info->c64_ir_max = c64_ir;
    info->basis_columns = m;
    info->seconds = now_s() - t_start;
    info->host_seconds = out.host_s;
  }
  return out.nconv >= o.nev ? KORE_OK : KORE_WARN_NOT_CONVERGED;
}

int32_t kore_schur(int32_t n, double* T, int64_t ldt, double* Q, int64_t ldq) {
  if (n < 0 || !T || !Q || ldt < n || ldq < n) return err(KORE_ERR_ARG, "bad kore_schur arguments");
  const int bad = kore_dense::schur(n, reinterpret_cast<zc*>(T), ldt, reinterpret_cast<zc*>(Q), ldq);
  if (bad) return err(KORE_ERR_SINGULAR, "QR iteration left %d eigenvalues unconverged", bad);
  return KORE_OK;
}

int32_t kore_schur_reorder(int32_t n, double* S, int64_t lds, double* Q, int64_t ldq, const int32_t* select,
                           int32_t* m) {
  if (n < 0 || !S || !Q || !select || !m || lds < n || ldq < n) return err(KORE_ERR_ARG, "bad reorder arguments");
  int mm = 0;
  kore_dense::schur_reorder(n, reinterpret_cast<zc*>(S), lds, reinterpret_cast<zc*>(Q), ldq, select, &mm);
  *m = mm;
  return KORE_OK;
}

}  // extern "C"
