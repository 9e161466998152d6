// Host runtime of the Kore MI355X solver backend: C ABI, degree-major
// layout, block-cyclic-reduction plan, factor/solve, mixed-precision
// refinement and the device Krylov workspace.
//
// Division of labour: everything that streams sizmat-long data or touches
// the factor runs on the device (kernels_dense.hip, kernels_vec.hip);
// the host builds the permutation, the union CSR pattern of A and B in
// degree-major order, the reduction plan (descriptor lists uploaded once
// per factorisation) and the row equilibration.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <atomic>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>
#include <thread>
#include <type_traits>

#include <sched.h>

#include "kore_amd.h"
#include "kore_device.hpp"
#include "kore_kernels.hpp"

using namespace kore;
using zd = std::complex<double>;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int32_t fail(int32_t code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// the error text of this host thread, set from the library's other sources
int32_t kore_internal_set_error(int32_t code, const char* msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(...)                                                                             \
  do {                                                                                           \
    hipError_t _e = (__VA_ARGS__);                                                               \
    if (_e != hipSuccess)                                                                        \
      return fail(KORE_ERR_HIP, "%s failed: %s (%s:%d)", #__VA_ARGS__, hipGetErrorString(_e), __FILE__, \
                  __LINE__);                                                                     \
  } while (0)

#define TRY(...)                  \
  do {                            \
    int32_t _r = (__VA_ARGS__);   \
    if (_r < 0) return _r;        \
  } while (0)

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
// Host-side stage timer: prints "kore_amd: <stage> <ms>" to stderr when the
// environment variable KORE_AMD_TRACE is set (read once).
struct StageTrace {
  bool on;
  std::chrono::steady_clock::time_point t0;
  explicit StageTrace() : on(std::getenv("KORE_AMD_TRACE") != nullptr), t0(std::chrono::steady_clock::now()) {}
  void mark(const char* what) {
    if (!on) return;
    const auto t = std::chrono::steady_clock::now();
    fprintf(stderr, "kore_amd: %s %.2f ms\n", what, std::chrono::duration<double, std::milli>(t - t0).count());
    t0 = t;
  }
};

struct TimerRec {
  int cls;
  hipEvent_t a, b;
};

// ---------------------------------------------------------------------------
// memory pools
// ---------------------------------------------------------------------------
// Device and pinned-host allocations are cached for the life of the process
// and shared by every context on a device: a request stream allocates the
// factor, system and workspace buffers of each request (gigabytes at the
// serving sizes), and hipMalloc / hipFree of such blocks cost milliseconds to
// a second each, with hipFree waiting for the whole device, so that requests
// on other streams stall.  A block goes back to its pool only after the
// releasing stream has drained (objects are released after their stream is
// synchronised, so this wait is normally free), so any context may take it
// at once.  Best fit: a request takes the smallest cached block that is at
// least as large and at most twice as large (plus a little).  When the device
// runs out of memory the cached blocks are returned to the runtime and the
// allocation is tried once more.
struct PoolBlock {
  void* p;
  size_t bytes;
};

struct MemPool {
  std::mutex mu;
  std::multimap<size_t, PoolBlock> blocks;
  size_t cached = 0;
};

static constexpr int POOL_MAXDEV = 64;
static MemPool g_dev_pool[POOL_MAXDEV];
static MemPool g_host_pool;
static std::atomic<int64_t> g_pool_misses{0};

static size_t pool_round(size_t n) {
  const size_t g = n >= (size_t(64) << 20) ? (size_t(2) << 20) : 4096;   // 2 MiB granules for large blocks
  return (n + g - 1) / g * g;
}

static bool pool_pick(MemPool& mp, size_t need, PoolBlock* out) {
  std::lock_guard<std::mutex> lk(mp.mu);
  auto it = mp.blocks.lower_bound(need);
  if (it == mp.blocks.end() || it->first > 2 * need + (size_t(64) << 20)) return false;
  *out = it->second;
  mp.cached -= it->second.bytes;
  mp.blocks.erase(it);
  return true;
}

static void pool_put(MemPool& mp, void* p, size_t bytes, hipStream_t st) {
  if (st) hipStreamSynchronize(st);      // no work of the releasing stream may still use the block
  std::lock_guard<std::mutex> lk(mp.mu);
  mp.blocks.emplace(bytes, PoolBlock{p, bytes});
  mp.cached += bytes;
}

// Returns every cached block of the pool to the runtime.
static void pool_trim(MemPool& mp, bool host) {
  std::multimap<size_t, PoolBlock> all;
  {
    std::lock_guard<std::mutex> lk(mp.mu);
    all.swap(mp.blocks);
    mp.cached = 0;
  }
  for (auto& kv : all) {
    if (host)
      hipHostFree(kv.second.p);
    else
      hipFree(kv.second.p);
  }
}

// Device block of at least `need` bytes for work on `st` (device current).
static int32_t dev_take(int device, hipStream_t st, size_t need, void** p, size_t* got) {
  MemPool& mp = g_dev_pool[device % POOL_MAXDEV];
  PoolBlock b{};
  if (pool_pick(mp, need, &b)) {
    *p = b.p;
    *got = b.bytes;
    return KORE_OK;
  }
  ++g_pool_misses;
  const size_t sz = pool_round(need);
  hipError_t e = hipMalloc(p, sz);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    pool_trim(mp, false);
    e = hipMalloc(p, sz);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(KORE_ERR_OOM, "hipMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
  }
  *got = sz;
  return KORE_OK;
}

// Pinned host block of at least `need` bytes.
static int32_t host_take(size_t need, char** p, size_t* got) {
  PoolBlock b{};
  if (pool_pick(g_host_pool, need, &b)) {
    *p = static_cast<char*>(b.p);
    *got = b.bytes;
    return KORE_OK;
  }
  const size_t sz = pool_round(need);
  hipError_t e = hipHostMalloc((void**)p, sz, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    pool_trim(g_host_pool, true);
    e = hipHostMalloc((void**)p, sz, hipHostMallocDefault);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(KORE_ERR_OOM, "hipHostMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
  }
  *got = sz;
  return KORE_OK;
}

// A pinned host buffer from the pool, returned to it (after `st` drains)
// when released.
struct PinnedBuf {
  char* p = nullptr;
  size_t bytes = 0;
  hipStream_t st = nullptr;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  int32_t reserve(size_t need, hipStream_t stream) {
    if (need <= bytes) return KORE_OK;
    release();
    st = stream;
    return host_take(need, &p, &bytes);
  }
  void release() {
    if (p) pool_put(g_host_pool, p, bytes, st);
    p = nullptr;
    bytes = 0;
  }
};

// ---------------------------------------------------------------------------
// allocation-free residuals: scratch in a static device array
// ---------------------------------------------------------------------------
// kore_residual_noalloc keeps its staging, gathered and residual vectors and
// the partial sums of its norms in this file-scope array instead of device
// memory of the system or the library's pool.  Each device holds its own
// copy at its own address, found once per device by hipGetSymbolAddress (the
// array's name in host code is only a stand-in for the symbol).  Every call on
// a device shares that device's copy: calls that overlap on one device race.
static constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];
static void* g_scratch_ptr[POOL_MAXDEV];
static std::mutex g_scratch_mu;

// The device address of `scratch` on `device` (which must be current).
static int32_t scratch_address(int device, char** out) {
  if (device < 0 || device >= POOL_MAXDEV) return fail(KORE_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(KORE_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                  hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = static_cast<char*>(g_scratch_ptr[device]);
  return KORE_OK;
}

struct kore_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t budget = 0;
  std::atomic<int64_t> in_use{0};
  bool timing = false;
  // replay solve sweeps from captured HIP graphs: off by default, since the
  // replay of one sweep measured slower than its eager launches on MI355X
  // (43.2 against 39.7 ms median per forced request at N = 500)
  bool graphs = false;
  std::vector<TimerRec> pending;
  std::vector<hipEvent_t> pool;
  double tot_ms[KORE_TIMER_COUNT] = {0};
  int64_t cnt[KORE_TIMER_COUNT] = {0};
  // pinned host arena for building and uploading system matrices: kept for
  // the life of the context (pre-faulted, DMA-able), grown on demand
  PinnedBuf pin;
};

static int32_t ctx_pinned(kore_ctx* c, size_t bytes, char** out) {
  if (bytes > c->pin.bytes) TRY(c->pin.reserve(bytes + bytes / 4, c->stream));
  *out = c->pin.p;
  return KORE_OK;
}

static hipEvent_t get_event(kore_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct Timed {
  kore_ctx* c;
  int cls;
  hipEvent_t a = nullptr;
  Timed(kore_ctx* ctx, int k) : c(ctx), cls(k) {
    if (c->timing) {
      a = get_event(c);
      if (a) hipEventRecord(a, c->stream);
    }
  }
  ~Timed() {
    if (c->timing && a) {
      hipEvent_t b = get_event(c);
      if (b) {
        hipEventRecord(b, c->stream);
        c->pending.push_back({cls, a, b});
      }
    }
  }
};

static void resolve_timers(kore_ctx* c) {
  if (c->pending.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto& t : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      c->tot_ms[t.cls] += ms;
      c->cnt[t.cls] += 1;
    }
    c->pool.push_back(t.a);
    c->pool.push_back(t.b);
  }
  c->pending.clear();
}

// Device buffer from the pool of the context's device; the context's budget
// counts the bytes asked for.
struct DevBuf {
  kore_ctx* c = nullptr;
  void* p = nullptr;
  size_t bytes = 0;     // bytes asked for
  size_t block = 0;     // bytes of the pool block
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  int32_t alloc(kore_ctx* ctx, size_t nbytes) {
    release();
    c = ctx;
    if (nbytes == 0) return KORE_OK;
    if (c->budget > 0 && c->in_use.load() + (int64_t)nbytes > c->budget)
      return fail(KORE_ERR_OOM, "device memory budget exceeded: need %zu bytes, %lld of %lld in use", nbytes,
                  (long long)c->in_use.load(), (long long)c->budget);
    TRY(dev_take(c->device, c->stream, nbytes, &p, &block));
    bytes = nbytes;
    c->in_use += (int64_t)nbytes;
    return KORE_OK;
  }
  void release() {
    if (p) {
      pool_put(g_dev_pool[c->device % POOL_MAXDEV], p, block, c->stream);
      c->in_use -= (int64_t)bytes;
    }
    p = nullptr;
    bytes = 0;
    block = 0;
  }
  template <class U> U* as() const { return reinterpret_cast<U*>(p); }
};

// Pinned staging for small host->device uploads (descriptors, small
// matrices) so that they stay asynchronous on the library stream.
struct Uploader {
  kore_ctx* c = nullptr;
  PinnedBuf buf;
  size_t cap = 0, off = 0;
  int32_t init(kore_ctx* ctx, size_t bytes) {
    c = ctx;
    cap = bytes;
    return buf.reserve(cap, c->stream);
  }
  int32_t put(void* dst, const void* src, size_t nbytes) {
    if (nbytes > cap) {
      HIP_TRY(hipMemcpyAsync(dst, src, nbytes, hipMemcpyHostToDevice, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      return KORE_OK;
    }
    if (off + nbytes > cap) {
      HIP_TRY(hipStreamSynchronize(c->stream));
      off = 0;
    }
    memcpy(buf.p + off, src, nbytes);
    HIP_TRY(hipMemcpyAsync(dst, buf.p + off, nbytes, hipMemcpyHostToDevice, c->stream));
    off += (nbytes + 255) & ~size_t(255);
    return KORE_OK;
  }
};

// Pinned staging for downloads: device -> pinned at DMA rate, then a copy into
// the caller's (pageable) array split over a few host threads.
struct Downloader {
  PinnedBuf buf;
  int32_t reserve(size_t bytes, hipStream_t st) { return buf.reserve(bytes, st); }
  char* host() const { return buf.p; }
};

struct CopySeg {
  char* dst;
  const char* src;
  size_t bytes;
};

// Host threads one call of the library's parallel host loops may use:
// kore_set_host_threads, else KORE_AMD_HOST_THREADS, else the process's CPU
// share (the cgroup's CPU quota where one is set, the affinity mask
// otherwise), at most 16.  hardware_concurrency alone counts every core of
// the machine; a process that runs more threads than its quota is throttled
// as a whole, its launching threads included.
static std::atomic<int> g_host_threads{0};

static int cpu_share() {
  static const int share = [] {
    int n = (int)std::max(1u, std::thread::hardware_concurrency());
    cpu_set_t cs;
    CPU_ZERO(&cs);
    if (sched_getaffinity(0, sizeof(cs), &cs) == 0) n = std::min(n, std::max(1, CPU_COUNT(&cs)));
    if (FILE* f = fopen("/sys/fs/cgroup/cpu.max", "r")) {
      char q[32] = {0};
      long long per = 0;
      if (fscanf(f, "%31s %lld", q, &per) == 2 && strcmp(q, "max") != 0 && per > 0) {
        const long long quota = atoll(q);
        if (quota > 0) n = (int)std::min<long long>(n, std::max<long long>(1, quota / per));
      }
      fclose(f);
    }
    return n;
  }();
  return share;
}

static int host_threads() {
  const int v = g_host_threads.load(std::memory_order_relaxed);
  if (v > 0) return v;
  static const int dflt = [] {
    if (const char* e = std::getenv("KORE_AMD_HOST_THREADS")) {
      const int x = std::atoi(e);
      if (x > 0) return std::min(x, 64);
    }
    return std::min(cpu_share(), 16);
  }();
  return dflt;
}

// memcpy of a list of segments, the total split evenly over up to 8 threads
// (at least 1 MiB each)
static void copy_parallel(const std::vector<CopySeg>& segs) {
  size_t total = 0;
  for (const CopySeg& g : segs) total += g.bytes;
  const int nt = (int)std::min<size_t>(std::min(8, host_threads()), std::max<size_t>(1, total >> 20));
  auto range = [&](size_t lo, size_t hi) {
    size_t base = 0;
    for (const CopySeg& g : segs) {
      const size_t s0 = std::max(lo, base), s1 = std::min(hi, base + g.bytes);
      if (s0 < s1) memcpy(g.dst + (s0 - base), g.src + (s0 - base), s1 - s0);
      base += g.bytes;
    }
  };
  if (nt <= 1) {
    range(0, total);
    return;
  }
  const size_t per = (total + nt - 1) / nt;
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(range, t * per, std::min(total, (t + 1) * per));
  range(0, per);
  for (std::thread& x : th) x.join();
}

// ---------------------------------------------------------------------------
// layout
// ---------------------------------------------------------------------------
static void ell_lists(int m, int lmax, int vsymm, std::vector<int>& top, std::vector<int>& bot,
                      std::vector<int>& all) {
  const int lm1 = lmax - m + 1;
  const int s = vsymm > 0 ? 1 : 0;
  const int sg = m > 0 ? 1 : 0;
  all.clear();
  top.clear();
  bot.clear();
  for (int l = m + 1 - sg; l < lmax + 2 - sg; ++l) all.push_back(l);
  const int first = (sg + s) % 2;
  for (int i = first; i < lm1; i += 2) top.push_back(all[i]);
  for (int i = 1 - first; i < lm1; i += 2) bot.push_back(all[i]);
}

static int32_t layout_build(const kore_layout* L, int64_t* sizmat, std::vector<int>* sizes,
                            std::vector<int64_t>* perm) {
  if (!L) return fail(KORE_ERR_ARG, "layout is NULL");
  if (L->N1 <= 0 || L->m < 0 || L->lmax < L->m) return fail(KORE_ERR_ARG, "bad layout sizes");
  if ((L->lmax - L->m + 1) % 2 != 0)
    return fail(KORE_ERR_ARG, "lmax - m + 1 must be even (lmax=%d, m=%d)", L->lmax, L->m);
  if ((L->symm != 1 && L->symm != -1) || (L->symmB0 != 1 && L->symmB0 != -1))
    return fail(KORE_ERR_ARG, "symm and symmB0 must be +1 or -1");
  if (!(L->hydro || L->magnetic || L->thermal || L->compositional)) return fail(KORE_ERR_ARG, "no fields");
  std::vector<int> top, bot, all;
  ell_lists(L->m, L->lmax, L->symm, top, bot, all);
  const int64_t nb = (L->lmax - L->m + 1) / 2;
  const int64_t n = (int64_t)L->N1 * nb;
  struct Sec {
    int64_t off;
    const std::vector<int>* lst;
  };
  std::vector<Sec> secs;
  int64_t off = 0;
  if (L->hydro) {
    secs.push_back({off, &top});
    secs.push_back({off + n, &bot});
    off += 2 * n;
  }
  if (L->magnetic) {
    // the induced field's poloidal (f) and toroidal (g) sections use the flow's
    // degree lists: f on the toroidal ones and g on the poloidal ones when B0 is
    // equatorially antisymmetric, the other way round otherwise (bin/assemble.py:69-74)
    const std::vector<int>* f_l = L->symmB0 == 1 ? &top : &bot;
    const std::vector<int>* g_l = L->symmB0 == 1 ? &bot : &top;
    secs.push_back({off, f_l});
    secs.push_back({off + n, g_l});
    off += 2 * n;
  }
  if (L->thermal) {
    secs.push_back({off, &top});
    off += n;
  }
  if (L->compositional) {
    secs.push_back({off, &top});
    off += n;
  }
  *sizmat = off;
  if (sizes) sizes->clear();
  if (perm) {
    perm->clear();
    perm->reserve(off);
  }
  for (int l : all) {
    int sz = 0;
    for (auto& s : secs) {
      auto it = std::find(s.lst->begin(), s.lst->end(), l);
      if (it == s.lst->end()) continue;
      const int64_t idx = it - s.lst->begin();
      if (perm)
        for (int r = 0; r < L->N1; ++r) perm->push_back(s.off + idx * L->N1 + r);
      sz += L->N1;
    }
    if (sz && sizes) sizes->push_back(sz);
  }
  if (perm && (int64_t)perm->size() != off) return fail(KORE_ERR_ARG, "internal: permutation size mismatch");
  return KORE_OK;
}

// ---------------------------------------------------------------------------
// factor (block cyclic reduction)
// ---------------------------------------------------------------------------
struct FactorBase {
  virtual ~FactorBase() = default;
  int prec = 0;
  double flops = 0, flops_inv = 0, flops_gemm = 0, flops_sparse = 0, solve_bytes = 0;
  int solve_launches = 0;
  int64_t bytes = 0;
  int nlev = 0;
  // lean form: no E = L Dinv, F = U Dinv products; the kept blocks are
  // updated through G, H and a sweep's forward level takes two launches
  bool lean = false;
};

template <class T> struct Level {
  InvDesc<T>* inv = nullptr;
  int ninv = 0, maxn_inv = 0, minn_inv = 1 << 30;
  GemmDesc<T>* g1 = nullptr;
  int ng1 = 0, t1 = 0;
  GemmDesc<T>* g2 = nullptr;
  int ng2 = 0, t2 = 0;
  SkinnyDesc<T>* fw = nullptr;
  int nfw = 0, mfw = 0;
  SkinnyDesc<T>* fy = nullptr;          // lean form: y_odd = Dinv_odd b_odd before fw
  int nfy = 0, mfy = 0;
  SkinnyDesc<T>* bw = nullptr;
  int nbw = 0, mbw = 0;
  // level 0 through the sparse couplings
  SpColDesc<T>* sc = nullptr;
  size_t sc_off = 0;
  int nsc = 0, maxn_sc = 0;
  SpRowDesc<T>* sr = nullptr;
  size_t sr_off = 0;
  int nsr = 0, maxm_sr = 0, maxn_sr = 0;
};

template <class T> struct Factor : FactorBase {
  // solve sweeps captured as HIP graphs, one per column count (1..16): a
  // sweep's launches read only device-resident descriptors and this factor's
  // fixed vectors, so one capture serves every later sweep of the factor
  hipGraphExec_t sweep_graph[17] = {};
  int sprow_nc = 8;                     // column tile of the level-0 sparse-row products (tuned)
  hipEvent_t tune_ev[2] = {nullptr, nullptr};   // set while the tuner times the products
  ~Factor() override {
    for (hipGraphExec_t g : sweep_graph)
      if (g) hipGraphExecDestroy(g);
  }
  DevBuf mem, aux, desc, vecs;
  std::vector<cx<T>*> D, L, U, G, H;   // G = Dinv L, H = Dinv U of eliminated blocks
  std::vector<Level<T>> lev;
  InvDesc<T>* root_inv = nullptr;
  SkinnyDesc<T>* root_solve = nullptr;
  int root_n = 0;
  // Two-block tail: the last level (blocks k0, k1) and the root solved in
  // one launch, X0 = Dinv0 B0 - P B1 and X1 = Q B1 - R B0, with
  // P = Dinv0 F0, R = G1 Dinv0, Q = Dinv1 + G1 P formed after the root
  // inversion (P in G[k0], R in H[k1], Q over D[k1]).
  bool tail = false;
  GemmDesc<T>* tail1 = nullptr;     // P and R
  GemmDesc<T>* tail2 = nullptr;     // Q
  int tail_tiles1 = 0, tail_tiles2 = 0, tail_m = 0;
  SkinnyDesc<T>* tail_solve = nullptr;   // two descriptors
  // Level 0 of a sweep with one right-hand side and an FP64 factor (at least
  // two levels before the tail): its couplings L, U are the sparse
  // off-diagonal blocks of the matrix itself, so the sweep applies them from
  // the CSR and reads only the level's inverse blocks, twice:
  //   forward   y_odd = Dinv_odd b_odd;  b_even -= L y  (sparse)
  //   backward  b_odd -= L x_even (sparse);  x_odd = Dinv_odd b_odd
  // instead of E, F (forward) and Dinv, G, H (backward) once each.
  bool lev0_sparse = false;
  SkinnyDesc<T>* lev0_dinv = nullptr;    // x_odd (or y_odd) = Dinv_odd b_odd, one descriptor per odd block
  int lev0_n = 0, lev0_m = 0;
  cx<T>** dD = nullptr;
  cx<T>** dL = nullptr;
  cx<T>** dU = nullptr;
  int* info = nullptr;
  cx<T>* Bv = nullptr;  // solve vectors, sizmat x 16, degree-major
  cx<T>* Xv = nullptr;
  int nblk = 0;
  int64_t ndlu = 0;    // elements of the D/L/U region (cleared before the scatter)
};

struct kore_sys {
  kore_ctx* ctx = nullptr;
  kore_layout lay{};
  int64_t n = 0;
  int nblk = 0, nmax = 0;
  std::vector<int> bsz;
  std::vector<int64_t> boff;
  std::vector<int64_t> perm;
  std::vector<int32_t> h_ip, h_blk;
  std::vector<int32_t> h_amap;     // union-pattern position of each entry of the A given at creation (-1: dropped)
  int64_t nnz = 0;                 // entries of the union pattern (device arrays only)
  int64_t off_nnz[2] = {0, 0};     // entries outside the diagonal block, rows of even / odd blocks
  int64_t off_rows[2] = {0, 0};    // rows of even / odd blocks
  // per row: entries before its diagonal block (bits 0-14), after it (bits
  // 15-29), block parity (bit 30); the row's entries are sorted by column
  std::vector<int32_t> h_split;
  DevBuf d_split;
  bool has_b = false;
  DevBuf d_ip, d_ix, d_va, d_vb, d_perm, d_blk, d_boff, d_scale, d_scr, d_part;
  CsrDev csr{};
  DevBuf d_ipb, d_ixb, d_vbb;      // B alone (degree order): the B x of every OP application
  CsrDev csr_b{};
  // entries outside each row's diagonal block (the level-0 couplings) as a
  // compact CSR (values per shift) and per (block, side) column segments
  int64_t nnz_off = 0;
  DevBuf d_optr, d_onlo, d_ocol, d_oval, d_cptr, d_cq, d_crow;
  std::vector<int64_t> h_cbase;    // 2 per block: first column-pointer slot of (block, side), -1 if no neighbour
  std::vector<int64_t> h_side_nnz; // 2 per block: entries of (block, side)
  OffCsr off{};
  OffCsc offc{};
  // scratch vectors (complex128, sizmat x 16 each): bb (rhs), xr (solution),
  // rr (residual), st (staging / scaled residual)
  cx<double>* bb = nullptr;
  cx<double>* xr = nullptr;
  cx<double>* rr = nullptr;
  cx<double>* st = nullptr;
  double* part = nullptr;
  int npart = 64;
  int32_t form = KORE_FORM_AUTO;        // kore_sys_set_factor_form
  std::unique_ptr<FactorBase> f;
  std::unique_ptr<FactorBase> fspare;   // the plan of the other precision, kept for the next switch
  bool factored = false;     // f holds a valid factorisation of A - sigma B
  double sre = 0, sim = 0;
  std::vector<double> h_scale;
  Uploader up;
  Downloader down;           // pinned staging of kore_ws_get
  DevBuf gm;                 // GMRES basis (restart+1 vectors), dot partials, coefficients
  int gm_cols = 0;
  std::mutex ws_mu;
  std::vector<kore_ws*> wss; // live workspaces: destroyed with the system at the latest
  // stream-ordered entry points: the caller's stream and the context's stream
  // wait for each other through these events
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  PinnedBuf info_pin;        // pivot flags of a queued factorisation (read by kore_factor_check)
  bool info_pending = false;
  ~kore_sys() {
    if (ev_in) hipEventDestroy(ev_in);
    if (ev_out) hipEventDestroy(ev_out);
  }
};

struct kore_ws {
  kore_sys* sys = nullptr;
  int ncol = 0;
  DevBuf V, part, qbuf, descb;
  DevBuf hb;                 // device-side Gram-Schmidt / Cholesky-QR coefficients of the expansion
  cx<double>* hh = nullptr;  // pinned host copies of them, two slots (blocks in flight)
  hipEvent_t hev[2] = {nullptr, nullptr};
  PinnedBuf hhbuf;           // pinned pool block behind hh
  Uploader up;
  std::atomic<int> stop_at{INT32_MAX};  // set from another thread: the running expansion starts no block at or past it
  ~kore_ws() {
    for (hipEvent_t e : hev)
      if (e) hipEventDestroy(e);
  }
};

static constexpr int MAXR = 16;
// degree-block size from which KORE_FORM_AUTO factors in the lean form
static constexpr int LEAN_AUTO_MIN = 320;
static constexpr int DOT_MAXCHUNK = 1024;   // split-K chunks of the workspace inner products
static constexpr int DOT_ROWS = 64;        // minimum rows per chunk

// Element offsets of the factor's dense blocks (shared by the build and the
// byte prediction).  With the sparse level 0 (at least two levels before the
// two-block tail) the level's couplings are read from the matrix, so the
// eliminated blocks of level 0 hold no L, U and the level has no E, F.
struct FactorPlan {
  std::vector<std::vector<int>> ids;   // block ids per level, then the root
  int nlev = 0;
  bool sp0 = false;
  std::vector<int64_t> offD, offL, offU, offG, offH;
  std::vector<std::vector<int64_t>> offE, offF;
  int64_t cntD = 0, cntDLU = 0, cnt = 0;
};

static void factor_plan(const std::vector<int>& bsz, FactorPlan& P, bool allow_sparse0, bool lean) {
  const int nb = (int)bsz.size();
  int nmax = 0;
  for (int b : bsz) nmax = std::max(nmax, b);
  P.ids.clear();
  {
    std::vector<int> cur(nb);
    for (int k = 0; k < nb; ++k) cur[k] = k;
    while (cur.size() > 1) {
      P.ids.push_back(cur);
      std::vector<int> nx;
      for (size_t t = 0; t < cur.size(); t += 2) nx.push_back(cur[t]);
      cur.swap(nx);
    }
    P.ids.push_back(cur);  // root
  }
  P.nlev = (int)P.ids.size() - 1;
  P.sp0 = allow_sparse0 && P.nlev >= 2;
  auto nk = [&](int k) { return (int64_t)bsz[k]; };
  int64_t cnt = 0;
  P.offD.assign(nb, -1);
  P.offL.assign(nb, -1);
  P.offU.assign(nb, -1);
  P.offG.assign(nb, -1);
  P.offH.assign(nb, -1);
  for (int k = 0; k < nb; ++k) {
    P.offD[k] = cnt;
    cnt += nk(k) * nk(k);
  }
  P.cntD = cnt;
  for (int k = 0; k < nb; ++k) {
    if (P.sp0 && (k & 1)) continue;        // eliminated at level 0: couplings read from the matrix
    if (k > 0) {
      P.offL[k] = cnt;
      cnt += nk(k) * nmax;
    }
    if (k < nb - 1) {
      P.offU[k] = cnt;
      cnt += nk(k) * nmax;
    }
  }
  P.cntDLU = cnt;
  for (int k = 0; k < nb; ++k) {
    P.offG[k] = cnt;
    cnt += nk(k) * nmax;
    P.offH[k] = cnt;
    cnt += nk(k) * nmax;
  }
  P.offE.assign(P.nlev, {});
  P.offF.assign(P.nlev, {});
  for (int l = 0; l < P.nlev; ++l) {
    const auto& id = P.ids[l];
    const int T_ = (int)id.size();
    P.offE[l].assign(T_, -1);
    P.offF[l].assign(T_, -1);
    if (l == 0 && P.sp0) continue;
    // sweep form: E, F of the kept (even) positions.  Lean form: the
    // couplings L, U of every position of the levels after the first one on
    // dense couplings (that one reads them from the per-block L, U)
    if (lean && l == (P.sp0 ? 1 : 0)) continue;
    for (int t = 0; t < T_; t += lean ? 1 : 2) {
      if (t - 1 >= 0) {
        P.offE[l][t] = cnt;
        cnt += nk(id[t]) * nk(id[t - 1]);
      }
      if (t + 1 <= T_ - 1) {
        P.offF[l][t] = cnt;
        cnt += nk(id[t]) * nk(id[t + 1]);
      }
    }
  }
  P.cnt = cnt;
}

template <class T> static int32_t factor_build(kore_sys* s, Factor<T>* F) {
  kore_ctx* c = s->ctx;
  const int nb = s->nblk, nmax = s->nmax;
  F->nblk = nb;
  FactorPlan P;
  factor_plan(s->bsz, P, s->nnz_off > 0, F->lean);
  const auto& ids = P.ids;
  const int nlev = P.nlev;
  F->nlev = nlev;
  const bool sp0 = P.sp0;
  auto nk = [&](int k) { return (int64_t)s->bsz[k]; };
  TRY(F->mem.alloc(c, (size_t)P.cnt * sizeof(cx<T>)));
  cx<T>* base = F->mem.template as<cx<T>>();
  F->D.resize(nb);
  F->L.assign(nb, nullptr);
  F->U.assign(nb, nullptr);
  F->G.assign(nb, nullptr);
  F->H.assign(nb, nullptr);
  for (int k = 0; k < nb; ++k) {
    F->D[k] = base + P.offD[k];
    if (P.offL[k] >= 0) F->L[k] = base + P.offL[k];
    if (P.offU[k] >= 0) F->U[k] = base + P.offU[k];
    F->G[k] = base + P.offG[k];
    F->H[k] = base + P.offH[k];
  }
  // aux: ipiv (nb * nmax ints), info (nb ints), pointer arrays, solve vectors
  const size_t auxbytes = (size_t)nb * nmax * sizeof(int) + (size_t)nb * sizeof(int) + 3 * (size_t)nb * sizeof(void*) + 256;
  TRY(F->aux.alloc(c, auxbytes));
  char* ab = F->aux.template as<char>();
  int* ipiv = reinterpret_cast<int*>(ab);
  F->info = ipiv + (size_t)nb * nmax;
  char* pp = reinterpret_cast<char*>(F->info + nb);
  pp = reinterpret_cast<char*>(((uintptr_t)pp + 63) & ~uintptr_t(63));
  F->dD = reinterpret_cast<cx<T>**>(pp);
  F->dL = F->dD + nb;
  F->dU = F->dL + nb;
  TRY(F->vecs.alloc(c, 2 * (size_t)s->n * MAXR * sizeof(cx<T>)));
  F->Bv = F->vecs.template as<cx<T>>();
  F->Xv = F->Bv + (size_t)s->n * MAXR;
  F->bytes = (int64_t)(F->mem.bytes + F->aux.bytes + F->vecs.bytes);

  // descriptors
  std::vector<InvDesc<T>> vinv;
  std::vector<GemmDesc<T>> vg;
  std::vector<SkinnyDesc<T>> vs;
  std::vector<SpColDesc<T>> vsc;
  std::vector<SpRowDesc<T>> vsr;
  struct Rng {
    size_t off, n;
  };
  std::vector<Rng> rinv(nlev + 1), rg1(nlev), rg2(nlev), rfw(nlev), rbw(nlev), rfy(nlev, Rng{0, 0});
  F->lev.resize(nlev);
  double flops = 0;
  auto tiles = [](int m, int n) { return ((m + 63) / 64) * ((n + 63) / 64); };
  auto mkgemm = [&](cx<T>* C, int m, int n, const cx<T>* A1, const cx<T>* B1, int k1, int lda1, int ldb1,
                    const cx<T>* A2, const cx<T>* B2, int k2, int lda2, int ldb2, T alpha, T beta) {
    GemmDesc<T> g{};
    g.A[0] = A1; g.B[0] = B1; g.k[0] = A1 ? k1 : 0; g.lda[0] = lda1; g.ldb[0] = ldb1;
    g.A[1] = A2; g.B[1] = B2; g.k[1] = A2 ? k2 : 0; g.lda[1] = lda2; g.ldb[1] = ldb2;
    g.C = C; g.ldc = m; g.m = m; g.n = n; g.opA = 0;
    g.tiles_m = (m + 63) / 64; g.tiles_n = (n + 63) / 64;
    g.alpha = alpha; g.beta = beta;
    flops += 8.0 * m * n * ((A1 ? k1 : 0) + (A2 ? k2 : 0));
    F->flops_gemm += 8.0 * m * n * ((A1 ? k1 : 0) + (A2 ? k2 : 0));
    return g;
  };
  auto vecp = [&](cx<T>* v, int blk) { return v + s->boff[blk]; };
  const int ldv = (int)s->n;
  struct Prod {
    const cx<T>* A;
    int lda;
    const cx<T>* X;
    int k;
    T alpha;
  };
  auto mkskin = [&](cx<T>* Y, int m, std::initializer_list<Prod> prods, T beta) {
    SkinnyDesc<T> d{};
    int p = 0;
    for (const Prod& q : prods) {
      if (!q.A) continue;
      d.A[p] = q.A; d.lda[p] = q.lda; d.X[p] = q.X; d.ldx[p] = ldv; d.k[p] = q.k; d.alpha[p] = q.alpha;
      F->solve_bytes += (double)sizeof(cx<T>) * m * q.k;
      ++p;
    }
    d.Y = Y; d.ldy = ldv; d.m = m; d.beta = beta;
    return d;
  };
  // sparse level-0 products: C = X S (one side of block k) and C = beta C + alpha sum S_p X_p
  auto mkspcol = [&](cx<T>* C, const cx<T>* X, int k, int side, T alpha, T beta) {
    const int nbk = side == 0 ? k - 1 : k + 1;
    SpColDesc<T> d{};
    d.X = X; d.ldx = nk(k); d.C = C; d.ldc = nk(k); d.cbase = s->h_cbase[2 * k + side];
    d.m = (int)nk(k); d.n = (int)nk(nbk); d.alpha = alpha; d.beta = beta;
    const double fl = 8.0 * (double)s->h_side_nnz[2 * k + side] * nk(k);
    flops += fl;
    F->flops_sparse += fl;
    return d;
  };
  struct SProd {
    const cx<T>* X;       // rows: the neighbouring block's unknowns (leading dimension = its size)
    int side;
  };
  auto mksprow = [&](cx<T>* C, int k, int n, std::initializer_list<SProd> prods, T alpha, T beta) {
    SpRowDesc<T> d{};
    d.side[0] = d.side[1] = -1;
    int p = 0;
    for (const SProd& q : prods) {
      if (!q.X) continue;
      const int nbk = q.side == 0 ? k - 1 : k + 1;
      d.X[p] = q.X; d.ldx[p] = nk(nbk); d.xcol0[p] = s->boff[nbk]; d.side[p] = q.side; d.xrows[p] = (int)nk(nbk);
      const double fl = 8.0 * (double)s->h_side_nnz[2 * k + q.side] * n;
      flops += fl;
      F->flops_sparse += fl;
      ++p;
    }
    d.C = C; d.ldc = nk(k); d.row0 = s->boff[k]; d.m = (int)nk(k); d.n = n; d.alpha = alpha; d.beta = beta;
    return d;
  };
  // Couplings of level l in the lean form: the first level on dense
  // couplings reads them where the scatter (or level 0's sparse products)
  // left them, in L, U; every later level in the plan's E, F slots of that
  // level, which have the couplings' shapes and are otherwise unused there.
  const bool lean = F->lean;
  const int l0 = sp0 ? 1 : 0;
  bool slot_missing = false;
  auto slot = [&](cx<T>* blk, int64_t off, int l) -> cx<T>* {
    cx<T>* p = l == l0 ? blk : (off >= 0 ? base + off : nullptr);
    if (!p) slot_missing = true;
    return p;
  };
  auto Lp = [&](int l, int t) { return slot(F->L[ids[l][t]], P.offE[l][t], l); };
  auto Up = [&](int l, int t) { return slot(F->U[ids[l][t]], P.offF[l][t], l); };
  for (int l = 0; l < nlev; ++l) {
    const auto& id = ids[l];
    const int T_ = (int)id.size();
    Level<T>& L = F->lev[l];
    // inversions of odd positions
    rinv[l].off = vinv.size();
    for (int t = 1; t < T_; t += 2) {
      const int k = id[t];
      vinv.push_back(InvDesc<T>{F->D[k], nk(k), (int)nk(k), ipiv + (size_t)k * nmax, F->info + k, F->G[k]});
      flops += 8.0 * nk(k) * nk(k) * nk(k);
      F->flops_inv += 8.0 * nk(k) * nk(k) * nk(k);
      L.maxn_inv = std::max(L.maxn_inv, (int)nk(k));
      L.minn_inv = std::min(L.minn_inv, (int)nk(k));
    }
    rinv[l].n = vinv.size() - rinv[l].off;
    if (l == 0 && sp0) {
      // level 0 through the couplings of the matrix (id[t] = t):
      //   eliminated t:  G_t = Dinv_t L_t,  H_t = Dinv_t U_t              (dense x sparse)
      //   kept t:        D_t -= L_t H_{t-1} + U_t G_{t+1},
      //                  L_t = -L_t G_{t-1},  U_t = -U_t H_{t+1}          (sparse x dense)
      L.sc_off = vsc.size();
      for (int t = 1; t < T_; t += 2) {
        vsc.push_back(mkspcol(F->G[t], F->D[t], t, 0, T(1), T(0)));
        L.maxn_sc = std::max(L.maxn_sc, (int)nk(t - 1));
        if (t + 1 <= T_ - 1) {
          vsc.push_back(mkspcol(F->H[t], F->D[t], t, 1, T(1), T(0)));
          L.maxn_sc = std::max(L.maxn_sc, (int)nk(t + 1));
        }
      }
      L.nsc = (int)(vsc.size() - L.sc_off);
      L.sr_off = vsr.size();
      for (int t = 0; t < T_; t += 2) {
        const bool hasE = t - 1 >= 0, hasF = t + 1 <= T_ - 1;
        if (hasE || hasF) {
          vsr.push_back(mksprow(F->D[t], t, (int)nk(t),
                                {SProd{hasE ? F->H[t - 1] : nullptr, 0}, SProd{hasF ? F->G[t + 1] : nullptr, 1}},
                                T(-1), T(1)));
          L.maxn_sr = std::max(L.maxn_sr, (int)nk(t));
        }
        if (t - 2 >= 0) {
          vsr.push_back(mksprow(F->L[t], t, (int)nk(t - 2), {SProd{F->G[t - 1], 0}}, T(-1), T(0)));
          L.maxn_sr = std::max(L.maxn_sr, (int)nk(t - 2));
        }
        if (t + 2 <= T_ - 1) {
          vsr.push_back(mksprow(F->U[t], t, (int)nk(t + 2), {SProd{F->H[t + 1], 1}}, T(-1), T(0)));
          L.maxn_sr = std::max(L.maxn_sr, (int)nk(t + 2));
        }
        L.maxm_sr = std::max(L.maxm_sr, (int)nk(t));
      }
      L.nsr = (int)(vsr.size() - L.sr_off);
      rg1[l] = Rng{vg.size(), 0};
      rg2[l] = Rng{vg.size(), 0};
      rfw[l] = Rng{vs.size(), 0};
      rbw[l] = Rng{vs.size(), 0};
      continue;
    }
    if (lean) {
      // Lean form (no E, F): with G_t = Dinv_t L_t, H_t = Dinv_t U_t of the
      // eliminated blocks,
      //   kept t:  D_t -= L_t H_{t-1} + U_t G_{t+1},
      //            L_t' = -L_t G_{t-1},  U_t' = -U_t H_{t+1}   (into level l+1's slots)
      // and a sweep's forward level runs y_odd = Dinv_odd b_odd, then
      // b_even -= L y_{t-1} + U y_{t+1}; backward x_odd = y_odd - G x - H x.
      rg1[l].off = vg.size();
      for (int t = 1; t < T_; t += 2) {
        const int k = id[t];
        const int m = (int)nk(k);
        const int kp = id[t - 1];
        vg.push_back(mkgemm(F->G[k], m, (int)nk(kp), F->D[k], Lp(l, t), m, m, m, nullptr, nullptr, 0, 0, 0, T(1),
                            T(0)));
        L.t1 = std::max(L.t1, tiles(m, (int)nk(kp)));
        if (t + 1 <= T_ - 1) {
          const int kn = id[t + 1];
          vg.push_back(mkgemm(F->H[k], m, (int)nk(kn), F->D[k], Up(l, t), m, m, m, nullptr, nullptr, 0, 0, 0, T(1),
                              T(0)));
          L.t1 = std::max(L.t1, tiles(m, (int)nk(kn)));
        }
      }
      if (l == nlev - 1) {
        // the two-block tail's F0 = U_0 Dinv_1, kept in H of the root block
        // (the root block is never eliminated, so its H is free)
        const int k0 = id[0], k1 = id[1];
        const int n0 = (int)nk(k0), n1 = (int)nk(k1);
        vg.push_back(mkgemm(F->H[k0], n0, n1, Up(l, 0), F->D[k1], n1, n0, n1, nullptr, nullptr, 0, 0, 0, T(1), T(0)));
        L.t1 = std::max(L.t1, tiles(n0, n1));
      }
      rg1[l].n = vg.size() - rg1[l].off;
      rg2[l].off = vg.size();
      for (int t = 0; t < T_; t += 2) {
        const int k = id[t];
        const int m = (int)nk(k);
        const bool hasE = t - 1 >= 0, hasF = t + 1 <= T_ - 1;
        if (hasE || hasF) {
          const int kp = hasE ? id[t - 1] : -1, kn = hasF ? id[t + 1] : -1;
          vg.push_back(mkgemm(F->D[k], m, m, hasE ? Lp(l, t) : nullptr, hasE ? F->H[kp] : nullptr,
                              hasE ? (int)nk(kp) : 0, m, hasE ? (int)nk(kp) : 0, hasF ? Up(l, t) : nullptr,
                              hasF ? F->G[kn] : nullptr, hasF ? (int)nk(kn) : 0, m, hasF ? (int)nk(kn) : 0, T(-1),
                              T(1)));
          L.t2 = std::max(L.t2, tiles(m, m));
        }
        if (t - 2 >= 0) {
          const int kp = id[t - 1], kpp = id[t - 2];
          vg.push_back(mkgemm(Lp(l + 1, t / 2), m, (int)nk(kpp), Lp(l, t), F->G[kp], (int)nk(kp), m, (int)nk(kp),
                              nullptr, nullptr, 0, 0, 0, T(-1), T(0)));
          L.t2 = std::max(L.t2, tiles(m, (int)nk(kpp)));
        }
        if (t + 2 <= T_ - 1) {
          const int kn = id[t + 1], knn = id[t + 2];
          vg.push_back(mkgemm(Up(l + 1, t / 2), m, (int)nk(knn), Up(l, t), F->H[kn], (int)nk(kn), m, (int)nk(kn),
                              nullptr, nullptr, 0, 0, 0, T(-1), T(0)));
          L.t2 = std::max(L.t2, tiles(m, (int)nk(knn)));
        }
      }
      rg2[l].n = vg.size() - rg2[l].off;
      rfy[l].off = vs.size();
      for (int t = 1; t < T_; t += 2) {
        const int k = id[t];
        const int m = (int)nk(k);
        vs.push_back(mkskin(vecp(F->Xv, k), m, {Prod{F->D[k], m, vecp(F->Bv, k), m, T(1)}}, T(0)));
        L.mfy = std::max(L.mfy, m);
      }
      rfy[l].n = vs.size() - rfy[l].off;
      rfw[l].off = vs.size();
      for (int t = 0; t < T_; t += 2) {
        const int k = id[t];
        const int m = (int)nk(k);
        const bool hasE = t - 1 >= 0, hasF = t + 1 <= T_ - 1;
        if (!hasE && !hasF) continue;
        vs.push_back(mkskin(vecp(F->Bv, k), m,
                            {Prod{hasE ? Lp(l, t) : nullptr, m, hasE ? vecp(F->Xv, id[t - 1]) : nullptr,
                                  hasE ? (int)nk(id[t - 1]) : 0, T(-1)},
                             Prod{hasF ? Up(l, t) : nullptr, m, hasF ? vecp(F->Xv, id[t + 1]) : nullptr,
                                  hasF ? (int)nk(id[t + 1]) : 0, T(-1)}},
                            T(1)));
        L.mfw = std::max(L.mfw, m);
      }
      rfw[l].n = vs.size() - rfw[l].off;
      rbw[l].off = vs.size();
      for (int t = 1; t < T_; t += 2) {
        const int k = id[t];
        const int m = (int)nk(k);
        const bool hasR = t + 1 <= T_ - 1;
        vs.push_back(mkskin(vecp(F->Xv, k), m,
                            {Prod{F->G[k], m, vecp(F->Xv, id[t - 1]), (int)nk(id[t - 1]), T(-1)},
                             Prod{hasR ? F->H[k] : nullptr, m, hasR ? vecp(F->Xv, id[t + 1]) : nullptr,
                                  hasR ? (int)nk(id[t + 1]) : 0, T(-1)}},
                            T(1)));
        L.mbw = std::max(L.mbw, m);
      }
      rbw[l].n = vs.size() - rbw[l].off;
      continue;
    }
    // g1: E_t = L_t Dinv_{t-1}, F_t = U_t Dinv_{t+1}
    rg1[l].off = vg.size();
    for (int t = 0; t < T_; t += 2) {
      const int k = id[t];
      const int m = (int)nk(k);
      if (t - 1 >= 0) {
        const int kp = id[t - 1];
        vg.push_back(mkgemm(base + P.offE[l][t], m, (int)nk(kp), F->L[k], F->D[kp], (int)nk(kp), m, (int)nk(kp),
                            nullptr, nullptr, 0, 0, 0, T(1), T(0)));
        L.t1 = std::max(L.t1, tiles(m, (int)nk(kp)));
      }
      if (t + 1 <= T_ - 1) {
        const int kn = id[t + 1];
        vg.push_back(mkgemm(base + P.offF[l][t], m, (int)nk(kn), F->U[k], F->D[kn], (int)nk(kn), m, (int)nk(kn),
                            nullptr, nullptr, 0, 0, 0, T(1), T(0)));
        L.t1 = std::max(L.t1, tiles(m, (int)nk(kn)));
      }
    }
    // G_t = Dinv_t L_t, H_t = Dinv_t U_t for the eliminated blocks (fused backward solve)
    for (int t = 1; t < T_; t += 2) {
      const int k = id[t];
      const int m = (int)nk(k);
      const int kp = id[t - 1];
      vg.push_back(mkgemm(F->G[k], m, (int)nk(kp), F->D[k], F->L[k], m, m, m, nullptr, nullptr, 0, 0, 0, T(1),
                          T(0)));
      L.t1 = std::max(L.t1, tiles(m, (int)nk(kp)));
      if (t + 1 <= T_ - 1) {
        const int kn = id[t + 1];
        vg.push_back(mkgemm(F->H[k], m, (int)nk(kn), F->D[k], F->U[k], m, m, m, nullptr, nullptr, 0, 0, 0, T(1),
                            T(0)));
        L.t1 = std::max(L.t1, tiles(m, (int)nk(kn)));
      }
    }
    rg1[l].n = vg.size() - rg1[l].off;
    // g2: D_t -= E U_{t-1} + F L_{t+1};  L_t = -E L_{t-1};  U_t = -F U_{t+1}
    rg2[l].off = vg.size();
    for (int t = 0; t < T_; t += 2) {
      const int k = id[t];
      const int m = (int)nk(k);
      const bool hasE = t - 1 >= 0, hasF = t + 1 <= T_ - 1;
      if (hasE || hasF) {
        const int kp = hasE ? id[t - 1] : -1, kn = hasF ? id[t + 1] : -1;
        vg.push_back(mkgemm(F->D[k], m, m, hasE ? base + P.offE[l][t] : nullptr, hasE ? F->U[kp] : nullptr,
                            hasE ? (int)nk(kp) : 0, m, hasE ? (int)nk(kp) : 0, hasF ? base + P.offF[l][t] : nullptr,
                            hasF ? F->L[kn] : nullptr, hasF ? (int)nk(kn) : 0, m, hasF ? (int)nk(kn) : 0, T(-1),
                            T(1)));
        L.t2 = std::max(L.t2, tiles(m, m));
      }
      if (t - 2 >= 0) {
        const int kp = id[t - 1], kpp = id[t - 2];
        vg.push_back(mkgemm(F->L[k], m, (int)nk(kpp), base + P.offE[l][t], F->L[kp], (int)nk(kp), m, (int)nk(kp),
                            nullptr, nullptr, 0, 0, 0, T(-1), T(0)));
        L.t2 = std::max(L.t2, tiles(m, (int)nk(kpp)));
      }
      if (t + 2 <= T_ - 1) {
        const int kn = id[t + 1], knn = id[t + 2];
        vg.push_back(mkgemm(F->U[k], m, (int)nk(knn), base + P.offF[l][t], F->U[kn], (int)nk(kn), m, (int)nk(kn),
                            nullptr, nullptr, 0, 0, 0, T(-1), T(0)));
        L.t2 = std::max(L.t2, tiles(m, (int)nk(knn)));
      }
    }
    rg2[l].n = vg.size() - rg2[l].off;
    // solve: forward  B_t -= E_t B_{t-1} + F_t B_{t+1}   (t even)
    rfw[l].off = vs.size();
    for (int t = 0; t < T_; t += 2) {
      const int k = id[t];
      const int m = (int)nk(k);
      const bool hasE = t - 1 >= 0, hasF = t + 1 <= T_ - 1;
      if (!hasE && !hasF) continue;
      vs.push_back(mkskin(vecp(F->Bv, k), m,
                          {Prod{hasE ? base + P.offE[l][t] : nullptr, m, hasE ? vecp(F->Bv, id[t - 1]) : nullptr,
                                hasE ? (int)nk(id[t - 1]) : 0, T(-1)},
                           Prod{hasF ? base + P.offF[l][t] : nullptr, m, hasF ? vecp(F->Bv, id[t + 1]) : nullptr,
                                hasF ? (int)nk(id[t + 1]) : 0, T(-1)}},
                          T(1)));
      L.mfw = std::max(L.mfw, m);
    }
    rfw[l].n = vs.size() - rfw[l].off;
    // backward: X_t = Dinv_t B_t - G_t X_{t-1} - H_t X_{t+1}   (t odd)
    rbw[l].off = vs.size();
    for (int t = 1; t < T_; t += 2) {
      const int k = id[t];
      const int m = (int)nk(k);
      const bool hasR = t + 1 <= T_ - 1;
      vs.push_back(mkskin(vecp(F->Xv, k), m,
                          {Prod{F->D[k], m, vecp(F->Bv, k), m, T(1)},
                           Prod{F->G[k], m, vecp(F->Xv, id[t - 1]), (int)nk(id[t - 1]), T(-1)},
                           Prod{hasR ? F->H[k] : nullptr, m, hasR ? vecp(F->Xv, id[t + 1]) : nullptr,
                                hasR ? (int)nk(id[t + 1]) : 0, T(-1)}},
                          T(0)));
      L.mbw = std::max(L.mbw, m);
    }
    rbw[l].n = vs.size() - rbw[l].off;
  }
  // root
  const int kr = ids[nlev][0];
  rinv[nlev].off = vinv.size();
  vinv.push_back(InvDesc<T>{F->D[kr], nk(kr), (int)nk(kr), ipiv + (size_t)kr * nmax, F->info + kr, F->G[kr]});
  rinv[nlev].n = 1;
  flops += 8.0 * nk(kr) * nk(kr) * nk(kr);
  F->flops_inv += 8.0 * nk(kr) * nk(kr) * nk(kr);
  const size_t rroot = vs.size();
  vs.push_back(mkskin(vecp(F->Xv, kr), (int)nk(kr), {Prod{F->D[kr], (int)nk(kr), vecp(F->Bv, kr), (int)nk(kr), T(1)}},
                      T(0)));
  F->root_n = (int)nk(kr);
  // two-block tail (see Factor): replaces the forward and backward launches
  // of the last level and the root launch by one; the sweep reads the same
  // bytes (n0 n1 + n1^2 + n1 n0 + n0^2 either way)
  size_t rtail1 = 0, rtail2 = 0, rtails = 0;
  if (nlev >= 1) {
    const auto& id = ids[nlev - 1];
    const int k0 = id[0], k1 = id[1];
    const int n0 = (int)nk(k0), n1 = (int)nk(k1);
    cx<T>* F0 = lean ? F->H[k0] : base + P.offF[nlev - 1][0];
    rtail1 = vg.size();
    vg.push_back(mkgemm(F->G[k0], n0, n1, F->D[k0], F0, n0, n0, n0, nullptr, nullptr, 0, 0, 0, T(1), T(0)));
    vg.push_back(mkgemm(F->H[k1], n1, n0, F->G[k1], F->D[k0], n0, n1, n0, nullptr, nullptr, 0, 0, 0, T(1), T(0)));
    rtail2 = vg.size();
    vg.push_back(mkgemm(F->D[k1], n1, n1, F->G[k1], F->G[k0], n0, n1, n0, nullptr, nullptr, 0, 0, 0, T(1), T(1)));
    F->tail_tiles1 = std::max(tiles(n0, n1), tiles(n1, n0));
    F->tail_tiles2 = tiles(n1, n1);
    F->tail_m = std::max(n0, n1);
    const double sb = F->solve_bytes;      // the replaced launches already counted these bytes
    rtails = vs.size();
    vs.push_back(mkskin(vecp(F->Xv, k0), n0, {Prod{F->D[k0], n0, vecp(F->Bv, k0), n0, T(1)},
                                              Prod{F->G[k0], n0, vecp(F->Bv, k1), n1, T(-1)}}, T(0)));
    vs.push_back(mkskin(vecp(F->Xv, k1), n1, {Prod{F->D[k1], n1, vecp(F->Bv, k1), n1, T(1)},
                                              Prod{F->H[k1], n1, vecp(F->Bv, k0), n0, T(-1)}}, T(0)));
    F->solve_bytes = sb;
    F->tail = true;
  }
  if (slot_missing) return fail(KORE_ERR_ARG, "internal: lean factor plan misses a coupling slot");
  F->flops = flops;
  F->solve_launches = F->tail ? 2 * nlev - 1 : 2 * nlev + 1;
  if (lean) F->solve_launches += std::max(0, (F->tail ? nlev - 1 : nlev) - l0);   // the y launches
  size_t rlev0 = 0;
  if (sp0) {
    // level 0 of a sweep: y_odd = Dinv_odd b_odd (forward; x_odd backward),
    // and the couplings from the compact off-block rows
    const auto& id = ids[0];
    const double sb = F->solve_bytes;
    rlev0 = vs.size();
    for (int t = 1; t < (int)id.size(); t += 2) {
      const int k = id[t];
      const int m = (int)nk(k);
      vs.push_back(mkskin(vecp(F->Xv, k), m, {Prod{F->D[k], m, vecp(F->Bv, k), m, T(1)}}, T(0)));
      F->lev0_m = std::max(F->lev0_m, m);
    }
    F->lev0_n = (int)(vs.size() - rlev0);
    const double dinv_bytes = F->solve_bytes - sb;
    // couplings: each off-block entry's value, column and the vector entry it
    // multiplies; per row its pointer, block id and y (read and written)
    const double sparse_bytes = (20.0 + sizeof(cx<T>)) * (double)s->nnz_off +
                                (double)s->n * (8.0 + 2.0 * sizeof(cx<T>));
    F->solve_bytes = sb + 2 * dinv_bytes + sparse_bytes;
    F->solve_launches += 2;
    F->lev0_sparse = true;
  }

  // upload descriptors
  const size_t bi = vinv.size() * sizeof(InvDesc<T>);
  const size_t bg = vg.size() * sizeof(GemmDesc<T>);
  const size_t bsk = vs.size() * sizeof(SkinnyDesc<T>);
  const size_t bsc = vsc.size() * sizeof(SpColDesc<T>);
  const size_t bsr = vsr.size() * sizeof(SpRowDesc<T>);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  TRY(F->desc.alloc(c, al(bi) + al(bg) + al(bsk) + al(bsc) + al(bsr) + 256));
  char* db = F->desc.template as<char>();
  InvDesc<T>* dinv = reinterpret_cast<InvDesc<T>*>(db);
  GemmDesc<T>* dg = reinterpret_cast<GemmDesc<T>*>(db + al(bi));
  SkinnyDesc<T>* dsk = reinterpret_cast<SkinnyDesc<T>*>(db + al(bi) + al(bg));
  SpColDesc<T>* dsc = reinterpret_cast<SpColDesc<T>*>(db + al(bi) + al(bg) + al(bsk));
  SpRowDesc<T>* dsr = reinterpret_cast<SpRowDesc<T>*>(db + al(bi) + al(bg) + al(bsk) + al(bsc));
  HIP_TRY(hipMemcpyAsync(dinv, vinv.data(), bi, hipMemcpyHostToDevice, c->stream));
  if (bg) HIP_TRY(hipMemcpyAsync(dg, vg.data(), bg, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dsk, vs.data(), bsk, hipMemcpyHostToDevice, c->stream));
  if (bsc) HIP_TRY(hipMemcpyAsync(dsc, vsc.data(), bsc, hipMemcpyHostToDevice, c->stream));
  if (bsr) HIP_TRY(hipMemcpyAsync(dsr, vsr.data(), bsr, hipMemcpyHostToDevice, c->stream));
  // the level-0 scatter fills D, and L, U only when level 0 runs on dense couplings
  std::vector<cx<T>*> hp(3 * (size_t)nb, nullptr);
  for (int k = 0; k < nb; ++k) {
    hp[k] = F->D[k];
    hp[nb + k] = sp0 ? nullptr : F->L[k];
    hp[2 * nb + k] = sp0 ? nullptr : F->U[k];
  }
  HIP_TRY(hipMemcpyAsync(F->dD, hp.data(), hp.size() * sizeof(void*), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int l = 0; l < nlev; ++l) {
    Level<T>& L = F->lev[l];
    L.inv = dinv + rinv[l].off; L.ninv = (int)rinv[l].n;
    L.g1 = dg + rg1[l].off; L.ng1 = (int)rg1[l].n;
    L.g2 = dg + rg2[l].off; L.ng2 = (int)rg2[l].n;
    L.fw = dsk + rfw[l].off; L.nfw = (int)rfw[l].n;
    L.fy = dsk + rfy[l].off; L.nfy = (int)rfy[l].n;
    L.bw = dsk + rbw[l].off; L.nbw = (int)rbw[l].n;
    L.sc = dsc + L.sc_off;
    L.sr = dsr + L.sr_off;
  }
  F->root_inv = dinv + rinv[nlev].off;
  F->root_solve = dsk + rroot;
  if (F->lev0_sparse) F->lev0_dinv = dsk + rlev0;
  if (F->tail) {
    F->tail1 = dg + rtail1;
    F->tail2 = dg + rtail2;
    F->tail_solve = dsk + rtails;
  }
  F->ndlu = sp0 ? P.cntD : P.cntDLU;
  return KORE_OK;
}

template <class T> static int64_t factor_count_bytes(const std::vector<int>& bsz, int64_t n) {
  const int nb = (int)bsz.size();
  int nmax = 0;
  for (int b : bsz) nmax = std::max(nmax, b);
  // degree blocks couple through the matrix's off-diagonal blocks; the
  // larger of the two forms' plans
  FactorPlan P, Q;
  factor_plan(bsz, P, true, false);
  factor_plan(bsz, Q, true, true);
  return std::max(P.cnt, Q.cnt) * (int64_t)sizeof(cx<T>) + 2 * n * MAXR * (int64_t)sizeof(cx<T>) +
         (int64_t)nb * (nmax + 1) * 4 + 3 * nb * 8 + 4096;
}

template <class T> static int32_t factor_run(kore_sys* s, Factor<T>* F) {
  kore_ctx* c = s->ctx;
  hipStream_t st = c->stream;
  {
    Timed tm(c, KORE_TIMER_OTHER);
    HIP_TRY(launch_zero<T>(F->mem.template as<cx<T>>(), F->ndlu, st));
    HIP_TRY(launch_build_blocks<T>(s->csr, s->sre, s->sim, s->d_scale.as<double>(), s->d_blk.as<int32_t>(),
                                   s->d_boff.as<int64_t>(), F->dD, F->dL, F->dU, st));
  }
  for (int l = 0; l < F->nlev; ++l) {
    Level<T>& L = F->lev[l];
    {
      Timed tm(c, KORE_TIMER_INV);
      HIP_TRY(launch_gj_inverse<T>(L.inv, L.ninv, L.minn_inv, L.maxn_inv, st));
    }
    if (l == 0 && F->lev0_sparse) {
      Timed tm(c, KORE_TIMER_SPARSE);
      HIP_TRY(launch_spcol<T>(s->offc, s->off.val, L.sc, L.nsc, L.maxn_sc, st));
      if (F->tune_ev[0]) HIP_TRY(hipEventRecord(F->tune_ev[0], st));
      HIP_TRY(launch_sprow<T>(s->off, L.sr, L.nsr, L.maxm_sr, L.maxn_sr, F->sprow_nc, st));
      if (F->tune_ev[1]) HIP_TRY(hipEventRecord(F->tune_ev[1], st));
      continue;
    }
    {
      Timed tm(c, KORE_TIMER_GEMM);
      HIP_TRY(launch_gemm_batched<T>(L.g1, L.ng1, L.t1, st));
    }
    {
      Timed tm(c, KORE_TIMER_GEMM);
      HIP_TRY(launch_gemm_batched<T>(L.g2, L.ng2, L.t2, st));
    }
  }
  {
    Timed tm(c, KORE_TIMER_INV);
    HIP_TRY(launch_gj_inverse<T>(F->root_inv, 1, F->root_n, F->root_n, st));
  }
  if (F->tail) {
    {
      Timed tm(c, KORE_TIMER_GEMM);
      HIP_TRY(launch_gemm_batched<T>(F->tail1, 2, F->tail_tiles1, st));
    }
    Timed tm(c, KORE_TIMER_GEMM);
    HIP_TRY(launch_gemm_batched<T>(F->tail2, 1, F->tail_tiles2, st));
  }
  return KORE_OK;
}

template <class T> static int32_t bcr_sweep_launches(kore_sys* s, Factor<T>* F, int nrhs);

// One solve sweep of the factor on F->Bv -> F->Xv.  Without per-kernel
// timing, the sweep for a column count is captured into a HIP graph the
// first time and replayed afterwards (2 levels + 1 launches become one graph
// launch); the sweep records no memset and no event, so the graph holds
// kernel nodes only.
template <class T> static int32_t bcr_solve(kore_sys* s, Factor<T>* F, int nrhs) {
  kore_ctx* c = s->ctx;
  if (!c->graphs || c->timing || nrhs < 1 || nrhs > MAXR) return bcr_sweep_launches<T>(s, F, nrhs);
  hipStream_t st = c->stream;
  if (!F->sweep_graph[nrhs]) {
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int32_t rc = bcr_sweep_launches<T>(s, F, nrhs);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &g);
    if (rc < 0) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess || !g) return fail(KORE_ERR_HIP, "sweep capture failed: %s", hipGetErrorString(ec));
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(KORE_ERR_HIP, "sweep graph instantiation failed: %s", hipGetErrorString(ei));
    F->sweep_graph[nrhs] = ex;
  }
  HIP_TRY(hipGraphLaunch(F->sweep_graph[nrhs], st));
  return KORE_OK;
}

template <class T> static int32_t bcr_sweep_launches(kore_sys* s, Factor<T>* F, int nrhs) {
  kore_ctx* c = s->ctx;
  hipStream_t st = c->stream;
  const int ltop = F->tail ? F->nlev - 1 : F->nlev;   // levels swept one by one
  const bool sp0 = F->lev0_sparse;                     // level 0 through the sparse couplings
  const int64_t n = s->n;
  // one column, blocks up to 256 wide: level 0's inverse blocks (one product
  // each) take the kernel with registers for one product (a two-product
  // kernel for the forward levels measured no faster)
  const bool fewp = std::is_same<T, double>::value && nrhs == 1 && s->nmax <= 256;
  auto inverse_blocks = [&](const SkinnyDesc<T>* d, int nd, int m) -> int32_t {
    Timed tm(c, KORE_TIMER_SKINNY);
    if constexpr (std::is_same<T, double>::value) {
      if (fewp) {
        HIP_TRY(launch_skinny_np<T>(d, nd, m, 1, st));
        return KORE_OK;
      }
    }
    HIP_TRY(launch_skinny<T>(d, nd, m, nrhs, st));
    return KORE_OK;
  };
  auto lev0_inverse = [&]() { return inverse_blocks(F->lev0_dinv, F->lev0_n, F->lev0_m); };
  for (int l = 0; l < ltop; ++l) {
    Level<T>& L = F->lev[l];
    if (l == 0 && sp0) {
      TRY(lev0_inverse());   // y_odd = Dinv_odd b_odd, in Xv
      Timed tm(c, KORE_TIMER_SKINNY);
      HIP_TRY(launch_offsweep<T>(s->off, s->d_blk.as<int32_t>(), n, 0, F->Xv, n, F->Bv, n, nrhs, st));
      continue;
    }
    if (L.nfy) TRY(inverse_blocks(L.fy, L.nfy, L.mfy));   // lean form: y_odd = Dinv_odd b_odd
    Timed tm(c, KORE_TIMER_SKINNY);
    HIP_TRY(launch_skinny<T>(L.fw, L.nfw, L.mfw, nrhs, st));
  }
  {
    Timed tm(c, KORE_TIMER_SKINNY);
    if (F->tail)
      HIP_TRY(launch_skinny<T>(F->tail_solve, 2, F->tail_m, nrhs, st));
    else
      HIP_TRY(launch_skinny<T>(F->root_solve, 1, F->root_n, nrhs, st));
  }
  for (int l = ltop - 1; l >= 0; --l) {
    Level<T>& L = F->lev[l];
    if (l == 0 && sp0) {
      {
        Timed tm(c, KORE_TIMER_SKINNY);
        HIP_TRY(launch_offsweep<T>(s->off, s->d_blk.as<int32_t>(), n, 1, F->Xv, n, F->Bv, n, nrhs, st));
      }
      TRY(lev0_inverse());   // x_odd = Dinv_odd b_odd
      continue;
    }
    Timed tm(c, KORE_TIMER_SKINNY);
    HIP_TRY(launch_skinny<T>(L.bw, L.nbw, L.mbw, nrhs, st));
  }
  return KORE_OK;
}

// Column 2-norms of one or two blocks (degree-major, length n) -> host, with a
// single synchronisation.  b may be null.
template <class Ta, class Tb>
static int32_t colnorms2(kore_sys* s, const cx<Ta>* a, int64_t lda, const cx<Tb>* b, int64_t ldb, int nrhs,
                         double* outa, double* outb) {
  kore_ctx* c = s->ctx;
  HIP_TRY(launch_colnorm2<Ta>(a, lda, s->n, nrhs, s->part, s->npart, c->stream));
  if (b) HIP_TRY(launch_colnorm2<Tb>(b, ldb, s->n, nrhs, s->part + (size_t)s->npart * MAXR, s->npart, c->stream));
  const int nh = b ? 2 : 1;
  std::vector<double> h((size_t)nh * s->npart * MAXR);
  HIP_TRY(hipMemcpyAsync(h.data(), s->part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int j = 0; j < nrhs; ++j) {
    double sa = 0, sb = 0;
    for (int q = 0; q < s->npart; ++q) {
      sa += h[(size_t)q * nrhs + j];
      if (b) sb += h[(size_t)s->npart * MAXR + (size_t)q * nrhs + j];
    }
    outa[j] = std::sqrt(sa);
    if (b) outb[j] = std::sqrt(sb);
  }
  return KORE_OK;
}

template <class T>
static int32_t refine_sweep(kore_sys* s, Factor<T>* F, const cx<double>* rhs, int64_t ldr, cx<double>* x,
                            int64_t ldx, int nrhs, bool first) {
  kore_ctx* c = s->ctx;
  hipStream_t st = c->stream;
  const int64_t n = s->n;
  {
    Timed tm(c, KORE_TIMER_OTHER);
    HIP_TRY(launch_gather_scale<double, T>(rhs, ldr, F->Bv, n, n, nrhs, nullptr, s->d_scale.as<double>(), 1.0, st));
  }
  TRY(bcr_solve<T>(s, F, nrhs));
  if (first) {
    Timed tm(c, KORE_TIMER_OTHER);
    HIP_TRY(launch_gather_scale<T, double>(F->Xv, n, x, ldx, n, nrhs, nullptr, nullptr, 1.0, st));
  }
  return KORE_OK;
}

// Solve (A - sigma B) x = b for nrhs <= 16 columns; b, x complex128 on the
// device in degree-major order (leading dimensions ldb, ldx).
//
// Refinement (always with a complex64 factor; ir_maxit extra sweeps with an
// FP64 one): r = b - (A - sigma B) x in FP64 by SpMV, d = factor^{-1} R r,
// x += d; stop when max_c ||d_c|| / ||x_c|| <= ir_tol (Wilkinson's test) or
// when the correction stops shrinking (ratio > 1/2 between sweeps).
static int32_t solve_dev(kore_sys* s, const cx<double>* b, int64_t ldb, cx<double>* x, int64_t ldx, int nrhs,
                         double ir_tol, int ir_maxit, kore_solve_info* info) {
  kore_ctx* c = s->ctx;
  hipStream_t st = c->stream;
  const int64_t n = s->n;
  if (!s->f || !s->factored) return fail(KORE_ERR_STATE, "no factorisation: call kore_factor first");
  if (nrhs < 1 || nrhs > MAXR) return fail(KORE_ERR_ARG, "nrhs must be 1..16 per device solve");
  if (info) {
    info->iterations = 0;
    info->converged = 1;
    info->rel_correction = -1.0;
  }
  const bool c64 = s->f->prec == KORE_PREC_C64;
  const int maxsweep = c64 ? std::max(2, 1 + ir_maxit) : 1 + std::max(0, ir_maxit);
  Factor<float>* F32 = c64 ? static_cast<Factor<float>*>(s->f.get()) : nullptr;
  Factor<double>* F64 = c64 ? nullptr : static_cast<Factor<double>*>(s->f.get());
  if (c64) TRY(refine_sweep<float>(s, F32, b, ldb, x, ldx, nrhs, true));
  else TRY(refine_sweep<double>(s, F64, b, ldb, x, ldx, nrhs, true));
  int it = 1;
  double prev = 1e300;
  std::vector<double> nd(nrhs), nx(nrhs);
  for (it = 2; it <= maxsweep; ++it) {
    {
      Timed tm(c, KORE_TIMER_SPMV);
      HIP_TRY(launch_copy_cols(b, ldb, s->rr, n, n, nrhs, st));
      HIP_TRY(launch_spmv(s->csr, -1.0, 0.0, s->sre, s->sim, x, ldx, s->rr, n, nrhs, 1.0, st));
    }
    if (c64) {
      TRY(refine_sweep<float>(s, F32, s->rr, n, x, ldx, nrhs, false));
      TRY(colnorms2<float, double>(s, F32->Xv, n, x, ldx, nrhs, nd.data(), nx.data()));
      Timed tm(c, KORE_TIMER_OTHER);
      HIP_TRY(launch_axpy<float>(x, ldx, F32->Xv, n, n, nrhs, 1.0, st));
    } else {
      TRY(refine_sweep<double>(s, F64, s->rr, n, x, ldx, nrhs, false));
      TRY(colnorms2<double, double>(s, F64->Xv, n, x, ldx, nrhs, nd.data(), nx.data()));
      Timed tm(c, KORE_TIMER_OTHER);
      HIP_TRY(launch_axpy<double>(x, ldx, F64->Xv, n, n, nrhs, 1.0, st));
    }
    double ratio = 0;
    for (int j = 0; j < nrhs; ++j) ratio = std::max(ratio, nx[j] > 0 ? nd[j] / nx[j] : 0.0);
    if (info) info->rel_correction = ratio;
    if (ratio <= ir_tol) break;
    if (it >= 3 && ratio > 0.5 * prev) {  // stagnated
      if (info) info->converged = 0;
      break;
    }
    prev = ratio;
    if (it == maxsweep && info) info->converged = 0;
  }
  if (info) info->iterations = std::min(it, maxsweep);
  return KORE_OK;
}

// ---------------------------------------------------------------------------
// GMRES-based iterative refinement (one right-hand side)
// ---------------------------------------------------------------------------
// FP64 restarted GMRES on (A - sigma B) x = b, right-preconditioned by the
// factor: every Krylov step is one factor solve (complex64 or FP64) and one
// FP64 sparse product; the basis lives on the device, the (restart+1) x
// restart Hessenberg matrix and its Givens rotations on the host.  Converges
// where plain refinement with a complex64 factor stalls (condition numbers up
// to about 1/eps_fp32), because the factor only has to be a preconditioner.
static constexpr int GM_MAXRESTART = 30, GM_CHUNKS = 1024;

static int32_t prec_apply(kore_sys* s, const cx<double>* v, cx<double>* z) {
  if (s->f->prec == KORE_PREC_C64)
    return refine_sweep<float>(s, static_cast<Factor<float>*>(s->f.get()), v, s->n, z, s->n, 1, true);
  return refine_sweep<double>(s, static_cast<Factor<double>*>(s->f.get()), v, s->n, z, s->n, 1, true);
}

// out (na complex, device) = V[:, 0:na]^H w
static int32_t gm_dot(kore_sys* s, const cx<double>* V, int na, const cx<double>* w, cx<double>* part,
                      cx<double>* out) {
  const int64_t n = s->n;
  const int nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(GM_CHUNKS, (n + 63) / 64));
  const int64_t chunk = (n + nchunk - 1) / nchunk;
  Timed tm(s->ctx, KORE_TIMER_ORTHO);
  HIP_TRY(launch_tsdot(V, n, na, w, n, 1, n, nchunk, chunk, part, s->ctx->stream));
  HIP_TRY(launch_sum_parts(reinterpret_cast<const double*>(part), nchunk, 2 * na, reinterpret_cast<double*>(out),
                           s->ctx->stream));
  return KORE_OK;
}

// y = beta y + alpha V[:, 0:k] c   (c: k complex on the device)
static int32_t gm_comb(kore_sys* s, const cx<double>* V, int k, const cx<double>* c, cx<double>* y, double alpha,
                       double beta) {
  SkinnyDesc<double> d{};
  d.A[0] = V; d.lda[0] = s->n;
  d.X[0] = c; d.ldx[0] = k;
  d.k[0] = k; d.alpha[0] = alpha;
  d.Y = y; d.ldy = s->n;
  d.m = (int)s->n; d.beta = beta;
  Timed tm(s->ctx, KORE_TIMER_ORTHO);
  HIP_TRY(launch_skinny1<double>(d, 1, s->ctx->stream));
  return KORE_OK;
}

// b, x: device vectors (degree-major); x receives the solution.
static int32_t gmres_dev(kore_sys* s, const cx<double>* b, cx<double>* x, double tol, int maxit, int m, double nA,
                         kore_solve_info* info) {
  kore_ctx* c = s->ctx;
  hipStream_t st = c->stream;
  const int64_t n = s->n;
  if (m < 1 || m > GM_MAXRESTART) return fail(KORE_ERR_ARG, "GMRES restart length must be 1..%d", GM_MAXRESTART);
  if (s->gm_cols < m + 1) {
    TRY(s->gm.alloc(c, ((size_t)(GM_MAXRESTART + 1) * n + (size_t)GM_CHUNKS * (GM_MAXRESTART + 1) + 4 * 64) *
                           sizeof(zd)));
    s->gm_cols = GM_MAXRESTART + 1;
  }
  cx<double>* V = s->gm.as<cx<double>>();
  cx<double>* part = V + (size_t)(GM_MAXRESTART + 1) * n;
  cx<double>* h1 = part + (size_t)GM_CHUNKS * (GM_MAXRESTART + 1);
  cx<double>* h2 = h1 + 64;
  cx<double>* yv = h2 + 64;
  cx<double>* z = s->st;       // preconditioned vector
  cx<double>* r = s->rr;       // residual
  double nb = 0;
  TRY(colnorms2<double, double>(s, b, n, nullptr, 0, 1, &nb, nullptr));
  // x0 = M^{-1} b
  TRY(prec_apply(s, b, x));
  int its = 1;
  bool conv = false;
  double berr = 1.0;
  std::vector<zd> H((size_t)(m + 1) * m), g(m + 1), cs(m), sn(m), hh(2 * 64);
  for (;;) {
    // r = b - (A - sigma B) x
    {
      Timed tm(c, KORE_TIMER_SPMV);
      HIP_TRY(launch_copy_cols(b, n, r, n, n, 1, st));
      HIP_TRY(launch_spmv(s->csr, -1.0, 0.0, s->sre, s->sim, x, n, r, n, 1, 1.0, st));
    }
    double beta = 0, nx = 0;
    TRY(colnorms2<double, double>(s, r, n, x, n, 1, &beta, &nx));
    const double thr = tol * (nA * nx + nb);
    berr = (nA * nx + nb) > 0 ? beta / (nA * nx + nb) : 0.0;
    if (beta <= thr) {
      conv = true;
      break;
    }
    if (its >= maxit) break;
    HIP_TRY(launch_gather_scale<double, double>(r, n, V, n, n, 1, nullptr, nullptr, 1.0 / beta, st));
    std::fill(H.begin(), H.end(), zd(0));
    std::fill(g.begin(), g.end(), zd(0));
    g[0] = beta;
    int j = 0;
    for (; j < m && its < maxit; ++j) {
      cx<double>* vj = V + (size_t)j * n;
      cx<double>* w = V + (size_t)(j + 1) * n;
      TRY(prec_apply(s, vj, z));
      ++its;
      {
        Timed tm(c, KORE_TIMER_SPMV);
        HIP_TRY(launch_spmv(s->csr, 1.0, 0.0, -s->sre, -s->sim, z, n, w, n, 1, 0.0, st));
      }
      // classical Gram-Schmidt, twice
      TRY(gm_dot(s, V, j + 1, w, part, h1));
      TRY(gm_comb(s, V, j + 1, h1, w, -1.0, 1.0));
      TRY(gm_dot(s, V, j + 1, w, part, h2));
      TRY(gm_comb(s, V, j + 1, h2, w, -1.0, 1.0));
      HIP_TRY(hipMemcpyAsync(hh.data(), h1, 2 * 64 * sizeof(zd), hipMemcpyDeviceToHost, st));
      double hn = 0;
      TRY(colnorms2<double, double>(s, w, n, nullptr, 0, 1, &hn, nullptr));   // synchronises
      zd* Hj = H.data() + (size_t)j * (m + 1);
      for (int i = 0; i <= j; ++i) Hj[i] = hh[i] + hh[64 + i];
      Hj[j + 1] = hn;
      if (hn > 0) HIP_TRY(launch_gather_scale<double, double>(w, n, w, n, n, 1, nullptr, nullptr, 1.0 / hn, st));
      // previous rotations, then a new one that zeroes H[j+1, j]
      for (int i = 0; i < j; ++i) {
        const zd t = cs[i] * Hj[i] + sn[i] * Hj[i + 1];
        Hj[i + 1] = -std::conj(sn[i]) * Hj[i] + cs[i] * Hj[i + 1];
        Hj[i] = t;
      }
      const double a = std::abs(Hj[j]), bb = std::abs(Hj[j + 1]);
      const double rho = std::hypot(a, bb);
      if (rho == 0) {
        cs[j] = 1;
        sn[j] = 0;
      } else if (a == 0) {
        cs[j] = 0;
        sn[j] = std::conj(Hj[j + 1]) / bb;
      } else {
        cs[j] = a / rho;
        sn[j] = (Hj[j] / a) * std::conj(Hj[j + 1]) / rho;
      }
      Hj[j] = cs[j] * Hj[j] + sn[j] * Hj[j + 1];
      Hj[j + 1] = 0;
      g[j + 1] = -std::conj(sn[j]) * g[j];
      g[j] = cs[j] * g[j];
      if (std::abs(g[j + 1]) <= 0.5 * thr || hn == 0) {
        ++j;
        break;
      }
    }
    // y = R^{-1} g, u = V y (into r), x += M^{-1} u
    const int k = j;
    std::vector<zd> y(k);
    for (int i = k - 1; i >= 0; --i) {
      zd t = g[i];
      for (int q = i + 1; q < k; ++q) t -= H[(size_t)q * (m + 1) + i] * y[q];
      const zd d = H[(size_t)i * (m + 1) + i];
      y[i] = d != zd(0) ? t / d : zd(0);
    }
    if (k > 0) {
      HIP_TRY(hipMemcpyAsync(yv, y.data(), k * sizeof(zd), hipMemcpyHostToDevice, st));
      TRY(gm_comb(s, V, k, yv, r, 1.0, 0.0));
      TRY(prec_apply(s, r, z));
      HIP_TRY(launch_axpy<double>(x, n, z, n, n, 1, 1.0, st));
      HIP_TRY(hipStreamSynchronize(st));   // y leaves scope
    }
  }
  if (info) {
    info->iterations = its;
    info->converged = conv ? 1 : 0;
    info->rel_correction = berr;
  }
  return KORE_OK;
}

// ---------------------------------------------------------------------------
// tuner: tile shapes chosen by measurement
// ---------------------------------------------------------------------------
// The level-0 sparse-row products take one of a few column tiles per
// workgroup: a wider tile reads each row's entries fewer times but needs
// more LDS and registers per workgroup.  Which is fastest depends on the
// block size and the precision, so the first factorisation of each
// (precision, block-size class of 64) runs the whole factorisation once per
// candidate tile with the products timed by HIP events and keeps the
// fastest.  The choice is cached for the process and, with
// KORE_AMD_TUNE_FILE set, read from and appended to that file (lines
// "sprow <precision> <class> <tile>") so that later processes skip the
// measurement.  KORE_AMD_SPROW_NC forces a tile (no measurement).
struct Tuner {
  std::mutex mu;
  std::map<std::pair<int, int>, int> sprow;
  bool loaded = false;
};
static Tuner g_tuner;

static std::vector<int> sprow_tiles(int prec) {
  return std::vector<int>{4, 8, 16};
}

static int tuner_lookup(int prec, int nmax) {
  std::lock_guard<std::mutex> lk(g_tuner.mu);
  if (!g_tuner.loaded) {
    g_tuner.loaded = true;
    if (const char* path = std::getenv("KORE_AMD_TUNE_FILE")) {
      if (FILE* f = fopen(path, "r")) {
        char what[32];
        int pr, cl, nc;
        while (fscanf(f, "%31s %d %d %d", what, &pr, &cl, &nc) == 4)
          if (!strcmp(what, "sprow")) g_tuner.sprow[{pr, cl}] = nc;
        fclose(f);
      }
    }
  }
  auto it = g_tuner.sprow.find({prec, (nmax + 63) / 64});
  return it == g_tuner.sprow.end() ? -1 : it->second;
}

static void tuner_store(int prec, int nmax, int nc) {
  std::lock_guard<std::mutex> lk(g_tuner.mu);
  const int cl = (nmax + 63) / 64;
  g_tuner.sprow[{prec, cl}] = nc;
  if (const char* path = std::getenv("KORE_AMD_TUNE_FILE")) {
    if (FILE* f = fopen(path, "a")) {
      fprintf(f, "sprow %d %d %d\n", prec, cl, nc);
      fclose(f);
    }
  }
}

// Chooses the factor's level-0 product tile: forced, cached, or measured
// (one full factorisation per candidate; the caller's factorisation follows).
template <class T> static int32_t tune_factor(kore_sys* s, Factor<T>* F, int prec) {
  if (!F->lev0_sparse) return KORE_OK;
  const std::vector<int> cands = sprow_tiles(prec);
  if (const char* e = std::getenv("KORE_AMD_SPROW_NC")) {
    const int nc = std::atoi(e);
    if (std::find(cands.begin(), cands.end(), nc) == cands.end())
      return fail(KORE_ERR_ARG, "KORE_AMD_SPROW_NC=%d is not a tile of this precision", nc);
    F->sprow_nc = nc;
    return KORE_OK;
  }
  const int known = tuner_lookup(prec, s->nmax);
  if (known > 0) {
    F->sprow_nc = known;
    return KORE_OK;
  }
  kore_ctx* c = s->ctx;
  HIP_TRY(hipEventCreate(&F->tune_ev[0]));
  HIP_TRY(hipEventCreate(&F->tune_ev[1]));
  // the measured factorisations are not the caller's work: per-class timing
  // is suspended while they run
  const bool timing = c->timing;
  c->timing = false;
  float best = 1e30f;
  int best_nc = cands[0];
  int32_t rc = KORE_OK;
  for (int nc : cands) {
    F->sprow_nc = nc;
    rc = factor_run<T>(s, F);
    if (rc < 0) break;
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
      rc = fail(KORE_ERR_HIP, "factorisation failed while tuning");
      break;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, F->tune_ev[0], F->tune_ev[1]);
    if (getenv("KORE_AMD_TRACE")) fprintf(stderr, "kore_amd: tune sprow prec %d nmax %d tile %d: %.3f ms\n", prec,
                                          s->nmax, nc, ms);
    if (ms < best) {
      best = ms;
      best_nc = nc;
    }
  }
  hipEventDestroy(F->tune_ev[0]);
  hipEventDestroy(F->tune_ev[1]);
  F->tune_ev[0] = F->tune_ev[1] = nullptr;
  c->timing = timing;
  if (rc < 0) return rc;
  F->sprow_nc = best_nc;
  tuner_store(prec, s->nmax, best_nc);
  return KORE_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int32_t kore_version(void) { return 100; }
const char* kore_last_error(void) { return g_err.c_str(); }

int32_t kore_ctx_create(int32_t device, int64_t mem_budget_bytes, kore_ctx** out) {
  if (!out) return fail(KORE_ERR_ARG, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(KORE_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
  HIP_TRY(hipSetDevice(device));
  auto c = std::make_unique<kore_ctx>();
  c->device = device;
  c->budget = mem_budget_bytes;
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  if (const char* g = std::getenv("KORE_AMD_GRAPHS")) c->graphs = std::atoi(g) != 0;   // default off
  char* scr = nullptr;
  TRY(scratch_address(device, &scr));     // once per device, with the device current
  *out = c.release();
  return KORE_OK;
}

void kore_ctx_destroy(kore_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  resolve_timers(ctx);
  for (auto e : ctx->pool) hipEventDestroy(e);
  ctx->pin.release();
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

int32_t kore_ctx_sync(kore_ctx* ctx) {
  if (!ctx) return fail(KORE_ERR_ARG, "ctx is NULL");
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KORE_OK;
}

int64_t kore_ctx_bytes_in_use(const kore_ctx* ctx) { return ctx ? ctx->in_use.load() : 0; }

int32_t kore_mem_trim(int32_t device) {
  if (device < 0) return fail(KORE_ERR_ARG, "bad device %d", device);
  HIP_TRY(hipSetDevice(device));
  pool_trim(g_dev_pool[device % POOL_MAXDEV], false);
  pool_trim(g_host_pool, true);
  return KORE_OK;
}

int32_t kore_mem_stats(int32_t device, int64_t* cached_bytes, int64_t* device_allocs) {
  if (device < 0) return fail(KORE_ERR_ARG, "bad device %d", device);
  MemPool& mp = g_dev_pool[device % POOL_MAXDEV];
  {
    std::lock_guard<std::mutex> lk(mp.mu);
    if (cached_bytes) *cached_bytes = (int64_t)mp.cached;
  }
  if (device_allocs) *device_allocs = g_pool_misses.load();
  return KORE_OK;
}

int32_t kore_ctx_set_timing(kore_ctx* ctx, int32_t enable) {
  if (!ctx) return fail(KORE_ERR_ARG, "ctx is NULL");
  resolve_timers(ctx);
  ctx->timing = enable != 0;
  return KORE_OK;
}

int32_t kore_ctx_set_graphs(kore_ctx* ctx, int32_t enable) {
  if (!ctx) return fail(KORE_ERR_ARG, "ctx is NULL");
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->graphs = enable != 0;
  return KORE_OK;
}

int32_t kore_ctx_timing(kore_ctx* ctx, int32_t cls, int64_t* count, double* total_ms) {
  if (!ctx || cls < 0 || cls >= KORE_TIMER_COUNT) return fail(KORE_ERR_ARG, "bad timing query");
  resolve_timers(ctx);
  if (count) *count = ctx->cnt[cls];
  if (total_ms) *total_ms = ctx->tot_ms[cls];
  return KORE_OK;
}

int32_t kore_ctx_timing_reset(kore_ctx* ctx) {
  if (!ctx) return fail(KORE_ERR_ARG, "ctx is NULL");
  resolve_timers(ctx);
  for (int i = 0; i < KORE_TIMER_COUNT; ++i) {
    ctx->cnt[i] = 0;
    ctx->tot_ms[i] = 0;
  }
  return KORE_OK;
}

int32_t kore_layout_blocks(const kore_layout* lay, int64_t* sizmat, int32_t* nblk, int32_t* blk_sizes,
                           int32_t blk_cap, int64_t* perm) {
  std::vector<int> sizes;
  std::vector<int64_t> p;
  int64_t n = 0;
  TRY(layout_build(lay, &n, &sizes, perm ? &p : nullptr));
  if (sizmat) *sizmat = n;
  if (nblk) *nblk = (int32_t)sizes.size();
  if (blk_sizes) {
    if (blk_cap < (int32_t)sizes.size()) return fail(KORE_ERR_ARG, "blk_sizes capacity too small");
    for (size_t i = 0; i < sizes.size(); ++i) blk_sizes[i] = sizes[i];
  }
  if (perm) memcpy(perm, p.data(), p.size() * sizeof(int64_t));
  return KORE_OK;
}

int64_t kore_factor_bytes(const kore_layout* lay, int32_t precision) {
  std::vector<int> sizes;
  int64_t n = 0;
  if (layout_build(lay, &n, &sizes, nullptr) < 0) return -1;
  return precision == KORE_PREC_C64 ? factor_count_bytes<float>(sizes, n) : factor_count_bytes<double>(sizes, n);
}

}  // extern "C"

// A system prepared on the host (kore_sys_prepare): the layout, the union
// pattern and values staged in pinned memory, the level-0 coupling views;
// kore_sys_create_prepared allocates the device buffers and uploads it.
struct kore_sys_plan {
  std::unique_ptr<kore_sys> s;
  PinnedBuf arena;                        // staged pattern and values, by row range
  int nth = 1;
  std::vector<size_t> ub, ubb, tcnt, tcntb;
  std::vector<int32_t> rowcntb, optr, onlo, ocol, cptr, cq, crow;
  zd *pva = nullptr, *pvb = nullptr, *pvbb = nullptr;
  int32_t *pix = nullptr, *pixb = nullptr;
};

// With a1_data, A's values are a_data + e a1_data (A(E) = A0 + E A1 on one
// pattern), formed as each entry is staged.
static int32_t sys_prepare(const kore_layout* lay, int64_t sizmat, const int32_t* a_indptr, const int32_t* a_indices,
                           const double* a_data, const int32_t* b_indptr, const int32_t* b_indices,
                           const double* b_data, int32_t b_is_complex, std::unique_ptr<kore_sys_plan>& P,
                           const double* a1_data = nullptr, double e = 0.0) {
  if (!lay || !a_indptr || !a_indices || !a_data) return fail(KORE_ERR_ARG, "NULL argument");
  P = std::make_unique<kore_sys_plan>();
  P->s = std::make_unique<kore_sys>();
  auto& s = P->s;
  s->lay = *lay;
  int64_t n = 0;
  StageTrace trace;
  TRY(layout_build(lay, &n, &s->bsz, &s->perm));
  if (n != sizmat) return fail(KORE_ERR_ARG, "matrix size %lld does not match the layout (%lld)", (long long)sizmat, (long long)n);
  s->n = n;
  s->nblk = (int)s->bsz.size();
  s->boff.assign(s->nblk + 1, 0);
  for (int k = 0; k < s->nblk; ++k) s->boff[k + 1] = s->boff[k] + s->bsz[k];
  s->nmax = *std::max_element(s->bsz.begin(), s->bsz.end());
  if (s->nmax > 512) return fail(KORE_ERR_UNSUPPORTED, "degree block of %d unknowns exceeds the 512 supported", s->nmax);
  s->h_blk.resize(n);
  for (int k = 0; k < s->nblk; ++k)
    for (int64_t i = s->boff[k]; i < s->boff[k + 1]; ++i) s->h_blk[i] = k;
  std::vector<int64_t> inv(n);
  for (int64_t i = 0; i < n; ++i) inv[s->perm[i]] = i;
  // validate CSR
  auto check_csr = [&](const int32_t* ip, const int32_t* ix, const char* nm) -> int32_t {
    if (ip[0] != 0) return fail(KORE_ERR_ARG, "%s: indptr[0] != 0", nm);
    for (int64_t r = 0; r < n; ++r) {
      if (ip[r + 1] < ip[r]) return fail(KORE_ERR_ARG, "%s: indptr not monotone at row %lld", nm, (long long)r);
    }
    const int64_t nnz = ip[n];
    for (int64_t k = 0; k < nnz; ++k)
      if (ix[k] < 0 || ix[k] >= n) return fail(KORE_ERR_ARG, "%s: column index out of range", nm);
    return KORE_OK;
  };
  trace.mark("sys_create layout");
  TRY(check_csr(a_indptr, a_indices, "A"));
  s->has_b = b_indptr != nullptr;
  if (s->has_b) {
    if (!b_indices || !b_data) return fail(KORE_ERR_ARG, "B arrays incomplete");
    TRY(check_csr(b_indptr, b_indices, "B"));
  }
  // union pattern in degree-major order, built by row ranges on host threads
  // (each range into its own arrays, then concatenated in row order)
  struct Ent {
    int32_t col;
    int32_t src;   // k >= 0: A entry k; k < 0: B entry -(k + 1)
  };
  const int64_t nnz_a = a_indptr[n];
  s->h_amap.assign((size_t)nnz_a, -1);
  const int nth = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)host_threads(), std::max<int64_t>(1, n / 4096)));
  // per-range upper bounds (entries of A plus B on the range's rows) place
  // each range's output in the pinned arena before anything is built
  std::vector<size_t>& ub = P->ub;
  std::vector<size_t>& ubb = P->ubb;
  ub.assign(nth + 1, 0);
  ubb.assign(nth + 1, 0);
  for (int t = 0; t < nth; ++t) {
    const int64_t r0 = n * t / nth, r1 = n * (t + 1) / nth;
    size_t u = 0, ubv = 0;
    for (int64_t i = r0; i < r1; ++i) {
      const int64_t r = s->perm[i];
      u += (size_t)(a_indptr[r + 1] - a_indptr[r]);
      if (s->has_b) ubv += (size_t)(b_indptr[r + 1] - b_indptr[r]);
    }
    ub[t + 1] = ub[t] + u + ubv;
    ubb[t + 1] = ubb[t] + ubv;
  }
  const size_t cap = std::max<size_t>(ub[nth], 1), capb = std::max<size_t>(ubb[nth], 1);
  TRY(P->arena.reserve(cap * (sizeof(int32_t) + 2 * sizeof(zd)) + capb * (sizeof(int32_t) + sizeof(zd)) + 512,
                       nullptr));
  char* arena = P->arena.p;
  zd* pva = reinterpret_cast<zd*>(arena);
  zd* pvb = pva + cap;
  zd* pvbb = pvb + cap;                     // B alone
  int32_t* pix = reinterpret_cast<int32_t*>(pvbb + capb);
  int32_t* pixb = pix + cap;
  P->pva = pva;
  P->pvb = pvb;
  P->pvbb = pvbb;
  P->pix = pix;
  P->pixb = pixb;
  P->nth = nth;
  std::vector<size_t>& tcnt = P->tcnt;
  std::vector<size_t>& tcntb = P->tcntb;
  tcnt.assign(nth, 0);
  tcntb.assign(nth, 0);
  std::vector<std::vector<int32_t>> tocol(nth);   // off-block columns of each range, in row order
  std::vector<int32_t> rowcnt(n, 0), rowoff(n, 0), rowlo(n, 0);
  std::vector<int32_t>& rowcntb = P->rowcntb;
  rowcntb.assign(n, 0);
  std::vector<int64_t> bad_row(nth, -1);
  std::vector<int> bad_kb(nth, 0), bad_kc(nth, 0);
  // A row (or B row) in Kore order maps to degree order as a few increasing
  // runs (one per section and degree), disjoint in their columns: ordering
  // the runs by their first column sorts the row.  Rows whose runs overlap
  // (unsorted input) fall back to an insertion sort.
  auto sort_runs = [](std::vector<Ent>& in, std::vector<Ent>& out, std::vector<std::pair<int32_t, int32_t>>& runs) {
    out.clear();
    runs.clear();
    const int32_t m = (int32_t)in.size();
    for (int32_t q = 0; q < m;) {
      int32_t e = q + 1;
      while (e < m && in[e].col > in[e - 1].col) ++e;
      runs.push_back({q, e});
      q = e;
    }
    if (runs.size() > 1) {
      std::sort(runs.begin(), runs.end(), [&](const std::pair<int32_t, int32_t>& x, const std::pair<int32_t, int32_t>& y) {
        return in[x.first].col < in[y.first].col;
      });
      for (size_t k = 1; k < runs.size(); ++k)
        if (in[runs[k - 1].second - 1].col >= in[runs[k].first].col) {   // overlapping runs
          out = in;
          for (size_t q = 1; q < out.size(); ++q) {
            const Ent e = out[q];
            size_t p = q;
            while (p > 0 && out[p - 1].col > e.col) {
              out[p] = out[p - 1];
              --p;
            }
            out[p] = e;
          }
          return;
        }
    }
    for (const auto& rg : runs) out.insert(out.end(), in.begin() + rg.first, in.begin() + rg.second);
  };
  auto build_rows = [&](int t) {
    const int64_t r0 = n * t / nth, r1 = n * (t + 1) / nth;
    std::vector<Ent> row, ra, rb, sa, sb;
    std::vector<std::pair<int32_t, int32_t>> runs;
    int32_t* oix = pix + ub[t];
    zd* ova = pva + ub[t];
    zd* ovb = pvb + ub[t];
    int32_t* oixb = pixb + ubb[t];
    zd* ovbb = pvbb + ubb[t];
    size_t cntt = 0, cnttb = 0;
    for (int64_t i = r0; i < r1; ++i) {
      const int64_t r = s->perm[i];
      ra.clear();
      rb.clear();
      for (int32_t k = a_indptr[r]; k < a_indptr[r + 1]; ++k) ra.push_back({(int32_t)inv[a_indices[k]], k});
      if (s->has_b)
        for (int32_t k = b_indptr[r]; k < b_indptr[r + 1]; ++k) rb.push_back({(int32_t)inv[b_indices[k]], -(k + 1)});
      // each part sorted by its runs, then the two merged (A before B on equal
      // columns); the buffers are reused from row to row
      sort_runs(ra, sa, runs);
      sort_runs(rb, sb, runs);
      row.clear();
      {
        size_t p = 0, q = 0;
        while (p < sa.size() || q < sb.size()) {
          if (q == sb.size() || (p < sa.size() && sa[p].col <= sb[q].col)) row.push_back(sa[p++]);
          else row.push_back(sb[q++]);
        }
      }
      const int kb = s->h_blk[i];
      int32_t cnt = 0, cntb = 0;
      for (size_t q = 0; q < row.size();) {
        const int32_t col = row[q].col;
        const size_t q0 = q;
        zd a = 0, b = 0;
        while (q < row.size() && row[q].col == col) {
          const int64_t k = row[q].src;
          if (k >= 0) {
            if (a1_data) {
              // e a1, then + a0, each rounded (no fused multiply-add): the
              // values numpy's a1 * e + a0 gives
#pragma clang fp contract(off)
              const double tr = e * a1_data[2 * k], ti = e * a1_data[2 * k + 1];
              a += zd(a_data[2 * k] + tr, a_data[2 * k + 1] + ti);
            } else {
              a += zd(a_data[2 * k], a_data[2 * k + 1]);
            }
          } else {
            const int64_t kb = -(k + 1);
            b += b_is_complex ? zd(b_data[2 * kb], b_data[2 * kb + 1]) : zd(b_data[kb], 0.0);
          }
          ++q;
        }
        const int kc = s->h_blk[col];
        const bool inband = kc >= kb - 1 && kc <= kb + 1;
        // explicit zeros inside the three block diagonals stay in the pattern
        // (kore_sys_update_a may give them values); outside it they are dropped
        if (a == zd(0, 0) && b == zd(0, 0) && !inband) continue;
        for (size_t qq = q0; qq < q; ++qq)
          if (row[qq].src >= 0) s->h_amap[row[qq].src] = (int32_t)cntt;   // range-local position, fixed below
        if (!inband && bad_row[t] < 0) {
          bad_row[t] = i;
          bad_kb[t] = kb;
          bad_kc[t] = kc;
        }
        oix[cntt] = col;
        ova[cntt] = a;
        ovb[cntt] = b;
        ++cntt;
        ++cnt;
        if (kc != kb) {
          ++rowoff[i];
          tocol[t].push_back(col);
        }
        if (kc < kb) ++rowlo[i];
        if (b != zd(0, 0)) {
          oixb[cnttb] = col;
          ovbb[cnttb] = b;
          ++cnttb;
          ++cntb;
        }
      }
      rowcnt[i] = cnt;
      rowcntb[i] = cntb;
    }
    tcnt[t] = cntt;
    tcntb[t] = cnttb;
  };
  {
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; ++t) pool.emplace_back(build_rows, t);
    trace.mark("sys_create checks");
    build_rows(0);
    for (auto& th : pool) th.join();
  }
  trace.mark("sys_create union pattern");
  // positions of the A entries: range-local -> global
  {
    std::vector<int64_t> rbase(nth + 1, 0);
    for (int t = 0; t < nth; ++t) rbase[t + 1] = rbase[t] + (int64_t)tcnt[t];
    auto fix = [&](int t) {
      const int64_t r0 = n * t / nth, r1 = n * (t + 1) / nth;
      for (int64_t i = r0; i < r1; ++i) {
        const int64_t r = s->perm[i];
        for (int32_t k = a_indptr[r]; k < a_indptr[r + 1]; ++k)
          if (s->h_amap[k] >= 0) s->h_amap[k] = (int32_t)(s->h_amap[k] + rbase[t]);
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; ++t) pool.emplace_back(fix, t);
    fix(0);
    for (auto& th : pool) th.join();
  }
  for (int t = 0; t < nth; ++t)
    if (bad_row[t] >= 0)
      return fail(KORE_ERR_UNSUPPORTED,
                  "matrix is not block-tridiagonal in degree order (row %lld couples block %d to %d)",
                  (long long)bad_row[t], bad_kb[t], bad_kc[t]);
  s->h_ip.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) s->h_ip[i + 1] = s->h_ip[i] + rowcnt[i];
  s->h_split.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    s->off_nnz[s->h_blk[i] & 1] += rowoff[i];
    s->off_rows[s->h_blk[i] & 1] += 1;
    const int32_t lo = rowlo[i], hi = rowoff[i] - rowlo[i];
    if (lo >= (1 << 15) || hi >= (1 << 15))
      return fail(KORE_ERR_UNSUPPORTED, "row %lld has more than 32767 entries in a neighbouring block", (long long)i);
    s->h_split[i] = lo | (hi << 15) | ((s->h_blk[i] & 1) << 30);
  }
  // compact off-block CSR: rows in degree order, each row's previous-block
  // entries first (rowlo of them), then its next-block entries
  std::vector<int32_t>& optr = P->optr;
  std::vector<int32_t>& onlo = P->onlo;
  std::vector<int32_t>& ocol = P->ocol;
  optr.assign(n + 1, 0);
  onlo.assign(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    optr[i + 1] = optr[i] + rowoff[i];
    onlo[i] = rowlo[i];
  }
  s->nnz_off = optr[n];
  ocol.reserve((size_t)s->nnz_off);
  for (int t = 0; t < nth; ++t) ocol.insert(ocol.end(), tocol[t].begin(), tocol[t].end());
  if ((int64_t)ocol.size() != s->nnz_off) return fail(KORE_ERR_ARG, "internal: off-block entry count mismatch");
  // column segments per (block, side): for column c of the neighbouring
  // block, the compact entries q of the block's rows in that column
  s->h_cbase.assign(2 * (size_t)s->nblk, -1);
  s->h_side_nnz.assign(2 * (size_t)s->nblk, 0);
  std::vector<int64_t> ebase(2 * (size_t)s->nblk, 0);
  int64_t ncp = 0, ne = 0;
  for (int k = 0; k < s->nblk; ++k) {
    for (int sd = 0; sd < 2; ++sd) {
      const int nbk = sd == 0 ? k - 1 : k + 1;
      int64_t cnt = 0;
      for (int64_t i = s->boff[k]; i < s->boff[k + 1]; ++i) cnt += sd == 0 ? onlo[i] : rowoff[i] - onlo[i];
      s->h_side_nnz[2 * k + sd] = cnt;
      if (nbk < 0 || nbk >= s->nblk) continue;
      s->h_cbase[2 * k + sd] = ncp;
      ebase[2 * k + sd] = ne;
      ncp += s->bsz[nbk] + 1;
      ne += cnt;
    }
  }
  if (ne != s->nnz_off) return fail(KORE_ERR_ARG, "internal: off-block side counts mismatch");
  std::vector<int32_t>& cptr = P->cptr;
  std::vector<int32_t>& cq = P->cq;
  std::vector<int32_t>& crow = P->crow;
  cptr.assign((size_t)std::max<int64_t>(ncp, 1), 0);
  cq.assign((size_t)std::max<int64_t>(ne, 1), 0);
  crow.assign((size_t)std::max<int64_t>(ne, 1), 0);
  {
    auto fill_blocks = [&](int k0, int k1) {
      std::vector<int32_t> fillp;
      for (int k = k0; k < k1; ++k)
        for (int sd = 0; sd < 2; ++sd) {
          if (s->h_cbase[2 * k + sd] < 0) continue;
          const int nbk = sd == 0 ? k - 1 : k + 1;
          const int64_t c0 = s->boff[nbk], nc = s->bsz[nbk];
          int32_t* cp = cptr.data() + s->h_cbase[2 * k + sd];
          for (int64_t c = 0; c <= nc; ++c) cp[c] = 0;
          for (int64_t i = s->boff[k]; i < s->boff[k + 1]; ++i) {
            const int32_t qa = sd == 0 ? optr[i] : optr[i] + onlo[i], qb = sd == 0 ? optr[i] + onlo[i] : optr[i + 1];
            for (int32_t q = qa; q < qb; ++q) ++cp[ocol[q] - c0 + 1];
          }
          cp[0] = (int32_t)ebase[2 * k + sd];
          for (int64_t c = 0; c < nc; ++c) cp[c + 1] += cp[c];
          fillp.assign(cp, cp + nc);
          for (int64_t i = s->boff[k]; i < s->boff[k + 1]; ++i) {
            const int32_t qa = sd == 0 ? optr[i] : optr[i] + onlo[i], qb = sd == 0 ? optr[i] + onlo[i] : optr[i + 1];
            for (int32_t q = qa; q < qb; ++q) {
              const int32_t at = fillp[ocol[q] - c0]++;
              cq[at] = q;
              crow[at] = (int32_t)(i - s->boff[k]);
            }
          }
        }
    };
    std::vector<std::thread> pool;
    const int nthb = std::max(1, std::min(nth, s->nblk));
    for (int t = 1; t < nthb; ++t) pool.emplace_back(fill_blocks, s->nblk * t / nthb, s->nblk * (t + 1) / nthb);
    fill_blocks(0, s->nblk / nthb);
    for (auto& th : pool) th.join();
  }
  trace.mark("sys_create off-block views");
  return KORE_OK;
}

// Device buffers and upload of a prepared system on ctx's stream.
static int32_t sys_upload(kore_ctx* ctx, kore_sys_plan* P, kore_sys** out) {
  StageTrace trace;
  std::unique_ptr<kore_sys>& s = P->s;
  s->ctx = ctx;
  const int64_t n = s->n;
  const int nth = P->nth;
  const auto& tcnt = P->tcnt;
  const auto& tcntb = P->tcntb;
  const auto& ub = P->ub;
  const auto& ubb = P->ubb;
  const auto& rowcntb = P->rowcntb;
  const auto& optr = P->optr;
  const auto& onlo = P->onlo;
  const auto& ocol = P->ocol;
  const auto& cptr = P->cptr;
  const auto& cq = P->cq;
  const auto& crow = P->crow;
  const zd* pva = P->pva;
  const zd* pvb = P->pvb;
  const zd* pvbb = P->pvbb;
  const int32_t* pix = P->pix;
  const int32_t* pixb = P->pixb;
  const size_t nnz = (size_t)s->h_ip[n];
  s->nnz = (int64_t)nnz;
  TRY(s->d_ip.alloc(ctx, (n + 1) * sizeof(int32_t)));
  TRY(s->d_ix.alloc(ctx, std::max<size_t>(nnz, 1) * sizeof(int32_t)));
  TRY(s->d_va.alloc(ctx, std::max<size_t>(nnz, 1) * sizeof(zd)));
  TRY(s->d_vb.alloc(ctx, std::max<size_t>(nnz, 1) * sizeof(zd)));
  std::vector<int32_t> ipb(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) ipb[i + 1] = ipb[i] + rowcntb[i];
  const size_t nnzb = (size_t)ipb[n];
  if (s->has_b) {
    TRY(s->d_ipb.alloc(ctx, (n + 1) * sizeof(int32_t)));
    TRY(s->d_ixb.alloc(ctx, std::max<size_t>(nnzb, 1) * sizeof(int32_t)));
    TRY(s->d_vbb.alloc(ctx, std::max<size_t>(nnzb, 1) * sizeof(zd)));
  }
  TRY(s->d_perm.alloc(ctx, n * sizeof(int64_t)));
  TRY(s->d_blk.alloc(ctx, n * sizeof(int32_t)));
  TRY(s->d_split.alloc(ctx, n * sizeof(int32_t)));
  TRY(s->d_boff.alloc(ctx, (s->nblk + 1) * sizeof(int64_t)));
  TRY(s->d_scale.alloc(ctx, n * sizeof(double)));
  TRY(s->d_optr.alloc(ctx, (n + 1) * sizeof(int32_t)));
  TRY(s->d_onlo.alloc(ctx, n * sizeof(int32_t)));
  TRY(s->d_ocol.alloc(ctx, std::max<size_t>(ocol.size(), 1) * sizeof(int32_t)));
  TRY(s->d_oval.alloc(ctx, std::max<size_t>(ocol.size(), 1) * sizeof(zd)));
  TRY(s->d_cptr.alloc(ctx, cptr.size() * sizeof(int32_t)));
  TRY(s->d_cq.alloc(ctx, cq.size() * sizeof(int32_t)));
  TRY(s->d_crow.alloc(ctx, crow.size() * sizeof(int32_t)));
  TRY(s->d_scr.alloc(ctx, 4 * (size_t)n * MAXR * sizeof(cx<double>)));
  TRY(s->d_part.alloc(ctx, 2 * (size_t)s->npart * MAXR * sizeof(double)));
  hipStream_t st = ctx->stream;
  trace.mark("sys_create device alloc");
  HIP_TRY(hipMemcpyAsync(s->d_ip.p, s->h_ip.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
  // each row range goes from its place in the pinned arena to its place on the device
  {
    size_t off = 0;
    for (int t = 0; t < nth; ++t) {
      const size_t cnt = tcnt[t];
      if (cnt) {
        HIP_TRY(hipMemcpyAsync(s->d_ix.as<int32_t>() + off, pix + ub[t], cnt * sizeof(int32_t),
                               hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(s->d_va.as<zd>() + off, pva + ub[t], cnt * sizeof(zd), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(s->d_vb.as<zd>() + off, pvb + ub[t], cnt * sizeof(zd), hipMemcpyHostToDevice, st));
      }
      off += cnt;
    }
  }
  if (s->has_b) {
    HIP_TRY(hipMemcpyAsync(s->d_ipb.p, ipb.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    size_t off = 0;
    for (int t = 0; t < nth; ++t) {
      const size_t cnt = tcntb[t];
      if (cnt) {
        HIP_TRY(hipMemcpyAsync(s->d_ixb.as<int32_t>() + off, pixb + ubb[t], cnt * sizeof(int32_t),
                               hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(s->d_vbb.as<zd>() + off, pvbb + ubb[t], cnt * sizeof(zd), hipMemcpyHostToDevice, st));
      }
      off += cnt;
    }
  }
  HIP_TRY(hipMemcpyAsync(s->d_perm.p, s->perm.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_blk.p, s->h_blk.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_split.p, s->h_split.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_boff.p, s->boff.data(), (s->nblk + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_optr.p, optr.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_onlo.p, onlo.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (!ocol.empty())
    HIP_TRY(hipMemcpyAsync(s->d_ocol.p, ocol.data(), ocol.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_cptr.p, cptr.data(), cptr.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_cq.p, cq.data(), cq.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_crow.p, crow.data(), crow.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  s->off = OffCsr{s->d_optr.as<int32_t>(), s->d_onlo.as<int32_t>(), s->d_ocol.as<int32_t>(), s->d_oval.as<cx<double>>()};
  s->offc = OffCsc{s->d_cptr.as<int32_t>(), s->d_cq.as<int32_t>(), s->d_crow.as<int32_t>()};
  s->csr = CsrDev{n, s->d_ip.as<int32_t>(), s->d_ix.as<int32_t>(), s->d_va.as<cx<double>>(), s->d_vb.as<cx<double>>()};
  // B alone: its A values are never read (ca = 0 skips them)
  if (s->has_b)
    s->csr_b = CsrDev{n, s->d_ipb.as<int32_t>(), s->d_ixb.as<int32_t>(), s->d_vbb.as<cx<double>>(),
                      s->d_vbb.as<cx<double>>()};
  cx<double>* scr = s->d_scr.as<cx<double>>();
  s->bb = scr;
  s->xr = scr + (size_t)n * MAXR;
  s->rr = scr + 2 * (size_t)n * MAXR;
  s->st = scr + 3 * (size_t)n * MAXR;
  s->part = s->d_part.as<double>();
  TRY(s->up.init(ctx, 1 << 20));
  trace.mark("sys_create upload");
  *out = s.release();
  return KORE_OK;
}

extern "C" {

int32_t kore_sys_prepare(const kore_layout* lay, int64_t sizmat, const int32_t* a_indptr, const int32_t* a_indices,
                         const double* a_data, const int32_t* b_indptr, const int32_t* b_indices,
                         const double* b_data, int32_t b_is_complex, kore_sys_plan** out) {
  if (!out) return fail(KORE_ERR_ARG, "NULL argument");
  *out = nullptr;
  std::unique_ptr<kore_sys_plan> P;
  TRY(sys_prepare(lay, sizmat, a_indptr, a_indices, a_data, b_indptr, b_indices, b_data, b_is_complex, P));
  *out = P.release();
  return KORE_OK;
}

int32_t kore_sys_prepare_affine(const kore_layout* lay, int64_t sizmat, const int32_t* a_indptr,
                                const int32_t* a_indices, const double* a0_data, const double* a1_data, double e,
                                const int32_t* b_indptr, const int32_t* b_indices, const double* b_data,
                                int32_t b_is_complex, kore_sys_plan** out) {
  if (!out || !a1_data) return fail(KORE_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (!std::isfinite(e)) return fail(KORE_ERR_ARG, "the coefficient of A1 is not finite");
  std::unique_ptr<kore_sys_plan> P;
  TRY(sys_prepare(lay, sizmat, a_indptr, a_indices, a0_data, b_indptr, b_indices, b_data, b_is_complex, P, a1_data,
                  e));
  *out = P.release();
  return KORE_OK;
}

int32_t kore_sys_create_prepared(kore_ctx* ctx, kore_sys_plan* plan, kore_sys** out) {
  if (!ctx || !plan || !out || !plan->s) return fail(KORE_ERR_ARG, "NULL argument or a plan already used");
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  const int32_t rc = sys_upload(ctx, plan, out);
  if (rc < 0) hipStreamSynchronize(ctx->stream);   // no copy may still read the staging
  plan->arena.release();
  return rc;
}

void kore_sys_plan_destroy(kore_sys_plan* plan) { delete plan; }

int32_t kore_sys_create(kore_ctx* ctx, const kore_layout* lay, int64_t sizmat, const int32_t* a_indptr,
                        const int32_t* a_indices, const double* a_data, const int32_t* b_indptr,
                        const int32_t* b_indices, const double* b_data, int32_t b_is_complex, kore_sys** out) {
  if (!ctx || !out) return fail(KORE_ERR_ARG, "NULL argument");
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  std::unique_ptr<kore_sys_plan> P;
  TRY(sys_prepare(lay, sizmat, a_indptr, a_indices, a_data, b_indptr, b_indices, b_data, b_is_complex, P));
  const int32_t rc = sys_upload(ctx, P.get(), out);
  if (rc < 0) hipStreamSynchronize(ctx->stream);   // no copy may still read the staging
  P->arena.release();
  return rc;
}

int32_t kore_sys_update_a(kore_sys* s, const double* a_data) {
  if (!s || !a_data) return fail(KORE_ERR_ARG, "NULL argument");
  kore_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  char* arena = nullptr;
  TRY(ctx_pinned(c, (size_t)std::max<int64_t>(s->nnz, 1) * sizeof(zd), &arena));
  zd* va = reinterpret_cast<zd*>(arena);
  std::fill(va, va + s->nnz, zd(0, 0));
  for (size_t k = 0; k < s->h_amap.size(); ++k) {
    const zd v(a_data[2 * k], a_data[2 * k + 1]);
    const int32_t pos = s->h_amap[k];
    if (pos < 0) {
      if (v != zd(0, 0))
        return fail(KORE_ERR_UNSUPPORTED, "A entry %zu was a zero outside the block-tridiagonal pattern at "
                    "creation and cannot take a value now", k);
      continue;
    }
    va[pos] += v;
  }
  if (s->nnz > 0)
    HIP_TRY(hipMemcpyAsync(s->d_va.p, va, (size_t)s->nnz * sizeof(zd), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  s->factored = false;            // the factorisation (if any) belongs to the old values
  return KORE_OK;
}

void kore_sys_destroy(kore_sys* sys) {
  if (!sys) return;
  hipSetDevice(sys->ctx->device);
  hipStreamSynchronize(sys->ctx->stream);
  std::vector<kore_ws*> left;
  {
    std::lock_guard<std::mutex> lk(sys->ws_mu);
    left.swap(sys->wss);
  }
  for (kore_ws* w : left) delete w;
  delete sys;
}

int32_t kore_sys_query(const kore_sys* s, kore_sys_state* out) {
  if (!s || !out) return fail(KORE_ERR_ARG, "NULL argument");
  out->sizmat = s->n;
  out->nblk = s->nblk;
  out->has_b = s->has_b ? 1 : 0;
  out->factored = (s->f && s->factored) ? 1 : 0;
  out->precision = s->f ? s->f->prec : -1;
  out->sre = s->sre;
  out->sim = s->sim;
  return KORE_OK;
}

int32_t kore_set_host_threads(int32_t n) {
  if (n > 64) return fail(KORE_ERR_ARG, "host threads %d: at most 64", n);
  g_host_threads.store(n > 0 ? n : 0, std::memory_order_relaxed);
  return KORE_OK;
}

int32_t kore_host_threads(void) { return host_threads(); }

int32_t kore_cpu_share(void) { return cpu_share(); }

int32_t kore_sys_set_factor_form(kore_sys* s, int32_t form) {
  if (!s) return fail(KORE_ERR_ARG, "NULL argument");
  if (form != KORE_FORM_AUTO && form != KORE_FORM_SWEEP && form != KORE_FORM_LEAN)
    return fail(KORE_ERR_ARG, "unknown factor form %d", form);
  s->form = form;
  return KORE_OK;
}

int32_t kore_sys_norm(kore_sys* s, double sre, double sim, double* out) {
  if (!s || !out) return fail(KORE_ERR_ARG, "NULL argument");
  kore_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int nblk = std::min(2 * s->npart * MAXR, 1024);
  if (s->nnz > 0) HIP_TRY(launch_csr_fro2(s->csr.va, s->csr.vb, s->nnz, sre, sim, s->part, nblk, c->stream));
  std::vector<double> h(nblk, 0.0);
  if (s->nnz > 0) {
    HIP_TRY(hipMemcpyAsync(h.data(), s->part, nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  double acc = 0;
  for (double v : h) acc += v;
  *out = std::sqrt(acc);
  return KORE_OK;
}

}  // extern "C"

// Lean form (no E, F products) for this system's next factorisation.  Auto:
// lean from degree blocks of LEAN_AUTO_MIN unknowns, where the quarter of the
// GEMM flops it saves outweighs its extra sweep launch per level.
static bool factor_lean(const kore_sys* s) {
  if (s->form == KORE_FORM_LEAN) return true;
  if (s->form == KORE_FORM_SWEEP) return false;
  return s->nmax >= LEAN_AUTO_MIN;
}

// Queues the factorisation of A - sigma B on the context's stream (plan and
// tile choice first, which may wait for the device); e0/e1 bracket the
// factorisation proper when given.  *dinfo: the device pivot flags.
static int32_t factor_enqueue(kore_sys* s, double sre, double sim, int32_t precision, hipEvent_t e0, hipEvent_t e1,
                              const int** dinfo) {
  kore_ctx* c = s->ctx;
  // keep the other precision's plan (and its device memory) when the
  // precision changes, so that alternating precisions does not rebuild plans
  const bool lean = factor_lean(s);
  auto other = [&](const std::unique_ptr<FactorBase>& f) { return f && (f->prec != precision || f->lean != lean); };
  if (other(s->f)) {
    std::swap(s->f, s->fspare);
    if (other(s->f)) s->f.reset();
  }
  s->sre = sre;
  s->sim = sim;
  s->factored = false;
  s->info_pending = false;
  // power-of-two row equilibration of A - sigma B (device), then the
  // level-0 couplings' values R (A - sigma B) for the compact views
  HIP_TRY(launch_row_scale(s->csr, sre, sim, s->d_scale.as<double>(), c->stream));
  if (s->nnz_off > 0)
    HIP_TRY(launch_offvals(s->csr, sre, sim, s->d_scale.as<double>(), s->d_split.as<int32_t>(), s->off, c->stream));
  int32_t rc = KORE_OK;
  if (precision == KORE_PREC_FP64) {
    if (!s->f) {
      auto F = std::make_unique<Factor<double>>();
      F->prec = precision;
      F->lean = lean;
      rc = factor_build<double>(s, F.get());
      if (rc >= 0) s->f = std::move(F);
    }
    if (rc >= 0) {
      Factor<double>* F = static_cast<Factor<double>*>(s->f.get());
      rc = tune_factor<double>(s, F, precision);
      if (e0) hipEventRecord(e0, c->stream);
      if (rc >= 0) rc = factor_run<double>(s, F);
      if (e1) hipEventRecord(e1, c->stream);
      *dinfo = F->info;
    }
  } else {
    if (!s->f) {
      auto F = std::make_unique<Factor<float>>();
      F->prec = precision;
      F->lean = lean;
      rc = factor_build<float>(s, F.get());
      if (rc >= 0) s->f = std::move(F);
    }
    if (rc >= 0) {
      Factor<float>* F = static_cast<Factor<float>*>(s->f.get());
      rc = tune_factor<float>(s, F, precision);
      if (e0) hipEventRecord(e0, c->stream);
      if (rc >= 0) rc = factor_run<float>(s, F);
      if (e1) hipEventRecord(e1, c->stream);
      *dinfo = F->info;
    }
  }
  if (rc < 0) s->f.reset();
  return rc;
}

// Pivot flags (host copy) -> status; a zero pivot leaves the system unfactored.
static int32_t factor_check_info(kore_sys* s, const int* hinfo) {
  s->factored = true;
  for (int k = 0; k < s->nblk; ++k)
    if (hinfo[k] != 0) {
      s->factored = false;
      return fail(KORE_ERR_SINGULAR, "zero pivot in degree block %d (column %d) during cyclic reduction", k, hinfo[k]);
    }
  return KORE_OK;
}

// The context's stream waits for the work queued so far on `user`.
static int32_t stream_join_in(kore_sys* s, hipStream_t user) {
  kore_ctx* c = s->ctx;
  if (user == c->stream) return KORE_OK;
  if (!s->ev_in) HIP_TRY(hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(s->ev_in, user));
  HIP_TRY(hipStreamWaitEvent(c->stream, s->ev_in, 0));
  return KORE_OK;
}

// `user` waits for the work queued so far on the context's stream.
static int32_t stream_join_out(kore_sys* s, hipStream_t user) {
  kore_ctx* c = s->ctx;
  if (user == c->stream) return KORE_OK;
  if (!s->ev_out) HIP_TRY(hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(s->ev_out, c->stream));
  HIP_TRY(hipStreamWaitEvent(user, s->ev_out, 0));
  return KORE_OK;
}

extern "C" {

int32_t kore_factor(kore_sys* s, double sre, double sim, int32_t precision, kore_factor_info* info) {
  if (!s) return fail(KORE_ERR_ARG, "sys is NULL");
  if (precision != KORE_PREC_FP64 && precision != KORE_PREC_C64) return fail(KORE_ERR_ARG, "bad precision");
  kore_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  std::vector<int> hinfo(s->nblk, 0);
  const int* dinfo = nullptr;
  int32_t rc = factor_enqueue(s, sre, sim, precision, e0, e1, &dinfo);
  if (rc >= 0 && dinfo) hipMemcpyAsync(hinfo.data(), dinfo, s->nblk * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  hipError_t se = hipStreamSynchronize(c->stream);
  float ms = 0.f;
  if (rc >= 0 && se == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc < 0) return rc;
  if (se != hipSuccess) {
    s->f.reset();
    return fail(KORE_ERR_HIP, "factorisation failed: %s", hipGetErrorString(se));
  }
  TRY(factor_check_info(s, hinfo.data()));
  if (info) {
    info->levels = s->f->nlev;
    info->nblk = s->nblk;
    info->flops = s->f->flops;
    info->flops_inv = s->f->flops_inv;
    info->flops_gemm = s->f->flops_gemm;
    info->flops_sparse = s->f->flops_sparse;
    info->solve_bytes = s->f->solve_bytes;
    info->solve_launches = s->f->solve_launches;
    info->form = s->f->lean ? KORE_FORM_LEAN : KORE_FORM_SWEEP;
    info->bytes = s->f->bytes;
    info->ms = ms;
  }
  return KORE_OK;
}

int32_t kore_factor_async(kore_sys* s, double sre, double sim, int32_t precision, void* stream) {
  if (!s) return fail(KORE_ERR_ARG, "sys is NULL");
  if (precision != KORE_PREC_FP64 && precision != KORE_PREC_C64) return fail(KORE_ERR_ARG, "bad precision");
  kore_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const hipStream_t user = static_cast<hipStream_t>(stream);
  TRY(stream_join_in(s, user));
  TRY(s->info_pin.reserve(s->nblk * sizeof(int), c->stream));
  const int* dinfo = nullptr;
  TRY(factor_enqueue(s, sre, sim, precision, nullptr, nullptr, &dinfo));
  HIP_TRY(hipMemcpyAsync(s->info_pin.p, dinfo, s->nblk * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  TRY(stream_join_out(s, user));
  // solves queued behind this call use the factor; its pivots are checked
  // by kore_factor_check
  s->factored = true;
  s->info_pending = true;
  return KORE_OK;
}

int32_t kore_factor_check(kore_sys* s) {
  if (!s) return fail(KORE_ERR_ARG, "sys is NULL");
  kore_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const hipError_t se = hipStreamSynchronize(c->stream);
  if (se != hipSuccess) {
    s->factored = false;
    return fail(KORE_ERR_HIP, "factorisation failed: %s", hipGetErrorString(se));
  }
  if (!s->info_pending) return s->factored ? KORE_OK : fail(KORE_ERR_STATE, "system is not factored");
  s->info_pending = false;
  return factor_check_info(s, reinterpret_cast<const int*>(s->info_pin.p));
}

int32_t kore_factor_release(kore_sys* s) {
  if (!s) return fail(KORE_ERR_ARG, "sys is NULL");
  hipStreamSynchronize(s->ctx->stream);
  s->f.reset();
  s->fspare.reset();
  s->factored = false;
  return KORE_OK;
}

int32_t kore_solve(kore_sys* s, int32_t nrhs, const double* b, int64_t ldb, double* x, int64_t ldx, double ir_tol,
                   int32_t ir_maxit, kore_solve_info* info) {
  if (!s || !b || !x || nrhs < 0 || ldb < s->n || ldx < s->n) return fail(KORE_ERR_ARG, "bad solve arguments");
  kore_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int64_t n = s->n;
  if (info) {
    info->iterations = 0;
    info->converged = 1;
    info->rel_correction = -1;
  }
  for (int c0 = 0; c0 < nrhs; c0 += MAXR) {
    const int nr = std::min(MAXR, nrhs - c0);
    for (int j = 0; j < nr; ++j)
      HIP_TRY(hipMemcpyAsync(s->st + (size_t)j * n, b + 2 * (ldb * (int64_t)(c0 + j)), n * sizeof(zd),
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_gather_scale<double, double>(s->st, n, s->bb, n, n, nr, s->d_perm.as<int64_t>(), nullptr, 1.0,
                                                c->stream));
    kore_solve_info li{};
    TRY(solve_dev(s, s->bb, n, s->xr, n, nr, ir_tol, ir_maxit, &li));
    HIP_TRY(launch_scatter(s->xr, n, s->st, n, n, nr, s->d_perm.as<int64_t>(), c->stream));
    for (int j = 0; j < nr; ++j)
      HIP_TRY(hipMemcpyAsync(x + 2 * (ldx * (int64_t)(c0 + j)), s->st + (size_t)j * n, n * sizeof(zd),
                             hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (info) {
      info->iterations = std::max(info->iterations, li.iterations);
      info->converged = info->converged && li.converged;
      info->rel_correction = std::max(info->rel_correction, li.rel_correction);
    }
  }
  return (info && !info->converged) ? KORE_WARN_IR_NOT_CONVERGED : KORE_OK;
}

int32_t kore_solve_async(kore_sys* s, int32_t nrhs, const double* d_b, int64_t ldb, double* d_x, int64_t ldx,
                         int32_t ir_sweeps, void* stream) {
  if (!s || !d_b || !d_x || nrhs < 0 || ldb < s->n || ldx < s->n || ir_sweeps < 0)
    return fail(KORE_ERR_ARG, "bad solve arguments");
  if (!s->f || !s->factored) return fail(KORE_ERR_STATE, "no factorisation: call kore_factor first");
  kore_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const hipStream_t user = static_cast<hipStream_t>(stream);
  const hipStream_t st = c->stream;
  const int64_t n = s->n;
  const bool c64 = s->f->prec == KORE_PREC_C64;
  TRY(stream_join_in(s, user));
  const cx<double>* b = reinterpret_cast<const cx<double>*>(d_b);
  cx<double>* x = reinterpret_cast<cx<double>*>(d_x);
  for (int c0 = 0; c0 < nrhs; c0 += MAXR) {
    const int nr = std::min(MAXR, nrhs - c0);
    // Kore order -> degree order, the factor solve, then fixed refinement
    // sweeps (r = b - (A - sigma B) x by FP64 SpMV, x += factor^{-1} R r)
    HIP_TRY(launch_gather_scale<double, double>(b + (size_t)c0 * ldb, ldb, s->bb, n, n, nr, s->d_perm.as<int64_t>(),
                                                nullptr, 1.0, st));
    if (c64) TRY(refine_sweep<float>(s, static_cast<Factor<float>*>(s->f.get()), s->bb, n, s->xr, n, nr, true));
    else TRY(refine_sweep<double>(s, static_cast<Factor<double>*>(s->f.get()), s->bb, n, s->xr, n, nr, true));
    for (int it = 0; it < ir_sweeps; ++it) {
      HIP_TRY(launch_copy_cols(s->bb, n, s->rr, n, n, nr, st));
      HIP_TRY(launch_spmv(s->csr, -1.0, 0.0, s->sre, s->sim, s->xr, n, s->rr, n, nr, 1.0, st));
      if (c64) {
        Factor<float>* F = static_cast<Factor<float>*>(s->f.get());
        TRY(refine_sweep<float>(s, F, s->rr, n, s->xr, n, nr, false));
        HIP_TRY(launch_axpy<float>(s->xr, n, F->Xv, n, n, nr, 1.0, st));
      } else {
        Factor<double>* F = static_cast<Factor<double>*>(s->f.get());
        TRY(refine_sweep<double>(s, F, s->rr, n, s->xr, n, nr, false));
        HIP_TRY(launch_axpy<double>(s->xr, n, F->Xv, n, n, nr, 1.0, st));
      }
    }
    HIP_TRY(launch_scatter(s->xr, n, x + (size_t)c0 * ldx, ldx, n, nr, s->d_perm.as<int64_t>(), st));
  }
  TRY(stream_join_out(s, user));
  return KORE_OK;
}

int32_t kore_residual(kore_sys* s, int32_t nrhs, const double* x, int64_t ldx, const double* lam, const double* b,
                      int64_t ldb, double sre, double sim, double* rnorm) {
  if (!s || !x || !rnorm || nrhs < 0 || ldx < s->n || (b && ldb < s->n))
    return fail(KORE_ERR_ARG, "bad residual arguments");
  if (!s->has_b && (lam || sre != 0.0 || sim != 0.0)) return fail(KORE_ERR_STATE, "system has no B matrix");
  kore_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int64_t n = s->n;
  for (int c0 = 0; c0 < nrhs; c0 += MAXR) {
    const int nr = std::min(MAXR, nrhs - c0);
    for (int j = 0; j < nr; ++j)
      HIP_TRY(hipMemcpyAsync(s->st + (size_t)j * n, x + 2 * (ldx * (int64_t)(c0 + j)), n * sizeof(zd),
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_gather_scale<double, double>(s->st, n, s->bb, n, n, nr, s->d_perm.as<int64_t>(), nullptr, 1.0,
                                                c->stream));
    {
      Timed tm(c, KORE_TIMER_SPMV);
      for (int j = 0; j < nr; ++j) {
        const double lre = lam ? lam[2 * (c0 + j)] : 0.0, lim = lam ? lam[2 * (c0 + j) + 1] : 0.0;
        HIP_TRY(launch_spmv(s->csr, 1.0, 0.0, -(sre + lre), -(sim + lim), s->bb + (size_t)j * n, n,
                            s->rr + (size_t)j * n, n, 1, 0.0, c->stream));
      }
    }
    if (b) {
      for (int j = 0; j < nr; ++j)
        HIP_TRY(hipMemcpyAsync(s->st + (size_t)j * n, b + 2 * (ldb * (int64_t)(c0 + j)), n * sizeof(zd),
                               hipMemcpyHostToDevice, c->stream));
      HIP_TRY(launch_gather_scale<double, double>(s->st, n, s->bb, n, n, nr, s->d_perm.as<int64_t>(), nullptr, 1.0,
                                                  c->stream));
      HIP_TRY(launch_axpy<double>(s->rr, n, s->bb, n, n, nr, -1.0, c->stream));
    }
    TRY(colnorms2<double, double>(s, s->rr, n, nullptr, 0, nr, rnorm + c0, nullptr));
  }
  return KORE_OK;
}

int32_t kore_residual_noalloc(kore_sys* s, int32_t nrhs, const double* x, int64_t ldx, const double* lam,
                              const double* b, int64_t ldb, double sre, double sim, double* rnorm) {
  if (!s || !x || !rnorm || nrhs < 0 || ldx < s->n || (b && ldb < s->n))
    return fail(KORE_ERR_ARG, "bad residual arguments");
  if (!s->has_b && (lam || sre != 0.0 || sim != 0.0)) return fail(KORE_ERR_STATE, "system has no B matrix");
  kore_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int64_t n = s->n;
  // layout: staging | gathered | residual (n x nr complex each), then the
  // column partial sums of the norms
  const int npart = s->npart;
  const size_t part_bytes = (size_t)npart * MAXR * sizeof(double);
  const size_t col_bytes = 3 * (size_t)n * sizeof(zd);
  if (col_bytes + part_bytes > NOALLOC_SCRATCH_BYTES)
    return fail(KORE_ERR_ARG,
                "kore_residual_noalloc: %lld unknowns need %zu bytes of scratch per column, more than the %zu "
                "of the static scratch array", (long long)n, col_bytes + part_bytes, NOALLOC_SCRATCH_BYTES);
  char* base = nullptr;
  TRY(scratch_address(c->device, &base));
  const int cols = (int)std::min<size_t>(MAXR, (NOALLOC_SCRATCH_BYTES - part_bytes) / col_bytes);
  double* part = reinterpret_cast<double*>(base);
  cx<double>* stg = reinterpret_cast<cx<double>*>(base + ((part_bytes + 255) & ~size_t(255)));
  cx<double>* gat = stg + (size_t)n * cols;
  cx<double>* res = gat + (size_t)n * cols;
  double h[64 * MAXR];                    // the partial sums on the host (npart <= 64)
  if (npart * MAXR > (int)(sizeof(h) / sizeof(h[0]))) return fail(KORE_ERR_STATE, "internal: too many partials");
  for (int c0 = 0; c0 < nrhs; c0 += cols) {
    const int nr = std::min(cols, nrhs - c0);
    for (int j = 0; j < nr; ++j)
      HIP_TRY(hipMemcpyAsync(stg + (size_t)j * n, x + 2 * (ldx * (int64_t)(c0 + j)), n * sizeof(zd),
                             hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_gather_scale<double, double>(stg, n, gat, n, n, nr, s->d_perm.as<int64_t>(), nullptr, 1.0,
                                                c->stream));
    for (int j = 0; j < nr; ++j) {
      const double lre = lam ? lam[2 * (c0 + j)] : 0.0, lim = lam ? lam[2 * (c0 + j) + 1] : 0.0;
      HIP_TRY(launch_spmv(s->csr, 1.0, 0.0, -(sre + lre), -(sim + lim), gat + (size_t)j * n, n, res + (size_t)j * n,
                          n, 1, 0.0, c->stream));
    }
    if (b) {
      for (int j = 0; j < nr; ++j)
        HIP_TRY(hipMemcpyAsync(stg + (size_t)j * n, b + 2 * (ldb * (int64_t)(c0 + j)), n * sizeof(zd),
                               hipMemcpyHostToDevice, c->stream));
      HIP_TRY(launch_gather_scale<double, double>(stg, n, gat, n, n, nr, s->d_perm.as<int64_t>(), nullptr, 1.0,
                                                  c->stream));
      HIP_TRY(launch_axpy<double>(res, n, gat, n, n, nr, -1.0, c->stream));
    }
    HIP_TRY(launch_colnorm2<double>(res, n, n, nr, part, npart, c->stream));
    HIP_TRY(hipMemcpyAsync(h, part, (size_t)npart * nr * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int j = 0; j < nr; ++j) {
      double sa = 0;
      for (int q = 0; q < npart; ++q) sa += h[(size_t)q * nr + j];
      rnorm[c0 + j] = std::sqrt(sa);
    }
  }
  return KORE_OK;
}

int32_t kore_solve_gmres(kore_sys* s, const double* b, double* x, double tol, int32_t maxit, int32_t restart,
                         kore_solve_info* info) {
  if (!s || !b || !x || maxit < 1) return fail(KORE_ERR_ARG, "bad GMRES arguments");
  if (!s->f || !s->factored) return fail(KORE_ERR_STATE, "no factorisation: call kore_factor first");
  kore_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int64_t n = s->n;
  double nA = 0;
  TRY(kore_sys_norm(s, s->sre, s->sim, &nA));
  HIP_TRY(hipMemcpyAsync(s->st, b, n * sizeof(zd), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(launch_gather_scale<double, double>(s->st, n, s->bb, n, n, 1, s->d_perm.as<int64_t>(), nullptr, 1.0,
                                              c->stream));
  kore_solve_info li{};
  TRY(gmres_dev(s, s->bb, s->xr, tol, maxit, restart, nA, &li));
  HIP_TRY(launch_scatter(s->xr, n, s->st, n, n, 1, s->d_perm.as<int64_t>(), c->stream));
  HIP_TRY(hipMemcpyAsync(x, s->st, n * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (info) *info = li;
  return li.converged ? KORE_OK : KORE_WARN_IR_NOT_CONVERGED;
}

// ---- workspace ---------------------------------------------------------------
}  // extern "C"

static inline cx<double>* wcol(kore_ws* w, int j) { return w->V.as<cx<double>>() + (size_t)j * w->sys->n; }

// Gram-Schmidt updates fused with the inner products that follow them
// (default); KORE_AMD_ORTHO_FUSED=0 selects separate update and product
// launches, read at every expansion step so that one process can compare both.
static bool ortho_fused() {
  const char* e = std::getenv("KORE_AMD_ORTHO_FUSED");
  return !e || std::atoi(e) != 0;
}

// Coefficient block of the expansion (complex entries): h1, h2 (ncol x MAXR
// each), G, R1, Rinv1, R2, Rinv2 (MAXR x MAXR each), then two int flags.
static size_t ws_hb_elems(int ncol) { return 2 * (size_t)ncol * MAXR + 5 * MAXR * MAXR + 1; }

// out (na x nb, column-major, ld na, device) = V[:, a0:a0+na]^H V[:, b0:b0+nb], nb <= MAXR.
static int32_t ws_dot_dev(kore_ws* w, int a0, int na, int b0, int nb, cx<double>* out) {
  kore_sys* s = w->sys;
  kore_ctx* c = s->ctx;
  const int64_t n = s->n;
  // split-K over row chunks so that every CU streams a slice of the basis
  int nchunk = (int)std::min<int64_t>(DOT_MAXCHUNK, (n + DOT_ROWS - 1) / DOT_ROWS);
  if (nchunk < 1) nchunk = 1;
  const int64_t chunk = (n + nchunk - 1) / nchunk;
  cx<double>* part = w->part.as<cx<double>>();
  Timed tm(c, KORE_TIMER_ORTHO);
  HIP_TRY(launch_tsdot(wcol(w, a0), n, na, wcol(w, b0), n, nb, n, nchunk, chunk, part, c->stream));
  HIP_TRY(launch_sum_parts(reinterpret_cast<const double*>(part), nchunk, 2 * na * nb, reinterpret_cast<double*>(out),
                           c->stream));
  return KORE_OK;
}

// V[:, d0:d0+nd] = beta V[:, d0:d0+nd] + alpha V[:, s0:s0+ns] Q, Q on the device (ldq).
static int32_t ws_gemm_dev(kore_ws* w, int d0, int nd, int s0, int ns, const cx<double>* Q, int ldq, double alpha,
                           double beta) {
  kore_sys* s = w->sys;
  kore_ctx* c = s->ctx;
  const int64_t n = s->n;
  if (nd <= MAXR) {
    // tall-skinny product (Gram-Schmidt update, Cholesky-QR): the row-sliced
    // streaming kernel reads V once and wastes no matrix-core tile on
    // padding; the descriptor travels as a kernel argument
    SkinnyDesc<double> d{};
    d.A[0] = wcol(w, s0); d.lda[0] = n;
    d.X[0] = Q; d.ldx[0] = ldq;
    d.k[0] = ns; d.alpha[0] = alpha;
    d.Y = wcol(w, d0); d.ldy = n;
    d.m = (int)n; d.beta = beta;
    Timed tm(c, KORE_TIMER_ORTHO);
    HIP_TRY(launch_skinny1<double>(d, nd, c->stream));
    return KORE_OK;
  }
  GemmDesc<double> g{};
  g.A[0] = wcol(w, s0); g.lda[0] = n;
  g.B[0] = Q; g.ldb[0] = ldq;
  g.k[0] = ns;
  g.C = wcol(w, d0); g.ldc = n;
  g.m = (int)n; g.n = nd; g.opA = 0;
  g.tiles_m = (int)((n + 63) / 64); g.tiles_n = (nd + 63) / 64;
  g.alpha = alpha; g.beta = beta;
  GemmDesc<double>* dd = w->descb.as<GemmDesc<double>>() + DOT_MAXCHUNK + 1;
  TRY(w->up.put(dd, &g, sizeof(g)));
  Timed tm(c, KORE_TIMER_ORTHO);
  HIP_TRY(launch_gemm_batched<double>(dd, 1, std::min(g.tiles_m * g.tiles_n, 65535), c->stream));
  return KORE_OK;
}

static int32_t ws_apply_op_dev(kore_ws* w, int src, int dst, int nc, double ir_tol, int ir_maxit,
                               kore_solve_info* info) {
  kore_sys* s = w->sys;
  const int64_t n = s->n;
  info->iterations = 0;
  info->converged = 1;
  info->rel_correction = -1;
  const bool direct = s->f && s->factored && s->f->prec == KORE_PREC_FP64 && ir_maxit <= 0;
  for (int a = 0; a < nc; a += MAXR) {
    const int nr = std::min(MAXR, nc - a);
    if (direct) {
      // FP64 factor, no refinement: B x lands row-scaled in the factor's
      // right-hand side (the equilibration of the first refinement sweep,
      // fused), then one cyclic-reduction solve
      Factor<double>* F = static_cast<Factor<double>*>(s->f.get());
      {
        Timed tm(s->ctx, KORE_TIMER_SPMV);
        HIP_TRY(launch_spmv(s->csr_b, 0.0, 0.0, 1.0, 0.0, wcol(w, src + a), n, F->Bv, n, nr, 0.0, s->ctx->stream,
                            s->d_scale.as<double>()));
      }
      TRY(bcr_solve<double>(s, F, nr));
      {
        Timed tm(s->ctx, KORE_TIMER_OTHER);
        HIP_TRY(launch_gather_scale<double, double>(F->Xv, n, wcol(w, dst + a), n, n, nr, nullptr, nullptr, 1.0,
                                                    s->ctx->stream));
      }
      info->iterations = std::max(info->iterations, 1);
      continue;
    }
    {
      Timed tm(s->ctx, KORE_TIMER_SPMV);
      HIP_TRY(launch_spmv(s->csr_b, 0.0, 0.0, 1.0, 0.0, wcol(w, src + a), n, s->bb, n, nr, 0.0, s->ctx->stream));
    }
    kore_solve_info li{};
    TRY(solve_dev(s, s->bb, n, wcol(w, dst + a), n, nr, ir_tol, ir_maxit, &li));
    info->iterations = std::max(info->iterations, li.iterations);
    info->converged = info->converged && li.converged;
    info->rel_correction = std::max(info->rel_correction, li.rel_correction);
  }
  return KORE_OK;
}

extern "C" {
int32_t kore_ws_create(kore_sys* s, int32_t ncol, kore_ws** out) {
  if (!s || !out || ncol <= 0) return fail(KORE_ERR_ARG, "bad workspace arguments");
  *out = nullptr;
  auto w = std::make_unique<kore_ws>();
  w->sys = s;
  w->ncol = ncol;
  kore_ctx* c = s->ctx;
  TRY(w->V.alloc(c, (size_t)s->n * ncol * sizeof(zd)));
  TRY(w->part.alloc(c, ((size_t)DOT_MAXCHUNK + 1) * ncol * MAXR * sizeof(zd)));
  TRY(w->qbuf.alloc(c, (size_t)ncol * ncol * sizeof(zd)));
  TRY(w->descb.alloc(c, (DOT_MAXCHUNK + 64) * sizeof(GemmDesc<double>)));
  TRY(w->hb.alloc(c, ws_hb_elems(ncol) * sizeof(zd)));
  TRY(w->hhbuf.reserve(2 * ws_hb_elems(ncol) * sizeof(zd), c->stream));
  w->hh = reinterpret_cast<cx<double>*>(w->hhbuf.p);
  for (hipEvent_t& e : w->hev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  TRY(w->up.init(c, 1 << 22));
  HIP_TRY(launch_zero<double>(w->V.as<cx<double>>(), (int64_t)s->n * ncol, c->stream));
  {
    std::lock_guard<std::mutex> lk(s->ws_mu);
    s->wss.push_back(w.get());
  }
  *out = w.release();
  return KORE_OK;
}

void kore_ws_destroy(kore_ws* w) {
  if (!w) return;
  kore_sys* s = w->sys;
  hipStreamSynchronize(s->ctx->stream);
  {
    std::lock_guard<std::mutex> lk(s->ws_mu);
    auto it = std::find(s->wss.begin(), s->wss.end(), w);
    if (it == s->wss.end()) return;   // already destroyed with its system
    s->wss.erase(it);
  }
  delete w;
}

static int32_t ws_check(kore_ws* w, int a, int na) {
  if (!w) return fail(KORE_ERR_ARG, "workspace is NULL");
  if (a < 0 || na < 0 || a + na > w->ncol) return fail(KORE_ERR_ARG, "column range [%d, %d) outside workspace of %d", a, a + na, w->ncol);
  return KORE_OK;
}

int32_t kore_ws_set(kore_ws* w, int32_t c0, int32_t nc, const double* host, int64_t ld) {
  TRY(ws_check(w, c0, nc));
  kore_sys* s = w->sys;
  const int64_t n = s->n;
  if (ld < n) return fail(KORE_ERR_ARG, "ld < sizmat");
  for (int a = 0; a < nc; a += MAXR) {
    const int nr = std::min(MAXR, nc - a);
    for (int j = 0; j < nr; ++j)
      HIP_TRY(hipMemcpyAsync(s->st + (size_t)j * n, host + 2 * ld * (int64_t)(a + j), n * sizeof(zd),
                             hipMemcpyHostToDevice, s->ctx->stream));
    HIP_TRY(launch_gather_scale<double, double>(s->st, n, wcol(w, c0 + a), n, n, nr, s->d_perm.as<int64_t>(), nullptr,
                                                1.0, s->ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  return KORE_OK;
}

int32_t kore_ws_get(kore_ws* w, int32_t c0, int32_t nc, double* host, int64_t ld) {
  TRY(ws_check(w, c0, nc));
  kore_sys* s = w->sys;
  const int64_t n = s->n;
  if (ld < n) return fail(KORE_ERR_ARG, "ld < sizmat");
  const size_t colb = (size_t)n * sizeof(zd);
  TRY(s->down.reserve(colb * MAXR, s->ctx->stream));   // pooled pinned staging
  std::vector<CopySeg> segs;
  for (int a = 0; a < nc; a += MAXR) {
    const int nr = std::min(MAXR, nc - a);
    HIP_TRY(launch_scatter(wcol(w, c0 + a), n, s->st, n, n, nr, s->d_perm.as<int64_t>(), s->ctx->stream));
    HIP_TRY(hipMemcpyAsync(s->down.host(), s->st, nr * colb, hipMemcpyDeviceToHost, s->ctx->stream));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    segs.clear();
    for (int j = 0; j < nr; ++j)
      segs.push_back(CopySeg{reinterpret_cast<char*>(host + 2 * ld * (int64_t)(a + j)), s->down.host() + j * colb, colb});
    copy_parallel(segs);
  }
  return KORE_OK;
}

int32_t kore_ws_apply_op(kore_ws* w, int32_t src, int32_t dst, int32_t nc, double ir_tol, int32_t ir_maxit,
                         kore_solve_info* info) {
  TRY(ws_check(w, src, nc));
  TRY(ws_check(w, dst, nc));
  kore_sys* s = w->sys;
  if (!s->has_b) return fail(KORE_ERR_STATE, "system has no B matrix");
  if (src < dst + nc && dst < src + nc) return fail(KORE_ERR_ARG, "source and destination overlap");
  kore_solve_info li{};
  TRY(ws_apply_op_dev(w, src, dst, nc, ir_tol, ir_maxit, &li));
  if (info) *info = li;
  return li.converged ? KORE_OK : KORE_WARN_IR_NOT_CONVERGED;
}

int32_t kore_ws_spmv(kore_ws* w, int32_t src, int32_t dst, int32_t nc, double ca_re, double ca_im, double cb_re,
                     double cb_im) {
  TRY(ws_check(w, src, nc));
  TRY(ws_check(w, dst, nc));
  if (src < dst + nc && dst < src + nc) return fail(KORE_ERR_ARG, "source and destination overlap");
  kore_sys* s = w->sys;
  Timed tm(s->ctx, KORE_TIMER_SPMV);
  HIP_TRY(launch_spmv(s->csr, ca_re, ca_im, cb_re, cb_im, wcol(w, src), s->n, wcol(w, dst), s->n, nc, 0.0,
                      s->ctx->stream));
  return KORE_OK;
}

int32_t kore_ws_dot(kore_ws* w, int32_t a0, int32_t na, int32_t b0, int32_t nb, double* H) {
  TRY(ws_check(w, a0, na));
  TRY(ws_check(w, b0, nb));
  if (!H) return fail(KORE_ERR_ARG, "H is NULL");
  kore_ctx* c = w->sys->ctx;
  cx<double>* res = w->part.as<cx<double>>() + (size_t)DOT_MAXCHUNK * w->ncol * MAXR;
  std::vector<zd> h((size_t)na * nb);
  for (int j0 = 0; j0 < nb; j0 += MAXR) {
    const int nbj = std::min(MAXR, nb - j0);
    TRY(ws_dot_dev(w, a0, na, b0 + j0, nbj, res));
    HIP_TRY(hipMemcpyAsync(h.data() + (size_t)j0 * na, res, (size_t)na * nbj * sizeof(zd), hipMemcpyDeviceToHost,
                           c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  memcpy(H, h.data(), h.size() * sizeof(zd));
  return KORE_OK;
}

int32_t kore_ws_gemm(kore_ws* w, int32_t d0, int32_t nd, int32_t s0, int32_t ns, const double* Q, double alpha,
                     double beta) {
  TRY(ws_check(w, d0, nd));
  TRY(ws_check(w, s0, ns));
  if (!Q) return fail(KORE_ERR_ARG, "Q is NULL");
  if (s0 < d0 + nd && d0 < s0 + ns) return fail(KORE_ERR_ARG, "source and destination overlap");
  if (nd == 0) return KORE_OK;
  cx<double>* q = w->qbuf.as<cx<double>>();
  TRY(w->up.put(q, Q, (size_t)ns * nd * sizeof(zd)));
  return ws_gemm_dev(w, d0, nd, s0, ns, q, ns, alpha, beta);
}

int32_t kore_ws_random(kore_ws* w, int32_t c0, int32_t nc, uint64_t seed) {
  TRY(ws_check(w, c0, nc));
  kore_sys* s = w->sys;
  HIP_TRY(launch_rand_cols(wcol(w, c0), s->n, s->n, nc, seed, s->ctx->stream));
  return KORE_OK;
}

int32_t kore_ws_copy(kore_ws* w, int32_t src, int32_t dst, int32_t nc) {
  TRY(ws_check(w, src, nc));
  TRY(ws_check(w, dst, nc));
  kore_sys* s = w->sys;
  HIP_TRY(launch_copy_cols(wcol(w, src), s->n, wcol(w, dst), s->n, s->n, nc, s->ctx->stream));
  return KORE_OK;
}

// Block Arnoldi expansion with everything but the bookkeeping on the device:
// per block, OP, two classical Gram-Schmidt passes and two Cholesky-QR
// passes (one for a single column).  The coefficients of block j travel to a pinned slot behind an
// event; block j+1 is queued before the host waits for block j, so the device
// never drains between blocks.  A breakdown found at block j discards block
// j+1 (everything it wrote lies in columns the recovery overwrites).
static int32_t ws_expand_finish(kore_ws* w, int j, int bs, int slot, double* H, int64_t ldh, bool* breakdown) {
  const bool one_pass = bs == 1;   // a single column needs one Cholesky-QR pass (R2 = 1)
  const int ncol = w->ncol;
  HIP_TRY(hipEventSynchronize(w->hev[slot]));
  const zd* h1 = reinterpret_cast<const zd*>(w->hh) + (size_t)slot * ws_hb_elems(ncol);
  const zd* h2 = h1 + (size_t)ncol * MAXR;
  const zd* R1 = h2 + (size_t)ncol * MAXR + MAXR * MAXR;
  const zd* R2 = R1 + 2 * MAXR * MAXR;
  const int* flag = reinterpret_cast<const int*>(h1 + (size_t)2 * ncol * MAXR + 5 * MAXR * MAXR);
  const int nb = j + bs;
  zd* Hz = reinterpret_cast<zd*>(H);
  for (int q = 0; q < bs; ++q)
    for (int i = 0; i < nb; ++i) Hz[(size_t)(j + q) * ldh + i] = h1[(size_t)q * nb + i] + h2[(size_t)q * nb + i];
  *breakdown = flag[0] || (!one_pass && flag[1]);
  if (*breakdown) return KORE_OK;
  if (one_pass) {
    Hz[(size_t)j * ldh + nb] = R1[0];
    return KORE_OK;
  }
  // R = R2 R1 (both upper triangular, ld bs)
  for (int q = 0; q < bs; ++q)
    for (int i = 0; i < bs; ++i) {
      zd acc = 0;
      for (int p = i; p <= q; ++p) acc += R2[(size_t)p * bs + i] * R1[(size_t)q * bs + p];
      Hz[(size_t)(j + q) * ldh + nb + i] = acc;
    }
  return KORE_OK;
}

int32_t kore_ws_stop_expand(kore_ws* w, int32_t j_end) {
  if (!w) return fail(KORE_ERR_ARG, "NULL workspace");
  w->stop_at.store(j_end < 0 ? INT32_MAX : j_end, std::memory_order_release);
  return KORE_OK;
}

int32_t kore_ws_expand(kore_ws* w, int32_t j0, int32_t j1, int32_t bs, int32_t cw, double ir_tol, int32_t ir_maxit,
                       int32_t ir_stop, double* H, int64_t ldh, int32_t* j_done, kore_solve_info* info) {
  if (!w || !H || !j_done) return fail(KORE_ERR_ARG, "NULL argument");
  *j_done = j0;
  kore_sys* s = w->sys;
  if (!s->has_b) return fail(KORE_ERR_STATE, "system has no B matrix");
  if (bs < 1 || bs > MAXR) return fail(KORE_ERR_ARG, "block size must be 1..%d", MAXR);
  if (j0 < 0 || j1 < j0 || (j1 - j0) % bs != 0)
    return fail(KORE_ERR_ARG, "expansion range [%d, %d) is not a whole number of blocks of %d", j0, j1, bs);
  TRY(ws_check(w, 0, j1 + bs));
  TRY(ws_check(w, cw, bs));
  if (cw < j1 + bs) return fail(KORE_ERR_ARG, "scratch block at %d overlaps the basis [0, %d)", cw, j1 + bs);
  if (ldh < j1 + bs) return fail(KORE_ERR_ARG, "ldh < %d", j1 + bs);
  kore_ctx* c = s->ctx;
  const int ncol = w->ncol;
  cx<double>* d_h1 = w->hb.as<cx<double>>();
  cx<double>* d_h2 = d_h1 + (size_t)ncol * MAXR;
  cx<double>* d_G = d_h2 + (size_t)ncol * MAXR;
  cx<double>* d_R1 = d_G + MAXR * MAXR;
  cx<double>* d_Ri1 = d_R1 + MAXR * MAXR;
  cx<double>* d_R2 = d_Ri1 + MAXR * MAXR;
  cx<double>* d_Ri2 = d_R2 + MAXR * MAXR;
  int* d_flag = reinterpret_cast<int*>(d_Ri2 + MAXR * MAXR);
  if (info) {
    info->iterations = 0;
    info->converged = 1;
    info->rel_correction = -1;
  }
  int pending = -1, pslot = 0;   // block queued on the device whose coefficients are not yet read
  bool bd = false;
  for (int j = j0, slot = 0; j < j1; j += bs, slot ^= 1) {
    if (j >= w->stop_at.load(std::memory_order_acquire)) break;   // the caller no longer needs the rest
    kore_solve_info li{};
    TRY(ws_apply_op_dev(w, j, cw, bs, ir_tol, ir_maxit, &li));
    if (info) {
      info->iterations = std::max(info->iterations, li.iterations);
      info->converged = info->converged && li.converged;
      info->rel_correction = std::max(info->rel_correction, li.rel_correction);
    }
    if (ir_stop > 0 && (li.iterations > ir_stop || !li.converged)) {
      if (pending >= 0) {
        TRY(ws_expand_finish(w, pending, bs, pslot, H, ldh, &bd));
        if (bd) {
          HIP_TRY(hipStreamSynchronize(c->stream));
          *j_done = pending;
          return KORE_WARN_BREAKDOWN;
        }
      }
      HIP_TRY(hipStreamSynchronize(c->stream));
      *j_done = j;
      return KORE_WARN_IR_NOT_CONVERGED;
    }
    const int nb = j + bs;
    const int64_t n = s->n;
    int nchunk = (int)std::min<int64_t>(DOT_MAXCHUNK, (n + DOT_ROWS - 1) / DOT_ROWS);
    if (nchunk < 1) nchunk = 1;
    const int64_t chunk = (n + nchunk - 1) / nchunk;
    cx<double>* part = w->part.as<cx<double>>();
    // two passes of classical Gram-Schmidt against V[:, 0:nb], then the Gram
    // partials of the block W = V[:, cw:cw+bs]
    TRY(ws_dot_dev(w, 0, nb, cw, bs, d_h1));
    if (bs <= 2 && nb <= ORTHO_STEP_MAXA && ortho_fused()) {
      // each update fused with the inner products that follow it: the basis
      // is read three times per block instead of five
      Timed tm(c, KORE_TIMER_ORTHO);
      HIP_TRY(launch_ortho_step(wcol(w, 0), n, nb, wcol(w, cw), n, bs, d_h1, nb, n, nchunk, chunk, 0, part,
                                c->stream));
      HIP_TRY(launch_sum_parts(reinterpret_cast<const double*>(part), nchunk, 2 * nb * bs,
                               reinterpret_cast<double*>(d_h2), c->stream));
      HIP_TRY(launch_ortho_step(wcol(w, 0), n, nb, wcol(w, cw), n, bs, d_h2, nb, n, nchunk, chunk, 1, part,
                                c->stream));
    } else {
      TRY(ws_gemm_dev(w, cw, bs, 0, nb, d_h1, nb, -1.0, 1.0));
      TRY(ws_dot_dev(w, 0, nb, cw, bs, d_h2));
      TRY(ws_gemm_dev(w, cw, bs, 0, nb, d_h2, nb, -1.0, 1.0));
      Timed tm(c, KORE_TIMER_ORTHO);
      HIP_TRY(launch_tsdot(wcol(w, cw), n, bs, wcol(w, cw), n, bs, n, nchunk, chunk, part, c->stream));
    }
    // Cholesky QR: V[:, nb:nb+bs] (R2 R1) = W, two passes for a block; a
    // single column is normalised once (its norm is already exact to
    // rounding, so a second pass would only return R2 = 1)
    if (bs == 1) {
      // one kernel sums the Gram partials and scales the column
      Timed tm(c, KORE_TIMER_ORTHO);
      HIP_TRY(launch_normalize1(reinterpret_cast<const double*>(part), nchunk, wcol(w, cw), wcol(w, nb), n, d_R1,
                                d_Ri1, d_flag, c->stream));
    } else {
      {
        Timed tm(c, KORE_TIMER_ORTHO);
        HIP_TRY(launch_sum_parts(reinterpret_cast<const double*>(part), nchunk, 2 * bs * bs,
                                 reinterpret_cast<double*>(d_G), c->stream));
      }
      HIP_TRY(launch_chol_inv(d_G, bs, bs, d_R1, d_Ri1, d_flag, c->stream));
      TRY(ws_gemm_dev(w, nb, bs, cw, bs, d_Ri1, bs, 1.0, 0.0));
    }
    if (bs > 1) {
      TRY(ws_dot_dev(w, nb, bs, nb, bs, d_G));
      HIP_TRY(launch_chol_inv(d_G, bs, bs, d_R2, d_Ri2, d_flag + 1, c->stream));
      TRY(ws_gemm_dev(w, cw, bs, nb, bs, d_Ri2, bs, 1.0, 0.0));
      HIP_TRY(launch_copy_cols(wcol(w, cw), s->n, wcol(w, nb), s->n, s->n, bs, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(w->hh + (size_t)slot * ws_hb_elems(ncol), d_h1, ws_hb_elems(ncol) * sizeof(zd),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(w->hev[slot], c->stream));
    if (pending >= 0) {
      TRY(ws_expand_finish(w, pending, bs, pslot, H, ldh, &bd));
      if (bd) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        *j_done = pending;
        return KORE_WARN_BREAKDOWN;
      }
      *j_done = pending + bs;
    }
    pending = j;
    pslot = slot;
  }
  if (pending >= 0) {
    TRY(ws_expand_finish(w, pending, bs, pslot, H, ldh, &bd));
    if (bd) {
      *j_done = pending;
      return KORE_WARN_BREAKDOWN;
    }
    *j_done = pending + bs;
  }
  return KORE_OK;
}

}  // extern "C"
