// Host side of liblowdin_fock.so: the LibintInterface_* C ABI of openLOWDIN
// (src/ints/Libint2Iface.cpp:3454-3589) implemented over HIP kernels.
//
// One handle per quantum species holds its shells (raw coefficients, as
// LibintInterface::add_shell receives them), the per-function norma, the
// significant shell-pair list, and the device-resident data of the Fock
// build: class-sorted shell pairs with precomputed primitive-pair records,
// the raw Schwarz bounds, and the wavefront task lists of every quartet
// class.  Those depend on the geometry only and are rebuilt when a shell is
// added or the pair list changes; every build uploads D, evaluates the
// screened unique quartets on the GPU and downloads G.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "gen/dispatch.h"
#include "lowdin_device.h"
#include "lowdin_fock.h"
#include "lowdin_kernels.h"

namespace lowdin {

static thread_local std::string g_last_error;

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  T* ensure(size_t n) {
    if (n > cap) {
      release();
      HIPCHK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
      cap = n;
    }
    return p;
  }
  void upload(const T* src, size_t n, hipStream_t s) {
    ensure(n);
    if (n) HIPCHK(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

// ------------------------------------------------------------ process state
constexpr int kSideStreams = 4;  // class launches overlap on these (default priority)

struct Context {
  bool ready = false;
  hipStream_t stream = nullptr;
  hipStream_t side[kSideStreams] = {};
  hipEvent_t fork = nullptr, join[kSideStreams] = {};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf<double> boys;
  std::mutex mu;  // one process stream: calls are serialised
};

static Context& ctx() {
  static Context c;
  return c;
}

// F_m(T_i) on the grid: power series for the highest order, then downward
// recursion (all terms positive, no cancellation).
static std::vector<double> make_boys_table() {
  std::vector<double> t((size_t)kBoysRows * kBoysCols);
  const int M = kBoysCols - 2;
  double F[kBoysCols];
  for (int i = 0; i < kBoysRows; ++i) {
    const double T = i * kBoysH;
    const double eT = std::exp(-T);
    double term = 1.0 / (2 * M + 1), sum = term;
    for (int k = 1; k < 4000; ++k) {
      term *= 2.0 * T / (2 * M + 2 * k + 1);
      sum += term;
      if (term < 1e-18 * sum) break;
    }
    F[M] = eT * sum;
    for (int m = M - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + eT) / (2 * m + 1);
    t[(size_t)i * kBoysCols] = eT;
    for (int m = 0; m <= M; ++m) t[(size_t)i * kBoysCols + 1 + m] = F[m];
  }
  return t;
}

static void ensure_context() {
  Context& c = ctx();
  if (c.ready) return;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (ndev < 1) throw std::runtime_error("no HIP device visible");
  HIPCHK(hipStreamCreate(&c.stream));  // default priority
  for (int k = 0; k < kSideStreams; ++k) {
    HIPCHK(hipStreamCreate(&c.side[k]));
    HIPCHK(hipEventCreateWithFlags(&c.join[k], hipEventDisableTiming));
  }
  HIPCHK(hipEventCreateWithFlags(&c.fork, hipEventDisableTiming));
  HIPCHK(hipEventCreate(&c.ev0));
  HIPCHK(hipEventCreate(&c.ev1));
  std::vector<double> tab = make_boys_table();
  c.boys.upload(tab.data(), tab.size(), c.stream);
  HIPCHK(hipStreamSynchronize(c.stream));
  c.ready = true;
}

// --------------------------------------------------------------- species
constexpr int kNumPairClasses = 6;  // ss, ps, pp, ds, dp, dd
static const int kPairClassL[kNumPairClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
static int pair_class(int la, int lb) {
  for (int c = 0; c < kNumPairClasses; ++c)
    if (kPairClassL[c][0] == la && kPairClassL[c][1] == lb) return c;
  return -1;
}
static inline int ncart_h(int l) { return (l + 1) * (l + 2) / 2; }

struct PairSet {
  bool valid = false;
  std::vector<PairInfo> pairs;      // grouped by pair class, canonical orientation
  int off[kNumPairClasses + 1] = {0};
  std::vector<double> prims;        // PrimPair records
  DevBuf<PairInfo> d_pairs;
  DevBuf<double> d_prims;
  bool q_ready = false;
};

struct TaskSet {
  bool valid = false;
  std::vector<Task> tasks;
  struct Launch {
    int la, lb, lc, ld;
    int cbra, cket;   // pair classes of bra / ket lists
    int bra_is_a;     // coupling: bra list belongs to the first species
    size_t off, count;
    int strip;        // intra: one workgroup per task with LDS row strips
  };
  std::vector<Launch> launches;
  std::vector<size_t> order;   // launch order: estimated cost, largest first
  DevBuf<Task> d_tasks;
};

struct Species {
  int stack_size = 0, species_id = 0;
  bool is_electron = false, parallel = true;
  std::vector<ShellInfo> shells;
  std::vector<double> pexp, pcoef;  // exponents; raw coefficient * N(alpha, l)
  std::vector<double> norma;
  std::vector<std::array<double, 4>> particles;  // z, x, y, z
  int nbasis = 0;
  uint64_t generation = 0;          // bumps whenever a shell is added

  std::vector<std::pair<int, int>> pairlist;  // (s1, s2 <= s1), sorted
  bool pairlist_built = false;

  uint64_t dev_generation = ~0ull;
  DevBuf<ShellInfo> d_shells;
  DevBuf<double> d_pexp, d_pcoef, d_norma;

  PairSet sig;        // significant pairs (intra builds), with Q
  PairSet all;        // every s1 <= s2 pair (coupling builds)
  TaskSet intra;
  uint64_t uid = 0;   // process-unique identity (keys the coupling task cache)
  struct CouplingKey {
    uint64_t gen_self, gen_other;
    TaskSet* tasks;
  };
  std::map<uint64_t, CouplingKey> coupling_tasks;  // keyed by the other species' uid

  DevBuf<double> d_D, d_Dn, d_Graw, d_G, d_Dother, d_R;
  DevBuf<unsigned long long> d_cnt;
  DevBuf<unsigned int> d_task_counts;
  lowdin_fock_stats stats{};
  // per class launch of the last intra build: device ms, quartets, primitive quartets
  std::vector<float> launch_ms;
  std::vector<unsigned long long> launch_cnt;
  std::vector<hipEvent_t> launch_ev;
  bool time_launches = false;
  bool serialize = false;   // all class launches on one side stream (clean per-class timing)

  ~Species() {
    for (auto& kv : coupling_tasks) delete kv.second.tasks;
    for (hipEvent_t e : launch_ev) (void)hipEventDestroy(e);
  }
};

// N(alpha, l) = sqrt(2^l (2 alpha)^(l+3/2) / (pi^(3/2) (2l-1)!!)), the
// normalisation of x^l exp(-alpha r^2) that libint2 embeds in raw
// coefficients (Libint2Iface.cpp:96-104 constructs the shells that way).
static double prim_norm(double alpha, int l) {
  double df = 1.0;
  for (int k = 2 * l - 1; k > 1; k -= 2) df *= k;
  return std::sqrt(std::pow(2.0, l) * std::pow(2.0 * alpha, l + 1.5) / (std::pow(M_PI, 1.5) * df));
}

static void ensure_device_basis(Species& s, hipStream_t st) {
  if (s.dev_generation == s.generation) return;
  s.d_shells.upload(s.shells.data(), s.shells.size(), st);
  s.d_pexp.upload(s.pexp.data(), s.pexp.size(), st);
  s.d_pcoef.upload(s.pcoef.data(), s.pcoef.size(), st);
  s.d_norma.upload(s.norma.data(), s.norma.size(), st);
  s.dev_generation = s.generation;
}

// Canonical pair records (higher-l shell first) grouped by pair class,
// sorted within a class by surviving primitive-pair count (descending) so
// that neighbouring lanes run equal primitive loops.
static void build_pairset(const Species& s, const std::vector<std::pair<int, int>>& list, PairSet& ps,
                          hipStream_t st) {
  struct Tmp {
    PairInfo info;
    std::vector<double> prim;
    size_t order;
  };
  std::vector<Tmp> tmp;
  tmp.reserve(list.size());
  const double kfac = std::sqrt(2.0) * std::pow(M_PI, 1.25);
  for (size_t k = 0; k < list.size(); ++k) {
    int sa = list[k].first, sb = list[k].second;
    if (s.shells[sa].l < s.shells[sb].l) std::swap(sa, sb);
    const ShellInfo& A = s.shells[sa];
    const ShellInfo& B = s.shells[sb];
    Tmp t;
    t.order = k;
    PairInfo& P = t.info;
    P.sa = sa;
    P.sb = sb;
    P.bfa = A.bf;
    P.bfb = B.bf;
    P.la = A.l;
    P.lb = B.l;
    P.Q = 0.0;
    P.ABx = A.x - B.x;
    P.ABy = A.y - B.y;
    P.ABz = A.z - B.z;
    const double ab2 = P.ABx * P.ABx + P.ABy * P.ABy + P.ABz * P.ABz;
    for (int i = 0; i < A.nprim; ++i) {
      const double a = s.pexp[A.poff + i], ca = s.pcoef[A.poff + i];
      for (int j = 0; j < B.nprim; ++j) {
        const double b = s.pexp[B.poff + j], cb = s.pcoef[B.poff + j];
        const double p = a + b, mu = a * b / p;
        if (mu * ab2 > kPrimPairDrop) continue;
        const double K = ca * cb * std::exp(-mu * ab2) * kfac / p;
        const double Px = (a * A.x + b * B.x) / p, Py = (a * A.y + b * B.y) / p, Pz = (a * A.z + b * B.z) / p;
        const double rec[kPrimPairStride] = {p, 0.5 / p, Px, Py, Pz, Px - A.x, Py - A.y, Pz - A.z, K, 0.0};
        t.prim.insert(t.prim.end(), rec, rec + kPrimPairStride);
      }
    }
    P.nprim = (int)(t.prim.size() / kPrimPairStride);
    tmp.push_back(std::move(t));
  }
  std::stable_sort(tmp.begin(), tmp.end(), [](const Tmp& x, const Tmp& y) {
    const int cx = pair_class(x.info.la, x.info.lb), cy = pair_class(y.info.la, y.info.lb);
    if (cx != cy) return cx < cy;
    if (x.info.nprim != y.info.nprim) return x.info.nprim > y.info.nprim;
    return x.order < y.order;
  });
  ps.pairs.clear();
  ps.prims.clear();
  for (int c = 0; c <= kNumPairClasses; ++c) ps.off[c] = 0;
  for (auto& t : tmp) {
    t.info.poff = (int)(ps.prims.size() / kPrimPairStride);
    ps.prims.insert(ps.prims.end(), t.prim.begin(), t.prim.end());
    ps.pairs.push_back(t.info);
    ps.off[pair_class(t.info.la, t.info.lb) + 1]++;
  }
  for (int c = 0; c < kNumPairClasses; ++c) ps.off[c + 1] += ps.off[c];
  ps.d_pairs.upload(ps.pairs.data(), ps.pairs.size(), st);
  ps.d_prims.upload(ps.prims.data(), ps.prims.size(), st);
  ps.valid = true;
  ps.q_ready = false;
}

static void compute_schwarz(PairSet& ps, hipStream_t st) {
  Context& c = ctx();
  for (int pc = 0; pc < kNumPairClasses; ++pc) {
    const int n = ps.off[pc + 1] - ps.off[pc];
    if (n == 0) continue;
    if (!dispatch_schwarz(kPairClassL[pc][0], kPairClassL[pc][1], ps.d_pairs.p + ps.off[pc], n, ps.d_prims.p,
                          c.boys.p, st))
      throw std::runtime_error("no Schwarz kernel for pair class");
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(ps.pairs.data(), ps.d_pairs.p, ps.pairs.size() * sizeof(PairInfo), hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipStreamSynchronize(st));
  ps.q_ready = true;
}

// Split the ket range [j0, j1) of one bra into wavefront tasks of at most
// 64 lanes that never straddle a change of the ket primitive count.
static void emit_tasks(std::vector<Task>& out, int bra, int j0, int j1, const PairInfo* kets, int flags) {
  int j = j0;
  while (j < j1) {
    int e = std::min(j + 64, j1);
    const int np = kets[j].nprim;
    int k = j + 1;
    while (k < e && kets[k].nprim == np) ++k;
    out.push_back(Task{bra, j, k, flags});
    j = k;
  }
}

static void sort_by_cost(std::vector<Task>& t, size_t from, const PairInfo* bras, const PairInfo* kets) {
  std::stable_sort(t.begin() + from, t.end(), [&](const Task& x, const Task& y) {
    const long cx = (long)bras[x.bra].nprim * kets[x.j0].nprim * (x.j1 - x.j0);
    const long cy = (long)bras[y.bra].nprim * kets[y.j0].nprim * (y.j1 - y.j0);
    return cx > cy;
  });
}

// Largest LDS row-strip footprint a J/K workgroup may use.
constexpr size_t kStripLdsMax = 65536;

// Intra-species tasks.  Where the bra's two row strips fit in LDS, a task is
// one workgroup: a bra pair and a run of up to `chunk` kets (chunk sized so
// that a class launch has at least ~2048 workgroups); otherwise a task is
// one wavefront of up to 64 kets with direct global accumulation.
static void build_intra_tasks(Species& s, hipStream_t st) {
  TaskSet& ts = s.intra;
  ts.tasks.clear();
  ts.launches.clear();
  const PairSet& ps = s.sig;
  for (int cx = 0; cx < kNumPairClasses; ++cx) {
    for (int cy = 0; cy <= cx; ++cy) {
      const int nx = ps.off[cx + 1] - ps.off[cx], ny = ps.off[cy + 1] - ps.off[cy];
      if (nx == 0 || ny == 0) continue;
      const PairInfo* bras = ps.pairs.data() + ps.off[cx];
      const PairInfo* kets = ps.pairs.data() + ps.off[cy];
      const size_t from = ts.tasks.size();
      const int la = kPairClassL[cx][0], lb = kPairClassL[cx][1];
      const bool strip = (size_t)(ncart_h(la) + ncart_h(lb)) * s.nbasis * sizeof(double) <= kStripLdsMax;
      if (strip) {
        long total = 0;
        for (int i = 0; i < nx; ++i) total += (cx == cy) ? i + 1 : ny;
        long chunk = (total + 2047) / 2048;
        chunk = std::max(256L, std::min(8192L, ((chunk + 255) / 256) * 256));
        for (int i = 0; i < nx; ++i) {
          const int nk = (cx == cy) ? i + 1 : ny;
          const int pieces = (int)((nk + chunk - 1) / chunk);
          for (int p = 0; p < pieces; ++p) {
            const int j0 = (int)((long)nk * p / pieces), j1 = (int)((long)nk * (p + 1) / pieces);
            if (j1 > j0) ts.tasks.push_back(Task{i, j0, j1, 1});
          }
        }
      } else {
        for (int i = 0; i < nx; ++i) emit_tasks(ts.tasks, i, 0, cx == cy ? i + 1 : ny, kets, 0);
      }
      sort_by_cost(ts.tasks, from, bras, kets);
      ts.launches.push_back({la, lb, kPairClassL[cy][0], kPairClassL[cy][1], cx, cy, 1, from,
                             ts.tasks.size() - from, strip ? 1 : 0});
    }
  }
  // estimated cost of each class launch: primitive quartets x class weight
  std::vector<double> cost(ts.launches.size(), 0.0);
  for (size_t k = 0; k < ts.launches.size(); ++k) {
    const TaskSet::Launch& L = ts.launches[k];
    const PairInfo* bras = ps.pairs.data() + ps.off[L.cbra];
    const PairInfo* kets = ps.pairs.data() + ps.off[L.cket];
    const double w = 1.0 + (L.la + L.lb + L.lc + L.ld) * (L.la + L.lb + L.lc + L.ld);
    for (size_t t = L.off; t < L.off + L.count; ++t) {
      const Task& tk = ts.tasks[t];
      cost[k] += w * (double)bras[tk.bra].nprim * kets[tk.j0].nprim * (tk.j1 - tk.j0);
    }
  }
  ts.order.resize(ts.launches.size());
  for (size_t k = 0; k < ts.order.size(); ++k) ts.order[k] = k;
  std::stable_sort(ts.order.begin(), ts.order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
  ts.d_tasks.upload(ts.tasks.data(), ts.tasks.size(), st);
  ts.valid = true;
}

static void ensure_sig(Species& s, hipStream_t st) {
  ensure_device_basis(s, st);
  if (!s.sig.valid) {
    build_pairset(s, s.pairlist, s.sig, st);
    compute_schwarz(s.sig, st);
    s.intra.valid = false;
  }
  if (!s.intra.valid) build_intra_tasks(s, st);
}

static void ensure_all(Species& s, hipStream_t st) {
  ensure_device_basis(s, st);
  if (!s.all.valid) {
    std::vector<std::pair<int, int>> list;
    const int ns = (int)s.shells.size();
    for (int s1 = 0; s1 < ns; ++s1)
      for (int s2 = s1; s2 < ns; ++s2) list.push_back({s1, s2});
    build_pairset(s, list, s.all, st);
  }
}

// Coupling quartets between species a and b: for every pair of pair
// classes the larger class is the bra (the kernels' VRR builds on the bra).
static TaskSet* coupling_taskset(Species& a, Species& b, hipStream_t st) {
  auto it = a.coupling_tasks.find(b.uid);
  if (it != a.coupling_tasks.end() && it->second.gen_self == a.generation && it->second.gen_other == b.generation &&
      it->second.tasks->valid)
    return it->second.tasks;
  TaskSet* ts = (it != a.coupling_tasks.end()) ? it->second.tasks : new TaskSet();
  a.coupling_tasks[b.uid] = {a.generation, b.generation, ts};
  ts->tasks.clear();
  ts->launches.clear();
  for (int ca = 0; ca < kNumPairClasses; ++ca)
    for (int cb = 0; cb < kNumPairClasses; ++cb) {
      const int na = a.all.off[ca + 1] - a.all.off[ca], nb = b.all.off[cb + 1] - b.all.off[cb];
      if (na == 0 || nb == 0) continue;
      const bool a_bra = ca >= cb;
      const PairSet& B = a_bra ? a.all : b.all;
      const PairSet& K = a_bra ? b.all : a.all;
      const int cbra = a_bra ? ca : cb, cket = a_bra ? cb : ca;
      const int nbra = a_bra ? na : nb, nket = a_bra ? nb : na;
      const PairInfo* bras = B.pairs.data() + B.off[cbra];
      const PairInfo* kets = K.pairs.data() + K.off[cket];
      const size_t from = ts->tasks.size();
      for (int i = 0; i < nbra; ++i) emit_tasks(ts->tasks, i, 0, nket, kets, 0);
      sort_by_cost(ts->tasks, from, bras, kets);
      ts->launches.push_back({kPairClassL[cbra][0], kPairClassL[cbra][1], kPairClassL[cket][0],
                              kPairClassL[cket][1], cbra, cket, a_bra ? 1 : 0, from, ts->tasks.size() - from, 0});
    }
  ts->d_tasks.upload(ts->tasks.data(), ts->tasks.size(), st);
  ts->valid = true;
  return ts;
}

void set_unsupported(const char* name) {
  g_last_error = std::string(name) + ": not provided by liblowdin_fock (outside the direct Fock path)";
  fprintf(stderr, "lowdin_fock: %s\n", g_last_error.c_str());
}

static void fill_nan(double* out, size_t n) {
  if (!out) return;
  const double nan = std::nan("");
  for (size_t k = 0; k < n; ++k) out[k] = nan;
}

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------ operations
// Enqueue one intra-species build on stream st: D (device, n x n) -> G
// (device, n x n).  Nothing is allocated once the species' buffers exist.
static void enqueue_2body(Species& s, const double* dD, double* dG, double factor, hipStream_t st) {
  Context& c = ctx();
  const int n = s.nbasis, ns = (int)s.shells.size();
  const size_t nn = (size_t)n * n;
  ensure_sig(s, st);
  const bool screen = s.is_electron;
  if (screen) {
    s.d_Dn.ensure((size_t)ns * ns);
    launch_dnorm(dD, n, s.d_shells.p, ns, s.d_Dn.p, st);
  }
  s.d_Graw.ensure(nn);
  HIPCHK(hipMemsetAsync(s.d_Graw.p, 0, nn * sizeof(double), st));
  const size_t nl = s.intra.launches.size();
  s.d_cnt.ensure(2 * std::max<size_t>(nl, 1));
  s.d_task_counts.ensure(2 * std::max<size_t>(s.intra.tasks.size(), 1));
  if (s.time_launches && s.launch_ev.size() < 2 * nl) {
    while (s.launch_ev.size() < 2 * nl) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      s.launch_ev.push_back(e);
    }
  }
  // Class launches fan out over the side streams, largest first, so that
  // small latency-bound classes overlap the large ones.
  HIPCHK(hipEventRecord(c.ev0, st));
  HIPCHK(hipEventRecord(c.fork, st));
  for (int q = 0; q < kSideStreams; ++q) HIPCHK(hipStreamWaitEvent(c.side[q], c.fork, 0));
  for (size_t r = 0; r < nl; ++r) {
    const size_t k = s.intra.order[r];
    const TaskSet::Launch& L = s.intra.launches[k];
    hipStream_t ls = c.side[s.serialize ? 0 : r % kSideStreams];
    JKArgs a{};
    a.tasks = s.intra.d_tasks.p + L.off;
    a.ntasks = (int)L.count;
    a.bra = s.sig.d_pairs.p + s.sig.off[L.cbra];
    a.ket = s.sig.d_pairs.p + s.sig.off[L.cket];
    a.prim = s.sig.d_prims.p;
    a.D = dD;
    a.Dn = s.d_Dn.p;
    a.norma = s.d_norma.p;
    a.boys = c.boys.p;
    a.Graw = s.d_Graw.p;
    a.counts = s.d_task_counts.p + 2 * L.off;
    a.n = n;
    a.nshell = ns;
    a.screen = screen ? 1 : 0;
    a.same_class = L.cbra == L.cket ? 1 : 0;
    a.eps = 2.220446049250313080847e-16;  // DBL_EPSILON, Libint2Iface.h:269-271
    a.half_factor = 0.5 * factor;
    if (s.time_launches) HIPCHK(hipEventRecord(s.launch_ev[2 * k], ls));
    const bool ok = L.strip ? dispatch_jk_strip(L.la, L.lb, L.lc, L.ld, a, ls) : dispatch_jk(L.la, L.lb, L.lc, L.ld, a, ls);
    if (!ok) throw std::runtime_error("no J/K kernel for class");
    if (s.time_launches) HIPCHK(hipEventRecord(s.launch_ev[2 * k + 1], ls));
  }
  for (int q = 0; q < kSideStreams; ++q) {
    HIPCHK(hipEventRecord(c.join[q], c.side[q]));
    HIPCHK(hipStreamWaitEvent(st, c.join[q], 0));
  }
  for (size_t k = 0; k < nl; ++k) {
    const TaskSet::Launch& L = s.intra.launches[k];
    launch_reduce_counts(s.d_task_counts.p + 2 * L.off, (int)L.count, s.d_cnt.p + 2 * k, st);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c.ev1, st));
  launch_symmetrize(s.d_Graw.p, n, 1.0, dG, st);
  HIPCHK(hipGetLastError());
}

// After the stream has drained: fold the counters and event times into stats.
static void collect_2body_stats(Species& s, hipStream_t st, double t0) {
  Context& c = ctx();
  const size_t nl = s.intra.launches.size();
  std::vector<unsigned long long> cnt(2 * std::max<size_t>(nl, 1));
  HIPCHK(hipMemcpyAsync(cnt.data(), s.d_cnt.p, cnt.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float kms = 0.f;
  HIPCHK(hipEventElapsedTime(&kms, c.ev0, c.ev1));
  lowdin_fock_stats stt{};
  s.launch_ms.assign(nl, 0.f);
  s.launch_cnt = cnt;
  for (size_t k = 0; k < nl; ++k) {
    const TaskSet::Launch& L = s.intra.launches[k];
    stt.quartets += cnt[2 * k];
    stt.prim_quartets += cnt[2 * k + 1];
    stt.eris += cnt[2 * k] * (uint64_t)(ncart_h(L.la) * ncart_h(L.lb) * ncart_h(L.lc) * ncart_h(L.ld));
    if (s.time_launches) HIPCHK(hipEventElapsedTime(&s.launch_ms[k], s.launch_ev[2 * k], s.launch_ev[2 * k + 1]));
  }
  stt.kernel_ms = kms;
  stt.total_ms = now_ms() - t0;
  s.stats = stt;
}

static void compute_2body(Species& s, const double* dens, double* result, double factor) {
  Context& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const double t0 = now_ms();
  ensure_context();
  hipStream_t st = c.stream;
  const size_t nn = (size_t)s.nbasis * s.nbasis;
  ensure_sig(s, st);
  s.d_D.upload(dens, nn, st);
  s.d_G.ensure(nn);
  enqueue_2body(s, s.d_D.p, s.d_G.p, factor, st);
  HIPCHK(hipMemcpyAsync(result, s.d_G.p, nn * sizeof(double), hipMemcpyDeviceToHost, st));
  collect_2body_stats(s, st, t0);
}

static void compute_coupling(Species& a, Species* b, const double* db, double* ra, const double* da, double* rb) {
  // ra (on a, from db) and/or rb (on b, from da)
  Context& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const double t0 = now_ms();
  ensure_context();
  hipStream_t st = c.stream;
  ensure_all(a, st);
  ensure_all(*b, st);
  TaskSet* ts = coupling_taskset(a, *b, st);
  const size_t na2 = (size_t)a.nbasis * a.nbasis, nb2 = (size_t)b->nbasis * b->nbasis;
  double* dRa = nullptr;
  double* dRb = nullptr;
  const double* dDa = nullptr;
  const double* dDb = nullptr;
  if (ra) {
    a.d_Dother.upload(db, nb2, st);
    dDb = a.d_Dother.p;
    dRa = a.d_R.ensure(na2);
    HIPCHK(hipMemsetAsync(dRa, 0, na2 * sizeof(double), st));
  }
  if (rb) {
    b->d_Dother.upload(da, na2, st);
    dDa = b->d_Dother.p;
    dRb = b->d_R.ensure(nb2);
    HIPCHK(hipMemsetAsync(dRb, 0, nb2 * sizeof(double), st));
  }
  HIPCHK(hipEventRecord(c.ev0, st));
  uint64_t quartets = 0, eris = 0, prims = 0;
  for (const TaskSet::Launch& L : ts->launches) {
    const bool a_bra = L.bra_is_a != 0;
    Species& sb = a_bra ? a : *b;
    Species& sk = a_bra ? *b : a;
    CouplingArgs g{};
    g.tasks = ts->d_tasks.p + L.off;
    g.ntasks = (int)L.count;
    g.bra = sb.all.d_pairs.p + sb.all.off[L.cbra];
    g.ket = sk.all.d_pairs.p + sk.all.off[L.cket];
    g.bprim = sb.all.d_prims.p;
    g.kprim = sk.all.d_prims.p;
    g.nbra_norma = sb.d_norma.p;
    g.nket_norma = sk.d_norma.p;
    g.nbra = sb.nbasis;
    g.nket = sk.nbasis;
    g.boys = c.boys.p;
    g.drop = 1.0e-10;  // Libint2Iface.cpp:1226-1228
    if (a_bra) {
      g.rbra = dRa;  g.dket = dDb;
      g.rket = dRb;  g.dbra = dDa;
    } else {
      g.rbra = dRb;  g.dket = dDa;
      g.rket = dRa;  g.dbra = dDb;
    }
    if (!dispatch_cp(L.la, L.lb, L.lc, L.ld, g, st)) throw std::runtime_error("no coupling kernel for class");
    for (size_t k = L.off; k < L.off + L.count; ++k) {
      const Task& t = ts->tasks[k];
      const int nq = t.j1 - t.j0;
      quartets += nq;
      eris += (uint64_t)nq * ncart_h(L.la) * ncart_h(L.lb) * ncart_h(L.lc) * ncart_h(L.ld);
      prims += (uint64_t)nq * sb.all.pairs[sb.all.off[L.cbra] + t.bra].nprim *
               sk.all.pairs[sk.all.off[L.cket] + t.j0].nprim;
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c.ev1, st));
  if (ra) HIPCHK(hipMemcpyAsync(ra, dRa, na2 * sizeof(double), hipMemcpyDeviceToHost, st));
  if (rb) HIPCHK(hipMemcpyAsync(rb, dRb, nb2 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float kms = 0.f;
  HIPCHK(hipEventElapsedTime(&kms, c.ev0, c.ev1));
  lowdin_fock_stats stt{quartets, eris, prims, kms, now_ms() - t0};
  a.stats = stt;
}

}  // namespace lowdin

// ================================================================ C ABI
using lowdin::Species;

extern "C" {

void* LibintInterface_new(int stack_size, int species_id, bool is_electron, bool parallel) {
  static std::atomic<uint64_t> next_uid{1};
  Species* s = new Species();
  s->uid = next_uid++;
  s->stack_size = stack_size;
  s->species_id = species_id;
  s->is_electron = is_electron;
  s->parallel = parallel;
  return s;
}

void LibintInterface_del(void* lint) { delete static_cast<Species*>(lint); }

void LibintInterface_add_particle(void* lint, int z, const double* center) {
  Species* s = static_cast<Species*>(lint);
  s->particles.push_back({(double)z, center[0], center[1], center[2]});
}

void LibintInterface_add_shell(void* lint, double* alpha, double* coeff, double* origin, int l, int nprim) {
  Species* s = static_cast<Species*>(lint);
  try {
    if (l < 0 || l > lowdin::kMaxL) throw std::runtime_error("angular momentum above d is not supported");
    if (nprim < 1) throw std::runtime_error("shell without primitives");
    lowdin::ShellInfo sh;
    sh.l = l;
    sh.bf = s->nbasis;
    sh.poff = (int)s->pexp.size();
    sh.nprim = nprim;
    sh.x = origin[0];
    sh.y = origin[1];
    sh.z = origin[2];
    for (int k = 0; k < nprim; ++k) {
      s->pexp.push_back(alpha[k]);
      s->pcoef.push_back(coeff[k] * lowdin::prim_norm(alpha[k], l));
    }
    s->shells.push_back(sh);
    // norma = 1 / sqrt(S_raw(i, i)) per Cartesian component
    double blk[36];
    lowdin::overlap_block(sh, sh, s->pexp.data(), s->pcoef.data(), blk);
    const int nc = lowdin::ncart_h(l);
    for (int i = 0; i < nc; ++i) s->norma.push_back(1.0 / std::sqrt(blk[i * nc + i]));
    s->nbasis += nc;
    s->generation++;
    s->sig.valid = false;
    s->all.valid = false;
    s->intra.valid = false;
    lowdin::g_last_error.clear();
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("add_shell: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
  }
}

int LibintInterface_get_nbasis(void* lint) { return static_cast<Species*>(lint)->nbasis; }

void LibintInterface_init_2body_ints(void* lint) {
  Species* s = static_cast<Species*>(lint);
  lowdin::Context& c = lowdin::ctx();
  try {
    std::lock_guard<std::mutex> lk(c.mu);
    lowdin::ensure_context();
    hipStream_t st = c.stream;
    lowdin::ensure_device_basis(*s, st);
    const int ns = (int)s->shells.size();
    const size_t npair = (size_t)ns * (ns + 1) / 2;
    lowdin::DevBuf<double> d_norms;
    d_norms.ensure(npair);
    lowdin::launch_pair_norms(s->d_shells.p, s->d_pexp.p, s->d_pcoef.p, ns, d_norms.p, st);
    HIPCHK(hipGetLastError());
    std::vector<double> norms(npair);
    if (npair)
      HIPCHK(hipMemcpyAsync(norms.data(), d_norms.p, npair * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<std::pair<int, int>> list;
    size_t k = 0;
    for (int s1 = 0; s1 < ns; ++s1)
      for (int s2 = 0; s2 <= s1; ++s2, ++k)
        if (norms[k] < 0.0 || norms[k] >= 1e-12) list.push_back({s1, s2});
    if (!s->pairlist_built || list != s->pairlist) {
      s->pairlist = std::move(list);
      s->sig.valid = false;
      s->intra.valid = false;
    }
    s->pairlist_built = true;
    lowdin::g_last_error.clear();
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("init_2body_ints: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
  }
}

void LibintInterface_compute_2body_direct(void* lint, double* dens, double* result, double* factor) {
  Species* s = static_cast<Species*>(lint);
  try {
    lowdin::compute_2body(*s, dens, result, *factor);
    lowdin::g_last_error.clear();
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("compute_2body_direct: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
    lowdin::fill_nan(result, (size_t)s->nbasis * s->nbasis);
  }
}

void LibintInterface_compute_coupling_direct(void* lint, void* olint, double* dens, double* result) {
  Species* s = static_cast<Species*>(lint);
  Species* o = static_cast<Species*>(olint);
  try {
    lowdin::compute_coupling(*s, o, dens, result, nullptr, nullptr);
    lowdin::g_last_error.clear();
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("compute_coupling_direct: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
    lowdin::fill_nan(result, (size_t)s->nbasis * s->nbasis);
  }
}

const char* lowdin_fock_last_error(void) { return lowdin::g_last_error.c_str(); }

int lowdin_fock_compute_2body_device(void* lint, const double* d_dens, double* d_result, double factor,
                                     void* stream, int synchronize) {
  Species* s = static_cast<Species*>(lint);
  lowdin::Context& c = lowdin::ctx();
  try {
    std::lock_guard<std::mutex> lk(c.mu);
    const double t0 = lowdin::now_ms();
    lowdin::ensure_context();
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
    lowdin::enqueue_2body(*s, d_dens, d_result, factor, st);
    if (synchronize) lowdin::collect_2body_stats(*s, st, t0);
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("compute_2body_device: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
    return -1;
  }
}

int lowdin_fock_time_launches(void* lint, int enable) {
  static_cast<Species*>(lint)->time_launches = enable != 0;
  return 0;
}

int lowdin_fock_serialize_launches(void* lint, int enable) {
  static_cast<Species*>(lint)->serialize = enable != 0;
  return 0;
}

int lowdin_fock_launch_profile(void* lint, int cap, int* classes, float* ms, uint64_t* quartets,
                               uint64_t* prim_quartets) {
  Species* s = static_cast<Species*>(lint);
  const int nl = (int)s->intra.launches.size();
  for (int k = 0; k < nl && k < cap; ++k) {
    const lowdin::TaskSet::Launch& L = s->intra.launches[k];
    classes[k] = L.la * 1000 + L.lb * 100 + L.lc * 10 + L.ld;
    ms[k] = k < (int)s->launch_ms.size() ? s->launch_ms[k] : 0.f;
    quartets[k] = 2 * k < (int)s->launch_cnt.size() ? s->launch_cnt[2 * k] : 0;
    prim_quartets[k] = 2 * k + 1 < (int)s->launch_cnt.size() ? s->launch_cnt[2 * k + 1] : 0;
  }
  return nl;
}

int lowdin_fock_get_stats(void* lint, lowdin_fock_stats* out) {
  if (!lint || !out) return -1;
  *out = static_cast<Species*>(lint)->stats;
  return 0;
}

int lowdin_fock_shellpair_list(void* lint, int* s1, int* s2, int cap) {
  Species* s = static_cast<Species*>(lint);
  if (!s->pairlist_built) return -1;
  const int n = (int)s->pairlist.size();
  for (int k = 0; k < n && k < cap; ++k) {
    s1[k] = s->pairlist[k].first;
    s2[k] = s->pairlist[k].second;
  }
  return n;
}

int lowdin_fock_schwarz(void* lint, double* Q) {
  Species* s = static_cast<Species*>(lint);
  lowdin::Context& c = lowdin::ctx();
  try {
    {
      std::lock_guard<std::mutex> lk(c.mu);
      lowdin::ensure_context();
      lowdin::ensure_sig(*s, c.stream);
    }
    const int ns = (int)s->shells.size();
    std::fill(Q, Q + (size_t)ns * ns, 0.0);
    for (const lowdin::PairInfo& p : s->sig.pairs) {
      Q[(size_t)p.sa * ns + p.sb] = p.Q;
      Q[(size_t)p.sb * ns + p.sa] = p.Q;
    }
    return 0;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("schwarz: ") + e.what();
    return -1;
  }
}

int lowdin_fock_coupling_pair(void* a, void* b, const double* da, const double* db, double* ra, double* rb) {
  Species* sa = static_cast<Species*>(a);
  Species* sb = static_cast<Species*>(b);
  try {
    lowdin::compute_coupling(*sa, sb, db, ra, da, rb);
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("coupling_pair: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
    lowdin::fill_nan(ra, (size_t)sa->nbasis * sa->nbasis);
    lowdin::fill_nan(rb, (size_t)sb->nbasis * sb->nbasis);
    return -1;
  }
}

}  // extern "C"
