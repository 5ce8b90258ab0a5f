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
// Kernels of the classes with total angular momentum >= 6 spill to a large
// private (scratch) segment (up to ~15 KB per lane for (dd|dd)).  The
// runtime provisions scratch per hardware queue for the device-wide wave
// count, so such kernels run on one stream only (side stream 0): J/K class
// launches, the Schwarz kernels and the coupling kernels alike.
constexpr int kHeavyL = 6;

struct Context {
  bool ready = false;
  hipStream_t stream = nullptr;
  hipStream_t side[kSideStreams] = {};
  hipEvent_t fork = nullptr, join[kSideStreams] = {};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf<double> boys;    // full table (one-body integrals)
  DevBuf<float> boysf;    // FP32 table of its own grid (mixed precision)
  std::mutex mu;  // one process stream: calls are serialised
};

static Context& ctx() {
  static Context c;
  return c;
}

// F_m(T) for m = 0..M at one point: power series for the highest order,
// then downward recursion (all terms positive, no cancellation).
static void boys_point(double T, int M, double* F) {
  const double eT = std::exp(-T);
  double term = 1.0 / (2 * M + 1), sum = term;
  for (int k = 1; k < 20000; ++k) {
    term *= 2.0 * T / (2 * M + 2 * k + 1);
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  F[M] = eT * sum;
  for (int m = M - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + eT) / (2 * m + 1);
}

// Full table (one-body integrals): per row T_i, e^{-T_i} then F_0..F_15.
static std::vector<double> make_boys_table() {
  std::vector<double> t((size_t)kBoysRows * kBoysCols);
  const int M = kBoysCols - 2;
  double F[kBoysCols];
  for (int i = 0; i < kBoysRows; ++i) {
    const double T = i * kBoysH;
    boys_point(T, M, F);
    t[(size_t)i * kBoysCols] = std::exp(-T);
    for (int m = 0; m <= M; ++m) t[(size_t)i * kBoysCols + 1 + m] = F[m];
  }
  return t;
}

// FP32 table of the mixed-precision quartets: rows T_i = i * kBoysF32H,
// e^{-T_i} then F_0..F_{kBoysF32Cols-2}, each rounded from FP64.
static std::vector<float> make_boys_table_f32() {
  std::vector<float> t((size_t)kBoysF32Rows * kBoysF32Cols);
  const int M = kBoysF32Cols - 2;
  double F[kBoysF32Cols];
  for (int i = 0; i < kBoysF32Rows; ++i) {
    const double T = i * (double)kBoysF32H;
    boys_point(T, M, F);
    t[(size_t)i * kBoysF32Cols] = (float)std::exp(-T);
    for (int m = 0; m <= M; ++m) t[(size_t)i * kBoysF32Cols + 1 + m] = (float)F[m];
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
  std::vector<float> tabf = make_boys_table_f32();
  if (tabf.size() != (size_t)kBoysF32Rows * kBoysF32Cols)
    throw std::runtime_error("FP32 Boys table size does not match its grid constants");
  c.boysf.upload(tabf.data(), tabf.size(), c.stream);
  HIPCHK(hipStreamSynchronize(c.stream));
  coop_attr_all();
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
  DevBuf<float> d_primsf;           // the same records in FP32
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
    int strip;        // intra: one workgroup per task with LDS row strips (1: 256 threads, 2: 64);
                      // 3: wave-cooperative kernel with row strips, 4: the same without strips
  };
  std::vector<Launch> launches;
  std::vector<size_t> order;   // launch order: estimated cost, largest first
  DevBuf<Task> d_tasks;
};

struct BuildPlan;

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
  uint64_t sig_gen = 0;   // bumps whenever sig is rebuilt
  BuildPlan* own = nullptr;  // the one-molecule build plan of this species
  uint64_t uid = 0;   // process-unique identity (keys the coupling task cache)
  struct CouplingKey {
    uint64_t gen_self, gen_other;
    TaskSet* tasks;
  };
  std::map<uint64_t, CouplingKey> coupling_tasks;  // keyed by the other species' uid

  DevBuf<double> d_D, d_G, d_Dother, d_R;
  TaskSet store_ts;                  // wavefront tasks of the integral store / transform
  uint64_t store_gen = ~0ull;        // sig_gen the store tasks were built from
  DevBuf<double> d_Dn1, d_cvec, d_store, d_Graw1, d_blk;
  lowdin_fock_stats stats{};
  int mixed = 0;            // mixed precision: FP32 quartets below tau
  int sig_mixed = 0;        // the precision mode the pair order of sig was chosen for
  double tau = 0.0;
  double screen_eps = 2.220446049250313080847e-16;   // quartet screen (DBL_EPSILON, Libint2Iface.h:269-271)
  bool setup_each_build = false;   // recompute pair records and Schwarz bounds before every build

  ~Species();
};

// The class launches of an intra-species build over one or many species
// (the molecules of a batch): one launch per quartet class, whose tasks carry
// the index of their molecule (Task.flags) into a device array of MolDesc.
// Densities and results are packed: molecule m's n_m x n_m block starts at
// off[m] elements.
constexpr int kNumClassPairs = 21;
struct BuildPlan {
  std::vector<Species*> mols;
  std::vector<uint64_t> sig_gen;        // each molecule's sig_gen when the tasks were built
  bool valid = false;
  TaskSet ts;
  std::vector<size_t> off, shoff;       // packed n^2 / nshell^2 offsets; last = total
  int nmax = 0, nshmax = 0;
  bool any_dn = false;                  // some molecule screens or runs mixed
  bool any_mixed = false;               // some molecule runs mixed precision
  std::vector<MolDesc> desc, desc_dev;  // wanted / on the device
  std::vector<SymDesc> sym, sym_dev;
  DevBuf<MolDesc> d_desc;
  DevBuf<SymDesc> d_sym;
  DevBuf<double> d_Graw, d_Dn;
  DevBuf<unsigned long long> d_cnt;
  DevBuf<unsigned int> d_task_counts;
  DevBuf<int> d_ranges;                 // per class launch: first task, task count
  uint64_t layout = 0;                  // bumps whenever tasks or buffers move
  int target_wg[kNumClassPairs];        // workgroups a strip launch aims at (autotuned)
  int kind[kNumClassPairs];             // 0: default, 1: LDS strips, 2: wavefront tasks,
                                        // 3: LDS strips with one-wavefront workgroups,
                                        // 4: wave-cooperative kernel with row strips (where LDS allows),
                                        // 5: wave-cooperative kernel without strips
  lowdin_fock_stats stats{};
  std::vector<float> launch_ms;         // per class launch of the last build
  std::vector<unsigned long long> launch_cnt;
  std::vector<hipEvent_t> launch_ev;
  bool time_launches = false;
  bool serialize = false;               // all class launches on one side stream (clean per-class timing)
  bool inline_launch = false;           // class launches on the caller's stream (no side streams)
  bool use_graph = false;
  double eps = 2.220446049250313080847e-16;   // quartet screen threshold of the plan's launches
  hipGraphExec_t graph = nullptr;
  uint64_t graph_layout = ~0ull;
  hipStream_t graph_stream = nullptr;
  bool graph_any_dn = false;
  bool graph_any_mixed = false;
  double graph_eps = 0.0;

  BuildPlan() {
    for (int k = 0; k < kNumClassPairs; ++k) {
      target_wg[k] = 512;
      kind[k] = 0;
    }
  }
  ~BuildPlan() {
    for (hipEvent_t e : launch_ev) (void)hipEventDestroy(e);
    if (graph) (void)hipGraphExecDestroy(graph);
  }
};

Species::~Species() {
  for (auto& kv : coupling_tasks) delete kv.second.tasks;
  delete own;
}

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
    P.Qn = 0.0;
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
  std::vector<float> pf(ps.prims.begin(), ps.prims.end());
  ps.d_primsf.upload(pf.data(), pf.size(), st);
  HIPCHK(hipStreamSynchronize(st));
  ps.valid = true;
  ps.q_ready = false;
}

// Morton (Z-order) key of a point: coordinates quantised on a 0.5 bohr grid
// from the box corner lo, 21 bits per axis interleaved.
static uint64_t morton_key(const double* P, const double* lo) {
  uint64_t key = 0;
  uint32_t c[3];
  for (int d = 0; d < 3; ++d) {
    const double q = std::max(0.0, (P[d] - lo[d]) * 2.0);
    c[d] = (uint32_t)std::min(q, 2097151.0);
  }
  for (int b = 20; b >= 0; --b)
    for (int d = 0; d < 3; ++d) key = (key << 1) | ((c[d] >> b) & 1u);
  return key;
}

// After Q is known: within each pair class order pairs by primitive count
// (descending), so that lanes of a wavefront run equal primitive loops, then
//   FP64 builds: along a Z-order curve of the pair centres (first primitive
//     pair's P), so that the 64 kets of a wavefront lie close together: for
//     a distant bra every lane then takes the Boys function's asymptotic
//     branch and the wavefront skips the table branch;
//   mixed precision (by_q): by decade of the normalised bound Qn
//     (descending), so that the FP32 (small-bound) quartets form contiguous
//     runs of kets, then along the Z-order curve within a decade.
// Rewrites the primitive-pair records in that order.
static void resort_by_q(PairSet& ps, hipStream_t st, bool by_q) {
  std::vector<size_t> idx(ps.pairs.size());
  for (size_t k = 0; k < idx.size(); ++k) idx[k] = k;
  double lo[3] = {1e300, 1e300, 1e300};
  std::vector<uint64_t> key(ps.pairs.size(), 0);
  for (const PairInfo& P : ps.pairs)
    if (P.nprim > 0)
      for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], ps.prims[(size_t)P.poff * kPrimPairStride + 2 + d]);
  for (size_t k = 0; k < ps.pairs.size(); ++k)
    if (ps.pairs[k].nprim > 0) key[k] = morton_key(&ps.prims[(size_t)ps.pairs[k].poff * kPrimPairStride + 2], lo);
  std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
    const PairInfo& a = ps.pairs[x];
    const PairInfo& b = ps.pairs[y];
    const int ca = pair_class(a.la, a.lb), cb = pair_class(b.la, b.lb);
    if (ca != cb) return ca < cb;
    if (a.nprim != b.nprim) return a.nprim > b.nprim;
    if (by_q) {
      const int da = (int)std::floor(std::log10(std::max(a.Qn, 1e-300)));
      const int db = (int)std::floor(std::log10(std::max(b.Qn, 1e-300)));
      if (da != db) return da > db;
    }
    return key[x] < key[y];
  });
  std::vector<PairInfo> pairs;
  std::vector<double> prims;
  pairs.reserve(ps.pairs.size());
  prims.reserve(ps.prims.size());
  for (size_t k : idx) {
    PairInfo P = ps.pairs[k];
    const size_t o = (size_t)P.poff * kPrimPairStride;
    P.poff = (int)(prims.size() / kPrimPairStride);
    prims.insert(prims.end(), ps.prims.begin() + o, ps.prims.begin() + o + (size_t)P.nprim * kPrimPairStride);
    pairs.push_back(P);
  }
  ps.pairs.swap(pairs);
  ps.prims.swap(prims);
  ps.d_pairs.upload(ps.pairs.data(), ps.pairs.size(), st);
  ps.d_prims.upload(ps.prims.data(), ps.prims.size(), st);
  std::vector<float> pf(ps.prims.begin(), ps.prims.end());
  ps.d_primsf.upload(pf.data(), pf.size(), st);
  HIPCHK(hipStreamSynchronize(st));
}

// Runs on the heavy-kernel stream (the (dd|dd) diagonal spills); the pair
// records were uploaded and their stream drained by build_pairset.
static void compute_schwarz(PairSet& ps, const double* d_norma, hipStream_t st) {
  Context& c = ctx();
  HIPCHK(hipStreamSynchronize(st));
  hipStream_t hs = c.side[0];
  for (int pc = 0; pc < kNumPairClasses; ++pc) {
    const int n = ps.off[pc + 1] - ps.off[pc];
    if (n == 0) continue;
    if (!dispatch_schwarz(kPairClassL[pc][0], kPairClassL[pc][1], ps.d_pairs.p + ps.off[pc], n, ps.d_prims.p,
                          c.boys.p, d_norma, hs))
      throw std::runtime_error("no Schwarz kernel for pair class");
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(ps.pairs.data(), ps.d_pairs.p, ps.pairs.size() * sizeof(PairInfo), hipMemcpyDeviceToHost,
                        hs));
  HIPCHK(hipStreamSynchronize(hs));
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
// Classes from this total angular momentum up run only the wave-cooperative
// kernel (the per-lane kernels would spill kilobytes per lane to scratch):
// by default without row strips (exchange updates as global atomics) and
// cut into about kCoopDefaultWg workgroups of four wavefronts, at least
// kCoopMinRun kets each; the autotuner may choose strips and another
// workgroup count.  A wavefront evaluates one quartet at a time, so a class
// of 10^5 quartets needs thousands of wavefronts to fill the 1024 SIMDs.
constexpr int kCoopDefaultL = 6;
constexpr int kCoopDefaultWg = 4096;
constexpr long kCoopMinRun = 4;

static inline int class_pair_id(int cx, int cy) { return cx * (cx + 1) / 2 + cy; }

static void ensure_sig(Species& s, hipStream_t st) {
  ensure_device_basis(s, st);
  if (!s.sig.valid) {
    build_pairset(s, s.pairlist, s.sig, st);
    compute_schwarz(s.sig, s.d_norma.p, st);
    resort_by_q(s.sig, st, s.mixed != 0);
    s.sig_mixed = s.mixed;
    s.sig_gen++;
  }
}

// Tasks of a plan, one class launch per (bra class, ket class <= bra class)
// present in any molecule.  Where the bra's two row strips fit in LDS (at
// the largest n of the plan) a task is one workgroup: a bra pair and a run of
// kets, the run length chosen so that the launch has about target_wg
// workgroups; otherwise a task is one wavefront of up to 64 kets with direct
// global accumulation.  Within a launch tasks run in order of estimated cost.
static void build_plan_tasks(BuildPlan& p, hipStream_t st) {
  TaskSet& ts = p.ts;
  ts.tasks.clear();
  ts.launches.clear();
  const int nm = (int)p.mols.size();
  p.off.assign(nm + 1, 0);
  p.shoff.assign(nm + 1, 0);
  p.nmax = p.nshmax = 0;
  p.any_dn = false;
  for (int m = 0; m < nm; ++m) {
    const Species& s = *p.mols[m];
    const int ns = (int)s.shells.size();
    p.off[m + 1] = p.off[m] + (size_t)s.nbasis * s.nbasis;
    p.shoff[m + 1] = p.shoff[m] + (size_t)ns * ns;
    p.nmax = std::max(p.nmax, s.nbasis);
    p.nshmax = std::max(p.nshmax, ns);
  }
  std::vector<double> cost;
  auto task_cost = [&](const Task& t, int cx, int cy) {
    const PairSet& ps = p.mols[t.flags]->sig;
    return (double)ps.pairs[ps.off[cx] + t.bra].nprim * ps.pairs[ps.off[cy] + t.j0].nprim * (t.j1 - t.j0);
  };
  for (int cx = 0; cx < kNumPairClasses; ++cx) {
    for (int cy = 0; cy <= cx; ++cy) {
      const int la = kPairClassL[cx][0], lb = kPairClassL[cx][1];
      const int lc = kPairClassL[cy][0], ld = kPairClassL[cy][1];
      long total = 0, nbras = 0;
      int maxbra = 0;   // largest bra primitive-pair count (staged in LDS by the strip kernel)
      for (int m = 0; m < nm; ++m) {
        const PairSet& ps = p.mols[m]->sig;
        const int nx = ps.off[cx + 1] - ps.off[cx], ny = ps.off[cy + 1] - ps.off[cy];
        if (nx && ny) {
          total += (cx == cy) ? (long)nx * (nx + 1) / 2 : (long)nx * ny;
          nbras += nx;
          for (int i = ps.off[cx]; i < ps.off[cx + 1]; ++i) maxbra = std::max(maxbra, ps.pairs[i].nprim);
        }
      }
      if (total == 0) continue;
      // LDS strips pay off when a bra has enough kets to fill a workgroup
      // (256 threads) for a while; short ket runs (small molecules) go to the
      // wavefront kernel, one bra x 64 kets per wave.
      int kd = p.kind[class_pair_id(cx, cy)];
      // wave-cooperative kernel: by default from total L = kCoopDefaultL up
      const bool coop_ok = has_jk_coop(la, lb, lc, ld);
      const bool coop_default = kd < 4 && la + lb + lc + ld >= kCoopDefaultL;
      const int coop = !coop_ok ? 0 : kd == 4 ? 3 : (kd == 5 || coop_default) ? 4 : 0;
      if (kd >= 4) kd = 0;
      const bool strip_ok = (size_t)(ncart_h(la) + ncart_h(lb)) * p.nmax * sizeof(double) <= kStripLdsMax &&
                            maxbra <= kMaxStripPrims;
      // untuned: strips wherever they fit, of 256 threads when a bra has on
      // average >= 192 kets, else of one wavefront (short ket runs of small
      // molecules would leave most of a 256-thread workgroup idle)
      const bool strip = coop || (strip_ok && (kd == 1 || kd == 3 || kd == 0));
      const bool wave64 = kd == 3 || (kd == 0 && total < 192 * nbras);
      const size_t from = ts.tasks.size();
      // run length: at least enough kets to amortise zeroing and flushing the
      // strips (which grow with n), at most 32768
      const long cmin = coop ? kCoopMinRun : std::max(64L, std::min(256L, (long)((p.nmax / 2 + 63) / 64) * 64));
      const long tgt = coop_default ? kCoopDefaultWg : p.target_wg[class_pair_id(cx, cy)];
      long chunk = (total + tgt - 1) / tgt;
      chunk = std::max(cmin, std::min(32768L, coop ? ((chunk + 3) / 4) * 4 : ((chunk + 63) / 64) * 64));
      for (int m = 0; m < nm; ++m) {
        const PairSet& ps = p.mols[m]->sig;
        const int nx = ps.off[cx + 1] - ps.off[cx], ny = ps.off[cy + 1] - ps.off[cy];
        if (nx == 0 || ny == 0) continue;
        const PairInfo* kets = ps.pairs.data() + ps.off[cy];
        for (int i = 0; i < nx; ++i) {
          const int nk = (cx == cy) ? i + 1 : ny;
          if (strip) {
            const int pieces = (int)((nk + chunk - 1) / chunk);
            for (int q = 0; q < pieces; ++q) {
              const int j0 = (int)((long)nk * q / pieces), j1 = (int)((long)nk * (q + 1) / pieces);
              if (j1 > j0) ts.tasks.push_back(Task{i, j0, j1, m});
            }
          } else {
            emit_tasks(ts.tasks, i, 0, nk, kets, m);
          }
        }
      }
      std::stable_sort(ts.tasks.begin() + from, ts.tasks.end(),
                       [&](const Task& x, const Task& y) { return task_cost(x, cx, cy) > task_cost(y, cx, cy); });
      ts.launches.push_back({la, lb, lc, ld, cx, cy, 1, from, ts.tasks.size() - from,
                             coop ? coop : (strip ? (wave64 ? 2 : 1) : 0)});
      // estimated cost of the launch: primitive quartets x class weight
      const double w = 1.0 + (la + lb + lc + ld) * (la + lb + lc + ld);
      double c = 0.0;
      for (size_t t = from; t < ts.tasks.size(); ++t) c += w * task_cost(ts.tasks[t], cx, cy);
      cost.push_back(c);
    }
  }
  ts.order.resize(ts.launches.size());
  for (size_t k = 0; k < ts.order.size(); ++k) ts.order[k] = k;
  std::stable_sort(ts.order.begin(), ts.order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
  ts.d_tasks.upload(ts.tasks.data(), ts.tasks.size(), st);
  const size_t nl = ts.launches.size();
  std::vector<int> ranges(2 * std::max<size_t>(nl, 1), 0);
  for (size_t k = 0; k < nl; ++k) {
    ranges[2 * k] = (int)ts.launches[k].off;
    ranges[2 * k + 1] = (int)ts.launches[k].count;
  }
  p.d_ranges.upload(ranges.data(), ranges.size(), st);
  p.d_cnt.ensure(3 * std::max<size_t>(nl, 1));
  p.d_task_counts.ensure(3 * std::max<size_t>(ts.tasks.size(), 1));
  p.d_Graw.ensure(std::max<size_t>(p.off[nm], 1));
  p.d_Dn.ensure(std::max<size_t>(p.shoff[nm], 1));
  p.d_desc.ensure(std::max(nm, 1));
  p.d_sym.ensure(std::max(nm, 1));
  p.desc_dev.clear();
  p.sym_dev.clear();
  p.sig_gen.resize(nm);
  for (int m = 0; m < nm; ++m) p.sig_gen[m] = p.mols[m]->sig_gen;
  HIPCHK(hipStreamSynchronize(st));
  p.layout++;
  ts.valid = true;
  p.valid = true;
}

static void ensure_plan(BuildPlan& p, hipStream_t st) {
  bool stale = !p.valid || p.sig_gen.size() != p.mols.size();
  for (size_t m = 0; m < p.mols.size(); ++m) {
    ensure_sig(*p.mols[m], st);
    if (!stale && p.sig_gen[m] != p.mols[m]->sig_gen) stale = true;
  }
  if (stale) build_plan_tasks(p, st);
}

static BuildPlan& own_plan(Species& s) {
  if (!s.own) {
    s.own = new BuildPlan();
    s.own->mols.push_back(&s);
  }
  return *s.own;
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
  std::vector<double> cost;
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
      // estimated cost: primitive quartets x class weight, as for the J/K launches
      const int lt = kPairClassL[cbra][0] + kPairClassL[cbra][1] + kPairClassL[cket][0] + kPairClassL[cket][1];
      double c = 0.0;
      for (size_t t = from; t < ts->tasks.size(); ++t) {
        const Task& k = ts->tasks[t];
        c += (double)bras[k.bra].nprim * kets[k.j0].nprim * (k.j1 - k.j0);
      }
      cost.push_back(c * (1.0 + lt * lt));
    }
  ts->order.resize(ts->launches.size());
  for (size_t k = 0; k < ts->order.size(); ++k) ts->order[k] = k;
  std::stable_sort(ts->order.begin(), ts->order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
  ts->d_tasks.upload(ts->tasks.data(), ts->tasks.size(), st);
  ts->valid = true;
  return ts;
}

void set_last_error(const std::string& msg) {
  g_last_error = msg;
  fprintf(stderr, "lowdin_fock: %s\n", g_last_error.c_str());
}
void clear_last_error() { g_last_error.clear(); }

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
static void ensure_launch_events(BuildPlan& p) {
  const size_t nl = p.ts.launches.size();
  while (p.launch_ev.size() < 2 * nl) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    p.launch_ev.push_back(e);
  }
}

// Per-lane kernels of the high-L classes spill to a large private segment;
// the wave-cooperative kernels keep their state in LDS.
static inline bool heavy_launch(const TaskSet::Launch& L) {
  return L.strip < 3 && L.la + L.lb + L.lc + L.ld >= kHeavyL;
}

// Class launches of one build on the side streams, forked from and joined
// back to st.  The scratch-heavy classes (few quartets, long per-lane
// latency) go first, all on side stream 0, so that they run beside the
// large classes instead of after them; the other classes follow largest
// first, round robin over the remaining side streams (all four when the
// plan has no heavy class).  which: 0 all, 1 all but the scratch-heavy
// ones, 2 only the scratch-heavy ones.
static void join_side_streams(hipStream_t st) {
  Context& c = ctx();
  for (int q = 0; q < kSideStreams; ++q) {
    HIPCHK(hipEventRecord(c.join[q], c.side[q]));
    HIPCHK(hipStreamWaitEvent(st, c.join[q], 0));
  }
}

// mols / counts: descriptor array and per-task counter array the kernels
// use (null: the plan's own; counts_off: no counters).
static void enqueue_classes(BuildPlan& p, hipStream_t st, bool timed, int which, bool join = true,
                            const MolDesc* mols = nullptr, bool counts_off = false) {
  Context& c = ctx();
  const size_t nl = p.ts.launches.size();
  bool any = false, any_heavy = false;
  for (size_t k = 0; k < nl; ++k) {
    any = any || which == 0 || (which == 1) != heavy_launch(p.ts.launches[k]);
    any_heavy = any_heavy || heavy_launch(p.ts.launches[k]);
  }
  if (!any) return;
  // inline plans launch on the caller's stream itself (several batched
  // SCF lanes then overlap without sharing the library's side streams)
  const bool fork = !p.inline_launch;
  if (fork) {
    HIPCHK(hipEventRecord(c.fork, st));
    for (int q = 0; q < kSideStreams; ++q) HIPCHK(hipStreamWaitEvent(c.side[q], c.fork, 0));
  }
  const int first_light = (any_heavy && which == 0) ? 1 : 0;
  size_t nlight = 0;
  for (size_t pass = 0; pass < 2 * nl; ++pass) {
    const size_t r = pass % nl;
    const size_t k = p.ts.order[r];
    const TaskSet::Launch& L = p.ts.launches[k];
    const bool heavy = heavy_launch(L);
    if ((pass < nl) != heavy) continue;   // first pass: heavy classes, second: the rest
    if ((which == 1 && heavy) || (which == 2 && !heavy)) continue;
    int q = 0;
    if (!p.serialize && !heavy) q = first_light + (int)(nlight++ % (size_t)(kSideStreams - first_light));
    hipStream_t ls = fork ? c.side[q] : st;
    JKArgs a{};
    a.tasks = p.ts.d_tasks.p + L.off;
    a.ntasks = (int)L.count;
    a.mols = mols ? mols : p.d_desc.p;
    a.boys = c.boys.p;
    a.boysf = c.boysf.p;
    a.counts = counts_off ? nullptr : p.d_task_counts.p + 3 * L.off;
    a.cbra = L.cbra;
    a.cket = L.cket;
    a.same_class = L.cbra == L.cket ? 1 : 0;
    a.lds_n = p.nmax;
    a.block = L.strip == 2 ? 64 : 256;
    if (L.strip == 4) a.lds_n = 0;
    a.eps = p.eps;  // DBL_EPSILON by default (Libint2Iface.h:269-271)
    a.any_mixed = p.any_mixed ? 1 : 0;
    if (timed) HIPCHK(hipEventRecord(p.launch_ev[2 * k], ls));
    const bool ok = L.strip >= 3 ? dispatch_jk_coop(L.la, L.lb, L.lc, L.ld, a, ls)
                    : L.strip    ? dispatch_jk_strip(L.la, L.lb, L.lc, L.ld, a, ls)
                                 : dispatch_jk(L.la, L.lb, L.lc, L.ld, a, ls);
    if (!ok) throw std::runtime_error("no J/K kernel for class");
    if (timed) HIPCHK(hipEventRecord(p.launch_ev[2 * k + 1], ls));
  }
  if (join && fork) join_side_streams(st);
}

// Dnorm and zeroed G_raw (no memset: capturable).
static void enqueue_prologue(BuildPlan& p, hipStream_t st) {
  const int nm = (int)p.mols.size();
  if (p.any_dn) launch_dnorm_batch(p.d_desc.p, nm, p.nshmax, st);
  launch_fill_zero(p.d_Graw.p, p.off[nm], st);
}

// Per-launch counters and G = 0.25 (G_raw + G_raw^T).
static void enqueue_epilogue(BuildPlan& p, hipStream_t st) {
  const int nm = (int)p.mols.size();
  launch_reduce_counts_multi(p.d_task_counts.p, p.d_ranges.p, (int)p.ts.launches.size(), p.d_cnt.p, st);
  launch_symmetrize_batch(p.d_sym.p, nm, p.nmax, st);
  HIPCHK(hipGetLastError());
}

// Enqueue one build of plan p on stream st: packed densities dD -> packed
// results dG, factors[m] the exchange factor of molecule m.  Nothing is
// allocated once the plan's buffers exist; the molecule descriptors are
// re-uploaded only when a pointer, factor or precision setting changed.
static void enqueue_plan(BuildPlan& p, const double* dD, double* dG, const double* factors, hipStream_t st,
                         const double* const* dD_each = nullptr, double* const* dG_each = nullptr,
                         bool record = true) {
  Context& c = ctx();
  ensure_plan(p, st);
  const int nm = (int)p.mols.size();
  p.desc.resize(nm);
  p.sym.resize(nm);
  p.any_dn = false;
  p.any_mixed = false;
  p.eps = nm > 0 ? p.mols[0]->screen_eps : p.eps;
  for (int m = 1; m < nm; ++m) p.eps = std::min(p.eps, p.mols[m]->screen_eps);
  for (int m = 0; m < nm; ++m) {
    Species& s = *p.mols[m];
    MolDesc& M = p.desc[m];
    std::memset(&M, 0, sizeof(M));
    M.pairs = s.sig.d_pairs.p;
    M.prim = s.sig.d_prims.p;
    M.primf = s.sig.d_primsf.p;
    M.D = dD_each ? dD_each[m] : dD + p.off[m];
    M.Dn = p.d_Dn.p + p.shoff[m];
    M.shells = s.d_shells.p;
    M.norma = s.d_norma.p;
    M.Graw = p.d_Graw.p + p.off[m];
    for (int k = 0; k <= kNumPairClasses; ++k) M.off[k] = s.sig.off[k];
    M.n = s.nbasis;
    M.nshell = (int)s.shells.size();
    M.screen = s.is_electron ? 1 : 0;
    M.mixed = s.mixed;
    M.half_factor = 0.5 * factors[m];
    M.tau = s.tau;
    p.any_dn = p.any_dn || M.screen || M.mixed;
    p.any_mixed = p.any_mixed || M.mixed;
    SymDesc& Y = p.sym[m];
    std::memset(&Y, 0, sizeof(Y));
    Y.Graw = M.Graw;
    Y.G = dG_each ? dG_each[m] : dG + p.off[m];
    Y.n = s.nbasis;
  }
  if (p.desc_dev.size() != p.desc.size() || std::memcmp(p.desc_dev.data(), p.desc.data(), nm * sizeof(MolDesc))) {
    HIPCHK(hipMemcpyAsync(p.d_desc.p, p.desc.data(), nm * sizeof(MolDesc), hipMemcpyHostToDevice, st));
    p.desc_dev = p.desc;
  }
  if (p.sym_dev.size() != p.sym.size() || std::memcmp(p.sym_dev.data(), p.sym.data(), nm * sizeof(SymDesc))) {
    HIPCHK(hipMemcpyAsync(p.d_sym.p, p.sym.data(), nm * sizeof(SymDesc), hipMemcpyHostToDevice, st));
    p.sym_dev = p.sym;
  }
  const bool timed = p.time_launches && !p.use_graph;
  if (timed) ensure_launch_events(p);
  if (record) HIPCHK(hipEventRecord(c.ev0, st));
  if (p.use_graph) {
    // The graph holds the classes without a large private segment.  Kernels
    // in a graph run on the runtime's own streams (one per branch), so the
    // scratch-heavy classes stay outside it, on side stream 0: they are
    // enqueued after the prologue and before the graph launch, and joined
    // with it before the epilogue, so that they run beside the graph.
    if (!p.graph || p.graph_layout != p.layout || p.graph_stream != st || p.graph_any_dn != p.any_dn ||
        p.graph_any_mixed != p.any_mixed ||
        p.graph_eps != p.eps) {
      if (p.graph) HIPCHK(hipGraphExecDestroy(p.graph));
      p.graph = nullptr;
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        enqueue_classes(p, st, false, 1);
      } catch (...) {
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIPCHK(hipStreamEndCapture(st, &g));
      const hipError_t e = hipGraphInstantiate(&p.graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHK(e);
      p.graph_layout = p.layout;
      p.graph_stream = st;
      p.graph_any_dn = p.any_dn;
      p.graph_any_mixed = p.any_mixed;
      p.graph_eps = p.eps;
    }
    enqueue_prologue(p, st);
    enqueue_classes(p, st, false, 2, false);
    HIPCHK(hipGraphLaunch(p.graph, st));
    join_side_streams(st);
    enqueue_epilogue(p, st);
  } else {
    enqueue_prologue(p, st);
    enqueue_classes(p, st, timed, 0);
    enqueue_epilogue(p, st);
  }
  if (record) HIPCHK(hipEventRecord(c.ev1, st));
}

// ------------------------------------------------ build without allocation
// lowdin_fock_compute_2body_device_noalloc keeps the scratch of a build --
// the molecule and symmetrisation descriptors, the shell-block norms of D
// and the raw accumulator G_raw -- in this file-scope device array instead
// of library allocations.  Every GPU holds its own copy until the process
// ends; calls that overlap on one GPU share it and race.
constexpr int kNoAllocMaxN = 768;        // basis functions
constexpr int kNoAllocMaxShells = 384;   // shells
constexpr size_t kNoAllocDescBytes = 512;
constexpr size_t kNoAllocScratchBytes =
    kNoAllocDescBytes + sizeof(double) * ((size_t)kNoAllocMaxShells * kNoAllocMaxShells +
                                          (size_t)kNoAllocMaxN * kNoAllocMaxN);   // 5.8 MB
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static_assert(sizeof(MolDesc) + sizeof(SymDesc) <= kNoAllocDescBytes, "descriptor slots");

// Device address of `scratch` on the current device (hipGetSymbolAddress
// once per device; the host-side name of the array is not a device address).
constexpr int kMaxDevices = 64;
static void* g_scratch_dev[kMaxDevices];
static char* scratch_ptr() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = ptr;
  }
  return static_cast<char*>(g_scratch_dev[dev]);
}

__global__ void store_desc_kernel(MolDesc* md, MolDesc m, SymDesc* sd, SymDesc y) {
  *md = m;
  *sd = y;
}

// The plan's tasks and every molecule's device data are current (what
// ensure_plan would otherwise rebuild, allocating).
static bool plan_ready(const BuildPlan& p) {
  if (!p.valid || !p.ts.valid || p.sig_gen.size() != p.mols.size()) return false;
  for (size_t m = 0; m < p.mols.size(); ++m) {
    const Species& s = *p.mols[m];
    if (!s.sig.valid || s.dev_generation != s.generation || p.sig_gen[m] != s.sig_gen) return false;
  }
  return true;
}

// One build of a single-species plan with its scratch in `scr` (the device
// address of `scratch`): descriptors stored by a kernel, Dnorm, G_raw
// zeroed by a fill kernel, the class launches (no counters, no events),
// symmetrisation into dG.  Nothing is allocated and nothing waits.
static void enqueue_noalloc(BuildPlan& p, Species& s, const double* dD, double* dG, double factor, hipStream_t st,
                            char* scr) {
  MolDesc* md = reinterpret_cast<MolDesc*>(scr);
  SymDesc* sd = reinterpret_cast<SymDesc*>(scr + 256);
  double* Dn = reinterpret_cast<double*>(scr + kNoAllocDescBytes);
  double* Graw = Dn + (size_t)kNoAllocMaxShells * kNoAllocMaxShells;
  MolDesc M;
  std::memset(&M, 0, sizeof(M));
  M.pairs = s.sig.d_pairs.p;
  M.prim = s.sig.d_prims.p;
  M.primf = s.sig.d_primsf.p;
  M.D = dD;
  M.Dn = Dn;
  M.shells = s.d_shells.p;
  M.norma = s.d_norma.p;
  M.Graw = Graw;
  for (int k = 0; k <= kNumPairClasses; ++k) M.off[k] = s.sig.off[k];
  M.n = s.nbasis;
  M.nshell = (int)s.shells.size();
  M.screen = s.is_electron ? 1 : 0;
  M.mixed = s.mixed;
  M.half_factor = 0.5 * factor;
  M.tau = s.tau;
  SymDesc Y;
  std::memset(&Y, 0, sizeof(Y));
  Y.Graw = Graw;
  Y.G = dG;
  Y.n = s.nbasis;
  p.eps = s.screen_eps;
  hipLaunchKernelGGL(store_desc_kernel, dim3(1), dim3(1), 0, st, md, M, sd, Y);
  if (M.screen || M.mixed) launch_dnorm_batch(md, 1, M.nshell, st);
  launch_fill_zero(Graw, (size_t)M.n * M.n, st);
  enqueue_classes(p, st, false, 0, true, md, true);
  launch_symmetrize_batch(sd, 1, M.n, st);
  HIPCHK(hipGetLastError());
}

// After the stream has drained: fold the counters and event times into stats.
static void collect_plan_stats(BuildPlan& p, hipStream_t st, double t0) {
  Context& c = ctx();
  const size_t nl = p.ts.launches.size();
  std::vector<unsigned long long> cnt(3 * std::max<size_t>(nl, 1));
  HIPCHK(hipMemcpyAsync(cnt.data(), p.d_cnt.p, cnt.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float kms = 0.f;
  HIPCHK(hipEventElapsedTime(&kms, c.ev0, c.ev1));
  lowdin_fock_stats stt{};
  const bool timed = p.time_launches && !p.use_graph;
  p.launch_ms.assign(nl, 0.f);
  p.launch_cnt = cnt;
  for (size_t k = 0; k < nl; ++k) {
    const TaskSet::Launch& L = p.ts.launches[k];
    const uint64_t ne = (uint64_t)(ncart_h(L.la) * ncart_h(L.lb) * ncart_h(L.lc) * ncart_h(L.ld));
    stt.quartets += cnt[3 * k];
    stt.prim_quartets += cnt[3 * k + 1];
    stt.quartets_fp32 += cnt[3 * k + 2];
    stt.eris += cnt[3 * k] * ne;
    stt.eris_fp32 += cnt[3 * k + 2] * ne;
    if (timed) HIPCHK(hipEventElapsedTime(&p.launch_ms[k], p.launch_ev[2 * k], p.launch_ev[2 * k + 1]));
  }
  stt.kernel_ms = kms;
  stt.total_ms = now_ms() - t0;
  p.stats = stt;
}

static void compute_2body(Species& s, const double* dens, double* result, double factor) {
  Context& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const double t0 = now_ms();
  ensure_context();
  hipStream_t st = c.stream;
  const size_t nn = (size_t)s.nbasis * s.nbasis;
  if (s.setup_each_build) s.sig.valid = false;
  BuildPlan& p = own_plan(s);
  s.d_D.upload(dens, nn, st);
  s.d_G.ensure(nn);
  enqueue_plan(p, s.d_D.p, s.d_G.p, &factor, st);
  HIPCHK(hipMemcpyAsync(result, s.d_G.p, nn * sizeof(double), hipMemcpyDeviceToHost, st));
  collect_plan_stats(p, st, t0);
  s.stats = p.stats;
}

// Enqueue the coupling quartets of species a and b on st: dRa (n_a^2) =
// coupling on a from dDb and/or dRb (n_b^2) = coupling on b from dDa (device
// pointers; either result may be null).  Results are zeroed here by a fill
// kernel; the classes run on the heavy-kernel stream, joined back to st.
static void enqueue_coupling(Species& a, Species& b, const double* dDb, double* dRa, const double* dDa, double* dRb,
                             hipStream_t st, lowdin_fock_stats* stt, bool inline_launch = false) {
  Context& c = ctx();
  ensure_all(a, st);
  ensure_all(b, st);
  TaskSet* ts = coupling_taskset(a, b, st);
  if (dRa) launch_fill_zero(dRa, (size_t)a.nbasis * a.nbasis, st);
  if (dRb) launch_fill_zero(dRb, (size_t)b.nbasis * b.nbasis, st);
  // class launches largest first, round robin over the side streams (the
  // cooperative kernels have no private segment, the per-lane ones at most
  // a few hundred bytes per lane); inline: all on the caller's stream
  if (!inline_launch) {
    HIPCHK(hipEventRecord(c.fork, st));
    for (int q = 0; q < kSideStreams; ++q) HIPCHK(hipStreamWaitEvent(c.side[q], c.fork, 0));
  }
  uint64_t quartets = 0, eris = 0, prims = 0;
  for (size_t r = 0; r < ts->launches.size(); ++r) {
    const TaskSet::Launch& L = ts->launches[ts->order[r]];
    hipStream_t hs = inline_launch ? st : c.side[r % kSideStreams];
    const bool a_bra = L.bra_is_a != 0;
    Species& sb = a_bra ? a : b;
    Species& sk = a_bra ? b : a;
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
    // from total L = kCoopDefaultL up one wavefront per quartet (no private
    // segment); below, one lane per quartet
    const bool coop = L.la + L.lb + L.lc + L.ld >= kCoopDefaultL && has_jk_coop(L.la, L.lb, L.lc, L.ld);
    if (!(coop ? dispatch_cp_coop(L.la, L.lb, L.lc, L.ld, g, hs) : dispatch_cp(L.la, L.lb, L.lc, L.ld, g, hs)))
      throw std::runtime_error("no coupling kernel for class");
    if (stt) {
      for (size_t k = L.off; k < L.off + L.count; ++k) {
        const Task& t = ts->tasks[k];
        const int nq = t.j1 - t.j0;
        quartets += nq;
        eris += (uint64_t)nq * ncart_h(L.la) * ncart_h(L.lb) * ncart_h(L.lc) * ncart_h(L.ld);
        prims += (uint64_t)nq * sb.all.pairs[sb.all.off[L.cbra] + t.bra].nprim *
                 sk.all.pairs[sk.all.off[L.cket] + t.j0].nprim;
      }
    }
  }
  HIPCHK(hipGetLastError());
  if (!inline_launch) join_side_streams(st);
  if (stt) {
    stt->quartets = quartets;
    stt->eris = eris;
    stt->prim_quartets = prims;
  }
}

static void compute_coupling(Species& a, Species* b, const double* db, double* ra, const double* da, double* rb) {
  // ra (on a, from db) and/or rb (on b, from da)
  Context& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const double t0 = now_ms();
  ensure_context();
  hipStream_t st = c.stream;
  const size_t na2 = (size_t)a.nbasis * a.nbasis, nb2 = (size_t)b->nbasis * b->nbasis;
  double* dRa = nullptr;
  double* dRb = nullptr;
  const double* dDa = nullptr;
  const double* dDb = nullptr;
  if (ra) {
    a.d_Dother.upload(db, nb2, st);
    dDb = a.d_Dother.p;
    dRa = a.d_R.ensure(na2);
  }
  if (rb) {
    b->d_Dother.upload(da, na2, st);
    dDa = b->d_Dother.p;
    dRb = b->d_R.ensure(nb2);
  }
  HIPCHK(hipEventRecord(c.ev0, st));
  lowdin_fock_stats stt{};
  enqueue_coupling(a, *b, dDb, dRa, dDa, dRb, st, &stt);
  HIPCHK(hipEventRecord(c.ev1, st));
  if (ra) HIPCHK(hipMemcpyAsync(ra, dRa, na2 * sizeof(double), hipMemcpyDeviceToHost, st));
  if (rb) HIPCHK(hipMemcpyAsync(rb, dRb, nb2 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float kms = 0.f;
  HIPCHK(hipEventElapsedTime(&kms, c.ev0, c.ev1));
  stt.kernel_ms = kms;
  stt.total_ms = now_ms() - t0;
  a.stats = stt;
}

// ------------------------------------------ integral store / first-quarter transform
// Wavefront tasks over the significant pairs of one species: the quartet
// set of the Fock build (bra class >= ket class, ket <= bra within a class).
static TaskSet& store_taskset(Species& s, hipStream_t st) {
  TaskSet& ts = s.store_ts;
  if (ts.valid && s.store_gen == s.sig_gen) return ts;
  ts.tasks.clear();
  ts.launches.clear();
  const PairSet& ps = s.sig;
  for (int cx = 0; cx < kNumPairClasses; ++cx)
    for (int cy = 0; cy <= cx; ++cy) {
      const int nx = ps.off[cx + 1] - ps.off[cx], ny = ps.off[cy + 1] - ps.off[cy];
      if (nx == 0 || ny == 0) continue;
      const PairInfo* kets = ps.pairs.data() + ps.off[cy];
      const size_t from = ts.tasks.size();
      for (int i = 0; i < nx; ++i) emit_tasks(ts.tasks, i, 0, cx == cy ? i + 1 : ny, kets, 0);
      sort_by_cost(ts.tasks, from, ps.pairs.data() + ps.off[cx], kets);
      ts.launches.push_back({kPairClassL[cx][0], kPairClassL[cx][1], kPairClassL[cy][0], kPairClassL[cy][1], cx, cy,
                             1, from, ts.tasks.size() - from, 0});
    }
  ts.d_tasks.upload(ts.tasks.data(), ts.tasks.size(), st);
  ts.valid = true;
  s.store_gen = s.sig_gen;
  return ts;
}

// Launch one store class; the classes whose raw block is staged in HBM
// (kStoreStagedNQ integrals per quartet) go in chunks of kStoreStagedTasks
// tasks over the species' staging buffer, one after another on hs.
template <typename Args, typename Dispatch>
static void launch_store_class(const TaskSet::Launch& L, Args g, Species& s, hipStream_t hs, Dispatch dispatch) {
  const int nq = ncart_h(L.la) * ncart_h(L.lb) * ncart_h(L.lc) * ncart_h(L.ld);
  if (nq < kStoreStagedNQ) {
    if (!dispatch(L.la, L.lb, L.lc, L.ld, g, hs)) throw std::runtime_error("no store kernel for class");
    return;
  }
  const Task* tasks = g.tasks;
  const int ntasks = g.ntasks;
  g.blk = s.d_blk.ensure((size_t)kStoreStagedTasks * 64 * nq);
  for (int t0 = 0; t0 < ntasks; t0 += kStoreStagedTasks) {
    g.tasks = tasks + t0;
    g.ntasks = std::min(kStoreStagedTasks, ntasks - t0);
    if (!dispatch(L.la, L.lb, L.lc, L.ld, g, hs)) throw std::runtime_error("no store kernel for class");
  }
}

// Candidate quartets / ERIs of a task range (before the density screen).
static void count_tasks(const TaskSet& ts, const TaskSet::Launch& L, const PairInfo* bras, const PairInfo* kets,
                        lowdin_fock_stats* stt) {
  if (!stt) return;
  const uint64_t per = (uint64_t)ncart_h(L.la) * ncart_h(L.lb) * ncart_h(L.lc) * ncart_h(L.ld);
  for (size_t k = L.off; k < L.off + L.count; ++k) {
    const Task& t = ts.tasks[k];
    const int nq = t.j1 - t.j0;
    stt->quartets += nq;
    stt->eris += nq * per;
    stt->prim_quartets += (uint64_t)nq * bras[t.bra].nprim * kets[t.j0].nprim;
  }
}

// Intra-species integral store (mode 0: dOut = A, n^4, only kept elements
// written) or first-quarter transform (mode 1: dOut = GG, n^3, zeroed,
// accumulated and mirrored here; dC = row p of the coefficients), enqueued
// on st from the device density dD (its block norms drive the screen).
static void enqueue_store(Species& s, const double* dD, const double* dC, double* dOut, int mode, hipStream_t st,
                          lowdin_fock_stats* stt) {
  Context& c = ctx();
  ensure_sig(s, st);
  TaskSet& ts = store_taskset(s, st);
  const int n = s.nbasis, ns = (int)s.shells.size();
  const bool screen = s.is_electron;
  if (screen) launch_dnorm(dD, n, s.d_shells.p, ns, s.d_Dn1.ensure((size_t)ns * ns), st);
  if (mode == 1) launch_fill_zero(dOut, (size_t)n * n * n, st);
  HIPCHK(hipEventRecord(c.fork, st));
  for (int q = 0; q < kSideStreams; ++q) HIPCHK(hipStreamWaitEvent(c.side[q], c.fork, 0));
  size_t nlight = 0;
  for (const TaskSet::Launch& L : ts.launches) {
    // scratch-heavy classes on side stream 0, the others round robin over 1..3
    hipStream_t hs = heavy_launch(L) ? c.side[0] : c.side[1 + nlight++ % (kSideStreams - 1)];
    StoreArgs g{};
    g.tasks = ts.d_tasks.p + L.off;
    g.ntasks = (int)L.count;
    g.bra = s.sig.d_pairs.p + s.sig.off[L.cbra];
    g.ket = s.sig.d_pairs.p + s.sig.off[L.cket];
    g.prim = s.sig.d_prims.p;
    g.norma = s.d_norma.p;
    g.Dn = s.d_Dn1.p;
    g.c = dC;
    g.out = dOut;
    g.boys = c.boys.p;
    g.n = n;
    g.nshell = ns;
    g.same_class = L.cbra == L.cket ? 1 : 0;
    g.screen = screen ? 1 : 0;
    g.mode = mode;
    g.eps = 2.220446049250313080847e-16;  // DBL_EPSILON, Libint2Iface.h:273-274, 316-317
    g.drop = 1.0e-10;                      // Libint2Iface.cpp:794, 2701
    launch_store_class(L, g, s, hs, dispatch_store);
    count_tasks(ts, L, s.sig.pairs.data() + s.sig.off[L.cbra], s.sig.pairs.data() + s.sig.off[L.cket], stt);
  }
  HIPCHK(hipGetLastError());
  join_side_streams(st);
  if (mode == 1) launch_mirror_upper(dOut, n, n, st);
  HIPCHK(hipGetLastError());
}

// Inter-species store / transform between the calling species a and the
// other species b (LibintInterface_compute_coupling_directAll/IT with
// lint = a, olint = b; permuted = a's species id > b's).
//   mode 0: A[x1][x2][y1][y2], X the lower species id, strides
//           (n_a n_b^2, n_b^2, n_b, 1) as the reference indexes it
//           (Libint2Iface.cpp:2748-2749, 2877-2880); with permuted and
//           n_a != n_b those indices leave the n_a^2 n_b^2 array, which is
//           refused here.
//   mode 1: GG[v][l][s] (n_a x n_b x n_b) = sum_u C(p,u) (uv|ls), u, v of a
//           (Libint2Iface.cpp:1318-1330, 1495-1545).
static void enqueue_coupling_store(Species& a, Species& b, const double* dC, double* dOut, int mode, hipStream_t st,
                                   lowdin_fock_stats* stt) {
  Context& c = ctx();
  const bool permuted = a.species_id > b.species_id;
  if (mode == 0 && permuted && a.nbasis != b.nbasis)
    throw std::runtime_error("coupling_directAll: called from the higher species id with a different basis size "
                             "(the reference layout indexes outside the n_a^2 n_b^2 result)");
  ensure_all(a, st);
  ensure_all(b, st);
  TaskSet* ts = coupling_taskset(a, b, st);
  const size_t na = a.nbasis, nb = b.nbasis;
  if (mode == 1) launch_fill_zero(dOut, na * nb * nb, st);
  HIPCHK(hipEventRecord(c.fork, st));
  for (int q = 0; q < kSideStreams; ++q) HIPCHK(hipStreamWaitEvent(c.side[q], c.fork, 0));
  size_t nlight = 0;
  for (const TaskSet::Launch& L : ts->launches) {
    hipStream_t hs = heavy_launch(L) ? c.side[0] : c.side[1 + nlight++ % (kSideStreams - 1)];
    const bool a_bra = L.bra_is_a != 0;
    Species& sb = a_bra ? a : b;
    Species& sk = a_bra ? b : a;
    CouplingStoreArgs g{};
    g.tasks = ts->d_tasks.p + L.off;
    g.ntasks = (int)L.count;
    g.bra = sb.all.d_pairs.p + sb.all.off[L.cbra];
    g.ket = sk.all.d_pairs.p + sk.all.off[L.cket];
    g.bprim = sb.all.d_prims.p;
    g.kprim = sk.all.d_prims.p;
    g.nbra_norma = sb.d_norma.p;
    g.nket_norma = sk.d_norma.p;
    g.c = dC;
    g.out = dOut;
    g.boys = c.boys.p;
    g.bra_is_x = (a_bra != permuted) ? 1 : 0;   // X = b when permuted
    g.bra_is_t = a_bra ? 1 : 0;
    g.mode = mode;
    g.n_o = (int)nb;
    g.sx1 = na * nb * nb;
    g.sx2 = nb * nb;
    g.sy1 = nb;
    g.drop = 1.0e-10;  // Libint2Iface.cpp:1456, 2865
    launch_store_class(L, g, a, hs, dispatch_cstore);
    count_tasks(*ts, L, sb.all.pairs.data() + sb.all.off[L.cbra], sk.all.pairs.data() + sk.all.off[L.cket], stt);
  }
  HIPCHK(hipGetLastError());
  join_side_streams(st);
  if (mode == 1) launch_mirror_upper(dOut, (int)na, (int)nb, st);
  HIPCHK(hipGetLastError());
}

// ------------------------------------------ Gaussian-geminal (G12) builds
static G12Params make_gem(const double* coefficients, const double* exponents, int pot_size) {
  if (pot_size < 0 || pot_size > kMaxGeminal)
    throw std::runtime_error("pot_size must be in 0.." + std::to_string(kMaxGeminal));
  if (pot_size > 0 && (!coefficients || !exponents)) throw std::runtime_error("null geminal parameters");
  G12Params g{};
  g.n = pot_size;
  for (int k = 0; k < pot_size; ++k) {
    g.c[k] = coefficients[k];
    g.g[k] = exponents[k];
    if (!(exponents[k] >= 0.0)) throw std::runtime_error("geminal exponents must be >= 0");
  }
  return g;
}

// dG (n x n) = 0.25 (Graw + Graw^T), Graw the six J/K updates of every
// quartet of the pair list with the geminal operator (compute_g12_direct,
// Libint2Iface.cpp:2193-2365: no density screen), enqueued on st.
static void enqueue_g12(Species& s, const double* dD, double* dG, double factor, const G12Params& gem,
                        hipStream_t st, lowdin_fock_stats* stt) {
  Context& c = ctx();
  ensure_sig(s, st);
  TaskSet& ts = store_taskset(s, st);
  const int n = s.nbasis;
  double* graw = s.d_Graw1.ensure(std::max<size_t>((size_t)n * n, 1));
  launch_fill_zero(graw, (size_t)n * n, st);
  HIPCHK(hipEventRecord(c.fork, st));
  for (int q = 0; q < kSideStreams; ++q) HIPCHK(hipStreamWaitEvent(c.side[q], c.fork, 0));
  size_t nlight = 0;
  for (const TaskSet::Launch& L : ts.launches) {
    // scratch-heavy classes on side stream 0, the others round robin over 1..3
    hipStream_t hs = heavy_launch(L) ? c.side[0] : c.side[1 + nlight++ % (kSideStreams - 1)];
    G12Args g{};
    g.tasks = ts.d_tasks.p + L.off;
    g.ntasks = (int)L.count;
    g.n = n;
    g.bra = s.sig.d_pairs.p + s.sig.off[L.cbra];
    g.ket = s.sig.d_pairs.p + s.sig.off[L.cket];
    g.prim = s.sig.d_prims.p;
    g.norma = s.d_norma.p;
    g.D = dD;
    g.Graw = graw;
    g.same_class = L.cbra == L.cket ? 1 : 0;
    g.half_factor = 0.5 * factor;
    g.gem = gem;
    if (!dispatch_g12(L.la, L.lb, L.lc, L.ld, g, hs)) throw std::runtime_error("no G12 kernel for class");
    count_tasks(ts, L, s.sig.pairs.data() + s.sig.off[L.cbra], s.sig.pairs.data() + s.sig.off[L.cket], stt);
  }
  HIPCHK(hipGetLastError());
  join_side_streams(st);
  launch_symmetrize(graw, n, 1.0, dG, st);
  HIPCHK(hipGetLastError());
}

// dRa (n_a x n_a) = G12 coupling on species a from b's density dDb
// (compute_g12inter_direct, Libint2Iface.cpp:2367-2550; the layout and
// element sets of the Coulomb coupling), zeroed here, enqueued on st.
static void enqueue_g12_coupling(Species& a, Species& b, const double* dDb, double* dRa, const G12Params& gem,
                                 hipStream_t st, lowdin_fock_stats* stt) {
  Context& c = ctx();
  ensure_all(a, st);
  ensure_all(b, st);
  TaskSet* ts = coupling_taskset(a, b, st);
  launch_fill_zero(dRa, (size_t)a.nbasis * a.nbasis, st);
  HIPCHK(hipEventRecord(c.fork, st));
  for (int q = 0; q < kSideStreams; ++q) HIPCHK(hipStreamWaitEvent(c.side[q], c.fork, 0));
  size_t nlight = 0;
  for (const TaskSet::Launch& L : ts->launches) {
    hipStream_t hs = heavy_launch(L) ? c.side[0] : c.side[1 + nlight++ % (kSideStreams - 1)];
    const bool a_bra = L.bra_is_a != 0;
    Species& sb = a_bra ? a : b;
    Species& sk = a_bra ? b : a;
    G12CouplingArgs ga{};
    CouplingArgs& g = ga.cp;
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
    g.drop = 1.0e-10;  // Libint2Iface.cpp:2496-2497
    if (a_bra) {
      g.rbra = dRa;
      g.dket = dDb;
    } else {
      g.rket = dRa;
      g.dbra = dDb;
    }
    ga.gem = gem;
    if (!dispatch_g12cp(L.la, L.lb, L.lc, L.ld, ga, hs)) throw std::runtime_error("no G12 coupling kernel for class");
    count_tasks(*ts, L, sb.all.pairs.data() + sb.all.off[L.cbra], sk.all.pairs.data() + sk.all.off[L.cket], stt);
  }
  HIPCHK(hipGetLastError());
  join_side_streams(st);
}

// Host-pointer G12 build (b null: intra-species on a with a's density;
// otherwise the coupling on a from b's density).
static void run_g12(Species& a, Species* b, const double* dens, double* result, double factor,
                    const G12Params& gem) {
  Context& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const double t0 = now_ms();
  ensure_context();
  hipStream_t st = c.stream;
  const size_t na2 = (size_t)a.nbasis * a.nbasis;
  const size_t nd2 = b ? (size_t)b->nbasis * b->nbasis : na2;
  a.d_D.upload(dens, nd2, st);
  double* dR = a.d_G.ensure(std::max<size_t>(na2, 1));
  lowdin_fock_stats stt{};
  HIPCHK(hipEventRecord(c.ev0, st));
  if (b)
    enqueue_g12_coupling(a, *b, a.d_D.p, dR, gem, st, &stt);
  else
    enqueue_g12(a, a.d_D.p, dR, factor, gem, st, &stt);
  HIPCHK(hipEventRecord(c.ev1, st));
  HIPCHK(hipMemcpyAsync(result, dR, na2 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float kms = 0.f;
  HIPCHK(hipEventElapsedTime(&kms, c.ev0, c.ev1));
  stt.kernel_ms = kms;
  stt.total_ms = now_ms() - t0;
  a.stats = stt;
}

// Row p of an n x n column-major coefficient matrix, C(p, u) = coeff[p + u n]
// (the Eigen copy of Libint2Iface.cpp:3535-3537, read as C(p, bf)).
static std::vector<double> coeff_row(const double* coeff, int n, int p) {
  if (p < 0 || p >= n) throw std::runtime_error("orbital index p out of range");
  std::vector<double> r(n);
  for (int u = 0; u < n; ++u) r[u] = coeff[p + (size_t)u * n];
  return r;
}

// Host-pointer store / transform: uploads, runs, downloads, records stats.
//   which 0: intra All, 1: intra IT, 2: coupling All, 3: coupling IT.
static void run_store(Species& a, Species* b, int which, const double* dens, const double* coeff, int p,
                      double* result) {
  // arguments are checked before the device is touched
  const int mode = which & 1;
  std::vector<double> row;
  if (mode == 1) row = coeff_row(coeff, a.nbasis, p);
  if (which >= 2 && !b) throw std::runtime_error("coupling variants need the other species");
  if (which == 2 && a.species_id > b->species_id && a.nbasis != b->nbasis)
    throw std::runtime_error("coupling_directAll: called from the higher species id with a different basis size "
                             "(the reference layout indexes outside the n_a^2 n_b^2 result)");
  Context& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const double t0 = now_ms();
  ensure_context();
  hipStream_t st = c.stream;
  const size_t na = a.nbasis, nb = b ? b->nbasis : 0;
  size_t count = 0;
  switch (which) {
    case 0: count = na * na * na * na; break;
    case 1: count = na * na * na; break;
    case 2: count = na * na * nb * nb; break;
    default: count = na * nb * nb; break;
  }
  double* dOut = a.d_store.ensure(count);
  // the store writes only the elements it keeps: the rest keep the caller's values
  if (mode == 0) a.d_store.upload(result, count, st);
  const double* dD = nullptr;
  if (which < 2) {
    a.d_D.upload(dens, na * na, st);
    dD = a.d_D.p;
  }
  const double* dC = nullptr;
  if (mode == 1) {
    a.d_cvec.upload(row.data(), na, st);
    HIPCHK(hipStreamSynchronize(st));   // row is a host buffer of this call
    dC = a.d_cvec.p;
  }
  lowdin_fock_stats stt{};
  HIPCHK(hipEventRecord(c.ev0, st));
  if (which < 2)
    enqueue_store(a, dD, dC, dOut, mode, st, &stt);
  else
    enqueue_coupling_store(a, *b, dC, dOut, mode, st, &stt);
  HIPCHK(hipEventRecord(c.ev1, st));
  HIPCHK(hipMemcpyAsync(result, dOut, count * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float kms = 0.f;
  HIPCHK(hipEventElapsedTime(&kms, c.ev0, c.ev1));
  stt.kernel_ms = kms;
  stt.total_ms = now_ms() - t0;
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

// Libint2Iface.cpp:3736-3750: every permutationally unique integral of the
// Fock build's quartet set, normalised, into results[i n^3 + j n^2 + k n + l]
// (all eight images; elements below the drop keep the caller's values).
void LibintInterface_compute_2body_directAll(void* lint, double* dens, double* results) {
  Species* s = static_cast<Species*>(lint);
  try {
    lowdin::run_store(*s, nullptr, 0, dens, nullptr, 0, results);
    lowdin::g_last_error.clear();
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("compute_2body_directAll: ") + e.what());
  }
}

// Libint2Iface.cpp:3528-3549: result[j n^2 + k n + l] = sum_i C(p,i) (ij|kl),
// C(p,i) = coeff[p + i n].
void LibintInterface_compute_2body_directIT(void* lint, double* dens, double* coeff, double* result, int* p) {
  Species* s = static_cast<Species*>(lint);
  try {
    lowdin::run_store(*s, nullptr, 1, dens, coeff, *p, result);
    lowdin::g_last_error.clear();
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("compute_2body_directIT: ") + e.what());
    const size_t n = s->nbasis;
    lowdin::fill_nan(result, n * n * n);
  }
}

// Libint2Iface.cpp:3752-3773 (dens is not read, as in the reference).
void LibintInterface_compute_coupling_directAll(void* lint, void* olint, double* dens, double* results) {
  (void)dens;
  Species* s = static_cast<Species*>(lint);
  try {
    lowdin::run_store(*s, static_cast<Species*>(olint), 2, nullptr, nullptr, 0, results);
    lowdin::g_last_error.clear();
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("compute_coupling_directAll: ") + e.what());
  }
}

// Libint2Iface.cpp:3591-3623: results[v n_o^2 + l n_o + s] = sum_u C(p,u)
// (uv|ls), u, v of lint's species (dens is not read, as in the reference).
void LibintInterface_compute_coupling_directIT(void* lint, void* olint, double* dens, double* coeff, double* results,
                                               int* p) {
  (void)dens;
  Species* s = static_cast<Species*>(lint);
  Species* o = static_cast<Species*>(olint);
  try {
    lowdin::run_store(*s, o, 3, nullptr, coeff, *p, results);
    lowdin::g_last_error.clear();
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("compute_coupling_directIT: ") + e.what());
    const size_t n = s->nbasis, on = o->nbasis;
    lowdin::fill_nan(results, n * on * on);
  }
}

// Libint2Iface.cpp:3687-3706: result = J[D] + factor K[D] with the
// contracted Gaussian geminal sum_k coefficients[k] exp(-exponents[k] r12^2).
void libintinterface_compute_g12_direct(void* lint, double* dens, double* result, double* factor,
                                        const double* coefficients, const double* exponents, const int pot_size) {
  Species* s = static_cast<Species*>(lint);
  try {
    const lowdin::G12Params gem = lowdin::make_gem(coefficients, exponents, pot_size);
    lowdin::run_g12(*s, nullptr, dens, result, *factor, gem);
    lowdin::g_last_error.clear();
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("compute_g12_direct: ") + e.what());
    lowdin::fill_nan(result, (size_t)s->nbasis * s->nbasis);
  }
}

// Libint2Iface.cpp:3708-3734: the G12 coupling on lint's species from the
// density of olint's species, in the layout of compute_coupling_direct.
void libintinterface_compute_g12inter_direct(void* lint, void* olint, double* dens, double* result,
                                             const double* coefficients, const double* exponents,
                                             const int pot_size) {
  Species* s = static_cast<Species*>(lint);
  try {
    const lowdin::G12Params gem = lowdin::make_gem(coefficients, exponents, pot_size);
    lowdin::run_g12(*s, static_cast<Species*>(olint), dens, result, 0.0, gem);
    lowdin::g_last_error.clear();
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("compute_g12inter_direct: ") + e.what());
    lowdin::fill_nan(result, (size_t)s->nbasis * s->nbasis);
  }
}

// Device-resident G12 build: olint NULL = intra-species (d_dens n x n of
// lint, factor), else the coupling on lint from olint's density d_dens.
int lowdin_fock_g12_device(void* lint, void* olint, const double* d_dens, double* d_result, double factor,
                           const double* coefficients, const double* exponents, int pot_size) {
  Species* s = static_cast<Species*>(lint);
  Species* o = static_cast<Species*>(olint);
  lowdin::Context& c = lowdin::ctx();
  try {
    const lowdin::G12Params gem = lowdin::make_gem(coefficients, exponents, pot_size);
    std::lock_guard<std::mutex> lk(c.mu);
    const double t0 = lowdin::now_ms();
    lowdin::ensure_context();
    hipStream_t st = c.stream;
    lowdin_fock_stats stt{};
    HIPCHK(hipEventRecord(c.ev0, st));
    if (o)
      lowdin::enqueue_g12_coupling(*s, *o, d_dens, d_result, gem, st, &stt);
    else
      lowdin::enqueue_g12(*s, d_dens, d_result, factor, gem, st, &stt);
    HIPCHK(hipEventRecord(c.ev1, st));
    HIPCHK(hipStreamSynchronize(st));
    float kms = 0.f;
    HIPCHK(hipEventElapsedTime(&kms, c.ev0, c.ev1));
    stt.kernel_ms = kms;
    stt.total_ms = lowdin::now_ms() - t0;
    s->stats = stt;
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("g12_device: ") + e.what());
    return -1;
  }
}

// Device-resident forms of the four calls above (device pointers, work on
// the library's stream, synchronised before return).
int lowdin_fock_store_device(void* lint, void* olint, int which, const double* d_dens, const double* d_crow,
                             double* d_out) {
  Species* s = static_cast<Species*>(lint);
  Species* o = static_cast<Species*>(olint);
  lowdin::Context& c = lowdin::ctx();
  try {
    if (which < 0 || which > 3) throw std::runtime_error("which must be 0..3");
    if (which >= 2 && !o) throw std::runtime_error("coupling variants need the other species");
    if ((which & 1) && !d_crow) throw std::runtime_error("transform needs the coefficient row");
    if (which < 2 && s->is_electron && !d_dens) throw std::runtime_error("screened store needs the density");
    std::lock_guard<std::mutex> lk(c.mu);
    const double t0 = lowdin::now_ms();
    lowdin::ensure_context();
    hipStream_t st = c.stream;
    lowdin_fock_stats stt{};
    HIPCHK(hipEventRecord(c.ev0, st));
    if (which < 2)
      lowdin::enqueue_store(*s, d_dens, d_crow, d_out, which & 1, st, &stt);
    else
      lowdin::enqueue_coupling_store(*s, *o, d_crow, d_out, which & 1, st, &stt);
    HIPCHK(hipEventRecord(c.ev1, st));
    HIPCHK(hipStreamSynchronize(st));
    float kms = 0.f;
    HIPCHK(hipEventElapsedTime(&kms, c.ev0, c.ev1));
    stt.kernel_ms = kms;
    stt.total_ms = lowdin::now_ms() - t0;
    s->stats = stt;
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("store_device: ") + e.what());
    return -1;
  }
}

const char* lowdin_fock_last_error(void) { return lowdin::g_last_error.c_str(); }

int lowdin_fock_1body(void* lint, int kind, int ncharge, const double* Z, const double* xyz, double* result) {
  Species* s = static_cast<Species*>(lint);
  lowdin::Context& c = lowdin::ctx();
  try {
    if (kind < 1 || kind > 3) throw std::runtime_error("kind must be 1 (overlap), 2 (kinetic) or 3 (attraction)");
    std::lock_guard<std::mutex> lk(c.mu);
    lowdin::ensure_context();
    hipStream_t st = c.stream;
    lowdin::ensure_device_basis(*s, st);
    const int n = s->nbasis, ns = (int)s->shells.size();
    lowdin::DevBuf<double> dM, dZ, dX;
    dM.ensure((size_t)n * n);
    const int nc = kind == 3 ? ncharge : 0;
    if (nc > 0) {
      dZ.upload(Z, nc, st);
      dX.upload(xyz, 3 * (size_t)nc, st);
    }
    lowdin::launch_onebody(kind, s->d_shells.p, s->d_pexp.p, s->d_pcoef.p, s->d_norma.p, ns, n, nc, dZ.p, dX.p,
                           c.boys.p, dM.p, st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(result, dM.p, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("1body: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
    return -1;
  }
}

// Libint2Iface.cpp:3482-3503: 1 overlap, 2 kinetic, 3 attraction to the
// particles added with add_particle (charge z at their centres).
void LibintInterface_compute_1body_ints(void* lint, int integral_kind, double* result) {
  Species* s = static_cast<Species*>(lint);
  std::vector<double> Z, X;
  for (const auto& p : s->particles) {
    Z.push_back(p[0]);
    X.push_back(p[1]);
    X.push_back(p[2]);
    X.push_back(p[3]);
  }
  if (lowdin_fock_1body(lint, integral_kind, (int)Z.size(), Z.data(), X.data(), result) != 0)
    lowdin::fill_nan(result, (size_t)s->nbasis * s->nbasis);
}

int lowdin_fock_compute_2body_device(void* lint, const double* d_dens, double* d_result, double factor,
                                     void* stream, int synchronize) {
  Species* s = static_cast<Species*>(lint);
  lowdin::Context& c = lowdin::ctx();
  try {
    std::lock_guard<std::mutex> lk(c.mu);
    const double t0 = lowdin::now_ms();
    lowdin::ensure_context();
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
    if (s->setup_each_build) s->sig.valid = false;
    lowdin::BuildPlan& p = lowdin::own_plan(*s);
    lowdin::enqueue_plan(p, d_dens, d_result, &factor, st);
    if (synchronize) {
      lowdin::collect_plan_stats(p, st, t0);
      s->stats = p.stats;
    }
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("compute_2body_device: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
    return -1;
  }
}

int lowdin_fock_compute_2body_device_noalloc(void* lint, const double* d_dens, double* d_result, double factor,
                                             void* stream) {
  Species* s = static_cast<Species*>(lint);
  lowdin::Context& c = lowdin::ctx();
  try {
    std::lock_guard<std::mutex> lk(c.mu);
    const int n = s->nbasis, ns = (int)s->shells.size();
    if (n > lowdin::kNoAllocMaxN || ns > lowdin::kNoAllocMaxShells)
      throw std::runtime_error("basis of " + std::to_string(n) + " functions / " + std::to_string(ns) +
                               " shells exceeds the no-allocation scratch (at most " +
                               std::to_string(lowdin::kNoAllocMaxN) + " functions and " +
                               std::to_string(lowdin::kNoAllocMaxShells) + " shells)");
    if (!c.ready || !s->own || !lowdin::plan_ready(*s->own))
      throw std::runtime_error("the species has no current build plan: call LibintInterface_compute_2body_direct or "
                               "lowdin_fock_compute_2body_device on it once first (that call prepares the geometry "
                               "data; this entry point allocates nothing)");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
    lowdin::enqueue_noalloc(*s->own, *s, d_dens, d_result, factor, st, lowdin::scratch_ptr());
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("compute_2body_device_noalloc: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
    return -1;
  }
}

int lowdin_fock_time_launches(void* lint, int enable) {
  lowdin::own_plan(*static_cast<Species*>(lint)).time_launches = enable != 0;
  return 0;
}

int lowdin_fock_set_precision(void* lint, int mode, double tau) {
  Species* s = static_cast<Species*>(lint);
  if (mode != 0 && mode != 1) {
    lowdin::g_last_error = "set_precision: mode must be 0 (FP64) or 1 (mixed)";
    return -1;
  }
  s->mixed = mode;
  s->tau = tau;
  // the pair order depends on the mode: re-sort on the next build
  if (s->sig.valid && s->sig_mixed != mode) s->sig.valid = false;
  lowdin::g_last_error.clear();
  return 0;
}

int lowdin_fock_set_screen(void* lint, double eps) {
  Species* s = static_cast<Species*>(lint);
  if (!(eps >= 0.0)) {
    lowdin::g_last_error = "set_screen: the threshold must be >= 0";
    return -1;
  }
  s->screen_eps = eps;
  lowdin::g_last_error.clear();
  return 0;
}

int lowdin_fock_set_setup_each_build(void* lint, int enable) {
  static_cast<Species*>(lint)->setup_each_build = enable != 0;
  return 0;
}

int lowdin_fock_serialize_launches(void* lint, int enable) {
  lowdin::own_plan(*static_cast<Species*>(lint)).serialize = enable != 0;
  return 0;
}

static int plan_profile(const lowdin::BuildPlan& p, int cap, int* classes, float* ms, uint64_t* quartets,
                        uint64_t* prim_quartets) {
  const int nl = (int)p.ts.launches.size();
  for (int k = 0; k < nl && k < cap; ++k) {
    const lowdin::TaskSet::Launch& L = p.ts.launches[k];
    classes[k] = L.la * 1000 + L.lb * 100 + L.lc * 10 + L.ld;
    ms[k] = k < (int)p.launch_ms.size() ? p.launch_ms[k] : 0.f;
    quartets[k] = 3 * k < (int)p.launch_cnt.size() ? p.launch_cnt[3 * k] : 0;
    prim_quartets[k] = 3 * k + 1 < (int)p.launch_cnt.size() ? p.launch_cnt[3 * k + 1] : 0;
  }
  return nl;
}

int lowdin_fock_launch_profile(void* lint, int cap, int* classes, float* ms, uint64_t* quartets,
                               uint64_t* prim_quartets) {
  return plan_profile(lowdin::own_plan(*static_cast<Species*>(lint)), cap, classes, ms, quartets, prim_quartets);
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

int lowdin_fock_coupling_pair_device(void* a, void* b, const double* d_da, const double* d_db, double* d_ra,
                                     double* d_rb, void* stream, int synchronize) {
  Species* sa = static_cast<Species*>(a);
  Species* sb = static_cast<Species*>(b);
  lowdin::Context& c = lowdin::ctx();
  try {
    std::lock_guard<std::mutex> lk(c.mu);
    lowdin::ensure_context();
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
    lowdin::enqueue_coupling(*sa, *sb, d_db, d_ra, d_da, d_rb, st, nullptr);
    if (synchronize) HIPCHK(hipStreamSynchronize(st));
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("coupling_pair_device: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
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

// ---------------------------------------------------------------- batches
struct lowdin_fock_batch_t {
  lowdin::BuildPlan plan;
};

void* lowdin_fock_batch_new(int nmol, void* const* lints) {
  try {
    if (nmol < 1 || !lints) throw std::runtime_error("a batch needs at least one species handle");
    auto* b = new lowdin_fock_batch_t();
    for (int m = 0; m < nmol; ++m) {
      if (!lints[m]) {
        delete b;
        throw std::runtime_error("null species handle");
      }
      b->plan.mols.push_back(static_cast<Species*>(lints[m]));
    }
    lowdin::g_last_error.clear();
    return b;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("batch_new: ") + e.what();
    return nullptr;
  }
}

void lowdin_fock_batch_del(void* batch) { delete static_cast<lowdin_fock_batch_t*>(batch); }

int64_t lowdin_fock_batch_layout(void* batch, int64_t* offsets) {
  const lowdin::BuildPlan& p = static_cast<lowdin_fock_batch_t*>(batch)->plan;
  int64_t o = 0;
  for (size_t m = 0; m < p.mols.size(); ++m) {
    if (offsets) offsets[m] = o;
    o += (int64_t)p.mols[m]->nbasis * p.mols[m]->nbasis;
  }
  if (offsets) offsets[p.mols.size()] = o;
  return o;
}

int lowdin_fock_batch_compute_2body_device(void* batch, const double* d_dens, double* d_result,
                                           const double* factors, void* stream, int synchronize) {
  lowdin::BuildPlan& p = static_cast<lowdin_fock_batch_t*>(batch)->plan;
  lowdin::Context& c = lowdin::ctx();
  try {
    std::lock_guard<std::mutex> lk(c.mu);
    const double t0 = lowdin::now_ms();
    lowdin::ensure_context();
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
    lowdin::enqueue_plan(p, d_dens, d_result, factors, st);
    if (synchronize) lowdin::collect_plan_stats(p, st, t0);
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("batch_compute_2body_device: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
    return -1;
  }
}

int lowdin_fock_batch_compute_2body_device_ptrs(void* batch, const double* const* d_dens, double* const* d_result,
                                                const double* factors, void* stream, int synchronize) {
  lowdin::BuildPlan& p = static_cast<lowdin_fock_batch_t*>(batch)->plan;
  lowdin::Context& c = lowdin::ctx();
  try {
    std::lock_guard<std::mutex> lk(c.mu);
    const double t0 = lowdin::now_ms();
    lowdin::ensure_context();
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c.stream;
    lowdin::enqueue_plan(p, nullptr, nullptr, factors, st, d_dens, d_result);
    if (synchronize) lowdin::collect_plan_stats(p, st, t0);
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("batch_compute_2body_device_ptrs: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
    return -1;
  }
}

int lowdin_fock_batch_options(void* batch, int time_launches, int serialize, int use_graph) {
  lowdin::BuildPlan& p = static_cast<lowdin_fock_batch_t*>(batch)->plan;
  std::lock_guard<std::mutex> lk(lowdin::ctx().mu);
  p.time_launches = time_launches != 0;
  p.serialize = serialize != 0;
  p.use_graph = use_graph != 0;
  return 0;
}

int lowdin_fock_batch_get_stats(void* batch, lowdin_fock_stats* out) {
  if (!batch || !out) return -1;
  *out = static_cast<lowdin_fock_batch_t*>(batch)->plan.stats;
  return 0;
}

int lowdin_fock_batch_launch_profile(void* batch, int cap, int* classes, float* ms, uint64_t* quartets,
                                     uint64_t* prim_quartets) {
  return plan_profile(static_cast<lowdin_fock_batch_t*>(batch)->plan, cap, classes, ms, quartets, prim_quartets);
}

// Per class launch: try each candidate workgroup target on serialised,
// event-timed builds and keep the fastest (ties to the smaller target).
static int plan_autotune(lowdin::BuildPlan& p, const double* d_dens, double* d_result, const double* factors,
                         hipStream_t stream, int reps, const double* const* dD_each = nullptr,
                         double* const* dG_each = nullptr) {
  lowdin::Context& c = lowdin::ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  lowdin::ensure_context();
  hipStream_t st = stream ? stream : c.stream;
  // candidates: the wavefront kernel, LDS strips aiming at that many
  // workgroups per launch, and LDS strips with one-wavefront workgroups
  static const int cand[][2] = {{2, 2048}, {1, 512}, {1, 2048}, {1, 8192}, {3, 512}, {4, 2048},
                                {4, 8192}, {5, 2048}, {5, 8192}, {5, 32768}};   // {kind, workgroups}
  const bool tl = p.time_launches, se = p.serialize, ug = p.use_graph;
  p.time_launches = true;
  p.serialize = true;
  p.use_graph = false;
  double best[lowdin::kNumClassPairs];
  int pick[lowdin::kNumClassPairs], pick_kind[lowdin::kNumClassPairs];
  for (int k = 0; k < lowdin::kNumClassPairs; ++k) {
    best[k] = 1e300;
    pick[k] = p.target_wg[k];
    pick_kind[k] = p.kind[k];
  }
  try {
    for (const auto& cd : cand) {
      const int kind = cd[0], wg = cd[1];
      for (int k = 0; k < lowdin::kNumClassPairs; ++k) {
        p.kind[k] = kind;
        p.target_wg[k] = wg;
      }
      p.valid = false;
      std::vector<double> ms;
      for (int r = 0; r < reps + 1; ++r) {
        lowdin::enqueue_plan(p, d_dens, d_result, factors, st, dD_each, dG_each);
        lowdin::collect_plan_stats(p, st, lowdin::now_ms());
        if (r == 0) {
          ms.assign(p.ts.launches.size(), 0.0);
          continue;  // warm-up
        }
        for (size_t k = 0; k < p.ts.launches.size(); ++k) ms[k] += p.launch_ms[k];
      }
      for (size_t k = 0; k < p.ts.launches.size(); ++k) {
        const lowdin::TaskSet::Launch& L = p.ts.launches[k];
        const int id = lowdin::class_pair_id(L.cbra, L.cket);
        if (kind != 2 && !L.strip) continue;   // strips impossible for this class: candidate ran as waves
        if ((kind >= 4) != (L.strip >= 3)) continue;   // no cooperative kernel, or a default one ran
        if (ms[k] < 0.97 * best[id]) {
          best[id] = ms[k];
          pick[id] = wg;
          pick_kind[id] = kind;
        }
      }
    }
  } catch (...) {
    for (int k = 0; k < lowdin::kNumClassPairs; ++k) {
      p.target_wg[k] = pick[k];
      p.kind[k] = pick_kind[k];
    }
    p.valid = false;
    p.time_launches = tl;
    p.serialize = se;
    p.use_graph = ug;
    throw;
  }
  for (int k = 0; k < lowdin::kNumClassPairs; ++k) {
    p.target_wg[k] = pick[k];
    p.kind[k] = pick_kind[k];
  }
  p.valid = false;
  p.time_launches = tl;
  p.serialize = se;
  p.use_graph = ug;
  return 0;
}

int lowdin_fock_batch_autotune(void* batch, const double* d_dens, double* d_result, const double* factors,
                               void* stream, int reps) {
  try {
    plan_autotune(static_cast<lowdin_fock_batch_t*>(batch)->plan, d_dens, d_result, factors,
                  static_cast<hipStream_t>(stream), std::max(reps, 1));
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::g_last_error = std::string("batch_autotune: ") + e.what();
    fprintf(stderr, "lowdin_fock: %s\n", lowdin::g_last_error.c_str());
    return -1;
  }
}

int lowdin_fock_autotune(void* lint, const double* d_dens, double* d_result, double factor, void* stream, int reps) {
  try {
    plan_autotune(lowdin::own_plan(*static_cast<Species*>(lint)), d_dens, d_result, &factor,
                  static_cast<hipStream_t>(stream), std::max(reps, 1));
    lowdin::g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("autotune: ") + e.what());
    return -1;
  }
}

static int encode_target(int kind, int wg) {
  switch (kind) {
    case 2: return 0;
    case 3: return -wg;
    case 4: return 1000000 + wg;
    case 5: return 2000000 + wg;
    default: return wg;
  }
}

int lowdin_fock_targets(void* lint, int* targets) {
  const lowdin::BuildPlan& p = lowdin::own_plan(*static_cast<Species*>(lint));
  for (int k = 0; k < lowdin::kNumClassPairs; ++k)
    targets[k] = encode_target(p.kind[k], p.target_wg[k]);
  return lowdin::kNumClassPairs;
}

int lowdin_fock_batch_targets(void* batch, int* targets) {
  const lowdin::BuildPlan& p = static_cast<lowdin_fock_batch_t*>(batch)->plan;
  for (int k = 0; k < lowdin::kNumClassPairs; ++k)
    targets[k] = encode_target(p.kind[k], p.target_wg[k]);
  return lowdin::kNumClassPairs;
}

}  // extern "C"

namespace lowdin {
// Capturable enqueue paths for the batched SCF (lowdin_scf_batch.cpp): the
// batch's class launches on per-molecule device buffers, and one fused
// coupling pass; no waits, no events of their own, no statistics.  The
// first call after a change of the molecules allocates and must not be
// under graph capture.
void internal_enqueue_batch_build(void* batch, const double* const* dD, double* const* dG, const double* factors,
                                  hipStream_t st, bool inline_launch) {
  Context& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  ensure_context();
  BuildPlan& p = static_cast<lowdin_fock_batch_t*>(batch)->plan;
  p.inline_launch = inline_launch;
  if (p.use_graph || p.time_launches) throw std::runtime_error("batched SCF builds run without graph or timing");
  enqueue_plan(p, nullptr, nullptr, factors, st, dD, dG, false);
}

// The autotuner on a batch whose densities and results are per-molecule
// device buffers (the batched SCF's slots); the results' contents are
// overwritten.  Waits.
void internal_batch_autotune(void* batch, const double* const* dD, double* const* dG, const double* factors,
                             hipStream_t st, int reps) {
  BuildPlan& p = static_cast<lowdin_fock_batch_t*>(batch)->plan;
  plan_autotune(p, nullptr, nullptr, factors, st, std::max(reps, 1), dD, dG);
}

// Per-class kernel choices of a batch's plan (kNumClassPairs each): read, and
// set on another batch of the same kind of molecules.
void internal_batch_get_tuning(void* batch, int* kind, int* target) {
  const BuildPlan& p = static_cast<lowdin_fock_batch_t*>(batch)->plan;
  for (int k = 0; k < kNumClassPairs; ++k) {
    kind[k] = p.kind[k];
    target[k] = p.target_wg[k];
  }
}
void internal_batch_set_tuning(void* batch, const int* kind, const int* target) {
  BuildPlan& p = static_cast<lowdin_fock_batch_t*>(batch)->plan;
  for (int k = 0; k < kNumClassPairs; ++k) {
    p.kind[k] = kind[k];
    p.target_wg[k] = target[k];
  }
  p.valid = false;
}

void internal_enqueue_coupling_pair(void* a, void* b, const double* dDa, const double* dDb, double* dRa, double* dRb,
                                    hipStream_t st, bool inline_launch) {
  Context& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  ensure_context();
  enqueue_coupling(*static_cast<Species*>(a), *static_cast<Species*>(b), dDb, dRa, dDa, dRb, st, nullptr,
                   inline_launch);
}
}  // namespace lowdin
