// Batched device-resident SCF: many independent jobs (molecules, each one or
// more quantum species) iterated together, every step of an iteration one
// batched operation over all jobs.
//
// The iteration is openLOWDIN's (MultiSCF.f90:276-440; SingleSCF.f90:146-252;
// WaveFunction.f90:394-467, 817-1535, 1768-2003), as in lowdin_scf.cpp:
//   G = q^2 (J + f K) through the batched direct build (one class launch per
//   quartet class over every species of every job), couplings q_i q_j J_ij,
//   F = H + G + C, E = e_pc + sum_s Tr(D (H + G/2 + C/2)), DIIS(8) on
//   X^T (F D S - S D F) X over the species of a job, diagonalisation of
//   X^T F X, D = eta C_occ C_occ^T.
// Species of equal basis size n (and orthogonal dimension m) share a bucket:
// strided arrays with one slot per species, so that every matrix product of
// the iteration is one rocblas strided-batched GEMM per bucket, the
// diagonalisation one rocsolver dsyevd_strided_batched per bucket, and the
// small steps (Fock assembly, traces, error maxima, the DIIS history, its
// B matrix, the solve of the DIIS equations and the extrapolation) are
// batched kernels of this file.  The DIIS state lives on the device.  The
// host sees one copy of the per-slot scalars per iteration.
//
// While the set of jobs is unchanged the iteration is replayed as two HIP
// graphs around the diagonalisation: (build, couplings, assembly, traces,
// error, DIIS, X^T F X) and (C = X C', D = C diag(occ) C^T, scalar copy).
// The first iteration after a change runs eagerly (it also prepares the
// build's task lists); the next one captures.  rocsolver runs outside the
// graphs.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lowdin_fock.h"
#include "lowdin_kernels.h"

namespace lowdin {
void set_last_error(const std::string& msg);
void clear_last_error();
// capturable enqueue paths of lowdin_fock.cpp (no waits, events or statistics)
void internal_enqueue_batch_build(void* batch, const double* const* dD, double* const* dG, const double* factors,
                                  hipStream_t st, bool inline_launch);
void internal_enqueue_coupling_pair(void* a, void* b, const double* dDa, const double* dDb, double* dRa, double* dRb,
                                    hipStream_t st, bool inline_launch);
void internal_batch_autotune(void* batch, const double* const* dD, double* const* dG, const double* factors,
                             hipStream_t st, int reps);
void internal_batch_get_tuning(void* batch, int* kind, int* target);
void internal_batch_set_tuning(void* batch, const int* kind, const int* target);
}  // namespace lowdin

namespace {

using lowdin::kMaxDiis;
constexpr int K = kMaxDiis;

#define BCHK(x)                                                                                      \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define BLCHK(x)                                                                                     \
  do {                                                                                               \
    rocblas_status s_ = (x);                                                                         \
    if (s_ != rocblas_status_success)                                                                \
      throw std::runtime_error(std::string(#x) + ": " + rocblas_status_to_string(s_));               \
  } while (0)

template <typename T>
struct DArr {
  T* p = nullptr;
  size_t n = 0;
  DArr() = default;
  DArr(const DArr&) = delete;
  DArr& operator=(const DArr&) = delete;
  ~DArr() {
    if (p) (void)hipFree(p);
  }
  // grow to at least count elements, keeping the first `keep` elements
  void grow(size_t count, size_t keep, hipStream_t st) {
    if (count <= n) return;
    T* q = nullptr;
    BCHK(hipMalloc((void**)&q, count * sizeof(T)));
    if (p && keep) BCHK(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, st));
    BCHK(hipStreamSynchronize(st));
    if (p) (void)hipFree(p);
    p = q;
    n = count;
  }
};

// ------------------------------------------------------------ kernels
// Per slot scalars: [0] Tr(D H), [1] Tr(D G), [2] Tr(D C), [3] max |error|.
constexpr int kSlotScal = 4;

// Job state on the device: ring of kMaxDiis Fock / error matrices per slot,
// B matrix, new row, coefficients.
struct JobDev {
  int nhist;
  int head;
  int pos;            // ring slot of the current iteration (set by diis_begin)
  int pad;
  double B[K][K];
  double row[K];
  double c[K];        // extrapolation weights by ring slot
};

// F = H + qq G + C over nslot slots of nn elements.
__global__ void b_assemble_kernel(double* F, const double* H, const double* G, const double* C, const double* qq,
                                  size_t nn, int nslot) {
  const int s = blockIdx.y;
  if (s >= nslot) return;
  const double q = qq[s];
  const size_t o = (size_t)s * nn;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x)
    F[o + e] = H[o + e] + q * G[o + e] + C[o + e];
}

// Tr(D X) = sum D .* X for symmetric D: one block per (slot, term).
__global__ __launch_bounds__(256) void b_traces_kernel(const double* D, const double* H, const double* G,
                                                       const double* C, size_t nn, double* sc) {
  const int s = blockIdx.x, t = blockIdx.y;
  const double* X = t == 0 ? H : (t == 1 ? G : C);
  const size_t o = (size_t)s * nn;
  double acc = 0.0;
  for (size_t e = threadIdx.x; e < nn; e += blockDim.x) acc += D[o + e] * X[o + e];
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sc[s * kSlotScal + t] = red[0];
}

// E = T - T^T (n x n per slot).
__global__ void b_antisym_kernel(double* E, const double* T, int n, int nslot) {
  const int s = blockIdx.z;
  const int i = blockIdx.y * blockDim.y + threadIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot || i >= n || j >= n) return;
  const size_t o = (size_t)s * n * n;
  E[o + (size_t)i * n + j] = T[o + (size_t)i * n + j] - T[o + (size_t)j * n + i];
}

// max |E| per slot (m x m), into sc[s][3].
__global__ __launch_bounds__(256) void b_absmax_kernel(const double* E, size_t mm, double* sc) {
  const int s = blockIdx.x;
  const size_t o = (size_t)s * mm;
  double v = 0.0;
  for (size_t e = threadIdx.x; e < mm; e += blockDim.x) v = fmax(v, fabs(E[o + e]));
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) sc[s * kSlotScal + 3] = red[0];
}

// Ring slot of this iteration for every job; zero the B row.
__global__ void b_diis_begin_kernel(JobDev* jobs, const int* active, int njobs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs || !active[j]) return;
  JobDev& J = jobs[j];
  J.pos = J.nhist < K ? J.nhist : J.head;
  for (int i = 0; i < K; ++i) J.row[i] = 0.0;
}

// Store F and E of every slot into its job's ring slot.
__global__ void b_diis_store_kernel(double* Fh, double* Eh, const double* F, const double* E, const int* slot_job,
                                    const JobDev* jobs, size_t nn, size_t mm, int nslot) {
  const int s = blockIdx.y;
  if (s >= nslot) return;
  const int j = slot_job[s];
  if (j < 0) return;
  const int pos = jobs[j].pos;
  double* fh = Fh + ((size_t)s * K + pos) * nn;
  double* eh = Eh + ((size_t)s * K + pos) * mm;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    fh[e] = F[(size_t)s * nn + e];
    if (e < mm) eh[e] = E[(size_t)s * mm + e];
  }
}

// <E_i, E_new> of every slot for every stored ring slot i, summed per job.
__global__ __launch_bounds__(256) void b_diis_row_kernel(const double* Eh, const double* E, const int* slot_job,
                                                         JobDev* jobs, size_t mm, int nslot) {
  const int s = blockIdx.x, i = blockIdx.y;
  if (s >= nslot) return;
  const int j = slot_job[s];
  if (j < 0) return;
  const int nstored = min(jobs[j].nhist + 1, K);
  if (i >= nstored) return;
  const double* a = Eh + ((size_t)s * K + i) * mm;
  const double* b = E + (size_t)s * mm;
  double acc = 0.0;
  for (size_t e = threadIdx.x; e < mm; e += blockDim.x) acc += a[e] * b[e];
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&jobs[j].row[i], red[0]);
}

// Per job: enter the new row into B, advance the ring, solve
// [B -1; -1 0][c; l] = [0; -1] over the stored matrices in chronological
// order (Gaussian elimination with partial pivoting), weights by ring slot
// (only the newest matrix when the system is singular).
__global__ void b_diis_solve_kernel(JobDev* jobs, const int* active, int njobs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs || !active[j]) return;
  JobDev& J = jobs[j];
  const int pos = J.pos;
  const int nstored = min(J.nhist + 1, K);
  for (int i = 0; i < nstored; ++i) {
    J.B[pos][i] = J.row[i];
    J.B[i][pos] = J.row[i];
  }
  if (J.nhist < K)
    J.nhist++;
  else
    J.head = (J.head + 1) % K;
  const int m = J.nhist;
  int idx[K];
  for (int i = 0; i < m; ++i) idx[i] = (m < K) ? i : (J.head + i) % K;
  const int N = m + 1;
  double A[K + 1][K + 2];
  for (int r = 0; r < N; ++r)
    for (int q = 0; q <= N; ++q) A[r][q] = 0.0;
  for (int r = 0; r < m; ++r) {
    for (int q = 0; q < m; ++q) A[r][q] = J.B[idx[r]][idx[q]];
    A[r][m] = -1.0;
    A[m][r] = -1.0;
  }
  A[m][N] = -1.0;
  bool ok = true;
  for (int col = 0; col < N && ok; ++col) {
    int piv = col;
    for (int r = col + 1; r < N; ++r)
      if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
    if (A[piv][col] == 0.0) {
      ok = false;
      break;
    }
    if (piv != col)
      for (int q = 0; q <= N; ++q) {
        const double t = A[piv][q];
        A[piv][q] = A[col][q];
        A[col][q] = t;
      }
    for (int r = col + 1; r < N; ++r) {
      const double f = A[r][col] / A[col][col];
      for (int q = col; q <= N; ++q) A[r][q] -= f * A[col][q];
    }
  }
  double x[K + 1];
  if (ok) {
    for (int r = N - 1; r >= 0; --r) {
      double v = A[r][N];
      for (int q = r + 1; q < N; ++q) v -= A[r][q] * x[q];
      x[r] = v / A[r][r];
    }
  }
  for (int i = 0; i < K; ++i) J.c[i] = 0.0;
  if (ok) {
    for (int i = 0; i < m; ++i) J.c[idx[i]] = x[i];
  } else {
    J.c[pos] = 1.0;
  }
}

// F = sum_i c_i F_i over the job's ring.
__global__ void b_diis_extrapolate_kernel(double* F, const double* Fh, const int* slot_job, const JobDev* jobs,
                                          size_t nn, int nslot) {
  const int s = blockIdx.y;
  if (s >= nslot) return;
  const int j = slot_job[s];
  if (j < 0) return;
  const JobDev& J = jobs[j];
  const double* fh = Fh + (size_t)s * K * nn;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int i = 0; i < K; ++i) {
      const double c = J.c[i];
      if (c != 0.0) v += c * fh[(size_t)i * nn + e];
    }
    F[(size_t)s * nn + e] = v;
  }
}

// T = C diag(occ) (n x m per slot, occupation weights per slot and column).
__global__ void b_scale_cols_kernel(double* T, const double* C, const double* occ, int n, int m, int nslot) {
  const int s = blockIdx.z;
  const int col = blockIdx.y, row = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot || col >= m || row >= n) return;
  const size_t o = (size_t)s * n * n;
  T[o + (size_t)col * n + row] = C[o + (size_t)col * n + row] * occ[(size_t)s * m + col];
}

__global__ void b_fill_kernel(double* p, size_t count, double v) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
    p[e] = v;
}

// F -= Cg over every slot (the ghost species' coupling leaves the Fock matrix).
__global__ void b_sub_kernel(double* F, const double* Cg, size_t count) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
    F[e] -= Cg[e];
}

__global__ void b_job_reset_kernel(JobDev* jobs, int j) {
  JobDev& J = jobs[j];
  J.nhist = 0;
  J.head = 0;
  J.pos = 0;
  for (int a = 0; a < K; ++a) {
    J.row[a] = 0.0;
    J.c[a] = 0.0;
    for (int b = 0; b < K; ++b) J.B[a][b] = 0.0;
  }
}

unsigned grid1(size_t count) { return (unsigned)std::max<size_t>(1, std::min<size_t>((count + 255) / 256, 1024)); }

// ------------------------------------------------------------ host side
struct SlotInfo {
  int job = -1;        // -1: free slot
  int species = 0;
  void* lint = nullptr;
  double factor = -0.5, charge = -1.0, eta = 2.0;
  int nocc = 0;
  bool build_two = true;
  bool ghost = false;   // SCF_GHOST_SPECIES: its coupling stays out of the job's other Fock matrices
};

struct Bucket {
  int n = 0, m = 0;
  int cap = 0;
  std::vector<SlotInfo> slots;
  DArr<double> H, S, X, D, G, Cs, Cg, F, T, R, C, Fp, E, W, occ, qq, Fh, Eh, sc;   // Cg: coupling from ghost species
  DArr<int> info, slot_job;
  size_t nn() const { return (size_t)n * n; }
  size_t mm() const { return (size_t)m * m; }
  int used() const {   // highest occupied slot + 1
    int u = 0;
    for (int k = 0; k < (int)slots.size(); ++k)
      if (slots[k].job >= 0) u = k + 1;
    return u;
  }
};

struct JobHost {
  bool active = false;
  double e_pc = 0.0;
  std::vector<std::pair<int, int>> where;   // (bucket, slot) per species
  double energy = 0.0, error = 0.0;
  int iterations = 0;
  int status = 0;     // 0 running, 1 failed (rocsolver info), 2 removed
  std::string message;
};

struct Engine {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  rocblas_handle blas = nullptr;
  std::vector<std::unique_ptr<Bucket>> buckets;
  std::vector<JobHost> jobs;
  DArr<JobDev> d_jobs;
  DArr<int> d_active;
  int job_cap = 0;
  double* h_sc = nullptr;   // pinned: all slots' scalars + info, per bucket in order
  size_t h_cap = 0;
  hipEvent_t done = nullptr;
  void* fock_batch = nullptr;
  // per-class kernel choices of the autotuner, applied to every later
  // version of the batched build (molecules leave as they converge)
  bool tuned = false;
  int tune_kind[21] = {}, tune_target[21] = {};
  // version counts changes of the job set; graphs are captured only for a
  // version whose eager iteration has run (that one prepares the build's
  // task lists and buffers, which allocate and must not be captured)
  uint64_t version = 1, batch_version = 0, graph_version = 0, warm_version = 0;
  bool use_graph = true;
  bool inline_launch = false;   // class launches on this engine's stream, not the library's side streams
  double overlap_threshold = 1e-8;   // OVERLAP_EIGEN_THRESHOLD for the jobs added from now on
  bool in_flight = false;       // an iteration enqueued and not yet collected
  hipGraphExec_t g1 = nullptr, g2 = nullptr;
  std::vector<void*> b_lints;
  std::vector<const double*> b_D;
  std::vector<double*> b_G;
  std::vector<double> b_fac;
  DArr<double> tmpA, tmpW, tmpE;
  DArr<int> tmpInfo;
  size_t bytes = 0;

  ~Engine() {
    if (g1) (void)hipGraphExecDestroy(g1);
    if (g2) (void)hipGraphExecDestroy(g2);
    if (fock_batch) lowdin_fock_batch_del(fock_batch);
    if (blas) (void)rocblas_destroy_handle(blas);
    if (h_sc) (void)hipHostFree(h_sc);
    if (done) (void)hipEventDestroy(done);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
  }
};

void gemm_sb(Engine& e, rocblas_operation ta, rocblas_operation tb, int m, int n, int k, const double* A, int lda,
             size_t sa, const double* B, int ldb, size_t sb, double beta, double* C, int ldc, size_t sc, int count,
             double alpha = 1.0) {
  if (count <= 0) return;
  BLCHK(rocblas_dgemm_strided_batched(e.blas, ta, tb, m, n, k, &alpha, A, lda, (rocblas_stride)sa, B, ldb,
                                      (rocblas_stride)sb, &beta, C, ldc, (rocblas_stride)sc, count));
}

Bucket& bucket_for(Engine& e, int n, int m) {
  for (auto& b : e.buckets)
    if (b->n == n && b->m == m) return *b;
  e.buckets.emplace_back(new Bucket());
  Bucket& b = *e.buckets.back();
  b.n = n;
  b.m = m;
  return b;
}

// A free slot of bucket b, growing every array when none is left.
int take_slot(Engine& e, Bucket& b) {
  for (int k = 0; k < (int)b.slots.size(); ++k)
    if (b.slots[k].job < 0) return k;
  const int old = b.cap;
  const int cap = std::max(4, 2 * old);
  const size_t nn = b.nn(), mm = b.mm();
  hipStream_t st = e.stream;
  for (DArr<double>* a : {&b.H, &b.S, &b.X, &b.D, &b.G, &b.Cs, &b.Cg, &b.F, &b.T, &b.R, &b.C})
    a->grow((size_t)cap * nn, (size_t)old * nn, st);
  b.Fp.grow((size_t)cap * mm, (size_t)old * mm, st);
  b.E.grow((size_t)cap * mm, (size_t)old * mm, st);
  b.W.grow((size_t)cap * b.m, (size_t)old * b.m, st);
  b.occ.grow((size_t)cap * b.m, (size_t)old * b.m, st);
  b.qq.grow(cap, old, st);
  b.sc.grow((size_t)cap * kSlotScal, (size_t)old * kSlotScal, st);
  b.Fh.grow((size_t)cap * K * nn, (size_t)old * K * nn, st);
  b.Eh.grow((size_t)cap * K * mm, (size_t)old * K * mm, st);
  b.info.grow(cap, old, st);
  b.slot_job.grow(cap, old, st);
  e.bytes += (size_t)(cap - old) * ((11 + K) * nn + (2 + K) * mm + 2 * b.m + kSlotScal + 1) * sizeof(double);
  b.cap = cap;
  b.slots.resize(cap);
  // new slots: nobody's, zero G (species without a two-particle build keep it)
  std::vector<int> owners(cap - old, -1);
  BCHK(hipMemcpyAsync(b.slot_job.p + old, owners.data(), owners.size() * sizeof(int), hipMemcpyHostToDevice, st));
  for (DArr<double>* a : {&b.G, &b.Cs, &b.Cg, &b.D, &b.H, &b.S, &b.X, &b.F, &b.Fp, &b.E})
    hipLaunchKernelGGL(b_fill_kernel, dim3(grid1((size_t)(cap - old) * (a == &b.Fp || a == &b.E ? mm : nn))),
                       dim3(256), 0, st, a->p + (size_t)old * (a == &b.Fp || a == &b.E ? mm : nn),
                       (size_t)(cap - old) * (a == &b.Fp || a == &b.E ? mm : nn), 0.0);
  hipLaunchKernelGGL(b_fill_kernel, dim3(grid1((size_t)(cap - old) * K * nn)), dim3(256), 0, st,
                     b.Fh.p + (size_t)old * K * nn, (size_t)(cap - old) * K * nn, 0.0);
  hipLaunchKernelGGL(b_fill_kernel, dim3(grid1((size_t)(cap - old) * K * mm)), dim3(256), 0, st,
                     b.Eh.p + (size_t)old * K * mm, (size_t)(cap - old) * K * mm, 0.0);
  hipLaunchKernelGGL(b_fill_kernel, dim3(grid1((size_t)(cap - old) * b.m)), dim3(256), 0, st,
                     b.occ.p + (size_t)old * b.m, (size_t)(cap - old) * b.m, 0.0);
  BCHK(hipGetLastError());
  return old;
}

void drop_graphs(Engine& e) {
  if (e.g1) (void)hipGraphExecDestroy(e.g1);
  if (e.g2) (void)hipGraphExecDestroy(e.g2);
  e.g1 = e.g2 = nullptr;
  e.graph_version = 0;
}

// The batched direct build's molecule list (every slot with a two-particle build).
void ensure_fock_batch(Engine& e) {
  if (e.batch_version == e.version) return;
  if (e.fock_batch) lowdin_fock_batch_del(e.fock_batch);
  e.fock_batch = nullptr;
  e.b_lints.clear();
  e.b_D.clear();
  e.b_G.clear();
  e.b_fac.clear();
  for (auto& bp : e.buckets) {
    Bucket& b = *bp;
    for (int k = 0; k < (int)b.slots.size(); ++k) {
      const SlotInfo& s = b.slots[k];
      if (s.job < 0 || !s.build_two) continue;
      e.b_lints.push_back(s.lint);
      e.b_D.push_back(b.D.p + (size_t)k * b.nn());
      e.b_G.push_back(b.G.p + (size_t)k * b.nn());
      e.b_fac.push_back(s.factor);
    }
  }
  if (!e.b_lints.empty()) {
    e.fock_batch = lowdin_fock_batch_new((int)e.b_lints.size(), e.b_lints.data());
    if (!e.fock_batch) throw std::runtime_error(std::string("batch build: ") + lowdin_fock_last_error());
    if (e.tuned) lowdin::internal_batch_set_tuning(e.fock_batch, e.tune_kind, e.tune_target);
  }
  e.batch_version = e.version;
}

// Everything before the diagonalisation.
void enqueue_part1(Engine& e) {
  hipStream_t st = e.stream;
  if (e.fock_batch)
    lowdin::internal_enqueue_batch_build(e.fock_batch, e.b_D.data(), e.b_G.data(), e.b_fac.data(), st,
                                         e.inline_launch);
  // a job with a ghost species: its partners' Fock matrices leave out that
  // coupling (Cg), the energy keeps it (WaveFunction.f90:1502-1503, 1980-1986)
  bool any_ghost = false;
  for (const auto& bp : e.buckets)
    for (const SlotInfo& si : bp->slots) any_ghost = any_ghost || (si.job >= 0 && si.ghost);
  for (auto& bp : e.buckets) {
    Bucket& b = *bp;
    const int u = b.used();
    if (u) hipLaunchKernelGGL(b_fill_kernel, dim3(grid1((size_t)u * b.nn())), dim3(256), 0, st, b.Cs.p, (size_t)u * b.nn(), 0.0);
    if (u && any_ghost)
      hipLaunchKernelGGL(b_fill_kernel, dim3(grid1((size_t)u * b.nn())), dim3(256), 0, st, b.Cg.p, (size_t)u * b.nn(), 0.0);
  }
  // couplings of the jobs with several species (one fused pass per pair)
  for (const JobHost& J : e.jobs) {
    if (!J.active || J.where.size() < 2) continue;
    for (size_t a = 0; a < J.where.size(); ++a)
      for (size_t c = a + 1; c < J.where.size(); ++c) {
        Bucket& ba = *e.buckets[J.where[a].first];
        Bucket& bc = *e.buckets[J.where[c].first];
        const int sa = J.where[a].second, sc = J.where[c].second;
        const SlotInfo& ia = ba.slots[sa];
        const SlotInfo& ic = bc.slots[sc];
        double* Ra = ba.R.p + (size_t)sa * ba.nn();
        double* Rc = bc.R.p + (size_t)sc * bc.nn();
        lowdin::internal_enqueue_coupling_pair(ia.lint, ic.lint, ba.D.p + (size_t)sa * ba.nn(),
                                               bc.D.p + (size_t)sc * bc.nn(), Ra, Rc, st, e.inline_launch);
        const double qq = ia.charge * ic.charge;
        lowdin::launch_sym_add(ba.Cs.p + (size_t)sa * ba.nn(), Ra, ba.n, qq, st);
        lowdin::launch_sym_add(bc.Cs.p + (size_t)sc * bc.nn(), Rc, bc.n, qq, st);
        if (ic.ghost) lowdin::launch_sym_add(ba.Cg.p + (size_t)sa * ba.nn(), Ra, ba.n, qq, st);
        if (ia.ghost) lowdin::launch_sym_add(bc.Cg.p + (size_t)sc * bc.nn(), Rc, bc.n, qq, st);
      }
  }
  hipLaunchKernelGGL(b_diis_begin_kernel, dim3((e.job_cap + 63) / 64), dim3(64), 0, st, e.d_jobs.p, e.d_active.p,
                     e.job_cap);
  const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
  for (auto& bp : e.buckets) {
    Bucket& b = *bp;
    const int u = b.used();
    if (!u) continue;
    const int n = b.n, m = b.m;
    const size_t nn = b.nn(), mm = b.mm();
    hipLaunchKernelGGL(b_assemble_kernel, dim3(grid1(nn), u), dim3(256), 0, st, b.F.p, b.H.p, b.G.p, b.Cs.p, b.qq.p,
                       nn, u);
    if (any_ghost)
      hipLaunchKernelGGL(b_sub_kernel, dim3(grid1((size_t)u * nn)), dim3(256), 0, st, b.F.p, b.Cg.p, (size_t)u * nn);
    hipLaunchKernelGGL(b_traces_kernel, dim3(u, 3), dim3(256), 0, st, b.D.p, b.H.p, b.G.p, b.Cs.p, nn, b.sc.p);
    // error X^T (F D S - S D F) X
    gemm_sb(e, N, N, n, n, n, b.F.p, n, nn, b.D.p, n, nn, 0.0, b.T.p, n, nn, u);
    gemm_sb(e, N, N, n, n, n, b.T.p, n, nn, b.S.p, n, nn, 0.0, b.R.p, n, nn, u);
    hipLaunchKernelGGL(b_antisym_kernel, dim3((n + 31) / 32, (n + 7) / 8, u), dim3(32, 8), 0, st, b.T.p, b.R.p, n, u);
    gemm_sb(e, T, N, m, n, n, b.X.p, n, nn, b.T.p, n, nn, 0.0, b.C.p, m, nn, u);      // X^T A (m x n)
    gemm_sb(e, N, N, m, m, n, b.C.p, m, nn, b.X.p, n, nn, 0.0, b.E.p, m, mm, u);      // (X^T A) X
    hipLaunchKernelGGL(b_absmax_kernel, dim3(u), dim3(256), 0, st, b.E.p, mm, b.sc.p);
    hipLaunchKernelGGL(b_diis_store_kernel, dim3(grid1(nn), u), dim3(256), 0, st, b.Fh.p, b.Eh.p, b.F.p, b.E.p,
                       b.slot_job.p, e.d_jobs.p, nn, mm, u);
    hipLaunchKernelGGL(b_diis_row_kernel, dim3(u, K), dim3(256), 0, st, b.Eh.p, b.E.p, b.slot_job.p, e.d_jobs.p, mm,
                       u);
  }
  hipLaunchKernelGGL(b_diis_solve_kernel, dim3((e.job_cap + 63) / 64), dim3(64), 0, st, e.d_jobs.p, e.d_active.p,
                     e.job_cap);
  for (auto& bp : e.buckets) {
    Bucket& b = *bp;
    const int u = b.used();
    if (!u) continue;
    const int n = b.n, m = b.m;
    const size_t nn = b.nn(), mm = b.mm();
    hipLaunchKernelGGL(b_diis_extrapolate_kernel, dim3(grid1(nn), u), dim3(256), 0, st, b.F.p, b.Fh.p, b.slot_job.p,
                       e.d_jobs.p, nn, u);
    gemm_sb(e, N, N, n, m, n, b.F.p, n, nn, b.X.p, n, nn, 0.0, b.T.p, n, nn, u);      // F X
    gemm_sb(e, T, N, m, m, n, b.X.p, n, nn, b.T.p, n, nn, 0.0, b.Fp.p, m, mm, u);     // X^T F X
  }
  BCHK(hipGetLastError());
}

void enqueue_syevd(Engine& e) {
  for (auto& bp : e.buckets) {
    Bucket& b = *bp;
    const int u = b.used();
    if (!u) continue;
    BLCHK(rocsolver_dsyevd_strided_batched(e.blas, rocblas_evect_original, rocblas_fill_upper, b.m, b.Fp.p, b.m,
                                           (rocblas_stride)b.mm(), b.W.p, b.m, b.E.p, b.m, b.info.p, u));
  }
}

// C = X C', D = C diag(occ) C^T, scalars to the host.
void enqueue_part2(Engine& e) {
  hipStream_t st = e.stream;
  const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
  size_t off = 0;
  for (auto& bp : e.buckets) {
    Bucket& b = *bp;
    const int u = b.used();
    if (u) {
      const int n = b.n, m = b.m;
      const size_t nn = b.nn(), mm = b.mm();
      gemm_sb(e, N, N, n, m, m, b.X.p, n, nn, b.Fp.p, m, mm, 0.0, b.C.p, n, nn, u);
      hipLaunchKernelGGL(b_scale_cols_kernel, dim3((n + 255) / 256, m, u), dim3(256), 0, st, b.T.p, b.C.p, b.occ.p,
                         n, m, u);
      gemm_sb(e, N, T, n, n, m, b.T.p, n, nn, b.C.p, n, nn, 0.0, b.D.p, n, nn, u);
      BCHK(hipMemcpyAsync(e.h_sc + off, b.sc.p, (size_t)u * kSlotScal * sizeof(double), hipMemcpyDeviceToHost, st));
      BCHK(hipMemcpyAsync(e.h_sc + off + (size_t)u * kSlotScal, b.info.p, (size_t)u * sizeof(int),
                          hipMemcpyDeviceToHost, st));
    }
    off += (size_t)b.cap * (kSlotScal + 1);
  }
  BCHK(hipGetLastError());
}

void ensure_host_scalars(Engine& e) {
  size_t need = 0;
  for (auto& b : e.buckets) need += (size_t)b->cap * (kSlotScal + 1);
  if (need > e.h_cap) {
    if (e.h_sc) (void)hipHostFree(e.h_sc);
    e.h_sc = nullptr;
    BCHK(hipHostMalloc((void**)&e.h_sc, std::max<size_t>(need, 1) * sizeof(double), hipHostMallocDefault));
    e.h_cap = need;
  }
}

// Enqueue one iteration of every active job (nothing waits).
void step_begin(Engine& e) {
  if (e.in_flight) throw std::runtime_error("an iteration is already in flight");
  hipStream_t st = e.stream;
  ensure_fock_batch(e);
  ensure_host_scalars(e);
  if (e.use_graph && e.graph_version == e.version && e.g1 && e.g2) {
    BCHK(hipGraphLaunch(e.g1, st));
    enqueue_syevd(e);
    BCHK(hipGraphLaunch(e.g2, st));
  } else if (e.use_graph && e.warm_version == e.version) {
    drop_graphs(e);
    hipGraph_t g = nullptr;
    BCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_part1(e);
    } catch (...) {
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    BCHK(hipStreamEndCapture(st, &g));
    hipError_t r = hipGraphInstantiate(&e.g1, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BCHK(r);
    g = nullptr;
    BCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_part2(e);
    } catch (...) {
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    BCHK(hipStreamEndCapture(st, &g));
    r = hipGraphInstantiate(&e.g2, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BCHK(r);
    e.graph_version = e.version;
    BCHK(hipGraphLaunch(e.g1, st));
    enqueue_syevd(e);
    BCHK(hipGraphLaunch(e.g2, st));
  } else {
    enqueue_part1(e);
    enqueue_syevd(e);
    enqueue_part2(e);
    e.warm_version = e.version;
  }
  BCHK(hipEventRecord(e.done, st));
  e.in_flight = true;
}

// Wait for the iteration in flight; per job: energy and error of the
// densities it started from.
void step_end(Engine& e) {
  if (!e.in_flight) throw std::runtime_error("no iteration in flight");
  BCHK(hipEventSynchronize(e.done));
  e.in_flight = false;
  std::vector<size_t> boff;
  size_t off = 0;
  for (auto& b : e.buckets) {
    boff.push_back(off);
    off += (size_t)b->cap * (kSlotScal + 1);
  }
  for (JobHost& J : e.jobs) {
    if (!J.active) continue;
    double E = J.e_pc, err = 0.0;
    for (auto& w : J.where) {
      Bucket& b = *e.buckets[w.first];
      const int u = b.used();
      const double* sc = e.h_sc + boff[w.first] + (size_t)w.second * kSlotScal;
      const SlotInfo& s = b.slots[w.second];
      E += sc[0] + 0.5 * s.charge * s.charge * sc[1] + 0.5 * sc[2];
      err = std::max(err, sc[3]);
      int info = 0;
      std::memcpy(&info, reinterpret_cast<const int*>(e.h_sc + boff[w.first] + (size_t)u * kSlotScal) + w.second,
                  sizeof(int));
      if (info != 0 && J.status == 0) {
        J.status = 1;
        J.message = "rocsolver_dsyevd did not converge (info " + std::to_string(info) + ")";
      }
    }
    J.energy = E;
    J.error = err;
    J.iterations++;
  }
}

// Orthogonaliser and H-core guess of one species (eager, once per job).
void setup_species(Engine& e, const double* H, const double* S, int n, std::vector<double>& hX, int& m) {
  hipStream_t st = e.stream;
  const size_t nn = (size_t)n * n;
  e.tmpA.grow(nn, 0, st);
  e.tmpW.grow(n, 0, st);
  e.tmpE.grow(n, 0, st);
  e.tmpInfo.grow(1, 0, st);
  BCHK(hipMemcpyAsync(e.tmpA.p, S, nn * sizeof(double), hipMemcpyHostToDevice, st));
  BLCHK(rocsolver_dsyevd(e.blas, rocblas_evect_original, rocblas_fill_upper, n, e.tmpA.p, n, e.tmpW.p, e.tmpE.p,
                         e.tmpInfo.p));
  std::vector<double> w(n), U(nn);
  int info = 0;
  BCHK(hipMemcpyAsync(w.data(), e.tmpW.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
  BCHK(hipMemcpyAsync(U.data(), e.tmpA.p, nn * sizeof(double), hipMemcpyDeviceToHost, st));
  BCHK(hipMemcpyAsync(&info, e.tmpInfo.p, sizeof(int), hipMemcpyDeviceToHost, st));
  BCHK(hipStreamSynchronize(st));
  if (info != 0) throw std::runtime_error("overlap diagonalisation failed");
  int col0 = 0;
  while (col0 < n && !(w[col0] >= e.overlap_threshold)) ++col0;
  m = n - col0;
  hX.assign(nn, 0.0);   // X = U_kept s^-1/2, n x m in an n x n slot (column-major)
  for (int c = 0; c < m; ++c) {
    const double f = 1.0 / std::sqrt(w[col0 + c]);
    for (int r = 0; r < n; ++r) hX[(size_t)c * n + r] = U[(size_t)(col0 + c) * n + r] * f;
  }
}

// Density of slot k of bucket b from F' = X^T M X of the host matrix M held
// in the slot's F (one slot; used for the H-core guess).
void density_one(Engine& e, Bucket& b, int k) {
  hipStream_t st = e.stream;
  const int n = b.n, m = b.m;
  const size_t nn = b.nn(), mm = b.mm();
  const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
  double* F = b.F.p + (size_t)k * nn;
  double* X = b.X.p + (size_t)k * nn;
  double* Tm = b.T.p + (size_t)k * nn;
  double* Fp = b.Fp.p + (size_t)k * mm;
  gemm_sb(e, N, N, n, m, n, F, n, nn, X, n, nn, 0.0, Tm, n, nn, 1);
  gemm_sb(e, T, N, m, m, n, X, n, nn, Tm, n, nn, 0.0, Fp, m, mm, 1);
  e.tmpE.grow(m, 0, st);
  e.tmpInfo.grow(1, 0, st);
  BLCHK(rocsolver_dsyevd(e.blas, rocblas_evect_original, rocblas_fill_upper, m, Fp, m, b.W.p + (size_t)k * m,
                         e.tmpE.p, e.tmpInfo.p));
  double* C = b.C.p + (size_t)k * nn;
  gemm_sb(e, N, N, n, m, m, X, n, nn, Fp, m, mm, 0.0, C, n, nn, 1);
  hipLaunchKernelGGL(b_scale_cols_kernel, dim3((n + 255) / 256, m, 1), dim3(256), 0, st, Tm, C, b.occ.p + (size_t)k * m,
                     n, m, 1);
  gemm_sb(e, N, T, n, n, m, Tm, n, nn, C, n, nn, 0.0, b.D.p + (size_t)k * nn, n, nn, 1);
  BCHK(hipGetLastError());
}

int add_job(Engine& e, int nspecies, void* const* lints, const lowdin_scf_species* desc, double e_pc,
            const double* const* H, const double* const* S) {
  hipStream_t st = e.stream;
  if (e.in_flight) throw std::runtime_error("jobs change only between iterations");
  if (nspecies < 1) throw std::runtime_error("a job needs at least one species");
  int j = -1;
  for (int k = 0; k < (int)e.jobs.size(); ++k)
    if (!e.jobs[k].active && e.jobs[k].status == 2) {
      j = k;
      break;
    }
  if (j < 0) {
    j = (int)e.jobs.size();
    e.jobs.emplace_back();
  }
  if ((int)e.jobs.size() > e.job_cap) {
    const int cap = std::max(16, 2 * e.job_cap);
    e.d_jobs.grow(cap, e.job_cap, st);
    e.d_active.grow(cap, e.job_cap, st);
    std::vector<int> zero(cap - e.job_cap, 0);
    BCHK(hipMemcpyAsync(e.d_active.p + e.job_cap, zero.data(), zero.size() * sizeof(int), hipMemcpyHostToDevice, st));
    e.job_cap = cap;
  }
  JobHost J;
  J.e_pc = e_pc;
  std::vector<std::pair<int, int>> placed;
  try {
    for (int s = 0; s < nspecies; ++s) {
      const int n = LibintInterface_get_nbasis(lints[s]);
      std::vector<double> hX;
      int m = 0;
      setup_species(e, H[s], S[s], n, hX, m);
      if (m < desc[s].nocc) throw std::runtime_error("fewer orthogonal functions than occupied orbitals");
      int bi = -1;
      Bucket& b = bucket_for(e, n, m);
      for (int q = 0; q < (int)e.buckets.size(); ++q)
        if (e.buckets[q].get() == &b) bi = q;
      const int k = take_slot(e, b);
      SlotInfo& si = b.slots[k];
      si.job = j;
      si.species = s;
      si.lint = lints[s];
      si.factor = desc[s].factor;
      si.charge = desc[s].charge;
      si.eta = desc[s].eta;
      si.nocc = desc[s].nocc;
      si.build_two = desc[s].build_two != 0;
      si.ghost = desc[s].ghost != 0;
      placed.push_back({bi, k});
      const size_t nn = b.nn();
      BCHK(hipMemcpyAsync(b.H.p + k * nn, H[s], nn * sizeof(double), hipMemcpyHostToDevice, st));
      BCHK(hipMemcpyAsync(b.S.p + k * nn, S[s], nn * sizeof(double), hipMemcpyHostToDevice, st));
      BCHK(hipMemcpyAsync(b.X.p + k * nn, hX.data(), nn * sizeof(double), hipMemcpyHostToDevice, st));
      BCHK(hipMemcpyAsync(b.F.p + k * nn, H[s], nn * sizeof(double), hipMemcpyHostToDevice, st));
      std::vector<double> occ(m, 0.0);
      for (int c = 0; c < desc[s].nocc; ++c) occ[c] = desc[s].eta;
      BCHK(hipMemcpyAsync(b.occ.p + (size_t)k * m, occ.data(), m * sizeof(double), hipMemcpyHostToDevice, st));
      const double qq = desc[s].charge * desc[s].charge;
      BCHK(hipMemcpyAsync(b.qq.p + k, &qq, sizeof(double), hipMemcpyHostToDevice, st));
      BCHK(hipMemcpyAsync(b.slot_job.p + k, &j, sizeof(int), hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(b_fill_kernel, dim3(grid1(nn)), dim3(256), 0, st, b.G.p + k * nn, nn, 0.0);
      density_one(e, b, k);
      BCHK(hipStreamSynchronize(st));   // host vectors above go out of scope
    }
  } catch (...) {
    for (auto& w : placed) {
      Bucket& b = *e.buckets[w.first];
      b.slots[w.second].job = -1;
      const int none = -1;
      (void)hipMemcpyAsync(b.slot_job.p + w.second, &none, sizeof(int), hipMemcpyHostToDevice, st);
    }
    (void)hipStreamSynchronize(st);
    e.jobs[j].status = 2;
    throw;
  }
  J.where = placed;
  J.active = true;
  e.jobs[j] = J;
  hipLaunchKernelGGL(b_job_reset_kernel, dim3(1), dim3(1), 0, st, e.d_jobs.p, j);
  const int one = 1;
  BCHK(hipMemcpyAsync(e.d_active.p + j, &one, sizeof(int), hipMemcpyHostToDevice, st));
  BCHK(hipStreamSynchronize(st));
  e.version++;
  return j;
}

void remove_job(Engine& e, int j) {
  if (j < 0 || j >= (int)e.jobs.size() || !e.jobs[j].active) throw std::runtime_error("no such active job");
  if (e.in_flight) throw std::runtime_error("jobs change only between iterations");
  JobHost& J = e.jobs[j];
  hipStream_t st = e.stream;
  for (auto& w : J.where) {
    Bucket& b = *e.buckets[w.first];
    b.slots[w.second] = SlotInfo();
    const int none = -1;
    BCHK(hipMemcpyAsync(b.slot_job.p + w.second, &none, sizeof(int), hipMemcpyHostToDevice, st));
    // a free slot stays harmless in the batched products: zero matrices
    hipLaunchKernelGGL(b_fill_kernel, dim3(grid1(b.nn())), dim3(256), 0, st, b.D.p + (size_t)w.second * b.nn(),
                       b.nn(), 0.0);
    hipLaunchKernelGGL(b_fill_kernel, dim3(grid1(b.nn())), dim3(256), 0, st, b.G.p + (size_t)w.second * b.nn(),
                       b.nn(), 0.0);
  }
  const int zero = 0;
  BCHK(hipMemcpyAsync(e.d_active.p + j, &zero, sizeof(int), hipMemcpyHostToDevice, st));
  BCHK(hipStreamSynchronize(st));
  J.active = false;
  J.status = 2;
  J.where.clear();
  e.version++;
}

Engine* as_engine(void* h) { return static_cast<Engine*>(h); }

}  // namespace

extern "C" {

void* lowdin_scf_batch_new(void* stream) {
  Engine* e = nullptr;
  try {
    e = new Engine();
    e->stream = static_cast<hipStream_t>(stream);
    if (!e->stream) {
      BCHK(hipStreamCreate(&e->stream));   // default priority
      e->own_stream = true;
    }
    BLCHK(rocblas_create_handle(&e->blas));
    BLCHK(rocblas_set_stream(e->blas, e->stream));
    BLCHK(rocblas_set_pointer_mode(e->blas, rocblas_pointer_mode_host));
    BCHK(hipEventCreateWithFlags(&e->done, hipEventDisableTiming));
    lowdin::clear_last_error();
    return e;
  } catch (const std::exception& ex) {
    delete e;
    lowdin::set_last_error(std::string("scf_batch_new: ") + ex.what());
    return nullptr;
  }
}

void lowdin_scf_batch_del(void* eng) { delete as_engine(eng); }

int lowdin_scf_batch_add(void* eng, int nspecies, void* const* lints, const lowdin_scf_species* desc, double e_pc,
                         const double* const* H, const double* const* S) {
  try {
    const int j = add_job(*as_engine(eng), nspecies, lints, desc, e_pc, H, S);
    lowdin::clear_last_error();
    return j;
  } catch (const std::exception& ex) {
    lowdin::set_last_error(std::string("scf_batch_add: ") + ex.what());
    return -1;
  }
}

int lowdin_scf_batch_remove(void* eng, int job) {
  try {
    remove_job(*as_engine(eng), job);
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& ex) {
    lowdin::set_last_error(std::string("scf_batch_remove: ") + ex.what());
    return -1;
  }
}

int lowdin_scf_batch_autotune(void* eng, int reps) {
  try {
    Engine& e = *as_engine(eng);
    if (e.in_flight) throw std::runtime_error("an iteration is in flight");
    ensure_fock_batch(e);
    if (e.fock_batch) {
      lowdin::internal_batch_autotune(e.fock_batch, e.b_D.data(), e.b_G.data(), e.b_fac.data(), e.stream, reps);
      lowdin::internal_batch_get_tuning(e.fock_batch, e.tune_kind, e.tune_target);
      e.tuned = true;
      // the slots' G buffers were overwritten (the next iteration rebuilds
      // them first); graphs hold the old task lists
      drop_graphs(e);
      e.warm_version = 0;
    }
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& ex) {
    lowdin::set_last_error(std::string("scf_batch_autotune: ") + ex.what());
    return -1;
  }
}

int lowdin_scf_batch_set_overlap_threshold(void* eng, double threshold) {
  if (!(threshold > 0.0)) {
    lowdin::set_last_error("scf_batch_set_overlap_threshold: threshold must be positive");
    return -1;
  }
  as_engine(eng)->overlap_threshold = threshold;
  lowdin::clear_last_error();
  return 0;
}

int lowdin_scf_batch_options(void* eng, int use_graph, int inline_launch) {
  Engine& e = *as_engine(eng);
  e.use_graph = use_graph != 0;
  e.inline_launch = inline_launch != 0;
  drop_graphs(e);
  e.warm_version = 0;
  return 0;
}

int lowdin_scf_batch_step_begin(void* eng) {
  try {
    step_begin(*as_engine(eng));
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& ex) {
    lowdin::set_last_error(std::string("scf_batch_step_begin: ") + ex.what());
    return -1;
  }
}

int lowdin_scf_batch_step_query(void* eng) {
  Engine& e = *as_engine(eng);
  if (!e.in_flight) return 1;
  return hipEventQuery(e.done) == hipSuccess ? 1 : 0;
}

int lowdin_scf_batch_step_end(void* eng) {
  try {
    step_end(*as_engine(eng));
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& ex) {
    lowdin::set_last_error(std::string("scf_batch_step_end: ") + ex.what());
    return -1;
  }
}

int lowdin_scf_batch_step(void* eng) {
  try {
    step_begin(*as_engine(eng));
    step_end(*as_engine(eng));
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& ex) {
    lowdin::set_last_error(std::string("scf_batch_step: ") + ex.what());
    return -1;
  }
}

int lowdin_scf_batch_status(void* eng, int job, double* energy, double* error, int* iterations) {
  Engine& e = *as_engine(eng);
  if (job < 0 || job >= (int)e.jobs.size()) {
    lowdin::set_last_error("scf_batch_status: no such job");
    return -1;
  }
  const JobHost& J = e.jobs[job];
  if (energy) *energy = J.energy;
  if (error) *error = J.error;
  if (iterations) *iterations = J.iterations;
  if (J.status == 1) lowdin::set_last_error("scf_batch: job " + std::to_string(job) + ": " + J.message);
  return J.status;
}

int lowdin_scf_batch_get(void* eng, int job, int species, int what, double* out) {
  try {
    Engine& e = *as_engine(eng);
    if (job < 0 || job >= (int)e.jobs.size() || !e.jobs[job].active) throw std::runtime_error("no such active job");
    const JobHost& J = e.jobs[job];
    if (species < 0 || species >= (int)J.where.size()) throw std::runtime_error("no such species");
    Bucket& b = *e.buckets[J.where[species].first];
    const int k = J.where[species].second;
    const double* src = what == 0 ? b.D.p : what == 1 ? b.F.p : what == 2 ? b.G.p : what == 3 ? b.Cs.p : nullptr;
    if (!src) throw std::runtime_error("what: 0 density, 1 Fock, 2 two-particle G, 3 coupling");
    BCHK(hipMemcpyAsync(out, src + (size_t)k * b.nn(), b.nn() * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    BCHK(hipStreamSynchronize(e.stream));
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& ex) {
    lowdin::set_last_error(std::string("scf_batch_get: ") + ex.what());
    return -1;
  }
}

int64_t lowdin_scf_batch_device_bytes(void* eng) { return (int64_t)as_engine(eng)->bytes; }

}  // extern "C"
