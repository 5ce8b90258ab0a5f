// Device-resident multi-species Hartree-Fock iteration around the direct
// Fock build: the caller side of the hot path kept on the GPU.
//
// Follows openLOWDIN's SCF (src/scf/MultiSCF.f90:276-440, 730-1002;
// SingleSCF.f90:146-252; WaveFunction.f90:394-467, 817-1535, 1768-2003):
//   * orthogonaliser X = U s^{-1/2} over the overlap eigenpairs with
//     s >= OVERLAP_EIGEN_THRESHOLD (1e-8 unless set; canonical
//     orthogonalisation, WaveFunction.f90:394-467, 419);
//   * H-core guess: D from the eigenvectors of X^T H X;
//   * per iteration and species: G = J[D] + f K[D] (the direct build,
//     skipped for one particle unless requested), times q^2; the couplings
//     sum_j q_i q_j J_{i<-j}[D_j], symmetrised from the upper triangle;
//     F = H + q^2 G + C; E = E_pc + sum_s Tr(D_s (H_s + G_s/2 + C_s/2));
//   * DIIS on the orthogonalised commutator X^T (F D S - S D F) X, history 8,
//     over all species at once; diagonalisation of X^T F X (rocSOLVER
//     dsyevd), C = X C', D = eta C_occ C_occ^T.
// Every matrix stays in HBM; the host sees only scalars (traces, the DIIS
// B-matrix row, the error maximum) and solves the (k+1) x (k+1) DIIS system.
// Matrices are column-major n x n (all symmetric except X and C).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "lowdin_fock.h"
#include "lowdin_kernels.h"

namespace lowdin {
void set_last_error(const std::string& msg);
void clear_last_error();
}  // namespace lowdin

namespace {

using lowdin::kMaxDiis;

#define SCFCHK(x)                                                                                    \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define BLASCHK(x)                                                                                   \
  do {                                                                                               \
    rocblas_status s_ = (x);                                                                         \
    if (s_ != rocblas_status_success)                                                                \
      throw std::runtime_error(std::string(#x) + ": " + rocblas_status_to_string(s_));               \
  } while (0)

struct Buf {
  double* p = nullptr;
  size_t n = 0;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  Buf(Buf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  double* alloc(size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = count;
    SCFCHK(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(double)));
    return p;
  }
};

struct ScfSpecies {
  void* lint = nullptr;
  int n = 0, m = 0;
  int nocc = 0;
  double eta = 2.0, charge = -1.0, factor = -0.5;
  bool build_two = true;
  bool ghost = false;   // SCF_GHOST_SPECIES: its coupling stays out of the others' Fock matrices
  Buf H, S, X, D, G, Cs, R, F, T, Fp, W, Ework, C, E, Fh, Eh;
  Buf Cg;               // coupling from ghost species (allocated only when there is one)
  bool have_onebody = false;
};

// Per species, the scalars of one Fock step land in a device array and come
// back with one asynchronous copy: Tr(D H), Tr(D G), Tr(D C), max |error|,
// then the DIIS row <E_i, E_new> (kMaxDiis).
constexpr int kScal = 5 + kMaxDiis;   // ..., [4] info of the last diagonalisation, [5..] DIIS row

struct Scf {
  std::vector<ScfSpecies> sp;
  double e_pc = 0.0;
  double overlap_threshold = 1e-8;   // OVERLAP_EIGEN_THRESHOLD (CONTROL.f90:1513)
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  int* d_info = nullptr;
  double* d_scratch = nullptr;   // kMaxDiis doubles (DIIS row)
  double* d_sc = nullptr;        // kScal per species
  double* h_sc = nullptr;        // pinned copy
  hipEvent_t sc_ready = nullptr;
  int pending_slot = -1;         // history slot of the Fock step in flight
  // DIIS ring
  int nhist = 0, head = 0;
  double B[kMaxDiis][kMaxDiis] = {};
  double energy = 0.0, error = 0.0;
  bool started = false;
  bool own_stream = false;
  size_t bytes = 0;

  ~Scf() {
    if (blas) (void)rocblas_destroy_handle(blas);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    if (d_info) (void)hipFree(d_info);
    if (d_scratch) (void)hipFree(d_scratch);
    if (d_sc) (void)hipFree(d_sc);
    if (h_sc) (void)hipHostFree(h_sc);
    if (sc_ready) (void)hipEventDestroy(sc_ready);
  }
};

void gemm(Scf& s, rocblas_operation ta, rocblas_operation tb, int m, int n, int k, double alpha, const double* A,
          int lda, const double* B, int ldb, double beta, double* C, int ldc) {
  BLASCHK(rocblas_dgemm(s.blas, ta, tb, m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc));
}

double dot(Scf& s, size_t nn, const double* x, const double* y) {
  double r = 0.0;
  BLASCHK(rocblas_ddot(s.blas, (rocblas_int)nn, x, 1, y, 1, &r));
  return r;
}

// Symmetric eigenproblem in place: A (n x n) -> eigenvectors, W ascending.
void syevd(Scf& s, ScfSpecies& p, int n, double* A, double* W) {
  BLASCHK(rocsolver_dsyevd(s.blas, rocblas_evect_original, rocblas_fill_upper, n, A, n, W, p.Ework.p, s.d_info));
  int info = 0;
  SCFCHK(hipMemcpyAsync(&info, s.d_info, sizeof(int), hipMemcpyDeviceToHost, s.stream));
  SCFCHK(hipStreamSynchronize(s.stream));
  if (info != 0) throw std::runtime_error("rocsolver_dsyevd did not converge (info " + std::to_string(info) + ")");
}

// F' = X^T F X into Fp (m x m, leading dimension m).
void orthogonal_fock(Scf& s, ScfSpecies& p, const double* F, double* Fp) {
  const int n = p.n, m = p.m;
  const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
  gemm(s, N, N, n, m, n, 1.0, F, n, p.X.p, n, 0.0, p.T.p, n);        // T = F X (n x m)
  gemm(s, T, N, m, m, n, 1.0, p.X.p, n, p.T.p, n, 0.0, Fp, m);       // F' = X^T T
}

// C = X C', D = eta C_occ C_occ^T from the eigenvectors C' (m x m).
void density_from_vectors(Scf& s, ScfSpecies& p, const double* Cp) {
  const int n = p.n, m = p.m;
  const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
  gemm(s, N, N, n, m, m, 1.0, p.X.p, n, Cp, m, 0.0, p.C.p, n);
  if (p.nocc > 0)
    gemm(s, N, T, n, n, p.nocc, p.eta, p.C.p, n, p.C.p, n, 0.0, p.D.p, n);
  else
    lowdin::launch_fill_zero(p.D.p, (size_t)n * n, s.stream);
}

// D from F: orthogonal F', rocSOLVER dsyevd (its info lands in the species'
// scalar slot and is checked with the next Fock step), C, D.  No host wait.
void density_from_fock(Scf& s, size_t k, const double* F) {
  ScfSpecies& p = s.sp[k];
  orthogonal_fock(s, p, F, p.Fp.p);
  BLASCHK(rocsolver_dsyevd(s.blas, rocblas_evect_original, rocblas_fill_upper, p.m, p.Fp.p, p.m, p.W.p, p.Ework.p,
                           s.d_info));
  SCFCHK(hipMemcpyAsync(s.d_sc + k * kScal + 4, s.d_info, sizeof(int), hipMemcpyDeviceToDevice, s.stream));
  density_from_vectors(s, p, p.Fp.p);
}

void start(Scf& s) {
  for (ScfSpecies& p : s.sp) {
    if (!p.have_onebody) throw std::runtime_error("one-body matrices not set for every species");
    const int n = p.n;
    const size_t nn = (size_t)n * n;
    // orthogonaliser from the overlap eigenpairs
    SCFCHK(hipMemcpyAsync(p.T.p, p.S.p, nn * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
    syevd(s, p, n, p.T.p, p.W.p);
    std::vector<double> w(n);
    SCFCHK(hipMemcpyAsync(w.data(), p.W.p, n * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    SCFCHK(hipStreamSynchronize(s.stream));
    int col0 = 0;
    while (col0 < n && !(w[col0] >= s.overlap_threshold)) ++col0;
    p.m = n - col0;
    if (p.m < p.nocc) throw std::runtime_error("fewer orthogonal functions than occupied orbitals");
    lowdin::launch_scale_cols(p.X.p, p.T.p, p.W.p, n, p.m, col0, s.stream);
  }
  // H-core guess
  SCFCHK(hipMemsetAsync(s.d_sc, 0, s.sp.size() * kScal * sizeof(double), s.stream));
  for (size_t k = 0; k < s.sp.size(); ++k) density_from_fock(s, k, s.sp[k].H.p);
  s.nhist = 0;
  s.head = 0;
  s.started = true;
}

// Enqueue F, the energy terms and the DIIS error of the current densities;
// the error enters the DIIS history (with F).  external_g: the species' G
// buffers were filled by the caller (a batched build) for the current
// densities.  Nothing waits for the device; fock_end collects the scalars.
void fock_begin(Scf& s, bool external_g) {
  const size_t ns = s.sp.size();
  for (size_t k = 0; k < ns; ++k) {
    ScfSpecies& p = s.sp[k];
    const size_t nn = (size_t)p.n * p.n;
    if (!p.build_two) {
      lowdin::launch_fill_zero(p.G.p, nn, s.stream);
    } else if (!external_g) {
      if (lowdin_fock_compute_2body_device(p.lint, p.D.p, p.G.p, p.factor, s.stream, 0) != 0)
        throw std::runtime_error(std::string("two-particle build: ") + lowdin_fock_last_error());
    }
    lowdin::launch_fill_zero(p.Cs.p, nn, s.stream);
    if (p.Cg.p) lowdin::launch_fill_zero(p.Cg.p, nn, s.stream);
  }
  // couplings: one pass per species pair gives both directions
  for (size_t a = 0; a < ns; ++a)
    for (size_t b = a + 1; b < ns; ++b) {
      ScfSpecies& pa = s.sp[a];
      ScfSpecies& pb = s.sp[b];
      if (lowdin_fock_coupling_pair_device(pa.lint, pb.lint, pa.D.p, pb.D.p, pa.R.p, pb.R.p, s.stream, 0) != 0)
        throw std::runtime_error(std::string("coupling build: ") + lowdin_fock_last_error());
      const double qq = pa.charge * pb.charge;
      lowdin::launch_sym_add(pa.Cs.p, pa.R.p, pa.n, qq, s.stream);
      lowdin::launch_sym_add(pb.Cs.p, pb.R.p, pb.n, qq, s.stream);
      // the energy keeps every coupling (WaveFunction.f90:1980-1986); the
      // Fock matrix of a species leaves out the ghost's (:1502-1503)
      if (pb.ghost) lowdin::launch_sym_add(pa.Cg.p, pa.R.p, pa.n, qq, s.stream);
      if (pa.ghost) lowdin::launch_sym_add(pb.Cg.p, pb.R.p, pb.n, qq, s.stream);
    }
  const int slot = s.nhist < kMaxDiis ? s.nhist : s.head;
  const int nstored = std::min(s.nhist + 1, kMaxDiis);
  BLASCHK(rocblas_set_pointer_mode(s.blas, rocblas_pointer_mode_device));
  for (size_t k = 0; k < ns; ++k) {
    ScfSpecies& p = s.sp[k];
    const int n = p.n, m = p.m;
    const size_t nn = (size_t)n * n, mm = (size_t)m * m;
    double* sc = s.d_sc + k * kScal;
    const double qq = p.charge * p.charge;
    lowdin::launch_fock_assemble(p.F.p, p.H.p, p.G.p, qq, p.Cs.p, nn, s.stream);
    if (p.Cg.p) lowdin::launch_sym_add(p.F.p, p.Cg.p, n, -1.0, s.stream);
    BLASCHK(rocblas_ddot(s.blas, (rocblas_int)nn, p.D.p, 1, p.H.p, 1, sc + 0));
    BLASCHK(rocblas_ddot(s.blas, (rocblas_int)nn, p.D.p, 1, p.G.p, 1, sc + 1));
    BLASCHK(rocblas_ddot(s.blas, (rocblas_int)nn, p.D.p, 1, p.Cs.p, 1, sc + 2));
    // error = X^T (F D S - S D F) X; (F D S)^T = S D F for symmetric F, D, S
    const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
    const double one = 1.0, zero = 0.0;
    BLASCHK(rocblas_set_pointer_mode(s.blas, rocblas_pointer_mode_host));
    gemm(s, N, N, n, n, n, 1.0, p.F.p, n, p.D.p, n, 0.0, p.T.p, n);
    gemm(s, N, N, n, n, n, 1.0, p.T.p, n, p.S.p, n, 0.0, p.R.p, n);
    lowdin::launch_antisym(p.T.p, p.R.p, n, s.stream);
    gemm(s, T, N, m, n, n, 1.0, p.X.p, n, p.T.p, n, 0.0, p.R.p, m);     // X^T A (m x n)
    double* e = p.Eh.p + (size_t)slot * mm;
    gemm(s, N, N, m, m, n, 1.0, p.R.p, m, p.X.p, n, 0.0, e, m);        // (X^T A) X
    lowdin::launch_absmax(e, mm, sc + 3, s.stream);
    SCFCHK(hipMemcpyAsync(p.Fh.p + (size_t)slot * nn, p.F.p, nn * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
    // DIIS row: <E_i, E_slot> for every stored i (one gemv over the history)
    BLASCHK(rocblas_dgemv(s.blas, rocblas_operation_transpose, (rocblas_int)mm, nstored, &one, p.Eh.p,
                          (rocblas_int)mm, e, 1, &zero, sc + 5, 1));
    BLASCHK(rocblas_set_pointer_mode(s.blas, rocblas_pointer_mode_device));
  }
  BLASCHK(rocblas_set_pointer_mode(s.blas, rocblas_pointer_mode_host));
  SCFCHK(hipMemcpyAsync(s.h_sc, s.d_sc, ns * kScal * sizeof(double), hipMemcpyDeviceToHost, s.stream));
  SCFCHK(hipEventRecord(s.sc_ready, s.stream));
  s.pending_slot = slot;
}

// Wait for the scalars of fock_begin; energy, error, DIIS bookkeeping.
void fock_end(Scf& s) {
  if (s.pending_slot < 0) throw std::runtime_error("no Fock step in flight");
  SCFCHK(hipEventSynchronize(s.sc_ready));
  const int slot = s.pending_slot;
  s.pending_slot = -1;
  const size_t ns = s.sp.size();
  const int nstored = std::min(s.nhist + 1, kMaxDiis);
  double energy = s.e_pc, err = 0.0;
  std::vector<double> row(kMaxDiis, 0.0);
  for (size_t k = 0; k < ns; ++k) {
    const ScfSpecies& p = s.sp[k];
    const double* sc = s.h_sc + k * kScal;
    const double qq = p.charge * p.charge;
    energy += sc[0] + 0.5 * qq * sc[1] + 0.5 * sc[2];
    err = std::max(err, sc[3]);
    int info = 0;
    std::memcpy(&info, sc + 4, sizeof(int));
    if (info != 0)
      throw std::runtime_error("rocsolver_dsyevd did not converge (info " + std::to_string(info) + ")");
    for (int i = 0; i < nstored; ++i) row[i] += sc[5 + i];
  }
  for (int i = 0; i < nstored; ++i) {
    s.B[slot][i] = row[i];
    s.B[i][slot] = row[i];
  }
  if (s.nhist < kMaxDiis)
    s.nhist++;
  else
    s.head = (s.head + 1) % kMaxDiis;
  s.energy = energy;
  s.error = err;
}

void fock(Scf& s, bool external_g) {
  fock_begin(s, external_g);
  fock_end(s);
}

// Solve [B -1; -1 0] [c; l] = [0; -1] (Gaussian elimination, partial
// pivoting); false if singular.
bool diis_solve(int m, const double B[kMaxDiis][kMaxDiis], const int* idx, double* c) {
  const int N = m + 1;
  std::vector<double> A((size_t)N * (N + 1), 0.0);
  auto at = [&](int r, int col) -> double& { return A[(size_t)r * (N + 1) + col]; };
  for (int i = 0; i < m; ++i) {
    for (int j = 0; j < m; ++j) at(i, j) = B[idx[i]][idx[j]];
    at(i, m) = -1.0;
    at(m, i) = -1.0;
  }
  at(m, m) = 0.0;
  at(m, N) = -1.0;
  for (int col = 0; col < N; ++col) {
    int piv = col;
    for (int r = col + 1; r < N; ++r)
      if (std::fabs(at(r, col)) > std::fabs(at(piv, col))) piv = r;
    if (at(piv, col) == 0.0) return false;
    if (piv != col)
      for (int q = 0; q <= N; ++q) std::swap(at(piv, q), at(col, q));
    for (int r = col + 1; r < N; ++r) {
      const double f = at(r, col) / at(col, col);
      for (int q = col; q <= N; ++q) at(r, q) -= f * at(col, q);
    }
  }
  std::vector<double> x(N);
  for (int r = N - 1; r >= 0; --r) {
    double v = at(r, N);
    for (int q = r + 1; q < N; ++q) v -= at(r, q) * x[q];
    x[r] = v / at(r, r);
  }
  for (int i = 0; i < m; ++i) c[i] = x[i];
  return true;
}

// DIIS extrapolation over the stored Fock matrices into F.
void extrapolate(Scf& s) {
  const int m = s.nhist;
  int idx[kMaxDiis];
  // chronological order: oldest first
  for (int i = 0; i < m; ++i) idx[i] = (m < kMaxDiis) ? i : (s.head + i) % kMaxDiis;
  double c[kMaxDiis] = {};
  if (!diis_solve(m, s.B, idx, c)) {
    for (int i = 0; i < m; ++i) c[i] = 0.0;
    c[m - 1] = 1.0;
  }
  lowdin::LinComb lc{};
  lc.k = m;
  for (int i = 0; i < m; ++i) {
    lc.c[i] = c[i];
    lc.idx[i] = idx[i];
  }
  for (size_t k = 0; k < s.sp.size(); ++k) {
    ScfSpecies& p = s.sp[k];
    lowdin::launch_lincomb(p.F.p, p.Fh.p, (size_t)p.n * p.n, lc, s.stream);
  }
}

void update(Scf& s) {
  extrapolate(s);
  for (size_t k = 0; k < s.sp.size(); ++k) density_from_fock(s, k, s.sp[k].F.p);
}

// Grouped diagonalisation for many SCF objects on one stream: the
// orthogonalised Fock matrices of equal dimension m are packed into one
// strided array and diagonalised by one rocsolver_dsyevd_strided_batched.
struct GroupWork {
  Buf A, W, E;
  int* info = nullptr;
  int cap = 0;
  ~GroupWork() {
    if (info) (void)hipFree(info);
  }
};

void update_many(const std::vector<Scf*>& scfs) {
  static std::map<int, GroupWork> work;   // per dimension m, kept for the process
  hipStream_t st = scfs[0]->stream;
  for (Scf* s : scfs)
    if (s->stream != st) {
      for (Scf* t : scfs) update(*t);
      return;
    }
  std::map<int, std::vector<std::pair<Scf*, size_t>>> groups;
  for (Scf* s : scfs) {
    extrapolate(*s);
    for (size_t k = 0; k < s->sp.size(); ++k) groups[s->sp[k].m].push_back({s, k});
  }
  Scf& s0 = *scfs[0];
  for (auto& g : groups) {
    const int m = g.first;
    const int cnt = (int)g.second.size();
    const size_t mm = (size_t)m * m;
    GroupWork& w = work[m];
    if (w.cap < cnt) {
      SCFCHK(hipStreamSynchronize(st));
      w.A.alloc((size_t)cnt * mm);
      w.W.alloc((size_t)cnt * m);
      w.E.alloc((size_t)cnt * m);
      if (w.info) (void)hipFree(w.info);
      w.info = nullptr;
      SCFCHK(hipMalloc((void**)&w.info, cnt * sizeof(int)));
      w.cap = cnt;
    }
    for (int i = 0; i < cnt; ++i) {
      Scf& s = *g.second[i].first;
      ScfSpecies& p = s.sp[g.second[i].second];
      orthogonal_fock(s, p, p.F.p, w.A.p + (size_t)i * mm);
    }
    BLASCHK(rocsolver_dsyevd_strided_batched(s0.blas, rocblas_evect_original, rocblas_fill_upper, m, w.A.p, m,
                                             (rocblas_stride)mm, w.W.p, m, w.E.p, m, w.info, cnt));
    for (int i = 0; i < cnt; ++i) {
      Scf& s = *g.second[i].first;
      const size_t k = g.second[i].second;
      ScfSpecies& p = s.sp[k];
      SCFCHK(hipMemcpyAsync(s.d_sc + k * kScal + 4, w.info + i, sizeof(int), hipMemcpyDeviceToDevice, st));
      density_from_vectors(s, p, w.A.p + (size_t)i * mm);
    }
  }
}

Scf* as_scf(void* h) { return static_cast<Scf*>(h); }

}  // namespace

extern "C" {

void* lowdin_scf_new(int nspecies, void* const* lints, const lowdin_scf_species* desc, double e_pc, void* stream) {
  Scf* s = nullptr;
  try {
    if (nspecies < 1 || !lints || !desc) throw std::runtime_error("an SCF needs at least one species");
    s = new Scf();
    s->e_pc = e_pc;
    s->stream = static_cast<hipStream_t>(stream);
    if (!s->stream) {
      SCFCHK(hipStreamCreate(&s->stream));  // default priority
      s->own_stream = true;
    }
    BLASCHK(rocblas_create_handle(&s->blas));
    BLASCHK(rocblas_set_stream(s->blas, s->stream));
    BLASCHK(rocblas_set_pointer_mode(s->blas, rocblas_pointer_mode_host));
    SCFCHK(hipMalloc((void**)&s->d_info, sizeof(int)));
    SCFCHK(hipMalloc((void**)&s->d_scratch, kMaxDiis * sizeof(double)));
    SCFCHK(hipMalloc((void**)&s->d_sc, (size_t)nspecies * kScal * sizeof(double)));
    SCFCHK(hipHostMalloc((void**)&s->h_sc, (size_t)nspecies * kScal * sizeof(double), hipHostMallocDefault));
    SCFCHK(hipEventCreateWithFlags(&s->sc_ready, hipEventDisableTiming));
    s->sp.resize(nspecies);
    for (int k = 0; k < nspecies; ++k) {
      ScfSpecies& p = s->sp[k];
      p.lint = lints[k];
      p.n = LibintInterface_get_nbasis(lints[k]);
      p.m = p.n;
      p.nocc = desc[k].nocc;
      p.eta = desc[k].eta;
      p.charge = desc[k].charge;
      p.factor = desc[k].factor;
      p.build_two = desc[k].build_two != 0;
      p.ghost = desc[k].ghost != 0;
      const size_t nn = (size_t)p.n * p.n;
      for (Buf* b : {&p.H, &p.S, &p.X, &p.D, &p.G, &p.Cs, &p.R, &p.F, &p.T, &p.Fp, &p.C, &p.E}) b->alloc(nn);
      p.W.alloc(p.n);
      p.Ework.alloc(p.n);
      p.Fh.alloc((size_t)kMaxDiis * nn);
      p.Eh.alloc((size_t)kMaxDiis * nn);
      s->bytes += (12 + 2 * kMaxDiis) * nn * sizeof(double) + 2 * p.n * sizeof(double);
    }
    bool any_ghost = false;
    for (const ScfSpecies& p : s->sp) any_ghost = any_ghost || p.ghost;
    if (any_ghost)
      for (ScfSpecies& p : s->sp) {
        p.Cg.alloc((size_t)p.n * p.n);
        s->bytes += (size_t)p.n * p.n * sizeof(double);
      }
    lowdin::clear_last_error();
    return s;
  } catch (const std::exception& e) {
    delete s;
    lowdin::set_last_error(std::string("scf_new: ") + e.what());
    return nullptr;
  }
}

void lowdin_scf_del(void* scf) { delete as_scf(scf); }

int lowdin_scf_set_onebody(void* scf, int species, const double* H, const double* S) {
  try {
    Scf& s = *as_scf(scf);
    if (species < 0 || species >= (int)s.sp.size()) throw std::runtime_error("species index out of range");
    ScfSpecies& p = s.sp[species];
    const size_t nn = (size_t)p.n * p.n;
    SCFCHK(hipMemcpyAsync(p.H.p, H, nn * sizeof(double), hipMemcpyHostToDevice, s.stream));
    SCFCHK(hipMemcpyAsync(p.S.p, S, nn * sizeof(double), hipMemcpyHostToDevice, s.stream));
    SCFCHK(hipStreamSynchronize(s.stream));
    p.have_onebody = true;
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("scf_set_onebody: ") + e.what());
    return -1;
  }
}

int lowdin_scf_set_overlap_threshold(void* scf, double threshold) {
  try {
    if (!(threshold > 0.0)) throw std::runtime_error("threshold must be positive");
    as_scf(scf)->overlap_threshold = threshold;
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("scf_set_overlap_threshold: ") + e.what());
    return -1;
  }
}

int lowdin_scf_start(void* scf) {
  try {
    start(*as_scf(scf));
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("scf_start: ") + e.what());
    return -1;
  }
}

int lowdin_scf_fock(void* scf, int external_g, double* energy, double* error) {
  try {
    Scf& s = *as_scf(scf);
    if (!s.started) throw std::runtime_error("lowdin_scf_start first");
    fock(s, external_g != 0);
    if (energy) *energy = s.energy;
    if (error) *error = s.error;
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("scf_fock: ") + e.what());
    return -1;
  }
}

int lowdin_scf_fock_begin(void* scf, int external_g) {
  try {
    Scf& s = *as_scf(scf);
    if (!s.started) throw std::runtime_error("lowdin_scf_start first");
    fock_begin(s, external_g != 0);
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("scf_fock_begin: ") + e.what());
    return -1;
  }
}

int lowdin_scf_fock_end(void* scf, double* energy, double* error) {
  try {
    Scf& s = *as_scf(scf);
    fock_end(s);
    if (energy) *energy = s.energy;
    if (error) *error = s.error;
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("scf_fock_end: ") + e.what());
    return -1;
  }
}

int lowdin_scf_update_many(void* const* scfs, int count) {
  try {
    if (count <= 0) return 0;
    std::vector<Scf*> v;
    for (int i = 0; i < count; ++i) v.push_back(as_scf(scfs[i]));
    update_many(v);
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("scf_update_many: ") + e.what());
    return -1;
  }
}

int lowdin_scf_update(void* scf) {
  try {
    update(*as_scf(scf));
    lowdin::clear_last_error();
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("scf_update: ") + e.what());
    return -1;
  }
}

int lowdin_scf_run(void* scf, int max_iter, double e_tol, double d_tol, double* energy, int* iterations) {
  try {
    Scf& s = *as_scf(scf);
    if (!s.started) start(s);
    double e_old = 0.0;
    int it = 0;
    bool have_old = false;
    for (; it < max_iter; ++it) {
      fock(s, false);
      if (have_old && std::fabs(s.energy - e_old) < e_tol && s.error < d_tol) break;
      e_old = s.energy;
      have_old = true;
      update(s);
    }
    if (energy) *energy = s.energy;
    if (iterations) *iterations = it + 1;
    lowdin::clear_last_error();
    return it < max_iter ? 0 : 1;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("scf_run: ") + e.what());
    return -1;
  }
}

// what: 0 density, 1 Fock, 2 two-particle G (unscaled), 3 coupling sum.
int lowdin_scf_get(void* scf, int species, int what, double* out) {
  try {
    Scf& s = *as_scf(scf);
    if (species < 0 || species >= (int)s.sp.size()) throw std::runtime_error("species index out of range");
    ScfSpecies& p = s.sp[species];
    const double* src = what == 0 ? p.D.p : what == 1 ? p.F.p : what == 2 ? p.G.p : what == 3 ? p.Cs.p : nullptr;
    if (!src) throw std::runtime_error("unknown matrix");
    SCFCHK(hipMemcpyAsync(out, src, (size_t)p.n * p.n * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    SCFCHK(hipStreamSynchronize(s.stream));
    return 0;
  } catch (const std::exception& e) {
    lowdin::set_last_error(std::string("scf_get: ") + e.what());
    return -1;
  }
}

int lowdin_scf_buffers(void* scf, int species, double** d_density, double** d_two, double* factor) {
  Scf& s = *as_scf(scf);
  if (species < 0 || species >= (int)s.sp.size()) return -1;
  ScfSpecies& p = s.sp[species];
  if (d_density) *d_density = p.D.p;
  if (d_two) *d_two = p.G.p;
  if (factor) *factor = p.factor;
  return p.build_two ? 1 : 0;
}

int64_t lowdin_scf_device_bytes(void* scf) { return (int64_t)as_scf(scf)->bytes; }

void* lowdin_scf_stream(void* scf) { return as_scf(scf)->stream; }

}  // extern "C"
