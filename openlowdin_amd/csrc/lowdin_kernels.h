// Kernel argument blocks and launchers shared by the host library and the
// HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "lowdin_device.h"

namespace lowdin {

// One molecule (species) of a batched intra-species build.
struct MolDesc {
  const PairInfo* pairs;   // significant pairs, grouped by pair class
  const double* prim;      // PrimPair records
  const float* primf;      // the same records in FP32 (mixed precision)
  const double* D;         // n x n density
  double* Dn;              // nshell x nshell block infinity norms (dnorm_batch writes them)
  const ShellInfo* shells;
  const double* norma;
  double* Graw;            // n x n raw accumulator (symmetrised afterwards)
  int off[8];              // pair-class offsets into pairs
  int n;
  int nshell;
  int screen;              // 1: Dnorm*Q*Q < eps screen (electrons)
  int mixed;               // 1: quartets with Dnorm*Q*Q < tau in FP32
  double half_factor;      // 0.5 * exchange factor
  double tau;
};

// A class launch over the tasks of one or many molecules; Task.flags is the
// task's molecule index into mols.
struct JKArgs {
  const Task* tasks;
  int ntasks;
  const MolDesc* mols;
  const double* boys;
  const float* boysf;      // FP32 Boys table
  unsigned int* counts;    // per task: quartets, primitive quartets, FP32 quartets (optional)
  int cbra, cket;          // pair classes of the bra and ket lists
  int same_class;          // bra and ket lists are the same class list
  int lds_n;               // largest n of the launch (LDS strip sizing)
  int block;               // strip kernel: threads per workgroup (0: 256)
  double eps;
  int any_mixed;           // some molecule of the launch runs mixed precision
};

// Per molecule of a batched symmetrisation.
struct SymDesc {
  const double* Graw;
  double* G;
  int n;
  int pad;
};

struct CouplingArgs {
  const Task* tasks;
  int ntasks;
  const PairInfo* bra;     // pairs of the bra species
  const PairInfo* ket;     // pairs of the ket species
  const double* bprim;
  const double* kprim;
  const double* nbra_norma;
  const double* nket_norma;
  const double* dbra;      // density of the bra species (for rket)
  const double* dket;      // density of the ket species (for rbra)
  double* rbra;            // result on the bra species (may be null)
  double* rket;            // result on the ket species (may be null)
  const double* boys;
  int nbra;
  int nket;
  double drop;             // raw-integral drop (1e-10)
};

// Store kernels of classes with at least this many integrals per quartet
// ((dd|dd)) keep the raw block in a buffer in HBM owned by the library (one
// block per launched thread, StoreArgs::blk) instead of private memory, so
// that their private segment stays within that of the Fock-build kernels;
// launches of those classes are split into chunks of at most this many
// tasks (256 tasks: 16384 threads, 170 MB of staging).
constexpr int kStoreStagedNQ = 1296;
constexpr int kStoreStagedTasks = 256;

// Integral store (mode 0) or first-quarter transform (mode 1) over the
// intra-species quartets of one species.
//   mode 0: A[i][j][k][l] (n^4) = (ij|kl) * norma^4 for |raw| > drop, all
//           eight index images; other elements are left as they are.
//   mode 1: GG[j][k][l] (n^3) += sum_i c[i] (ij|kl), k <= l only (the other
//           triangle is mirrored afterwards); GG must start at zero.
struct StoreArgs {
  const Task* tasks;
  int ntasks;
  const PairInfo* bra;     // pairs of the bra class list
  const PairInfo* ket;     // pairs of the ket class list
  const double* prim;
  const double* norma;
  const double* Dn;        // nshell x nshell block norms of D (screen only)
  const double* c;         // mode 1: row p of the coefficient matrix (n)
  double* out;
  const double* boys;
  double* blk;             // classes with kStoreStagedNQ integrals per block: one raw block per thread
  int n;
  int nshell;
  int same_class;          // bra and ket lists are the same class list
  int screen;              // Dnorm * Q * Q < eps screen (electrons)
  int mode;
  int pad;
  double eps;
  double drop;             // kept if |raw| > drop
};

// The same over the quartets (pair of species X | pair of species Y) of an
// inter-species pair, X the species whose shells run the outer loops.
//   mode 0: A[x1][x2][y1][y2] = (x1x2|y1y2) * norma^4 for |raw| >= drop,
//           strides (sx1, sx2, sy1, 1); four images (x1 <-> x2, y1 <-> y2).
//   mode 1: GG[v][l][s] (n_T x n_O x n_O) += sum_u c[u] (uv|ls), u, v of
//           the transformed species T, l <= s of the other species O.
struct CouplingStoreArgs {
  const Task* tasks;
  int ntasks;
  const PairInfo* bra;
  const PairInfo* ket;
  const double* bprim;
  const double* kprim;
  const double* nbra_norma;
  const double* nket_norma;
  const double* c;         // mode 1: row p of T's coefficient matrix
  double* out;
  const double* boys;
  double* blk;             // as StoreArgs::blk
  int bra_is_x;            // mode 0: bra pair belongs to X
  int bra_is_t;            // mode 1: bra pair belongs to T
  int mode;
  int n_o;                 // mode 1: basis size of O
  size_t sx1, sx2, sy1;    // mode 0 strides
  double drop;             // dropped if |raw| < drop
};

// Contracted Gaussian-type geminal sum_k c[k] exp(-g[k] r12^2).
constexpr int kMaxGeminal = 64;
struct G12Params {
  int n;
  int pad;
  double c[kMaxGeminal];
  double g[kMaxGeminal];
};

// Intra-species G12 J/K build (compute_g12_direct): the Fock build's
// quartet set without the density screen; Graw (n x n) accumulates the six
// updates with degeneracy and 0.5 * factor exchange weight, symmetrised
// afterwards as 0.25 (Graw + Graw^T).
struct G12Args {
  const Task* tasks;
  int ntasks;
  int n;
  const PairInfo* bra;
  const PairInfo* ket;
  const double* prim;
  const double* norma;
  const double* D;
  double* Graw;
  int same_class;
  int pad;
  double half_factor;
  G12Params gem;
};

// Inter-species G12 coupling (compute_g12inter_direct): the coupling
// build's quartets, element sets and drop with the geminal operator.
struct G12CouplingArgs {
  CouplingArgs cp;
  G12Params gem;
};

// Raw overlap block <A|B> (row-major, A's components slowest) of two shells
// with l <= 2, from primitives whose coefficients already carry N(alpha,l).
// Obara-Saika 1-D recursion S(i+1,j) = PA S(i,j) + (i S(i-1,j) + j S(i,j-1))/(2p).
__host__ __device__ inline void overlap_block(const ShellInfo& A, const ShellInfo& B, const double* pexp,
                                              const double* pcoef, double* blk) {
  const int la = A.l, lb = B.l;
  const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
  int ca[6][3], cb[6][3];
  {
    int k = 0;
    for (int lx = la; lx >= 0; --lx)
      for (int ly = la - lx; ly >= 0; --ly) { ca[k][0] = lx; ca[k][1] = ly; ca[k][2] = la - lx - ly; ++k; }
    k = 0;
    for (int lx = lb; lx >= 0; --lx)
      for (int ly = lb - lx; ly >= 0; --ly) { cb[k][0] = lx; cb[k][1] = ly; cb[k][2] = lb - lx - ly; ++k; }
  }
  for (int k = 0; k < na * nb; ++k) blk[k] = 0.0;
  const double Ax[3] = {A.x, A.y, A.z}, Bx[3] = {B.x, B.y, B.z};
  double ab2 = 0.0;
  for (int d = 0; d < 3; ++d) ab2 += (Ax[d] - Bx[d]) * (Ax[d] - Bx[d]);
  for (int i = 0; i < A.nprim; ++i) {
    const double a = pexp[A.poff + i], ca_ = pcoef[A.poff + i];
    for (int j = 0; j < B.nprim; ++j) {
      const double b = pexp[B.poff + j], cb_ = pcoef[B.poff + j];
      const double p = a + b, o2p = 0.5 / p;
      const double pre = ca_ * cb_ * exp(-a * b / p * ab2) * pow(3.14159265358979323846 / p, 1.5);
      double S[3][3][3];  // [axis][i][j], i, j <= 2
      for (int d = 0; d < 3; ++d) {
        const double P = (a * Ax[d] + b * Bx[d]) / p;
        const double PA = P - Ax[d], PB = P - Bx[d];
        S[d][0][0] = 1.0;
        for (int ii = 0; ii < la; ++ii)
          S[d][ii + 1][0] = PA * S[d][ii][0] + (ii > 0 ? ii * o2p * S[d][ii - 1][0] : 0.0);
        for (int jj = 0; jj < lb; ++jj)
          for (int ii = 0; ii <= la; ++ii)
            S[d][ii][jj + 1] = PB * S[d][ii][jj] + (ii > 0 ? ii * o2p * S[d][ii - 1][jj] : 0.0) +
                               (jj > 0 ? jj * o2p * S[d][ii][jj - 1] : 0.0);
      }
      for (int x = 0; x < na; ++x)
        for (int y = 0; y < nb; ++y)
          blk[x * nb + y] += pre * S[0][ca[x][0]][cb[y][0]] * S[1][ca[x][1]][cb[y][1]] * S[2][ca[x][2]][cb[y][2]];
    }
  }
}

void launch_pair_norms(const ShellInfo* sh, const double* pexp, const double* pcoef, int nsh, double* norms,
                       hipStream_t s);
void launch_dnorm(const double* D, int n, const ShellInfo* sh, int nsh, double* Dn, hipStream_t s);
void launch_onebody(int kind, const ShellInfo* sh, const double* pexp, const double* pcoef, const double* norma,
                    int nsh, int n, int ncharge, const double* Z, const double* Cxyz, const double* boys, double* M,
                    hipStream_t s);
void launch_symmetrize(const double* Graw, int n, double scale, double* G, hipStream_t s);
// DIIS extrapolation weights: out = sum_k c[k] * hist[idx[k]].
constexpr int kMaxDiis = 8;
struct LinComb {
  double c[kMaxDiis];
  int idx[kMaxDiis];
  int k;
};
void launch_fock_assemble(double* F, const double* H, const double* G, double qq, const double* Cs, size_t nn,
                          hipStream_t s);
void launch_antisym(double* E, const double* T, int n, hipStream_t s);
void launch_sym_add(double* Acc, const double* R, int n, double scale, hipStream_t s);
void launch_lincomb(double* out, const double* hist, size_t nn, const LinComb& lc, hipStream_t s);
void launch_absmax(const double* x, size_t n, double* out, hipStream_t s);
void launch_scale_cols(double* X, const double* U, const double* w, int n, int m, int col0, hipStream_t s);
void launch_symmetrize_batch(const SymDesc* d, int nmol, int nmax, hipStream_t s);
void launch_fill_zero(double* p, size_t count, hipStream_t s);
// GG[i][k][j] = GG[i][j][k] for j < k, i < n1, j, k < n2.
void launch_mirror_upper(double* GG, int n1, int n2, hipStream_t s);
void launch_dnorm_batch(const MolDesc* mols, int nmol, int nshmax, hipStream_t s);
void launch_reduce_counts(const unsigned int* counts, int ntasks, unsigned long long* out, hipStream_t s);  // 3 per task
void launch_reduce_counts_multi(const unsigned int* counts, const int* ranges, int nlaunch, unsigned long long* out,
                                hipStream_t s);

}  // namespace lowdin
