// Setup kernels of the direct Fock build: shell-pair significance (the
// reference's compute_shellpair_list, Libint2Iface.cpp:3248-3342), the
// shell-block infinity norms of D (143-161) and the final symmetrisation
// G = 0.25 (G + G^T) (616).  The class kernels live in gen/k_<class>.hip.
#include <hip/hip_runtime.h>

#include "lowdin_device.h"
#include "lowdin_kernels.h"

namespace lowdin {

__host__ __device__ constexpr int ncart_(int l) { return (l + 1) * (l + 2) / 2; }

// ----------------------------------------------- pair-list (a10) kernel
// Frobenius norm of the raw overlap block of every (s1, s2 <= s1), -1 for
// same-centre pairs (significant without a test).
__global__ void pair_norm_kernel(const ShellInfo* sh, const double* pexp, const double* pcoef, int nsh,
                                 double* norms) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long npair = (long)nsh * (nsh + 1) / 2;
  if (idx >= npair) return;
  // idx -> (s1, s2 <= s1)
  int s1 = (int)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
  while ((long)s1 * (s1 + 1) / 2 > idx) --s1;
  while ((long)(s1 + 1) * (s1 + 2) / 2 <= idx) ++s1;
  const int s2 = (int)(idx - (long)s1 * (s1 + 1) / 2);
  const ShellInfo A = sh[s1], B = sh[s2];
  if (A.x == B.x && A.y == B.y && A.z == B.z) {
    norms[idx] = -1.0;
    return;
  }
  double blk[36];
  overlap_block(A, B, pexp, pcoef, blk);
  double s = 0.0;
  for (int k = 0; k < ncart_(A.l) * ncart_(B.l); ++k) s += blk[k] * blk[k];
  norms[idx] = sqrt(s);
}

// ----------------------------------------------------- one-body integrals
// Overlap, kinetic (-1/2 nabla^2) and point-charge attraction
// (sum_C -Z_C/|r - C|) blocks of two shells with l <= 2, Obara-Saika:
// 1-D overlap recursions for S and T, the (e|0)^(m) vertical recursion plus
// horizontal transfer for V.  Normalised with norma like the reference's
// compute_1body_ints (Libint2Iface.cpp:163-217).
__device__ inline int cart_index(int lx, int ly, int lz) {
  const int l = lx + ly + lz;
  return l * (l + 1) * (l + 2) / 6 + ((l - lx) * (l - lx + 1)) / 2 + (l - lx - ly);
}

__device__ void boys_generic(int L, double T, const double* tab, double* F) {
  if (T < kBoysTmax) {
    const int i = (int)(T * kBoysInvH + 0.5);
    const double x = (double)i * kBoysH - T;
    const double* r = tab + i * kBoysCols;
    double s = r[1 + L + 7];
    for (int k = 6; k >= 0; --k) s = fma(s, x / (double)(k + 1), r[1 + L + k]);
    F[L] = s;
    double e = 1.0;
    for (int k = 6; k >= 0; --k) e = fma(e, x / (double)(k + 1), 1.0);
    const double eT = r[0] * e;
    for (int m = L - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + eT) / (double)(2 * m + 1);
  } else {
    F[0] = 0.88622692545275801365 / sqrt(T);
    const double eT = exp(-T);
    for (int m = 0; m < L; ++m) F[m + 1] = ((double)(2 * m + 1) * F[m] - eT) * (0.5 / T);
  }
}

__global__ void onebody_kernel(int kind, const ShellInfo* sh, const double* pexp, const double* pcoef,
                               const double* norma, int nsh, int n, int ncharge, const double* Z,
                               const double* Cxyz, const double* boys, double* M) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nsh * nsh) return;
  const ShellInfo A = sh[idx / nsh], B = sh[idx % nsh];
  const int la = A.l, lb = B.l, na = ncart_(la), nb = ncart_(lb);
  int ca[6][3], cb[6][3];
  {
    int k = 0;
    for (int lx = la; lx >= 0; --lx)
      for (int ly = la - lx; ly >= 0; --ly) { ca[k][0] = lx; ca[k][1] = ly; ca[k][2] = la - lx - ly; ++k; }
    k = 0;
    for (int lx = lb; lx >= 0; --lx)
      for (int ly = lb - lx; ly >= 0; --ly) { cb[k][0] = lx; cb[k][1] = ly; cb[k][2] = lb - lx - ly; ++k; }
  }
  double blk[36];
  for (int k = 0; k < 36; ++k) blk[k] = 0.0;
  const double Ax[3] = {A.x, A.y, A.z}, Bx[3] = {B.x, B.y, B.z};
  double AB[3], ab2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    AB[d] = Ax[d] - Bx[d];
    ab2 += AB[d] * AB[d];
  }
  for (int i = 0; i < A.nprim; ++i) {
    const double a = pexp[A.poff + i], cA = pcoef[A.poff + i];
    for (int j = 0; j < B.nprim; ++j) {
      const double b = pexp[B.poff + j], cB = pcoef[B.poff + j];
      const double p = a + b, o2p = 0.5 / p, kab = cA * cB * exp(-a * b / p * ab2);
      double P[3];
      for (int d = 0; d < 3; ++d) P[d] = (a * Ax[d] + b * Bx[d]) / p;
      if (kind == 1 || kind == 2) {
        double S[3][3][5];  // [axis][i <= 2][j <= 4]
        for (int d = 0; d < 3; ++d) {
          const double PA = P[d] - Ax[d], PB = P[d] - Bx[d];
          S[d][0][0] = 1.0;
          for (int ii = 0; ii < la; ++ii)
            S[d][ii + 1][0] = PA * S[d][ii][0] + (ii > 0 ? ii * o2p * S[d][ii - 1][0] : 0.0);
          for (int jj = 0; jj < lb + 2; ++jj)
            for (int ii = 0; ii <= la; ++ii)
              S[d][ii][jj + 1] = PB * S[d][ii][jj] + (ii > 0 ? ii * o2p * S[d][ii - 1][jj] : 0.0) +
                                 (jj > 0 ? jj * o2p * S[d][ii][jj - 1] : 0.0);
        }
        const double pre = kab * pow(3.14159265358979323846 / p, 1.5);
        for (int x = 0; x < na; ++x)
          for (int y = 0; y < nb; ++y) {
            double s1[3], t1[3];
            for (int d = 0; d < 3; ++d) {
              const int ii = ca[x][d], jj = cb[y][d];
              s1[d] = S[d][ii][jj];
              double d2 = -2.0 * b * (2 * jj + 1) * S[d][ii][jj] + 4.0 * b * b * S[d][ii][jj + 2];
              if (jj >= 2) d2 += jj * (jj - 1) * S[d][ii][jj - 2];
              t1[d] = -0.5 * d2;
            }
            const double v = (kind == 1) ? s1[0] * s1[1] * s1[2]
                                         : t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2];
            blk[x * nb + y] += pre * v;
          }
      } else {
        const int L = la + lb;
        for (int cI = 0; cI < ncharge; ++cI) {
          double PC[3], pc2 = 0.0;
          for (int d = 0; d < 3; ++d) {
            PC[d] = P[d] - Cxyz[3 * cI + d];
            pc2 += PC[d] * PC[d];
          }
          double F[5];
          boys_generic(L, p * pc2, boys, F);
          double V[35][5];  // (e|0)^(m), e over all components with |e| <= 4
          const double pre = -Z[cI] * 2.0 * 3.14159265358979323846 / p * kab;
          for (int m = 0; m <= L; ++m) V[0][m] = pre * F[m];
          for (int l = 1; l <= L; ++l)
            for (int ex = l; ex >= 0; --ex)
              for (int ey = l - ex; ey >= 0; --ey) {
                const int ez = l - ex - ey;
                const int e[3] = {ex, ey, ez};
                int dir = 0;
                if (ey > e[dir]) dir = 1;
                if (ez > e[dir]) dir = 2;
                int e1[3] = {ex, ey, ez};
                e1[dir] -= 1;
                const int i1 = cart_index(e1[0], e1[1], e1[2]);
                const int tgt = cart_index(ex, ey, ez);
                for (int m = 0; m <= L - l; ++m) {
                  double v = (P[dir] - Ax[dir]) * V[i1][m] - PC[dir] * V[i1][m + 1];
                  if (e1[dir] > 0) {
                    int e2[3] = {e1[0], e1[1], e1[2]};
                    e2[dir] -= 1;
                    const int i2 = cart_index(e2[0], e2[1], e2[2]);
                    v += e1[dir] * o2p * (V[i2][m] - V[i2][m + 1]);
                  }
                  V[tgt][m] = v;
                }
              }
          for (int x = 0; x < na; ++x)
            for (int y = 0; y < nb; ++y) {
              const int* av = ca[x];
              const int* bv = cb[y];
              double val;
              if (lb == 0) {
                val = V[cart_index(av[0], av[1], av[2])][0];
              } else {
                int i = 0, jdir = 0;
                // directions of the ket's unit vectors
                int dirs[2], nd = 0;
                for (int d = 0; d < 3; ++d)
                  for (int r = 0; r < bv[d]; ++r) dirs[nd++] = d;
                (void)i;
                (void)jdir;
                auto g = [&](int dx, int dy, int dz) { return V[cart_index(av[0] + dx, av[1] + dy, av[2] + dz)][0]; };
                if (nd == 1) {
                  const int d0 = dirs[0];
                  const int u[3] = {d0 == 0, d0 == 1, d0 == 2};
                  val = g(u[0], u[1], u[2]) + AB[d0] * g(0, 0, 0);
                } else {
                  const int di = dirs[0], dj = dirs[1];
                  const int ui[3] = {di == 0, di == 1, di == 2}, uj[3] = {dj == 0, dj == 1, dj == 2};
                  val = g(ui[0] + uj[0], ui[1] + uj[1], ui[2] + uj[2]) + AB[dj] * g(ui[0], ui[1], ui[2]) +
                        AB[di] * (g(uj[0], uj[1], uj[2]) + AB[dj] * g(0, 0, 0));
                }
              }
              blk[x * nb + y] += val;
            }
        }
      }
    }
  }
  for (int x = 0; x < na; ++x)
    for (int y = 0; y < nb; ++y)
      M[(size_t)(A.bf + x) * n + B.bf + y] = blk[x * nb + y] * norma[A.bf + x] * norma[B.bf + y];
}

// Shell-block infinity norms of D (Libint2Iface.cpp:143-161).
__global__ void dnorm_kernel(const double* D, int n, const ShellInfo* sh, int nsh, double* Dn) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nsh * nsh) return;
  const int s1 = idx / nsh, s2 = idx % nsh;
  const ShellInfo A = sh[s1], B = sh[s2];
  double m = 0.0;
  for (int i = 0; i < ncart_(A.l); ++i)
    for (int k = 0; k < ncart_(B.l); ++k) m = fmax(m, fabs(D[(size_t)(A.bf + i) * n + B.bf + k]));
  Dn[idx] = m;
}

// G = 0.25 (Graw + Graw^T) (Libint2Iface.cpp:616), optionally scaled.
__global__ void symmetrize_kernel(const double* Graw, int n, double scale, double* G) {
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || k >= n) return;
  G[(size_t)i * n + k] = scale * 0.25 * (Graw[(size_t)i * n + k] + Graw[(size_t)k * n + i]);
}

// Sum the per-task (quartets, primitive quartets, FP32 quartets) triples of
// one launch.
// Dnorm of every molecule of a batch: blockIdx.y is the molecule.
__global__ void dnorm_batch_kernel(const MolDesc* mols) {
  const MolDesc& M = mols[blockIdx.y];
  const int nsh = M.nshell;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nsh * nsh) return;
  const int s1 = idx / nsh, s2 = idx % nsh;
  const ShellInfo A = M.shells[s1], B = M.shells[s2];
  const int n = M.n;
  double m = 0.0;
  for (int i = 0; i < ncart_(A.l); ++i)
    for (int k = 0; k < ncart_(B.l); ++k) m = fmax(m, fabs(M.D[(size_t)(A.bf + i) * n + B.bf + k]));
  M.Dn[idx] = m;
}

// G = 0.25 (Graw + Graw^T) of every molecule of a batch (blockIdx.z).
__global__ void symmetrize_batch_kernel(const SymDesc* d) {
  const SymDesc S = d[blockIdx.z];
  const int i = blockIdx.y * blockDim.y + threadIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.n || j >= S.n) return;
  S.G[(size_t)i * S.n + j] = 0.25 * (S.Graw[(size_t)i * S.n + j] + S.Graw[(size_t)j * S.n + i]);
}

// Vector-store fill; used instead of hipMemsetAsync so that a build can be
// captured in a HIP graph.
__global__ void fill_zero_kernel(double* p, size_t count) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) p[k] = 0.0;
}

// Lower triangle of every (k, l) slice from the upper one: the
// first-quarter-transform kernels accumulate k <= l only.
__global__ void mirror_upper_kernel(double* GG, int n1, int n2) {
  const size_t nn = (size_t)n2 * n2, total = (size_t)n1 * nn;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const size_t i = e / nn, r = e - i * nn;
    const int j = (int)(r / n2), k = (int)(r - (size_t)j * n2);
    if (k < j) GG[e] = GG[i * nn + (size_t)k * n2 + j];
  }
}

// ------------------------------------------------------------ SCF helpers
// F = H + qq G (+ Cs): the Fock matrix of a species from its core
// Hamiltonian, two-particle matrix (scaled by the charge squared) and the
// summed coupling matrices.
__global__ void fock_assemble_kernel(double* F, const double* H, const double* G, double qq, const double* Cs,
                                     size_t nn) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < nn; k += stride) {
    double v = H[k];
    if (G) v += qq * G[k];
    if (Cs) v += Cs[k];
    F[k] = v;
  }
}

// E = T - T^T (n x n).
__global__ void antisym_kernel(double* E, const double* T, int n) {
  const int i = blockIdx.y * blockDim.y + threadIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || j >= n) return;
  E[(size_t)i * n + j] = T[(size_t)i * n + j] - T[(size_t)j * n + i];
}

// Acc += scale * sym(R), R holding one triangle: element (i, j) of the
// symmetric matrix is R[max(i,j) * n + min(i,j)] (the coupling build's layout).
__global__ void sym_add_kernel(double* Acc, const double* R, int n, double scale) {
  const int i = blockIdx.y * blockDim.y + threadIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || j >= n) return;
  const int r = i > j ? i : j, c = i > j ? j : i;
  Acc[(size_t)i * n + j] += scale * R[(size_t)r * n + c];
}

// out = sum_k c[k] * hist[idx[k]] over nn elements.
__global__ void lincomb_kernel(double* out, const double* hist, size_t nn, LinComb lc) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += stride) {
    double v = 0.0;
    for (int k = 0; k < lc.k; ++k) v += lc.c[k] * hist[(size_t)lc.idx[k] * nn + e];
    out[e] = v;
  }
}

// X(:, j) = U(:, col0 + j) / sqrt(w[col0 + j]), column-major n x m.
__global__ void scale_cols_kernel(double* X, const double* U, const double* w, int n, int m, int col0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n || j >= m) return;
  X[(size_t)j * n + i] = U[(size_t)(col0 + j) * n + i] / sqrt(w[col0 + j]);
}

// out[0] = max |x| over n elements (one workgroup).
__global__ __launch_bounds__(1024) void absmax_kernel(const double* x, size_t n, double* out) {
  __shared__ double s[1024];
  double m = 0.0;
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, fabs(x[i]));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// Per class launch (blockIdx.x): sum the per-task counters of its task range
// (ranges[2k], ranges[2k+1] = first task, task count) into out[3k .. 3k+2].
__global__ __launch_bounds__(1024) void reduce_counts_multi_kernel(const unsigned int* counts, const int* ranges,
                                                                   unsigned long long* out) {
  __shared__ unsigned long long sq[1024], sp[1024], s3[1024];
  const int first = ranges[2 * blockIdx.x], ntasks = ranges[2 * blockIdx.x + 1];
  const unsigned int* c = counts + 3 * (size_t)first;
  unsigned long long q = 0, p = 0, q3 = 0;
  for (int i = threadIdx.x; i < ntasks; i += blockDim.x) {
    q += c[3 * i];
    p += c[3 * i + 1];
    q3 += c[3 * i + 2];
  }
  sq[threadIdx.x] = q;
  sp[threadIdx.x] = p;
  s3[threadIdx.x] = q3;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sq[threadIdx.x] += sq[threadIdx.x + w];
      sp[threadIdx.x] += sp[threadIdx.x + w];
      s3[threadIdx.x] += s3[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[3 * blockIdx.x] = sq[0];
    out[3 * blockIdx.x + 1] = sp[0];
    out[3 * blockIdx.x + 2] = s3[0];
  }
}

__global__ __launch_bounds__(1024) void reduce_counts_kernel(const unsigned int* counts, int ntasks,
                                                             unsigned long long* out) {
  __shared__ unsigned long long sq[1024], sp[1024], s3[1024];
  unsigned long long q = 0, p = 0, q3 = 0;
  for (int i = threadIdx.x; i < ntasks; i += blockDim.x) {
    q += counts[3 * i];
    p += counts[3 * i + 1];
    q3 += counts[3 * i + 2];
  }
  sq[threadIdx.x] = q;
  sp[threadIdx.x] = p;
  s3[threadIdx.x] = q3;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sq[threadIdx.x] += sq[threadIdx.x + w];
      sp[threadIdx.x] += sp[threadIdx.x + w];
      s3[threadIdx.x] += s3[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sq[0];
    out[1] = sp[0];
    out[2] = s3[0];
  }
}

// ------------------------------------------------------------ launchers
void launch_reduce_counts_multi(const unsigned int* counts, const int* ranges, int nlaunch, unsigned long long* out,
                                hipStream_t s) {
  if (nlaunch > 0)
    hipLaunchKernelGGL(reduce_counts_multi_kernel, dim3(nlaunch), dim3(1024), 0, s, counts, ranges, out);
}

void launch_reduce_counts(const unsigned int* counts, int ntasks, unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(reduce_counts_kernel, dim3(1), dim3(1024), 0, s, counts, ntasks, out);
}

void launch_pair_norms(const ShellInfo* sh, const double* pexp, const double* pcoef, int nsh, double* norms,
                       hipStream_t s) {
  const long npair = (long)nsh * (nsh + 1) / 2;
  if (npair == 0) return;
  hipLaunchKernelGGL(pair_norm_kernel, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, s, sh, pexp, pcoef,
                     nsh, norms);
}

void launch_onebody(int kind, const ShellInfo* sh, const double* pexp, const double* pcoef, const double* norma,
                    int nsh, int n, int ncharge, const double* Z, const double* Cxyz, const double* boys, double* M,
                    hipStream_t s) {
  const int tot = nsh * nsh;
  if (tot == 0) return;
  hipLaunchKernelGGL(onebody_kernel, dim3((tot + 63) / 64), dim3(64), 0, s, kind, sh, pexp, pcoef, norma, nsh, n,
                     ncharge, Z, Cxyz, boys, M);
}

void launch_dnorm(const double* D, int n, const ShellInfo* sh, int nsh, double* Dn, hipStream_t s) {
  const int tot = nsh * nsh;
  if (tot == 0) return;
  hipLaunchKernelGGL(dnorm_kernel, dim3((tot + 255) / 256), dim3(256), 0, s, D, n, sh, nsh, Dn);
}

void launch_symmetrize(const double* Graw, int n, double scale, double* G, hipStream_t s) {
  if (n == 0) return;
  dim3 blk(16, 16), grd((n + 15) / 16, (n + 15) / 16);
  hipLaunchKernelGGL(symmetrize_kernel, grd, blk, 0, s, Graw, n, scale, G);
}

void launch_dnorm_batch(const MolDesc* mols, int nmol, int nshmax, hipStream_t s) {
  if (nmol <= 0 || nshmax <= 0) return;
  const int tot = nshmax * nshmax;
  hipLaunchKernelGGL(dnorm_batch_kernel, dim3((tot + 255) / 256, nmol), dim3(256), 0, s, mols);
}

void launch_symmetrize_batch(const SymDesc* d, int nmol, int nmax, hipStream_t s) {
  if (nmol <= 0 || nmax <= 0) return;
  const dim3 blk(32, 8), grd((nmax + 31) / 32, (nmax + 7) / 8, nmol);
  hipLaunchKernelGGL(symmetrize_batch_kernel, grd, blk, 0, s, d);
}

void launch_fill_zero(double* p, size_t count, hipStream_t s) {
  if (count == 0) return;
  const size_t blocks = std::min<size_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, count);
}

static unsigned grid_for(size_t count) { return (unsigned)std::min<size_t>((count + 255) / 256, 4096); }

void launch_mirror_upper(double* GG, int n1, int n2, hipStream_t s) {
  const size_t total = (size_t)n1 * n2 * n2;
  if (total) hipLaunchKernelGGL(mirror_upper_kernel, dim3(grid_for(total)), dim3(256), 0, s, GG, n1, n2);
}

void launch_fock_assemble(double* F, const double* H, const double* G, double qq, const double* Cs, size_t nn,
                          hipStream_t s) {
  if (nn) hipLaunchKernelGGL(fock_assemble_kernel, dim3(grid_for(nn)), dim3(256), 0, s, F, H, G, qq, Cs, nn);
}

void launch_antisym(double* E, const double* T, int n, hipStream_t s) {
  if (n <= 0) return;
  const dim3 blk(32, 8), grd((n + 31) / 32, (n + 7) / 8);
  hipLaunchKernelGGL(antisym_kernel, grd, blk, 0, s, E, T, n);
}

void launch_sym_add(double* Acc, const double* R, int n, double scale, hipStream_t s) {
  if (n <= 0) return;
  const dim3 blk(32, 8), grd((n + 31) / 32, (n + 7) / 8);
  hipLaunchKernelGGL(sym_add_kernel, grd, blk, 0, s, Acc, R, n, scale);
}

void launch_lincomb(double* out, const double* hist, size_t nn, const LinComb& lc, hipStream_t s) {
  if (nn) hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(nn)), dim3(256), 0, s, out, hist, nn, lc);
}

void launch_scale_cols(double* X, const double* U, const double* w, int n, int m, int col0, hipStream_t s) {
  if (n <= 0 || m <= 0) return;
  hipLaunchKernelGGL(scale_cols_kernel, dim3((n + 255) / 256, m), dim3(256), 0, s, X, U, w, n, m, col0);
}

void launch_absmax(const double* x, size_t n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(absmax_kernel, dim3(1), dim3(1024), 0, s, x, n, out);
}

}  // namespace lowdin
