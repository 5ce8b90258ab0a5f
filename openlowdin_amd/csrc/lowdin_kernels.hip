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

// Sum the per-task (quartets, primitive quartets) pairs of one launch.
__global__ __launch_bounds__(1024) void reduce_counts_kernel(const unsigned int* counts, int ntasks,
                                                             unsigned long long* out) {
  __shared__ unsigned long long sq[1024], sp[1024];
  unsigned long long q = 0, p = 0;
  for (int i = threadIdx.x; i < ntasks; i += blockDim.x) {
    q += counts[2 * i];
    p += counts[2 * i + 1];
  }
  sq[threadIdx.x] = q;
  sp[threadIdx.x] = p;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sq[threadIdx.x] += sq[threadIdx.x + w];
      sp[threadIdx.x] += sp[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sq[0];
    out[1] = sp[0];
  }
}

// ------------------------------------------------------------ launchers
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

}  // namespace lowdin
