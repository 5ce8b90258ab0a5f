// Templates of the integral-store and first-quarter-transform kernels: the
// MEMORY and integral-transformation variants of the two-particle path.
// They share the generated quartet evaluation (Obara-Saika VRR + HGP HRR)
// and the work decomposition of the Fock-build kernels: a wavefront owns
// one bra shell pair and up to 64 ket shell pairs, one per lane.
//
// Reference behaviour reproduced (openLOWDIN src/ints/Libint2Iface.cpp):
//   compute_2body_directAll (2552-2742): the Fock build's quartet set and
//     screen (2629-2669), integrals with |raw| > 1e-10 normalised by norma
//     and written to all eight index images (2698-2717);
//   compute_2body_directIT (620-921): GG[j][k][l] += C(p,i) (ij|kl) over
//     the distinct index permutations, accumulated into k <= l (803-851),
//     then the k > l triangle mirrored (864-874);
//   compute_coupling_directAll (2744-2908): every pair s1 <= s2 of both
//     species, no screen, elements bf1 <= bf2 / obf1 <= obf2 with
//     |raw| >= 1e-10 and their four images (2864-2880);
//   compute_coupling_directIT (1296-1622): the same quartets transformed
//     on the index of the calling species (1455-1545), mirrored (1562-1572).
#pragma once
#include "eri_core.h"

namespace lowdin {

// Blocks up to this many integrals get fully unrolled element loops.
constexpr int kStoreFullUnroll = 216;

// ------------------------------------------------- intra-species store / IT
template <int LA, int LB, int LC, int LD, int MODE>
__global__ __launch_bounds__(256) void store_kernel(StoreArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  // Element loops: fully unrolled for small blocks; for large ones (the
  // block lives in private memory anyway) the two outer levels stay loops,
  // which keeps code size and compile time bounded.
  constexpr bool kFull = NA * NB * NC * ND <= kStoreFullUnroll;
  constexpr int U1 = kFull ? 6 : 1;
  const int wid = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (wid >= a.ntasks) return;
  const Task t = a.tasks[wid];
  const int j = t.j0 + (int)(threadIdx.x & 63);
  if (j >= t.j1) return;
  const PairInfo P = a.bra[t.bra];
  const PairInfo K = a.ket[j];
  if (a.screen && quartet_dnorm((gcptr<double>)a.Dn, a.nshell, P, K) * P.Q * K.Q < a.eps) return;
  // the raw block: private memory, or for the largest class this thread's
  // slot of the library's staging buffer (written and read by this lane only)
  constexpr int NQ = NA * NB * NC * ND;
  double priv[NQ >= kStoreStagedNQ ? 1 : NQ];
  double* out = NQ >= kStoreStagedNQ ? a.blk + ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * NQ : priv;
  contracted_quartet<LA, LB, LC, LD>(a.prim + (size_t)P.poff * kPrimPairStride, P.nprim,
                                     a.prim + (size_t)K.poff * kPrimPairStride, K.nprim, a.boys,
                                     P.ABx, P.ABy, P.ABz, K.ABx, K.ABy, K.ABz, out);
  double na[NA], nb[NB], nc[NC], nd[ND];
#pragma unroll
  for (int i = 0; i < NA; ++i) na[i] = a.norma[P.bfa + i];
#pragma unroll
  for (int i = 0; i < NB; ++i) nb[i] = a.norma[P.bfb + i];
#pragma unroll
  for (int i = 0; i < NC; ++i) nc[i] = a.norma[K.bfa + i];
#pragma unroll
  for (int i = 0; i < ND; ++i) nd[i] = a.norma[K.bfb + i];
  // drop (|raw| > drop kept) and normalise, applied where each element is
  // read (the raw block is not rewritten: a second pass over it would keep
  // it live in private memory beside the recursion's own)
  const double drop = a.drop;
  auto kept = [&](int ia, int ib, int ic, int id) {
    const double v = out[((ia * NB + ib) * NC + ic) * ND + id];
    return fabs(v) > drop ? v * na[ia] * nb[ib] * nc[ic] * nd[id] : 0.0;
  };
  const size_t n = (size_t)a.n, n2 = n * n;
  if constexpr (MODE == 0) {
    // All eight images of every kept integral.  Images that coincide are
    // written by this lane only, so the result does not depend on timing.
    double* A = a.out;
    const size_t n3 = n2 * n;
#pragma unroll U1
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll U1
      for (int ib = 0; ib < NB; ++ib)
#pragma unroll
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
          for (int id = 0; id < ND; ++id) {
            const double v = kept(ia, ib, ic, id);
            if (v == 0.0) continue;
            const size_t i = P.bfa + ia, jj = P.bfb + ib, k = K.bfa + ic, l = K.bfb + id;
            A[i * n3 + jj * n2 + k * n + l] = v;
            A[jj * n3 + i * n2 + k * n + l] = v;
            A[i * n3 + jj * n2 + l * n + k] = v;
            A[jj * n3 + i * n2 + l * n + k] = v;
            A[k * n3 + l * n2 + i * n + jj] = v;
            A[k * n3 + l * n2 + jj * n + i] = v;
            A[l * n3 + k * n2 + i * n + jj] = v;
            A[l * n3 + k * n2 + jj * n + i] = v;
          }
  } else {
    // First-quarter transform.  Each distinct shell-level permutation of the
    // quartet contributes once; within a pair of one shell the block already
    // holds both orders, so only its k <= l elements feed a (k, l) slot.
    const gptr<double> GG = (gptr<double>)a.out;
    const gcptr<double> c = (gcptr<double>)a.c;
    const bool sab = P.sa != P.sb, scd = K.sa != K.sb;
    const bool spk = !(a.same_class && j == t.bra);
    // (ab|cd): i in A, j in B, {k, l} = {c, d}
#pragma unroll U1
    for (int ib = 0; ib < NB; ++ib)
#pragma unroll U1
      for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
          const int fc = K.bfa + ic, fd = K.bfb + id;
          if (!scd && fc > fd) continue;
          double s = 0.0;
#pragma unroll
          for (int ia = 0; ia < NA; ++ia) s += c[P.bfa + ia] * kept(ia, ib, ic, id);
          if (s != 0.0)
            global_add(GG + (size_t)(P.bfb + ib) * n2 + (size_t)min(fc, fd) * n + max(fc, fd), s);
        }
    // (ba|cd)
    if (sab) {
#pragma unroll U1
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll U1
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
          for (int id = 0; id < ND; ++id) {
            const int fc = K.bfa + ic, fd = K.bfb + id;
            if (!scd && fc > fd) continue;
            double s = 0.0;
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) s += c[P.bfb + ib] * kept(ia, ib, ic, id);
            if (s != 0.0)
              global_add(GG + (size_t)(P.bfa + ia) * n2 + (size_t)min(fc, fd) * n + max(fc, fd), s);
          }
    }
    if (!spk) return;
    // (cd|ab)
#pragma unroll U1
    for (int id = 0; id < ND; ++id)
#pragma unroll U1
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
          const int fa = P.bfa + ia, fb = P.bfb + ib;
          if (!sab && fa > fb) continue;
          double s = 0.0;
#pragma unroll
          for (int ic = 0; ic < NC; ++ic) s += c[K.bfa + ic] * kept(ia, ib, ic, id);
          if (s != 0.0)
            global_add(GG + (size_t)(K.bfb + id) * n2 + (size_t)min(fa, fb) * n + max(fa, fb), s);
        }
    // (dc|ab)
    if (scd) {
#pragma unroll U1
      for (int ic = 0; ic < NC; ++ic)
#pragma unroll U1
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
          for (int ib = 0; ib < NB; ++ib) {
            const int fa = P.bfa + ia, fb = P.bfb + ib;
            if (!sab && fa > fb) continue;
            double s = 0.0;
#pragma unroll
            for (int id = 0; id < ND; ++id) s += c[K.bfb + id] * kept(ia, ib, ic, id);
            if (s != 0.0)
              global_add(GG + (size_t)(K.bfa + ic) * n2 + (size_t)min(fa, fb) * n + max(fa, fb), s);
          }
    }
  }
}

// ------------------------------------------------- inter-species store / IT
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void coupling_store_kernel(CouplingStoreArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  // Element loops: fully unrolled for small blocks; for large ones (the
  // block lives in private memory anyway) the two outer levels stay loops,
  // which keeps code size and compile time bounded.
  constexpr bool kFull = NA * NB * NC * ND <= kStoreFullUnroll;
  constexpr int U1 = kFull ? 6 : 1;
  const int wid = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (wid >= a.ntasks) return;
  const Task t = a.tasks[wid];
  const int j = t.j0 + (int)(threadIdx.x & 63);
  if (j >= t.j1) return;
  const PairInfo P = a.bra[t.bra];
  const PairInfo K = a.ket[j];
  // the raw block: private memory, or for the largest class this thread's
  // slot of the library's staging buffer (written and read by this lane only)
  constexpr int NQ = NA * NB * NC * ND;
  double priv[NQ >= kStoreStagedNQ ? 1 : NQ];
  double* out = NQ >= kStoreStagedNQ ? a.blk + ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * NQ : priv;
  contracted_quartet<LA, LB, LC, LD>(a.bprim + (size_t)P.poff * kPrimPairStride, P.nprim,
                                     a.kprim + (size_t)K.poff * kPrimPairStride, K.nprim, a.boys,
                                     P.ABx, P.ABy, P.ABz, K.ABx, K.ABy, K.ABz, out);
  double na[NA], nb[NB], nc[NC], nd[ND];
#pragma unroll
  for (int i = 0; i < NA; ++i) na[i] = a.nbra_norma[P.bfa + i];
#pragma unroll
  for (int i = 0; i < NB; ++i) nb[i] = a.nbra_norma[P.bfb + i];
#pragma unroll
  for (int i = 0; i < NC; ++i) nc[i] = a.nket_norma[K.bfa + i];
#pragma unroll
  for (int i = 0; i < ND; ++i) nd[i] = a.nket_norma[K.bfb + i];
  // drop (|raw| < drop dropped) and normalise
#pragma unroll U1
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll U1
    for (int ib = 0; ib < NB; ++ib)
#pragma unroll
      for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
          const int e = ((ia * NB + ib) * NC + ic) * ND + id;
          const double v = out[e];
          out[e] = fabs(v) < a.drop ? 0.0 : v * na[ia] * nb[ib] * nc[ic] * nd[id];
        }
  const bool bdiag = P.sa == P.sb, kdiag = K.sa == K.sb;
  if (a.mode == 0) {
    double* A = a.out;
    const size_t s1 = a.sx1, s2 = a.sx2, s3 = a.sy1;
#pragma unroll U1
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll U1
      for (int ib = 0; ib < NB; ++ib)
#pragma unroll
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
          for (int id = 0; id < ND; ++id) {
            const int fa = P.bfa + ia, fb = P.bfb + ib, fc = K.bfa + ic, fd = K.bfb + id;
            if ((bdiag && fa > fb) || (kdiag && fc > fd)) continue;
            const double v = out[((ia * NB + ib) * NC + ic) * ND + id];
            if (v == 0.0) continue;
            const size_t x1 = a.bra_is_x ? fa : fc, x2 = a.bra_is_x ? fb : fd;
            const size_t y1 = a.bra_is_x ? fc : fa, y2 = a.bra_is_x ? fd : fb;
            A[x1 * s1 + x2 * s2 + y1 * s3 + y2] = v;
            A[x2 * s1 + x1 * s2 + y1 * s3 + y2] = v;
            A[x1 * s1 + x2 * s2 + y2 * s3 + y1] = v;
            A[x2 * s1 + x1 * s2 + y2 * s3 + y1] = v;
          }
    return;
  }
  const gptr<double> GG = (gptr<double>)a.out;
  const gcptr<double> c = (gcptr<double>)a.c;
  const size_t no = (size_t)a.n_o, no2 = no * no;
  if (a.bra_is_t) {
    // transformed pair (a b), other pair (c d)
#pragma unroll U1
    for (int ib = 0; ib < NB; ++ib)
#pragma unroll U1
      for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
          const int fc = K.bfa + ic, fd = K.bfb + id;
          if (kdiag && fc > fd) continue;
          double s = 0.0;
#pragma unroll
          for (int ia = 0; ia < NA; ++ia) s += c[P.bfa + ia] * out[((ia * NB + ib) * NC + ic) * ND + id];
          if (s != 0.0)
            global_add(GG + (size_t)(P.bfb + ib) * no2 + (size_t)min(fc, fd) * no + max(fc, fd), s);
        }
    if (!bdiag) {
#pragma unroll U1
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll U1
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
          for (int id = 0; id < ND; ++id) {
            const int fc = K.bfa + ic, fd = K.bfb + id;
            if (kdiag && fc > fd) continue;
            double s = 0.0;
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) s += c[P.bfb + ib] * out[((ia * NB + ib) * NC + ic) * ND + id];
            if (s != 0.0)
              global_add(GG + (size_t)(P.bfa + ia) * no2 + (size_t)min(fc, fd) * no + max(fc, fd), s);
          }
    }
  } else {
    // transformed pair (c d), other pair (a b)
#pragma unroll U1
    for (int id = 0; id < ND; ++id)
#pragma unroll U1
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
          const int fa = P.bfa + ia, fb = P.bfb + ib;
          if (bdiag && fa > fb) continue;
          double s = 0.0;
#pragma unroll
          for (int ic = 0; ic < NC; ++ic) s += c[K.bfa + ic] * out[((ia * NB + ib) * NC + ic) * ND + id];
          if (s != 0.0)
            global_add(GG + (size_t)(K.bfb + id) * no2 + (size_t)min(fa, fb) * no + max(fa, fb), s);
        }
    if (!kdiag) {
#pragma unroll U1
      for (int ic = 0; ic < NC; ++ic)
#pragma unroll U1
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
          for (int ib = 0; ib < NB; ++ib) {
            const int fa = P.bfa + ia, fb = P.bfb + ib;
            if (bdiag && fa > fb) continue;
            double s = 0.0;
#pragma unroll
            for (int id = 0; id < ND; ++id) s += c[K.bfb + id] * out[((ia * NB + ib) * NC + ic) * ND + id];
            if (s != 0.0)
              global_add(GG + (size_t)(K.bfa + ic) * no2 + (size_t)min(fa, fb) * no + max(fa, fb), s);
          }
    }
  }
}

template <int LA, int LB, int LC, int LD>
static void launch_store_t(const StoreArgs& a, hipStream_t s) {
  if (a.ntasks <= 0) return;
  const unsigned blocks = (unsigned)((a.ntasks + 3) / 4);
  if (a.mode == 0)
    hipLaunchKernelGGL((store_kernel<LA, LB, LC, LD, 0>), dim3(blocks), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((store_kernel<LA, LB, LC, LD, 1>), dim3(blocks), dim3(256), 0, s, a);
}

template <int LA, int LB, int LC, int LD>
static void launch_cstore_t(const CouplingStoreArgs& a, hipStream_t s) {
  if (a.ntasks <= 0) return;
  const unsigned blocks = (unsigned)((a.ntasks + 3) / 4);
  hipLaunchKernelGGL((coupling_store_kernel<LA, LB, LC, LD>), dim3(blocks), dim3(256), 0, s, a);
}

}  // namespace lowdin
