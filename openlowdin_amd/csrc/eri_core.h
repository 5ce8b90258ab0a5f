// Templates of the HIP kernels of the direct two-particle Fock build and the inter-species
// coupling build, written for gfx950 (CDNA4, 64-wide wavefronts).
//
// Work decomposition: a wavefront owns one bra shell pair and up to 64 ket
// shell pairs of one quartet class (one ket pair per lane).  All bra data
// (pair record, primitive pairs) are wave-uniform and are fetched through
// the scalar unit; the per-lane ket data are vector loads.  Each lane
// evaluates its contracted quartet with Obara-Saika VRR over its primitive
// quartets, contracts to [e0|f0] in registers, runs the HGP HRR once, and
// folds the integrals straight into the J/K (or coupling) blocks.  The J
// block of the shared bra pair is reduced across the wavefront before one
// atomic per element; the other five blocks differ per lane and go to HBM
// with hardware FP64 atomics.
//
// Reference behaviour reproduced (openLOWDIN src/ints/Libint2Iface.cpp):
//   quartet screen Dnorm * Q_ab * Q_cd < eps for electrons only (523-534),
//   permutational degeneracy (543-546), the six J/K updates with the
//   0.5*factor exchange weight (562-586), G = 0.25 (G + G^T) (616);
//   coupling: |raw| < 1e-10 drop, f1 <= f2 / of1 <= of2 element sets, the
//   doubling of off-diagonal density elements, lower-triangle (row-major)
//   result (1179-1266).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/amd_detail/amd_hip_unsafe_atomics.h>

#include "lowdin_device.h"
#include "lowdin_kernels.h"

namespace lowdin {

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Global-address-space views of pointers read from the molecule descriptors
// (a generic pointer loaded from memory would otherwise be accessed with
// flat instructions).
template <typename T>
using gcptr = const __attribute__((address_space(1))) T*;
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

// Member-wise copy of a pair record read through a global view (a struct
// copy from an address-space-qualified object has no implicit constructor).
__device__ __forceinline__ PairInfo load_pair(gcptr<PairInfo> p) {
  PairInfo r;
  r.sa = p->sa;
  r.sb = p->sb;
  r.bfa = p->bfa;
  r.bfb = p->bfb;
  r.poff = p->poff;
  r.nprim = p->nprim;
  r.la = p->la;
  r.lb = p->lb;
  r.Q = p->Q;
  r.ABx = p->ABx;
  r.ABy = p->ABy;
  r.ABz = p->ABz;
  r.Qn = p->Qn;
  return r;
}

__device__ __forceinline__ void global_add(gptr<double> p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 1/sqrt(x) for x > 0: hardware v_rsq_f64 refined by one Newton step.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
  return r * (1.5 - 0.5 * x * r * r);
}

// -------------------------------------------------------------- Boys F_m(T)
// Taylor expansion of F_L around the nearest tabulated point, exp(-T) the
// same way, then downward recursion F_m = (2T F_{m+1} + e^{-T}) / (2m+1).
// The table row of T_i holds e^{-T_i} at column 0 and F_L(T_i) .. F_{L+7}(T_i)
// at columns FO .. FO+7; RS doubles per row.  The full table has RS =
// kBoysCols, FO = 1 + L; a per-class copy in LDS keeps only the nine
// columns a class reads (RS = 9, FO = 1).
template <int L, int RS = kBoysCols, int FO = 1 + L, typename TP>
__device__ __forceinline__ void boys_f64(double T, TP tab, double* F) {
  if (T < kBoysTmax) {
    const int i = (int)(T * kBoysInvH + 0.5);
    const double x = (double)i * kBoysH - T;
    const auto r = tab + i * RS;
    double s = r[FO + 7];
    s = fma(s, x * (1.0 / 7.0), r[FO + 6]);
    s = fma(s, x * (1.0 / 6.0), r[FO + 5]);
    s = fma(s, x * (1.0 / 5.0), r[FO + 4]);
    s = fma(s, x * (1.0 / 4.0), r[FO + 3]);
    s = fma(s, x * (1.0 / 3.0), r[FO + 2]);
    s = fma(s, x * (1.0 / 2.0), r[FO + 1]);
    s = fma(s, x, r[FO + 0]);
    F[L] = s;
    if (L > 0) {
      double e = 1.0;
      e = fma(e, x * (1.0 / 7.0), 1.0);
      e = fma(e, x * (1.0 / 6.0), 1.0);
      e = fma(e, x * (1.0 / 5.0), 1.0);
      e = fma(e, x * (1.0 / 4.0), 1.0);
      e = fma(e, x * (1.0 / 3.0), 1.0);
      e = fma(e, x * (1.0 / 2.0), 1.0);
      e = fma(e, x, 1.0);
      const double eT = r[0] * e;
      const double t2 = 2.0 * T;
#pragma unroll
      for (int m = L - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + eT) * (1.0 / (2 * m + 1));
    }
  } else {
    // erf(sqrt(T)) = 1 to 4e-19 here, and 1/(2T) = r^2 / 2.  The upward
    // recursion keeps its exp(-T) term, which is not negligible against F_m
    // for m ~ T/5; for L <= 2 it is below F_L's rounding (e^-40 = 4.2e-18
    // against F_2(40) = 6.6e-5: 8e-16 relative) and is left out.
    const double r = rsqrt_nr(T);
    F[0] = 0.88622692545275801365 * r;
    if (L > 0) {
      const double o2t = 0.5 * r * r;
      if constexpr (L <= 2) {
#pragma unroll
        for (int m = 0; m < L; ++m) F[m + 1] = (double)(2 * m + 1) * F[m] * o2t;
      } else {
        const double eT = exp(-T);
#pragma unroll
        for (int m = 0; m < L; ++m) F[m + 1] = ((double)(2 * m + 1) * F[m] - eT) * o2t;
      }
    }
  }
}

// ---------------------------------------------------------- class dispatch
template <int LA, int LB, int LC, int LD>
struct EriClass;
#define LOWDIN_ERI_CLASS(a, b, c, d)                                                                 \
  template <>                                                                                        \
  struct EriClass<a, b, c, d> {                                                                      \
    static constexpr int nacc = NACC_##a##b##c##d;                                                   \
    template <typename R>                                                                            \
    __device__ static __forceinline__ void vrr(const R* F, R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, \
                                               R QCx, R QCy, R QCz, R WQx, R WQy, R WQz, R o2p,      \
                                               R o2q, R o2pq, R rop, R roq, R* acc) {                \
      vrr_##a##b##c##d<R>(F, PAx, PAy, PAz, WPx, WPy, WPz, QCx, QCy, QCz, WQx, WQy, WQz, o2p, o2q,   \
                          o2pq, rop, roq, acc);                                                      \
    }                                                                                                \
    template <typename R>                                                                            \
    __device__ static __forceinline__ void hrr(const R* acc, R ABx, R ABy, R ABz, R CDx, R CDy,      \
                                               R CDz, R* out) {                                      \
      hrr_##a##b##c##d<R>(acc, ABx, ABy, ABz, CDx, CDy, CDz, out);                                   \
    }                                                                                                \
  };


// Contracted raw quartet (canonical orientation) for one lane.  The ket's
// primitive pairs (per lane, vector loads) form the outer loop and are read
// once; the bra's primitive pairs (wave-uniform, scalar loads) the inner
// one.  VRR accumulates [e0|f0] over all primitive quartets, then one HRR.
template <int LA, int LB, int LC, int LD, int RS = kBoysCols, int FO = 1 + LA + LB + LC + LD, typename BP,
          typename KP, typename TP>
__device__ __forceinline__ void contracted_quartet_body(BP bprim, int nbp, KP kprim, int nkp, TP boys,
                                                   double ABx, double ABy, double ABz, double CDx,
                                                   double CDy, double CDz, double* out) {
  using C = EriClass<LA, LB, LC, LD>;
  constexpr int L = LA + LB + LC + LD;
  double acc[C::nacc];
#pragma unroll
  for (int k = 0; k < C::nacc; ++k) acc[k] = 0.0;
  for (int kp = 0; kp < nkp; ++kp) {
    const auto kq = kprim + kp * kPrimPairStride;
    const double q = kq[0], o2q = kq[1];
    const double Qx = kq[2], Qy = kq[3], Qz = kq[4];
    const double QCx = kq[5], QCy = kq[6], QCz = kq[7];
    const double Kcd = kq[8];
    for (int ip = 0; ip < nbp; ++ip) {
      const auto bp = bprim + ip * kPrimPairStride;
      const double p = bp[0], o2p = bp[1];
      const double PQx = bp[2] - Qx, PQy = bp[3] - Qy, PQz = bp[4] - Qz;
      const double rs = rsqrt_nr(p + q);
      const double inv = rs * rs;
      const double rop = q * inv, roq = p * inv;
      const double rho = p * rop;
      const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
      const double pref = bp[8] * Kcd * rs;
      double F[L + 1];
      boys_f64<L, RS, FO>(T, boys, F);
#pragma unroll
      for (int m = 0; m <= L; ++m) F[m] *= pref;
      C::template vrr<double>(F, bp[5], bp[6], bp[7], -rop * PQx, -rop * PQy, -rop * PQz, QCx, QCy, QCz,
                              roq * PQx, roq * PQy, roq * PQz, o2p, o2q, 0.5 * inv, rop, roq, acc);
    }
  }
  C::template hrr<double>(acc, ABx, ABy, ABz, CDx, CDy, CDz, out);
}

// FP32 Boys: fifth-order Taylor on the full table's grid (|dT| <= 0.05,
// error ~3e-9 relative), exp(-T) the same way, downward recursion; above
// T = 40 the asymptotic form with upward recursion.
// RS floats per row, F_L .. F_{L+4} at columns FO .. FO+4 (the full table:
// RS = kBoysF32Cols, FO = 1 + L; a per-class copy in LDS: RS = 6, FO = 1).
template <int L, int RS = kBoysF32Cols, int FO = 1 + L, typename TP>
__device__ __forceinline__ void boys_f32(float T, TP tab, float* F) {
  static_assert(RS != kBoysF32Cols || 1 + L + kBoysF32Taylor - 1 < kBoysF32Cols,
                "FP32 Boys table lacks the Taylor columns of this L");
  if (T < kBoysF32Tmax) {
    // T < Tmax rounds to a row <= Tmax * InvH, the table's last row
    const int i = (int)(T * kBoysF32InvH + 0.5f);
    const float x = (float)i * kBoysF32H - T;
    const auto r = tab + i * RS;
    float s = r[FO + 4];
    s = fmaf(s, x * 0.25f, r[FO + 3]);
    s = fmaf(s, x * (1.0f / 3.0f), r[FO + 2]);
    s = fmaf(s, x * 0.5f, r[FO + 1]);
    s = fmaf(s, x, r[FO + 0]);
    F[L] = s;
    if (L > 0) {
      float e = 1.0f;
      e = fmaf(e, x * 0.25f, 1.0f);
      e = fmaf(e, x * (1.0f / 3.0f), 1.0f);
      e = fmaf(e, x * 0.5f, 1.0f);
      e = fmaf(e, x, 1.0f);
      const float eT = r[0] * e;
      const float t2 = 2.0f * T;
#pragma unroll
      for (int m = L - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + eT) * (1.0f / (2 * m + 1));
    }
  } else {
    F[0] = 0.88622692545275801365f * __builtin_amdgcn_rsqf(T);
    if (L > 0) {
      const float o2t = 0.5f / T;
      const float eT = __expf(-T);
#pragma unroll
      for (int m = 0; m < L; ++m) F[m + 1] = ((float)(2 * m + 1) * F[m] - eT) * o2t;
    }
  }
}

// FP32 contracted quartet (mixed precision): same loop structure as the
// FP64 one over FP32 copies of the primitive-pair records; the integrals
// come back as FP64 for the FP64 contraction and accumulation.
template <int LA, int LB, int LC, int LD, int RS = kBoysF32Cols, int FO = 1 + LA + LB + LC + LD, typename BP,
          typename KP, typename TP>
__device__ __forceinline__ void contracted_quartet_f32_body(BP bprim, int nbp, KP kprim, int nkp, TP boys,
                                                       double ABx, double ABy, double ABz, double CDx,
                                                       double CDy, double CDz, double* out) {
  using C = EriClass<LA, LB, LC, LD>;
  constexpr int L = LA + LB + LC + LD;
  constexpr int NOUT = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
  float acc[C::nacc];
#pragma unroll
  for (int k = 0; k < C::nacc; ++k) acc[k] = 0.0f;
  for (int kp = 0; kp < nkp; ++kp) {
    const auto kq = kprim + kp * kPrimPairStride;
    const float q = kq[0], o2q = kq[1];
    const float Qx = kq[2], Qy = kq[3], Qz = kq[4];
    const float QCx = kq[5], QCy = kq[6], QCz = kq[7];
    const float Kcd = kq[8];
    for (int ip = 0; ip < nbp; ++ip) {
      const auto bp = bprim + ip * kPrimPairStride;
      const float p = bp[0], o2p = bp[1];
      const float PQx = bp[2] - Qx, PQy = bp[3] - Qy, PQz = bp[4] - Qz;
      const float rs = __builtin_amdgcn_rsqf(p + q);
      const float inv = rs * rs;
      const float rop = q * inv, roq = p * inv;
      const float rho = p * rop;
      const float T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
      const float pref = bp[8] * Kcd * rs;
      float F[L + 1];
      boys_f32<L, RS, FO>(T, boys, F);
#pragma unroll
      for (int m = 0; m <= L; ++m) F[m] *= pref;
      C::template vrr<float>(F, bp[5], bp[6], bp[7], -rop * PQx, -rop * PQy, -rop * PQz, QCx, QCy, QCz,
                             roq * PQx, roq * PQy, roq * PQz, o2p, o2q, 0.5f * inv, rop, roq, acc);
    }
  }
  float o[NOUT];
  C::template hrr<float>(acc, (float)ABx, (float)ABy, (float)ABz, (float)CDx, (float)CDy, (float)CDz, o);
#pragma unroll
  for (int k = 0; k < NOUT; ++k) out[k] = (double)o[k];
}

// Quartet evaluation is inlined into every kernel of a class: an
// out-of-line body for the largest classes doubled their scratch (the
// integral array then lives in the caller's private segment as well) and
// the runtime could not provision it at full occupancy.  The two largest
// classes ((dd|dp), (dd|dd): single-primitive d shells in every workload
// here) keep FP64 in mixed mode, which halves their code and compile time.
constexpr int kMaxNaccFp32 = 300;

template <int LA, int LB, int LC, int LD, typename BP, typename KP, typename TP>
__device__ __forceinline__ void contracted_quartet(BP bprim, int nbp, KP kprim, int nkp, TP boys, double ABx,
                                                   double ABy, double ABz, double CDx, double CDy, double CDz,
                                                   double* out) {
  contracted_quartet_body<LA, LB, LC, LD>(bprim, nbp, kprim, nkp, boys, ABx, ABy, ABz, CDx, CDy, CDz, out);
}

template <int LA, int LB, int LC, int LD>
struct Fp32Allowed {
  static constexpr bool value = EriClass<LA, LB, LC, LD>::nacc <= kMaxNaccFp32;
};

template <int LA, int LB, int LC, int LD, typename BP, typename KP>
__device__ __forceinline__ void contracted_quartet_f32(BP bprim, int nbp, KP kprim, int nkp,
                                                       const float* __restrict__ boys, double ABx, double ABy,
                                                       double ABz, double CDx, double CDy, double CDz, double* out) {
  contracted_quartet_f32_body<LA, LB, LC, LD>(bprim, nbp, kprim, nkp, boys, ABx, ABy, ABz, CDx, CDy, CDz, out);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Largest block infinity norm of D over the six shell blocks a quartet
// touches (ab, cd, ac, bd, ad, bc).
__device__ __forceinline__ double quartet_dnorm(gcptr<double> Dn, int ns, const PairInfo& P, const PairInfo& K) {
  double dn = Dn[P.sa * ns + P.sb];
  dn = fmax(dn, Dn[K.sa * ns + K.sb]);
  dn = fmax(dn, Dn[P.sa * ns + K.sa]);
  dn = fmax(dn, Dn[P.sb * ns + K.sb]);
  dn = fmax(dn, Dn[P.sa * ns + K.sb]);
  dn = fmax(dn, Dn[P.sb * ns + K.sa]);
  return dn;
}

// The J/K updates of one quartet's raw block (out), normalised and weighted
// by deg: J_cd, K_ac, K_bd, K_ad and K_bc go to G with FP64 atomics; J_ab is
// added into jab (unscaled by deg and the bra norms) for the caller to reduce.
// Libint2Iface.cpp:562-586 (exchange weight 0.5 * factor = half_factor).
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void jk_quartet_update(const double* out, const PairInfo& P, const PairInfo& K, int n,
                                                  gcptr<double> D, gcptr<double> nrm, gptr<double> G,
                                                  double deg, double half_factor, double* jab) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  double na[NA], nb[NB], nc[NC], nd[ND];
#pragma unroll
  for (int i = 0; i < NA; ++i) na[i] = nrm[P.bfa + i];
#pragma unroll
  for (int i = 0; i < NB; ++i) nb[i] = nrm[P.bfb + i];
#pragma unroll
  for (int i = 0; i < NC; ++i) nc[i] = nrm[K.bfa + i];
#pragma unroll
  for (int i = 0; i < ND; ++i) nd[i] = nrm[K.bfb + i];
  // normalised density blocks
  double dab[NA * NB], dcd[NC * ND], dac[NA * NC], dbd[NB * ND], dad[NA * ND], dbc[NB * NC];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int k = 0; k < NB; ++k) dab[i * NB + k] = D[(size_t)(P.bfa + i) * n + P.bfb + k] * na[i] * nb[k];
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int k = 0; k < ND; ++k) dcd[i * ND + k] = D[(size_t)(K.bfa + i) * n + K.bfb + k] * nc[i] * nd[k];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int k = 0; k < NC; ++k) dac[i * NC + k] = D[(size_t)(P.bfa + i) * n + K.bfa + k] * na[i] * nc[k];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int k = 0; k < ND; ++k) dbd[i * ND + k] = D[(size_t)(P.bfb + i) * n + K.bfb + k] * nb[i] * nd[k];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int k = 0; k < ND; ++k) dad[i * ND + k] = D[(size_t)(P.bfa + i) * n + K.bfb + k] * na[i] * nd[k];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int k = 0; k < NC; ++k) dbc[i * NC + k] = D[(size_t)(P.bfb + i) * n + K.bfa + k] * nb[i] * nc[k];
  double jcd[NC * ND], kac[NA * NC], kbd[NB * ND], kad[NA * ND], kbc[NB * NC];
#pragma unroll
  for (int k = 0; k < NC * ND; ++k) jcd[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NA * NC; ++k) kac[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NB * ND; ++k) kbd[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NA * ND; ++k) kad[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NB * NC; ++k) kbc[k] = 0.0;
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib)
#pragma unroll
      for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
          const double v = out[((ia * NB + ib) * NC + ic) * ND + id];
          jab[ia * NB + ib] += dcd[ic * ND + id] * v;
          jcd[ic * ND + id] += dab[ia * NB + ib] * v;
          kac[ia * NC + ic] += dbd[ib * ND + id] * v;
          kbd[ib * ND + id] += dac[ia * NC + ic] * v;
          kad[ia * ND + id] += dbc[ib * NC + ic] * v;
          kbc[ib * NC + ic] += dad[ia * ND + id] * v;
        }
  const double sj = deg, sk = deg * half_factor;
#pragma unroll
  for (int ic = 0; ic < NC; ++ic)
#pragma unroll
    for (int id = 0; id < ND; ++id)
      global_add(G + (size_t)(K.bfa + ic) * n + K.bfb + id, jcd[ic * ND + id] * sj * nc[ic] * nd[id]);
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
      global_add(G + (size_t)(P.bfa + ia) * n + K.bfa + ic, kac[ia * NC + ic] * sk * na[ia] * nc[ic]);
#pragma unroll
  for (int ib = 0; ib < NB; ++ib)
#pragma unroll
    for (int id = 0; id < ND; ++id)
      global_add(G + (size_t)(P.bfb + ib) * n + K.bfb + id, kbd[ib * ND + id] * sk * nb[ib] * nd[id]);
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int id = 0; id < ND; ++id)
      global_add(G + (size_t)(P.bfa + ia) * n + K.bfb + id, kad[ia * ND + id] * sk * na[ia] * nd[id]);
#pragma unroll
  for (int ib = 0; ib < NB; ++ib)
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
      global_add(G + (size_t)(P.bfb + ib) * n + K.bfa + ic, kbc[ib * NC + ic] * sk * nb[ib] * nc[ic]);
}

// ------------------------------------------------------------ J/K kernel
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void jk_kernel(JKArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  const int wid = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (wid >= a.ntasks) return;
  const Task t = a.tasks[wid];
  const MolDesc& M = a.mols[t.flags];
  const gcptr<PairInfo> pairs = (gcptr<PairInfo>)M.pairs;
  const gcptr<double> prim = (gcptr<double>)M.prim;
  const gcptr<double> D = (gcptr<double>)M.D;
  const gcptr<double> nrm = (gcptr<double>)M.norma;
  const gptr<double> G = (gptr<double>)M.Graw;
  const int lane = threadIdx.x & 63;
  const PairInfo P = load_pair(pairs + M.off[a.cbra] + t.bra);
  const int j = t.j0 + lane;
  bool active = j < t.j1;
  const PairInfo K = active ? load_pair(pairs + M.off[a.cket] + j) : P;
  const int n = M.n, ns = M.nshell;
  bool lo = false;
  if (active && (M.screen || M.mixed)) {
    const double dn = quartet_dnorm((gcptr<double>)M.Dn, ns, P, K);
    if (M.screen && dn * P.Q * K.Q < a.eps) active = false;
    lo = Fp32Allowed<LA, LB, LC, LD>::value && M.mixed && dn * P.Qn * K.Qn < M.tau;
  }
  const uint64_t amask = __ballot(active);
  if (a.counts) {
    // per-wave statistics: plain stores into this task's slot (summed by
    // reduce_counts_kernel), no same-address atomics
    unsigned int np = active ? (unsigned int)(P.nprim * K.nprim) : 0u;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) np += __shfl_xor(np, o, 64);
    const uint64_t lmask = __ballot(active && lo);
    if (lane == 0) {
      a.counts[3 * wid] = (unsigned int)__popcll(amask);
      a.counts[3 * wid + 1] = np;
      a.counts[3 * wid + 2] = (unsigned int)__popcll(lmask);
    }
  }
  double jab[NA * NB];
#pragma unroll
  for (int k = 0; k < NA * NB; ++k) jab[k] = 0.0;
  if (active) {
    double out[NA * NB * NC * ND];
    if constexpr (Fp32Allowed<LA, LB, LC, LD>::value) {
      if (lo)
        contracted_quartet_f32<LA, LB, LC, LD>((gcptr<float>)M.primf + (size_t)P.poff * kPrimPairStride, P.nprim,
                                               (gcptr<float>)M.primf + (size_t)K.poff * kPrimPairStride, K.nprim,
                                               a.boysf,
                                               P.ABx, P.ABy, P.ABz, K.ABx, K.ABy, K.ABz, out);
    }
    if (!lo)
      contracted_quartet<LA, LB, LC, LD>(prim + (size_t)P.poff * kPrimPairStride, P.nprim,
                                         prim + (size_t)K.poff * kPrimPairStride, K.nprim, a.boys,
                                         P.ABx, P.ABy, P.ABz, K.ABx, K.ABy, K.ABz, out);
    const double deg = (P.sa != P.sb ? 2.0 : 1.0) * (K.sa != K.sb ? 2.0 : 1.0) *
                       ((a.same_class && j == t.bra) ? 1.0 : 2.0);
    jk_quartet_update<LA, LB, LC, LD>(out, P, K, n, D, nrm, G, deg, M.half_factor, jab);
#pragma unroll
    for (int k = 0; k < NA * NB; ++k) jab[k] *= deg;
  }
  // the bra pair is shared by the whole wavefront: reduce its J block
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) {
      const double s = wave_sum(jab[ia * NB + ib]);
      if (lane == 0 && amask)
        global_add(G + (size_t)(P.bfa + ia) * n + P.bfb + ib, s * nrm[P.bfa + ia] * nrm[P.bfb + ib]);
    }
}

// Waves per SIMD the strip kernel of a class is compiled for (the register
// budget: 512 / waves).  Default 1 (no limit).  The classes listed measured
// faster with the tighter budget on C20H42 / cc-pVDZ: (ps|ss) at 5 waves
// (85 VGPRs, no spill; 4 waves unrestricted), (ds|ds) and (dd|ss) at 2
// (a few hundred bytes of spill against occupancy 1).  Others measured
// slower: (ps|ps), (pp|ss), (ds|ss) at 5 waves by 6-22 %, (ps|ss) at 6,
// (ds|ps), (dp|ss), (pp|ps) at 3, and the other classes with more than
// 256 registers at 2 waves (e.g. (dp|pp) 1.6 -> 3.6 ms).
template <int LA, int LB, int LC, int LD>
struct StripWaves {
  static constexpr int value = 1;
};
#define LOWDIN_STRIP_WAVES(a, b, c, d, w) \
  template <>                              \
  struct StripWaves<a, b, c, d> {          \
    static constexpr int value = w;        \
  };
LOWDIN_STRIP_WAVES(1, 0, 0, 0, 5)
LOWDIN_STRIP_WAVES(2, 0, 2, 0, 2)
LOWDIN_STRIP_WAVES(2, 2, 0, 0, 2)
#undef LOWDIN_STRIP_WAVES

// Classes whose strip kernel copies the nine Boys-table columns it reads
// (e^-T and F_L .. F_{L+7} at the 401 grid points, 28.9 KB) into LDS once per
// workgroup when that fits beside the row strips in 64 KB: the table gathers
// then leave the texture-address unit, which these classes keep busy
// (measured on C20H42 / cc-pVDZ, serialised: (ps|ss) 3.39 -> 3.23 ms,
// (pp|ss) 1.16 -> 1.00, (ss|ss) 2.79 -> 2.74; (ps|ps) and (ds|ss) unchanged).
template <int LA, int LB, int LC, int LD>
struct StripLdsBoys {
  static constexpr bool value = false;
};
template <> struct StripLdsBoys<0, 0, 0, 0> { static constexpr bool value = true; };
template <> struct StripLdsBoys<1, 0, 0, 0> { static constexpr bool value = true; };
template <> struct StripLdsBoys<1, 1, 0, 0> { static constexpr bool value = true; };
constexpr int kLdsBoysCols = 9;
constexpr size_t kLdsBoysBytes = (size_t)kBoysRows * kLdsBoysCols * sizeof(double);
// FP32 copy for the mixed-precision quartets: e^-T and F_L .. F_{L+4}.
constexpr int kLdsBoys32Cols = 6;
constexpr size_t kLdsBoys32Bytes = (size_t)kBoysF32Rows * kLdsBoys32Cols * sizeof(float);

// Whether a strip launch of this class with row strips of width lds_n and
// workgroups of `block` threads keeps the Boys columns in LDS (host and
// device agree through this function).  Only 256-thread workgroups, which
// run long ket streams: a one-wavefront workgroup of a small molecule
// (a few dozen kets) would spend more on copying the table than it saves.
template <int LA, int LB, int LC, int LD>
__host__ __device__ constexpr bool strip_lds_boys(int lds_n, int block) {
  return StripLdsBoys<LA, LB, LC, LD>::value && (block <= 0 || block >= 256) &&
         (size_t)(ncart(LA) + ncart(LB)) * (size_t)lds_n * sizeof(double) + kLdsBoysBytes <= 65536;
}
// The same for the FP32 columns, when some molecule of the launch runs mixed
// precision and both copies fit.
template <int LA, int LB, int LC, int LD>
__host__ __device__ constexpr bool strip_lds_boys32(int lds_n, int block, int any_mixed) {
  return strip_lds_boys<LA, LB, LC, LD>(lds_n, block) && any_mixed != 0 &&
         (size_t)(ncart(LA) + ncart(LB)) * (size_t)lds_n * sizeof(double) + kLdsBoysBytes + kLdsBoys32Bytes <=
             65536;
}

// ------------------------------------------------ J/K kernel, LDS strips
// A workgroup owns one bra pair (a b) and streams up to t.j1 - t.j0 ket
// pairs, one per thread per round (256 threads, or one wavefront of 64 for
// the short ket runs of small molecules).  Every J/K update of a quartet except
// J_cd lands in row a or row b of G (J_ab; K_ac, K_ad in row a; K_bc, K_bd
// in row b), so those are accumulated with LDS atomics in two row strips
// (NA + NB) x n held for the whole ket stream and flushed to HBM once, with
// one global atomic per touched element.  J_cd, distinct for every ket, is
// the only per-quartet global atomic.  The bra's primitive-pair records
// (at most kMaxStripPrims, guaranteed by the host; FP64, and in mixed mode
// also the FP32 copies) are staged in LDS once: every lane reads the same
// record, an LDS broadcast.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256, (StripWaves<LA, LB, LC, LD>::value)) void jk_strip_kernel(JKArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr bool kF32 = Fp32Allowed<LA, LB, LC, LD>::value;
  extern __shared__ double strip[];  // rows of a (NA x n), then rows of b (NB x n)
  __shared__ double bpd[kMaxStripPrims * kPrimPairStride];
  __shared__ float bpf[kF32 ? kMaxStripPrims * kPrimPairStride : 1];   // FP32 bra records (mixed)
  __shared__ unsigned int wg_q, wg_p, wg_q32;
  const int tid = threadIdx.x;
  const Task t = a.tasks[blockIdx.x];
  const MolDesc& M = a.mols[t.flags];
  const gcptr<PairInfo> pairs = (gcptr<PairInfo>)M.pairs;
  const gcptr<double> prim = (gcptr<double>)M.prim;
  const gcptr<double> D = (gcptr<double>)M.D;
  const gcptr<double> nrm = (gcptr<double>)M.norma;
  const gptr<double> G = (gptr<double>)M.Graw;
  const PairInfo P = load_pair(pairs + M.off[a.cbra] + t.bra);
  const int n = M.n, ns = M.nshell;
  double* sa = strip;
  double* sb = strip + NA * n;
  for (int k = tid; k < (NA + NB) * n; k += blockDim.x) strip[k] = 0.0;
  const bool ldsb = strip_lds_boys<LA, LB, LC, LD>(a.lds_n, a.block);
  double* btab = strip + (NA + NB) * a.lds_n;
  const bool ldsb32 = kF32 && strip_lds_boys32<LA, LB, LC, LD>(a.lds_n, a.block, a.any_mixed);
  float* btab32 = reinterpret_cast<float*>(btab + kBoysRows * kLdsBoysCols);
  if constexpr (StripLdsBoys<LA, LB, LC, LD>::value) {
    constexpr int L = LA + LB + LC + LD;
    if (ldsb) {
      for (int k = tid; k < kBoysRows * kLdsBoysCols; k += blockDim.x) {
        const int r = k / kLdsBoysCols, c = k % kLdsBoysCols;
        btab[k] = a.boys[r * kBoysCols + (c == 0 ? 0 : L + c)];
      }
    }
    if (ldsb32) {
      for (int k = tid; k < kBoysF32Rows * kLdsBoys32Cols; k += blockDim.x) {
        const int r = k / kLdsBoys32Cols, c = k % kLdsBoys32Cols;
        btab32[k] = a.boysf[r * kBoysF32Cols + (c == 0 ? 0 : L + c)];
      }
    }
  }
  {
    const gcptr<double> src = prim + (size_t)P.poff * kPrimPairStride;
    for (int k = tid; k < P.nprim * kPrimPairStride; k += blockDim.x) bpd[k] = src[k];
    if constexpr (kF32) {
      if (M.mixed) {
        const gcptr<float> srcf = (gcptr<float>)M.primf + (size_t)P.poff * kPrimPairStride;
        for (int k = tid; k < P.nprim * kPrimPairStride; k += blockDim.x) bpf[k] = srcf[k];
      }
    }
  }
  if (tid == 0) {
    wg_q = 0;
    wg_p = 0;
    wg_q32 = 0;
  }
  __syncthreads();
  double na[NA], nb[NB];
#pragma unroll
  for (int i = 0; i < NA; ++i) na[i] = nrm[P.bfa + i];
#pragma unroll
  for (int i = 0; i < NB; ++i) nb[i] = nrm[P.bfb + i];
  double dab[NA * NB];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int k = 0; k < NB; ++k) dab[i * NB + k] = D[(size_t)(P.bfa + i) * n + P.bfb + k] * na[i] * nb[k];
  double jab[NA * NB];
#pragma unroll
  for (int k = 0; k < NA * NB; ++k) jab[k] = 0.0;
  unsigned int myq = 0, myp = 0, myq32 = 0;
  const double sbra = (P.sa != P.sb ? 2.0 : 1.0);
  const double hf = M.half_factor;
  for (int j = t.j0 + tid; j < t.j1; j += blockDim.x) {
    const PairInfo K = load_pair(pairs + M.off[a.cket] + j);
    bool lo = false;
    if (M.screen || M.mixed) {
      const double dn = quartet_dnorm((gcptr<double>)M.Dn, ns, P, K);
      if (M.screen && dn * P.Q * K.Q < a.eps) continue;
      lo = kF32 && M.mixed && dn * P.Qn * K.Qn < M.tau;
    }
    myq += 1;
    myp += (unsigned int)(P.nprim * K.nprim);
    double out[NA * NB * NC * ND];
    if constexpr (kF32) {
      if (lo) {
        myq32 += 1;
        bool done = false;
        if constexpr (StripLdsBoys<LA, LB, LC, LD>::value) {
          if (ldsb32) {
            contracted_quartet_f32_body<LA, LB, LC, LD, kLdsBoys32Cols, 1>(
                (const float*)bpf, P.nprim, (gcptr<float>)M.primf + (size_t)K.poff * kPrimPairStride, K.nprim,
                (const float*)btab32, P.ABx, P.ABy, P.ABz, K.ABx, K.ABy, K.ABz, out);
            done = true;
          }
        }
        if (!done)
          contracted_quartet_f32<LA, LB, LC, LD>((const float*)bpf, P.nprim,
                                                 (gcptr<float>)M.primf + (size_t)K.poff * kPrimPairStride,
                                                 K.nprim, a.boysf, P.ABx, P.ABy, P.ABz, K.ABx, K.ABy, K.ABz, out);
      }
    }
    if (!lo) {
      bool done = false;
      if constexpr (StripLdsBoys<LA, LB, LC, LD>::value) {
        if (ldsb) {
          contracted_quartet_body<LA, LB, LC, LD, kLdsBoysCols, 1>((const double*)bpd, P.nprim,
                                                                   prim + (size_t)K.poff * kPrimPairStride,
                                                                   K.nprim, (const double*)btab, P.ABx, P.ABy,
                                                                   P.ABz, K.ABx, K.ABy, K.ABz, out);
          done = true;
        }
      }
      if (!done)
        contracted_quartet<LA, LB, LC, LD>((const double*)bpd, P.nprim,
                                           prim + (size_t)K.poff * kPrimPairStride, K.nprim, a.boys,
                                           P.ABx, P.ABy, P.ABz, K.ABx, K.ABy, K.ABz, out);
    }
    const double deg = sbra * (K.sa != K.sb ? 2.0 : 1.0) * ((a.same_class && j == t.bra) ? 1.0 : 2.0);
    double nc[NC], nd[ND];
#pragma unroll
    for (int i = 0; i < NC; ++i) nc[i] = nrm[K.bfa + i];
#pragma unroll
    for (int i = 0; i < ND; ++i) nd[i] = nrm[K.bfb + i];
    double dcd[NC * ND], dac[NA * NC], dbd[NB * ND], dad[NA * ND], dbc[NB * NC];
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int k = 0; k < ND; ++k) dcd[i * ND + k] = D[(size_t)(K.bfa + i) * n + K.bfb + k] * nc[i] * nd[k];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int k = 0; k < NC; ++k) dac[i * NC + k] = D[(size_t)(P.bfa + i) * n + K.bfa + k] * na[i] * nc[k];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int k = 0; k < ND; ++k) dbd[i * ND + k] = D[(size_t)(P.bfb + i) * n + K.bfb + k] * nb[i] * nd[k];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int k = 0; k < ND; ++k) dad[i * ND + k] = D[(size_t)(P.bfa + i) * n + K.bfb + k] * na[i] * nd[k];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int k = 0; k < NC; ++k) dbc[i * NC + k] = D[(size_t)(P.bfb + i) * n + K.bfa + k] * nb[i] * nc[k];
    double jcd[NC * ND], kac[NA * NC], kbd[NB * ND], kad[NA * ND], kbc[NB * NC];
#pragma unroll
    for (int k = 0; k < NC * ND; ++k) jcd[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NA * NC; ++k) kac[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NB * ND; ++k) kbd[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NA * ND; ++k) kad[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NB * NC; ++k) kbc[k] = 0.0;
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib)
#pragma unroll
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
          for (int id = 0; id < ND; ++id) {
            const double v = deg * out[((ia * NB + ib) * NC + ic) * ND + id];
            jab[ia * NB + ib] += dcd[ic * ND + id] * v;
            jcd[ic * ND + id] += dab[ia * NB + ib] * v;
            kac[ia * NC + ic] += dbd[ib * ND + id] * v;
            kbd[ib * ND + id] += dac[ia * NC + ic] * v;
            kad[ia * ND + id] += dbc[ib * NC + ic] * v;
            kbc[ib * NC + ic] += dad[ia * ND + id] * v;
          }
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
      for (int id = 0; id < ND; ++id)
        global_add(G + (size_t)(K.bfa + ic) * n + K.bfb + id, jcd[ic * ND + id] * nc[ic] * nd[id]);
#pragma unroll
    for (int ia = 0; ia < NA; ++ia) {
#pragma unroll
      for (int ic = 0; ic < NC; ++ic) atomicAdd(&sa[ia * n + K.bfa + ic], kac[ia * NC + ic] * hf * nc[ic]);
#pragma unroll
      for (int id = 0; id < ND; ++id) atomicAdd(&sa[ia * n + K.bfb + id], kad[ia * ND + id] * hf * nd[id]);
    }
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) {
#pragma unroll
      for (int ic = 0; ic < NC; ++ic) atomicAdd(&sb[ib * n + K.bfa + ic], kbc[ib * NC + ic] * hf * nc[ic]);
#pragma unroll
      for (int id = 0; id < ND; ++id) atomicAdd(&sb[ib * n + K.bfb + id], kbd[ib * ND + id] * hf * nd[id]);
    }
  }
  // J_ab of this thread into row a, column block b
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib)
      if (jab[ia * NB + ib] != 0.0) atomicAdd(&sa[ia * n + P.bfb + ib], jab[ia * NB + ib] * nb[ib]);
  if (a.counts) {
    atomicAdd(&wg_q, myq);
    atomicAdd(&wg_p, myp);
    atomicAdd(&wg_q32, myq32);
  }
  __syncthreads();
  // flush: one global atomic per touched strip element (row norma applied here)
  for (int r = 0; r < NA; ++r)
    for (int c = tid; c < n; c += blockDim.x) {
      const double v = sa[r * n + c];
      if (v != 0.0) global_add(G + (size_t)(P.bfa + r) * n + c, v * na[r]);
    }
  for (int r = 0; r < NB; ++r)
    for (int c = tid; c < n; c += blockDim.x) {
      const double v = sb[r * n + c];
      if (v != 0.0) global_add(G + (size_t)(P.bfb + r) * n + c, v * nb[r]);
    }
  if (a.counts && tid == 0) {
    a.counts[3 * blockIdx.x] = wg_q;
    a.counts[3 * blockIdx.x + 1] = wg_p;
    a.counts[3 * blockIdx.x + 2] = wg_q32;
  }
}

// ------------------------------------------------------- coupling kernel
// Quartets (bra pair of species X | ket pair of species Y) over ALL pairs
// s1 <= s2 of both species.  Computes the result on the bra species from the
// ket species' density (rbra) and/or on the ket species from the bra
// species' density (rket), each with the reference's element sets and drop.
// Drop, density weights and scatter of one coupling quartet's raw block
// (shared by the Coulomb and the Gaussian-geminal coupling kernels).
// kWaveBra: every lane of the wavefront calls this with the same bra pair
// (lanes without a ket pass a zero block), and the bra-side sums are reduced
// over the wavefront before one atomic per element.
template <int LA, int LB, int LC, int LD, bool kWaveBra = false>
__device__ __forceinline__ void coupling_scatter(const CouplingArgs& a, const PairInfo& P, const PairInfo& K,
                                                 double* out) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  const double drop = a.drop;
#pragma unroll
  for (int k = 0; k < NA * NB * NC * ND; ++k)
    if (fabs(out[k]) < drop) out[k] = 0.0;
  const bool bdiag = P.sa == P.sb, kdiag = K.sa == K.sb;
  if (a.rbra) {
    // weights of the ket density: element set of0 <= of1 within a diagonal
    // shell pair, doubled off the function diagonal; normalised
    double w[NC * ND];
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
      for (int id = 0; id < ND; ++id) {
        const int fc = K.bfa + ic, fd = K.bfb + id;
        double v = 0.0;
        if (!kdiag || ic <= id)
          v = (fc != fd ? 2.0 : 1.0) * a.dket[(size_t)fc * a.nket + fd] * a.nket_norma[fc] * a.nket_norma[fd];
        w[ic * ND + id] = v;
      }
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) {
        if (bdiag && ia > ib) continue;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NC * ND; ++k) s += w[k] * out[(ia * NB + ib) * NC * ND + k];
        const int fa = P.bfa + ia, fb = P.bfb + ib;
        const int r = fa > fb ? fa : fb, c = fa > fb ? fb : fa;
        if constexpr (kWaveBra) {
          s = wave_sum(s);
          if ((threadIdx.x & 63) == 0 && s != 0.0)
            unsafeAtomicAdd(&a.rbra[(size_t)r * a.nbra + c], s * a.nbra_norma[fa] * a.nbra_norma[fb]);
        } else {
          unsafeAtomicAdd(&a.rbra[(size_t)r * a.nbra + c], s * a.nbra_norma[fa] * a.nbra_norma[fb]);
        }
      }
  }
  if (a.rket) {
    double w[NA * NB];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) {
        const int fa = P.bfa + ia, fb = P.bfb + ib;
        double v = 0.0;
        if (!bdiag || ia <= ib)
          v = (fa != fb ? 2.0 : 1.0) * a.dbra[(size_t)fa * a.nbra + fb] * a.nbra_norma[fa] * a.nbra_norma[fb];
        w[ia * NB + ib] = v;
      }
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
      for (int id = 0; id < ND; ++id) {
        if (kdiag && ic > id) continue;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NA * NB; ++k) s += w[k] * out[k * NC * ND + ic * ND + id];
        const int fc = K.bfa + ic, fd = K.bfb + id;
        const int r = fc > fd ? fc : fd, c = fc > fd ? fd : fc;
        unsafeAtomicAdd(&a.rket[(size_t)r * a.nket + c], s * a.nket_norma[fc] * a.nket_norma[fd]);
      }
  }
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void coupling_kernel(CouplingArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  const int wid = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (wid >= a.ntasks) return;
  const Task t = a.tasks[wid];
  const int lane = threadIdx.x & 63;
  const int j = t.j0 + lane;
  // every lane stays for the wavefront's bra-side reduction; lanes without
  // a ket contribute a zero block and no ket-side update
  const bool active = j < t.j1;
  const PairInfo& P = a.bra[t.bra];
  const PairInfo K = a.ket[active ? j : t.j0];
  double out[NA * NB * NC * ND];
  if (active) {
    contracted_quartet<LA, LB, LC, LD>(a.bprim + (size_t)P.poff * kPrimPairStride, P.nprim,
                                       a.kprim + (size_t)K.poff * kPrimPairStride, K.nprim, a.boys,
                                       P.ABx, P.ABy, P.ABz, K.ABx, K.ABy, K.ABz, out);
  } else {
#pragma unroll
    for (int k = 0; k < NA * NB * NC * ND; ++k) out[k] = 0.0;
  }
  CouplingArgs b = a;
  if (!active) b.rket = nullptr;
  coupling_scatter<LA, LB, LC, LD, true>(b, P, K, out);
}

// --------------------------------------------------- Schwarz (a11) kernel
// Q(pair) = sqrt(max_i |(ij|ij)_raw|) over the diagonal of the pair's
// (ab|ab) block, no primitive screening (Libint2Iface.cpp:3361-3404).
template <int LA, int LB>
__global__ __launch_bounds__(256) void schwarz_kernel(PairInfo* pairs, int npairs, const double* prim,
                                                      const double* boys, const double* norma) {
  constexpr int NA = ncart(LA), NB = ncart(LB), N2 = NA * NB;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  PairInfo P = pairs[i];
  double out[N2 * N2];
  const double* pp = prim + (size_t)P.poff * kPrimPairStride;
  contracted_quartet<LA, LB, LA, LB>(pp, P.nprim, pp, P.nprim, boys, P.ABx, P.ABy, P.ABz, P.ABx,
                                     P.ABy, P.ABz, out);
  double mx = 0.0, mxn = 0.0;
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) {
      const int k = ia * NB + ib;
      const double v = fabs(out[k * N2 + k]);
      const double nn = norma[P.bfa + ia] * norma[P.bfb + ib];
      mx = fmax(mx, v);
      mxn = fmax(mxn, v * nn * nn);
    }
  pairs[i].Q = sqrt(mx);
  pairs[i].Qn = sqrt(mxn);
}

template <int LA, int LB, int LC, int LD>
static void launch_jk_t(const JKArgs& a, hipStream_t s) {
  if (a.ntasks <= 0) return;
  const int blocks = (a.ntasks + 3) / 4;
  hipLaunchKernelGGL((jk_kernel<LA, LB, LC, LD>), dim3(blocks), dim3(256), 0, s, a);
}

template <int LA, int LB, int LC, int LD>
static void launch_jk_strip_t(const JKArgs& a, hipStream_t s) {
  if (a.ntasks <= 0) return;
  const size_t lds = (size_t)(ncart(LA) + ncart(LB)) * a.lds_n * sizeof(double) +
                     (strip_lds_boys<LA, LB, LC, LD>(a.lds_n, a.block) ? kLdsBoysBytes : 0) +
                     (strip_lds_boys32<LA, LB, LC, LD>(a.lds_n, a.block, a.any_mixed) ? kLdsBoys32Bytes : 0);
  hipLaunchKernelGGL((jk_strip_kernel<LA, LB, LC, LD>), dim3(a.ntasks), dim3(a.block > 0 ? a.block : 256), lds, s,
                     a);
}

template <int LA, int LB, int LC, int LD>
static void launch_cp_t(const CouplingArgs& a, hipStream_t s) {
  if (a.ntasks <= 0) return;
  const int blocks = (a.ntasks + 3) / 4;
  hipLaunchKernelGGL((coupling_kernel<LA, LB, LC, LD>), dim3(blocks), dim3(256), 0, s, a);
}

template <int LA, int LB>
static void launch_schwarz_t(PairInfo* pairs, int npairs, const double* prim, const double* boys, const double* norma,
                             hipStream_t s) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL((schwarz_kernel<LA, LB>), dim3((npairs + 255) / 256), dim3(256), 0, s, pairs, npairs, prim, boys,
                     norma);
}

}  // namespace lowdin
