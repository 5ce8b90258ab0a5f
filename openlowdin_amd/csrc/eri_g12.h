// Templates of the Gaussian-geminal (G12) kernels: the intra-species Fock
// build and the inter-species coupling with the operator
// sum_k c_k exp(-g_k r12^2) in place of 1/r12.
//
// Reference behaviour reproduced (openLOWDIN src/ints/Libint2Iface.cpp):
//   compute_g12_direct (2193-2365): the Fock build's quartet loops over the
//     pair list without a density screen (2263-2281), degeneracy
//     (2291-2294), the six J/K updates with the 0.5*factor exchange weight
//     (2324-2329), G = 0.25 (G + G^T) (2363);
//   compute_g12inter_direct (2367-2550): every pair s1 <= s2 of both
//     species, |raw| < 1e-10 dropped, the coupling build's element sets and
//     off-diagonal doubling (2495-2519).
// Primitive integrals: one-dimensional Gaussian moments (gen_eri.py,
// G12Gen) per geminal term, then the same HGP HRR as the Coulomb kernels.
#pragma once
#include "eri_core.h"

namespace lowdin {

template <int LA, int LB, int LC, int LD>
struct G12Class;
#define LOWDIN_G12_CLASS(a, b, c, d)                                                                 \
  template <>                                                                                        \
  struct G12Class<a, b, c, d> {                                                                      \
    template <typename R>                                                                            \
    __device__ static __forceinline__ void vrr(R XAx, R XAy, R XAz, R XCx, R XCy, R XCz, R s11,      \
                                               R s12, R s22, R pref, R* acc) {                       \
      g12vrr_##a##b##c##d<R>(XAx, XAy, XAz, XCx, XCy, XCz, s11, s12, s22, pref, acc);                \
    }                                                                                                \
  };

// Contracted raw G12 quartet (canonical orientation) for one lane.  Per
// primitive quartet and geminal term g (weight c), with D = pq + g(p+q):
//   [00|00] = c K_ab K_cd (pq sqrt(pi) / 2) D^(-3/2) exp(-pq g |PQ|^2 / D)
// (the pair records' K carry sqrt(2) pi^(5/4) / p), then the moment
// recursion of G12Gen into [e0|f0] and one HRR.
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void g12_quartet(const double* bprim, int nbp, const double* kprim, int nkp,
                                            const G12Params& gem, double ABx, double ABy, double ABz,
                                            double CDx, double CDy, double CDz, double* out) {
  using C = EriClass<LA, LB, LC, LD>;
  using G = G12Class<LA, LB, LC, LD>;
  double acc[C::nacc];
#pragma unroll
  for (int k = 0; k < C::nacc; ++k) acc[k] = 0.0;
  const double half_sqrt_pi = 0.88622692545275801365;
  for (int kp = 0; kp < nkp; ++kp) {
    const double* kq = kprim + kp * kPrimPairStride;
    const double q = kq[0];
    const double Qx = kq[2], Qy = kq[3], Qz = kq[4];
    const double QCx = kq[5], QCy = kq[6], QCz = kq[7];
    const double Kcd = kq[8];
    for (int ip = 0; ip < nbp; ++ip) {
      const double* bp = bprim + ip * kPrimPairStride;
      const double p = bp[0];
      const double PQx = bp[2] - Qx, PQy = bp[3] - Qy, PQz = bp[4] - Qz;
      const double R2 = PQx * PQx + PQy * PQy + PQz * PQz;
      const double pq = p * q, spq = p + q;
      const double base = bp[8] * Kcd * pq * half_sqrt_pi;
      for (int k = 0; k < gem.n; ++k) {
        const double g = gem.g[k];
        const double r = rsqrt_nr(pq + g * spq);   // D^(-1/2)
        const double iD = r * r;
        const double pref = gem.c[k] * base * iD * r * exp(-pq * g * R2 * iD);
        const double gqD = g * q * iD, gpD = g * p * iD;
        G::template vrr<double>(bp[5] - gqD * PQx, bp[6] - gqD * PQy, bp[7] - gqD * PQz, QCx + gpD * PQx,
                                QCy + gpD * PQy, QCz + gpD * PQz, 0.5 * (q + g) * iD, 0.5 * g * iD,
                                0.5 * (p + g) * iD, pref, acc);
      }
    }
  }
  C::template hrr<double>(acc, ABx, ABy, ABz, CDx, CDy, CDz, out);
}

// One lane per quartet (a wavefront: one bra pair, up to 64 kets), the J/K
// updates of the Coulomb wavefront kernel: the bra pair's J block reduced
// across the wavefront, the other five blocks as FP64 atomics.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void g12_jk_kernel(G12Args a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  const int wid = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (wid >= a.ntasks) return;
  const Task t = a.tasks[wid];
  const int lane = threadIdx.x & 63;
  const int j = t.j0 + lane;
  const bool active = j < t.j1;
  const PairInfo P = a.bra[t.bra];
  const PairInfo K = active ? a.ket[j] : P;
  const gcptr<double> nrm = (gcptr<double>)a.norma;
  const gptr<double> G = (gptr<double>)a.Graw;
  double jab[NA * NB];
#pragma unroll
  for (int k = 0; k < NA * NB; ++k) jab[k] = 0.0;
  if (active) {
    double out[NA * NB * NC * ND];
    g12_quartet<LA, LB, LC, LD>(a.prim + (size_t)P.poff * kPrimPairStride, P.nprim,
                                a.prim + (size_t)K.poff * kPrimPairStride, K.nprim, a.gem, P.ABx, P.ABy, P.ABz,
                                K.ABx, K.ABy, K.ABz, out);
    const double deg = (P.sa != P.sb ? 2.0 : 1.0) * (K.sa != K.sb ? 2.0 : 1.0) *
                       ((a.same_class && j == t.bra) ? 1.0 : 2.0);
    jk_quartet_update<LA, LB, LC, LD>(out, P, K, a.n, (gcptr<double>)a.D, nrm, G, deg, a.half_factor, jab);
#pragma unroll
    for (int k = 0; k < NA * NB; ++k) jab[k] *= deg;
  }
  const uint64_t amask = __ballot(active);
#pragma unroll
  for (int ia = 0; ia < NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) {
      const double s = wave_sum(jab[ia * NB + ib]);
      if (lane == 0 && amask)
        global_add(G + (size_t)(P.bfa + ia) * a.n + P.bfb + ib, s * nrm[P.bfa + ia] * nrm[P.bfb + ib]);
    }
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void g12_coupling_kernel(G12CouplingArgs ga) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  const CouplingArgs& a = ga.cp;
  const int wid = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (wid >= a.ntasks) return;
  const Task t = a.tasks[wid];
  const int j = t.j0 + (int)(threadIdx.x & 63);
  if (j >= t.j1) return;
  const PairInfo P = a.bra[t.bra];
  const PairInfo K = a.ket[j];
  double out[NA * NB * NC * ND];
  g12_quartet<LA, LB, LC, LD>(a.bprim + (size_t)P.poff * kPrimPairStride, P.nprim,
                              a.kprim + (size_t)K.poff * kPrimPairStride, K.nprim, ga.gem, P.ABx, P.ABy, P.ABz,
                              K.ABx, K.ABy, K.ABz, out);
  coupling_scatter<LA, LB, LC, LD>(a, P, K, out);
}

template <int LA, int LB, int LC, int LD>
static void launch_g12_t(const G12Args& a, hipStream_t s) {
  if (a.ntasks <= 0) return;
  hipLaunchKernelGGL((g12_jk_kernel<LA, LB, LC, LD>), dim3((unsigned)((a.ntasks + 3) / 4)), dim3(256), 0, s, a);
}

template <int LA, int LB, int LC, int LD>
static void launch_g12cp_t(const G12CouplingArgs& a, hipStream_t s) {
  if (a.cp.ntasks <= 0) return;
  hipLaunchKernelGGL((g12_coupling_kernel<LA, LB, LC, LD>), dim3((unsigned)((a.cp.ntasks + 3) / 4)), dim3(256), 0,
                     s, a);
}

}  // namespace lowdin
