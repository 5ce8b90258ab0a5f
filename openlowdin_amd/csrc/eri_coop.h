// Wave-cooperative J/K kernel for the high angular momentum classes.
//
// The per-lane kernels of eri_core.h hold a whole quartet's recursion in
// one lane's registers.  From total L = 6 up ((dp|dp), (dd|pp), (dd|ds),
// (dd|dp), (dd|dd)) that state no longer fits: the kernels spill kilobytes
// per lane to scratch, and the few quartets of those classes (3959 (dd|dd)
// of C20H42 / cc-pVDZ: 62 wavefronts for 1024 SIMDs) leave the GPU idle.
// Here one wavefront evaluates one quartet and the 64 lanes share each
// recursion level:
//   * the Obara-Saika VRR ([e0|f0]^(m) from the scaled Boys values) and the
//     Head-Gordon-Pople HRR ((ab|cd) from [e0|f0]) run as precomputed
//     rounds of at most 64 independent operations, one per lane, whose
//     operands live in the wavefront's LDS region (gen_eri.py, CoopGen,
//     writes each class's round table and LDS layout);
//   * the contracted [e0|f0] accumulate in LDS over the primitive quartets;
//   * the J/K contraction of Libint2Iface.cpp:562-586 splits the six
//     updates over the lanes (one output element per lane, its sum over
//     the other two indices in LDS).
// A workgroup owns one bra pair and a run of kets (the strip kernel's
// task), its wavefronts take kets in turn.  With LDS room for the bra's two
// row strips ((NA + NB) x n doubles), J_ab and the four exchange blocks
// accumulate there and are flushed once per workgroup; J_cd is one global
// atomic per element; without strips every update is a global atomic.
// Within one wavefront LDS operations complete in issue order, so a round
// reads what the previous round wrote without a barrier.
#pragma once
#include "eri_core.h"

namespace lowdin {

template <int LA, int LB, int LC, int LD>
struct CoopProgram;

// Fixed slots of a wavefront's LDS region (doubles).
constexpr int kCoopZero = 18;   // 18 and 19: always 0.0 (m and m+1 operands of absent terms)
constexpr int kCoopAcc = 22;    // first contracted accumulator (20: target of idle lanes)

__device__ __forceinline__ double lds_at(const double* R, unsigned byteoff) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(R) + byteoff);
}
__device__ __forceinline__ double& lds_ref(double* R, unsigned byteoff) {
  return *reinterpret_cast<double*>(reinterpret_cast<char*>(R) + byteoff);
}

// Independent rounds a lane takes at once (COOP_GROUP of gen_eri.py).
constexpr int kCoopGroup = 4;

// Group g of a stage: 4 bits (rounds - 1, kind << 3) packed 16 per word.
template <unsigned long long G0, unsigned long long G1>
__device__ __forceinline__ unsigned coop_code(int g) {
  return (unsigned)(((g < 16 ? G0 : G1) >> (4 * (g & 15))) & 15ull);
}

// This lane's descriptors of group g, whose first round q points at.
template <unsigned long long G0, unsigned long long G1>
__device__ __forceinline__ void coop_load(const uint2* q, int g, uint2* d) {
  const int nr = (int)(coop_code<G0, G1>(g) & 7u) + 1;
#pragma unroll
  for (int k = 0; k < kCoopGroup; ++k)
    if (k < nr) d[k] = q[k * 64];
}

// Doubles of one wavefront's region: the program region, then the six
// normalised density blocks (ab, cd, ac, bd, ad, bc), rounded up to even.
template <int LA, int LB, int LC, int LD>
__host__ __device__ constexpr int coop_wave_doubles() {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  return (CoopProgram<LA, LB, LC, LD>::kRegion + NA * NB + NC * ND + NA * NC + NB * ND + NA * ND + NB * NC + 1) & ~1;
}

// F_m(T) for one order m per lane (the cooperative kernels need F_0..F_L
// on lanes 0..L): below kBoysTmax the seventh-order Taylor expansion of F_m
// itself around the nearest tabulated point (columns 1 + m .. 8 + m of the
// row), F_m(T) = sum_k F_{m+k}(T_i) (T_i - T)^k / k!; above, the asymptotic
// F_0 = sqrt(pi / T) / 2 and the upward recursion
// F_{k+1} = ((2k + 1) F_k - e^-T) / (2T) up to m.  No array of orders per
// lane, no downward recursion.
__device__ __forceinline__ double boys_one(double T, int m, const double* __restrict__ tab) {
  if (T < kBoysTmax) {
    const int i = (int)(T * kBoysInvH + 0.5);
    const double x = (double)i * kBoysH - T;
    const double* r = tab + i * kBoysCols + 1 + m;
    double s = r[7];
    s = fma(s, x * (1.0 / 7.0), r[6]);
    s = fma(s, x * (1.0 / 6.0), r[5]);
    s = fma(s, x * (1.0 / 5.0), r[4]);
    s = fma(s, x * (1.0 / 4.0), r[3]);
    s = fma(s, x * (1.0 / 3.0), r[2]);
    s = fma(s, x * (1.0 / 2.0), r[1]);
    return fma(s, x, r[0]);
  }
  const double rt = rsqrt_nr(T);
  const double o2t = 0.5 * rt * rt;
  const double eT = exp(-T);
  double f = 0.88622692545275801365 * rt;
  for (int k = 0; k < m; ++k) f = ((double)(2 * k + 1) * f - eT) * o2t;
  return f;
}

// One contracted quartet of class (LA LB|LC LD), raw (unnormalised
// functions), by one wavefront: the [e0|f0] accumulators are cleared, every
// primitive quartet (bra primitive pairs bprim[0..nbp), ket kprim[0..nkp))
// runs the VRR rounds and adds its [e0|f0]^(0) targets, then the HRR rounds
// leave (ab|cd) at R + kOut.  pg: this lane's column of the round table.
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void coop_quartet(double* R, const uint2* pg, gcptr<double> bprim, int nbp,
                                             gcptr<double> kprim, int nkp, const double* boys, const PairInfo& P,
                                             const PairInfo& K, int lane) {
  using Pg = CoopProgram<LA, LB, LC, LD>;
  constexpr int L = LA + LB + LC + LD;
  constexpr int NACC = EriClass<LA, LB, LC, LD>::nacc;
  for (int k = lane; k < NACC; k += 64) R[kCoopAcc + k] = 0.0;
  for (int kp = 0; kp < nkp; ++kp) {
    const gcptr<double> kq = kprim + kp * kPrimPairStride;
    const double q = kq[0], o2q = kq[1];
    const double Qx = kq[2], Qy = kq[3], Qz = kq[4];
    const double Kcd = kq[8];
    for (int ip = 0; ip < nbp; ++ip) {
      const gcptr<double> bp = bprim + ip * kPrimPairStride;
      const double p = bp[0], o2p = bp[1];
      const double PQx = bp[2] - Qx, PQy = bp[3] - Qy, PQz = bp[4] - Qz;
      const double rs = rsqrt_nr(p + q);
      const double inv = rs * rs;
      const double rop = q * inv, roq = p * inv;
      const double T = p * rop * (PQx * PQx + PQy * PQy + PQz * PQz);
      const double pref = bp[8] * Kcd * rs;
      // lane m <= L: the scaled Boys value F_m(T) * pref
      if (lane <= L) R[Pg::kBase + lane] = boys_one(T, lane, boys) * pref;
      if (lane == 0) {
        R[0] = bp[5];
        R[1] = -rop * PQx;
        R[2] = bp[6];
        R[3] = -rop * PQy;
        R[4] = bp[7];
        R[5] = -rop * PQz;
        R[6] = kq[5];
        R[7] = roq * PQx;
        R[8] = kq[6];
        R[9] = roq * PQy;
        R[10] = kq[7];
        R[11] = roq * PQz;
      }
      const double o2pq = 0.5 * inv;
      // VRR: groups of independent rounds; a lane loads every descriptor
      // and operand of a group before it writes, and the next group's
      // descriptors while it computes
      {
        const uint2* q = pg;
        uint2 dn[kCoopGroup];
        coop_load<Pg::kVrrG[0], Pg::kVrrG[1]>(q, 0, dn);
#pragma unroll 1
        for (int g = 0; g < Pg::kVrrGroups; ++g) {
          const unsigned c = coop_code<Pg::kVrrG[0], Pg::kVrrG[1]>(g);
          const int nr = (int)(c & 7u) + 1;
          uint2 d[kCoopGroup];
#pragma unroll
          for (int k = 0; k < kCoopGroup; ++k) d[k] = dn[k];
          q += nr * 64;
          if (g + 1 < Pg::kVrrGroups) coop_load<Pg::kVrrG[0], Pg::kVrrG[1]>(q, g + 1, dn);
          double v[kCoopGroup];
          if (c & 8u) {
#pragma unroll
            for (int k = 0; k < kCoopGroup; ++k)
              if (k < nr) {
                const unsigned s0 = d[k].x & 0xfff8u, k1 = d[k].x & 7u;
                const unsigned s2 = (d[k].x >> 16) & 0xfff8u, k2 = (d[k].x >> 16) & 7u;
                const unsigned i = (d[k].y >> 16) & 3u;
                v[k] = R[6 + 2 * i] * lds_at(R, s0) + R[7 + 2 * i] * lds_at(R, s0 + 8) +
                       ((double)k1 * o2q) * (lds_at(R, s2) - roq * lds_at(R, s2 + 8)) +
                       ((double)k2 * o2pq) * lds_at(R, d[k].y & 0xfff8u);
              }
          } else {
#pragma unroll
            for (int k = 0; k < kCoopGroup; ++k)
              if (k < nr) {
                const unsigned s0 = d[k].x & 0xfff8u, k1 = d[k].x & 7u;
                const unsigned s2 = (d[k].x >> 16) & 0xfff8u;
                const unsigned i = (d[k].y >> 16) & 3u;
                v[k] = R[2 * i] * lds_at(R, s0) + R[1 + 2 * i] * lds_at(R, s0 + 8) +
                       ((double)k1 * o2p) * (lds_at(R, s2) - rop * lds_at(R, s2 + 8));
              }
          }
#pragma unroll
          for (int k = 0; k < kCoopGroup; ++k)
            if (k < nr) lds_ref(R, (d[k].y >> 16) & 0xfff8u) = v[k];
        }
        // accumulate the [e0|f0]^(0) targets
        coop_load<Pg::kAccG[0], Pg::kAccG[1]>(q, 0, dn);
#pragma unroll 1
        for (int g = 0; g < Pg::kAccGroups; ++g) {
          const int nr = (int)(coop_code<Pg::kAccG[0], Pg::kAccG[1]>(g) & 7u) + 1;
          uint2 d[kCoopGroup];
#pragma unroll
          for (int k = 0; k < kCoopGroup; ++k) d[k] = dn[k];
          q += nr * 64;
          if (g + 1 < Pg::kAccGroups) coop_load<Pg::kAccG[0], Pg::kAccG[1]>(q, g + 1, dn);
          double v[kCoopGroup];
#pragma unroll
          for (int k = 0; k < kCoopGroup; ++k)
            if (k < nr) v[k] = lds_at(R, d[k].y) + lds_at(R, d[k].x);
#pragma unroll
          for (int k = 0; k < kCoopGroup; ++k)
            if (k < nr) lds_ref(R, d[k].y) = v[k];
        }
      }
    }
  }
  // HRR groups
  if (lane == 0) {
    R[12] = P.ABx;
    R[13] = P.ABy;
    R[14] = P.ABz;
    R[15] = K.ABx;
    R[16] = K.ABy;
    R[17] = K.ABz;
  }
  {
    const uint2* q = pg + (Pg::kVrrRounds + Pg::kAccRounds) * 64;
    uint2 dn[kCoopGroup];
    coop_load<Pg::kHrrG[0], Pg::kHrrG[1]>(q, 0, dn);
#pragma unroll 1
    for (int g = 0; g < Pg::kHrrGroups; ++g) {
      const unsigned c = coop_code<Pg::kHrrG[0], Pg::kHrrG[1]>(g);
      const int nr = (int)(c & 7u) + 1;
      const int cb = (c & 8u) ? 15 : 12;
      uint2 d[kCoopGroup];
#pragma unroll
      for (int k = 0; k < kCoopGroup; ++k) d[k] = dn[k];
      q += nr * 64;
      if (g + 1 < Pg::kHrrGroups) coop_load<Pg::kHrrG[0], Pg::kHrrG[1]>(q, g + 1, dn);
      double v[kCoopGroup];
#pragma unroll
      for (int k = 0; k < kCoopGroup; ++k)
        if (k < nr) v[k] = lds_at(R, d[k].x & 0xffffu) + R[cb + (d[k].y & 3u)] * lds_at(R, d[k].x >> 16);
#pragma unroll
      for (int k = 0; k < kCoopGroup; ++k)
        if (k < nr) lds_ref(R, d[k].y & 0xfff8u) = v[k];
    }
  }
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256, 2) void jk_coop_kernel(JKArgs a) {
  using Pg = CoopProgram<LA, LB, LC, LD>;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND;
  constexpr int kWave = coop_wave_doubles<LA, LB, LC, LD>();
  constexpr int oDab = Pg::kRegion, oDcd = oDab + NAB, oDac = oDcd + NCD, oDbd = oDac + NA * NC,
                oDad = oDbd + NB * ND, oDbc = oDad + NA * ND;
  static_assert(NAB <= 64, "one J_ab element per lane");
  extern __shared__ double lds[];
  __shared__ unsigned int wg_q, wg_p;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nw = blockDim.x >> 6;
  const Task t = a.tasks[blockIdx.x];
  const MolDesc& M = a.mols[t.flags];
  const gcptr<PairInfo> pairs = (gcptr<PairInfo>)M.pairs;
  const gcptr<double> prim = (gcptr<double>)M.prim;
  const gcptr<double> D = (gcptr<double>)M.D;
  const gcptr<double> nrm = (gcptr<double>)M.norma;
  const gptr<double> G = (gptr<double>)M.Graw;
  const PairInfo P = load_pair(pairs + M.off[a.cbra] + t.bra);
  const int n = M.n, ns = M.nshell;
  const int sn = a.lds_n;                 // strip width (0: no strips)
  double* sa = lds;
  double* sb = lds + NA * sn;
  double* R = lds + (NA + NB) * sn + w * kWave;
  for (int k = tid; k < (NA + NB) * sn; k += blockDim.x) lds[k] = 0.0;
  if (tid == 0) {
    wg_q = 0;
    wg_p = 0;
  }
  if (lane < 2) R[kCoopZero + lane] = 0.0;
  // normalised bra density block; lane k < NAB owns J_ab element k
  for (int k = lane; k < NAB; k += 64) {
    const int ia = k / NB, ib = k % NB;
    R[oDab + k] = D[(size_t)(P.bfa + ia) * n + P.bfb + ib] * nrm[P.bfa + ia] * nrm[P.bfb + ib];
  }
  double jab = 0.0;
  unsigned int myq = 0, myp = 0;
  const double sbra = (P.sa != P.sb ? 2.0 : 1.0);
  const double hf = M.half_factor;
  const gcptr<double> bprim = prim + (size_t)P.poff * kPrimPairStride;
  const uint2* pg = Pg::prog() + lane;   // this lane's column of the round table
  __syncthreads();
  for (int j = t.j0 + w; j < t.j1; j += nw) {
    const PairInfo K = load_pair(pairs + M.off[a.cket] + j);
    if (M.screen) {
      const double dn = quartet_dnorm((gcptr<double>)M.Dn, ns, P, K);
      if (dn * P.Q * K.Q < a.eps) continue;
    }
    myq += 1;
    myp += (unsigned int)(P.nprim * K.nprim);
    const gcptr<double> kprim = prim + (size_t)K.poff * kPrimPairStride;
    coop_quartet<LA, LB, LC, LD>(R, pg, bprim, P.nprim, kprim, K.nprim, a.boys, P, K, lane);
    // J/K contraction of the (ab|cd) block (Libint2Iface.cpp:562-586)
    const double* out = R + Pg::kOut;
    const double deg = sbra * (K.sa != K.sb ? 2.0 : 1.0) * ((a.same_class && j == t.bra) ? 1.0 : 2.0);
    for (int k = lane; k < NCD + NA * NC + NB * ND + NA * ND + NB * NC; k += 64) {
      int r0, c0, o;
      if (k < NCD) {
        r0 = K.bfa + k / ND; c0 = K.bfb + k % ND; o = oDcd + k;
      } else if (k < NCD + NA * NC) {
        const int u = k - NCD;
        r0 = P.bfa + u / NC; c0 = K.bfa + u % NC; o = oDac + u;
      } else if (k < NCD + NA * NC + NB * ND) {
        const int u = k - NCD - NA * NC;
        r0 = P.bfb + u / ND; c0 = K.bfb + u % ND; o = oDbd + u;
      } else if (k < NCD + NA * NC + NB * ND + NA * ND) {
        const int u = k - NCD - NA * NC - NB * ND;
        r0 = P.bfa + u / ND; c0 = K.bfb + u % ND; o = oDad + u;
      } else {
        const int u = k - NCD - NA * NC - NB * ND - NA * ND;
        r0 = P.bfb + u / NC; c0 = K.bfa + u % NC; o = oDbc + u;
      }
      R[o] = D[(size_t)r0 * n + c0] * nrm[r0] * nrm[c0];
    }
    if (lane < NAB) {
      double s = 0.0;
#pragma unroll 6
      for (int cd = 0; cd < NCD; ++cd) s += out[lane * NCD + cd] * R[oDcd + cd];
      jab += deg * s;
    }
    if (lane < NCD) {
      double s = 0.0;
#pragma unroll 6
      for (int ab = 0; ab < NAB; ++ab) s += out[ab * NCD + lane] * R[oDab + ab];
      const int fc = K.bfa + lane / ND, fd = K.bfb + lane % ND;
      global_add(G + (size_t)fc * n + fd, deg * s * nrm[fc] * nrm[fd]);
    }
    const double sk = deg * hf;
    if (lane < NA * NC) {  // K_ac = sum_bd (ab|cd) D_bd
      const int ia = lane / NC, ic = lane % NC;
      double s = 0.0;
#pragma unroll 6
      for (int ib = 0; ib < NB; ++ib)
#pragma unroll 6
        for (int id = 0; id < ND; ++id) s += out[((ia * NB + ib) * NC + ic) * ND + id] * R[oDbd + ib * ND + id];
      const int fc = K.bfa + ic;
      if (sn) atomicAdd(&sa[ia * sn + fc], sk * s * nrm[fc]);
      else global_add(G + (size_t)(P.bfa + ia) * n + fc, sk * s * nrm[fc] * nrm[P.bfa + ia]);
    }
    if (lane < NB * ND) {  // K_bd = sum_ac (ab|cd) D_ac
      const int ib = lane / ND, id = lane % ND;
      double s = 0.0;
#pragma unroll 6
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll 6
        for (int ic = 0; ic < NC; ++ic) s += out[((ia * NB + ib) * NC + ic) * ND + id] * R[oDac + ia * NC + ic];
      const int fd = K.bfb + id;
      if (sn) atomicAdd(&sb[ib * sn + fd], sk * s * nrm[fd]);
      else global_add(G + (size_t)(P.bfb + ib) * n + fd, sk * s * nrm[fd] * nrm[P.bfb + ib]);
    }
    if (lane < NA * ND) {  // K_ad = sum_bc (ab|cd) D_bc
      const int ia = lane / ND, id = lane % ND;
      double s = 0.0;
#pragma unroll 6
      for (int ib = 0; ib < NB; ++ib)
#pragma unroll 6
        for (int ic = 0; ic < NC; ++ic) s += out[((ia * NB + ib) * NC + ic) * ND + id] * R[oDbc + ib * NC + ic];
      const int fd = K.bfb + id;
      if (sn) atomicAdd(&sa[ia * sn + fd], sk * s * nrm[fd]);
      else global_add(G + (size_t)(P.bfa + ia) * n + fd, sk * s * nrm[fd] * nrm[P.bfa + ia]);
    }
    if (lane < NB * NC) {  // K_bc = sum_ad (ab|cd) D_ad
      const int ib = lane / NC, ic = lane % NC;
      double s = 0.0;
#pragma unroll 6
      for (int ia = 0; ia < NA; ++ia)
#pragma unroll 6
        for (int id = 0; id < ND; ++id) s += out[((ia * NB + ib) * NC + ic) * ND + id] * R[oDad + ia * ND + id];
      const int fc = K.bfa + ic;
      if (sn) atomicAdd(&sb[ib * sn + fc], sk * s * nrm[fc]);
      else global_add(G + (size_t)(P.bfb + ib) * n + fc, sk * s * nrm[fc] * nrm[P.bfb + ib]);
    }
  }
  // J_ab of this wavefront (lane-owned elements)
  if (lane < NAB && jab != 0.0) {
    const int ia = lane / NB, ib = lane % NB;
    const double v = jab * nrm[P.bfb + ib];
    if (sn) atomicAdd(&sa[ia * sn + P.bfb + ib], v);
    else global_add(G + (size_t)(P.bfa + ia) * n + P.bfb + ib, v * nrm[P.bfa + ia]);
  }
  if (a.counts && lane == 0) {
    atomicAdd(&wg_q, myq);
    atomicAdd(&wg_p, myp);
  }
  __syncthreads();
  if (sn) {
    for (int r = 0; r < NA; ++r) {
      const double nr = nrm[P.bfa + r];
      for (int c = tid; c < n; c += blockDim.x) {
        const double v = sa[r * sn + c];
        if (v != 0.0) global_add(G + (size_t)(P.bfa + r) * n + c, v * nr);
      }
    }
    for (int r = 0; r < NB; ++r) {
      const double nr = nrm[P.bfb + r];
      for (int c = tid; c < n; c += blockDim.x) {
        const double v = sb[r * sn + c];
        if (v != 0.0) global_add(G + (size_t)(P.bfb + r) * n + c, v * nr);
      }
    }
  }
  if (a.counts && tid == 0) {
    a.counts[3 * blockIdx.x] = wg_q;
    a.counts[3 * blockIdx.x + 1] = wg_p;
    a.counts[3 * blockIdx.x + 2] = 0;
  }
}

// Inter-species Coulomb coupling (Libint2Iface.cpp:1111-1294) of one class
// with one wavefront per quartet: the quartet as above, then the drop of
// raw integrals below a.drop and the two contractions of coupling_scatter
// (eri_core.h) split over the lanes -- lane k < NA NB sums result element
// (a b) on the bra species over the ket density's element set, lane k <
// NC ND element (c d) on the ket species over the bra density's.  A
// workgroup owns one coupling task (a bra pair and up to 64 ket pairs);
// its wavefronts take the kets in turn.  No private segment: the state of
// a quartet lives in the wavefront's LDS region.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256, 2) void coupling_coop_kernel(CouplingArgs a) {
  using Pg = CoopProgram<LA, LB, LC, LD>;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND;
  constexpr int kWave = coop_wave_doubles<LA, LB, LC, LD>();
  constexpr int oWb = Pg::kRegion, oWk = oWb + NAB;   // bra and ket density weights
  static_assert(NAB <= 64 && NCD <= 64, "one result element per lane");
  static_assert(oWk + NCD <= kWave, "weights inside the wavefront's region");
  extern __shared__ double lds[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nw = blockDim.x >> 6;
  const Task t = a.tasks[blockIdx.x];
  const PairInfo P = load_pair((gcptr<PairInfo>)a.bra + t.bra);
  double* R = lds + w * kWave;
  if (lane < 2) R[kCoopZero + lane] = 0.0;
  const bool bdiag = P.sa == P.sb;
  // weights of the bra density (result on the ket species): element set
  // ia <= ib within a diagonal shell pair, doubled off the function
  // diagonal, normalised
  if (a.rket && lane < NAB) {
    const int ia = lane / NB, ib = lane % NB;
    const int fa = P.bfa + ia, fb = P.bfb + ib;
    double v = 0.0;
    if (!bdiag || ia <= ib)
      v = (fa != fb ? 2.0 : 1.0) * a.dbra[(size_t)fa * a.nbra + fb] * a.nbra_norma[fa] * a.nbra_norma[fb];
    R[oWb + lane] = v;
  }
  const gcptr<double> bprim = (gcptr<double>)a.bprim + (size_t)P.poff * kPrimPairStride;
  const uint2* pg = Pg::prog() + lane;
  const double drop = a.drop;
  for (int j = t.j0 + w; j < t.j1; j += nw) {
    const PairInfo K = load_pair((gcptr<PairInfo>)a.ket + j);
    const gcptr<double> kprim = (gcptr<double>)a.kprim + (size_t)K.poff * kPrimPairStride;
    coop_quartet<LA, LB, LC, LD>(R, pg, bprim, P.nprim, kprim, K.nprim, a.boys, P, K, lane);
    const double* out = R + Pg::kOut;
    const bool kdiag = K.sa == K.sb;
    if (a.rbra) {
      if (lane < NCD) {
        const int ic = lane / ND, id = lane % ND;
        const int fc = K.bfa + ic, fd = K.bfb + id;
        double v = 0.0;
        if (!kdiag || ic <= id)
          v = (fc != fd ? 2.0 : 1.0) * a.dket[(size_t)fc * a.nket + fd] * a.nket_norma[fc] * a.nket_norma[fd];
        R[oWk + lane] = v;
      }
      if (lane < NAB) {
        const int ia = lane / NB, ib = lane % NB;
        if (!(bdiag && ia > ib)) {
          double s = 0.0;
          for (int k = 0; k < NCD; ++k) {
            const double x = out[lane * NCD + k];
            s += R[oWk + k] * (fabs(x) < drop ? 0.0 : x);
          }
          const int fa = P.bfa + ia, fb = P.bfb + ib;
          const int r = fa > fb ? fa : fb, c = fa > fb ? fb : fa;
          unsafeAtomicAdd(&a.rbra[(size_t)r * a.nbra + c], s * a.nbra_norma[fa] * a.nbra_norma[fb]);
        }
      }
    }
    if (a.rket && lane < NCD) {
      const int ic = lane / ND, id = lane % ND;
      if (!(kdiag && ic > id)) {
        double s = 0.0;
        for (int k = 0; k < NAB; ++k) {
          const double x = out[k * NCD + lane];
          s += R[oWb + k] * (fabs(x) < drop ? 0.0 : x);
        }
        const int fc = K.bfa + ic, fd = K.bfb + id;
        const int r = fc > fd ? fc : fd, c = fc > fd ? fd : fc;
        unsafeAtomicAdd(&a.rket[(size_t)r * a.nket + c], s * a.nket_norma[fc] * a.nket_norma[fd]);
      }
    }
  }
}

// Dynamic LDS a cooperative workgroup may use (gfx950: 160 KB per workgroup,
// less the kernel's static counters).
constexpr size_t kCoopLdsMax = 163840 - 256;

template <int LA, int LB, int LC, int LD>
static void coop_attr_t() {
  (void)hipFuncSetAttribute((const void*)jk_coop_kernel<LA, LB, LC, LD>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kCoopLdsMax);
  (void)hipFuncSetAttribute((const void*)coupling_coop_kernel<LA, LB, LC, LD>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCoopLdsMax);
}

// One workgroup of up to four wavefronts per coupling task.
template <int LA, int LB, int LC, int LD>
static bool launch_cp_coop_t(const CouplingArgs& a, hipStream_t s) {
  if (a.ntasks <= 0) return true;
  constexpr size_t wave = (size_t)coop_wave_doubles<LA, LB, LC, LD>() * sizeof(double);
  int W = 4;
  while (W > 1 && W * wave > kCoopLdsMax) --W;
  if (W * wave > kCoopLdsMax) return false;
  hipLaunchKernelGGL((coupling_coop_kernel<LA, LB, LC, LD>), dim3(a.ntasks), dim3(64 * W), W * wave, s, a);
  return true;
}

// One workgroup per task.  a.lds_n: row-strip width wanted (0: none);
// a.block: threads wanted (64, 128, 192 or 256).  Fewer wavefronts are used
// when strips and regions would not fit in LDS, no strips when even one
// wavefront would not fit beside them.
template <int LA, int LB, int LC, int LD>
static bool launch_jk_coop_t(JKArgs a, hipStream_t s) {
  if (a.ntasks <= 0) return true;
  constexpr size_t wave = (size_t)coop_wave_doubles<LA, LB, LC, LD>() * sizeof(double);
  size_t strip = (size_t)(ncart(LA) + ncart(LB)) * (size_t)a.lds_n * sizeof(double);
  int W = (a.block > 0 ? a.block : 256) / 64;
  W = W < 1 ? 1 : (W > 4 ? 4 : W);
  while (W > 1 && strip + W * wave > kCoopLdsMax) --W;
  if (strip + W * wave > kCoopLdsMax) {
    a.lds_n = 0;
    strip = 0;
    W = (int)(kCoopLdsMax / wave);
    W = W > 4 ? 4 : W;
  }
  if (W < 1) return false;
  hipLaunchKernelGGL((jk_coop_kernel<LA, LB, LC, LD>), dim3(a.ntasks), dim3(64 * W), strip + W * wave, s, a);
  return true;
}

}  // namespace lowdin
