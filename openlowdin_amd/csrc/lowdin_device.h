// Shared host/device data layout of the direct Fock build.
//
// A species' basis lives in HBM as
//   * shells:     ShellInfo[nshell]   (l, first function, primitive range, centre)
//   * primitives: exponent and coefficient*N(alpha,l) per primitive
//   * norma:      1/sqrt(S_raw,ii) per Cartesian function
// and its significant shell pairs (the reference's obs_shellpair_list,
// Libint2Iface.cpp:3248-3342) are grouped by pair class (la, lb), la >= lb,
// each pair in canonical orientation (the higher-l shell first) with its
// surviving primitive pairs precomputed in PrimPair records:
//   PrimPair = {p, 1/(2p), P(3), P-A(3), K, 0}
//   K = c_a c_b N_a N_b exp(-ab/p |AB|^2) sqrt(2) pi^(5/4) / p
// so that a primitive quartet's prefactor 2 pi^(5/2)/(p q sqrt(p+q)) K_ab K_cd
// is K_bra K_ket / sqrt(p+q).
#pragma once
#include <stdint.h>

namespace lowdin {

constexpr int kMaxL = 2;                 // s, p, d shells (Cartesian)
constexpr int kPrimPairStride = 10;      // doubles per PrimPair record
constexpr int kMaxStripPrims = 64;       // bra primitive pairs an LDS-strip workgroup stages
constexpr double kPrimPairDrop = 69.0;   // drop primitive pairs with ab/p |AB|^2 > 69 (exp < 1e-30)

// Boys table: F_m(T_i), m = 0..kBoysCols-2, at T_i = i * kBoysH, plus
// exp(-T_i) in column 0.  Seventh-order Taylor expansion around the nearest
// grid point (|dT| <= kBoysH/2) gives ~1e-15 relative error; T >= kBoysTmax
// uses F_0 = sqrt(pi/T)/2 and upward recursion with its exp(-T) term.
constexpr int kBoysMmax = 4 * kMaxL;     // highest order a quartet needs
constexpr int kBoysTaylor = 8;           // Taylor terms
constexpr int kBoysCols = 1 + kBoysMmax + kBoysTaylor;  // exp + F_0..F_{mmax+7}
constexpr double kBoysH = 0.1;
constexpr double kBoysInvH = 10.0;
constexpr double kBoysTmax = 40.0;
constexpr int kBoysRows = 401;
static_assert(kBoysRows == (int)(kBoysTmax * kBoysInvH) + 1, "FP64 Boys table rows must cover [0, kBoysTmax]");

// FP32 Boys table (mixed precision): its own grid constants and its own
// layout (exp, then F_0..F_{mmax+4} for a fifth-order Taylor step), built
// separately from the FP64 one, so that a change of the FP64 table's range
// or columns cannot move FP32 reads off their table.
constexpr int kBoysF32Taylor = 5;
constexpr int kBoysF32Cols = 1 + kBoysMmax + kBoysF32Taylor;
constexpr float kBoysF32H = 0.1f;
constexpr float kBoysF32InvH = 10.0f;
constexpr float kBoysF32Tmax = 40.0f;
constexpr int kBoysF32Rows = 401;
static_assert(kBoysF32Rows == (int)(kBoysF32Tmax * kBoysF32InvH) + 1, "FP32 Boys table rows must cover [0, Tmax]");
           // T = 0 .. 40 inclusive

struct ShellInfo {
  int l;
  int bf;      // first basis function
  int poff;    // first primitive
  int nprim;
  double x, y, z;
};

struct PairInfo {   // one significant shell pair, canonical orientation
  int sa, sb;       // shells, l[sa] >= l[sb]
  int bfa, bfb;     // first functions of sa, sb
  int poff;         // first PrimPair record
  int nprim;        // surviving primitive pairs
  int la, lb;
  double Q;         // raw Schwarz bound of the pair (the reference's screening convention)
  double ABx, ABy, ABz;
  double Qn;        // normalised Schwarz bound: max over the pair's functions (mixed-precision test)
};

struct Task {       // one wavefront: bra pair x a run of <= 64 ket pairs
  int bra;
  int j0, j1;
  int flags;
};

}  // namespace lowdin
