/*
 * lowdin_oracle.c -- CPU FP64 restatement of openLOWDIN's direct two-particle
 * Fock build and inter-species coupling build.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or the timed CPU baseline.  The product library never
 * links or calls it.
 *
 * Algorithms (restated from the mathematics, not from libint2's code):
 *   - Boys function F_m(T): power series for T < 35 followed by downward
 *     recursion; erf closed form plus upward recursion for T >= 35.
 *   - Cartesian Gaussian integrals: McMurchie-Davidson Hermite expansion
 *     (Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure Theory",
 *     ch. 9).  The GPU kernels use Obara-Saika/Head-Gordon-Pople instead, so
 *     the two implementations do not share a recursion.
 *
 * Reference behaviour followed (file:line of openLOWDIN):
 *   - primitive normalisation embedded in raw coefficients and per-Cartesian
 *     "norma" = 1/sqrt(S_raw,ii):         src/ints/Libint2Iface.cpp:79-128
 *   - shell-pair list (same centre, or Frobenius norm of the raw overlap
 *     block >= 1e-12, lists ascending):   src/ints/Libint2Iface.cpp:3248-3342
 *   - Schwarz matrix Q = sqrt(max diag |(ab|ab)|_raw):
 *                                         src/ints/Libint2Iface.cpp:3345-3412
 *   - direct J/K build (unique quartets, Dnorm*Q*Q < eps screen for
 *     electrons only, degeneracy, 6 updates, 0.25 (G + G^T)):
 *                                         src/ints/Libint2Iface.cpp:418-618
 *   - coupling build (all s1<=s2 x os1<=os2, |raw v| < 1e-10 drop,
 *     permuted/non-permuted scatter, no symmetrisation):
 *                                         src/ints/Libint2Iface.cpp:1111-1294
 *   - 1-body S/T/V with norma:            src/ints/Libint2Iface.cpp:163-217
 *   - G12 (Gaussian-geminal) Fock and coupling builds, the same loops with
 *     sum_k c_k exp(-g_k r12^2) for 1/r12: src/ints/Libint2Iface.cpp:2193-2550
 *   - integral store over the Fock build's quartets (|raw| > 1e-10, eight
 *     images, norma):                     src/ints/Libint2Iface.cpp:2552-2742
 *   - inter-species integral store (all pairs, bf1 <= bf2 / obf1 <= obf2,
 *     |raw| >= 1e-10, four images):       src/ints/Libint2Iface.cpp:2744-2908
 *     (the first-quarter transforms of :620-921 and :1296-1622 are the
 *     contraction of these tensors with one coefficient row, done in
 *     oracle.py)
 *
 * Matrix layout: all n x n matrices are dense FP64 in the memory image the
 * reference's C wrappers produce (Libint2Iface.cpp:3510-3526, 3566-3589):
 * element (r, c) of the reference's row-major Eigen matrix at index r*n + c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXL 3                 /* per shell */
#define ORC_MAXLT (4 * ORC_MAXL)   /* total in a quartet */
#define ORC_MAXNC 10               /* Cartesians for l = 3 */
#define ORC_PI 3.14159265358979323846264338327950288

typedef struct {
  int nshell;
  const int *l;        /* [nshell] */
  const int *nprim;    /* [nshell] */
  const int *poff;     /* [nshell] offset into exps/coefs */
  const double *exps;  /* raw exponents */
  const double *coefs; /* raw (file) contraction coefficients */
  const double *xyz;   /* [3*nshell] shell origins, bohr */
} orc_basis;

/* Contracted Gaussian-type geminal sum_k c[k] exp(-g[k] r12^2) (NULL: 1/r12). */
typedef struct {
  int n;
  const double *c;
  const double *g;
} orc_gem;

static inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

/* Cartesian components in the standard order (xx, xy, xz, yy, yz, zz). */
static void cart_comps(int l, int c[][3]) {
  int k = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      c[k][0] = lx; c[k][1] = ly; c[k][2] = l - lx - ly; ++k;
    }
}

static double dfact(int n) { /* n!!, with (-1)!! = 1 */
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

/* ------------------------------------------------------------------ Boys */
void orc_boys(int mmax, double T, double *F) {
  if (T < 35.0) {
    const double eT = exp(-T);
    double term = 1.0 / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 2000; ++k) {
      term *= 2.0 * T / (2 * mmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    F[mmax] = eT * sum;
    for (int m = mmax - 1; m >= 0; --m)
      F[m] = (2.0 * T * F[m + 1] + eT) / (2 * m + 1);
  } else {
    const double eT = exp(-T), sT = sqrt(T);
    F[0] = 0.5 * sqrt(ORC_PI) * erf(sT) / sT;
    for (int m = 0; m < mmax; ++m)
      F[m + 1] = ((2 * m + 1) * F[m] - eT) / (2.0 * T);
  }
}

/* N(alpha, l) of libint2's convention: normalisation of x^l exp(-alpha r^2). */
double orc_prim_norm(double alpha, int l) {
  const double two_a = 2.0 * alpha;
  return sqrt(pow(2.0, l) * pow(two_a, l + 1.5) / (pow(ORC_PI, 1.5) * dfact(2 * l - 1)));
}

/* ------------------------------------------------ Hermite E coefficients */
/* E[i][j][t] for i <= la, j <= lb along one axis, without the exp(-mu X^2)
 * factor (applied once per primitive pair). */
#define EDIM (ORC_MAXL + 3) /* room for l+2 (kinetic) */
#define ETD (2 * ORC_MAXL + 4)
typedef double etab[EDIM][EDIM][ETD];

static void hermite_e(int la, int lb, double p, double xpa, double xpb, etab E) {
  const double o2p = 0.5 / p;
  memset(E, 0, sizeof(etab));
  E[0][0][0] = 1.0;
  for (int i = 0; i < la; ++i)
    for (int t = 0; t <= i + 1; ++t) {
      double v = xpa * E[i][0][t];
      if (t > 0) v += o2p * E[i][0][t - 1];
      if (t + 1 <= i) v += (t + 1) * E[i][0][t + 1];
      E[i + 1][0][t] = v;
    }
  for (int j = 0; j < lb; ++j)
    for (int i = 0; i <= la; ++i)
      for (int t = 0; t <= i + j + 1; ++t) {
        double v = xpb * E[i][j][t];
        if (t > 0) v += o2p * E[i][j][t - 1];
        if (t + 1 <= i + j) v += (t + 1) * E[i][j][t + 1];
        E[i][j + 1][t] = v;
      }
}

/* ------------------------------------------------ Hermite R integrals */
#define RD (ORC_MAXLT + 1)
/* R[t][u][v] = R^{(0)}_{tuv}(alpha, PQ) for t+u+v <= L.  Uses ping-pong
 * buffers over the auxiliary index n. */
static void hermite_r(int L, double alpha, const double pq[3], double *Rout /* RD^3 */) {
  static const int S = RD * RD * RD;
  double F[ORC_MAXLT + 1];
  const double T = alpha * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
  orc_boys(L, T, F);
  double bufA[RD * RD * RD], bufB[RD * RD * RD];
  double *cur = bufA, *prev = bufB;
  (void)S;
#define IX(t, u, v) (((t) * RD + (u)) * RD + (v))
  double m2a = 1.0;
  double pw[ORC_MAXLT + 1];
  for (int n = 0; n <= L; ++n) { pw[n] = m2a * F[n]; m2a *= -2.0 * alpha; }
  for (int n = L; n >= 0; --n) {
    const int lev = L - n;
    for (int t = 0; t <= lev; ++t)
      for (int u = 0; u <= lev - t; ++u)
        for (int v = 0; v <= lev - t - u; ++v) {
          double r;
          if (t == 0 && u == 0 && v == 0) r = pw[n];
          else if (t > 0) {
            r = pq[0] * prev[IX(t - 1, u, v)];
            if (t > 1) r += (t - 1) * prev[IX(t - 2, u, v)];
          } else if (u > 0) {
            r = pq[1] * prev[IX(t, u - 1, v)];
            if (u > 1) r += (u - 1) * prev[IX(t, u - 2, v)];
          } else {
            r = pq[2] * prev[IX(t, u, v - 1)];
            if (v > 1) r += (v - 1) * prev[IX(t, u, v - 2)];
          }
          cur[IX(t, u, v)] = r;
        }
    double *tmp = prev; prev = cur; cur = tmp;
  }
  /* after the last swap, prev holds level n = 0 */
  for (int t = 0; t <= L; ++t)
    for (int u = 0; u <= L - t; ++u)
      for (int v = 0; v <= L - t - u; ++v) Rout[IX(t, u, v)] = prev[IX(t, u, v)];
}

/* The same table for the geminal operator: R[t][u][v] = d^t/dX^t d^u/dY^u
 * d^v/dZ^v of sum_k c_k (pi^2/D_k)^(3/2) exp(-rho_k |PQ|^2), D_k = pq +
 * g_k (p + q), rho_k = pq g_k / D_k -- the integral of the operator between
 * two unit s-type Hermite Gaussians at P and Q, differentiated.  Each axis
 * factor h_t = d^t/dX^t exp(-rho X^2) obeys h_{t+1} = -2 rho (X h_t + t h_{t-1}). */
static void hermite_r_gem(int L, double p, double q, const double pq[3], const orc_gem *gem, double *Rout) {
  for (int t = 0; t <= L; ++t)
    for (int u = 0; u <= L - t; ++u)
      for (int v = 0; v <= L - t - u; ++v) Rout[IX(t, u, v)] = 0.0;
  for (int k = 0; k < gem->n; ++k) {
    const double g = gem->g[k], D = p * q + g * (p + q), rho = p * q * g / D;
    const double w = gem->c[k] * pow(ORC_PI * ORC_PI / D, 1.5);
    double h[3][ORC_MAXLT + 2];
    for (int d = 0; d < 3; ++d) {
      const double X = pq[d];
      h[d][0] = exp(-rho * X * X);
      if (L >= 1) h[d][1] = -2.0 * rho * X * h[d][0];
      for (int t = 1; t < L; ++t) h[d][t + 1] = -2.0 * rho * (X * h[d][t] + t * h[d][t - 1]);
    }
    for (int t = 0; t <= L; ++t)
      for (int u = 0; u <= L - t; ++u)
        for (int v = 0; v <= L - t - u; ++v) Rout[IX(t, u, v)] += w * h[0][t] * h[1][u] * h[2][v];
  }
}

/* ------------------------------------------------ primitive ERIs */
/* Unnormalised primitive Cartesian ERIs [ab|cd] for x^l exp(-alpha r^2)
 * primitives, all components, index ((ia*nb + ib)*nc + ic)*nd + id. */
static void prim_two_body(int la, int lb, int lc, int ld, const double *A, const double *B, const double *C,
                          const double *D, double a, double b, double c, double d, const orc_gem *gem,
                          double *out);

void orc_prim_eri(int la, int lb, int lc, int ld, const double *A, const double *B,
                  const double *C, const double *D, double a, double b, double c,
                  double d, double *out) {
  prim_two_body(la, lb, lc, ld, A, B, C, D, a, b, c, d, NULL, out);
}

/* Primitive integrals of the geminal operator, same layout as orc_prim_eri. */
void orc_prim_g12(int la, int lb, int lc, int ld, const double *A, const double *B, const double *C,
                  const double *D, double a, double b, double c, double d, int ng, const double *gc,
                  const double *gg, double *out) {
  const orc_gem gem = {ng, gc, gg};
  prim_two_body(la, lb, lc, ld, A, B, C, D, a, b, c, d, &gem, out);
}

/* McMurchie-Davidson: (ab|op|cd) = sum E^ab_tuv E^cd_tau,nu,phi (-1)^(tau+nu+phi)
 * R_{t+tau,u+nu,v+phi}, R the operator's Hermite table (Coulomb: Boys-based
 * with prefactor 2 pi^(5/2) / (pq sqrt(p+q)); geminal: hermite_r_gem). */
static void prim_two_body(int la, int lb, int lc, int ld, const double *A, const double *B, const double *C,
                          const double *D, double a, double b, double c, double d, const orc_gem *gem,
                          double *out) {
  const double p = a + b, q = c + d, rho = p * q / (p + q);
  double P[3], Q[3], PQ[3];
  double ab2 = 0, cd2 = 0;
  for (int k = 0; k < 3; ++k) {
    P[k] = (a * A[k] + b * B[k]) / p;
    Q[k] = (c * C[k] + d * D[k]) / q;
    PQ[k] = P[k] - Q[k];
    ab2 += (A[k] - B[k]) * (A[k] - B[k]);
    cd2 += (C[k] - D[k]) * (C[k] - D[k]);
  }
  const double pref = (gem ? 1.0 : 2.0 * pow(ORC_PI, 2.5) / (p * q * sqrt(p + q))) *
                      exp(-a * b / p * ab2) * exp(-c * d / q * cd2);
  static etab Eab[3], Ecd[3];
#pragma omp threadprivate(Eab, Ecd)
  for (int k = 0; k < 3; ++k) {
    hermite_e(la, lb, p, P[k] - A[k], P[k] - B[k], Eab[k]);
    hermite_e(lc, ld, q, Q[k] - C[k], Q[k] - D[k], Ecd[k]);
  }
  const int L = la + lb + lc + ld;
  static double R[RD * RD * RD];
#pragma omp threadprivate(R)
  if (gem)
    hermite_r_gem(L, p, q, PQ, gem, R);
  else
    hermite_r(L, rho, PQ, R);
  int ca[ORC_MAXNC][3], cb[ORC_MAXNC][3], cc[ORC_MAXNC][3], cdd[ORC_MAXNC][3];
  cart_comps(la, ca); cart_comps(lb, cb); cart_comps(lc, cc); cart_comps(ld, cdd);
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
  const int Lab = la + lb;
  /* H[t][u][v] per ket component: sum_{tau,nu,phi} (-1)^(tau+nu+phi) E E E R */
  double H[RD * RD * RD];
  for (int ic = 0; ic < nc; ++ic)
    for (int id = 0; id < nd; ++id) {
      const int *xc = cc[ic], *xd = cdd[id];
      const int tmax = xc[0] + xd[0], umax = xc[1] + xd[1], vmax = xc[2] + xd[2];
      for (int t = 0; t <= Lab; ++t)
        for (int u = 0; u <= Lab - t; ++u)
          for (int v = 0; v <= Lab - t - u; ++v) {
            double s = 0.0;
            for (int tau = 0; tau <= tmax; ++tau) {
              const double ex = Ecd[0][xc[0]][xd[0]][tau];
              if (ex == 0.0) continue;
              for (int nu = 0; nu <= umax; ++nu) {
                const double ey = Ecd[1][xc[1]][xd[1]][nu];
                if (ey == 0.0) continue;
                for (int phi = 0; phi <= vmax; ++phi) {
                  const double ez = Ecd[2][xc[2]][xd[2]][phi];
                  const double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                  s += sg * ex * ey * ez * R[IX(t + tau, u + nu, v + phi)];
                }
              }
            }
            H[IX(t, u, v)] = s;
          }
      for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
          const int *xa = ca[ia], *xb = cb[ib];
          double s = 0.0;
          for (int t = 0; t <= xa[0] + xb[0]; ++t) {
            const double ex = Eab[0][xa[0]][xb[0]][t];
            for (int u = 0; u <= xa[1] + xb[1]; ++u) {
              const double ey = Eab[1][xa[1]][xb[1]][u];
              for (int v = 0; v <= xa[2] + xb[2]; ++v)
                s += ex * ey * Eab[2][xa[2]][xb[2]][v] * H[IX(t, u, v)];
            }
          }
          out[((ia * nb + ib) * nc + ic) * nd + id] = pref * s;
        }
    }
}

/* ------------------------------------------------ contracted shells */
static void shell_xyz(const orc_basis *bs, int s, double X[3]) {
  X[0] = bs->xyz[3 * s]; X[1] = bs->xyz[3 * s + 1]; X[2] = bs->xyz[3 * s + 2];
}

/* Raw contracted ERIs (libint-normalised primitives, raw coefficients, no
 * norma) for shells (s1 s2 | s3 s4); s1, s2 from basis b12; s3, s4 from b34.
 * With prim_eps > 0, primitive pairs whose Gaussian-product factor
 * exp(-ab/(a+b) |AB|^2) is below prim_eps are skipped (the same pair-level
 * truncation the GPU build applies at 1e-30); prim_eps = 0 keeps all. */
static void shell_block(const orc_basis *b12, int s1, int s2, const orc_basis *b34, int s3, int s4,
                        double prim_eps, const orc_gem *gem, double *out);

void orc_eri_shell(const orc_basis *b12, int s1, int s2, const orc_basis *b34, int s3,
                   int s4, double prim_eps, double *out) {
  shell_block(b12, s1, s2, b34, s3, s4, prim_eps, NULL, out);
}

/* Raw contracted integrals of the geminal operator for (s1 s2 | s3 s4). */
void orc_g12_shell(const orc_basis *b12, int s1, int s2, const orc_basis *b34, int s3, int s4, int ng,
                   const double *gc, const double *gg, double *out) {
  const orc_gem gem = {ng, gc, gg};
  shell_block(b12, s1, s2, b34, s3, s4, 0.0, &gem, out);
}

static void shell_block(const orc_basis *b12, int s1, int s2, const orc_basis *b34, int s3, int s4,
                        double prim_eps, const orc_gem *gem, double *out) {
  const int l1 = b12->l[s1], l2 = b12->l[s2], l3 = b34->l[s3], l4 = b34->l[s4];
  const int n = ncart(l1) * ncart(l2) * ncart(l3) * ncart(l4);
  double A[3], B[3], C[3], D[3];
  shell_xyz(b12, s1, A); shell_xyz(b12, s2, B); shell_xyz(b34, s3, C); shell_xyz(b34, s4, D);
  double prim[ORC_MAXNC * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC];
  for (int k = 0; k < n; ++k) out[k] = 0.0;
  double ab2 = 0, cd2 = 0;
  for (int k = 0; k < 3; ++k) {
    ab2 += (A[k] - B[k]) * (A[k] - B[k]);
    cd2 += (C[k] - D[k]) * (C[k] - D[k]);
  }
  const double cut = prim_eps > 0.0 ? -log(prim_eps) : INFINITY;
  for (int i = 0; i < b12->nprim[s1]; ++i) {
    const double a = b12->exps[b12->poff[s1] + i];
    const double ca = b12->coefs[b12->poff[s1] + i] * orc_prim_norm(a, l1);
    for (int j = 0; j < b12->nprim[s2]; ++j) {
      const double b = b12->exps[b12->poff[s2] + j];
      if (a * b / (a + b) * ab2 > cut) continue;
      const double cb = b12->coefs[b12->poff[s2] + j] * orc_prim_norm(b, l2);
      for (int k = 0; k < b34->nprim[s3]; ++k) {
        const double c = b34->exps[b34->poff[s3] + k];
        const double cc = b34->coefs[b34->poff[s3] + k] * orc_prim_norm(c, l3);
        for (int l = 0; l < b34->nprim[s4]; ++l) {
          const double d = b34->exps[b34->poff[s4] + l];
          if (c * d / (c + d) * cd2 > cut) continue;
          const double cd = b34->coefs[b34->poff[s4] + l] * orc_prim_norm(d, l4);
          prim_two_body(l1, l2, l3, l4, A, B, C, D, a, b, c, d, gem, prim);
          const double w = ca * cb * cc * cd;
          for (int m = 0; m < n; ++m) out[m] += w * prim[m];
        }
      }
    }
  }
}

/* Raw overlap block <s1|s2> (rows s1 components). */
void orc_overlap_shell(const orc_basis *b1, int s1, const orc_basis *b2, int s2, double *out) {
  const int l1 = b1->l[s1], l2 = b2->l[s2], n1 = ncart(l1), n2 = ncart(l2);
  int c1[ORC_MAXNC][3], c2[ORC_MAXNC][3];
  cart_comps(l1, c1); cart_comps(l2, c2);
  double A[3], B[3], ab2 = 0;
  shell_xyz(b1, s1, A); shell_xyz(b2, s2, B);
  for (int k = 0; k < 3; ++k) ab2 += (A[k] - B[k]) * (A[k] - B[k]);
  for (int m = 0; m < n1 * n2; ++m) out[m] = 0.0;
  etab E[3];
  for (int i = 0; i < b1->nprim[s1]; ++i) {
    const double a = b1->exps[b1->poff[s1] + i];
    const double ca = b1->coefs[b1->poff[s1] + i] * orc_prim_norm(a, l1);
    for (int j = 0; j < b2->nprim[s2]; ++j) {
      const double b = b2->exps[b2->poff[s2] + j];
      const double cb = b2->coefs[b2->poff[s2] + j] * orc_prim_norm(b, l2);
      const double p = a + b;
      const double w = ca * cb * exp(-a * b / p * ab2) * pow(ORC_PI / p, 1.5);
      for (int k = 0; k < 3; ++k) {
        const double Pk = (a * A[k] + b * B[k]) / p;
        hermite_e(l1, l2, p, Pk - A[k], Pk - B[k], E[k]);
      }
      for (int x = 0; x < n1; ++x)
        for (int y = 0; y < n2; ++y)
          out[x * n2 + y] += w * E[0][c1[x][0]][c2[y][0]][0] * E[1][c1[x][1]][c2[y][1]][0] *
                             E[2][c1[x][2]][c2[y][2]][0];
    }
  }
}

int orc_nbasis(const orc_basis *bs) {
  int n = 0;
  for (int s = 0; s < bs->nshell; ++s) n += ncart(bs->l[s]);
  return n;
}

static void shell_offsets(const orc_basis *bs, int *off) {
  int n = 0;
  for (int s = 0; s < bs->nshell; ++s) { off[s] = n; n += ncart(bs->l[s]); }
}

/* norma[bf] = 1/sqrt(S_raw(bf,bf)) -- Libint2Iface.cpp:115-127 */
void orc_norma(const orc_basis *bs, double *norma) {
  int k = 0;
  double buf[ORC_MAXNC * ORC_MAXNC];
  for (int s = 0; s < bs->nshell; ++s) {
    const int nc = ncart(bs->l[s]);
    orc_overlap_shell(bs, s, bs, s, buf);
    for (int i = 0; i < nc; ++i) norma[k++] = 1.0 / sqrt(buf[i * nc + i]);
  }
}

/* ------------------------------------------------ shell-pair list (a10) */
/* Fills pairs_s1/pairs_s2 (s2 <= s1, ascending s2 within each s1, s1
 * ascending) and returns the count.  norms_out (optional, [nshell*nshell])
 * receives the Frobenius norm of each raw overlap block (-1 for same-centre
 * pairs, which are significant without a test). */
int orc_shellpair_list(const orc_basis *bs, double threshold, int *pairs_s1, int *pairs_s2,
                       double *norms_out) {
  int np = 0;
  double buf[ORC_MAXNC * ORC_MAXNC];
  for (int s1 = 0; s1 < bs->nshell; ++s1)
    for (int s2 = 0; s2 <= s1; ++s2) {
      const int same = bs->xyz[3 * s1] == bs->xyz[3 * s2] &&
                       bs->xyz[3 * s1 + 1] == bs->xyz[3 * s2 + 1] &&
                       bs->xyz[3 * s1 + 2] == bs->xyz[3 * s2 + 2];
      int sig = same;
      double nrm = -1.0;
      if (!same) {
        const int n = ncart(bs->l[s1]) * ncart(bs->l[s2]);
        orc_overlap_shell(bs, s1, bs, s2, buf);
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += buf[k] * buf[k];
        nrm = sqrt(s);
        sig = nrm >= threshold;
      }
      if (norms_out) norms_out[s1 * bs->nshell + s2] = nrm;
      if (sig) { pairs_s1[np] = s1; pairs_s2[np] = s2; ++np; }
    }
  return np;
}

/* ------------------------------------------------ Schwarz matrix (a11) */
void orc_schwarz(const orc_basis *bs, double *Q) {
  const int ns = bs->nshell;
#pragma omp parallel for schedule(dynamic)
  for (int s1 = 0; s1 < ns; ++s1) {
    double *buf = (double *)malloc(sizeof(double) * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC);
    for (int s2 = 0; s2 <= s1; ++s2) {
      const int n1 = ncart(bs->l[s1]), n2 = ncart(bs->l[s2]), n12 = n1 * n2;
      orc_eri_shell(bs, s1, s2, bs, s1, s2, 0.0, buf);
      double mx = 0.0;
      for (int k = 0; k < n12; ++k) {
        const double v = fabs(buf[k * n12 + k]);
        if (v > mx) mx = v;
      }
      Q[s1 * ns + s2] = Q[s2 * ns + s1] = sqrt(mx);
    }
    free(buf);
  }
}

/* ------------------------------------------------ direct J/K (a8) */
/* G receives 0.25 (Graw + Graw^T) = J[D] + factor K[D].  Dmat is read as
 * D[i*n + j] (symmetric).  pairs: output of orc_shellpair_list.  Q: Schwarz
 * matrix (nshell^2).  Returns the reference's num_ints_computed.
 * counters (optional, 2 longs): [0] computed shell quartets, [1] primitive
 * quartets before primitive screening.  s1_stride > 1 restricts the outer
 * shell loop to s1 = 0 mod s1_stride (a bounded sample for CPU timing). */
static long direct_build(const orc_basis *bs, const int *pairs_s1, const int *pairs_s2, int npairs,
                         const double *Q, const double *Dmat, double factor, int screen, double precision,
                         double prim_eps, int nthreads, int s1_stride, const orc_gem *gem, double *G,
                         long *counters);

long orc_compute_2body_direct(const orc_basis *bs, const int *pairs_s1, const int *pairs_s2,
                              int npairs, const double *Q, const double *Dmat, double factor,
                              int screen, double precision, double prim_eps, int nthreads,
                              int s1_stride, double *G, long *counters) {
  return direct_build(bs, pairs_s1, pairs_s2, npairs, Q, Dmat, factor, screen, precision, prim_eps, nthreads,
                      s1_stride, NULL, G, counters);
}

/* Reference LibintInterface::compute_g12_direct (Libint2Iface.cpp:2193-2365):
 * the loops, degeneracy, updates and symmetrisation of the direct build with
 * the geminal operator and no density screen. */
long orc_compute_g12_direct(const orc_basis *bs, const int *pairs_s1, const int *pairs_s2, int npairs,
                            const double *Dmat, double factor, int ng, const double *gc, const double *gg,
                            int nthreads, double *G) {
  const orc_gem gem = {ng, gc, gg};
  return direct_build(bs, pairs_s1, pairs_s2, npairs, NULL, Dmat, factor, 0, 0.0, 0.0, nthreads, 1, &gem, G,
                      NULL);
}

static long direct_build(const orc_basis *bs, const int *pairs_s1, const int *pairs_s2, int npairs,
                         const double *Q, const double *Dmat, double factor, int screen, double precision,
                         double prim_eps, int nthreads, int s1_stride, const orc_gem *gem, double *G,
                         long *counters) {
  const int ns = bs->nshell, n = orc_nbasis(bs);
  int *off = (int *)malloc(sizeof(int) * ns);
  shell_offsets(bs, off);
  double *norma = (double *)malloc(sizeof(double) * n);
  orc_norma(bs, norma);
  /* pair list as CSR over s1 */
  int *start = (int *)calloc(ns + 1, sizeof(int));
  for (int k = 0; k < npairs; ++k) start[pairs_s1[k] + 1]++;
  for (int s = 0; s < ns; ++s) start[s + 1] += start[s];
  int *plist = (int *)malloc(sizeof(int) * (npairs > 0 ? npairs : 1));
  {
    int *fill = (int *)calloc(ns, sizeof(int));
    for (int k = 0; k < npairs; ++k) plist[start[pairs_s1[k]] + fill[pairs_s1[k]]++] = pairs_s2[k];
    free(fill);
  }
  /* shell-block infinity norms of D -- Libint2Iface.cpp:143-161 */
  double *Dn = (double *)calloc((size_t)ns * ns, sizeof(double));
  for (int s1 = 0; s1 < ns; ++s1)
    for (int s2 = 0; s2 < ns; ++s2) {
      double m = 0.0;
      for (int i = 0; i < ncart(bs->l[s1]); ++i)
        for (int j = 0; j < ncart(bs->l[s2]); ++j) {
          const double v = fabs(Dmat[(size_t)(off[s1] + i) * n + off[s2] + j]);
          if (v > m) m = v;
        }
      Dn[s1 * ns + s2] = m;
    }
#ifdef _OPENMP
  if (nthreads < 1) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  double *Gt = (double *)calloc((size_t)nthreads * n * n, sizeof(double));
  long nints = 0, nquart = 0, nprimq = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : nints, nquart, nprimq)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    double *g = Gt + (size_t)tid * n * n;
    double *buf = (double *)malloc(sizeof(double) * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC);
    long s1234 = 0;
    for (int s1 = 0; s1 < ns; ++s1) {
      if (s1_stride > 1 && s1 % s1_stride != 0) continue;
      const int n1 = ncart(bs->l[s1]), bf1f = off[s1];
      for (int k2 = start[s1]; k2 < start[s1 + 1]; ++k2) {
        const int s2 = plist[k2], n2 = ncart(bs->l[s2]), bf2f = off[s2];
        const double Dn12 = Dn[s1 * ns + s2];
        for (int s3 = 0; s3 <= s1; ++s3) {
          const int n3 = ncart(bs->l[s3]), bf3f = off[s3];
          double Dn123 = Dn[s1 * ns + s3];
          if (Dn[s2 * ns + s3] > Dn123) Dn123 = Dn[s2 * ns + s3];
          if (Dn12 > Dn123) Dn123 = Dn12;
          const int s4max = (s1 == s3) ? s2 : s3;
          for (int k4 = start[s3]; k4 < start[s3 + 1]; ++k4) {
            const int s4 = plist[k4];
            if (s4 > s4max) break;
            if ((s1234++) % nthreads != tid) continue;
            if (screen) {
              double Dn1234 = Dn[s1 * ns + s4];
              if (Dn[s2 * ns + s4] > Dn1234) Dn1234 = Dn[s2 * ns + s4];
              if (Dn[s3 * ns + s4] > Dn1234) Dn1234 = Dn[s3 * ns + s4];
              if (Dn123 > Dn1234) Dn1234 = Dn123;
              if (Dn1234 * Q[s1 * ns + s2] * Q[s3 * ns + s4] < precision) continue;
            }
            const int n4 = ncart(bs->l[s4]), bf4f = off[s4];
            nints += (long)n1 * n2 * n3 * n4;
            nquart += 1;
            nprimq += (long)bs->nprim[s1] * bs->nprim[s2] * bs->nprim[s3] * bs->nprim[s4];
            const double deg = (s1 == s2 ? 1.0 : 2.0) * (s3 == s4 ? 1.0 : 2.0) *
                               ((s1 == s3) ? (s2 == s4 ? 1.0 : 2.0) : 2.0);
            shell_block(bs, s1, s2, bs, s3, s4, prim_eps, gem, buf);
            for (int f1 = 0, f1234 = 0; f1 < n1; ++f1) {
              const int b1 = bf1f + f1;
              for (int f2 = 0; f2 < n2; ++f2) {
                const int b2 = bf2f + f2;
                for (int f3 = 0; f3 < n3; ++f3) {
                  const int b3 = bf3f + f3;
                  for (int f4 = 0; f4 < n4; ++f4, ++f1234) {
                    const int b4 = bf4f + f4;
                    const double v = buf[f1234] * deg * norma[b1] * norma[b2] * norma[b3] * norma[b4];
                    g[(size_t)b1 * n + b2] += Dmat[(size_t)b3 * n + b4] * v;
                    g[(size_t)b3 * n + b4] += Dmat[(size_t)b1 * n + b2] * v;
                    g[(size_t)b1 * n + b3] += 0.5 * factor * Dmat[(size_t)b2 * n + b4] * v;
                    g[(size_t)b2 * n + b4] += 0.5 * factor * Dmat[(size_t)b1 * n + b3] * v;
                    g[(size_t)b1 * n + b4] += 0.5 * factor * Dmat[(size_t)b2 * n + b3] * v;
                    g[(size_t)b2 * n + b3] += 0.5 * factor * Dmat[(size_t)b1 * n + b4] * v;
                  }
                }
              }
            }
          }
        }
      }
    }
    free(buf);
  }
  for (int t = 1; t < nthreads; ++t)
    for (size_t k = 0; k < (size_t)n * n; ++k) Gt[k] += Gt[(size_t)t * n * n + k];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      G[(size_t)i * n + j] = 0.25 * (Gt[(size_t)i * n + j] + Gt[(size_t)j * n + i]);
  if (counters) { counters[0] = nquart; counters[1] = nprimq; }
  free(Gt); free(Dn); free(plist); free(start); free(norma); free(off);
  return nints;
}

/* ------------------------------------------------ coupling (a9) */
/* Reference LibintInterface::compute_coupling_direct with `this` = outer
 * (the lower species id) and `other` = inner.  D is the density of the
 * species whose index is NOT the result's: non-permuted -> D of inner
 * (n_inner^2) and the result is n_outer^2; permuted -> D of outer and the
 * result is n_inner^2.  Output image: element (row r, col c) at r*n + c, the
 * lower triangle (row >= col) filled, exactly as the reference returns it.
 * counts (optional): number of raw ERIs kept after the 1e-10 drop, and the
 * number of ERIs inside 1e-3 relative of the cutoff. */
static void coupling_build(const orc_basis *outer, const orc_basis *inner, const double *Dmat, int permuted,
                           double drop, int nthreads, const orc_gem *gem, double *Bout, long *counts);

void orc_compute_coupling_direct(const orc_basis *outer, const orc_basis *inner,
                                 const double *Dmat, int permuted, double drop,
                                 int nthreads, double *Bout, long *counts) {
  coupling_build(outer, inner, Dmat, permuted, drop, nthreads, NULL, Bout, counts);
}

/* Reference LibintInterface::compute_g12inter_direct
 * (Libint2Iface.cpp:2367-2550): the coupling build with the geminal
 * operator. */
void orc_compute_g12inter_direct(const orc_basis *outer, const orc_basis *inner, const double *Dmat,
                                 int permuted, double drop, int ng, const double *gc, const double *gg,
                                 int nthreads, double *Bout) {
  const orc_gem gem = {ng, gc, gg};
  coupling_build(outer, inner, Dmat, permuted, drop, nthreads, &gem, Bout, NULL);
}

static void coupling_build(const orc_basis *outer, const orc_basis *inner, const double *Dmat, int permuted,
                           double drop, int nthreads, const orc_gem *gem, double *Bout, long *counts) {
  const int ns = outer->nshell, ons = inner->nshell;
  const int nA = orc_nbasis(outer), nB = orc_nbasis(inner);
  const int n = permuted ? nB : nA;
  int *off = (int *)malloc(sizeof(int) * ns), *ooff = (int *)malloc(sizeof(int) * ons);
  shell_offsets(outer, off); shell_offsets(inner, ooff);
  double *normA = (double *)malloc(sizeof(double) * nA), *normB = (double *)malloc(sizeof(double) * nB);
  orc_norma(outer, normA); orc_norma(inner, normB);
#ifdef _OPENMP
  if (nthreads < 1) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  double *Bt = (double *)calloc((size_t)nthreads * n * n, sizeof(double));
  long kept = 0, near = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : kept, near)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    double *b = Bt + (size_t)tid * n * n;
    double *buf = (double *)malloc(sizeof(double) * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC);
    long s1234 = 0;
    for (int s1 = 0; s1 < ns; ++s1) {
      const int n1 = ncart(outer->l[s1]);
      for (int s2 = s1; s2 < ns; ++s2) {
        const int n2 = ncart(outer->l[s2]);
        for (int os1 = 0; os1 < ons; ++os1) {
          const int on1 = ncart(inner->l[os1]);
          for (int os2 = os1; os2 < ons; ++os2) {
            if ((s1234++) % nthreads != tid) continue;
            const int on2 = ncart(inner->l[os2]);
            shell_block(outer, s1, s2, inner, os1, os2, 0.0, gem, buf);
            for (int f1 = 0, f = 0; f1 < n1; ++f1) {
              const int bf1 = off[s1] + f1;
              for (int f2 = 0; f2 < n2; ++f2) {
                const int bf2 = off[s2] + f2;
                for (int of1 = 0; of1 < on1; ++of1) {
                  const int obf1 = ooff[os1] + of1;
                  for (int of2 = 0; of2 < on2; ++of2, ++f) {
                    const int obf2 = ooff[os2] + of2;
                    if (!(bf1 <= bf2 && obf1 <= obf2)) continue;
                    const double raw = buf[f];
                    if (fabs(fabs(raw) - drop) < 1e-3 * drop) ++near;
                    if (fabs(raw) < drop) continue;
                    ++kept;
                    const double nrm = normA[bf1] * normA[bf2] * normB[obf1] * normB[obf2];
                    if (permuted) {
                      const double cval = Dmat[(size_t)bf1 * nA + bf2] * raw * nrm;
                      b[(size_t)obf2 * n + obf1] += cval;
                      if (bf1 != bf2) b[(size_t)obf2 * n + obf1] += cval;
                    } else {
                      const double cval = Dmat[(size_t)obf1 * nB + obf2] * raw * nrm;
                      b[(size_t)bf2 * n + bf1] += cval;
                      if (obf1 != obf2) b[(size_t)bf2 * n + bf1] += cval;
                    }
                  }
                }
              }
            }
          }
        }
      }
    }
    free(buf);
  }
  for (size_t k = 0; k < (size_t)n * n; ++k) {
    double s = 0.0;
    for (int t = 0; t < nthreads; ++t) s += Bt[(size_t)t * n * n + k];
    Bout[k] = s;
  }
  if (counts) { counts[0] = kept; counts[1] = near; }
  free(Bt); free(normA); free(normB); free(off); free(ooff);
}

/* ------------------------------------------------ integral store */
/* Reference LibintInterface::compute_2body_directAll
 * (Libint2Iface.cpp:2552-2742): the quartet loops and screen of the direct
 * build (2629-2669); each integral with |raw| > drop, times the four norma,
 * written to A[i n^3 + j n^2 + k n + l] and its seven other index images
 * (2701-2717).  Elements never written keep the caller's values.  The
 * reference visits each unique integral once (IntraIntsIt); writing the
 * images of every element of the shell block covers the same set.
 * Returns the number of shell quartets that passed the screen. */
long orc_compute_2body_all(const orc_basis *bs, const int *pairs_s1, const int *pairs_s2, int npairs,
                           const double *Q, const double *Dmat, int screen, double precision, double drop,
                           double *A) {
  const int ns = bs->nshell, n = orc_nbasis(bs);
  const size_t N = (size_t)n, N2 = N * N, N3 = N2 * N;
  int *off = (int *)malloc(sizeof(int) * ns);
  shell_offsets(bs, off);
  double *norma = (double *)malloc(sizeof(double) * n);
  orc_norma(bs, norma);
  int *start = (int *)calloc(ns + 1, sizeof(int));
  for (int k = 0; k < npairs; ++k) start[pairs_s1[k] + 1]++;
  for (int s = 0; s < ns; ++s) start[s + 1] += start[s];
  int *plist = (int *)malloc(sizeof(int) * (npairs > 0 ? npairs : 1));
  {
    int *fill = (int *)calloc(ns, sizeof(int));
    for (int k = 0; k < npairs; ++k) plist[start[pairs_s1[k]] + fill[pairs_s1[k]]++] = pairs_s2[k];
    free(fill);
  }
  double *Dn = (double *)calloc((size_t)ns * ns, sizeof(double));
  for (int s1 = 0; s1 < ns; ++s1)
    for (int s2 = 0; s2 < ns; ++s2) {
      double m = 0.0;
      for (int i = 0; i < ncart(bs->l[s1]); ++i)
        for (int j = 0; j < ncart(bs->l[s2]); ++j) {
          const double v = fabs(Dmat[(size_t)(off[s1] + i) * n + off[s2] + j]);
          if (v > m) m = v;
        }
      Dn[s1 * ns + s2] = m;
    }
  double *buf = (double *)malloc(sizeof(double) * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC);
  long nquart = 0;
  for (int s1 = 0; s1 < ns; ++s1) {
    const int n1 = ncart(bs->l[s1]);
    for (int k2 = start[s1]; k2 < start[s1 + 1]; ++k2) {
      const int s2 = plist[k2], n2 = ncart(bs->l[s2]);
      for (int s3 = 0; s3 <= s1; ++s3) {
        const int n3 = ncart(bs->l[s3]);
        const int s4max = (s1 == s3) ? s2 : s3;
        for (int k4 = start[s3]; k4 < start[s3 + 1]; ++k4) {
          const int s4 = plist[k4];
          if (s4 > s4max) break;
          if (screen) {
            double d = Dn[s1 * ns + s2];
            const double o[5] = {Dn[s1 * ns + s3], Dn[s2 * ns + s3], Dn[s1 * ns + s4], Dn[s2 * ns + s4],
                                 Dn[s3 * ns + s4]};
            for (int q = 0; q < 5; ++q) if (o[q] > d) d = o[q];
            if (d * Q[s1 * ns + s2] * Q[s3 * ns + s4] < precision) continue;
          }
          const int n4 = ncart(bs->l[s4]);
          ++nquart;
          orc_eri_shell(bs, s1, s2, bs, s3, s4, 0.0, buf);
          for (int f1 = 0, f = 0; f1 < n1; ++f1)
            for (int f2 = 0; f2 < n2; ++f2)
              for (int f3 = 0; f3 < n3; ++f3)
                for (int f4 = 0; f4 < n4; ++f4, ++f) {
                  if (!(fabs(buf[f]) > drop)) continue;
                  const size_t i = off[s1] + f1, j = off[s2] + f2, k = off[s3] + f3, l = off[s4] + f4;
                  const double v = buf[f] * norma[i] * norma[j] * norma[k] * norma[l];
                  A[i * N3 + j * N2 + k * N + l] = v;
                  A[j * N3 + i * N2 + k * N + l] = v;
                  A[i * N3 + j * N2 + l * N + k] = v;
                  A[j * N3 + i * N2 + l * N + k] = v;
                  A[k * N3 + l * N2 + i * N + j] = v;
                  A[k * N3 + l * N2 + j * N + i] = v;
                  A[l * N3 + k * N2 + i * N + j] = v;
                  A[l * N3 + k * N2 + j * N + i] = v;
                }
        }
      }
    }
  }
  free(buf); free(Dn); free(plist); free(start); free(norma); free(off);
  return nquart;
}

/* Reference LibintInterface::compute_coupling_directAll
 * (Libint2Iface.cpp:2744-2908) with `this` = outer: every s1 <= s2 of outer
 * and os1 <= os2 of inner, no screen, elements bf1 <= bf2 and obf1 <= obf2
 * with |raw| >= drop (2864-2866), normalised, written as
 * T[bf1][bf2][obf1][obf2] (dims n_outer^2 x n_inner^2) and the images
 * bf1 <-> bf2, obf1 <-> obf2 (2877-2880).  Unwritten elements keep the
 * caller's values. */
void orc_compute_coupling_all(const orc_basis *outer, const orc_basis *inner, double drop, double *T) {
  const int ns = outer->nshell, ons = inner->nshell;
  const int nA = orc_nbasis(outer), nB = orc_nbasis(inner);
  const size_t s1x = (size_t)nA * nB * nB, s2x = (size_t)nB * nB, s3x = (size_t)nB;
  int *off = (int *)malloc(sizeof(int) * ns), *ooff = (int *)malloc(sizeof(int) * ons);
  shell_offsets(outer, off); shell_offsets(inner, ooff);
  double *normA = (double *)malloc(sizeof(double) * nA), *normB = (double *)malloc(sizeof(double) * nB);
  orc_norma(outer, normA); orc_norma(inner, normB);
  double *buf = (double *)malloc(sizeof(double) * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC * ORC_MAXNC);
  for (int s1 = 0; s1 < ns; ++s1)
    for (int s2 = s1; s2 < ns; ++s2)
      for (int os1 = 0; os1 < ons; ++os1)
        for (int os2 = os1; os2 < ons; ++os2) {
          const int n1 = ncart(outer->l[s1]), n2 = ncart(outer->l[s2]);
          const int on1 = ncart(inner->l[os1]), on2 = ncart(inner->l[os2]);
          orc_eri_shell(outer, s1, s2, inner, os1, os2, 0.0, buf);
          for (int f1 = 0, f = 0; f1 < n1; ++f1)
            for (int f2 = 0; f2 < n2; ++f2)
              for (int of1 = 0; of1 < on1; ++of1)
                for (int of2 = 0; of2 < on2; ++of2, ++f) {
                  const size_t b1 = off[s1] + f1, b2 = off[s2] + f2, o1 = ooff[os1] + of1, o2 = ooff[os2] + of2;
                  if (!(b1 <= b2 && o1 <= o2)) continue;
                  if (fabs(buf[f]) < drop) continue;
                  const double v = buf[f] * normA[b1] * normA[b2] * normB[o1] * normB[o2];
                  T[b1 * s1x + b2 * s2x + o1 * s3x + o2] = v;
                  T[b2 * s1x + b1 * s2x + o1 * s3x + o2] = v;
                  T[b1 * s1x + b2 * s2x + o2 * s3x + o1] = v;
                  T[b2 * s1x + b1 * s2x + o2 * s3x + o1] = v;
                }
        }
  free(buf); free(normA); free(normB); free(off); free(ooff);
}

/* ------------------------------------------------ 1-body (S, T, V) */
/* kind: 1 overlap, 2 kinetic (-1/2 nabla^2), 3 point-charge attraction
 * sum_C -Z_C/|r-C|.  Result normalised with norma (Libint2Iface.cpp:163-217),
 * n x n at r*n + c. */
void orc_1body(const orc_basis *bs, int kind, int ncharge, const double *Z, const double *Cxyz,
               double *M) {
  const int ns = bs->nshell, n = orc_nbasis(bs);
  int *off = (int *)malloc(sizeof(int) * ns);
  shell_offsets(bs, off);
  double *norma = (double *)malloc(sizeof(double) * n);
  orc_norma(bs, norma);
#pragma omp parallel for schedule(dynamic)
  for (int s1 = 0; s1 < ns; ++s1) {
    for (int s2 = 0; s2 < ns; ++s2) {
      const int l1 = bs->l[s1], l2 = bs->l[s2], n1 = ncart(l1), n2 = ncart(l2);
      int c1[ORC_MAXNC][3], c2[ORC_MAXNC][3];
      cart_comps(l1, c1); cart_comps(l2, c2);
      double A[3], B[3], ab2 = 0;
      shell_xyz(bs, s1, A); shell_xyz(bs, s2, B);
      for (int k = 0; k < 3; ++k) ab2 += (A[k] - B[k]) * (A[k] - B[k]);
      double blk[ORC_MAXNC * ORC_MAXNC];
      for (int m = 0; m < n1 * n2; ++m) blk[m] = 0.0;
      etab E[3];
      for (int i = 0; i < bs->nprim[s1]; ++i) {
        const double a = bs->exps[bs->poff[s1] + i];
        const double ca = bs->coefs[bs->poff[s1] + i] * orc_prim_norm(a, l1);
        for (int j = 0; j < bs->nprim[s2]; ++j) {
          const double b = bs->exps[bs->poff[s2] + j];
          const double cb = bs->coefs[bs->poff[s2] + j] * orc_prim_norm(b, l2);
          const double p = a + b, kab = exp(-a * b / p * ab2);
          double P[3];
          for (int k = 0; k < 3; ++k) P[k] = (a * A[k] + b * B[k]) / p;
          if (kind == 1 || kind == 2) {
            /* need E up to lb + 2 for kinetic: build per axis with lb+2 */
            etab Ek[3];
            for (int k = 0; k < 3; ++k)
              hermite_e(l1, l2 + 2, p, P[k] - A[k], P[k] - B[k], Ek[k]);
            const double s0 = pow(ORC_PI / p, 1.5) * kab * ca * cb;
            for (int x = 0; x < n1; ++x)
              for (int y = 0; y < n2; ++y) {
                const int *ia = c1[x], *ib = c2[y];
                double S1[3];
                for (int k = 0; k < 3; ++k) S1[k] = Ek[k][ia[k]][ib[k]][0];
                if (kind == 1) {
                  blk[x * n2 + y] += s0 * S1[0] * S1[1] * S1[2];
                } else {
                  /* T = -1/2 <a| d^2/dx^2 |b> summed over axes.
                   * d^2/dx^2 x^j e^{-b x^2} = j(j-1) x^{j-2} - 2b(2j+1) x^j + 4b^2 x^{j+2} */
                  double t = 0.0;
                  for (int k = 0; k < 3; ++k) {
                    const int jb = ib[k];
                    double d2 = -2.0 * b * (2 * jb + 1) * Ek[k][ia[k]][jb][0];
                    d2 += 4.0 * b * b * Ek[k][ia[k]][jb + 2][0];
                    if (jb >= 2) d2 += jb * (jb - 1) * Ek[k][ia[k]][jb - 2][0];
                    double other = 1.0;
                    for (int kk = 0; kk < 3; ++kk)
                      if (kk != k) other *= S1[kk];
                    t += -0.5 * d2 * other;
                  }
                  blk[x * n2 + y] += s0 * t;
                }
              }
          } else {
            for (int k = 0; k < 3; ++k) hermite_e(l1, l2, p, P[k] - A[k], P[k] - B[k], E[k]);
            const int L = l1 + l2;
            double R[RD * RD * RD];
            for (int cI = 0; cI < ncharge; ++cI) {
              const double PC[3] = {P[0] - Cxyz[3 * cI], P[1] - Cxyz[3 * cI + 1], P[2] - Cxyz[3 * cI + 2]};
              hermite_r(L, p, PC, R);
              const double pre = -Z[cI] * 2.0 * ORC_PI / p * kab * ca * cb;
              for (int x = 0; x < n1; ++x)
                for (int y = 0; y < n2; ++y) {
                  const int *ia = c1[x], *ib = c2[y];
                  double s = 0.0;
                  for (int t = 0; t <= ia[0] + ib[0]; ++t)
                    for (int u = 0; u <= ia[1] + ib[1]; ++u)
                      for (int v = 0; v <= ia[2] + ib[2]; ++v)
                        s += E[0][ia[0]][ib[0]][t] * E[1][ia[1]][ib[1]][u] * E[2][ia[2]][ib[2]][v] *
                             R[IX(t, u, v)];
                  blk[x * n2 + y] += pre * s;
                }
            }
          }
        }
      }
      for (int x = 0; x < n1; ++x)
        for (int y = 0; y < n2; ++y)
          M[(size_t)(off[s1] + x) * n + off[s2] + y] =
              blk[x * n2 + y] * norma[off[s1] + x] * norma[off[s2] + y];
    }
  }
  free(norma); free(off);
}
