/*
 * lowdin_fock.h -- C ABI of the MI355X direct two-particle Fock build.
 *
 * liblowdin_fock.so is a drop-in for the native half of openLOWDIN's direct
 * SCF path.  The first block of entry points keeps the exact symbol names,
 * argument order and meaning of the `extern "C"` functions that openLOWDIN's
 * Fortran binds through ISO_C_BINDING (src/ints/Libint2Interface.f90:165-373),
 * declared in src/ints/Libint2Iface.h:350-453 and defined in
 * src/ints/Libint2Iface.cpp:3454-3589.  Behind them the integrals, the J/K
 * contraction and the coupling contraction run as HIP kernels on the GPU.
 *
 * Conventions (as in the reference):
 *   - one handle per quantum species; shells are added in particle order,
 *     then basis-file contraction order, with RAW (file) coefficients;
 *   - Cartesian functions only (xx, xy, xz, yy, yz, zz), each normalised to 1;
 *   - all n x n matrices are dense FP64, Fortran column-major; D is assumed
 *     symmetric; the callee overwrites every element of `result`;
 *   - the entry points return nothing.  A failure leaves `result` filled
 *     with NaN and is reported by lowdin_fock_last_error().
 */
#ifndef LOWDIN_FOCK_H
#define LOWDIN_FOCK_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------ reference entry points */

/* Libint2Iface.cpp:3454 (LibintInterface_new).  Creates the handle of one
 * species.  stack_size is accepted and unused (it sized the disk path's
 * integral stack); `parallel` is accepted and unused (the GPU replaces the
 * OpenMP thread pool). */
void* LibintInterface_new(int stack_size, int species_id, bool is_electron, bool parallel);

/* Libint2Iface.cpp:3460 (LibintInterface_del). */
void LibintInterface_del(void* lint);

/* Libint2Iface.cpp:3467 (LibintInterface_add_particle).  z = -(total charge)
 * of the particle; recorded, not used by the two-particle path. */
void LibintInterface_add_particle(void* lint, int z, const double* center);

/* Libint2Iface.cpp:3473 (LibintInterface_add_shell).  Exponents, raw
 * contraction coefficients, origin (bohr), angular momentum l <= 2 and the
 * number of primitives.  Computes norma = 1/sqrt(S_raw,ii) per Cartesian
 * function (Libint2Iface.cpp:79-128). */
void LibintInterface_add_shell(void* lint, double* alpha, double* coeff, double* origin, int l, int nprim);

/* Declared at Libint2Iface.h:368 but never defined by the reference; defined
 * here: the number of Cartesian basis functions of the species. */
int LibintInterface_get_nbasis(void* lint);

/* Libint2Iface.cpp:3505 (LibintInterface_init_2body_ints).  Builds the
 * significant shell-pair list (same centre, or raw overlap Frobenius norm
 * >= 1e-12; Libint2Iface.cpp:3248-3342) on the GPU. */
void LibintInterface_init_2body_ints(void* lint);

/* Libint2Iface.cpp:3510-3526 (LibintInterface_compute_2body_direct).
 * result = J[D] + factor K[D] (n x n), with the reference's screening:
 * unique shell quartets from the pair list, Dnorm * Q_ab * Q_cd < DBL_EPSILON
 * skipped for electrons only, G = 0.25 (G + G^T).  `factor` is passed by
 * address, as the Fortran binding does (real(c_double) without VALUE). */
void LibintInterface_compute_2body_direct(void* lint, double* dens, double* result, double* factor);

/* Libint2Iface.cpp:3566-3589 (LibintInterface_compute_coupling_direct).
 * result (n_lint x n_lint) = Coulomb matrix on lint's species from the
 * density `dens` (n_olint x n_olint) of olint's species, every shell pair of
 * both species (no screening), raw integrals below 1e-10 dropped, filled in
 * the upper triangle of the column-major result (the caller symmetrises). */
void LibintInterface_compute_coupling_direct(void* lint, void* olint, double* dens, double* result);

/* Libint2Iface.cpp:3482-3503 (LibintInterface_compute_1body_ints).
 * integral_kind 1: overlap, 2: kinetic (-1/2 nabla^2), 3: attraction to the
 * particles added with add_particle (sum -z/|r - R|); normalised with norma,
 * n x n.  Computed on the GPU. */
void LibintInterface_compute_1body_ints(void* lint, int integral_kind, double* result);

/* Libint2Iface.cpp:3736-3750 (LibintInterface_compute_2body_directAll,
 * compute_2body_directAll at :2552-2742).  results (n^4, element (i j|k l)
 * at i n^3 + j n^2 + k n + l): every integral of the Fock build's quartet set
 * (same pair list and Dnorm * Q * Q < DBL_EPSILON screen, dens giving the
 * density block norms) with |raw| > 1e-10, normalised with norma, written to
 * all eight index images.  Elements not written keep the caller's values. */
void LibintInterface_compute_2body_directAll(void* lint, double* dens, double* results);

/* Libint2Iface.cpp:3528-3549 (LibintInterface_compute_2body_directIT,
 * compute_2body_directIT at :620-921).  First-quarter transform on orbital
 * p: result[j n^2 + k n + l] = sum_i C(p, i) (i j|k l), C(p, i) = coeff[p +
 * i n] (coeff n x n column-major, p passed by address as the Fortran binding
 * does), over the same quartets, drop and screen as directAll.  Every
 * element of result is written. */
void LibintInterface_compute_2body_directIT(void* lint, double* dens, double* coeff, double* result, int* p);

/* Libint2Iface.cpp:3752-3773 (LibintInterface_compute_coupling_directAll,
 * compute_coupling_directAll at :2744-2908).  Every shell pair of both
 * species, no screen, |raw| >= 1e-10: results[x1 n n_o^2 + x2 n_o^2 + y1 n_o
 * + y2] = (x1 x2|y1 y2) normalised, x of the lower species id, y of the
 * other, n = lint's basis size, n_o = olint's, four images (x1 <-> x2,
 * y1 <-> y2).  Called from the higher species id the reference's indices
 * only fit when n == n_o; otherwise the call fails.  dens is not read. */
void LibintInterface_compute_coupling_directAll(void* lint, void* olint, double* dens, double* results);

/* Libint2Iface.cpp:3591-3623 (LibintInterface_compute_coupling_directIT,
 * compute_coupling_directIT at :1296-1622).  results[v n_o^2 + l n_o + s]
 * (n x n_o x n_o) = sum_u C(p, u) (u v|l s), u, v of lint's species, l, s of
 * olint's, C(p, u) = coeff[p + u n]; same quartets and drop as
 * coupling_directAll.  Every element is written; dens is not read. */
void LibintInterface_compute_coupling_directIT(void* lint, void* olint, double* dens, double* coeff, double* results,
                                               int* p);

/* Libint2Iface.cpp:3687-3706 (libintinterface_compute_g12_direct,
 * compute_g12_direct at :2193-2365).  result (n x n) = J[D] + factor K[D]
 * with the contracted Gaussian geminal sum_k coefficients[k]
 * exp(-exponents[k] r12^2) (k < pot_size <= 64, exponents >= 0) in place of
 * 1/r12: every quartet of the pair list (no density screen), degeneracy,
 * G = 0.25 (G + G^T).  factor by address, pot_size by value, as bound by
 * Libint2Interface.f90:305-316. */
void libintinterface_compute_g12_direct(void* lint, double* dens, double* result, double* factor,
                                        const double* coefficients, const double* exponents, const int pot_size);

/* Libint2Iface.cpp:3708-3734 (libintinterface_compute_g12inter_direct,
 * compute_g12inter_direct at :2367-2550).  result (n_lint x n_lint) = the
 * geminal coupling on lint's species from the density `dens` (n_olint x
 * n_olint) of olint's species, with the quartets, |raw| < 1e-10 drop and
 * layout of LibintInterface_compute_coupling_direct. */
void libintinterface_compute_g12inter_direct(void* lint, void* olint, double* dens, double* result,
                                             const double* coefficients, const double* exponents,
                                             const int pot_size);

/* ------------------------------------------------ link-level stubs
 * The rest of the reference's exported C ABI (Libint2Iface.h:381-445,
 * bound by Libint2Interface.f90:165-373): disk storage, the alpha-beta
 * build and the G12 store variants are outside this library's scope.  They
 * are exported so that openLOWDIN's Fortran links unchanged; a call records
 * an error (lowdin_fock_last_error), prints it and fills its result array,
 * where it has one, with NaN (alphabeta: 1 value; g12_directAll: n^4;
 * g12inter_directAll: n_lint^2 n_olint^2).  The disk variants write no
 * file. */
void LibintInterface_compute_2body_disk(void* lint, const char* filename, double* dens);
void LibintInterface_compute_alphabeta_direct(void* lint, void* olint, double* dens, double* otherdens,
                                              double* result);
void LibintInterface_compute_coupling_disk(void* lint, void* olint, const char* filename);
void libintinterface_compute_g12_disk(void* lint, const char* filename, const double* coefficients,
                                      const double* exponents, const int pot_size);
void libintinterface_compute_g12inter_disk(void* lint, void* olint, const char* filename,
                                           const double* coefficients, const double* exponents, const int pot_size);
void libintinterface_compute_g12_directAll(void* lint, double* dens, double* result, const double* coefficients,
                                           const double* exponents, const int pot_size);
void libintinterface_compute_g12inter_directAll(void* lint, void* olint, double* dens, double* result,
                                                const double* coefficients, const double* exponents,
                                                const int pot_size);

/* ------------------------------------------------ extensions (no reference
 * counterpart; used by this package's host code, tests and benchmarks) */

/* Message of the last failure in this thread ("" if none). */
const char* lowdin_fock_last_error(void);

/* One-body matrices on the GPU for arbitrary point charges: kind 1 overlap,
 * 2 kinetic (-1/2 nabla^2, unscaled by mass), 3 attraction sum_C -Z_C/|r - C|
 * over the ncharge charges Z at xyz (3 per charge, bohr).  n x n, normalised
 * with norma.  Returns 0 on success. */
int lowdin_fock_1body(void* lint, int kind, int ncharge, const double* Z, const double* xyz, double* result);

typedef struct {
  uint64_t quartets;        /* shell quartets evaluated by the last build */
  uint64_t eris;            /* contracted ERIs (the reference's num_ints_computed) */
  uint64_t prim_quartets;   /* primitive quartets evaluated */
  double kernel_ms;         /* device time of the integral kernels (HIP events) */
  double total_ms;          /* wall time of the call */
  uint64_t quartets_fp32;   /* of `quartets`, evaluated in FP32 (mixed precision) */
  uint64_t eris_fp32;       /* of `eris`, evaluated in FP32 */
} lowdin_fock_stats;

/* Counters of the last compute_2body_direct / compute_coupling_direct call
 * on this handle.  Returns 0 on success. */
int lowdin_fock_get_stats(void* lint, lowdin_fock_stats* out);

/* The shell-pair list built by init_2body_ints as (s1, s2 <= s1) pairs,
 * s1 ascending then s2 ascending.  Writes at most cap pairs; returns the
 * total count (or -1). */
int lowdin_fock_shellpair_list(void* lint, int* s1, int* s2, int cap);

/* The raw Schwarz matrix Q (nshell x nshell, row-major; zero for pairs not
 * in the pair list).  Returns 0 on success. */
int lowdin_fock_schwarz(void* lint, double* Q);

/* Both directions of one species pair from a single pass over the quartets:
 * ra (n_a^2) = coupling on a from db, rb (n_b^2) = coupling on b from da,
 * each in the layout of LibintInterface_compute_coupling_direct.  Either
 * result may be NULL.  Returns 0 on success. */
int lowdin_fock_coupling_pair(void* a, void* b, const double* da, const double* db, double* ra, double* rb);

/* Device-resident variant of lowdin_fock_coupling_pair: device pointers,
 * enqueued on `stream` (NULL = the library stream); results zeroed by the
 * call.  Returns 0 on success. */
int lowdin_fock_coupling_pair_device(void* a, void* b, const double* d_da, const double* d_db, double* d_ra,
                                     double* d_rb, void* stream, int synchronize);

/* Device-resident forms of the integral store / first-quarter transform:
 * which 0 = 2body_directAll (d_dens n x n, d_out n^4, written where kept),
 * 1 = 2body_directIT (d_dens, d_crow = C(p, :) as n contiguous values,
 * d_out n^3), 2 = coupling_directAll (olint, d_out n^2 n_o^2, written where
 * kept), 3 = coupling_directIT (olint, d_crow, d_out n n_o^2).  Layouts and
 * element sets as the host entry points above; runs on the library stream
 * and synchronises.  Counters in lowdin_fock_get_stats count the candidate
 * quartets before the density screen.  Returns 0 on success. */
int lowdin_fock_store_device(void* lint, void* olint, int which, const double* d_dens, const double* d_crow,
                             double* d_out);

/* Device-resident G12 builds: olint NULL = libintinterface_compute_g12_direct
 * on lint (d_dens, d_result n x n; factor), otherwise
 * libintinterface_compute_g12inter_direct (d_dens n_olint^2, d_result
 * n_lint^2).  Geminal parameters are host arrays.  Runs on the library
 * stream and synchronises.  Returns 0 on success. */
int lowdin_fock_g12_device(void* lint, void* olint, const double* d_dens, double* d_result, double factor,
                           const double* coefficients, const double* exponents, int pot_size);

/* Device-resident variant of LibintInterface_compute_2body_direct: d_dens
 * and d_result are device pointers (n x n FP64); the build is enqueued on
 * `stream` (a hipStream_t; NULL = the library's stream).  With synchronize
 * != 0 the call waits and updates the statistics.  Returns 0 on success. */
int lowdin_fock_compute_2body_device(void* lint, const double* d_dens, double* d_result, double factor,
                                     void* stream, int synchronize);

/* lowdin_fock_compute_2body_device without any allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The build's scratch -- descriptors, the shell-block norms of
 * D and the raw n x n accumulator -- lives in a file-scope device array of
 * the library (5.8 MB, one copy per GPU, held until the process ends), so
 * the basis is limited to 768 functions and 384 shells; larger ones are
 * refused with an error.  The species must have been built once by
 * LibintInterface_compute_2body_direct or lowdin_fock_compute_2body_device
 * (that call prepares the geometry data) and not changed since.  The build
 * is enqueued on `stream` (NULL: the library's stream) and not waited for;
 * no statistics are kept.  Calls that overlap on one GPU share the scratch
 * and race.  Returns 0 on success. */
int lowdin_fock_compute_2body_device_noalloc(void* lint, const double* d_dens, double* d_result, double factor,
                                             void* stream);

/* Record HIP events around every class launch of the intra build. */
int lowdin_fock_time_launches(void* lint, int enable);

/* Precision of the intra-species build.  mode 0: all quartets in FP64 (the
 * default, reference parity).  mode 1: mixed -- a quartet whose bound
 * Dnorm_abcd * Qn_ab * Qn_cd is below tau runs its integrals in FP32 with
 * FP64 contraction and accumulation; the rest in FP64.  Dnorm_abcd is the
 * largest |D| over the six density blocks the quartet touches, Qn the
 * Schwarz bound of the normalised functions (max over the pair's
 * components).  Classes above 300 Cartesian integrals per quartet (dd|dp),
 * (dd|dd) always run in FP64.  Returns 0 on success. */
int lowdin_fock_set_precision(void* lint, int mode, double tau);

/* Quartet screen threshold of the species' intra builds (Dnorm * Q_ab * Q_cd
 * < eps skips the quartet; electrons only).  Default DBL_EPSILON, the
 * reference's (Libint2Iface.cpp:523-534); 1e-12 is the looser threshold the
 * benchmark also reports.  Returns 0 on success. */
int lowdin_fock_set_screen(void* lint, double eps);

/* With enable != 0 every intra build first recomputes the pair records,
 * the Schwarz bounds and the task lists, as the reference recomputes its
 * pair list and Schwarz matrix on every call (Libint2Interface.f90:544,
 * Libint2Iface.cpp:3520); the default reuses them while the shells are
 * unchanged. */
int lowdin_fock_set_setup_each_build(void* lint, int enable);

/* Run the class launches of the intra build one after another on one
 * stream instead of overlapping them on several (for per-class timing). */
int lowdin_fock_serialize_launches(void* lint, int enable);

/* Per class launch of the last intra build: class code 1000*la+100*lb+10*lc+ld,
 * device ms (when timing is on), shell quartets and primitive quartets.
 * Returns the number of launches. */
int lowdin_fock_launch_profile(void* lint, int cap, int* classes, float* ms, uint64_t* quartets,
                               uint64_t* prim_quartets);

/* ------------------------------------------------------------ batches
 * Many independent intra-species builds (the molecules of a batch, each a
 * species handle with its shells added and init_2body_ints called) in one
 * class launch per quartet class.  No reference counterpart: the reference
 * builds one species per call (Libint2Iface.cpp:418-618); a batch gives
 * every molecule exactly the result of that call.  The species handles must
 * outlive the batch.  Densities and results are packed device arrays: the
 * n_m x n_m block of molecule m starts at element offsets[m]. */
void* lowdin_fock_batch_new(int nmol, void* const* lints);
void lowdin_fock_batch_del(void* batch);

/* Fills offsets[0..nmol] (may be NULL) and returns the packed length. */
int64_t lowdin_fock_batch_layout(void* batch, int64_t* offsets);

/* G_m = J[D_m] + factors[m] K[D_m] for every molecule (factors: host array
 * of nmol).  Enqueued on `stream` (NULL: the library stream); with
 * synchronize != 0 waits and updates the batch statistics.  No allocation
 * once the first call has sized the batch; with the graph option the
 * kernels of a build replay as one captured HIP graph. */
int lowdin_fock_batch_compute_2body_device(void* batch, const double* d_dens, double* d_result,
                                           const double* factors, void* stream, int synchronize);

/* As lowdin_fock_batch_compute_2body_device with one device pointer per
 * molecule (host arrays of nmol device addresses) instead of packed buffers. */
int lowdin_fock_batch_compute_2body_device_ptrs(void* batch, const double* const* d_dens, double* const* d_result,
                                                const double* factors, void* stream, int synchronize);

/* time_launches: HIP events around each class launch (not in graph mode);
 * serialize: class launches on one stream; use_graph: capture once, replay. */
int lowdin_fock_batch_options(void* batch, int time_launches, int serialize, int use_graph);

int lowdin_fock_batch_get_stats(void* batch, lowdin_fock_stats* out);

int lowdin_fock_batch_launch_profile(void* batch, int cap, int* classes, float* ms, uint64_t* quartets,
                                     uint64_t* prim_quartets);

/* Autotuner: for each class launch, the kernel shape is chosen among the
 * wavefront kernel (one bra x 64 kets per wave), the LDS-strip kernel cut
 * into about {512, 2048, 8192} workgroups of 256 threads, the LDS-strip
 * kernel with one-wavefront workgroups (about 512) and, for total angular
 * momentum >= 5, the wave-cooperative kernel (one quartet per wavefront,
 * with row strips at about {2048, 8192} workgroups or without at about
 * {2048, 8192, 32768}), by timing `reps` serialised builds per candidate on
 * the given inputs.  Classes of total angular momentum >= 6 run only the
 * wave-cooperative kernel (the per-lane kernels would spill to scratch).
 * Results are unchanged by the choice up to FP64 summation order. */
int lowdin_fock_batch_autotune(void* batch, const double* d_dens, double* d_result, const double* factors,
                               void* stream, int reps);

/* The 21 per-class choices (class pair id cx*(cx+1)/2+cy over the pair
 * classes ss, ps, pp, ds, dp, dd): 0 for the wavefront kernel, else the
 * strip kernel's workgroup target, negative for one-wavefront workgroups,
 * 1000000 + target for the wave-cooperative kernel with LDS row strips and
 * 2000000 + target for the wave-cooperative kernel without strips (classes
 * of total angular momentum >= 5).  Returns 21. */
int lowdin_fock_batch_targets(void* batch, int* targets);

/* The same autotuner and choices for one species' own build (the plan behind
 * LibintInterface_compute_2body_direct and lowdin_fock_compute_2body_device):
 * d_dens / d_result n x n device buffers, `factor` the exchange factor. */
int lowdin_fock_autotune(void* lint, const double* d_dens, double* d_result, double factor, void* stream, int reps);
int lowdin_fock_targets(void* lint, int* targets);

/* ------------------------------------------------ device-resident SCF
 * The caller side of the path on the GPU: a multi-species Hartree-Fock
 * iteration whose matrices stay in HBM (openLOWDIN src/scf/MultiSCF.f90:
 * 276-440, 730-1002; WaveFunction.f90:394-467, 817-1535, 1768-2003).  Per
 * species: H-core guess, canonical orthogonaliser (overlap eigenvalues
 * >= OVERLAP_EIGEN_THRESHOLD, 1e-8 unless set), G = q^2 (J + f K) through this library's direct build (skipped
 * when build_two = 0), couplings q_i q_j J_{i<-j} through the device coupling
 * build, F = H + G + C, DIIS (8) on X^T (F D S - S D F) X over all species,
 * rocSOLVER diagonalisation, D = eta C_occ C_occ^T.  E = e_pc +
 * sum_s Tr(D_s (H_s + G_s / 2 + C_s / 2)).  No reference counterpart as a C
 * entry point: openLOWDIN runs this loop in Fortran with LAPACK on the host. */
typedef struct {
  int nocc;        /* occupied orbitals (particles / eta) */
  double eta;      /* occupation per orbital (2 closed-shell electrons, 1 otherwise) */
  double charge;   /* particle charge q */
  double factor;   /* exchange factor kappa / eta of the direct build */
  int build_two;   /* 0: G = 0 (one particle, WaveFunction.f90:867) */
  int ghost;       /* 1: SCF_GHOST_SPECIES -- its coupling is left out of the other species'
                      Fock matrices (WaveFunction.f90:1502-1503), kept in the energy */
} lowdin_scf_species;

/* lints: species handles (shells added); e_pc: point-charge repulsion;
 * stream: hipStream_t (NULL: a stream of its own).  NULL on failure. */
void* lowdin_scf_new(int nspecies, void* const* lints, const lowdin_scf_species* desc, double e_pc, void* stream);
void lowdin_scf_del(void* scf);
/* Core Hamiltonian and overlap of a species (host, n x n column-major). */
int lowdin_scf_set_onebody(void* scf, int species, const double* H, const double* S);
/* Overlap eigenvalues kept by the canonical orthogonaliser from the next
 * lowdin_scf_start on (CONTROL OVERLAP_EIGEN_THRESHOLD, default 1e-8,
 * CONTROL.f90:1513, used at WaveFunction.f90:419).  threshold > 0. */
int lowdin_scf_set_overlap_threshold(void* scf, double threshold);
/* Orthogonalisers and H-core guess densities; clears the DIIS history. */
int lowdin_scf_start(void* scf);
/* Fock matrices, energy and DIIS error (max |X^T (F D S - S D F) X|) from the
 * current densities.  external_g != 0: the G buffers (lowdin_scf_buffers)
 * already hold J + f K of the current densities (a batched build). */
int lowdin_scf_fock(void* scf, int external_g, double* energy, double* error);
/* lowdin_scf_fock in two halves: _begin enqueues the step (nothing waits
 * for the device), _end waits for its scalars and returns energy and error.
 * Many SCF objects on one stream: begin all, then end all (one wait). */
int lowdin_scf_fock_begin(void* scf, int external_g);
int lowdin_scf_fock_end(void* scf, double* energy, double* error);
/* DIIS extrapolation and new densities. */
int lowdin_scf_update(void* scf);
/* lowdin_scf_update for many SCF objects on one stream: their orthogonalised
 * Fock matrices of equal dimension are diagonalised by one batched rocSOLVER
 * call per dimension. */
int lowdin_scf_update_many(void* const* scfs, int count);
/* Iterate fock/update until |dE| < e_tol and error < d_tol.  Returns 0 when
 * converged, 1 when max_iter was reached, -1 on failure. */
int lowdin_scf_run(void* scf, int max_iter, double e_tol, double d_tol, double* energy, int* iterations);
/* Copy a species matrix to the host: what 0 density, 1 Fock, 2 G, 3 couplings. */
int lowdin_scf_get(void* scf, int species, int what, double* out);
/* Device density and G buffers of a species and its exchange factor (for a
 * batched build).  Returns build_two (0/1) or -1. */
int lowdin_scf_buffers(void* scf, int species, double** d_density, double** d_two, double* factor);
/* HBM held by the SCF object (not counting the species handles). */
int64_t lowdin_scf_device_bytes(void* scf);
/* The stream the SCF object works on (hipStream_t). */
void* lowdin_scf_stream(void* scf);

/* ------------------------------------------------ batched SCF
 * Many independent SCF jobs (molecules of one or more species) iterated
 * together on one stream, each step of an iteration one batched operation
 * over all jobs: the direct builds of every species in one class launch per
 * quartet class, couplings per species pair, and -- per group of species of
 * equal basis size -- strided-batched GEMMs (error, X^T F X, C, D), one
 * rocsolver dsyevd_strided_batched, and batched kernels for the Fock
 * assembly, traces, error maxima and the DIIS history, B matrix, solve and
 * extrapolation (the DIIS state stays on the device).  While the jobs are
 * unchanged an iteration replays two captured HIP graphs around the
 * diagonalisation.  Same iteration as lowdin_scf_* (ghost species included).  No
 * reference counterpart: openLOWDIN iterates one molecule per process. */
void* lowdin_scf_batch_new(void* stream);
void lowdin_scf_batch_del(void* eng);
/* Add a job: species handles (shells added, init_2body_ints called),
 * descriptors, point-charge repulsion, per species host H and S (n x n,
 * column-major).  Computes the orthogonaliser and the H-core guess (the
 * call waits).  Returns the job id (>= 0) or -1. */
int lowdin_scf_batch_add(void* eng, int nspecies, void* const* lints, const lowdin_scf_species* desc, double e_pc,
                         const double* const* H, const double* const* S);
/* Remove a job (its slots are reused by later jobs).  Returns 0 on success. */
int lowdin_scf_batch_remove(void* eng, int job);
/* use_graph != 0 (default): replay captured graphs while the jobs are
 * unchanged.  inline_launch != 0: every kernel of the engine runs on its own
 * stream (the default forks the class launches over the library's side
 * streams); engines on different streams then overlap freely. */
int lowdin_scf_batch_options(void* eng, int use_graph, int inline_launch);
/* Overlap eigenvalues kept by the orthogonaliser of the jobs added from now
 * on (OVERLAP_EIGEN_THRESHOLD, default 1e-8; lowdin_scf_set_overlap_threshold
 * for one SCF object).  threshold > 0. */
int lowdin_scf_batch_set_overlap_threshold(void* eng, double threshold);
/* lowdin_fock_batch_autotune on the engine's own batched build (every
 * species with a two-particle build of every job, at the jobs' current
 * densities); the choices stay with the engine and apply to the build after
 * jobs join or leave.  Call between iterations; waits. */
int lowdin_scf_batch_autotune(void* eng, int reps);
/* One SCF iteration of every job; waits for the per-job scalars.  Returns 0
 * on success (a job whose diagonalisation failed is marked, the others go on). */
int lowdin_scf_batch_step(void* eng);
/* lowdin_scf_batch_step in halves: _begin enqueues (nothing waits), _query
 * returns 1 once the iteration has finished on the device (0 otherwise),
 * _end waits and collects.  Jobs are added and removed only between
 * iterations. */
int lowdin_scf_batch_step_begin(void* eng);
int lowdin_scf_batch_step_query(void* eng);
int lowdin_scf_batch_step_end(void* eng);
/* Energy and error of the densities the job's last iteration started from,
 * iterations done.  Returns 0 running, 1 failed (message in
 * lowdin_fock_last_error), 2 removed, -1 no such job. */
int lowdin_scf_batch_status(void* eng, int job, double* energy, double* error, int* iterations);
/* what: 0 density, 1 Fock (extrapolated), 2 two-particle G, 3 coupling sum. */
int lowdin_scf_batch_get(void* eng, int job, int species, int what, double* out);
int64_t lowdin_scf_batch_device_bytes(void* eng);

#ifdef __cplusplus
}
#endif

#endif /* LOWDIN_FOCK_H */
