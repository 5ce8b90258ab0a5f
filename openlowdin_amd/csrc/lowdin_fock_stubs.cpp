// Link-level completeness of the reference's C ABI.
//
// openLOWDIN's Fortran interface block (src/ints/Libint2Interface.f90:165-373)
// binds every entry point of src/ints/Libint2Iface.h:350-453.  The direct
// Fock path, its integral-store / first-quarter-transform variants and the
// direct G12 builds are implemented in lowdin_fock.cpp; the disk,
// alpha-beta and G12 store variants below are outside this library's
// scope.  They exist so that a Fortran build of openLOWDIN links against
// liblowdin_fock.so unchanged.  Calling one records an error
// (lowdin_fock_last_error), prints it and fills its result array, where it
// has one, with NaN: the reference's callers use these results without an
// error check (Libint2Interface.f90:1108 calls the G12 directAll entry
// whenever an interparticle potential is present), so a stale array would
// otherwise pass unnoticed.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

#include "lowdin_fock.h"

namespace lowdin {
void set_unsupported(const char* name);
}

#define LOWDIN_UNSUPPORTED(name) lowdin::set_unsupported(name)

static void nan_fill(double* out, int64_t count) {
  if (!out || count <= 0) return;
  const double nan = std::nan("");
  for (int64_t k = 0; k < count; ++k) out[k] = nan;
}

static int64_t nb(void* lint) {
  const int n = lint ? LibintInterface_get_nbasis(lint) : -1;
  return n > 0 ? (int64_t)n : 0;
}

extern "C" {

void LibintInterface_compute_2body_disk(void*, const char*, double*) {
  LOWDIN_UNSUPPORTED("LibintInterface_compute_2body_disk");
}
// Libint2Iface.cpp:3625-3660: the result is the 1 x 1 alpha-beta energy term.
void LibintInterface_compute_alphabeta_direct(void*, void*, double*, double*, double* result) {
  LOWDIN_UNSUPPORTED("LibintInterface_compute_alphabeta_direct");
  nan_fill(result, 1);
}
void LibintInterface_compute_coupling_disk(void*, void*, const char*) {
  LOWDIN_UNSUPPORTED("LibintInterface_compute_coupling_disk");
}
void libintinterface_compute_g12_disk(void*, const char*, const double*, const double*, const int) {
  LOWDIN_UNSUPPORTED("libintinterface_compute_g12_disk");
}
void libintinterface_compute_g12inter_disk(void*, void*, const char*, const double*, const double*, const int) {
  LOWDIN_UNSUPPORTED("libintinterface_compute_g12inter_disk");
}
// Libint2Iface.cpp:2910-3071: n^4 integrals of lint's species.
void libintinterface_compute_g12_directAll(void* lint, double*, double* result, const double*, const double*,
                                           const int) {
  LOWDIN_UNSUPPORTED("libintinterface_compute_g12_directAll");
  const int64_t n = nb(lint);
  nan_fill(result, n * n * n * n);
}
// Libint2Iface.cpp:3073-3243: n_lint^2 n_olint^2 integrals.
void libintinterface_compute_g12inter_directAll(void* lint, void* olint, double*, double* result, const double*,
                                                const double*, const int) {
  LOWDIN_UNSUPPORTED("libintinterface_compute_g12inter_directAll");
  const int64_t n = nb(lint), no = nb(olint);
  nan_fill(result, n * n * no * no);
}
void libintinterface_buildg12_(int*, int*, int*, int*, int*, int*, void*, double*) {
  LOWDIN_UNSUPPORTED("libintinterface_buildg12_");
}
void LibintInterface_setLibint(void*, void*) { LOWDIN_UNSUPPORTED("LibintInterface_setLibint"); }

}  // extern "C"
