// Link-level completeness of the reference's C ABI.
//
// openLOWDIN's Fortran interface block (src/ints/Libint2Interface.f90:165-373)
// binds every entry point of src/ints/Libint2Iface.h:350-453.  The direct
// Fock path, its integral-store / first-quarter-transform variants and the
// direct G12 builds are implemented in lowdin_fock.cpp; the disk,
// alpha-beta and G12 store variants below are outside this library's scope.  They exist so that a Fortran build of openLOWDIN
// links against liblowdin_fock.so unchanged; calling one records an error
// (lowdin_fock_last_error) and prints it, leaving outputs untouched.
#include <cstdio>
#include <string>

namespace lowdin {
void set_unsupported(const char* name);
}

#define LOWDIN_UNSUPPORTED(name) lowdin::set_unsupported(name)

extern "C" {

void LibintInterface_compute_2body_disk(void*, const char*, double*) {
  LOWDIN_UNSUPPORTED("LibintInterface_compute_2body_disk");
}
void LibintInterface_compute_alphabeta_direct(void*, void*, double*, double*, double*) {
  LOWDIN_UNSUPPORTED("LibintInterface_compute_alphabeta_direct");
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
void libintinterface_compute_g12_directAll(void*, double*, double*, const double*, const double*, const int) {
  LOWDIN_UNSUPPORTED("libintinterface_compute_g12_directAll");
}
void libintinterface_compute_g12inter_directAll(void*, void*, double*, double*, const double*, const double*,
                                                const int) {
  LOWDIN_UNSUPPORTED("libintinterface_compute_g12inter_directAll");
}
void libintinterface_buildg12_(int*, int*, int*, int*, int*, int*, void*, double*) {
  LOWDIN_UNSUPPORTED("libintinterface_buildg12_");
}
void LibintInterface_setLibint(void*, void*) { LOWDIN_UNSUPPORTED("LibintInterface_setLibint"); }

}  // extern "C"
