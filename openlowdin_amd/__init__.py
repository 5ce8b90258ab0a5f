"""MI355X-native direct two-particle Fock build for openLOWDIN's APMO/NEO
Hartree-Fock: intra-species J/K and inter-species coupling matrices on the
GPU behind openLOWDIN's LibintInterface_* C ABI (liblowdin_fock.so)."""
from . import system  # noqa: F401

__all__ = ["system", "libint2_interface", "wavefunction"]
