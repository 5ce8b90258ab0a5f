"""Python mirror of openLOWDIN's Fortran-side interface to the direct Fock
build, over the C ABI of liblowdin_fock.so.

* :class:`Libint2Interface` mirrors ``type(Libint2Interface)`` and its
  constructor / destructor / direct wrappers
  (``src/ints/Libint2Interface.f90:36-39, 382-460, 514-547, 637-675``):
  one handle per species, shells added in particle order then contraction
  order with raw coefficients, the pair list rebuilt before every intra
  build, result arrays allocated by the caller side.
* :func:`get_direct_intra_repulsion_matrix` and
  :func:`get_direct_inter_repulsion_matrix` mirror
  ``DirectIntegralManager_getDirect{Intra,Inter}RepulsionMatrix``
  (``src/ints/DirectIntegralManager.f90:66-179``), including the STOP on an
  integral scheme other than LIBINT (raised here as ValueError).

Matrices cross the boundary as Fortran column-major float64 arrays; the
library is loaded from the package directory and its absence is an error
(there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblowdin_fock.so")
_LIB = None


class LowdinFockError(RuntimeError):
    pass


class Stats(C.Structure):
    _fields_ = [("quartets", C.c_uint64), ("eris", C.c_uint64), ("prim_quartets", C.c_uint64),
                ("kernel_ms", C.c_double), ("total_ms", C.c_double), ("quartets_fp32", C.c_uint64),
                ("eris_fp32", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ScfSpeciesDesc(C.Structure):
    """lowdin_scf_species."""
    _fields_ = [("nocc", C.c_int), ("eta", C.c_double), ("charge", C.c_double), ("factor", C.c_double),
                ("build_two", C.c_int), ("ghost", C.c_int)]


def lib():
    """Load liblowdin_fock.so (built by ``__graft_entry__.build()``)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise LowdinFockError("liblowdin_fock.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.LibintInterface_new.argtypes = [C.c_int, C.c_int, C.c_bool, C.c_bool]
        L.LibintInterface_new.restype = vp
        L.LibintInterface_del.argtypes = [vp]
        L.LibintInterface_add_particle.argtypes = [vp, C.c_int, dp]
        L.LibintInterface_add_shell.argtypes = [vp, dp, dp, dp, C.c_int, C.c_int]
        L.LibintInterface_get_nbasis.argtypes = [vp]
        L.LibintInterface_get_nbasis.restype = C.c_int
        L.LibintInterface_init_2body_ints.argtypes = [vp]
        L.LibintInterface_compute_2body_direct.argtypes = [vp, dp, dp, dp]
        L.LibintInterface_compute_coupling_direct.argtypes = [vp, vp, dp, dp]
        L.LibintInterface_compute_2body_directAll.argtypes = [vp, dp, dp]
        L.LibintInterface_compute_2body_directIT.argtypes = [vp, dp, dp, dp, ip]
        L.LibintInterface_compute_coupling_directAll.argtypes = [vp, vp, dp, dp]
        L.LibintInterface_compute_coupling_directIT.argtypes = [vp, vp, dp, dp, dp, ip]
        L.lowdin_fock_store_device.argtypes = [vp, vp, C.c_int, vp, vp, vp]
        L.lowdin_fock_store_device.restype = C.c_int
        L.libintinterface_compute_g12_direct.argtypes = [vp, dp, dp, dp, dp, dp, C.c_int]
        L.libintinterface_compute_g12inter_direct.argtypes = [vp, vp, dp, dp, dp, dp, C.c_int]
        L.lowdin_fock_g12_device.argtypes = [vp, vp, vp, vp, C.c_double, dp, dp, C.c_int]
        L.lowdin_fock_g12_device.restype = C.c_int
        L.lowdin_fock_autotune.argtypes = [vp, vp, vp, C.c_double, vp, C.c_int]
        L.lowdin_fock_autotune.restype = C.c_int
        L.lowdin_fock_targets.argtypes = [vp, ip]
        L.lowdin_fock_targets.restype = C.c_int
        L.lowdin_fock_last_error.restype = C.c_char_p
        L.lowdin_fock_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.lowdin_fock_shellpair_list.argtypes = [vp, ip, ip, C.c_int]
        L.lowdin_fock_shellpair_list.restype = C.c_int
        L.lowdin_fock_schwarz.argtypes = [vp, dp]
        L.lowdin_fock_schwarz.restype = C.c_int
        L.lowdin_fock_coupling_pair.argtypes = [vp, vp, dp, dp, dp, dp]
        L.lowdin_fock_coupling_pair.restype = C.c_int
        L.lowdin_fock_compute_2body_device.argtypes = [vp, vp, vp, C.c_double, vp, C.c_int]
        L.lowdin_fock_compute_2body_device.restype = C.c_int
        L.lowdin_fock_compute_2body_device_noalloc.argtypes = [vp, vp, vp, C.c_double, vp]
        L.lowdin_fock_compute_2body_device_noalloc.restype = C.c_int
        L.lowdin_fock_set_screen.argtypes = [vp, C.c_double]
        L.lowdin_fock_set_screen.restype = C.c_int
        L.lowdin_fock_set_setup_each_build.argtypes = [vp, C.c_int]
        L.lowdin_fock_set_setup_each_build.restype = C.c_int
        L.lowdin_fock_time_launches.argtypes = [vp, C.c_int]
        L.lowdin_fock_serialize_launches.argtypes = [vp, C.c_int]
        L.lowdin_fock_1body.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp]
        L.lowdin_fock_1body.restype = C.c_int
        L.LibintInterface_compute_1body_ints.argtypes = [vp, C.c_int, dp]
        L.lowdin_fock_set_precision.argtypes = [vp, C.c_int, C.c_double]
        L.lowdin_fock_set_precision.restype = C.c_int
        L.lowdin_fock_launch_profile.argtypes = [vp, C.c_int, ip, C.POINTER(C.c_float),
                                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.lowdin_fock_launch_profile.restype = C.c_int
        L.lowdin_fock_batch_new.argtypes = [C.c_int, C.POINTER(vp)]
        L.lowdin_fock_batch_new.restype = vp
        L.lowdin_fock_batch_del.argtypes = [vp]
        L.lowdin_fock_batch_layout.argtypes = [vp, C.POINTER(C.c_int64)]
        L.lowdin_fock_batch_layout.restype = C.c_int64
        L.lowdin_fock_batch_compute_2body_device.argtypes = [vp, vp, vp, dp, vp, C.c_int]
        L.lowdin_fock_batch_compute_2body_device.restype = C.c_int
        L.lowdin_fock_batch_options.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.lowdin_fock_batch_options.restype = C.c_int
        L.lowdin_fock_batch_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.lowdin_fock_batch_get_stats.restype = C.c_int
        L.lowdin_fock_batch_launch_profile.argtypes = [vp, C.c_int, ip, C.POINTER(C.c_float),
                                                       C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.lowdin_fock_batch_launch_profile.restype = C.c_int
        L.lowdin_fock_batch_autotune.argtypes = [vp, vp, vp, dp, vp, C.c_int]
        L.lowdin_fock_batch_autotune.restype = C.c_int
        L.lowdin_fock_batch_targets.argtypes = [vp, ip]
        L.lowdin_fock_batch_targets.restype = C.c_int
        L.lowdin_fock_batch_compute_2body_device_ptrs.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), dp, vp, C.c_int]
        L.lowdin_fock_batch_compute_2body_device_ptrs.restype = C.c_int
        L.lowdin_fock_coupling_pair_device.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_int]
        L.lowdin_fock_coupling_pair_device.restype = C.c_int
        L.lowdin_scf_new.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(ScfSpeciesDesc), C.c_double, vp]
        L.lowdin_scf_new.restype = vp
        L.lowdin_scf_del.argtypes = [vp]
        L.lowdin_scf_set_onebody.argtypes = [vp, C.c_int, dp, dp]
        L.lowdin_scf_set_onebody.restype = C.c_int
        L.lowdin_scf_set_overlap_threshold.argtypes = [vp, C.c_double]
        L.lowdin_scf_set_overlap_threshold.restype = C.c_int
        L.lowdin_scf_start.argtypes = [vp]
        L.lowdin_scf_start.restype = C.c_int
        L.lowdin_scf_fock.argtypes = [vp, C.c_int, dp, dp]
        L.lowdin_scf_fock.restype = C.c_int
        L.lowdin_scf_update.argtypes = [vp]
        L.lowdin_scf_update.restype = C.c_int
        L.lowdin_scf_fock_begin.argtypes = [vp, C.c_int]
        L.lowdin_scf_fock_begin.restype = C.c_int
        L.lowdin_scf_fock_end.argtypes = [vp, dp, dp]
        L.lowdin_scf_fock_end.restype = C.c_int
        L.lowdin_scf_update_many.argtypes = [C.POINTER(vp), C.c_int]
        L.lowdin_scf_update_many.restype = C.c_int
        L.lowdin_scf_run.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp, ip]
        L.lowdin_scf_run.restype = C.c_int
        L.lowdin_scf_get.argtypes = [vp, C.c_int, C.c_int, dp]
        L.lowdin_scf_get.restype = C.c_int
        L.lowdin_scf_buffers.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(vp), dp]
        L.lowdin_scf_buffers.restype = C.c_int
        L.lowdin_scf_device_bytes.argtypes = [vp]
        L.lowdin_scf_device_bytes.restype = C.c_int64
        L.lowdin_scf_stream.argtypes = [vp]
        L.lowdin_scf_stream.restype = vp
        _LIB = L
    return _LIB


def _dptr(a: Optional[np.ndarray]):
    if a is None:
        return C.cast(None, C.POINTER(C.c_double))
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(where: str):
    msg = lib().lowdin_fock_last_error()
    if msg:
        raise LowdinFockError("%s: %s" % (where, msg.decode()))


def _fortran(m: np.ndarray, n: int) -> np.ndarray:
    a = np.asfortranarray(m, dtype=np.float64)
    if a.shape != (n, n):
        raise ValueError("expected an %d x %d matrix, got %s" % (n, n, a.shape))
    return a


def _geminal(coefficients, exponents):
    c = np.ascontiguousarray(coefficients, dtype=np.float64).reshape(-1)
    g = np.ascontiguousarray(exponents, dtype=np.float64).reshape(-1)
    if c.shape != g.shape:
        raise ValueError("geminal coefficients and exponents differ in length")
    return (c if len(c) else np.zeros(1)), (g if len(g) else np.zeros(1))


class Libint2Interface:
    """One species' handle (``type(Libint2Interface)``)."""

    def __init__(self, species, species_id: int, stack_size: int = 30000, parallel: bool = True):
        self.species = species
        self.species_id = species_id
        self.this = lib().LibintInterface_new(stack_size, species_id, bool(species.is_electron), bool(parallel))
        if not self.this:
            raise LowdinFockError("LibintInterface_new returned NULL")
        self._keep = []
        # Libint2Interface_constructor: per particle add_particle, then the
        # particle's contractions (Libint2Interface.f90:412-446).  The species
        # model here stores shells already in that order.
        for sh in species.shells:
            e = np.ascontiguousarray(sh.exps, dtype=np.float64)
            c = np.ascontiguousarray(sh.coefs, dtype=np.float64)
            o = np.ascontiguousarray(sh.origin, dtype=np.float64)
            lib().LibintInterface_add_shell(self.this, _dptr(e), _dptr(c), _dptr(o), int(sh.l), len(e))
            _check("add_shell")
        self.nbasis = lib().LibintInterface_get_nbasis(self.this)
        self.is_instanced = True

    def close(self):
        if getattr(self, "this", None):
            lib().LibintInterface_del(self.this)
            self.this = None
            self.is_instanced = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # Libint2Interface_compute2BodyIntraspecies_direct (Libint2Interface.f90:514-547)
    def compute_2body_intraspecies_direct(self, density: np.ndarray, factor: float) -> np.ndarray:
        n = self.nbasis
        D = _fortran(density, n)
        G = np.zeros((n, n), dtype=np.float64, order="F")
        f = C.c_double(factor)
        lib().LibintInterface_init_2body_ints(self.this)
        _check("init_2body_ints")
        lib().LibintInterface_compute_2body_direct(self.this, _dptr(D), _dptr(G), C.byref(f))
        _check("compute_2body_direct")
        return G

    # Libint2Interface_compute2BodyInterspecies_direct (Libint2Interface.f90:637-675)
    def compute_2body_interspecies_direct(self, other: "Libint2Interface", density_other: np.ndarray) -> np.ndarray:
        n = self.nbasis
        D = _fortran(density_other, other.nbasis)
        B = np.zeros((n, n), dtype=np.float64, order="F")
        lib().LibintInterface_compute_coupling_direct(self.this, other.this, _dptr(D), _dptr(B))
        _check("compute_coupling_direct")
        return B

    # Libint2Interface_compute2BodyIntraspecies_direct_all, Coulomb branch
    # (Libint2Interface.f90:1065-1115 -> Libint2Iface.cpp:3736-3750)
    def compute_2body_intraspecies_direct_all(self, density: np.ndarray, ints: np.ndarray = None) -> np.ndarray:
        """Every integral of the Fock build's quartet set (the pair list of
        the last init_2body_ints, the density-block screen for electrons),
        |raw| > 1e-10, normalised, all eight images.  Returns the
        (n, n, n, n) tensor, (i j|k l) at [i, j, k, l]; elements not written
        keep the values of ``ints`` (zeros when None)."""
        n = self.nbasis
        D = _fortran(density, n)
        A = np.zeros(n ** 4) if ints is None else np.array(ints, dtype=np.float64).reshape(-1).copy()
        lib().LibintInterface_compute_2body_directAll(self.this, _dptr(D), _dptr(A))
        _check("compute_2body_directAll")
        return A.reshape(n, n, n, n)

    # Libint2Interface_compute2BodyIntraspecies_direct_IT
    # (Libint2Interface.f90:551-579 -> Libint2Iface.cpp:3528-3549)
    def compute_2body_intraspecies_direct_IT(self, density: np.ndarray, coefficients: np.ndarray,
                                             p: int) -> np.ndarray:
        """First-quarter transform on orbital row p: GG[j, k, l] =
        sum_i C(p, i) (i j|k l), C(p, i) = coefficients[p, i] (n x n)."""
        n = self.nbasis
        D = _fortran(density, n)
        Cm = _fortran(coefficients, n)
        GG = np.zeros(n ** 3)
        pp = C.c_int(int(p))
        lib().LibintInterface_compute_2body_directIT(self.this, _dptr(D), _dptr(Cm), _dptr(GG), C.byref(pp))
        _check("compute_2body_directIT")
        return GG.reshape(n, n, n)

    # Libint2Interface_compute2BodyInterspecies_direct_all, Coulomb branch
    # (Libint2Interface.f90:1117-1179 -> Libint2Iface.cpp:3752-3773)
    def compute_2body_interspecies_direct_all(self, other: "Libint2Interface", density: np.ndarray,
                                              ints: np.ndarray = None) -> np.ndarray:
        """(x1 x2|y1 y2) over every pair of both species, x of the lower
        species id: the (n_x, n_x, n_y, n_y) tensor with the reference's
        strides (n n_o^2, n_o^2, n_o, 1), n = this species' basis size."""
        n, no = self.nbasis, other.nbasis
        D = np.asfortranarray(density, dtype=np.float64)   # passed through, not read
        A = np.zeros(n * n * no * no) if ints is None else np.array(ints, dtype=np.float64).reshape(-1).copy()
        lib().LibintInterface_compute_coupling_directAll(self.this, other.this, _dptr(D), _dptr(A))
        _check("compute_coupling_directAll")
        if self.species_id > other.species_id:
            return A.reshape(no, no, n, n)
        return A.reshape(n, n, no, no)

    # Libint2Interface_compute2BodyInterspecies_direct_IT
    # (Libint2Interface.f90:677-707 -> Libint2Iface.cpp:3591-3623)
    def compute_2body_interspecies_direct_IT(self, other: "Libint2Interface", density: np.ndarray,
                                             coefficients: np.ndarray, p: int) -> np.ndarray:
        """GG[v, l, s] (n x n_o x n_o) = sum_u C(p, u) (u v|l s), u, v of this
        species, l, s of the other."""
        n, no = self.nbasis, other.nbasis
        D = np.asfortranarray(density, dtype=np.float64)   # passed through, not read
        Cm = _fortran(coefficients, n)
        GG = np.zeros(n * no * no)
        pp = C.c_int(int(p))
        lib().LibintInterface_compute_coupling_directIT(self.this, other.this, _dptr(D), _dptr(Cm), _dptr(GG),
                                                        C.byref(pp))
        _check("compute_coupling_directIT")
        return GG.reshape(n, no, no)

    # Libint2Interface_computeG12Intraspecies_direct
    # (Libint2Interface.f90:935-992 -> Libint2Iface.cpp:3687-3706)
    def compute_g12_intraspecies_direct(self, density: np.ndarray, factor: float, coefficients,
                                        exponents) -> np.ndarray:
        """J + factor K with the contracted Gaussian geminal
        sum_k coefficients[k] exp(-exponents[k] r12^2) (the interparticle
        potential of the species with itself); the pair list is rebuilt
        first, as the Fortran wrapper does."""
        n = self.nbasis
        D = _fortran(density, n)
        c, g = _geminal(coefficients, exponents)
        G = np.zeros((n, n), dtype=np.float64, order="F")
        f = C.c_double(factor)
        lib().LibintInterface_init_2body_ints(self.this)
        _check("init_2body_ints")
        lib().libintinterface_compute_g12_direct(self.this, _dptr(D), _dptr(G), C.byref(f), _dptr(c), _dptr(g),
                                                 len(coefficients))
        _check("compute_g12_direct")
        return G

    # Libint2Interface_computeG12Interspecies_direct
    # (Libint2Interface.f90:995-1061 -> Libint2Iface.cpp:3708-3734)
    def compute_g12_interspecies_direct(self, other: "Libint2Interface", density_other: np.ndarray, coefficients,
                                        exponents) -> np.ndarray:
        """The geminal coupling on this species from the other's density, in
        the layout of compute_2body_interspecies_direct."""
        n = self.nbasis
        D = _fortran(density_other, other.nbasis)
        c, g = _geminal(coefficients, exponents)
        B = np.zeros((n, n), dtype=np.float64, order="F")
        lib().libintinterface_compute_g12inter_direct(self.this, other.this, _dptr(D), _dptr(B), _dptr(c), _dptr(g),
                                                      len(coefficients))
        _check("compute_g12inter_direct")
        return B

    def g12_device(self, d_dens: int, d_result: int, coefficients, exponents, factor: float = 0.0,
                   other: "Libint2Interface" = None):
        """Device-resident G12 build (other None: intra-species with factor;
        otherwise the coupling from the other species' density)."""
        c, g = _geminal(coefficients, exponents)
        rc = lib().lowdin_fock_g12_device(self.this, other.this if other is not None else None, d_dens, d_result,
                                          float(factor), _dptr(c), _dptr(g), len(coefficients))
        if rc != 0:
            _check("g12_device")
            raise LowdinFockError("g12_device failed")

    def store_device(self, which: int, d_out: int, d_dens: int = 0, d_crow: int = 0,
                     other: "Libint2Interface" = None):
        """Device-resident store / transform (which 0: intra all, 1: intra
        first quarter, 2: inter all, 3: inter first quarter); device
        addresses, synchronised on return."""
        rc = lib().lowdin_fock_store_device(self.this, other.this if other is not None else None, int(which),
                                            d_dens or None, d_crow or None, d_out)
        if rc != 0:
            _check("store_device")
            raise LowdinFockError("store_device failed")

    def coupling_pair(self, other: "Libint2Interface", d_self: np.ndarray, d_other: np.ndarray):
        """Both coupling directions from one pass over the quartets."""
        Da = _fortran(d_self, self.nbasis)
        Db = _fortran(d_other, other.nbasis)
        Ra = np.zeros((self.nbasis, self.nbasis), order="F")
        Rb = np.zeros((other.nbasis, other.nbasis), order="F")
        rc = lib().lowdin_fock_coupling_pair(self.this, other.this, _dptr(Da), _dptr(Db), _dptr(Ra), _dptr(Rb))
        if rc != 0:
            _check("coupling_pair")
            raise LowdinFockError("coupling_pair failed")
        return Ra, Rb

    def compute_2body_device(self, d_dens: int, d_result: int, factor: float, stream: int = 0,
                             synchronize: bool = True):
        """Device-resident build: d_dens / d_result are device addresses of
        n x n FP64 buffers; the work is enqueued on ``stream`` (hipStream_t
        address, 0 = the library's stream)."""
        rc = lib().lowdin_fock_compute_2body_device(self.this, d_dens, d_result, factor, stream or None,
                                                     1 if synchronize else 0)
        if rc != 0:
            _check("compute_2body_device")
            raise LowdinFockError("compute_2body_device failed")

    def autotune(self, d_dens: int, d_result: int, factor: float, stream: int = 0, reps: int = 1):
        """Per-class kernel shape of this species' build chosen by timing
        (wavefront kernel or LDS strips at 512 / 2048 / 8192 workgroups)."""
        if lib().lowdin_fock_autotune(self.this, d_dens, d_result, factor, stream or None, reps) != 0:
            _check("autotune")
            raise LowdinFockError("autotune failed")

    def targets(self):
        t = (C.c_int * 21)()
        lib().lowdin_fock_targets(self.this, t)
        return list(t)

    def time_launches(self, enable: bool = True):
        lib().lowdin_fock_time_launches(self.this, 1 if enable else 0)

    def one_body(self, kind: int, charges=()) -> np.ndarray:
        """kind 1 overlap, 2 kinetic (-1/2 nabla^2), 3 attraction to point
        charges [(Z, xyz bohr)]: sum_C -Z_C/|r - C|.  Normalised, n x n."""
        Z = np.ascontiguousarray([z for z, _ in charges] or [0.0], dtype=np.float64)
        X = np.ascontiguousarray([x for _, x in charges] or [[0.0, 0.0, 0.0]], dtype=np.float64).reshape(-1)
        M = np.zeros((self.nbasis, self.nbasis))
        nc = len(charges) if kind == 3 else 0
        if lib().lowdin_fock_1body(self.this, int(kind), nc, _dptr(Z), _dptr(X), _dptr(M)) != 0:
            _check("1body")
        return M

    def set_precision(self, mode: str = "fp64", tau: float = 0.0):
        """'fp64' (reference parity) or 'mixed': quartets with
        Dnorm_abcd * Qn_ab * Qn_cd < tau (normalised Schwarz bounds times the
        largest |D| of the quartet's density blocks) are evaluated in FP32."""
        m = {"fp64": 0, "mixed": 1}[mode]
        if lib().lowdin_fock_set_precision(self.this, m, float(tau)) != 0:
            _check("set_precision")

    def set_screen(self, eps: float):
        """Quartet screen threshold (default DBL_EPSILON, the reference's)."""
        if lib().lowdin_fock_set_screen(self.this, float(eps)) != 0:
            _check("set_screen")

    def set_setup_each_build(self, enable: bool = True):
        """Recompute pair records and Schwarz bounds before every build (the
        reference's per-call setup)."""
        lib().lowdin_fock_set_setup_each_build(self.this, 1 if enable else 0)

    def coupling_pair_device(self, other: "Libint2Interface", d_self: int, d_other: int, d_rself: int,
                             d_rother: int, stream: int = 0, synchronize: bool = True):
        """Both coupling directions in one pass on device buffers: d_rself =
        coupling on this species from d_other, d_rother = on the other from
        d_self (row-major images of the reference's results)."""
        rc = lib().lowdin_fock_coupling_pair_device(self.this, other.this, d_self, d_other, d_rself, d_rother,
                                                    stream or None, 1 if synchronize else 0)
        if rc != 0:
            _check("coupling_pair_device")
            raise LowdinFockError("coupling_pair_device failed")

    def serialize_launches(self, enable: bool = True):
        lib().lowdin_fock_serialize_launches(self.this, 1 if enable else 0)

    def launch_profile(self):
        """[(class code, ms, quartets, primitive quartets)] of the last intra build."""
        cap = 64
        cls = np.zeros(cap, dtype=np.int32)
        ms = np.zeros(cap, dtype=np.float32)
        q = np.zeros(cap, dtype=np.uint64)
        pq = np.zeros(cap, dtype=np.uint64)
        n = lib().lowdin_fock_launch_profile(self.this, cap, cls.ctypes.data_as(C.POINTER(C.c_int)),
                                             ms.ctypes.data_as(C.POINTER(C.c_float)),
                                             q.ctypes.data_as(C.POINTER(C.c_uint64)),
                                             pq.ctypes.data_as(C.POINTER(C.c_uint64)))
        return [("%04d" % cls[k], float(ms[k]), int(q[k]), int(pq[k])) for k in range(min(n, cap))]

    def stats(self) -> dict:
        s = Stats()
        lib().lowdin_fock_get_stats(self.this, C.byref(s))
        return s.as_dict()

    def shellpair_list(self) -> np.ndarray:
        ns = len(self.species.shells)
        cap = ns * (ns + 1) // 2
        s1 = np.zeros(cap, dtype=np.int32)
        s2 = np.zeros(cap, dtype=np.int32)
        n = lib().lowdin_fock_shellpair_list(self.this, s1.ctypes.data_as(C.POINTER(C.c_int)),
                                             s2.ctypes.data_as(C.POINTER(C.c_int)), cap)
        if n < 0:
            raise LowdinFockError("pair list not built (call init_2body_ints first)")
        return np.stack([s1[:n], s2[:n]], axis=1)

    def init_2body_ints(self):
        lib().LibintInterface_init_2body_ints(self.this)
        _check("init_2body_ints")

    def schwarz(self) -> np.ndarray:
        ns = len(self.species.shells)
        Q = np.zeros(ns * ns)
        if lib().lowdin_fock_schwarz(self.this, _dptr(Q)) != 0:
            _check("schwarz")
        return Q.reshape(ns, ns)


# ------------------------------------------------ DirectIntegralManager
def _instance(objs: Dict[int, Libint2Interface], system, species_id: int) -> Libint2Interface:
    """A species' handle, constructed with its pair list on first use
    (DirectIntegralManager_constructor, DirectIntegralManager.f90:457-460)."""
    if species_id not in objs or not objs[species_id].is_instanced:
        objs[species_id] = Libint2Interface(system.species[species_id - 1], species_id)
        objs[species_id].init_2body_ints()
    return objs[species_id]


def get_direct_intra_repulsion_matrix(species_id: int, scheme: str, density: np.ndarray, factor: float,
                                      system, objs: Dict[int, Libint2Interface]) -> np.ndarray:
    """DirectIntegralManager_getDirectIntraRepulsionMatrix (DirectIntegralManager.f90:66-130).
    species_id is 1-based, as in the reference."""
    if scheme.strip().upper() != "LIBINT":
        raise ValueError("The integral method selected has not been implemented")
    return _instance(objs, system, species_id).compute_2body_intraspecies_direct(np.array(density), factor)


def get_direct_inter_repulsion_matrix(species_id: int, other_species_id: int, scheme: str, density: np.ndarray,
                                      system, objs: Dict[int, Libint2Interface]) -> np.ndarray:
    """DirectIntegralManager_getDirectInterRepulsionMatrix (DirectIntegralManager.f90:139-179)."""
    if scheme.strip().upper() != "LIBINT":
        raise ValueError("The integral method selected has not been implemented")
    a = _instance(objs, system, species_id)
    b = _instance(objs, system, other_species_id)
    return a.compute_2body_interspecies_direct(b, np.array(density))


def get_direct_intra_repulsion_first_quarter(species_id: int, scheme: str, density: np.ndarray,
                                             coefficients: np.ndarray, p: int, system,
                                             objs: Dict[int, Libint2Interface]) -> np.ndarray:
    """DirectIntegralManager_getDirectIntraRepulsionFirstQuarter
    (DirectIntegralManager.f90:274-327): the (n, n, n) first-quarter
    transform on coefficient row p.  ``scheme`` is not consulted, as in the
    reference (its select case is commented out)."""
    return _instance(objs, system, species_id).compute_2body_intraspecies_direct_IT(
        np.array(density), np.array(coefficients), p)


def get_direct_inter_repulsion_first_quarter(species_id: int, other_species_id: int, scheme: str,
                                             density: np.ndarray, coefficients: np.ndarray, p: int, system,
                                             objs: Dict[int, Libint2Interface]) -> np.ndarray:
    """DirectIntegralManager_getDirectInterRepulsionFirstQuarter
    (DirectIntegralManager.f90:336-391)."""
    a = _instance(objs, system, species_id)
    b = _instance(objs, system, other_species_id)
    return a.compute_2body_interspecies_direct_IT(b, np.array(density), np.array(coefficients), p)


def get_direct_intra_repulsion_integrals_all(species_id: int, density: np.ndarray, system,
                                             objs: Dict[int, Libint2Interface], ints: np.ndarray = None):
    """DirectIntegralManager_getDirectIntraRepulsionIntegralsAll
    (DirectIntegralManager.f90:957-990): the species' (n, n, n, n) store."""
    return _instance(objs, system, species_id).compute_2body_intraspecies_direct_all(np.array(density), ints)


def get_direct_inter_repulsion_integrals_all(species_id: int, other_species_id: int, density: np.ndarray, system,
                                             objs: Dict[int, Libint2Interface], ints: np.ndarray = None):
    """DirectIntegralManager_getDirectInterRepulsionIntegralsAll
    (DirectIntegralManager.f90:999-1034)."""
    a = _instance(objs, system, species_id)
    b = _instance(objs, system, other_species_id)
    return a.compute_2body_interspecies_direct_all(b, np.array(density), ints)
