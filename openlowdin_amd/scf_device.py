"""Device-resident SCF: the multi-species Hartree-Fock iteration of
openLOWDIN (src/scf/MultiSCF.f90:276-440, 730-1002; WaveFunction.f90:394-467,
817-1535, 1768-2003) with every matrix in HBM (``lowdin_scf_*`` in
liblowdin_fock.so: rocBLAS/rocSOLVER for the dense algebra, this library's
direct Fock and coupling builds for G and C).  Same iteration, DIIS and
convergence test as :func:`openlowdin_amd.wavefunction.run_scf`, which keeps
the dense algebra on the host.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import libint2_interface as LI
from . import wavefunction as WF


class DeviceSCF:
    def __init__(self, system, objs: Optional[Dict[int, LI.Libint2Interface]] = None,
                 build_for_one_particle: bool = False, stream: int = 0, hcore=None, overlap=None,
                 ghost: str = "", overlap_threshold: float = 1e-8):
        self.system = system
        self.objs = {} if objs is None else objs
        ns = len(system.species)
        self.handles = [LI._instance(self.objs, system, k + 1) for k in range(ns)]
        for h in self.handles:
            h.init_2body_ints()
        desc = (LI.ScfSpeciesDesc * ns)()
        for k, sp in enumerate(system.species):
            desc[k].nocc = sp.n_occupied
            desc[k].eta = sp.eta
            desc[k].charge = sp.charge
            desc[k].factor = sp.exchange_factor
            desc[k].build_two = 1 if (sp.n_particles > 1 or build_for_one_particle) else 0
            desc[k].ghost = 1 if (ghost and sp.name == ghost) else 0
        arr = (C.c_void_p * ns)(*[h.this for h in self.handles])
        self.this = LI.lib().lowdin_scf_new(ns, arr, desc, float(system.point_charge_repulsion()), stream or None)
        if not self.this:
            LI._check("scf_new")
            raise LI.LowdinFockError("scf_new failed")
        if LI.lib().lowdin_scf_set_overlap_threshold(self.this, float(overlap_threshold)) != 0:
            LI._check("scf_set_overlap_threshold")
        if hcore is None or overlap is None:
            hcore, overlap = WF.one_body_matrices(system, self.objs)
        self.n = [h.nbasis for h in self.handles]
        for k in range(ns):
            H = np.asfortranarray(hcore[k], dtype=np.float64)
            S = np.asfortranarray(overlap[k], dtype=np.float64)
            if LI.lib().lowdin_scf_set_onebody(self.this, k, LI._dptr(H), LI._dptr(S)) != 0:
                LI._check("scf_set_onebody")
        self.energy = None
        self.iterations = 0

    def close(self):
        if getattr(self, "this", None):
            LI.lib().lowdin_scf_del(self.this)
            self.this = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def start(self):
        if LI.lib().lowdin_scf_start(self.this) != 0:
            LI._check("scf_start")

    def fock(self, external_g: bool = False):
        e, err = C.c_double(0.0), C.c_double(0.0)
        if LI.lib().lowdin_scf_fock(self.this, 1 if external_g else 0, C.byref(e), C.byref(err)) != 0:
            LI._check("scf_fock")
        return e.value, err.value

    def fock_begin(self, external_g: bool = False):
        if LI.lib().lowdin_scf_fock_begin(self.this, 1 if external_g else 0) != 0:
            LI._check("scf_fock_begin")

    def fock_end(self):
        e, err = C.c_double(0.0), C.c_double(0.0)
        if LI.lib().lowdin_scf_fock_end(self.this, C.byref(e), C.byref(err)) != 0:
            LI._check("scf_fock_end")
        return e.value, err.value

    def update(self):
        if LI.lib().lowdin_scf_update(self.this) != 0:
            LI._check("scf_update")

    def run(self, max_iter: int = 300, e_tol: float = 1e-11, d_tol: float = 1e-9):
        e = C.c_double(0.0)
        it = C.c_int(0)
        rc = LI.lib().lowdin_scf_run(self.this, max_iter, e_tol, d_tol, C.byref(e), C.byref(it))
        if rc < 0:
            LI._check("scf_run")
        self.energy, self.iterations = e.value, it.value
        self.converged = rc == 0
        return e.value

    def matrix(self, species: int, what: str = "density") -> np.ndarray:
        code = {"density": 0, "fock": 1, "two": 2, "coupling": 3}[what]
        n = self.n[species]
        out = np.zeros(n * n)
        if LI.lib().lowdin_scf_get(self.this, species, code, LI._dptr(out)) != 0:
            LI._check("scf_get")
        return out.reshape(n, n, order="F")

    def buffers(self, species: int):
        """(device density, device G, exchange factor, build_two)."""
        d, g = C.c_void_p(), C.c_void_p()
        f = C.c_double(0.0)
        bt = LI.lib().lowdin_scf_buffers(self.this, species, C.byref(d), C.byref(g), C.byref(f))
        return d.value, g.value, f.value, bt == 1

    @property
    def device_bytes(self) -> int:
        return int(LI.lib().lowdin_scf_device_bytes(self.this))

    @property
    def stream(self) -> int:
        return LI.lib().lowdin_scf_stream(self.this) or 0


def update_many(scfs):
    """DIIS + diagonalisation for many SCF objects on one stream, the
    diagonalisations batched per matrix dimension."""
    if not scfs:
        return
    arr = (C.c_void_p * len(scfs))(*[s.this for s in scfs])
    if LI.lib().lowdin_scf_update_many(arr, len(scfs)) != 0:
        LI._check("scf_update_many")
