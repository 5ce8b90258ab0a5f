"""Batched device-resident SCF over many jobs (``lowdin_scf_batch_*``).

Every iteration of every admitted job runs as one batched step on one
stream: the direct builds of all species in one class launch per quartet
class, then strided-batched products, one batched diagonalisation per group
of equal basis size and batched kernels for the Fock assembly, traces, error
and the DIIS state (kept on the device), replayed as captured HIP graphs
while the set of jobs is unchanged.  The host reads one array of scalars
per iteration.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import libint2_interface as LI
from . import wavefunction as WF


def _bind():
    L = LI.lib()
    if getattr(L, "_scf_batch_bound", False):
        return L
    vp = C.c_void_p
    L.lowdin_scf_batch_new.restype = vp
    L.lowdin_scf_batch_new.argtypes = [vp]
    L.lowdin_scf_batch_del.argtypes = [vp]
    L.lowdin_scf_batch_add.restype = C.c_int
    L.lowdin_scf_batch_add.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(LI.ScfSpeciesDesc), C.c_double,
                                       C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.POINTER(C.c_double))]
    L.lowdin_scf_batch_remove.restype = C.c_int
    L.lowdin_scf_batch_remove.argtypes = [vp, C.c_int]
    L.lowdin_scf_batch_options.argtypes = [vp, C.c_int, C.c_int]
    L.lowdin_scf_batch_set_overlap_threshold.restype = C.c_int
    L.lowdin_scf_batch_set_overlap_threshold.argtypes = [vp, C.c_double]
    L.lowdin_scf_batch_autotune.restype = C.c_int
    L.lowdin_scf_batch_autotune.argtypes = [vp, C.c_int]
    L.lowdin_scf_batch_step.restype = C.c_int
    L.lowdin_scf_batch_step.argtypes = [vp]
    for name in ("lowdin_scf_batch_step_begin", "lowdin_scf_batch_step_query", "lowdin_scf_batch_step_end"):
        getattr(L, name).restype = C.c_int
        getattr(L, name).argtypes = [vp]
    L.lowdin_scf_batch_status.restype = C.c_int
    L.lowdin_scf_batch_status.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int)]
    L.lowdin_scf_batch_get.restype = C.c_int
    L.lowdin_scf_batch_get.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.lowdin_scf_batch_device_bytes.restype = C.c_int64
    L.lowdin_scf_batch_device_bytes.argtypes = [vp]
    L._scf_batch_bound = True
    return L


class BatchSCF:
    def __init__(self, stream: int = 0, use_graph: bool = True, inline_launch: bool = False):
        L = _bind()
        self.this = L.lowdin_scf_batch_new(stream or None)
        if not self.this:
            LI._check("scf_batch_new")
            raise LI.LowdinFockError("scf_batch_new failed")
        L.lowdin_scf_batch_options(self.this, 1 if use_graph else 0, 1 if inline_launch else 0)
        self._keep: Dict[int, tuple] = {}    # job -> (handles, system) kept alive while the job is in

    def add(self, system, objs: Optional[dict] = None, mixed_tau: Optional[float] = None,
            build_for_one_particle: bool = False, hcore=None, overlap=None,
            overlap_threshold: float = 1e-8, ghost: str = "") -> int:
        """Admit a job; returns its id.  Raises LowdinFockError when the
        library refuses it (the engine is unchanged then).  overlap_threshold:
        overlap eigenvalues the job's orthogonalisers keep
        (OVERLAP_EIGEN_THRESHOLD); ghost: name of the job's SCF ghost species
        (its coupling stays out of the other species' Fock matrices, in the
        energy)."""
        objs = {} if objs is None else objs
        ns = len(system.species)
        handles = [LI._instance(objs, system, k + 1) for k in range(ns)]
        for h in handles:
            h.init_2body_ints()
            if mixed_tau is not None:
                h.set_precision("mixed", mixed_tau)
        if hcore is None or overlap is None:
            hcore, overlap = WF.one_body_matrices(system, objs)
        desc = (LI.ScfSpeciesDesc * ns)()
        for k, sp in enumerate(system.species):
            desc[k].nocc = sp.n_occupied
            desc[k].eta = sp.eta
            desc[k].charge = sp.charge
            desc[k].factor = sp.exchange_factor
            desc[k].build_two = 1 if (sp.n_particles > 1 or build_for_one_particle) else 0
            desc[k].ghost = 1 if (ghost and sp.name == ghost) else 0
        Hs = [np.asfortranarray(h, dtype=np.float64) for h in hcore]
        Ss = [np.asfortranarray(s, dtype=np.float64) for s in overlap]
        Hp = (C.POINTER(C.c_double) * ns)(*[LI._dptr(a) for a in Hs])
        Sp = (C.POINTER(C.c_double) * ns)(*[LI._dptr(a) for a in Ss])
        arr = (C.c_void_p * ns)(*[h.this for h in handles])
        if LI.lib().lowdin_scf_batch_set_overlap_threshold(self.this, float(overlap_threshold)) != 0:
            LI._check("scf_batch_set_overlap_threshold")
            raise LI.LowdinFockError("scf_batch_set_overlap_threshold failed")
        j = LI.lib().lowdin_scf_batch_add(self.this, ns, arr, desc, float(system.point_charge_repulsion()), Hp, Sp)
        if j < 0:
            LI._check("scf_batch_add")
            raise LI.LowdinFockError("scf_batch_add failed")
        self._keep[j] = (handles, objs, system)
        return j

    def autotune(self, reps: int = 1):
        """Per-class kernel shapes of the engine's batched build chosen by
        timing (kept for the builds after jobs join or leave)."""
        if LI.lib().lowdin_scf_batch_autotune(self.this, reps) != 0:
            LI._check("scf_batch_autotune")
            raise LI.LowdinFockError("scf_batch_autotune failed")

    def remove(self, job: int):
        if LI.lib().lowdin_scf_batch_remove(self.this, job) != 0:
            LI._check("scf_batch_remove")
        self._keep.pop(job, None)

    def step(self):
        if LI.lib().lowdin_scf_batch_step(self.this) != 0:
            LI._check("scf_batch_step")
            raise LI.LowdinFockError("scf_batch_step failed")

    def step_begin(self):
        if LI.lib().lowdin_scf_batch_step_begin(self.this) != 0:
            LI._check("scf_batch_step_begin")
            raise LI.LowdinFockError("scf_batch_step_begin failed")

    def step_done(self) -> bool:
        return LI.lib().lowdin_scf_batch_step_query(self.this) == 1

    def step_end(self):
        if LI.lib().lowdin_scf_batch_step_end(self.this) != 0:
            LI._check("scf_batch_step_end")
            raise LI.LowdinFockError("scf_batch_step_end failed")

    def status(self, job: int):
        """(code, energy, error, iterations); code 0 running, 1 failed, 2 removed."""
        e, err, it = C.c_double(0.0), C.c_double(0.0), C.c_int(0)
        code = LI.lib().lowdin_scf_batch_status(self.this, job, C.byref(e), C.byref(err), C.byref(it))
        msg = LI.lib().lowdin_fock_last_error() if code == 1 else None
        return code, e.value, err.value, it.value, (msg.decode() if msg else None)

    def matrix(self, job: int, species: int, what: str = "density") -> np.ndarray:
        code = {"density": 0, "fock": 1, "two": 2, "coupling": 3}[what]
        n = self._keep[job][0][species].nbasis
        out = np.zeros((n, n), order="F")
        if LI.lib().lowdin_scf_batch_get(self.this, job, species, code, LI._dptr(out)) != 0:
            LI._check("scf_batch_get")
        return out

    def device_bytes(self) -> int:
        return int(LI.lib().lowdin_scf_batch_device_bytes(self.this))

    def close(self):
        if getattr(self, "this", None):
            LI.lib().lowdin_scf_batch_del(self.this)
            self.this = None
        self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_batch(systems, e_tol: float = 1e-9, d_tol: float = 1e-7, max_iter: int = 200, use_graph: bool = True,
              mixed_tau: Optional[float] = None, options=None, autotune: bool = False):
    """Converge every system together; returns [(energy, iterations, converged)]
    in input order.  A job leaves the batch when converged or failed.
    options: per system, keyword arguments of BatchSCF.add (e.g.
    build_for_one_particle, overlap_threshold), or None.  autotune: tune
    the batched build's kernel shapes once the jobs are in."""
    eng = BatchSCF(use_graph=use_graph)
    opts = options or [None] * len(systems)
    jobs = [eng.add(s, mixed_tau=mixed_tau, **(o or {})) for s, o in zip(systems, opts)]
    if autotune:
        eng.autotune()
    res = [None] * len(systems)
    prev = {}
    live = dict(enumerate(jobs))
    while live:
        eng.step()
        for idx, j in list(live.items()):
            code, e, err, it, msg = eng.status(j)
            conv = code == 0 and idx in prev and abs(e - prev[idx]) < e_tol and err < d_tol
            prev[idx] = e
            if code == 1 or conv or it >= max_iter:
                res[idx] = (e, it, bool(conv))
                eng.remove(j)
                del live[idx]
    eng.close()
    return res
