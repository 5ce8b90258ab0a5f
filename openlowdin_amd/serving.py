"""Served SCF with continuous batching.

Requests (molecular systems) arrive over time; the engine admits them while
their estimated HBM footprint fits a budget, keeps every admitted job's
densities and Fock matrices resident on the device (one
:class:`openlowdin_amd.scf_device.DeviceSCF` per job), and advances all
admitted jobs together: each tick is ONE batched intra-species Fock build
over every species of every active job (``lowdin_fock_batch_*``: one class
launch per quartet class, replayed as a HIP graph while the set of jobs is
unchanged), then each job's couplings, Fock assembly, DIIS and
diagonalisation on the device.  Jobs that converge leave at the end of the
tick and free their slot and memory; waiting jobs join at the next tick.

A job's result does not depend on the company it keeps: the batched build
gives every molecule what a build of it alone gives (to FP64 summation
order), and the rest of the iteration is per job.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import hip as H
from . import libint2_interface as LI
from . import scf_device as SD


def estimate_device_bytes(system) -> int:
    """Upper estimate of a job's HBM footprint: the SCF matrices (12 n^2 +
    2 x 8 n^2 DIIS history per species) plus the Fock-build data (pair and
    primitive-pair records of every shell pair, task lists, build buffers)."""
    total = 0
    for sp in system.species:
        n = sp.nbasis
        S = len(sp.shells)
        prims = sum(len(s.exps) for s in sp.shells)
        npairs = S * (S + 1) // 2
        pp = (prims * prims) // 2 + prims
        total += (12 + 16) * n * n * 8 + 4 * n * n * 8                    # SCF + build buffers
        total += 2 * npairs * 80 + 2 * pp * (80 + 40) + S * S * 8        # pair sets, records, Dnorm
        total += 16 * max(npairs, 1) * 4                                   # task lists
    return int(total)


@dataclass
class Job:
    jid: int
    label: str
    system: object
    arrival: float                     # seconds (engine clock)
    admitted: Optional[float] = None
    finished: Optional[float] = None
    energy: Optional[float] = None
    iterations: int = 0
    converged: bool = False
    bytes: int = 0
    scf: Optional[SD.DeviceSCF] = None
    objs: dict = field(default_factory=dict)
    e_old: Optional[float] = None


class ScfEngine:
    def __init__(self, hbm_budget: float = 200e9, max_jobs: int = 64, mixed_tau: Optional[float] = None,
                 use_graph: bool = True, e_tol: float = 1e-9, d_tol: float = 1e-7, max_iter: int = 200):
        self.budget = float(hbm_budget)
        self.max_jobs = max_jobs
        self.mixed_tau = mixed_tau
        self.use_graph = use_graph
        self.e_tol, self.d_tol, self.max_iter = e_tol, d_tol, max_iter
        self.stream = H.Stream()
        self.waiting: List[Job] = []
        self.active: List[Job] = []
        self.done: List[Job] = []
        self.used = 0
        self._batch = None
        self._batch_key = None
        self.ticks = 0
        self.build_ms = []          # device time of each tick's batched build
        self._ev = (H.Event(), H.Event())
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, job: Job):
        job.bytes = estimate_device_bytes(job.system)
        if job.bytes > self.budget:
            raise ValueError("job %s needs %.1f GB, above the HBM budget" % (job.label, job.bytes / 1e9))
        self.waiting.append(job)

    # -- admission: first come first served while memory and slots allow
    def _admit(self):
        now = self.now()
        still = []
        for job in sorted(self.waiting, key=lambda j: j.arrival):
            if job.arrival > now:
                still.append(job)
                continue
            if len(self.active) < self.max_jobs and self.used + job.bytes <= self.budget:
                job.admitted = now
                job.scf = SD.DeviceSCF(job.system, objs=job.objs, stream=self.stream.handle)
                if self.mixed_tau is not None:
                    for h in job.scf.handles:
                        h.set_precision("mixed", self.mixed_tau)
                job.scf.start()
                self.used += job.bytes
                self.active.append(job)
            else:
                still.append(job)
        self.waiting = still

    def _build(self):
        """One batched build of every active species with two-particle terms."""
        from . import batch as B
        entries = []
        for job in self.active:
            for k in range(len(job.system.species)):
                d, g, f, bt = job.scf.buffers(k)
                if bt:
                    entries.append((job.scf.handles[k], d, g, f))
        if not entries:
            return
        key = tuple(id(e[0]) for e in entries)
        if key != self._batch_key:
            self._batch = B.FockBatch([e[0] for e in entries], [e[3] for e in entries], init=False)
            self._batch.options(use_graph=self.use_graph)
            self._batch_key = key
        self._batch.compute_device_ptrs([e[1] for e in entries], [e[2] for e in entries], self.stream.handle,
                                        synchronize=False)

    def tick(self):
        self._admit()
        if not self.active:
            return
        self._ev[0].record(self.stream)
        self._build()
        self._ev[1].record(self.stream)
        for job in self.active:
            job.scf.fock_begin(external_g=True)
        keep = []
        for k, job in enumerate(self.active):
            e, err = job.scf.fock_end()
            if k == 0:
                self.build_ms.append(self._ev[0].elapsed_ms(self._ev[1]))
            job.iterations += 1
            conv = job.e_old is not None and abs(e - job.e_old) < self.e_tol and err < self.d_tol
            if conv or job.iterations >= self.max_iter:
                job.energy = e
                job.converged = conv
                job.finished = self.now()
                job.scf.close()
                job.scf = None
                self.used -= job.bytes
                self.done.append(job)
            else:
                job.e_old = e
                keep.append(job)
        SD.update_many([j.scf for j in keep])
        self.active = keep
        self.ticks += 1

    def run(self, timeout: float = 600.0):
        while (self.waiting or self.active) and self.now() < timeout:
            if not self.active and self.waiting:
                nxt = min(j.arrival for j in self.waiting)
                if nxt > self.now():
                    time.sleep(min(nxt - self.now(), 0.05))
            self.tick()
        return self.done

    def report(self):
        lat = np.array([j.finished - j.arrival for j in self.done])
        span = max(j.finished for j in self.done) - min(j.arrival for j in self.done) if self.done else 0.0
        return {"jobs": len(self.done), "converged": int(sum(j.converged for j in self.done)),
                "jobs_per_s": len(self.done) / span if span > 0 else None,
                "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
                "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
                "mean_iterations": float(np.mean([j.iterations for j in self.done])) if self.done else None,
                "ticks": self.ticks,
                "fock_build_ms_per_tick_mean": float(np.mean(self.build_ms)) if self.build_ms else None}
