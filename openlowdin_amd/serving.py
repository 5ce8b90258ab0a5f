"""Served SCF with continuous batching.

Requests (molecular systems) arrive over time.  The engine admits each into
a lane by size while the estimated HBM footprint of the admitted jobs fits a
budget; a lane is a batched device-resident SCF (:mod:`scf_batch`) on a
stream of its own, so every iteration of every job in it is one batched
step (one class launch per quartet class for all its builds, batched
products, diagonalisation and DIIS on the device, replayed as HIP graphs
while its jobs are unchanged).  Lanes run asynchronously: a lane of small
molecules iterates at its own pace and does not wait for the builds of a
lane of large ones, and several lanes' iterations are in flight at once.
Jobs join a lane between its iterations and leave it when converged,
failed or out of iterations, freeing their memory for waiting requests.

Every request ends with a result: ``converged``, ``not_converged`` (the
iteration limit; the reference continues after its limit too,
MultiSCF.f90:808-811), ``failed`` (the library reported an error for this
job, e.g. a diagonalisation that did not converge; the other jobs go on) or
``rejected`` (larger than the memory budget, or refused at admission).

A job's result does not depend on the company it keeps: the batched build
gives every molecule what a build of it alone gives (to FP64 summation
order) and every other step is per job.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import hip as H
from . import libint2_interface as LI
from . import scf_batch as SB


def estimate_device_bytes(system) -> int:
    """Upper estimate of a job's HBM footprint: the batched SCF's matrices
    (10 n^2 + 8 n^2 Fock history + 8 n^2 error history per species) plus the
    Fock-build data (pair and primitive-pair records of every shell pair,
    task lists, build buffers)."""
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
    status: str = "waiting"            # waiting, running, converged, not_converged, failed, rejected
    message: str = ""
    bytes: int = 0
    lane: Optional[int] = None
    handle: Optional[int] = None       # job id inside the lane's engine
    e_old: Optional[float] = None
    objs: dict = field(default_factory=dict)
    options: dict = field(default_factory=dict)   # BatchSCF.add keywords (overlap_threshold, build_for_one_particle)


# lanes by total basis functions of a job: small, medium, large
LANE_LIMITS = (130, 330, 1 << 30)


class _Lane:
    def __init__(self, k: int, use_graph: bool, inline: bool):
        self.k = k
        self.stream = H.Stream()
        self.eng = SB.BatchSCF(stream=self.stream.handle, use_graph=use_graph, inline_launch=inline)
        self.jobs: Dict[int, Job] = {}
        self.in_flight = False
        self.iterations = 0


class ScfEngine:
    def __init__(self, hbm_budget: float = 200e9, max_jobs: int = 256, mixed_tau: Optional[float] = None,
                 use_graph: bool = True, e_tol: float = 1e-9, d_tol: float = 1e-7, max_iter: int = 200,
                 lanes: int = 3):
        self.budget = float(hbm_budget)
        self.max_jobs = max_jobs
        self.mixed_tau = mixed_tau
        self.e_tol, self.d_tol, self.max_iter = e_tol, d_tol, max_iter
        limits = LANE_LIMITS if lanes == 3 else (1 << 30,)
        self.limits = limits
        # the lane of the largest jobs forks its class launches over the
        # library's side streams; the others launch on their own streams
        self.lanes = [_Lane(k, use_graph, inline=(k < len(limits) - 1)) for k in range(len(limits))]
        self.waiting: List[Job] = []
        self.done: List[Job] = []
        self.used = 0
        self.ticks = 0
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    @property
    def active(self) -> List[Job]:
        return [j for lane in self.lanes for j in lane.jobs.values()]

    def _finish(self, job: Job, status: str, message: str = ""):
        job.status = status
        job.message = message
        job.converged = status == "converged"
        job.finished = self.now()
        self.done.append(job)

    def submit(self, job: Job):
        """Queue a request; one larger than the whole budget is rejected at once."""
        job.bytes = estimate_device_bytes(job.system)
        if job.bytes > self.budget:
            self._finish(job, "rejected", "needs %.1f GB, above the HBM budget" % (job.bytes / 1e9))
            return
        self.waiting.append(job)

    def _lane_of(self, job: Job) -> int:
        return self._lane_of_system(job.system)

    def _lane_of_system(self, system) -> int:
        n = sum(sp.nbasis for sp in system.species)
        return next(k for k, lim in enumerate(self.limits) if n <= lim)

    def tune(self, systems, reps: int = 1) -> float:
        """Before serving: the kernel shapes of every lane's batched build
        chosen by the autotuner on the given representative systems of the
        lane's size range (added, tuned, removed; the shapes stay with the
        lane's engine).  Restarts the engine clock.  Returns the seconds
        spent."""
        if self.waiting or self.active:
            raise RuntimeError("tune before submitting requests")
        t0 = time.perf_counter()
        for lane in self.lanes:
            mine = [s for s in systems if self._lane_of_system(s) == lane.k]
            if not mine:
                continue
            ids = [lane.eng.add(s, mixed_tau=self.mixed_tau) for s in mine]
            try:
                lane.eng.autotune(reps)
            finally:
                for j in ids:
                    lane.eng.remove(j)
        spent = time.perf_counter() - t0
        self.t0 = time.perf_counter()
        return spent

    def _admit(self, lane: _Lane):
        """First come first served into an idle lane while memory and slots allow."""
        now = self.now()
        still = []
        for job in sorted(self.waiting, key=lambda j: j.arrival):
            if job.arrival > now or self._lane_of(job) != lane.k:
                still.append(job)
                continue
            if len(self.active) < self.max_jobs and self.used + job.bytes <= self.budget:
                try:
                    job.handle = lane.eng.add(job.system, objs=job.objs, mixed_tau=self.mixed_tau, **job.options)
                except LI.LowdinFockError as exc:
                    job.admitted = now
                    self._finish(job, "rejected", str(exc))
                    continue
                job.admitted = now
                job.lane = lane.k
                job.status = "running"
                self.used += job.bytes
                lane.jobs[job.handle] = job
            else:
                still.append(job)
        self.waiting = still

    def _collect(self, lane: _Lane):
        lane.eng.step_end()
        lane.in_flight = False
        lane.iterations += 1
        self.ticks += 1
        for h, job in list(lane.jobs.items()):
            code, e, err, it, msg = lane.eng.status(h)
            job.iterations = it
            job.energy = e
            status = None
            if code == 1:
                status = "failed"
            elif job.e_old is not None and abs(e - job.e_old) < self.e_tol and err < self.d_tol:
                status = "converged"
            elif it >= self.max_iter:
                status = "not_converged"
            job.e_old = e
            if status:
                lane.eng.remove(h)
                del lane.jobs[h]
                self.used -= job.bytes
                job.handle = None
                self._finish(job, status, msg or "")

    def poll(self) -> bool:
        """Collect finished lane iterations, admit, start the next ones.
        Returns True when some lane made progress."""
        progress = False
        for lane in self.lanes:
            if lane.in_flight and lane.eng.step_done():
                self._collect(lane)
                progress = True
            if not lane.in_flight:
                self._admit(lane)
                if lane.jobs:
                    lane.eng.step_begin()
                    lane.in_flight = True
                    progress = True
        return progress

    def run(self, timeout: float = 600.0):
        while (self.waiting or any(l.jobs or l.in_flight for l in self.lanes)) and self.now() < timeout:
            if not self.poll():
                busy = any(l.in_flight for l in self.lanes)
                if not busy and self.waiting:
                    nxt = min(j.arrival for j in self.waiting)
                    time.sleep(max(0.0, min(nxt - self.now(), 0.01)))
                else:
                    time.sleep(0.0002)
        for lane in self.lanes:
            if lane.in_flight:
                self._collect(lane)
        return self.done

    def close(self):
        for lane in self.lanes:
            lane.eng.close()

    def report(self):
        fin = [j for j in self.done if j.status in ("converged", "not_converged", "failed")]
        lat = np.array([j.finished - j.arrival for j in fin])
        span = (max(j.finished for j in fin) - min(j.arrival for j in fin)) if fin else 0.0
        counts = {}
        for j in self.done:
            counts[j.status] = counts.get(j.status, 0) + 1
        return {"jobs": len(self.done), "converged": counts.get("converged", 0), "status_counts": counts,
                "jobs_per_s": len(fin) / span if span > 0 else None,
                "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
                "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
                "mean_iterations": float(np.mean([j.iterations for j in fin])) if fin else None,
                "lane_iterations": [l.iterations for l in self.lanes], "ticks": self.ticks}
