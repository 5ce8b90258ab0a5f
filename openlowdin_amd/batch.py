"""Batched intra-species Fock builds: many molecules per class launch.

A :class:`FockBatch` groups species handles (one per molecule) and builds
G_m = J[D_m] + f_m K[D_m] for all of them with one launch per quartet class
(``lowdin_fock_batch_*`` in liblowdin_fock.so).  Every molecule gets exactly
what ``LibintInterface_compute_2body_direct`` (src/ints/Libint2Iface.cpp:
3510-3526) returns for it alone.  Densities and results live in packed
device buffers: molecule m's n_m x n_m block starts at ``offsets[m]``.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import hip as H
from . import libint2_interface as LI


class FockBatch:
    def __init__(self, handles: Sequence[LI.Libint2Interface], factors: Sequence[float], init: bool = True):
        """init: build each handle's shell-pair list first (not needed when
        the caller already did for the current geometry)."""
        if len(handles) != len(factors) or not handles:
            raise ValueError("one exchange factor per species handle")
        self.handles = list(handles)   # keeps the handles alive for the batch's lifetime
        if init:
            for h in self.handles:
                h.init_2body_ints()
        arr = (C.c_void_p * len(handles))(*[h.this for h in handles])
        self.this = LI.lib().lowdin_fock_batch_new(len(handles), arr)
        if not self.this:
            LI._check("batch_new")
            raise LI.LowdinFockError("batch_new failed")
        off = np.zeros(len(handles) + 1, dtype=np.int64)
        LI.lib().lowdin_fock_batch_layout(self.this, off.ctypes.data_as(C.POINTER(C.c_int64)))
        self.offsets = off
        self.nbasis = [h.nbasis for h in handles]
        self.factors = np.ascontiguousarray(factors, dtype=np.float64)

    @classmethod
    def from_species(cls, species_list, factors=None):
        hs = [LI.Libint2Interface(sp, 1) for sp in species_list]
        if factors is None:
            factors = [sp.exchange_factor for sp in species_list]
        return cls(hs, factors)

    def close(self):
        if getattr(self, "this", None):
            LI.lib().lowdin_fock_batch_del(self.this)
            self.this = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def packed_size(self) -> int:
        return int(self.offsets[-1])

    def pack(self, mats: Sequence[np.ndarray]) -> np.ndarray:
        out = np.empty(self.packed_size)
        for m, (a, n) in enumerate(zip(mats, self.nbasis)):
            out[self.offsets[m]:self.offsets[m + 1]] = np.asarray(a, dtype=np.float64).reshape(n, n).ravel(order="F")
        return out

    def unpack(self, flat: np.ndarray) -> List[np.ndarray]:
        return [flat[self.offsets[m]:self.offsets[m + 1]].reshape(n, n, order="F").copy(order="F")
                for m, n in enumerate(self.nbasis)]

    def options(self, time_launches=False, serialize=False, use_graph=False):
        LI.lib().lowdin_fock_batch_options(self.this, int(time_launches), int(serialize), int(use_graph))

    def compute_device(self, d_dens: int, d_result: int, stream: int = 0, synchronize: bool = True):
        rc = LI.lib().lowdin_fock_batch_compute_2body_device(self.this, d_dens, d_result, LI._dptr(self.factors),
                                                             stream or None, 1 if synchronize else 0)
        if rc != 0:
            LI._check("batch_compute_2body_device")
            raise LI.LowdinFockError("batch_compute_2body_device failed")

    def compute_device_ptrs(self, d_dens: Sequence[int], d_result: Sequence[int], stream: int = 0,
                            synchronize: bool = True):
        """One device pointer per molecule instead of packed buffers."""
        nm = len(self.handles)
        dd = (C.c_void_p * nm)(*d_dens)
        gg = (C.c_void_p * nm)(*d_result)
        rc = LI.lib().lowdin_fock_batch_compute_2body_device_ptrs(self.this, dd, gg, LI._dptr(self.factors),
                                                                  stream or None, 1 if synchronize else 0)
        if rc != 0:
            LI._check("batch_compute_2body_device_ptrs")
            raise LI.LowdinFockError("batch_compute_2body_device_ptrs failed")

    def compute(self, densities: Sequence[np.ndarray]) -> List[np.ndarray]:
        """Host convenience: upload, build, download (one synchronous call)."""
        dD = H.DeviceArray.from_host(self.pack(densities))
        dG = H.DeviceArray((self.packed_size,))
        self.compute_device(dD.ptr, dG.ptr)
        return self.unpack(dG.download())

    def autotune(self, d_dens: int, d_result: int, reps: int = 2, stream: int = 0):
        rc = LI.lib().lowdin_fock_batch_autotune(self.this, d_dens, d_result, LI._dptr(self.factors),
                                                 stream or None, int(reps))
        if rc != 0:
            LI._check("batch_autotune")
            raise LI.LowdinFockError("batch_autotune failed")

    def targets(self) -> List[int]:
        t = np.zeros(21, dtype=np.int32)
        LI.lib().lowdin_fock_batch_targets(self.this, t.ctypes.data_as(C.POINTER(C.c_int)))
        return [int(x) for x in t]

    def stats(self) -> dict:
        s = LI.Stats()
        LI.lib().lowdin_fock_batch_get_stats(self.this, C.byref(s))
        return s.as_dict()

    def launch_profile(self):
        cap = 64
        cls = np.zeros(cap, dtype=np.int32)
        ms = np.zeros(cap, dtype=np.float32)
        q = np.zeros(cap, dtype=np.uint64)
        pq = np.zeros(cap, dtype=np.uint64)
        n = LI.lib().lowdin_fock_batch_launch_profile(self.this, cap, cls.ctypes.data_as(C.POINTER(C.c_int)),
                                                      ms.ctypes.data_as(C.POINTER(C.c_float)),
                                                      q.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                      pq.ctypes.data_as(C.POINTER(C.c_uint64)))
        return [("%04d" % cls[k], float(ms[k]), int(q[k]), int(pq[k])) for k in range(min(n, cap))]
