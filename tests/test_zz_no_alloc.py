"""lowdin_fock_compute_2body_device_noalloc (the build whose scratch lives in
the library's static device array) against lowdin_fock_compute_2body_device
(the same kernels over library-allocated scratch), through the C ABI.

Three calls in a row on one GPU, each on a smaller basis than the one
before (butane cc-pVDZ with d shells, the water dimer, water), so that a
call that read the previous call's leftovers in the shared scratch would
differ.  Bar: every element within 1e-12 Eh of the allocating entry point
(the two differ only in FP64 atomic summation order).  Error paths, which do
no GPU work: a basis larger than the scratch, and a species not yet built.
"""
import ctypes as C

import numpy as np
import pytest

from openlowdin_amd import system as S

pytestmark = pytest.mark.gpu


def _case(msg):
    print("\n[case] " + msg, flush=True)


def _density(n, nocc, seed):
    rng = np.random.default_rng(seed)
    Cm = rng.normal(size=(n, nocc)) / np.sqrt(n)
    return 2.0 * Cm @ Cm.T


def test_noalloc_matches_allocating_entry():
    from openlowdin_amd import hip as H
    from openlowdin_amd import libint2_interface as LI
    systems = [("butane cc-pVDZ", S.alkane(4, "cc-pVDZ").species[0]),
               ("water dimer 6-31G**", S.water_cluster(2, np.random.default_rng(11)).species[0]),
               ("water 6-31G**", S.water("6-31G**").species[0])]
    sizes = [sp.nbasis for _, sp in systems]
    assert sizes == sorted(sizes, reverse=True) and len(set(sizes)) == 3
    for k, (name, sp) in enumerate(systems):
        n = sp.nbasis
        _case("noalloc build %d: %s (n=%d)" % (k, name, n))
        h = LI.Libint2Interface(sp, 1)
        h.init_2body_ints()
        D = _density(n, max(1, sp.n_occupied), 100 + k)
        dD = H.DeviceArray.from_host(D)
        dG = H.DeviceArray((n, n))
        dG0 = H.DeviceArray((n, n))
        dG0.upload(np.full((n, n), 7.0))
        h.compute_2body_device(dD.ptr, dG.ptr, -0.5, synchronize=True)
        G = dG.download()
        rc = LI.lib().lowdin_fock_compute_2body_device_noalloc(h.this, C.c_void_p(dD.ptr), C.c_void_p(dG0.ptr),
                                                              -0.5, None)
        assert rc == 0, LI.lib().lowdin_fock_last_error()
        H.check(H.hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        G0 = dG0.download()
        err = float(np.max(np.abs(G0 - G)))
        print("   max|G_noalloc - G_device| = %.3e" % err, flush=True)
        assert np.all(np.isfinite(G0))
        assert err <= 1e-12
        assert float(np.max(np.abs(G))) > 1e-3


def test_noalloc_errors_without_gpu_work():
    from openlowdin_amd import libint2_interface as LI
    big = S.alkane(32, "cc-pVDZ").species[0]
    assert big.nbasis > 768
    _case("noalloc refusals: n=%d over the scratch; species never built" % big.nbasis)
    hb = LI.Libint2Interface(big, 1)
    rc = LI.lib().lowdin_fock_compute_2body_device_noalloc(hb.this, None, None, -0.5, None)
    assert rc != 0 and b"exceeds" in LI.lib().lowdin_fock_last_error()
    hw = LI.Libint2Interface(S.water("6-31G**").species[0], 1)
    rc = LI.lib().lowdin_fock_compute_2body_device_noalloc(hw.this, None, None, -0.5, None)
    assert rc != 0 and b"allocates nothing" in LI.lib().lowdin_fock_last_error()
