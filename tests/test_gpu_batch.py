"""Batched Fock builds (one class launch over many molecules) through the C
ABI: every molecule's G must equal its own single-species build and the
oracle; HIP-graph replays must equal eager builds; the autotuner must not
change results.  Bars: 1e-10 Eh against the oracle, 1e-12 between GPU
paths (atomic summation order only)."""
import os

import numpy as np
import pytest

import oracle as O
from openlowdin_amd import system as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "water_clusters_256.json")


def _case(msg):
    print("\n[case] " + msg, flush=True)


def _density(n, nocc, rng, eta=2.0):
    C = rng.normal(size=(n, nocc)) / np.sqrt(n)
    return eta * C @ C.T


def _mixed_species():
    """Water clusters, the FHF- electrons and the FHF- proton (no screening,
    factor -1): molecules of different n and screening in one batch."""
    rng = np.random.default_rng(3)
    sps = [S.water_cluster(k, rng).species[0] for k in (1, 3, 2, 1)]
    fhf = S.fhf_anion()
    sps += [fhf.species[0], fhf.species[1], S.water("6-31G**").species[0]]
    return sps


def test_batch_equals_single_builds_and_oracle():
    from openlowdin_amd import batch as B
    from openlowdin_amd import libint2_interface as LI
    sps = _mixed_species()
    rng = np.random.default_rng(5)
    Ds = [_density(sp.nbasis, max(1, sp.n_occupied), rng, sp.eta) for sp in sps]
    _case("batch of %d species, n = %s" % (len(sps), [sp.nbasis for sp in sps]))
    fb = B.FockBatch.from_species(sps)
    Gs = fb.compute(Ds)
    st = fb.stats()
    q_single = 0
    for sp, D, G, f in zip(sps, Ds, Gs, fb.factors):
        h = LI.Libint2Interface(sp, 1)
        Gi = h.compute_2body_intraspecies_direct(D, f)
        q_single += h.stats()["quartets"]
        assert np.max(np.abs(G - Gi)) < 1e-12
        assert np.max(np.abs(G - G.T)) < 1e-13
    assert st["quartets"] == q_single
    for k in (0, 4, 5):
        b = O.Basis(sps[k])
        Gr = O.compute_2body_direct(b, Ds[k], fb.factors[k])
        assert np.max(np.abs(Gs[k] - Gr)) < 1e-10


def test_batch_graph_replays_match_eager():
    from openlowdin_amd import batch as B
    from openlowdin_amd import hip as H
    sps = _mixed_species()
    rng = np.random.default_rng(7)
    Ds = [_density(sp.nbasis, max(1, sp.n_occupied), rng, sp.eta) for sp in sps]
    _case("batch graph replays, %d species" % len(sps))
    fb = B.FockBatch.from_species(sps)
    dD = H.DeviceArray.from_host(fb.pack(Ds))
    dG = H.DeviceArray((fb.packed_size,))
    fb.compute_device(dD.ptr, dG.ptr)
    eager = dG.download()
    fb.options(use_graph=True)
    for r in range(4):
        dG.upload(np.full(fb.packed_size, 7.0))          # eager work between replays
        fb.compute_device(dD.ptr, dG.ptr)
        got = dG.download()
        assert np.max(np.abs(got - eager)) < 1e-12, "replay %d" % r
    # new densities in the same buffer: the captured graph follows them
    Ds2 = [_density(sp.nbasis, max(1, sp.n_occupied), rng, sp.eta) for sp in sps]
    dD.upload(fb.pack(Ds2))
    fb.compute_device(dD.ptr, dG.ptr)
    g_graph = dG.download()
    fb.options(use_graph=False)
    fb.compute_device(dD.ptr, dG.ptr)
    assert np.max(np.abs(dG.download() - g_graph)) < 1e-12
    # other factors: descriptors re-uploaded, graph still valid
    fb.factors[:] = -1.0
    fb.options(use_graph=True)
    fb.compute_device(dD.ptr, dG.ptr)
    g_graph = dG.download()
    fb.options(use_graph=False)
    fb.compute_device(dD.ptr, dG.ptr)
    assert np.max(np.abs(dG.download() - g_graph)) < 1e-12


def test_batch_autotune_keeps_results():
    from openlowdin_amd import batch as B
    from openlowdin_amd import hip as H
    sps = [S.water_cluster(k, np.random.default_rng(k)).species[0] for k in (2, 4, 6)]
    rng = np.random.default_rng(9)
    Ds = [_density(sp.nbasis, sp.n_occupied, rng) for sp in sps]
    _case("batch autotune, n = %s" % [sp.nbasis for sp in sps])
    fb = B.FockBatch.from_species(sps)
    dD = H.DeviceArray.from_host(fb.pack(Ds))
    dG = H.DeviceArray((fb.packed_size,))
    fb.compute_device(dD.ptr, dG.ptr)
    before = dG.download()
    fb.autotune(dD.ptr, dG.ptr, reps=1)
    t = fb.targets()
    allowed = {0, 512, 2048, 8192, -512} | {1000000 + wg for wg in (2048, 8192)} | \
        {2000000 + wg for wg in (2048, 8192, 32768)}
    assert all(x in allowed for x in t), t
    fb.compute_device(dD.ptr, dG.ptr)
    assert np.max(np.abs(dG.download() - before)) < 1e-12


def test_water_batch_256():
    """The 256-cluster workload from its committed fixture: every cluster's
    G equals its single-species build (and the oracle for four of them)."""
    from openlowdin_amd import batch as B
    from openlowdin_amd import libint2_interface as LI
    systems = S.load_water_batch(FIXTURE)
    sps = [s.species[0] for s in systems]
    rng = np.random.default_rng(11)
    Ds = [_density(sp.nbasis, sp.n_occupied, rng) for sp in sps]
    _case("water batch: %d clusters, n %d..%d" % (len(sps), min(sp.nbasis for sp in sps),
                                                  max(sp.nbasis for sp in sps)))
    fb = B.FockBatch.from_species(sps)
    Gs = fb.compute(Ds)
    worst = 0.0
    for sp, D, G in zip(sps, Ds, Gs):
        h = LI.Libint2Interface(sp, 1)
        Gi = h.compute_2body_intraspecies_direct(D, -0.5)
        worst = max(worst, float(np.max(np.abs(G - Gi))))
        h.close()
    print("   max |G_batch - G_single| = %.2e" % worst)
    assert worst < 1e-12
    small = sorted(range(len(sps)), key=lambda k: sps[k].nbasis)[:4]
    for k in small:
        Gr = O.compute_2body_direct(O.Basis(sps[k]), Ds[k], -0.5)
        assert np.max(np.abs(Gs[k] - Gr)) < 1e-10
