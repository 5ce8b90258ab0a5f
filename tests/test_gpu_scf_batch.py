"""Batched device-resident SCF (lowdin_scf_batch_*, openlowdin_amd/scf_batch.py).

* Four of the reference's test systems in ONE batch -- different basis
  sizes, one to four species, couplings between species, open and closed
  shells -- each reach the energy the reference's checker holds, to 1e-9 Eh:
  test/H2O.BOA.DIRECT.py:18, test/FHF-.NOCI-RHF.py:19, test/HCN.py:19,
  test/H2O.APMO.UHF.DIRECT.py:19-20.
* Graph replays give the eager iteration: the same batch with and without
  captured graphs, iteration by iteration (energies to 1e-11 Eh), with eager
  work between the replays.
* A job's iterations are the same alone and in a batch.
* A job the engine refuses (more occupied orbitals than functions) leaves
  the other jobs and their results unchanged.
"""
import numpy as np
import pytest

from openlowdin_amd import system as S

pytestmark = pytest.mark.gpu

H2O = [("O", (0.0, 0.0, -0.066575)), ("H", (0.0, 0.754175, 0.528381)), ("H", (0.0, -0.754174, 0.528382))]

REFERENCE = [
    ("H2O.BOA.DIRECT", lambda: S.build_system(H2O, "6-311G"), -76.008843007734),
    ("FHF-.NOCI-RHF", lambda: S.build_system(
        [("F", (0, 0, -1.140079)), ("H", (0, 0, 0)), ("F", (0, 0, 1.140079))], "6-31G",
        quantum=[("H_1", "NAKAI-3-S", (0, 0, 0), 1.0078250321 * 1822.88850065855 - 1.0)], extra_electrons=1),
     -199.396255393556),
    ("HCN", lambda: S.build_system([("H", (0, 0, 0)), ("C", (0, 0, 1.059)), ("N", (0, 0, 2.186))], "6-311++G**"),
     -92.901743215468),
    ("H2O.APMO.UHF.DIRECT", lambda: S.build_system(
        H2O, "6-311G", quantum=[("H-A_1", "NAKAI-5-SPD", (0.0, 0.754175, 0.528381), 1836.15267247),
                                ("H-B_1", "NAKAI-5-SPD", (0.0, -0.754174, 0.528382), 1836.15267247)],
        open_shell=True), -75.93125555356),
]


def _case(msg):
    print("\n[case] " + msg, flush=True)


@pytest.mark.parametrize("autotune", [False, True], ids=["default", "autotuned"])
def test_reference_energies_in_one_batch(autotune):
    """Also with the engine's batched build autotuned: the kernel shapes
    change, the energies do not."""
    from openlowdin_amd import scf_batch as SB
    _case("batched SCF: 4 reference systems in one batch (autotune %s)" % autotune)
    res = SB.run_batch([b() for _, b, _ in REFERENCE], e_tol=1e-11, d_tol=1e-9, max_iter=300, autotune=autotune)
    for (name, _, eref), (E, it, conv) in zip(REFERENCE, res):
        print("   %-22s E = %.12f  diff %.2e  (%d iterations)" % (name, E, E - eref, it), flush=True)
        assert conv, name
        assert abs(E - eref) <= 1e-9, name


def test_reference_energies_mixed_options_in_one_batch():
    """Reference systems with their own orthogonaliser thresholds, ghost
    species and one-particle builds in one batch: HCN.e+ and H3+.FCI
    (overlap threshold 1e-6), HCN.e+.FT (the positron as SCF ghost species,
    the checker's 1e-8), PsH.FCI (two-particle matrices for one particle),
    CO.CISD-dipole (UHF) -- each to its reference energy."""
    from openlowdin_amd import scf_batch as SB
    from ref_cases import CASES
    hcn = [("H", (0, 0, 0)), ("C", (0, 0, 1.059)), ("N", (0, 0, 2.186))]
    jobs = [("HCN.e+", lambda: S.build_system(hcn, "6-311++G.D.P",
                                              quantum=[("E+", "GRIBAKIN-10S", x, 1.0) for _, x in hcn],
                                              add_particles={"E+": -2}),
             -92.901811855736, {"overlap_threshold": 1e-6}, 1e-9)]
    jobs.append(("HCN.e+.FT", jobs[0][1], -92.901807583131, {"overlap_threshold": 1e-6, "ghost": "E+"}, 1e-8))
    for name in ("H3+.FCI", "PsH.FCI", "CO.CISD-dipole"):
        build, eref, kw, tol = CASES[name]
        opts = {}
        if kw.get("build_for_one"):
            opts["build_for_one_particle"] = True
        if "overlap_threshold" in kw:
            opts["overlap_threshold"] = kw["overlap_threshold"]
        jobs.append((name, build, eref, opts, tol))
    _case("batched SCF: %d reference systems with per-job options" % len(jobs))
    res = SB.run_batch([b() for _, b, _, _, _ in jobs], e_tol=1e-11, d_tol=1e-8, max_iter=300,
                       options=[o for _, _, _, o, _ in jobs])
    for (name, _, eref, _, tol), (E, it, conv) in zip(jobs, res):
        print("   %-16s E = %.12f  diff %.2e  (%d iterations)" % (name, E, E - eref, it), flush=True)
        assert conv, name
        assert abs(E - eref) <= tol, name


def _energies(use_graph, systems, steps):
    from openlowdin_amd import scf_batch as SB
    eng = SB.BatchSCF(use_graph=use_graph)
    jobs = [eng.add(s) for s in systems]
    hist = []
    for _ in range(steps):
        eng.step()
        hist.append([eng.status(j)[1] for j in jobs])
        _ = [np.sum(np.ones(1000)) for _ in range(3)]   # host work between replays
    eng.close()
    return np.array(hist)


def test_graph_replays_equal_eager():
    build = lambda: [S.water_cluster(k, np.random.default_rng(40 + k)) for k in (1, 2, 3, 2)] + \
        [REFERENCE[1][1]()]
    _case("batched SCF: graph replays vs eager, 5 jobs, 12 iterations")
    eager = _energies(False, build(), 12)
    graph = _energies(True, build(), 12)
    print("   max |dE| graph - eager per iteration: %s" % np.max(np.abs(graph - eager), axis=1), flush=True)
    assert np.all(np.isfinite(graph))
    assert np.max(np.abs(graph - eager)) < 1e-11


def test_alone_equals_in_batch_and_refusal_isolated():
    from openlowdin_amd import scf_batch as SB
    from openlowdin_amd import libint2_interface as LI
    target = lambda: S.water_cluster(3, np.random.default_rng(7))
    _case("batched SCF: one job alone vs among others; a refused job")
    alone = _energies(True, [target()], 10)[:, 0]
    eng = SB.BatchSCF()
    others = [eng.add(S.water_cluster(k, np.random.default_rng(k))) for k in (1, 2, 4)]
    j = eng.add(target())
    bad = S.water("6-31G**")
    bad.species[0].n_particles = 200     # 100 occupied orbitals in 25 functions
    with pytest.raises(LI.LowdinFockError):
        eng.add(bad)
    mix = []
    for _ in range(10):
        eng.step()
        mix.append(eng.status(j)[1])
        assert all(eng.status(o)[0] == 0 for o in others)
    eng.close()
    print("   max |E_alone - E_mix| = %.2e" % np.max(np.abs(np.array(mix) - alone)), flush=True)
    assert np.max(np.abs(np.array(mix) - alone)) < 1e-11
