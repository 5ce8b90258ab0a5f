"""Served SCF with continuous batching in size lanes: every job's converged
energy is the same alone and in the mix (1e-8 Eh), jobs that arrive while
others iterate join their running lane, lanes iterate independently,
admission never exceeds the HBM budget, and a request that is rejected (over
the budget, or refused by the library) gets its own result while every other
job's result is unchanged."""
import numpy as np
import pytest

from openlowdin_amd import system as S

pytestmark = pytest.mark.gpu


def _case(msg):
    print("\n[case] " + msg, flush=True)


def _jobs():
    rng = np.random.default_rng(77)
    out = [("(H2O)1/6-31G**", S.build_system(S.water_cluster_geometry(1, rng), "6-31G**")),
           ("(H2O)3/6-31G**", S.build_system(S.water_cluster_geometry(3, rng), "6-31G**")),
           ("C2H6/cc-pVDZ", S.build_system(S.alkane_geometry(2), "cc-pVDZ")),
           ("C3H8/6-31G**+H_1", S.with_quantum_proton(S.alkane_geometry(3), "6-31G**", 4)),
           ("(H2O)2/cc-pVDZ", S.build_system(S.water_cluster_geometry(2, rng), "cc-pVDZ")),
           ("C4H10/6-31G**", S.build_system(S.alkane_geometry(4), "6-31G**")),
           ("C6H14/6-31G**", S.build_system(S.alkane_geometry(6), "6-31G**"))]   # 160 functions: 2nd lane
    return out


def _run(jobs, arrivals, **kw):
    from openlowdin_amd import serving as SV
    eng = SV.ScfEngine(**kw)
    for k, ((label, sysm), t) in enumerate(zip(jobs, arrivals)):
        eng.submit(SV.Job(k, label, sysm, t))
    peak = 0
    while eng.waiting or any(l.jobs or l.in_flight for l in eng.lanes):
        eng.poll()
        peak = max(peak, eng.used)
        assert eng.now() < 300
    return {j.jid: j for j in eng.done}, eng, peak


def test_alone_equals_mix():
    jobs = _jobs()
    _case("serving: %d jobs alone vs in the mix" % len(jobs))
    alone = {}
    for k, (label, sysm) in enumerate(jobs):
        res, _, _ = _run([(label, sysm)], [0.0], max_jobs=1)
        alone[k] = res[0].energy
        assert res[0].converged
    mix, eng, _ = _run(jobs, [0.0, 0.0, 0.02, 0.05, 0.1, 0.1, 0.0])
    for k, (label, _) in enumerate(jobs):
        print("   %-18s alone %.10f  mix %.10f  diff %.1e  it %d" %
              (label, alone[k], mix[k].energy, mix[k].energy - alone[k], mix[k].iterations))
        assert mix[k].converged
        assert abs(mix[k].energy - alone[k]) < 1e-8
    assert eng.ticks < sum(j.iterations for j in mix.values())   # iterations overlapped
    assert sum(1 for l in eng.lanes if l.iterations) >= 2           # more than one lane ran


def test_tuned_lanes_serve_the_same_energies():
    """Lanes tuned on representative jobs before serving: the kernel shapes
    change, every job's energy does not."""
    from openlowdin_amd import serving as SV
    jobs = _jobs()
    _case("serving: %d jobs with tuned lanes against untuned" % len(jobs))
    ref, _, _ = _run(jobs, [0.0] * len(jobs))
    eng = SV.ScfEngine()
    spent = eng.tune([s for _, s in jobs])
    for k, (label, sysm) in enumerate(jobs):
        eng.submit(SV.Job(k, label, sysm, 0.0))
    eng.run(timeout=300)
    res = {j.jid: j for j in eng.done}
    eng.close()
    print("   tuning %.2f s" % spent)
    for k, (label, _) in enumerate(jobs):
        assert res[k].converged, label
        assert abs(res[k].energy - ref[k].energy) < 1e-8, label


def test_rejected_and_refused_jobs_leave_others_unchanged():
    from openlowdin_amd import serving as SV
    jobs = _jobs()[:3]
    need = max(SV.estimate_device_bytes(s) for _, s in jobs)
    _case("serving: an over-budget request and a refused one among %d jobs" % len(jobs))
    ref, _, _ = _run(jobs, [0.0] * len(jobs), hbm_budget=4 * need)
    big = ("C20H42/cc-pVDZ", S.alkane(20, "cc-pVDZ"))
    bad = S.water("6-31G**")
    bad.species[0].n_particles = 200           # more occupied orbitals than functions
    res, eng, _ = _run(jobs + [big, ("bad", bad)], [0.0] * (len(jobs) + 2), hbm_budget=4 * need)
    assert res[len(jobs)].status == "rejected" and "budget" in res[len(jobs)].message
    assert res[len(jobs) + 1].status == "rejected"
    for k in range(len(jobs)):
        assert res[k].status == "converged"
        assert abs(res[k].energy - ref[k].energy) < 1e-10


def test_admission_respects_budget():
    from openlowdin_amd import serving as SV
    jobs = _jobs()
    need = [SV.estimate_device_bytes(s) for _, s in jobs]
    budget = max(need) + min(need)      # never room for everything at once
    _case("serving: admission under a %.1f MB budget" % (budget / 1e6))
    res, eng, peak = _run(jobs, [0.0] * len(jobs), hbm_budget=budget)
    assert peak <= budget
    assert len(res) == len(jobs) and all(j.converged for j in res.values())


def test_water_batch_scf_matches_single():
    """Eight clusters of the batched workload iterated together (one batched
    SCF step per iteration) reach the energies of their single-job SCFs."""
    import os
    from openlowdin_amd import scf_device as SD
    path = os.path.join(os.path.dirname(__file__), "golden", "water_clusters_256.json")
    systems = S.load_water_batch(path)[:8]
    _case("batched SCF: 8 water clusters, n = %s" % [s.species[0].nbasis for s in systems])
    res, eng, _ = _run([("c%d" % k, s) for k, s in enumerate(systems)], [0.0] * 8)
    for k, sysm in enumerate(systems):
        single = SD.DeviceSCF(sysm)
        e1 = single.run(e_tol=1e-9, d_tol=1e-7)
        assert res[k].converged
        assert abs(res[k].energy - e1) < 1e-8, (k, res[k].energy, e1)
