#!/usr/bin/env python3
"""Benchmark of the direct two-particle Fock build on one MI355X.

Workload: all-trans C20H42 / cc-pVDZ (510 Cartesian functions, 246 shells),
closed-shell electrons (exchange factor -0.5), the reference's screening
(overlap pair list at 1e-12, Schwarz x density at DBL_EPSILON).  One step is
one complete Fock build G = J[D] - 0.5 K[D] from a device-resident density to
a device-resident G, through lowdin_fock_compute_2body_device.  The density
is synthetic: 2 C C^T of 81 random orbitals (seed 17).

Printed: ONE JSON line with Fock builds/s (value), ERIs/s, the roofline of
the dominant class kernel (algorithmic FP64 flops / its HIP-event time) and
the CPU baseline (the oracle, a bounded sample, on this host's cores).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Fock builds/s + ERIs/s, FP64 vs mixed (|dE|<1e-6 Eh); served SCF p50/p99"
FP64_PEAK_TFLOPS = 78.6


def synthetic_density(n, nocc, seed=17):
    rng = np.random.default_rng(seed)
    C = rng.normal(size=(n, nocc)) / np.sqrt(n)
    return 2.0 * C @ C.T


def cpu_baseline(sp, D, stride, eris_full):
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    b = O.Basis(sp)
    pairs, _ = O.shellpair_list(b)
    Q = O.schwarz(b)
    t0 = time.perf_counter()
    _, cnt = O.compute_2body_direct(b, D, -0.5, pairs=pairs, Q=Q, prim_eps=1e-30, nthreads=threads,
                                    return_counts=True, s1_stride=stride)
    dt = time.perf_counter() - t0
    eris_per_s = cnt["eris"] / dt
    return {"value": eris_per_s / eris_full, "unit": "Fock builds/s", "cores": threads, "kind": "port",
            "eris_per_s": eris_per_s,
            "sample": "oracle (McMurchie-Davidson restatement of compute_2body_direct, OpenMP round-robin "
                      "over unique quartets) on every %d-th bra shell s1: %d of %d ERIs in %.1f s, scaled "
                      "by ERI count" % (stride, cnt["eris"], eris_full, dt)}


def kernel_name(code, target, tuned):
    """Kernel template that ran class `code` (e.g. "2121") under the
    autotuner choice `target` (lowdin_fock_targets encoding).  Untuned plans
    run the wave-cooperative kernel from total L = 6 and LDS strips below."""
    args = ", ".join(code)
    if target >= 1000000 or (not tuned and sum(int(c) for c in code) >= 6):
        return "jk_coop_kernel<%s>" % args
    if tuned and target == 0:
        return "jk_kernel<%s>" % args
    return "jk_strip_kernel<%s>" % args


def time_device_builds(h, dD, dG, factor, stream, steps, H):
    """Seconds per device-resident build, K builds back to back."""
    h.compute_2body_device(dD.ptr, dG.ptr, factor, stream.handle, synchronize=True)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        h.compute_2body_device(dD.ptr, dG.ptr, factor, stream.handle, synchronize=False)
    stream.synchronize()
    return (time.perf_counter() - t0) / steps


def mixed_sweep(system, steps, taus, H, LI, WF, FL, stream):
    """C20H42 SCF in FP64 and in mixed precision for each tau: total-energy
    error of the converged SCF, Fock (G) error and build rate at the
    converged FP64 density."""
    sp = system.species[0]
    n = sp.nbasis
    hcore, overlap = WF.one_body_matrices(system, {})
    h64 = LI.Libint2Interface(sp, 1)
    t0 = time.perf_counter()
    E64, wfs = WF.run_scf(system, hcore=hcore, overlap=overlap, objs={1: h64}, e_tol=1e-10, d_tol=1e-7)
    t_scf = time.perf_counter() - t0
    D = wfs[0].density
    G64 = h64.compute_2body_intraspecies_direct(D, sp.exchange_factor)
    dD = H.DeviceArray.from_host(np.asfortranarray(D).ravel(order="F"))
    dG = H.DeviceArray((n * n,))
    s64 = time_device_builds(h64, dD, dG, sp.exchange_factor, stream, steps, H)
    st64 = h64.stats()
    rows = []
    for tau in taus:
        hm = LI.Libint2Interface(sp, 1)
        hm.set_precision("mixed", tau)
        Em, _ = WF.run_scf(system, hcore=hcore, overlap=overlap, objs={1: hm}, e_tol=1e-10, d_tol=1e-7)
        Gm = hm.compute_2body_intraspecies_direct(D, sp.exchange_factor)
        st = hm.stats()
        sm = time_device_builds(hm, dD, dG, sp.exchange_factor, stream, steps, H)
        rows.append({"tau": tau, "builds_per_s": round(1.0 / sm, 3), "speedup_vs_fp64": round(s64 / sm, 3),
                     "dE_scf": Em - E64, "max_abs_dF": float(np.max(np.abs(Gm - G64))),
                     "dE_fixed_density": 0.5 * float(np.sum(D * (Gm - G64))),
                     "fp32_quartet_share": round(st["quartets_fp32"] / max(st["quartets"], 1), 4),
                     "fp32_eri_share": round(st["eris_fp32"] / max(st["eris"], 1), 4)})
        print("[bench] mixed tau %.0e: %.2f builds/s (x%.2f), dE %.2e, max|dF| %.2e" %
              (tau, 1.0 / sm, s64 / sm, Em - E64, rows[-1]["max_abs_dF"]), file=sys.stderr, flush=True)
    ok = [r for r in rows if abs(r["dE_scf"]) < 1e-6]
    best = max(ok, key=lambda r: r["builds_per_s"]) if ok else None
    # for other molecules (batches, served jobs) a tenfold margin: the
    # largest tau whose C20H42 error is below 1e-7
    safe = [r for r in rows if abs(r["dE_scf"]) < 1e-7]
    margin = max(safe, key=lambda r: r["tau"]) if safe else None
    return {"E_fp64": E64, "scf_seconds_fp64": round(t_scf, 3), "scf_iterations_note": "host DIIS + eigh, GPU builds",
            "fp64_builds_per_s": round(1.0 / s64, 3), "eris_per_build": st64["eris"], "sweep": rows,
            "best_tau_within_1e-6": best["tau"] if best else None,
            "best_mixed_builds_per_s": best["builds_per_s"] if best else None,
            "best_speedup_vs_fp64": best["speedup_vs_fp64"] if best else None,
            "tau_for_other_molecules": margin["tau"] if margin else None}


def batch_bench(steps, H, B, S, mixed_tau, stream):
    """The 256 water-cluster batch: one class launch per quartet class over
    all molecules, replayed as a HIP graph after autotuning; against the same
    256 builds one molecule per call."""
    from openlowdin_amd import libint2_interface as LI
    systems = S.water_cluster_batch(256)
    sps = [s.species[0] for s in systems]
    rng = np.random.default_rng(23)
    Ds = [synthetic_density(sp.nbasis, sp.n_occupied, seed=int(rng.integers(1 << 30))) for sp in sps]
    out = {"molecules": len(sps), "n_range": [min(sp.nbasis for sp in sps), max(sp.nbasis for sp in sps)],
           "density": "2 C C^T of random orbitals per molecule"}
    G64 = None
    for mode in ("fp64", "mixed"):
        fb = B.FockBatch.from_species(sps)
        if mode == "mixed":
            for h in fb.handles:
                h.set_precision("mixed", mixed_tau)
        dD = H.DeviceArray.from_host(fb.pack(Ds))
        dG = H.DeviceArray((fb.packed_size,))
        fb.compute_device(dD.ptr, dG.ptr, stream.handle)
        t0 = time.perf_counter()
        fb.autotune(dD.ptr, dG.ptr, reps=1, stream=stream.handle)
        t_tune = time.perf_counter() - t0
        fb.options(use_graph=True)
        fb.compute_device(dD.ptr, dG.ptr, stream.handle)
        st = fb.stats()
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fb.compute_device(dD.ptr, dG.ptr, stream.handle, synchronize=False)
        stream.synchronize()
        sb = (time.perf_counter() - t0) / steps
        out[mode] = {"batch_builds_per_s": round(1.0 / sb, 3), "ms_per_batch": round(sb * 1e3, 4),
                     "molecule_builds_per_s": round(len(sps) / sb, 1), "eris_per_s": round(st["eris"] / sb, 1),
                     "eris_per_batch": st["eris"], "fp32_eri_share": round(st["eris_fp32"] / max(st["eris"], 1), 4),
                     "autotune_s": round(t_tune, 2), "targets": fb.targets()}
        Gs = fb.unpack(dG.download())
        if mode == "fp64":
            G64 = Gs
        else:
            out[mode]["tau"] = mixed_tau
            # every molecule's result against its FP64 build: max |dG| and
            # the energy change at fixed density, 0.5 tr(D dG)
            dG_max = [float(np.max(np.abs(g - g0))) for g, g0 in zip(Gs, G64)]
            dE = [0.5 * float(np.sum(D * (g - g0))) for D, g, g0 in zip(Ds, Gs, G64)]
            out[mode]["max_abs_dG_vs_fp64"] = max(dG_max)
            out[mode]["max_abs_dE_fixed_density_vs_fp64"] = max(abs(x) for x in dE)
        fb.close()
    # the same molecules one build per call (device-resident, FP64)
    hs = [LI.Libint2Interface(sp, 1) for sp in sps]
    for h in hs:
        h.init_2body_ints()
    bufs = [(H.DeviceArray.from_host(np.asfortranarray(D).ravel(order="F")), H.DeviceArray((sp.nbasis ** 2,)))
            for sp, D in zip(sps, Ds)]
    for h, (d, g) in zip(hs, bufs):
        h.compute_2body_device(d.ptr, g.ptr, -0.5, stream.handle, synchronize=True)
    t0 = time.perf_counter()
    reps = max(1, steps // 4)
    for _ in range(reps):
        for h, (d, g) in zip(hs, bufs):
            h.compute_2body_device(d.ptr, g.ptr, -0.5, stream.handle, synchronize=False)
    stream.synchronize()
    s1 = (time.perf_counter() - t0) / reps
    out["one_per_call_fp64"] = {"ms_per_256": round(s1 * 1e3, 4), "molecule_builds_per_s": round(len(sps) / s1, 1)}
    out["fp64_batch_speedup_vs_one_per_call"] = round(s1 / (out["fp64"]["ms_per_batch"] * 1e-3), 3)
    return out


def batched_scf_bench(S, H, mixed_tau):
    """The 256 water clusters iterated to convergence together in one
    batched device-resident SCF (lowdin_scf_batch: every step of an
    iteration one batched operation over all molecules, DIIS on the device,
    HIP-graph replay while the set of molecules is unchanged); converged
    molecules leave.  Each engine's batched build is autotuned before its
    timed region.  Against one (autotuned) batched Fock build of all 256."""
    from openlowdin_amd import batch as B
    from openlowdin_amd import scf_batch as SB
    systems = S.water_cluster_batch(256)
    out = {"molecules": len(systems)}
    energies = {}
    for mode, tau in (("fp64", None), ("mixed", mixed_tau)):
        eng = SB.BatchSCF(use_graph=True)
        jobs = [eng.add(s, mixed_tau=tau) for s in systems]
        # kernel shapes of the batched build tuned on the guess densities,
        # before the timed region (as the main line's autotuner)
        t_tune = time.perf_counter()
        eng.autotune()
        t_tune = time.perf_counter() - t_tune
        live = dict(enumerate(jobs))
        prev, its, conv = {}, {}, 0
        energies[mode] = {}
        step_ms = []
        t0 = time.perf_counter()
        while live:
            ts = time.perf_counter()
            eng.step()
            step_ms.append((time.perf_counter() - ts) * 1e3)
            for idx, j in list(live.items()):
                code, e, err, it, _ = eng.status(j)
                ok = code == 0 and idx in prev and abs(e - prev[idx]) < 1e-9 and err < 1e-7
                prev[idx] = e
                if code != 0 or ok or it >= 200:
                    conv += int(ok)
                    its[idx] = it
                    energies[mode][idx] = e
                    eng.remove(j)
                    del live[idx]
        dt = time.perf_counter() - t0
        eng.close()
        full = step_ms[2:6] if len(step_ms) > 6 else step_ms
        out[mode] = {"seconds_to_converge_all": round(dt, 3), "steps": len(step_ms), "converged": conv,
                     "autotune_s": round(t_tune, 2),
                     "mean_iterations": round(float(np.mean(list(its.values()))), 2),
                     "molecule_scf_per_s": round(len(systems) / dt, 2),
                     "step_ms_all_256_active": round(float(np.median(full)), 3)}
        if tau is not None:
            out[mode]["tau"] = tau
            dE = [energies[mode][k] - energies["fp64"][k] for k in energies[mode]]
            out[mode]["max_abs_dE_vs_fp64"] = max(abs(x) for x in dE)
            out[mode]["mean_abs_dE_vs_fp64"] = float(np.mean(np.abs(dE)))
        print("[bench] batched SCF %s: 256 molecules in %.2f s (%d steps, %.1f ms per step with all active)" % (
            mode, dt, len(step_ms), out[mode]["step_ms_all_256_active"]), file=sys.stderr, flush=True)
    # one batched build of the same 256 molecules (graph replay) for comparison
    sps = [s.species[0] for s in systems]
    fb = B.FockBatch.from_species(sps)
    Ds = [synthetic_density(sp.nbasis, sp.n_occupied, seed=3 + k) for k, sp in enumerate(sps)]
    dD = H.DeviceArray.from_host(fb.pack(Ds))
    dG = H.DeviceArray((fb.packed_size,))
    stream = H.Stream()
    fb.compute_device(dD.ptr, dG.ptr, stream.handle)
    fb.autotune(dD.ptr, dG.ptr, reps=1, stream=stream.handle)
    fb.options(use_graph=True)
    fb.compute_device(dD.ptr, dG.ptr, stream.handle)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        fb.compute_device(dD.ptr, dG.ptr, stream.handle, synchronize=False)
    stream.synchronize()
    build_ms = (time.perf_counter() - t0) / 5 * 1e3
    fb.close()
    out["batched_build_ms_256"] = round(build_ms, 3)
    out["fp64_step_over_build"] = round(out["fp64"]["step_ms_all_256_active"] / build_ms, 3)
    return out


def serving_bench(S, SV, njobs, rel_rates, mixed_tau, seed=20261202):
    """Served SCF: Poisson arrivals of mixed jobs ((H2O)_1..24 / C2..C20
    alkanes, 6-31G** or cc-pVDZ, 20% with a quantum proton), continuous
    batching in size lanes under a 200 GB HBM budget.  The saturation rate is
    measured with every job present at t = 0; arrival rates are fractions of
    it.  Serial FP64 solves (one job at a time, one lane) on the arrivals of
    the first rate are the baseline."""
    rng = np.random.default_rng(seed)
    jobs = [S.serving_job(rng) for _ in range(njobs)]

    def run(arrivals, energies=None, **kw):
        eng = SV.ScfEngine(hbm_budget=200e9, **kw)
        for k, ((label, sysm), t) in enumerate(zip(jobs, arrivals)):
            eng.submit(SV.Job(k, label, sysm, float(t)))
        eng.run(timeout=900)
        rep = eng.report()
        if energies is not None:
            energies.update({j.jid: j.energy for j in eng.done if j.converged})
        eng.close()
        return rep

    sat = run([0.0] * njobs)
    lam_sat = sat["jobs_per_s"]
    out = {"jobs": njobs, "mix": "%d water clusters, %d alkanes, %d with H_1" % (
        sum(l.startswith("(H2O)") for l, _ in jobs), sum(l.startswith("C") for l, _ in jobs),
        sum(l.endswith("+H_1") for l, _ in jobs)),
        "saturation_fp64": sat, "rates": []}
    print("[bench] serving saturation: %.3f jobs/s" % lam_sat, file=sys.stderr, flush=True)
    arr_rng = np.random.default_rng(seed + 1)
    gaps = arr_rng.exponential(1.0, size=njobs)
    for rel in rel_rates:
        lam = rel * lam_sat
        arrivals = np.cumsum(gaps) / lam
        row = {"rate_rel_saturation": rel, "lambda_jobs_per_s": round(lam, 4)}
        e64, emx = {}, {}
        for mode, tau, en in (("fp64", None, e64), ("mixed", mixed_tau, emx)):
            row[mode] = run(arrivals, energies=en, mixed_tau=tau)
        # every job's mixed-precision result against its FP64 result
        dE = [emx[k] - e64[k] for k in emx if k in e64]
        if dE:
            row["mixed"]["max_abs_dE_vs_fp64"] = max(abs(x) for x in dE)
            row["mixed"]["jobs_compared"] = len(dE)
        out["rates"].append(row)
        print("[bench] serving at %.2fx: fp64 p50 %.2f s p99 %.2f s; mixed p50 %.2f s p99 %.2f s" % (
            rel, row["fp64"]["p50_s"], row["fp64"]["p99_s"], row["mixed"]["p50_s"], row["mixed"]["p99_s"]),
            file=sys.stderr, flush=True)
    # serial FP64 solves (one job at a time, one lane) on the same arrivals
    # as the first served rate
    rel0 = rel_rates[0]
    arrivals = np.cumsum(gaps) / (rel0 * lam_sat)
    ser = run(arrivals, max_jobs=1, lanes=1)
    out["serial_fp64"] = dict(ser, rate_rel_saturation=rel0)
    one = out["rates"][0]
    if ser.get("p99_s"):
        out["served_over_serial_p50"] = round(one["fp64"]["p50_s"] / ser["p50_s"], 3)
        out["served_over_serial_p99"] = round(one["fp64"]["p99_s"] / ser["p99_s"], 3)
        out["served_over_serial_jobs_per_s"] = round(one["fp64"]["jobs_per_s"] / ser["jobs_per_s"], 3)
    print("[bench] serial at %.2fx: p50 %.2f s p99 %.2f s" % (rel0, ser["p50_s"], ser["p99_s"]),
          file=sys.stderr, flush=True)
    if mixed_tau is not None:
        out["mixed_tau"] = mixed_tau
    return out


def small_configs_bench(H, LI, S, WF, steps, stream):
    """The reference's CPU-runnable cases on the GPU (device buffers):
      C0 water 6-31G** (n=25): build rate, and the first-iteration Fock
         matrix F = H + G[D_hcore] against the same through the host entry;
      C1 [F-H-F]- 6-311++G** electrons (n=46) + H_1 spd (n=40): the
         electron and proton builds and the coupling in both directions
         (one fused pass), per SCF iteration."""
    out = {}
    w = S.water("6-31G**")
    sp = w.species[0]
    n = sp.nbasis
    hcore, overlap = WF.one_body_matrices(w, {})
    X = WF._sym_orth(overlap[0])
    ev, Cp = np.linalg.eigh(X.T @ hcore[0] @ X)
    Cm = X @ Cp
    D0 = 2.0 * Cm[:, :sp.n_occupied] @ Cm[:, :sp.n_occupied].T
    h = LI.Libint2Interface(sp, 1)
    h.init_2body_ints()
    dD = H.DeviceArray.from_host(np.asfortranarray(D0).ravel(order="F"))
    dG = H.DeviceArray((n * n,))
    s0 = time_device_builds(h, dD, dG, -0.5, stream, max(steps, 20), H)
    G1 = dG.download().reshape(n, n, order="F")
    F1 = hcore[0] + G1
    Gh = h.compute_2body_intraspecies_direct(D0, -0.5)
    from openlowdin_amd import scf_device as SD
    t0 = time.perf_counter()
    scf = SD.DeviceSCF(S.water("6-31G**"))
    E0 = scf.run(e_tol=1e-10, d_tol=1e-8)
    t_scf = time.perf_counter() - t0
    out["C0_water"] = {"n": n, "builds_per_s": round(1.0 / s0, 1), "ms_per_build": round(s0 * 1e3, 4),
                       "eris_per_build": h.stats()["eris"],
                       "first_iteration_fock": {"trace": float(np.trace(F1)), "max_abs": float(np.max(np.abs(F1))),
                                                "max_abs_diff_vs_host_entry": float(np.max(np.abs(G1 - Gh)))},
                       "scf": {"energy": E0, "iterations": scf.iterations, "converged": bool(scf.converged),
                               "seconds": round(t_scf, 3), "note": "device-resident SCF, H-core guess, DIIS, "
                                                                   "|dE| < 1e-10 and density change < 1e-8"}}
    f = S.fhf_anion("6-311++G**")
    e, pr = f.species
    he, hp = LI.Libint2Interface(e, 1), LI.Libint2Interface(pr, 2)
    he.init_2body_ints()
    hp.init_2body_ints()
    De = synthetic_density(e.nbasis, e.n_occupied)
    Dp = synthetic_density(pr.nbasis, 1, seed=4) * 0.5
    dDe = H.DeviceArray.from_host(np.ascontiguousarray(De))
    dDp = H.DeviceArray.from_host(np.ascontiguousarray(Dp))
    dGe, dGp = H.DeviceArray((e.nbasis ** 2,)), H.DeviceArray((pr.nbasis ** 2,))
    dRe, dRp = H.DeviceArray((e.nbasis ** 2,)), H.DeviceArray((pr.nbasis ** 2,))
    se = time_device_builds(he, dDe, dGe, -0.5, stream, max(steps, 20), H)
    sp_ = time_device_builds(hp, dDp, dGp, -1.0, stream, max(steps, 20), H)
    he.coupling_pair_device(hp, dDe.ptr, dDp.ptr, dRe.ptr, dRp.ptr, stream.handle)
    reps = max(steps, 20)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        he.coupling_pair_device(hp, dDe.ptr, dDp.ptr, dRe.ptr, dRp.ptr, stream.handle, synchronize=False)
    stream.synchronize()
    sc = (time.perf_counter() - t0) / reps
    out["C1_fhf"] = {"n_e": e.nbasis, "n_p": pr.nbasis, "electron_builds_per_s": round(1.0 / se, 1),
                     "proton_builds_per_s": round(1.0 / sp_, 1), "coupling_pairs_per_s": round(1.0 / sc, 1),
                     "ms_per_iteration_builds": round((se + sp_ + sc) * 1e3, 4)}
    return out


def coupling_bench(H, LI, S, steps, stream):
    """The Coulomb coupling build (compute_coupling_direct,
    Libint2Iface.cpp:1111-1294; every pair of both species, no screening):
    C20H42 cc-pVDZ electrons with one quantum proton (H_1 spd basis of the
    [F-H-F]- case) at the first terminal hydrogen, both directions in one
    fused pass, device buffers."""
    atoms = S.alkane_geometry(20)
    hidx = next(k for k, (el, _) in enumerate(atoms) if el == "H")
    sysm = S.with_quantum_proton(atoms, "cc-pVDZ", hidx)
    e, pr = sysm.species
    he, hp = LI.Libint2Interface(e, 1), LI.Libint2Interface(pr, 2)
    De = synthetic_density(e.nbasis, e.n_occupied)
    Dp = synthetic_density(pr.nbasis, 1, seed=4) * 0.5
    dDe = H.DeviceArray.from_host(np.ascontiguousarray(De))
    dDp = H.DeviceArray.from_host(np.ascontiguousarray(Dp))
    dRe, dRp = H.DeviceArray((e.nbasis ** 2,)), H.DeviceArray((pr.nbasis ** 2,))
    he.coupling_pair_device(hp, dDe.ptr, dDp.ptr, dRe.ptr, dRp.ptr, stream.handle)
    st = he.stats()
    reps = max(3, steps // 2)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        he.coupling_pair_device(hp, dDe.ptr, dDp.ptr, dRe.ptr, dRp.ptr, stream.handle, synchronize=False)
    stream.synchronize()
    sc = (time.perf_counter() - t0) / reps
    ns_e, ns_p = len(e.shells), len(pr.shells)
    quartets = (ns_e * (ns_e + 1) // 2) * (ns_p * (ns_p + 1) // 2)
    return {"n_e": e.nbasis, "n_p": pr.nbasis, "ms_per_coupling_pair": round(sc * 1e3, 4),
            "coupling_pairs_per_s": round(1.0 / sc, 2), "shell_quartets": quartets,
            "eris": st.get("eris"), "eris_per_s": round(st["eris"] / sc, 1) if st.get("eris") else None}


def variants_bench(h, dD, dG, stream, steps, H, G_ref):
    """The headline build with the reference's per-call setup (shell-pair
    list and Schwarz bounds recomputed before every build, as
    Libint2Interface.f90:544 and Libint2Iface.cpp:3520 do) and with the
    BASELINE wording's 1e-12 quartet screen instead of DBL_EPSILON."""
    out = {}
    h.set_setup_each_build(True)
    h.compute_2body_device(dD.ptr, dG.ptr, -0.5, stream.handle, synchronize=True)
    reps = max(3, steps // 2)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        h.init_2body_ints()
        h.compute_2body_device(dD.ptr, dG.ptr, -0.5, stream.handle, synchronize=True)
    s_ref = (time.perf_counter() - t0) / reps
    h.set_setup_each_build(False)
    G_asref = dG.download()
    out["as_reference_setup"] = {"builds_per_s": round(1.0 / s_ref, 3), "ms_per_build": round(s_ref * 1e3, 3),
                                 "max_abs_dG_vs_amortized": float(np.max(np.abs(G_asref - G_ref)))}
    h.set_screen(1e-12)
    h.compute_2body_device(dD.ptr, dG.ptr, -0.5, stream.handle, synchronize=True)
    st = h.stats()
    s12 = time_device_builds(h, dD, dG, -0.5, stream, steps, H)
    G12 = dG.download()
    h.set_screen(2.220446049250313e-16)
    out["screen_1e-12"] = {"builds_per_s": round(1.0 / s12, 3), "ms_per_build": round(s12 * 1e3, 3),
                           "quartets": st["quartets"], "eris": st["eris"],
                           "max_abs_dG_vs_dbl_epsilon": float(np.max(np.abs(G12 - G_ref)))}
    return out


def store_bench(H, LI, S, steps, D510, h510, eris510, cpu=True):
    """Integral store and first-quarter transform (the MEMORY / integral-
    transformation variants): device-resident calls, each synchronised.
      transform: C20H42 cc-pVDZ (n=510), the Fock build's quartets and
        screen, GG = n^3 = 1.06 GB, orbital row 0 of a random C;
      store: water hexamer 6-31G** (n=150), A = n^4 = 4.05 GB;
      CPU: the oracle's transform on the water tetramer (n=100) beside the
        GPU's on the same input."""
    out = {}
    n = h510.nbasis
    rng = np.random.default_rng(31)
    dD = H.DeviceArray.from_host(np.asfortranarray(D510).ravel(order="F"))
    dc = H.DeviceArray.from_host(rng.normal(size=n) / np.sqrt(n))
    dGG = H.DeviceArray((n ** 3,))
    h510.store_device(1, dGG.ptr, dD.ptr, dc.ptr)
    t0 = time.perf_counter()
    for _ in range(steps):
        h510.store_device(1, dGG.ptr, dD.ptr, dc.ptr)
    dt = (time.perf_counter() - t0) / steps
    st = h510.stats()
    out["transform_c20h42"] = {"ms_per_transform": round(dt * 1e3, 3), "kernel_ms": round(st["kernel_ms"], 3),
                               "eris_per_s": round(eris510 / dt, 1), "output_bytes": n ** 3 * 8}
    dGG.free()
    dD.free()
    sp = S.water_cluster(6, np.random.default_rng(2)).species[0]
    m = sp.nbasis
    h = LI.Libint2Interface(sp, 1)
    h.init_2body_ints()
    Dm = synthetic_density(m, sp.n_occupied, seed=5)
    dDm = H.DeviceArray.from_host(np.ascontiguousarray(Dm))
    dA = H.DeviceArray.from_host(np.zeros(m ** 4))
    h.store_device(0, dA.ptr, dDm.ptr)
    t0 = time.perf_counter()
    for _ in range(steps):
        h.store_device(0, dA.ptr, dDm.ptr)
    dt = (time.perf_counter() - t0) / steps
    st = h.stats()
    out["store_water6"] = {"n": m, "ms_per_store": round(dt * 1e3, 3), "kernel_ms": round(st["kernel_ms"], 3),
                           "candidate_quartets": st["quartets"], "output_bytes": m ** 4 * 8}
    dA.free()
    h.close()
    if cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as O
        sp = S.water_cluster(4, np.random.default_rng(2)).species[0]
        m = sp.nbasis
        Dm = synthetic_density(m, sp.n_occupied, seed=5)
        Cm = np.random.default_rng(6).normal(size=(m, m)) / np.sqrt(m)
        h = LI.Libint2Interface(sp, 1)
        h.init_2body_ints()
        h.compute_2body_intraspecies_direct_IT(Dm, Cm, 0)
        t0 = time.perf_counter()
        for _ in range(steps):
            G = h.compute_2body_intraspecies_direct_IT(Dm, Cm, 0)
        tg = (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        Gr = O.compute_2body_directIT(O.Basis(sp), Dm, Cm, 0)
        tc = time.perf_counter() - t0
        out["transform_water4_vs_cpu"] = {"n": m, "gpu_ms_host_buffers": round(tg * 1e3, 3),
                                          "cpu_ms": round(tc * 1e3, 1), "cores": 1, "kind": "port",
                                          "max_abs_diff": float(np.max(np.abs(G - Gr)))}
        h.close()
    return out


def g12_bench(H, LI, S, steps, cpu=True):
    """Gaussian-geminal builds with a synthetic 26-term geminal (exponents
    10^(1 - k/8), k = 0..24, plus a constant term; alternating weights):
      Fock build (factor -1) of the water tetramer 6-31G** (n=100), device
        buffers;
      coupling of the FHF- 6-311++G** electrons from the proton density;
      CPU: the oracle's build on the water dimer beside the GPU's."""
    g = [10.0 ** (1.0 - k / 8.0) for k in range(25)] + [0.0]
    c = [(1.0 if k % 2 == 0 else -0.5) / (1.0 + k) for k in range(26)]
    out = {"geminal_terms": len(g)}
    sp = S.water_cluster(4, np.random.default_rng(2)).species[0]
    n = sp.nbasis
    h = LI.Libint2Interface(sp, 1)
    h.init_2body_ints()
    D = synthetic_density(n, sp.n_occupied, seed=3)
    dD = H.DeviceArray.from_host(np.ascontiguousarray(D))
    dG = H.DeviceArray((n * n,))
    h.g12_device(dD.ptr, dG.ptr, c, g, factor=-1.0)
    t0 = time.perf_counter()
    for _ in range(steps):
        h.g12_device(dD.ptr, dG.ptr, c, g, factor=-1.0)
    dt = (time.perf_counter() - t0) / steps
    st = h.stats()
    out["fock_water4"] = {"n": n, "ms_per_build": round(dt * 1e3, 3), "kernel_ms": round(st["kernel_ms"], 3),
                          "quartets": st["quartets"], "eris_per_s": round(st["eris"] / dt, 1)}
    h.close()
    f = S.fhf_anion("6-311++G**")
    e, p = f.species
    he, hp = LI.Libint2Interface(e, 1), LI.Libint2Interface(p, 2)
    Dp = synthetic_density(p.nbasis, 1, seed=4) * 0.5
    dDp = H.DeviceArray.from_host(np.ascontiguousarray(Dp))
    dB = H.DeviceArray((e.nbasis * e.nbasis,))
    he.g12_device(dDp.ptr, dB.ptr, c, g, other=hp)
    t0 = time.perf_counter()
    for _ in range(steps):
        he.g12_device(dDp.ptr, dB.ptr, c, g, other=hp)
    dt = (time.perf_counter() - t0) / steps
    out["coupling_fhf"] = {"n_e": e.nbasis, "n_p": p.nbasis, "ms_per_build": round(dt * 1e3, 3)}
    if cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as O
        sp = S.water_cluster(2, np.random.default_rng(11)).species[0]
        m = sp.nbasis
        Dm = synthetic_density(m, sp.n_occupied, seed=5)
        h = LI.Libint2Interface(sp, 1)
        Gg = h.compute_g12_intraspecies_direct(Dm, -1.0, c, g)
        t0 = time.perf_counter()
        for _ in range(steps):
            Gg = h.compute_g12_intraspecies_direct(Dm, -1.0, c, g)
        tg = (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        Gr = O.compute_g12_direct(O.Basis(sp), Dm, -1.0, c, g, nthreads=1)
        tc = time.perf_counter() - t0
        out["fock_water2_vs_cpu"] = {"n": m, "gpu_ms_host_buffers": round(tg * 1e3, 3), "cpu_ms": round(tc * 1e3, 1),
                                     "cores": 1, "kind": "port", "max_abs_diff": float(np.max(np.abs(Gg - Gr)))}
        h.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=4)
    ap.add_argument("--no-mixed", action="store_true", help="skip the C20H42 mixed-precision SCF sweep")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-cluster batched build")
    ap.add_argument("--taus", default="1e-2,1e-3,1e-4,1e-5,1e-6,1e-7")
    ap.add_argument("--no-batch-scf", action="store_true", help="skip the 256-cluster batched SCF")
    ap.add_argument("--no-serving", action="store_true", help="skip the served-SCF runs")
    ap.add_argument("--serving-jobs", type=int, default=96)
    ap.add_argument("--serving-rates", default="0.75")
    ap.add_argument("--no-store", action="store_true", help="skip the integral store / transform runs")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default per-class kernel shapes")
    ap.add_argument("--main-tau", type=float, default=None,
                    help="run the timed C20H42 builds in mixed precision with this tau (profiling; the headline "
                         "line is FP64)")
    ap.add_argument("--no-g12", action="store_true", help="skip the Gaussian-geminal builds")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the per-call-setup and 1e-12-screen variants of the headline build")
    ap.add_argument("--no-coupling", action="store_true", help="skip the coupling-build timing")
    ap.add_argument("--no-small", action="store_true", help="skip the water / [F-H-F]- build rates")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 256 MiB device-to-device copy (profiler byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from openlowdin_amd import flops as FL
    from openlowdin_amd import hip as H
    from openlowdin_amd import libint2_interface as LI
    from openlowdin_amd import system as S

    if args.copy_calibration:
        src, dst = H.DeviceArray((1 << 25,)), H.DeviceArray((1 << 25,))   # 256 MiB each
        dst.copy_from(src)
        H.hip().hipDeviceSynchronize()
        src.free()
        dst.free()

    sp = S.alkane(20, "cc-pVDZ").species[0]
    n = sp.nbasis
    D = synthetic_density(n, sp.n_occupied)
    h = LI.Libint2Interface(sp, 1)
    h.init_2body_ints()
    if args.main_tau is not None:
        h.set_precision("mixed", args.main_tau)
    stream = H.Stream()
    dD = H.DeviceArray.from_host(np.asfortranarray(D).ravel(order="F"))
    dG = H.DeviceArray((n * n,))
    for _ in range(args.warmup):
        h.compute_2body_device(dD.ptr, dG.ptr, -0.5, stream.handle, synchronize=True)
    tuned = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        h.autotune(dD.ptr, dG.ptr, -0.5, stream.handle, reps=1)
        tuned = {"seconds": round(time.perf_counter() - t0, 2), "targets": h.targets()}
        h.compute_2body_device(dD.ptr, dG.ptr, -0.5, stream.handle, synchronize=True)
    st = h.stats()
    print("[bench] warm: %d quartets, %d ERIs, %.2f ms kernels" % (st["quartets"], st["eris"], st["kernel_ms"]),
          file=sys.stderr, flush=True)

    # timed region: K builds back to back, inputs resident in HBM
    e0, e1 = H.Event(), H.Event()
    stream.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for k in range(args.steps):
        h.compute_2body_device(dD.ptr, dG.ptr, -0.5, stream.handle, synchronize=False)
    e1.record(stream)
    e1.synchronize()
    t1 = time.perf_counter()
    h.compute_2body_device(dD.ptr, dG.ptr, -0.5, stream.handle, synchronize=True)
    st = h.stats()
    wall = t1 - t0
    ev_ms = e0.elapsed_ms(e1)
    ms_per_step = wall * 1e3 / args.steps
    builds_per_s = args.steps / wall

    # per-class timing pass: the class launches serialised on one stream, so
    # each kernel's HIP-event duration is its own
    h.time_launches(True)
    h.serialize_launches(True)
    per_class = {}
    nprof = max(3, min(args.steps, 5))
    for k in range(nprof):
        h.compute_2body_device(dD.ptr, dG.ptr, -0.5, stream.handle, synchronize=True)
        for code, ms, q, p in h.launch_profile():
            acc = per_class.setdefault(code, [0.0, q, p])
            acc[0] += ms
    h.time_launches(False)
    h.serialize_launches(False)

    # roofline of the dominant class kernel
    code, (tot_ms, q, p) = max(per_class.items(), key=lambda kv: kv[1][0])
    avg_ms = tot_ms / nprof
    fl = FL.class_flops(code, q, p)
    achieved = fl / (avg_ms * 1e-3) / 1e12
    total_flops = sum(FL.class_flops(c, v[1], v[2]) for c, v in per_class.items())
    classes = {c: {"ms": round(v[0] / nprof, 4), "quartets": v[1], "prim_quartets": v[2],
                   "tflops": round(FL.class_flops(c, v[1], v[2]) / (v[0] / nprof * 1e-3) / 1e12, 3)
                   if v[0] > 0 else None}
               for c, v in sorted(per_class.items(), key=lambda kv: -kv[1][0])}

    ids = [(a, b) for a in range(6) for b in range(a + 1)]
    pcl = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
    tmap = {}
    targets = h.targets()
    for k, (cx, cy) in enumerate(ids):
        tmap["%d%d%d%d" % (pcl[cx] + pcl[cy])] = targets[k]
    kname = kernel_name(code, tmap.get(code, 0), tuned is not None)
    # HBM traffic of the dominant kernel: measured offline with rocprofv3
    # FETCH_SIZE / WRITE_SIZE (copy-calibrated) and committed under profiles/
    traffic = None
    tpath = os.path.join(HERE, "profiles", "jk_c20h42_fp64_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tk = json.load(fh)["kernels"].get("jk_strip_kernel<%s>" % ", ".join(code))
        if tk and kname.startswith("jk_strip_kernel"):
            traffic = tk["fetch_bytes"] + tk["write_bytes"]

    out = {
        "metric": METRIC,
        "value": round(builds_per_s, 4),
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.main_tau is None else "f64+f32",
        "data": "synthetic",
        "config": {"workload": "C20H42 cc-pVDZ direct RHF Fock build (J - 0.5 K), n=510 Cartesian, "
                               "pair list 1e-12, Schwarz x density screen at DBL_EPSILON",
                   "precision": "fp64" if args.main_tau is None else "mixed, tau %g" % args.main_tau,
                   "density": "2 C C^T, 81 random orbitals (seed 17)",
                   "setup": "pair list, Schwarz bounds and task lists kept while the geometry is unchanged "
                            "(the reference recomputes them per call: variants.as_reference_setup)"},
        "eris_per_s": round(st["eris"] * builds_per_s, 1),
        "eris_per_build": st["eris"],
        "quartets_per_build": st["quartets"],
        "prim_quartets_per_build": st["prim_quartets"],
        "event_ms_per_step": round(ev_ms / args.steps, 4),
        "build_tflops_fp64": round(total_flops * builds_per_s / 1e12, 3),
        "roofline": {"bound": "mfma", "pipe": "FP64 VALU: the recursions are not matrix products; 78.6 TFLOP/s is "
                                              "MI355X's FP64 peak for both the matrix and the vector units",
                     "kernel": kname,
                     "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic, "traffic_unit": "bytes/launch",
                     "traffic_source": "profiles/jk_c20h42_fp64_traffic.md" if traffic else None, "flops_per_launch": fl,
                     "avg_launch_ms": round(avg_ms, 4),
                     "timing": "HIP events around the kernel, class launches serialised (%d builds)" % nprof},
        "autotune": tuned,
        "serialized_ms_per_build": round(sum(v[0] for v in per_class.values()) / nprof, 4),
        "classes": classes,
    }
    G_ref = dG.download()
    if not args.no_variants:
        print("[bench] per-call setup and 1e-12 screen variants...", file=sys.stderr, flush=True)
        out["variants"] = variants_bench(h, dD, dG, stream, args.steps, H, G_ref)
    if not args.no_coupling:
        print("[bench] coupling build...", file=sys.stderr, flush=True)
        out["coupling"] = coupling_bench(H, LI, S, args.steps, stream)
    if not args.no_small:
        from openlowdin_amd import wavefunction as WF
        print("[bench] water / [F-H-F]- build rates...", file=sys.stderr, flush=True)
        out["small_configs"] = small_configs_bench(H, LI, S, WF, args.steps, stream)
    if not args.no_mixed:
        from openlowdin_amd import wavefunction as WF
        print("[bench] C20H42 FP64 vs mixed SCF sweep...", file=sys.stderr, flush=True)
        out["mixed"] = mixed_sweep(S.alkane(20, "cc-pVDZ"), max(3, args.steps // 2),
                                   [float(t) for t in args.taus.split(",")], H, LI, WF, FL, stream)
    if not args.no_batch:
        from openlowdin_amd import batch as B
        print("[bench] 256-cluster batch...", file=sys.stderr, flush=True)
        tau_b = (out.get("mixed") or {}).get("tau_for_other_molecules") or 1e-3
        out["batch"] = batch_bench(args.steps, H, B, S, tau_b, stream)
    tau_best = (out.get("mixed") or {}).get("tau_for_other_molecules") or 1e-3
    if not args.no_batch_scf or not args.no_serving:
        from openlowdin_amd import serving as SV
        if not args.no_batch_scf:
            print("[bench] 256-cluster batched SCF...", file=sys.stderr, flush=True)
            out["batched_scf"] = batched_scf_bench(S, H, tau_best)
        if not args.no_serving:
            print("[bench] served SCF...", file=sys.stderr, flush=True)
            out["serving"] = serving_bench(S, SV, args.serving_jobs,
                                           [float(r) for r in args.serving_rates.split(",")], tau_best)
    if not args.no_store:
        print("[bench] integral store / first-quarter transform...", file=sys.stderr, flush=True)
        out["store"] = store_bench(H, LI, S, max(2, min(args.steps, 5)), D, h, st["eris"],
                                   cpu=not args.no_cpu_baseline)
    if not args.no_g12:
        print("[bench] G12 builds...", file=sys.stderr, flush=True)
        out["g12"] = g12_bench(H, LI, S, max(2, min(args.steps, 5)), cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle sample)...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sp, D, args.cpu_stride, st["eris"])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
