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
    threads = min(16, os.cpu_count() or 1)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=4)
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
    stream = H.Stream()
    dD = H.DeviceArray.from_host(np.asfortranarray(D).ravel(order="F"))
    dG = H.DeviceArray((n * n,))
    for _ in range(args.warmup):
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

    # HBM traffic of the dominant kernel: measured offline with rocprofv3
    # FETCH_SIZE / WRITE_SIZE (copy-calibrated) and committed under profiles/
    traffic = None
    tpath = os.path.join(HERE, "profiles", "jk_c20h42_fp64_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tk = json.load(fh)["kernels"].get("jk_strip_kernel<%s>" % ", ".join(code))
        if tk:
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
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "C20H42 cc-pVDZ direct RHF Fock build (J - 0.5 K), n=510 Cartesian, "
                               "pair list 1e-12, Schwarz x density screen at DBL_EPSILON",
                   "precision": "fp64", "density": "2 C C^T, 81 random orbitals (seed 17)"},
        "eris_per_s": round(st["eris"] * builds_per_s, 1),
        "eris_per_build": st["eris"],
        "quartets_per_build": st["quartets"],
        "prim_quartets_per_build": st["prim_quartets"],
        "event_ms_per_step": round(ev_ms / args.steps, 4),
        "build_tflops_fp64": round(total_flops * builds_per_s / 1e12, 3),
        "roofline": {"bound": "valu", "kernel": "jk_strip_kernel<%s>" % ",".join(code),
                     "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic, "traffic_unit": "bytes/launch",
                     "traffic_source": "profiles/jk_c20h42_fp64_traffic.md" if traffic else None, "flops_per_launch": fl,
                     "avg_launch_ms": round(avg_ms, 4),
                     "timing": "HIP events around the kernel, class launches serialised (%d builds)" % nprof},
        "serialized_ms_per_build": round(sum(v[0] for v in per_class.values()) / nprof, 4),
        "classes": classes,
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle sample)...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sp, D, args.cpu_stride, st["eris"])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
