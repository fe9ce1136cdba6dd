"""Throughput of the WaveQ3D hot path on one MI355X.

Workload (default): the single-fan case at the size of the range-dependent
configuration - DE -80:0.25:80 x AZ 0:2:358 = 115,380 rays, 8 log-spaced
frequencies 100-10000 Hz, 100 x 100 targets, dt 0.1 s, 1201 steps - in the
analytic environment of the CPU-runnable case (Munk profile, flat 5000 m
Rayleigh-sand bottom, Thorp).  The gridded sound speed and bathymetry of the
range-dependent case are switched off on the GPU in this build.  A "step" is one wave_queue::step() of the whole
fan: reflections, AB3 propagation, eigenray search over all targets and
eigenray construction.  W warm-up steps run untimed, then exactly K steps are
timed; the default W + K = 1201 covers the whole 120 s run.

Prints one JSON line: ray-steps/s of the timed steps (eigenrays/s beside it),
the roofline of the dominant kernel measured with HIP events on the library's
stream, and the CPU oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
W_TGT = 13.0              # FP64 ops per (target, ray) pair of the window test: 6 FMA + 1 compare
W_BOX = 59.0              # FP64 ops per (ray, box) bound: the window form bounded over the box, 2 compares
# HBM bytes per step of the search stage (k_supmask + k_detect) at this
# workload: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from a
# 1 GiB device copy), means of 310 launches of each kernel;
# profiles/k_detect_single_fan_115k_rays_10k_targets.md holds the commands
K_DETECT_TRAFFIC_BYTES = ((3733.50 + 21876.55) * 2.0 + (72.06 + 431.15)) * 1024.0
KERNELS = ["k_init", "(id 1, not launched)", "k_advance", "k_detect", "k_build", "(id 5, not launched)", "k_bounce"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(steps=3):
    """Oracle (scalar C++ port, 1 thread) on `steps` steps of the same fan,
    environment and targets."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from underseamodelinglibrary_amd import scenarios
    sc = scenarios.config2_analytic()
    q = oracle.oracle_queue(*sc.args())
    t0 = time.perf_counter()
    q.step(steps)
    dt = time.perf_counter() - t0
    n = sc.num_rays()
    return {"value": n * steps / dt, "unit": "ray-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle, steps 1-{steps} of the full fan with all {sc.num_targets()} "
                      f"targets ({dt:.1f} s)"}


def batch_leg(lib, n_sources, chunk=64):
    """The production caller's form: wavefront_generator tasks for the first
    n_sources of config 4's seeded sources (uneven fans, 4-16 frequencies,
    90 s), no targets, in config 1's analytic ocean.  Timed once one after
    another (each fan alone, as the reference's thread pool would hand them
    to one device) and once as one wavefront_batch (every fan on its own
    stream, chunks of steps queued on all fans before any is waited on).
    Fan creation and release are inside both timings."""
    from underseamodelinglibrary_amd import scenarios
    from underseamodelinglibrary_amd.usml.wavegen import (wavefront_batch, wavefront_generator,
                                                          wavefront_source)
    ocean = scenarios.config1().ocean
    specs = scenarios.config4_sources(256)[:n_sources]
    ray_steps = sum(sp["de"].size() * sp["az"].size() * sp["steps"] for sp in specs)

    def make():
        return [wavefront_generator(wavefront_source(sp["pos"], keyID=k), None, None, sp["freq"],
                                    sp["de"], sp["az"], sp["time_step"], sp["time_max"], 300.0,
                                    999, 999, ocean=ocean)
                for k, sp in enumerate(specs)]

    log(f"[bench] wavefront_batch: {n_sources} fans, {ray_steps / 1e9:.2f} G ray-steps")
    gens = make()
    t0 = time.perf_counter()
    for g in gens:
        g.run(chunk=chunk)
    alone_s = time.perf_counter() - t0
    assert all(g.done() for g in gens)
    gens = make()
    t0 = time.perf_counter()
    wavefront_batch(gens, chunk=chunk).run()
    batch_s = time.perf_counter() - t0
    assert all(g.done() for g in gens)
    log(f"[bench] wavefront_batch: alone {alone_s:.2f} s, batch {batch_s:.2f} s")
    return {"workload": f"config 4 sources 1-{n_sources} of 256 (seeded), uneven fans "
                        "DE 61-721 x AZ 1-181, 4-16 freqs, 90 s, no targets, Munk + flat "
                        "5000 m bottom (gridded ocean off on GPU)",
            "fans": n_sources, "ray_steps": ray_steps,
            "alone_ray_steps_per_s": ray_steps / alone_s,
            "batch_ray_steps_per_s": ray_steps / batch_s,
            "speedup": alone_s / batch_s}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1181)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--batch-fans", type=int, default=32,
                    help="sources of the wavefront_batch leg (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(independent fans or target subsets would shard with no exchange)", file=sys.stderr)
        sys.exit(2)

    from underseamodelinglibrary_amd import _lib, scenarios
    from underseamodelinglibrary_amd.usml.waveq3d import wave_queue
    import ctypes as C

    lib = _lib.load()
    if lib.wq3d_device_count() < 1:
        raise RuntimeError("no HIP device")
    sc = scenarios.config2_analytic()
    log(f"[bench] fan {sc.num_rays()} rays, {sc.num_targets()} targets, "
        f"{sc.freq.size()} frequencies; warmup {args.warmup}, timed {args.steps}")
    q = wave_queue(*sc.args())
    q.step(args.warmup)
    q.sync()
    q.drain_eigenrays()
    lib.check(lib.wq3d_set_timing(q._fan, 2))  # HIP events around k_detect only
    lib.check(lib.wq3d_kernel_times(q._fan, (C.c_double * 8)(), (C.c_int64 * 8)(), 8))
    tested, boxes = C.c_int64(), C.c_int64()  # pairs tested, (ray, box) bounds tested
    lib.check(lib.wq3d_search_totals(q._fan, C.byref(tested), C.byref(boxes)))  # restart the counts
    q.sync()
    t0 = time.perf_counter()
    done = 0
    chunk = 100
    while done < args.steps:
        n = min(chunk, args.steps - done)
        q.step(n)
        done += n
        log(f"[bench] {args.warmup + done} steps")
    q.sync()
    elapsed = time.perf_counter() - t0
    ms = (C.c_double * 8)()
    launches = (C.c_int64 * 8)()
    lib.check(lib.wq3d_kernel_times(q._fan, ms, launches, 8))
    lib.check(lib.wq3d_search_totals(q._fan, C.byref(tested), C.byref(boxes)))
    rays, rays_int, rays_phase = q.drain_eigenrays()
    n_rays = sc.num_rays()
    value = n_rays * args.steps / elapsed
    kernel_ms = {KERNELS[k]: ms[k] for k in range(len(KERNELS)) if launches[k]}
    # dominant kernel: the target search.  Its work is the (ray, box) bounds
    # it tests and the window test of every (target, ray) pair in a block
    # the bounds kept, both counted on the device over the timed launches
    # (the neighbour-form bounds a box test may add are not counted).
    det_ms = ms[3]
    det_launches = launches[3]
    pairs_per_launch = tested.value / det_launches if det_launches else 0.0
    boxes_per_launch = boxes.value / det_launches if det_launches else 0.0
    flops = W_TGT * pairs_per_launch + W_BOX * boxes_per_launch
    achieved = flops / (det_ms * 1e-3) / 1e12 if det_ms > 0 else 0.0
    out = {
        "metric": "ray-steps/sec + eigenrays/sec on 1 MI355X, plain vs batched+adjoint at fixed tol",
        "form": "plain per-fan kernels",
        "value": value,
        "unit": "ray-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic Munk ocean, seeded target grid)",
        "config": {"workload": "single fan 641 DE x 180 AZ rays, 8 freqs 100-10000 Hz, "
                               "100x100 targets, Munk + flat 5000 m Rayleigh-sand bottom, "
                               "dt 0.1 s, 1201 steps (gridded SSP/bathymetry off on GPU)",
                   "rays": n_rays, "targets": sc.num_targets(), "frequencies": sc.freq.size()},
        "eigenrays": int(len(rays)),
        "eigenrays_per_s": len(rays) / elapsed,
        "kernel_ms_per_launch": kernel_ms,
        "roofline": {"bound": "mfma", "kernel": "k_supmask + k_detect (the search stage)", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": K_DETECT_TRAFFIC_BYTES,
                     "pairs_tested_per_launch": pairs_per_launch,
                     "box_tests_per_launch": boxes_per_launch,
                     "pairs_all_per_launch": float(sc.num_targets() * n_rays),
                     "note": "FP64 vector (no MFMA-shaped work); achieved = (13 FP64 ops per "
                             "(target, ray) pair put to the window test + 59 per (ray, box) bound "
                             "tested), both counted on the device, / mean time per step of the "
                             "search stage's two launches (HIP events around both); traffic in "
                             "bytes per step of the stage"},
    }
    q.close()
    # the path's next caller stage: eigenray_collection's per-target totals
    # (coherent) of the timed steps' eigenrays, on the GPU, upload and
    # download included; the oracle's time for the same sum beside it
    import numpy as np
    from underseamodelinglibrary_amd.usml.eigenrays import sum_eigenray_arrays
    freq = np.array([sc.freq(k) for k in range(sc.freq.size())])
    n1, n2 = sc.targets.size1(), sc.targets.size2()
    t0 = time.perf_counter()
    st, _, _, _ = sum_eigenray_arrays(lib.wq3d_sum_eigenrays, rays, rays_int, rays_phase, freq, n1, n2,
                                      True)
    lib.check(st)
    coll = {"rays": int(len(rays)), "targets": n1 * n2, "frequencies": len(freq),
            "gpu_ms": (time.perf_counter() - t0) * 1e3}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        t0 = time.perf_counter()
        st, _, _, _ = sum_eigenray_arrays(oracle.load().orc_sum_eigenrays, rays, rays_int, rays_phase,
                                          freq, n1, n2, True)
        coll["cpu_ms"] = (time.perf_counter() - t0) * 1e3
        coll["cpu_kind"] = "port, 1 core"
    out["eigenray_collection"] = coll
    if args.batch_fans > 0:
        out["wavefront_batch"] = batch_leg(lib, args.batch_fans)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
