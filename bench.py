"""Throughput of the WaveQ3D hot path on one MI355X.

Workload (default): config 5's forward runs - 32 seeded 64-node
piecewise-linear c(z) profiles (perturbed Munk, a 1-D profile_grid linear in
rho), range-independent, flat 4500 m Rayleigh-sand bottom; each profile one
fan of DE -45:0.1:45 x AZ 0:5:355 = 64,872 rays, 4 frequencies 500-4000 Hz,
50 x 50 targets, dt 0.05 s, t < 120 s (2401 steps).  A "step" is one
wave_queue::step() of all 32 fans: reflections, AB3 propagation, eigenray
search over the targets and eigenray construction.  W warm-up steps run
untimed, then exactly K steps are timed; the default W + K = 2401 covers the
whole run.

Two forms of the same work:
  batched  the 32 fans as one wq3d_batch: one fused launch per kernel of a
           step for all fans, the launch shape picked by the autotuner on a
           throw-away batch first (value, roofline);
  plain    each fan stepped alone with its own per-step launches, one after
           another.
Both produce the same eigenrays, bit for bit.

Prints one JSON line: ray-steps/s of the batched timed steps (eigenrays/s
beside it), the plain form's rate, the roofline of the dominant kernel (the
eigenray search stage) measured with HIP events on the batch's stream, the
CPU oracle on a bounded sample of the same workload (1 thread and all
cores), and legs on other workloads.
"""
import argparse
import ctypes as C
import json
import multiprocessing as mp
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
W_TGT = 13.0              # FP64 ops per (target, ray) pair of the window test: 6 FMA + 1 compare
W_BOX = 59.0              # FP64 ops per (ray, box) bound: the window form bounded over the box, 2 compares
# HBM bytes per step of the fused search stage (k_supmask_multi +
# k_detect_multi) on this workload: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE
# x 1.0 (the factors a 1 GiB device copy gives), means over the 4874 launches
# of a run; profiles/search_stage_config5_32_profiles.md holds the commands
SEARCH_TRAFFIC_BYTES = ((65107.0 + 75768.0) * 2.0 + (1249.0 + 2334.0)) * 1024.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    model = platform.processor() or "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return model, os.cpu_count() or 1


def _oracle_profile(args):
    """One process of the CPU baseline: the oracle on `steps` steps of
    config 5's profile `p` (all 2500 targets); returns (ray-steps, seconds)."""
    p, steps = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from underseamodelinglibrary_amd import scenarios
    oracle.load()
    sc = scenarios.config5(profile=p)
    q = oracle.oracle_queue(*sc.args())
    t0 = time.perf_counter()
    q._lib.orc_step(q._h, steps)
    return sc.num_rays() * steps, time.perf_counter() - t0


def cpu_baseline(steps, procs):
    """The oracle (scalar C++ restatement) on the GPU box's host cores: one
    thread on profile 0, then `procs` processes on profiles 0..procs-1 at once
    (independent fans in parallel, as config 5's profiles are)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build()
    model, ncpu = host_cpu()
    rs, dt = _oracle_profile((0, steps))
    one = rs / dt
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        res = pool.map(_oracle_profile, [(p, steps) for p in range(procs)])
    wall = time.perf_counter() - t0
    allc = sum(r[0] for r in res) / wall
    return {"value": allc, "unit": "ray-steps/s", "cores": procs, "kind": "port",
            "one_thread": one, "host_cpu": model, "host_cpus_seen": ncpu,
            "sample": f"oracle, steps 1-{steps} of config 5 fans with all 2500 targets: 1 thread on profile 0 "
                      f"({dt:.1f} s), then {procs} processes on profiles 0-{procs - 1} at once ({wall:.1f} s wall)"}


def make_fans(n, lib):
    from underseamodelinglibrary_amd import scenarios
    from underseamodelinglibrary_amd.usml.waveq3d import wave_queue
    qs = []
    for p in range(n):
        sc = scenarios.config5(profile=p)
        qs.append(wave_queue(*sc.args(), lib=lib))
    return qs, sc


def close_all(qs):
    for q in qs:
        q.close()


def run_plain(lib, n, warmup, steps):
    qs, sc = make_fans(n, lib)
    for q in qs:
        lib.check(lib.wq3d_step(q._fan, warmup))
        q.sync()
        q.drain_eigenrays()
    t0 = time.perf_counter()
    for q in qs:
        lib.check(lib.wq3d_step(q._fan, steps))
        q.sync()
    dt = time.perf_counter() - t0
    rays = sum(len(q.drain_eigenrays()[0]) for q in qs)
    close_all(qs)
    return sc, dt, rays


def run_batched(lib, n, warmup, steps, shape, timing):
    from underseamodelinglibrary_amd.usml.waveq3d import fan_batch
    qs, sc = make_fans(n, lib)
    b = fan_batch(qs, lib=lib)
    b.set_shape(**shape)
    lib.check(lib.wq3d_batch_step(b._h, warmup))
    for q in qs:
        q.drain_eigenrays()
        lib.check(lib.wq3d_search_totals(q._fan, C.byref(C.c_int64()), C.byref(C.c_int64())))
    if timing:
        lib.check(lib.wq3d_batch_set_timing(b._h, 1))
        lib.check(lib.wq3d_batch_stage_times(b._h, C.byref(C.c_double()), C.byref(C.c_int64())))
    t0 = time.perf_counter()
    lib.check(lib.wq3d_batch_step(b._h, steps))
    dt = time.perf_counter() - t0
    stage_ms, stage_n = C.c_double(), C.c_int64()
    if timing:
        lib.check(lib.wq3d_batch_stage_times(b._h, C.byref(stage_ms), C.byref(stage_n)))
    pairs = boxes = 0
    for q in qs:
        t, bx = C.c_int64(), C.c_int64()
        lib.check(lib.wq3d_search_totals(q._fan, C.byref(t), C.byref(bx)))
        pairs += t.value
        boxes += bx.value
    rays = sum(len(q.drain_eigenrays()[0]) for q in qs)
    b.close()
    close_all(qs)
    return sc, dt, rays, stage_ms.value, stage_n.value, pairs, boxes


def tune(lib, n, trial=8):
    """The autotuner on a throw-away batch of the same fans: returns the
    fastest shape (the steps it times are real, the batch is discarded)."""
    from underseamodelinglibrary_amd.usml.waveq3d import fan_batch
    qs, _ = make_fans(n, lib)
    b = fan_batch(qs, lib=lib)
    ms, taken = C.c_double(), C.c_int64()
    lib.check(lib.wq3d_batch_tune(b._h, trial, C.byref(ms), C.byref(taken)))
    shape = b.shape()
    b.close()
    close_all(qs)
    return shape, ms.value


def leg_config1(lib):
    """Config 1 at full size (the CPU-runnable case): one fan, plain
    per-step launches, the whole 600-step run timed after fan creation."""
    from underseamodelinglibrary_amd import scenarios
    from underseamodelinglibrary_amd.usml.waveq3d import wave_queue
    sc = scenarios.config1()
    q = wave_queue(*sc.args(), lib=lib)
    t0 = time.perf_counter()
    lib.check(lib.wq3d_step(q._fan, sc.steps))
    q.sync()
    dt = time.perf_counter() - t0
    n = len(q.drain_eigenrays()[0])
    q.close()
    return {"workload": "config 1: Munk, flat 5000 m Rayleigh-sand bottom, 121 DE x 36 AZ, 1000 Hz, 10x10 "
                        "targets, dt 0.1 s, 600 steps", "ray_steps_per_s": sc.num_rays() * sc.steps / dt,
            "eigenrays": n, "eigenrays_per_s": n / dt, "ms_per_step": dt * 1e3 / sc.steps}


def leg_config2_fan(lib):
    """Config 2's fan, frequencies and targets in config 1's analytic ocean
    (config 2's gridded ocean is not run on the GPU): one fan, plain."""
    from underseamodelinglibrary_amd import scenarios
    from underseamodelinglibrary_amd.usml.waveq3d import wave_queue
    sc = scenarios.config2_analytic()
    q = wave_queue(*sc.args(), lib=lib)
    lib.check(lib.wq3d_step(q._fan, 20))
    q.sync()
    q.drain_eigenrays()
    pairs, boxes = C.c_int64(), C.c_int64()
    lib.check(lib.wq3d_search_totals(q._fan, C.byref(pairs), C.byref(boxes)))  # from here on
    steps = sc.steps - 20
    t0 = time.perf_counter()
    lib.check(lib.wq3d_step(q._fan, steps))
    q.sync()
    dt = time.perf_counter() - t0
    lib.check(lib.wq3d_search_totals(q._fan, C.byref(pairs), C.byref(boxes)))
    recs, inten, phase = q.drain_eigenrays()
    q.close()
    return ({"workload": "config 2's fan in config 1's Munk ocean (off-config: the gridded ocean is not run on "
                         "the GPU): 641 DE x 180 AZ, 8 freqs 100-10000 Hz, 100x100 targets, steps 21-1201",
             "ray_steps_per_s": sc.num_rays() * steps / dt, "eigenrays": int(len(recs)),
             "eigenrays_per_s": len(recs) / dt, "ms_per_step": dt * 1e3 / steps,
             "pairs_tested_per_step": pairs.value / steps, "box_tests_per_step": boxes.value / steps},
            sc, recs, inten, phase)


def leg_collection(lib, sc, recs, inten, phase, cpu):
    """eigenray_collection's per-target coherent sums of a run's eigenrays on
    the GPU (upload and download included), the oracle's time beside it."""
    import numpy as np
    from underseamodelinglibrary_amd.usml.eigenrays import sum_eigenray_arrays
    freq = np.array([sc.freq(k) for k in range(sc.freq.size())])
    n1, n2 = sc.targets.size1(), sc.targets.size2()
    t0 = time.perf_counter()
    st, _, _, _ = sum_eigenray_arrays(lib.wq3d_sum_eigenrays, recs, inten, phase, freq, n1, n2, True)
    lib.check(st)
    out = {"rays": int(len(recs)), "targets": n1 * n2, "frequencies": len(freq),
           "gpu_ms": (time.perf_counter() - t0) * 1e3}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        t0 = time.perf_counter()
        st, _, _, _ = sum_eigenray_arrays(oracle.load().orc_sum_eigenrays, recs, inten, phase, freq, n1, n2, True)
        out["cpu_ms"] = (time.perf_counter() - t0) * 1e3
        out["cpu_kind"] = "port, 1 core"
    return out


def leg_sources(lib, n_sources):
    """Config 4's seeded sources (uneven fans DE 61-721 x AZ 1-181, 4-16
    frequencies, 90 s, no targets) in config 1's analytic ocean (config 4's
    gridded ocean is not run on the GPU): the fans one after another, then
    as one fused batch (autotuned shape).  Fan creation is outside both
    timings."""
    from underseamodelinglibrary_amd import scenarios
    from underseamodelinglibrary_amd.usml.waveq3d import fan_batch, wave_queue
    ocean = scenarios.config1().ocean
    specs = scenarios.config4_sources(256)[:n_sources]
    steps = specs[0]["steps"]
    ray_steps = sum(sp["de"].size() * sp["az"].size() * steps for sp in specs)

    def make():
        return [wave_queue(ocean, sp["freq"], sp["pos"], sp["de"], sp["az"], sp["time_step"], None, lib=lib)
                for sp in specs]

    qs = make()
    t0 = time.perf_counter()
    for q in qs:
        lib.check(lib.wq3d_step(q._fan, steps))
        q.sync()
    alone = time.perf_counter() - t0
    close_all(qs)
    qs = make()
    b = fan_batch(qs, lib=lib)
    shape, _ = b.tune(trial=4)
    b.close()
    close_all(qs)
    qs = make()
    b = fan_batch(qs, lib=lib)
    b.set_shape(**shape)
    t0 = time.perf_counter()
    lib.check(lib.wq3d_batch_step(b._h, steps))
    batched = time.perf_counter() - t0
    b.close()
    close_all(qs)
    return {"workload": f"config 4 sources 1-{n_sources} of 256 (seeded), uneven fans DE 61-721 x AZ 1-181, "
                        f"4-16 freqs, 90 s ({steps} steps), no targets, config 1's Munk ocean (off-config: "
                        "the gridded ocean is not run on the GPU)",
            "fans": n_sources, "ray_steps": ray_steps, "alone_ray_steps_per_s": ray_steps / alone,
            "batched_ray_steps_per_s": ray_steps / batched, "speedup": alone / batched, "shape": shape}


def leg_adjoint(lib, n_profiles, workers):
    """North_star's differentiation form on config 5: the forward runs of
    n_profiles seeded profiles as one fan_batch with a checkpoint of every
    fan every 100 steps, then the gradient of a loss over every eigenray's
    travel time and intensities with respect to each profile's 64 node
    speeds (wq3d_adjoint_gradient, `workers` fans at once).  Profile 0 alone
    too: its plain forward run and its backward pass.  The gradient's
    agreement with central differences of the oracle is checked in
    tests/test_adjoint.py."""
    import numpy as np
    from underseamodelinglibrary_amd import scenarios
    from underseamodelinglibrary_amd.usml.adjoint import profile_runs
    from underseamodelinglibrary_amd.usml.waveq3d import wave_queue

    def make(speeds):
        return scenarios.config5(speeds=speeds)

    speeds = np.stack([scenarios.config5_profile(p)[1] for p in range(n_profiles)])
    runs = profile_runs(make, checkpoint_every=100, batched=True, lib=lib)
    t0 = time.perf_counter()
    runs.run(speeds)
    fwd = time.perf_counter() - t0
    counts = runs.counts()
    nf = runs.nf
    rng = np.random.default_rng(1)
    gt = rng.normal(size=sum(counts))
    gi = rng.normal(size=(sum(counts), nf))
    t0 = time.perf_counter()
    grad = runs.gradient(gt, gi, workers=workers, release=True)
    bwd = time.perf_counter() - t0
    runs.close()
    # one profile alone: plain forward (no checkpoints), then with
    # checkpoints and its backward pass
    sc = make(speeds[0])
    q = wave_queue(*sc.args(), lib=lib)
    t0 = time.perf_counter()
    lib.check(lib.wq3d_step(q._fan, sc.steps))
    q.sync()
    one_fwd = time.perf_counter() - t0
    q.close()
    one = profile_runs(make, checkpoint_every=100, batched=False, lib=lib)
    one.run(speeds[:1])
    n1 = one.counts()[0]
    t0 = time.perf_counter()
    one.gradient(rng.normal(size=n1), rng.normal(size=(n1, nf)), workers=1, release=True)
    one_bwd = time.perf_counter() - t0
    one.close()
    return {"workload": f"config 5 differentiation: {n_profiles} profiles x 64 node speeds, loss over every "
                        "eigenray's travel time and 4 intensities (dB), checkpoints every 100 steps",
            "forward_batched_s": fwd, "backward_s": bwd, "backward_workers": workers,
            "backward_over_forward": bwd / fwd, "eigenrays": int(sum(counts)),
            "grad_abs_max": float(np.abs(grad).max()),
            "one_profile": {"forward_plain_s": one_fwd, "backward_s": one_bwd, "backward_over_forward":
                            one_bwd / one_fwd, "eigenrays": int(n1)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2400)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--profiles", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-procs", type=int, default=16)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--legs", default="config1,config2fan,collection,sources,adjoint")
    ap.add_argument("--adjoint-workers", type=int, default=8)
    ap.add_argument("--adjoint-profiles", type=int, default=None,
                    help="profiles in the adjoint leg (default: --profiles)")
    ap.add_argument("--sources", type=int, default=256, help="config 4 sources in the batch leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(independent fans or target subsets would shard with no exchange)", file=sys.stderr)
        sys.exit(2)

    from underseamodelinglibrary_amd import _lib
    lib = _lib.load()
    if lib.wq3d_device_count() < 1:
        raise RuntimeError("no HIP device")
    n = args.profiles
    log(f"[bench] config 5: {n} profiles, warmup {args.warmup}, timed {args.steps}")
    shape, tune_ms = tune(lib, n)
    log(f"[bench] autotuned shape {shape} ({tune_ms:.3f} ms/step in its trial)")
    sc, dt, rays, stage_ms, stage_n, pairs, boxes = run_batched(lib, n, args.warmup, args.steps, shape, True)
    n_rays = sc.num_rays()
    total_rs = n * n_rays * args.steps
    log(f"[bench] batched: {dt:.3f} s, {total_rs / dt / 1e9:.3f} G ray-steps/s, {rays} eigenrays")
    _, dt_graph, rays_graph, _, _, _, _ = run_batched(lib, n, args.warmup, args.steps, dict(shape, use_graph=True),
                                                     False)
    log(f"[bench] batched, graph replay: {dt_graph:.3f} s")
    plain = None
    if not args.no_plain:
        _, dt_plain, rays_plain, = run_plain(lib, n, args.warmup, args.steps)
        plain = {"ray_steps_per_s": total_rs / dt_plain, "eigenrays": rays_plain, "seconds": dt_plain}
        log(f"[bench] plain: {dt_plain:.3f} s")
        assert rays_plain == rays, (rays_plain, rays)
    # roofline of the search stage: FP64 ops counted on the device over the
    # timed steps / the stage's mean time per step (HIP events)
    flops_per_step = (W_TGT * pairs + W_BOX * boxes) / max(1, stage_n)
    achieved = flops_per_step / (stage_ms * 1e-3) / 1e12 if stage_ms > 0 else 0.0
    out = {
        "metric": "ray-steps/sec + eigenrays/sec on 1 MI355X, plain vs batched+adjoint at fixed tol",
        "form": "batched: one fused launch per kernel of a step for all 32 fans (wq3d_batch), autotuned shape",
        "value": total_rs / dt,
        "unit": "ray-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded perturbed Munk sound speed profiles, seeded target grid)",
        "config": {"workload": f"config 5 forward: {n} seeded 64-node piecewise-linear c(z) profiles "
                               "(1-D profile_grid, linear), flat 4500 m Rayleigh-sand bottom, "
                               "901 DE x 72 AZ rays, 4 freqs 500-4000 Hz, 50x50 targets, dt 0.05 s, 2401 steps",
                   "fans": n, "rays_per_fan": n_rays, "targets_per_fan": sc.num_targets(),
                   "frequencies": sc.freq.size()},
        "eigenrays": rays,
        "eigenrays_per_s": rays / dt,
        "plain": plain,
        "speedup_vs_plain": (plain["seconds"] / dt) if plain else None,
        "batched_graph": {"ray_steps_per_s": total_rs / dt_graph, "eigenrays": rays_graph,
                          "note": "the same batch with 64-step loops replayed as HIP graphs (the timed region "
                                  "above launches step by step so that HIP events can bracket the search stage)"},
        "autotuned_shape": shape,
        "roofline": {"bound": "fp64-vector", "kernel": "k_supmask_multi + k_detect_multi (the search stage)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": SEARCH_TRAFFIC_BYTES,
                     "stage_ms_per_step": stage_ms, "pairs_tested_per_step": pairs / max(1, stage_n),
                     "box_tests_per_step": boxes / max(1, stage_n),
                     "pairs_all_per_step": float(n * sc.num_targets() * n_rays),
                     "note": "FP64 vector (no MFMA-shaped work); achieved = (13 FP64 ops per (target, ray) pair "
                             "put to the window test + 59 per (ray, box) bound), both counted on the device "
                             "over the timed steps, / the stage's mean time per step (HIP events on the "
                             "batch's stream around its two fused launches); traffic in bytes per step of the "
                             "stage"},
    }
    legs = {}
    names = [x for x in args.legs.split(",") if x]
    if "config1" in names:
        legs["config1"] = leg_config1(lib)
        log(f"[bench] config 1: {legs['config1']['ray_steps_per_s'] / 1e6:.1f} M ray-steps/s")
    if "config2fan" in names:
        legs["config2_fan_munk"], sc2, recs, inten, phase = leg_config2_fan(lib)
        log(f"[bench] config 2 fan: {legs['config2_fan_munk']['ms_per_step']:.3f} ms/step")
        if "collection" in names:
            legs["eigenray_collection"] = leg_collection(lib, sc2, recs, inten, phase, not args.no_cpu_baseline)
    if "sources" in names and args.sources > 0:
        legs["config4_sources"] = leg_sources(lib, args.sources)
        log(f"[bench] config 4 sources: {legs['config4_sources']['speedup']:.2f}x batched over alone")
    if "adjoint" in names:
        legs["adjoint"] = leg_adjoint(lib, args.adjoint_profiles or n, args.adjoint_workers)
        log(f"[bench] adjoint: forward {legs['adjoint']['forward_batched_s']:.2f} s, backward "
            f"{legs['adjoint']['backward_s']:.2f} s; one profile {legs['adjoint']['one_profile']['backward_over_forward']:.2f}x")
    out["legs"] = legs
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps, args.cpu_procs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
