"""The kernel bodies of libwq3d run on the CPU (serial emulation build with
AddressSanitizer, device allocations poisoned with 0xff) against the oracle.

This checks every array access the kernels make and their exact agreement
with the oracle, including the gridded sound-speed / bathymetry path that is
switched off on the GPU.  Each case runs in a child process with the ASan
runtime preloaded.
"""
import os
import subprocess
import sys
import textwrap

import pytest

import emulation

COMMON = textwrap.dedent("""
    import sys
    import emulation, oracle
    from underseamodelinglibrary_amd import scenarios, _abi
    from underseamodelinglibrary_amd.usml import ocean as oc
    from underseamodelinglibrary_amd.usml.types import seq_linear, seq_log
    from underseamodelinglibrary_amd.usml.waveq3d import wave_queue
    from parity import compare_eigenrays, compare_fronts
    lib = emulation.load()
    oracle.build(); oracle.load()

    def run(sc, verbs=False, limits=None):
        g = wave_queue(*sc.args(), lib=lib)
        o = oracle.oracle_queue(*sc.args())
        if limits is not None:
            for q in (g, o):
                q.max_surface(limits[0])
                q.max_bottom(limits[1])
        if verbs:
            d = type("L", (), {"add_eigenverb": lambda *a: None, "reflect": lambda *a: None})()
            for q in (g, o):
                q.add_eigenverb_listener(d)
                q.add_reflection_listener(d)
        g._lib.check(g._lib.wq3d_step(g._fan, sc.steps))
        o._lib.orc_step(o._h, sc.steps)
        gr, gi, gp = g.drain_eigenrays()
        orr, oi, op = o.drain_eigenrays()
        n = compare_eigenrays(gr, gi, gp, orr, oi, op, sc.freq.size())
        for slot in (_abi.PAST, _abi.PREV, _abi.CURR, _abi.NEXT):
            compare_fronts(g.snapshot_arrays(slot), o.snapshot_arrays(slot), rtol=0.0)
        if verbs:
            gv, gpw = g.drain_eigenverbs()
            ov, opw = o.drain_eigenverbs()
            assert len(gv) == len(ov)
            for k in gv.dtype.names:
                if k != "pad_":
                    assert (gv[k] == ov[k]).all(), k
            assert (gpw == opw).all()
            gf, of = g.drain_reflections(), o.drain_reflections()
            assert len(gf) == len(of)
            for k in gf.dtype.names:
                assert (gf[k] == of[k]).all(), k
        return n
""")


def _run(body, timeout=900):
    res = emulation.run_script_under_asan(COMMON + textwrap.dedent(body), timeout=timeout)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-4000:]
    return res.stdout


@pytest.fixture(scope="module", autouse=True)
def emulation_build():
    from underseamodelinglibrary_amd import build
    build.build_emulation()


def test_emulated_config2_gridded_exact():
    """Gridded SSP (pchip x linear x linear) and bathymetry, bit-identical
    wavefronts and eigenrays."""
    out = _run("""
        sc = scenarios.config2(de=seq_linear(-80.0, 2.0, 80.0), az=seq_linear(0.0, 10.0, 350.0),
                               n_targets=(10, 10), time_max=30.0, freq=seq_log(100.0, 10.0, 3))
        print("eigenrays", run(sc))
    """)
    assert int(out.split("eigenrays")[-1]) > 500


GRIDDED_CASE = textwrap.dedent("""
    sc = scenarios.config2(de=seq_linear(-80.0, 2.0, 80.0), az=seq_linear(0.0, 10.0, 350.0),
                           n_targets=(10, 10), time_max=STEPS_S, freq=seq_log(100.0, 10.0, 3))
    print("eigenrays", run(sc, verbs=True))
""")


@pytest.mark.parametrize("order", ["reverse", "phased"])
def test_emulated_config2_gridded_launch_orders(order):
    """The gridded case with the rays of each launch visited last to first,
    and with every ray's caustic / vertex part before any ray's reflections:
    a kernel whose result depended on the order of its threads (a neighbour
    read racing a neighbour write) would differ.  Bit-identical to the
    oracle, eigenverbs and reflections included; every pointer each launch
    hands its kernel is checked against the emulated device allocations."""
    out = _run(f"import os\nos.environ['WQ3D_EMU_ORDER'] = '{order}'\n" +
               GRIDDED_CASE.replace("STEPS_S", "30.0"))
    assert int(out.split("eigenrays")[-1]) > 500


def test_emulated_config2_gridded_split_lanes():
    """The GPU form of eigenray construction (a group of 32 lanes per
    candidate, cross-lane shuffles, an LDS cube, walk cells fetched ahead by
    the lanes) with the lanes as host threads in lockstep: every group
    operation is a barrier that all lanes must reach from the same call site
    of the same candidate.  Bit-identical to the oracle on the gridded case."""
    out = _run("import os\nos.environ['WQ3D_EMU_SPLIT'] = '1'\n" + GRIDDED_CASE.replace("STEPS_S", "15.0"))
    assert int(out.split("eigenrays")[-1]) > 100


def test_emulated_config2_gridded_ubsan(tmp_path):
    """The gridded case under UndefinedBehaviorSanitizer with
    float-cast-overflow: no float-to-integer conversion of the kernels is
    out of range (the indices of grid lookups and table reads come from
    such conversions), both construction forms."""
    from underseamodelinglibrary_amd import build
    lib = str(tmp_path / "libwq3d_ubsan.so")
    subprocess.run([build.HIPCC, "--offload-host-only", "-O1", "-g", "-std=c++17", "-fPIC", "-shared",
                    "-ffp-contract=off", "-fno-fast-math", "-DWQ3D_EMULATE", "-Xarch_host",
                    "-fsanitize=undefined,float-cast-overflow", "-Xarch_host", "-fno-sanitize-recover=all",
                    "-shared-libsan", "-I", os.path.join(emulation.ROOT, "include"), "-o", lib, build.SRC],
                   check=True, capture_output=True)
    rt = subprocess.run(["/opt/rocm/bin/amdclang++", "-print-file-name=libclang_rt.ubsan_standalone-x86_64.so"],
                        capture_output=True, text=True, check=True).stdout.strip()
    script = COMMON.replace("lib = emulation.load()", f"import ctypes as C\nfrom underseamodelinglibrary_amd "
                            f"import _lib as _L\nlib = _L.bind(C.CDLL({lib!r}))")
    for split in ("0", "1"):
        env = dict(os.environ, LD_PRELOAD=rt, UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
                   WQ3D_EMU_SPLIT=split,
                   PYTHONPATH=os.pathsep.join([emulation.ROOT, os.path.join(emulation.ROOT, "oracle"),
                                               os.path.join(emulation.ROOT, "tests")]))
        res = subprocess.run([sys.executable, "-c", script + GRIDDED_CASE.replace("STEPS_S", "10.0")],
                             env=env, capture_output=True, text=True, timeout=900)
        assert res.returncode == 0 and "runtime error" not in res.stderr, res.stdout[-2000:] + res.stderr[-4000:]
        assert int(res.stdout.split("eigenrays")[-1]) > 10


def test_find_index_defined_for_every_input():
    """find_index of the kernels (seq_linear / seq_log / seq_data searches)
    for NaN, infinities and values far outside the axis: always an index in
    [0, size-2], equal to the reference's max(0, min(size-2,
    (int64)floor(q))) as x86-64 evaluates it (numpy's float64 -> int64 cast
    uses the same conversion, which gives INT64_MIN for NaN and |q| >= 2^63)."""
    out = _run("""
        import ctypes as C, warnings
        import numpy as np
        fn = lib.wq3d_emu_find_index
        fn.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32, C.c_int32, C.c_double]
        fn.restype = C.c_int32
        probes = [float("nan"), float("inf"), -float("inf"), 1e300, -1e300, 1e18, -1e18, 1e15, -1e15,
                  0.0, 1.0, -1.0]
        bad = 0
        checked = 0
        for axis, search in ((seq_linear(6371000.0, -128.2, 40), 0), (seq_linear(-0.7, 0.004, 41), 0),
                             (seq_log(100.0, 10.0 ** (2.0 / 7.0), 8), 1)):
            v, inc = axis.values, axis.increments
            n = v.size
            xs = probes + list(v) + [float(x) for x in (v[:-1] + v[1:]) / 2] + [v[0] * 1.0001, v[-1] * 1.0001]
            for x in xs:
                k = fn(v.ctypes.data_as(C.POINTER(C.c_double)), inc.ctypes.data_as(C.POINTER(C.c_double)),
                       n, search, x)
                with np.errstate(all="ignore"), warnings.catch_warnings():
                    warnings.simplefilter("ignore")
                    q = (np.floor((np.float64(x) - v[0]) / inc[0]) if search == 0 else
                         np.floor(np.log(np.float64(x) / v[0]) / np.log(inc[1] / inc[0])))
                    ref = max(0, min(n - 2, int(np.array([q]).astype(np.int64)[0])))
                checked += 1
                if not (0 <= k <= n - 2) or (search == 0 and k != ref):
                    bad += 1
                    print("mismatch", x, k, ref)
        print("counts", checked, bad)
    """)
    checked, bad = (int(x) for x in out.split("counts")[-1].split())
    assert checked > 100 and bad == 0


def test_emulated_classic_ray_exact():
    """CLASSIC_RAY spreading on config 1's ocean and on the branch-point fan
    closing on itself in AZ, bit-identical to the oracle."""
    out = _run("""
        sc = scenarios.config1(de=seq_linear(-60.0, 3.0, 60.0), az=seq_linear(0.0, 20.0, 340.0),
                               n_targets=(6, 6), time_max=40.0)
        sc.spreading = _abi.SPREADING_CLASSIC_RAY
        n = run(sc)
        sc = scenarios.eigenray_branch_pt()
        sc.spreading = _abi.SPREADING_CLASSIC_RAY
        print("eigenrays", n + run(sc))
    """)
    assert int(out.split("eigenrays")[-1]) > 40


def test_emulated_eigenverbs_volume_exact():
    out = _run("""
        sc = scenarios.config1(de=seq_linear(-60.0, 3.0, 60.0), az=seq_linear(0.0, 30.0, 330.0),
                               time_max=20.0)
        sc.ocean.add_volume(oc.volume_flat(300.0, 10.0))
        print("eigenrays", run(sc, verbs=True))
    """)
    assert int(out.split("eigenrays")[-1]) > 10


def test_emulated_bounce_threshold_exact():
    """waveq3d/test/reflection_test.cc:597-656 scenario with the reference's
    limits (3 surface, 5 bottom bounces) and with the bottom alone limited to
    2: eigenrays, eigenverbs and reflections bit-identical to the oracle."""
    out = _run("""
        n = run(scenarios.bounce_threshold(), verbs=True, limits=(3, 5))
        print("eigenrays", n + run(scenarios.bounce_threshold(), verbs=True, limits=(999, 2)))
    """)
    assert int(out.split("eigenrays")[-1]) > 10


def test_emulated_known_answer_scenarios_exact():
    """Lloyd's mirror in range (1961 targets in one column) and depth on a
    flat earth with tangent-spaced DE (proploss_test.cc:357-724) and the
    hybrid Gaussian known-answer case (proploss_test.cc:153-245),
    bit-identical to the oracle."""
    out = _run("""
        n = run(scenarios.lloyds_range()) + run(scenarios.lloyds_depth())
        print("eigenrays", n + run(scenarios.proploss_gaussian()))
    """)
    assert int(out.split("eigenrays")[-1]) > 2000


def test_emulated_az_wrap_fan_exact():
    """AZ fan closing on itself (0:15:360, _az_boundary) with targets."""
    out = _run("""
        sc = scenarios.config1(de=seq_linear(-60.0, 4.0, 60.0), az=seq_linear(0.0, 15.0, 360.0),
                               n_targets=(4, 4), time_max=25.0)
        print("eigenrays", run(sc))
    """)
    assert int(out.split("eigenrays")[-1]) > 10


def test_emulated_branch_point_exact():
    """Targets straight above / below the source (the source-branch rule),
    with device allocations filled with random bytes instead of 0xff."""
    out = _run("""
        import os
        os.environ["WQ3D_EMU_FILL"] = "random"
        sc = scenarios.eigenray_branch_pt()
        print("eigenrays", run(sc))
    """)
    assert int(out.split("eigenrays")[-1]) >= 36


def test_detect_prefilter_matches_exact_comparisons():
    """The linear-form pre-filter of the eigenray search never changes a
    decision: random rays and targets, plus targets bisected onto the exact
    boundary of every comparison of the pre-test, decided with and without
    it, and no block of targets the block bound rejects holds a target inside
    the window (emulation-only test hook)."""
    out = _run("""
        import ctypes as C
        counts = (C.c_int64 * 5)()
        fn = lib.wq3d_emu_check_detect
        fn.argtypes = [C.c_int64, C.c_int64, C.c_uint64, C.c_double, C.POINTER(C.c_int64)]
        fn.restype = C.c_int32
        total = [0, 0, 0, 0, 0]
        for seed, radius in ((1, 6371000.0), (2, 6378101.030201019), (3, 6371000.0)):
            assert fn(200, 200, seed, radius, counts) == 0
            total = [a + b for a, b in zip(total, counts)]
        print("counts", *total)
    """)
    pairs, mismatches, window, passes, blocks = (int(x) for x in out.split("counts")[-1].split())
    assert mismatches == 0
    assert pairs > 100000 and window > passes > 100 and blocks > 1000


def test_search_totals_count_and_reset():
    """wq3d_search_totals: the pairs put to the window test over a run lie
    between the last step's filter hits and rays x targets x steps, the
    block bounds tested within one per (ray, step, box), and a second call
    reads zero."""
    out = _run("""
        import ctypes as C
        sc = scenarios.config1(n_targets=(4, 4), time_max=3.0)
        g = wave_queue(*sc.args(), lib=lib)
        g._lib.check(g._lib.wq3d_step(g._fan, sc.steps))
        h, c, t, b, t2, b2 = (C.c_int64() for _ in range(6))
        lib.check(lib.wq3d_search_stats(g._fan, C.byref(h), C.byref(c)))
        lib.check(lib.wq3d_search_totals(g._fan, C.byref(t), C.byref(b)))
        lib.check(lib.wq3d_search_totals(g._fan, C.byref(t2), C.byref(b2)))
        nblk = (sc.num_targets() + 63) // 64
        print("counts", h.value, c.value, t.value, b.value, t2.value, b2.value,
              sc.num_rays() * sc.num_targets() * sc.steps,
              sc.num_rays() * sc.steps * (nblk + (nblk + 7) // 8))
    """)
    hits, cands, tested, forms, again, again_forms, bound, forms_bound = (
        int(x) for x in out.split("counts")[-1].split())
    assert cands <= hits <= tested and 0 < tested < bound
    # at most one bound per (ray, step, box)
    assert 0 < forms <= forms_bound
    assert again == 0 and again_forms == 0


def test_emulated_tiny_and_ragged_fans_exact():
    """Fans smaller than one 8 x 8 ray tile and fans whose sizes are not
    multiples of it, with one target and with several: 1 x 1, 1 x 5, 2 x 1,
    3 x 9 and 13 x 7 rays (partial tiles in the search, edge rows only)."""
    out = _run("""
        from underseamodelinglibrary_amd.usml.types import wposition
        total = 0
        for de, az, nt in ((seq_linear(-10.0, 1.0, 1), seq_linear(0.0, 1.0, 1), (1, 1)),
                           (seq_linear(-10.0, 1.0, 1), seq_linear(-2.0, 1.0, 5), (2, 2)),
                           (seq_linear(-10.0, 20.0, 2), seq_linear(0.0, 1.0, 1), (1, 1)),
                           (seq_linear(-20.0, 10.0, 3), seq_linear(-4.0, 1.0, 9), (3, 3)),
                           (seq_linear(-30.0, 5.0, 13), seq_linear(-6.0, 2.0, 7), (4, 4))):
            sc = scenarios.config1(de=de, az=az, n_targets=nt, time_max=15.0)
            total += run(sc)
        print("eigenrays", total)
    """)
    assert int(out.split("eigenrays")[-1]) > 0


def test_emulated_limits_and_errors():
    """The library's limits and argument checks: more than 4096 AZ rays or
    32 frequencies and CLASSIC_RAY spreading are refused as unsupported,
    an unknown spreading model as unsupported, a non-positive time step and
    a negative step count as bad arguments;
    zero steps is a no-op.  Each failure leaves a message."""
    out = _run("""
        import ctypes as C
        from underseamodelinglibrary_amd import _lib
        from underseamodelinglibrary_amd.usml.waveq3d import wave_queue

        def status_of(sc, **kw):
            try:
                q = wave_queue(*sc.args(), lib=lib, **kw)
            except _lib.WQ3DError as e:
                return e.status, lib.wq3d_last_error().decode()
            q.close()
            return 0, ""

        base = scenarios.config1(de=seq_linear(-10.0, 10.0, 3), az=seq_linear(0.0, 90.0, 4),
                                 n_targets=(2, 2), time_max=1.0)
        wide = scenarios.config1(de=seq_linear(-10.0, 10.0, 3), az=seq_linear(0.0, 360.0 / 4097, 4097),
                                 n_targets=(2, 2), time_max=1.0)
        many = scenarios.config1(de=seq_linear(-10.0, 10.0, 3), az=seq_linear(0.0, 90.0, 4),
                                 n_targets=(2, 2), time_max=1.0)
        many.freq = seq_log(100.0, 1.1, 33)
        zero_dt = scenarios.config1(de=seq_linear(-10.0, 10.0, 3), az=seq_linear(0.0, 90.0, 4),
                                    n_targets=(2, 2), time_max=1.0)
        zero_dt.time_step = 0.0
        print("wide", *status_of(wide))
        print("many", *status_of(many))
        unknown = scenarios.config1(de=seq_linear(-10.0, 10.0, 3), az=seq_linear(0.0, 90.0, 4),
                                    n_targets=(2, 2), time_max=1.0)
        unknown.spreading = 7
        print("classic", *status_of(unknown))
        print("zero_dt", *status_of(zero_dt))
        q = wave_queue(*base.args(), lib=lib)
        print("neg", lib.wq3d_step(q._fan, -1))
        print("zero", lib.wq3d_step(q._fan, 0), len(q.drain_eigenrays()[0]), q.time())
        q.close()
        print("done")
    """)
    lines = dict(l.split(" ", 1) for l in out.strip().splitlines() if " " in l)
    assert lines["wide"].startswith("2 ") and "4096" in lines["wide"]
    assert lines["many"].startswith("2 ") and "32" in lines["many"]
    assert lines["classic"].startswith("2 ") and "CLASSIC_RAY or HYBRID_GAUSSIAN" in lines["classic"]
    assert lines["zero_dt"].startswith("1 ") and "time_step" in lines["zero_dt"]
    assert lines["neg"] == "1"
    assert lines["zero"] == "0 0 0.0"
