"""The kernel bodies of libwq3d run on the CPU (serial emulation build with
AddressSanitizer, device allocations poisoned with 0xff) against the oracle.

This checks every array access the kernels make and their exact agreement
with the oracle, including the gridded sound-speed / bathymetry path that is
switched off on the GPU.  Each case runs in a child process with the ASan
runtime preloaded.
"""
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

    def run(sc, verbs=False):
        g = wave_queue(*sc.args(), lib=lib)
        o = oracle.oracle_queue(*sc.args())
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


def test_emulated_eigenverbs_volume_exact():
    out = _run("""
        sc = scenarios.config1(de=seq_linear(-60.0, 3.0, 60.0), az=seq_linear(0.0, 30.0, 330.0),
                               time_max=20.0)
        sc.ocean.add_volume(oc.volume_flat(300.0, 10.0))
        print("eigenrays", run(sc, verbs=True))
    """)
    assert int(out.split("eigenrays")[-1]) > 10


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
