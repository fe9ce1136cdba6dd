"""eigenray_collection's per-target sum (eigenrays/eigenray_collection.cc:
90-191): the oracle against the reference's known answer and closed forms
(CPU), the kernel body against the oracle in the CPU emulation build, and
the GPU library against the oracle through the C ABI (bit-exact: both use
the shared elementary functions)."""
import math
import textwrap

import numpy as np
import pytest

from underseamodelinglibrary_amd import _abi, scenarios
from underseamodelinglibrary_amd.usml.eigenrays import (eigenray_collection, pack_eigenrays,
                                                        sum_eigenray_arrays)
from underseamodelinglibrary_amd.usml.types import seq_linear, wposition, wposition1
from underseamodelinglibrary_amd.usml.waveq3d import eigenray_model, wave_queue


def known_answer_rays(nf=1):
    """eigenrays/test/eigenrays_test.cc:25-44 and :57-88: three eigenrays at
    target (0, 0), index n = 0, 1, 2."""
    out = []
    for n in range(3):
        r = eigenray_model()
        r.travel_time = 10.0 + n
        r.source_de, r.source_az = 20.0 + n, 30.0 + n
        r.target_de, r.target_az = 40.0 + n, 50.0 + n
        r.surface, r.bottom, r.caustic, r.upper, r.lower = 60 + n, 70 + n, 80 + n, 90 + n, 10 + n
        r.intensity = np.full(nf, 200.0 + n)
        r.phase = np.full(nf, (n + 1.0) / 10.0)
        out.append((0, 0, r))
    return out


def random_rays(rng, n, n1, n2, nf):
    out = []
    for _ in range(n):
        r = eigenray_model()
        r.travel_time = rng.uniform(1.0, 100.0)
        r.source_de, r.source_az = rng.uniform(-80, 80), rng.uniform(-180, 360)
        r.target_de, r.target_az = rng.uniform(-80, 80), rng.uniform(-180, 360)
        r.surface, r.bottom, r.caustic = (int(x) for x in rng.integers(0, 5, 3))
        r.upper, r.lower = (int(x) for x in rng.integers(0, 9, 2))
        r.intensity = rng.uniform(40.0, 140.0, nf)
        r.phase = rng.uniform(-math.pi, math.pi, nf)
        out.append((int(rng.integers(0, n1)), int(rng.integers(0, n2)), r))
    return out


def oracle_sum(oracle_mod, entries, freq, n1, n2, coherent):
    lib = oracle_mod.load()
    recs, inten, phase = pack_eigenrays(entries, len(freq))
    st, tot, ti, tp = sum_eigenray_arrays(lib.orc_sum_eigenrays, recs, inten, phase, freq, n1, n2,
                                          coherent)
    assert st == 0
    return tot, ti, tp


def closed_form(entries, freq, n1, n2, coherent):
    """numpy restatement of the sum for comparison to rounding."""
    nf = len(freq)
    res = {}
    for t1 in range(n1):
        for t2 in range(n2):
            rays = [r for (a, b, r) in entries if (a, b) == (t1, t2)]
            I, P = [], []
            num = np.zeros(5)
            wgt = 0.0
            for f in range(nf):
                a = np.array([10.0 ** (r.intensity[f] / -20.0) for r in rays])
                if coherent:
                    p = np.array([math.fmod(2 * math.pi * freq[f] * r.travel_time + r.phase[f],
                                            2 * math.pi) for r in rays])
                    z = np.sum(a * np.exp(1j * p)) if rays else 0j
                else:
                    z = complex(np.sum(a)) if rays else 0j
                I.append(-20.0 * math.log10(max(1e-15, abs(z))))
                P.append(math.atan2(z.imag, z.real))
                w = a * a
                wgt += float(np.sum(w))
                num += np.array([np.sum(w * np.array([r.travel_time for r in rays])),
                                 np.sum(w * np.array([r.source_de for r in rays])),
                                 np.sum(w * np.array([r.target_de for r in rays])),
                                 0.0, 0.0]) if rays else 0.0
            res[(t1, t2)] = (np.array(I), np.array(P), num[:3] / wgt if rays else None, len(rays))
    return res


def test_oracle_known_answer(oracle_mod):
    """eigenrays_test.cc:57-88: the coherent total of the three rays is
    191.4 dB within 0.1 %."""
    freq = np.array([3000.0])
    tot, ti, tp = oracle_sum(oracle_mod, known_answer_rays(), freq, 1, 1, True)
    assert abs(ti[0] - 191.4) <= 0.001 * 191.4
    assert tot[0]["count"] == 3
    assert tot[0]["initial_time"] == 10.0
    # the strongest ray (n = 0, 200 dB) sets the counters
    assert (tot[0]["surface"], tot[0]["bottom"], tot[0]["caustic"], tot[0]["upper"],
            tot[0]["lower"]) == (60, 70, 80, 90, 10)


@pytest.mark.parametrize("coherent", [True, False])
def test_oracle_closed_form(oracle_mod, coherent):
    rng = np.random.default_rng(7)
    freq = np.array([100.0, 1000.0, 3150.0])
    n1, n2 = 3, 4
    entries = random_rays(rng, 60, n1, n2, len(freq))
    tot, ti, tp = oracle_sum(oracle_mod, entries, freq, n1, n2, coherent)
    ref = closed_form(entries, freq, n1, n2, coherent)
    nf = len(freq)
    for (t1, t2), (I, P, avg, count) in ref.items():
        k = t1 * n2 + t2
        assert tot[k]["count"] == count
        np.testing.assert_allclose(ti[k * nf:(k + 1) * nf], I, rtol=0, atol=1e-9)
        if count:
            np.testing.assert_allclose(tp[k * nf:(k + 1) * nf], P, rtol=0, atol=1e-9)
            np.testing.assert_allclose([tot[k]["travel_time"], tot[k]["source_de"],
                                        tot[k]["target_de"]], avg, rtol=1e-12)
            times = [r.travel_time for (a, b, r) in entries if (a, b) == (t1, t2)]
            assert tot[k]["initial_time"] == min(times)


def test_oracle_incoherent_sums_amplitudes(oracle_mod):
    """The reference's incoherent total adds pressure amplitudes, not
    intensities: two equal rays give I - 20 log10(2)."""
    r1, r2 = eigenray_model(), eigenray_model()
    for r in (r1, r2):
        r.travel_time, r.source_de, r.source_az, r.target_de, r.target_az = 5.0, 0.0, 0.0, 0.0, 0.0
        r.surface = r.bottom = r.caustic = r.upper = r.lower = 0
        r.intensity, r.phase = np.array([80.0]), np.array([0.0])
    tot, ti, tp = oracle_sum(oracle_mod, [(0, 0, r1), (0, 0, r2)], np.array([1000.0]), 1, 1, False)
    assert abs(ti[0] - (80.0 - 20.0 * math.log10(2.0))) < 1e-12
    assert tp[0] == 0.0


def test_oracle_coherent_cancellation(oracle_mod):
    """Two equal rays half a cycle apart cancel to the 1e-15 floor (300 dB)."""
    f = 1000.0
    rays = []
    for k in range(2):
        r = eigenray_model()
        r.travel_time = 1.0 + k * 0.5 / f
        r.source_de = r.source_az = r.target_de = r.target_az = 0.0
        r.surface = r.bottom = r.caustic = r.upper = r.lower = 0
        r.intensity, r.phase = np.array([60.0]), np.array([0.0])
        rays.append((0, 0, r))
    tot, ti, tp = oracle_sum(oracle_mod, rays, np.array([f]), 1, 1, True)
    assert ti[0] > 60.0 + 150.0


def test_oracle_empty_targets(oracle_mod):
    """Targets without eigenrays keep the reference's empty-list values."""
    tot, ti, tp = oracle_sum(oracle_mod, known_answer_rays(), np.array([3000.0]), 2, 1, True)
    assert ti[1] == 300.0 and tp[1] == 0.0
    assert math.isnan(tot[1]["travel_time"]) and tot[1]["source_az"] == 90.0
    assert tot[1]["count"] == 0 and tot[1]["initial_time"] == 0.0 and tot[1]["surface"] == -1


def test_collection_listener_bookkeeping():
    """add_eigenray files rays per target and keeps the initial time."""
    sc = scenarios.eigenray_basic()
    col = eigenray_collection(sc.freq, sc.pos, sc.targets, lib=_FakeLib())
    for t1, t2, r in known_answer_rays():
        col.add_eigenray(t1, t2, r)
    assert col.num_eigenrays() == 3 and len(col.eigenrays(0, 0)) == 3
    assert col.initial_time(0, 0) == 10.0
    assert col.coherent()


class _FakeLib:
    pass


def test_emulated_sum_matches_oracle(oracle_mod):
    """The kernel body, built for the CPU with bounds checking, against the
    oracle on random rays (coherent and incoherent), bit for bit."""
    import emulation
    res = emulation.run_script_under_asan(textwrap.dedent("""
        import numpy as np
        import emulation, oracle
        import test_collection as tc
        from underseamodelinglibrary_amd.usml.eigenrays import pack_eigenrays, sum_eigenray_arrays
        lib = emulation.load()
        rng = np.random.default_rng(11)
        freq = np.array([100.0, 1000.0, 3150.0, 10000.0])
        entries = tc.random_rays(rng, 300, 5, 7, len(freq))
        recs, inten, phase = pack_eigenrays(entries, len(freq))
        for coherent in (True, False):
            st, tot, ti, tp = sum_eigenray_arrays(lib.wq3d_sum_eigenrays, recs, inten, phase, freq,
                                                  5, 7, coherent)
            assert st == 0
            ot, oi, op = tc.oracle_sum(oracle, entries, freq, 5, 7, coherent)
            assert np.array_equal(ti, oi) and np.array_equal(tp, op)
            for name in tot.dtype.names:
                assert np.array_equal(tot[name], ot[name], equal_nan=True), name
        print("ok")
    """), timeout=600)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-3000:]
    assert "ok" in res.stdout


def _gpu_vs_oracle(oracle_mod, entries, freq, n1, n2):
    from underseamodelinglibrary_amd import _lib
    lib = _lib.load()
    recs, inten, phase = pack_eigenrays(entries, len(freq))
    for coherent in (True, False):
        st, tot, ti, tp = sum_eigenray_arrays(lib.wq3d_sum_eigenrays, recs, inten, phase, freq,
                                              n1, n2, coherent)
        lib.check(st)
        ot, oi, op = oracle_sum(oracle_mod, entries, freq, n1, n2, coherent)
        assert np.array_equal(ti, oi) and np.array_equal(tp, op)
        for name in tot.dtype.names:
            assert np.array_equal(tot[name], ot[name], equal_nan=True), name
    return tot


@pytest.mark.gpu
def test_gpu_sum_known_answer_and_random(oracle_mod):
    print("[collection] known answer and random rays", flush=True)
    tot = _gpu_vs_oracle(oracle_mod, known_answer_rays(), np.array([3000.0]), 1, 1)
    assert tot[0]["count"] == 3
    rng = np.random.default_rng(5)
    freq = np.array([100.0, 1000.0, 3150.0, 10000.0])
    _gpu_vs_oracle(oracle_mod, random_rays(rng, 2000, 9, 11, len(freq)), freq, 9, 11)


@pytest.mark.gpu
def test_gpu_collection_on_a_fan(oracle_mod):
    """The collection attached to a wave_queue as its eigenray listener, on
    the benchmark's frequencies and environment (thinned fan, 20 x 20
    targets, 30 s): the GPU totals equal the oracle's on the same rays."""
    sc = scenarios.config2_analytic(n_targets=(20, 20), time_max=30.0)
    sc.de = seq_linear(-80.0, 1.0, 80.0)
    sc.az = seq_linear(0.0, 6.0, 354.0)
    print(f"[collection] fan {sc.num_rays()} rays, {sc.num_targets()} targets", flush=True)
    q = wave_queue(*sc.args())
    col = eigenray_collection(sc.freq, sc.pos, sc.targets)
    q.add_eigenray_listener(col)
    while q.time() < 30.0:
        q.step()
    assert col.num_eigenrays() > 1000
    col.sum_eigenrays()
    freq = np.array([sc.freq(k) for k in range(sc.freq.size())])
    ot, oi, op = oracle_sum(oracle_mod, col._entries, freq, col.size1(), col.size2(), True)
    nf = len(freq)
    for t1 in range(col.size1()):
        for t2 in range(col.size2()):
            k = t1 * col.size2() + t2
            m = col.total(t1, t2)
            assert np.array_equal(m.intensity, oi[k * nf:(k + 1) * nf])
            assert np.array_equal(m.phase, op[k * nf:(k + 1) * nf])
            assert (m.travel_time == ot[k]["travel_time"]
                    or (math.isnan(m.travel_time) and math.isnan(ot[k]["travel_time"])))


def test_write_netcdf_layout(tmp_path, oracle_mod):
    """write_netcdf (eigenray_collection.cc:196-356): read back with an
    independent netCDF classic reader (scipy.io.netcdf_file): dimensions,
    variable order, units, and the record layout with the summed record of
    each target first."""
    from scipy.io import netcdf_file
    from underseamodelinglibrary_amd.usml.types import seq_linear as sl
    freq = sl(3000.0, 1.0, 1)
    targets = wposition(2, 1, 12.0, 37.0)
    col = eigenray_collection(freq, wposition1(15.0, 35.0), targets, lib=_FakeLib())
    for t1, t2, r in known_answer_rays():
        col.add_eigenray(t1, t2, r)
    # totals from the oracle (sum_eigenrays needs the GPU library)
    tot, ti, tp = oracle_sum(oracle_mod, col._entries, np.array([3000.0]), 2, 1, True)
    for t1 in range(2):
        m = col._empty_total()
        m.intensity, m.phase = ti[t1:t1 + 1].copy(), tp[t1:t1 + 1].copy()
        for name in ("travel_time", "source_de", "source_az", "target_de", "target_az"):
            setattr(m, name, float(tot[t1][name]))
        for name in ("surface", "bottom", "caustic", "upper", "lower"):
            setattr(m, name, int(tot[t1][name]))
        col._total[t1][0] = m
    path = str(tmp_path / "eigenrays.nc")
    col.write_netcdf(path, "create_eigenray test")
    f = netcdf_file(path, "r", mmap=False)
    assert f.long_name == b"create_eigenray test" and f.Conventions == b"COARDS"
    assert f.dimensions == {"rows": 2, "cols": 1, "eigenrays": 5, "frequencies": 1}
    names = list(f.variables)
    assert names[:4] == ["sourceID", "source_latitude", "source_longitude", "source_altitude"]
    assert names[-5:] == ["surface", "bottom", "caustic", "upper", "lower"]
    v = f.variables
    assert v["intensity"].units == b"dB" and v["source_az"].positive == b"clockwise"
    assert list(v["proploss_index"][:, 0]) == [0, 4]
    assert list(v["eigenray_index"][:, 0]) == [1, 5]
    assert list(v["eigenray_num"][:, 0]) == [3, 0]
    assert v["intensity"][0, 0] == ti[0] and v["intensity"][4, 0] == 300.0
    assert list(v["intensity"][1:4, 0]) == [200.0, 201.0, 202.0]
    assert list(v["surface"][1:4]) == [60, 61, 62]
    assert list(v["travel_time"][1:4]) == [10.0, 11.0, 12.0]
    assert v["initial_time"][0, 0] == 10.0 and v["initial_time"][1, 0] == 0.0
    assert abs(v["source_latitude"].getValue() - 15.0) < 1e-12
    assert abs(v["latitude"][0, 0] - 12.0) < 1e-9 and abs(v["longitude"][0, 0] - 37.0) < 1e-9
    f.close()
