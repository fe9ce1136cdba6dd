"""The C++ wave_queue mirror (underseamodelinglibrary_amd/cpp) on the
reference's eigenray_basic scenario (waveq3d/test/eigenray_test.cc:86-198):
listener calls, check_eigenrays per step, snapshots, the classic ray model
and the rejection of an unknown spreading model, compared with the oracle run of the same inputs.
The CPU case links the program against the emulation build; the GPU case runs
the program build() linked against libwq3d.so."""
import os
import subprocess

import numpy as np
import pytest

import emulation
from underseamodelinglibrary_amd import build, scenarios


def _parse(out):
    rays, checks, classic, total, seq = [], None, None, None, None
    unknown = None
    for line in out.splitlines():
        f = line.split()
        if f[0] == "total":
            total = {"n": int(f[1]), "time": float(f[2]), "i0": float(f[3]), "i2": float(f[4]),
                     "p0": float(f[5]), "sde": float(f[6]), "srf": int(f[7]), "btm": int(f[8]),
                     "initial": float(f[9]), "id": int(f[10])}
        elif f[0] == "eigenray":
            rays.append({"t1": int(f[1]), "t2": int(f[2]), "run": int(f[3]), "time": float(f[4]),
                         "i0": float(f[5]), "p0": float(f[6]), "sde": float(f[7]),
                         "tde": float(f[8]), "srf": int(f[11]), "btm": int(f[12])})
        elif f[0] == "checks":
            checks = int(f[1])
        elif f[0] == "classic":
            classic = int(f[1])
        elif f[0] == "unknown":
            unknown = f[1]
        elif f[0] == "seq":
            seq = [float(x) for x in f[1:]]
    return rays, checks, (classic, unknown), total, seq


def _check(out, oracle_mod):
    rays, checks, (classic, unknown), total, seq = _parse(out)
    # the classic ray model runs (3 eigenrays on this scenario); an unknown
    # spreading type is refused with std::invalid_argument
    assert classic == 3 and unknown == "rejected"
    # DE fan -60:5:60: index -1 and 1000 both clamp to the last element (60),
    # index 0 is the first (-60)
    assert seq == [60.0, 5.0, 60.0, -60.0]
    sc = scenarios.eigenray_basic()
    assert checks == sc.steps
    o = oracle_mod.oracle_queue(*sc.args())
    o.step(sc.steps)
    orr, oi, op = o.drain_eigenrays()
    assert len(rays) == len(orr) == 3
    for k, r in enumerate(rays):
        assert r["run"] == 7
        assert abs(r["time"] - orr["travel_time"][k]) <= 1e-9 * orr["travel_time"][k]
        assert abs(r["i0"] - oi[k * 3]) <= 0.01
        assert abs(r["p0"] - op[k * 3]) <= 1e-9
        assert abs(r["sde"] - orr["source_de"][k]) <= 1e-6
        assert (r["srf"], r["btm"]) == (int(orr["surface"][k]), int(orr["bottom"][k]))
    # the C++ eigenray_collection's coherent total against the oracle's sum
    # of the oracle's eigenrays (eigenray_collection.cc:105-191)
    from underseamodelinglibrary_amd.usml.eigenrays import sum_eigenray_arrays
    freq = np.array([sc.freq(k) for k in range(sc.freq.size())])
    st, tot, ti, tp = sum_eigenray_arrays(oracle_mod.load().orc_sum_eigenrays, orr, oi, op, freq, 1, 1, True)
    assert st == 0
    assert total["n"] == 3 and total["id"] == 42
    assert abs(total["time"] - tot["travel_time"][0]) <= 1e-9 * tot["travel_time"][0]
    assert abs(total["i0"] - ti[0]) <= 0.01 and abs(total["i2"] - ti[2]) <= 0.01
    assert abs(total["p0"] - tp[0]) <= 1e-6
    assert abs(total["sde"] - tot["source_de"][0]) <= 1e-6
    assert (total["srf"], total["btm"]) == (int(tot["surface"][0]), int(tot["bottom"][0]))
    assert total["initial"] == min(r["time"] for r in rays)
    # the reference's own assertions on the direct and surface paths
    assert abs(rays[0]["time"] - 1.484018789) < 0.002
    assert abs(rays[1]["time"] - 1.995102731) < 0.002
    assert abs(rays[1]["p0"] + np.pi) < 1e-6


def test_cpp_wave_queue_emulated(oracle_mod, tmp_path):
    build.build_emulation()
    asan = emulation.asan_runtime()
    exe = build.build_cpp_test(lib=build.EMU_LIB, out=str(tmp_path / "wq_cpp"), extra=[asan])
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    res = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-3000:]
    _check(res.stdout, oracle_mod)


@pytest.mark.gpu
def test_cpp_wave_queue_gpu(oracle_mod):
    assert os.path.exists(build.CPP_TEST), "run __graft_entry__.build() first"
    print("[cpp] wave_queue eigenray_basic on the GPU", flush=True)
    res = subprocess.run([build.CPP_TEST], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-3000:]
    _check(res.stdout, oracle_mod)
