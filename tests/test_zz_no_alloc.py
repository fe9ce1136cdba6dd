"""wq3d_sum_eigenrays_noalloc, the allocation-free form of the eigenray
collection's per-target sums, against wq3d_sum_eigenrays through the C ABI:
bit-identical totals, intensities and phases on two calls in a row (the
second smaller, so the shared device array holds the first call's data),
and a clear capacity error for records larger than the array."""
import ctypes as C

import numpy as np
import pytest

import emulation
from underseamodelinglibrary_amd import _abi
from underseamodelinglibrary_amd.usml.eigenrays import sum_eigenray_arrays

_REC = _abi.record_dtype(_abi.EigenrayRec)


def _records(seed, n, n1, n2, nf):
    """Synthetic eigenrays in arrival order: random targets (some targets
    left empty), travel times, angles, counters, intensities and phases."""
    rng = np.random.default_rng(seed)
    recs = np.zeros(n, dtype=_REC)
    used = rng.choice(n1 * n2, size=max(1, (n1 * n2 * 3) // 4), replace=False)
    tgt = used[rng.integers(0, used.size, n)]
    recs["t1"], recs["t2"] = tgt // n2, tgt % n2
    recs["step"] = np.sort(rng.integers(1, 600, n))
    recs["travel_time"] = rng.uniform(0.5, 60.0, n)
    for k in ("source_de", "target_de"):
        recs[k] = rng.uniform(-80.0, 80.0, n)
    for k in ("source_az", "target_az"):
        recs[k] = rng.uniform(0.0, 360.0, n)
    for k in ("surface", "bottom", "caustic", "upper", "lower"):
        recs[k] = rng.integers(0, 4, n)
    inten = rng.uniform(40.0, 120.0, n * nf)
    phase = rng.uniform(-np.pi, np.pi, n * nf)
    freq = np.geomspace(100.0, 10000.0, nf)
    return recs, inten, phase, freq


def _compare(lib, cases):
    for seed, n, n1, n2, nf, coherent in cases:
        recs, inten, phase, freq = _records(seed, n, n1, n2, nf)
        print(f"[noalloc] seed {seed}: {n} eigenrays, {n1}x{n2} targets, {nf} freqs", flush=True)
        ref = sum_eigenray_arrays(lib.wq3d_sum_eigenrays, recs, inten, phase, freq, n1, n2, coherent)
        got = sum_eigenray_arrays(lib.wq3d_sum_eigenrays_noalloc, recs, inten, phase, freq, n1, n2, coherent)
        assert ref[0] == 0 and got[0] == 0, lib.wq3d_last_error()
        assert ref[1].tobytes() == got[1].tobytes()
        assert ref[2].tobytes() == got[2].tobytes()
        assert ref[3].tobytes() == got[3].tobytes()
        assert (got[1]["count"] == np.bincount(recs["t1"] * n2 + recs["t2"], minlength=n1 * n2)).all()


CASES = [(11, 40000, 100, 100, 8, True), (12, 3000, 10, 10, 3, False)]


def _capacity_error(lib):
    # 600,000 eigenrays of 32 frequencies need about 370 MB of device memory
    recs, inten, phase, freq = _records(13, 600000, 50, 50, 32)
    st = sum_eigenray_arrays(lib.wq3d_sum_eigenrays_noalloc, recs, inten, phase, freq, 50, 50, True)[0]
    assert st == _abi.WQ3D_ERR_CAPACITY
    msg = lib.wq3d_last_error().decode()
    assert "allocation-free" in msg and str(128 << 20) in msg


def test_noalloc_sum_matches_emulated():
    """CPU emulation build (AddressSanitizer, device memory checked against
    the emulated allocations, the scratch array registered as one): the
    device-side list building (count, scan, scatter, per-target sort) and the
    sums equal the host-sorted entry point bit for bit."""
    import os
    mod = os.path.splitext(os.path.basename(__file__))[0]
    res = emulation.run_script_under_asan(
        f"import emulation, {mod} as t\nlib = emulation.load()\n"
        f"t._compare(lib, t.CASES)\nt._capacity_error(lib)\nprint('noalloc ok')\n")
    assert res.returncode == 0 and "noalloc ok" in res.stdout, res.stdout[-2000:] + res.stderr[-4000:]


@pytest.mark.gpu
def test_noalloc_sum_matches_gpu():
    from underseamodelinglibrary_amd import _lib
    lib = _lib.load()
    assert lib.wq3d_device_count() >= 1
    _compare(lib, CASES)
    _capacity_error(lib)
