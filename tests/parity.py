"""Comparison helpers for HIP-vs-oracle parity tests.

Tolerances are those the hot path is held to: identical eigenray keys and
integer path counters (exact), travel time within 1e-9 relative, intensity
within 0.01 dB, launch/arrival angles within 1e-6 deg.
"""
import numpy as np

TIME_RTOL = 1e-9
DB_ATOL = 0.01
DEG_ATOL = 1e-6
PHASE_ATOL = 1e-9


def run_both(make_queue_gpu, make_queue_orc, steps, chunk=None):
    g = make_queue_gpu()
    o = make_queue_orc()
    g.step(steps)
    o.step(steps)
    return g, o


def compare_eigenrays(g_recs, g_int, g_ph, o_recs, o_int, o_ph, nf):
    """Assert identical eigenray sets and values within tolerance."""
    keys = ["step", "t1", "t2", "de", "az"]
    gk = [tuple(int(r[k]) for k in keys) for r in g_recs]
    ok = [tuple(int(r[k]) for k in keys) for r in o_recs]
    assert len(gk) == len(ok), f"eigenray count {len(gk)} (hip) vs {len(ok)} (oracle)"
    assert gk == ok, "eigenray keys differ: first mismatch at " + str(
        next(i for i, (a, b) in enumerate(zip(gk, ok)) if a != b))
    for name in ("surface", "bottom", "caustic", "upper", "lower"):
        np.testing.assert_array_equal(g_recs[name], o_recs[name], err_msg=name)
    if len(gk) == 0:
        return 0
    t_rel = np.abs(g_recs["travel_time"] - o_recs["travel_time"]) / np.abs(o_recs["travel_time"])
    assert t_rel.max() <= TIME_RTOL, f"travel time rel err {t_rel.max():.3e}"
    for name in ("source_de", "source_az", "target_de", "target_az"):
        err = np.abs(g_recs[name] - o_recs[name]).max()
        assert err <= DEG_ATOL, f"{name} err {err:.3e} deg"
    err = np.abs(g_int - o_int).max()
    assert err <= DB_ATOL, f"intensity err {err:.3e} dB"
    err = np.abs(g_ph - o_ph).max()
    assert err <= PHASE_ATOL, f"phase err {err:.3e}"
    return len(gk)


def compare_fronts(g_arrays, o_arrays, rtol=1e-10):
    for name, o in o_arrays.items():
        g = g_arrays[name]
        if o.dtype.kind in "iu":
            np.testing.assert_array_equal(g, o, err_msg=name)
        else:
            scale = np.maximum(np.abs(o), 1.0)
            err = (np.abs(g - o) / scale).max() if o.size else 0.0
            assert err <= rtol, f"{name}: max scaled err {err:.3e}"
