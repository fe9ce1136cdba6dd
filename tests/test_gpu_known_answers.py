"""The reference's known-answer tests run on the HIP library itself, through
the C ABI, with the same assertions the oracle meets
(test_oracle_known_answers.py), and the device's records against the
oracle's on the same inputs:

* Lloyd's mirror in range and in depth (waveq3d/test/proploss_test.cc:
  357-560, :562-724), totals summed on the device by wq3d_sum_eigenrays;
* hybrid Gaussian transmission loss against spherical spreading
  (proploss_test.cc:153-245);
* the phase bookkeeping of surface and bottom bounces and caustics in the
  Munk channel (reflection_model.cc:155-160, :230-235, wave_queue.cc:
  285-289);
* Munk cycle ranges and vertex counts on a flat earth (refraction_test.cc:
  827-955), which exercises the flat-earth sound speed correction
  (ocean/profile_model.cc:14-26) on the device;
* rays bouncing between a flat surface and a flat bottom following their
  image-source chains of straight lines (reflection_model.cc:33-360)."""
import pytest

from underseamodelinglibrary_amd import _lib, scenarios
from underseamodelinglibrary_amd.usml.waveq3d import wave_queue

import test_oracle_known_answers as ka
import test_oracle_pins as pins
from parity import compare_eigenrays

pytestmark = pytest.mark.gpu


def _device(sc):
    print(f"[known answer] {sc.name}: {sc.num_rays()} rays, {sc.num_targets()} targets, "
          f"{sc.steps} steps", flush=True)
    g = wave_queue(*sc.args())
    g._lib.check(g._lib.wq3d_step(g._fan, sc.steps))
    return g


def _oracle(sc, oracle_mod):
    o = oracle_mod.oracle_queue(*sc.args())
    o._lib.orc_step(o._h, sc.steps)
    return o


@pytest.mark.parametrize("case", ["range", "depth"])
def test_lloyds_mirror_gpu(oracle_mod, case):
    sc = scenarios.lloyds_range() if case == "range" else scenarios.lloyds_depth()
    g = _device(sc)
    recs, inten, phase = g.drain_eigenrays()
    stats = ka.lloyds_check(sc, recs, inten, phase, _lib.load().wq3d_sum_eigenrays,
                            0.5 if case == "range" else 0.7)
    print(f"[known answer] (freq, bias, dev, detcoef) {stats}", flush=True)
    sc = scenarios.lloyds_range() if case == "range" else scenarios.lloyds_depth()
    orr, oi, op = _oracle(sc, oracle_mod).drain_eigenrays()
    assert compare_eigenrays(recs, inten, phase, orr, oi, op, sc.freq.size()) > 100


def test_proploss_gaussian_gpu(oracle_mod):
    sc = scenarios.proploss_gaussian()
    recs, inten, phase = _device(sc).drain_eigenrays()
    ka.proploss_gaussian_check(recs, inten)
    sc = scenarios.proploss_gaussian()
    orr, oi, op = _oracle(sc, oracle_mod).drain_eigenrays()
    assert compare_eigenrays(recs, inten, phase, orr, oi, op, sc.freq.size()) == 3


def test_caustic_and_reflection_phase_gpu():
    sc = ka.caustic_scenario()
    print(f"[known answer] {sc.name}: {sc.num_rays()} rays, {sc.steps} steps", flush=True)
    g = wave_queue(*sc.args())
    ka.caustic_phase_check(sc, g, lambda n: g._lib.check(g._lib.wq3d_step(g._fan, n)))


def test_refraction_munk_range_flat_earth_gpu():
    print("[known answer] refraction_munk_range on a flat earth: 29 rays, 950 steps", flush=True)
    ka.munk_range_check(lambda *a: wave_queue(*a))


def test_reflection_reinit_image_source_gpu():
    print("[known answer] image-source chains: 21 rays, 120 steps", flush=True)
    pins.image_source_check(lambda *a: wave_queue(*a))


@pytest.mark.gpu
def test_de_offset_sign_straight_rays_gpu(oracle_mod):
    """The straight-ray DE sign check of test_oracle_pins on the device's
    eigenrays, which also equal the oracle's."""
    from test_oracle_pins import check_de_offsets_point_to_targets
    from parity import compare_eigenrays
    from underseamodelinglibrary_amd import scenarios
    from underseamodelinglibrary_amd.usml.waveq3d import wave_queue
    sc = scenarios.straight_ray_targets()
    g = wave_queue(*sc.args())
    o = oracle_mod.oracle_queue(*sc.args())
    g.step(sc.steps)
    o.step(sc.steps)
    gr, gi, gp = g.drain_eigenrays()
    orr, oi, op = o.drain_eigenrays()
    compare_eigenrays(gr, gi, gp, orr, oi, op, sc.freq.size())
    assert check_de_offsets_point_to_targets(sc, gr) > 100


@pytest.mark.gpu
def test_caustic_fold_test_matches_turning_points_gpu(oracle_mod):
    """The caustic / turning-point relation of test_oracle_pins on the
    device's wavefront, whose counters also equal the oracle's."""
    import numpy as np
    from test_oracle_pins import check_caustics_are_turning_points
    from underseamodelinglibrary_amd import _abi
    from underseamodelinglibrary_amd.usml.types import seq_linear
    from underseamodelinglibrary_amd.usml.waveq3d import wave_queue
    sc = scenarios.config1(de=seq_linear(-20.0, 0.25, 20.0), az=seq_linear(0.0, 90.0, 270.0), targets=False,
                           time_max=120.0)
    g = wave_queue(*sc.args())
    o = oracle_mod.oracle_queue(*sc.args())
    g.step(sc.steps)
    o.step(sc.steps)
    gf, of = g.snapshot_arrays(_abi.CURR), o.snapshot_arrays(_abi.CURR)
    for k in ("surface", "bottom", "caustic", "upper", "lower"):
        assert np.array_equal(np.asarray(gf[k]), np.asarray(of[k])), k
    assert check_caustics_are_turning_points(sc, gf) > 400
