"""Writes the golden eigenray fixtures under tests/golden/ from the oracle.

Each fixture holds, for one seeded scenario run to its end time, the
eigenray records in the order the queue reports them (step, t1, t2, de, az
keys, travel time, angles, path counters) with their per-frequency
intensity and phase.  The oracle that writes them is itself pinned by the
reference's known answers (eigenrays/test/eigenrays_test.cc:57-88,
waveq3d/test/reflection_test.cc:168-280) and the checks of
tests/test_oracle_known_answers.py and tests/test_oracle_pins.py; the
fixtures freeze its output so that the oracle cannot drift unseen and so
that the device can be checked against data on disk.

    python tests/golden/make_golden.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from underseamodelinglibrary_amd import scenarios  # noqa: E402
from underseamodelinglibrary_amd.usml.types import seq_linear  # noqa: E402

# scenario name -> constructor; config 1's ocean on a thinned fan and target grid
CASES = {
    "eigenray_basic": lambda: scenarios.eigenray_basic(),
    "eigenray_branch_pt": lambda: scenarios.eigenray_branch_pt(),
    "config1_thin": lambda: scenarios.config1(de=seq_linear(-60.0, 4.0, 60.0),
                                              az=seq_linear(0.0, 30.0, 330.0),
                                              n_targets=(6, 6), time_max=30.0),
}


# The benchmark fan with all of its 10,000 targets for its first 35 steps
# (config 2's fan, frequencies and targets in config 1's Munk ocean; about
# 2.5 minutes of the oracle).  Heavy: the CPU suite re-runs it only with
# WQ3D_SLOW=1; the device checks it on every GPU run.
HEAVY_CASES = {
    "config2_fan_munk_35": lambda: _steps(scenarios.config2_analytic(), 35),
}


def _steps(sc, n):
    sc.steps = n
    return sc


CASES.update(HEAVY_CASES)


def fixture_path(name):
    return os.path.join(HERE, f"{name}.npz")


def run_oracle(name):
    import oracle
    sc = CASES[name]()
    q = oracle.oracle_queue(*sc.args())
    q._lib.orc_step(q._h, sc.steps)
    recs, inten, phase = q.drain_eigenrays()
    return sc, recs, inten, phase


def main(names=None):
    for name in (names or CASES):
        t0 = time.time()
        sc, recs, inten, phase = run_oracle(name)
        np.savez_compressed(fixture_path(name), records=recs, intensity=inten, phase=phase,
                            steps=np.int64(sc.steps), nf=np.int64(sc.freq.size()))
        print(f"{name}: {len(recs)} eigenrays, {sc.steps} steps, {time.time() - t0:.1f} s, "
              f"{os.path.getsize(fixture_path(name))} B", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
