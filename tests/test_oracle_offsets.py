"""compute_offsets (waveq3d/wave_queue.cc:708-924, make_taylor_coeff :929-977)
pinned on neighbourhoods sampled from exact quadratics.

On a 3 x 3 x 3 neighbourhood of a quadratic d2(x) = c + g.x + x'Hx/2 (x the
offsets in time, DE and AZ), central differences give g and H exactly, so
what compute_offsets must return follows from the mathematics:

* stable branch: the minimum x* = -H^-1 g, scaled by det(H) / det'(H) where
  det' is the reference's determinant with its typo (the last cofactor uses
  H21 H20 for H11 H20), so det' = det + H02 H20 (H11 - H21) - with no
  time-AZ coupling (H02 = 0) exactly the minimum;
* a DE curvature below zero: unstable, DE offset half a DE step toward
  -g1, the time and AZ offsets the minimum of the quadratic with DE fixed
  (the 2 x 2 solve), the DE distance the remainder of the centre distance;
* a gradient too large for its curvature: the diagonal estimates clipped to
  half a step, the 2 x 2 solve skipped when its determinant is below |g|;
* per-axis distances sqrt(-g_n x_n - H_nn x_n^2 / 2) signed like x_n.

The device's compute_offsets (CPU emulation build) equals the oracle's bit
for bit on random neighbourhoods that reach every branch."""
import ctypes as C

import numpy as np

import emulation

DELTA = np.array([0.1, 0.25, 2.0])


def cube(c, g, H, delta=DELTA):
    d = np.empty(27)
    for i in range(3):
        for j in range(3):
            for k in range(3):
                x = np.array([(i - 1) * delta[0], (j - 1) * delta[1], (k - 1) * delta[2]])
                d[i * 9 + j * 3 + k] = c + g @ x + 0.5 * x @ H @ x
    return d


def core(fn, d27, delta=DELTA, counters_differ=False):
    dp = C.POINTER(C.c_double)
    off = np.zeros(3)
    dist = np.zeros(3)
    d27 = np.ascontiguousarray(d27, dtype=np.float64)
    delta = np.ascontiguousarray(delta, dtype=np.float64)
    u = fn(d27.ctypes.data_as(dp), delta.ctypes.data_as(dp), 1 if counters_differ else 0,
           off.ctypes.data_as(dp), dist.ctypes.data_as(dp))
    return bool(u), off, dist


def axis_distance(g, H, x):
    d = np.sqrt(np.maximum(0.0, -g * x - 0.5 * np.diag(H) * x * x))
    return np.where(x < 0.0, -d, d)


def spd(rng, couple_t_az):
    """A positive definite Hessian with large entries (the determinant well
    above |g|), optionally with time-AZ coupling."""
    H = np.diag(rng.uniform(2e3, 8e3, 3))
    H[0, 1] = H[1, 0] = rng.uniform(-300.0, 300.0)
    H[1, 2] = H[2, 1] = rng.uniform(-300.0, 300.0)
    if couple_t_az:
        H[0, 2] = H[2, 0] = rng.uniform(-500.0, 500.0)
    return H


def test_stable_branch_recovers_the_minimum(oracle_mod):
    fn = oracle_mod.load().orc_compute_offsets_core
    rng = np.random.default_rng(3)
    for _ in range(200):
        H = spd(rng, couple_t_az=False)
        xs = rng.uniform(-0.3, 0.3, 3) * DELTA
        g = -H @ xs
        c = 1e6
        unstable, off, dist = core(fn, cube(c, g, H), DELTA)
        assert not unstable
        assert np.allclose(off, xs, rtol=1e-9, atol=1e-12 * DELTA.max())
        assert np.allclose(dist, axis_distance(g, H, xs), rtol=1e-6, atol=1e-6)


def test_stable_branch_keeps_the_determinant_typo(oracle_mod):
    fn = oracle_mod.load().orc_compute_offsets_core
    rng = np.random.default_rng(4)
    seen = 0
    for _ in range(300):
        H = spd(rng, couple_t_az=True)
        xs = rng.uniform(-0.2, 0.2, 3) * DELTA
        g = -H @ xs
        unstable, off, _ = core(fn, cube(1e6, g, H), DELTA)
        if unstable:
            continue
        det = np.linalg.det(H)
        det_typo = det + H[0, 2] * H[2, 0] * (H[1, 1] - H[2, 1])
        assert np.allclose(off, xs * det / det_typo, rtol=1e-8, atol=1e-13)
        seen += 1
    assert seen > 200


def test_negative_de_curvature_takes_the_time_az_solve(oracle_mod):
    fn = oracle_mod.load().orc_compute_offsets_core
    rng = np.random.default_rng(5)
    for _ in range(200):
        H = spd(rng, couple_t_az=True)
        H[1, 1] = -rng.uniform(100.0, 5000.0)
        g = rng.uniform(-50.0, 50.0, 3)
        c = 1e6
        unstable, off, dist = core(fn, cube(c, g, H), DELTA)
        assert unstable
        two = np.linalg.solve(np.array([[H[0, 0], H[0, 2]], [H[2, 0], H[2, 2]]]), -np.array([g[0], g[2]]))
        assert np.allclose([off[0], off[2]], two, rtol=1e-8, atol=1e-13)
        # DE: half a step (its sign is set afterwards from the rays' geometry)
        assert abs(off[1]) == 0.5 * DELTA[1]
        x = np.array([off[0], 0.0, off[2]])
        d0, d2 = axis_distance(g, H, x)[[0, 2]]
        assert np.allclose([dist[0], dist[2]], [d0, d2], rtol=1e-8)
        assert np.isclose(dist[1], np.sqrt(c - d0 * d0 - d2 * d2), rtol=1e-8)


def test_large_gradient_clips_to_half_a_step(oracle_mod):
    fn = oracle_mod.load().orc_compute_offsets_core
    rng = np.random.default_rng(6)
    for _ in range(200):
        H = np.diag(rng.uniform(0.5, 2.0, 3))
        g = rng.choice([-1.0, 1.0], 3) * rng.uniform(1e3, 1e4, 3)
        unstable, off, _ = core(fn, cube(1e9, g, H), DELTA)
        assert unstable
        # every diagonal estimate -g/H is far outside half a step; the
        # time-AZ determinant (about 1) is below |g|, so nothing overrides them
        assert np.allclose(np.abs(off[[0, 2]]), 0.5 * DELTA[[0, 2]])
        assert (np.sign(off[[0, 2]]) == -np.sign(g[[0, 2]])).all()
        assert abs(off[1]) == 0.5 * DELTA[1]


def test_counter_change_forces_the_unstable_branch(oracle_mod):
    fn = oracle_mod.load().orc_compute_offsets_core
    rng = np.random.default_rng(7)
    H = spd(rng, couple_t_az=False)
    xs = np.array([0.1, -0.2, 0.15]) * DELTA
    g = -H @ xs
    u0, off0, _ = core(fn, cube(1e6, g, H), DELTA, counters_differ=False)
    u1, off1, _ = core(fn, cube(1e6, g, H), DELTA, counters_differ=True)
    assert not u0 and u1
    two = np.linalg.solve(np.array([[H[0, 0], H[0, 2]], [H[2, 0], H[2, 2]]]), -np.array([g[0], g[2]]))
    assert np.allclose([off1[0], off1[2]], two, rtol=1e-8)


def test_device_offsets_equal_oracle():
    """The kernels' compute_offsets core against the oracle's on 20,000
    random neighbourhoods (quadratic, perturbed and raw), bit for bit."""
    res = emulation.run_script_under_asan("""
import ctypes as C, numpy as np, oracle, emulation
import test_oracle_offsets as t
lib = emulation.load()
fd = lib.wq3d_emu_compute_offsets_core
dp = C.POINTER(C.c_double)
fd.argtypes = [dp, dp, C.c_int32, dp, dp]
fd.restype = C.c_int32
fo = oracle.load().orc_compute_offsets_core
rng = np.random.default_rng(11)
bad = 0
branches = set()
for n in range(20000):
    kind = n % 4
    H = t.spd(rng, couple_t_az=bool(n & 1))
    if kind == 1:
        H[1, 1] = -H[1, 1]
    g = rng.normal(0.0, [30.0, 300.0, 3e4][n % 3], 3)
    d = t.cube(rng.uniform(1e4, 1e8), g, H)
    if kind == 2:
        d = d * (1.0 + rng.normal(0.0, 1e-3, 27))
    if kind == 3:
        d = rng.uniform(0.0, 1e8, 27)
    delta = t.DELTA * rng.uniform(0.5, 2.0, 3)
    cd = bool(rng.uniform() < 0.1)
    a = t.core(fd, d, delta, cd)
    b = t.core(fo, d, delta, cd)
    branches.add((a[0], abs(a[1][1]) == 0.5 * delta[1]))
    if a[0] != b[0] or a[1].tobytes() != b[1].tobytes() or a[2].tobytes() != b[2].tobytes():
        bad += 1
print("result", bad, len(branches))
""")
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-3000:]
    bad, branches = (int(x) for x in res.stdout.split("result")[-1].split())
    assert bad == 0 and branches >= 3
