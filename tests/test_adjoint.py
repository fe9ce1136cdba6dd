"""The adjoint of eigenray travel time and transmission loss with respect
to the node speeds of a 1-D sound speed grid (include/wq3d.h,
wq3d_adjoint_*), against central differences of the oracle (see
adjoint_case.py), on the CPU emulation build (AddressSanitizer, every
pointer a launch hands its kernel checked) and on the device."""
import ctypes as C
import textwrap

import numpy as np
import pytest

import adjoint_case as case
import emulation

COMMON = textwrap.dedent("""
    import ctypes as C
    import numpy as np
    import emulation, oracle
    import adjoint_case as case
    from underseamodelinglibrary_amd import scenarios
    from underseamodelinglibrary_amd.usml.types import seq_linear
    from underseamodelinglibrary_amd.usml.waveq3d import wave_queue
    lib = emulation.load()
    oracle.build(); oracle.load()
""")


def _run(body, timeout=900):
    res = emulation.run_script_under_asan(COMMON + textwrap.dedent(body), timeout=timeout)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-4000:]
    return res.stdout


@pytest.fixture(scope="module")
def emulation_build():
    from underseamodelinglibrary_amd import build
    build.build_emulation()


def test_step_adjoint_matches_tape(emulation_build):
    """The hand-written adjoint of one ray's AB3 step, update() and Thorp
    absorption equals the same step recorded on a tape, at mid-run states of
    rays across the fan (input and node adjoints, relative 1e-12)."""
    out = _run("""
        lib.wq3d_emu_check_step_vjp.restype = C.c_double
        lib.wq3d_emu_check_step_vjp.argtypes = [C.c_void_p, C.c_int64, C.c_uint64]
        sc = scenarios.config5(0, de=seq_linear(-45.0, 3.0, 45.0), az=seq_linear(0.0, 30.0, 330.0),
                               n_targets=(5, 5), time_max=20.0)
        g = wave_queue(*sc.args(), lib=lib)
        worst = 0.0
        for n in (1, 57, 300):
            g._lib.check(g._lib.wq3d_step(g._fan, n))
            for i in (0, 5, 100, 200, 371):
                worst = max(worst, lib.wq3d_emu_check_step_vjp(g._fan, i, 977 * i + n))
        print("worst", worst)
    """)
    assert float(out.split("worst")[-1]) < 1e-12


def test_adjoint_matches_central_differences_emulated(emulation_build):
    """Adjoint gradient of a loss over travel times and intensities (64 node
    speeds, reflections, four checkpoint segments) against central
    differences of the oracle: within 1e-4 of the largest derivative."""
    out = _run("""
        sp = case.base_speeds()
        sc = case.make(sp)
        g = wave_queue(*sc.args(), lib=lib)
        recs, inten = case.run_forward(g, sc.steps)
        nf = sc.freq.size()
        wt, wi = case.weights(len(recs), nf)
        grad = g.adjoint_gradient(recs, inten, wt, wi, len(sp))
        fd = case.central_differences(oracle, sp, recs, wt, wi)
        print("eigenrays", len(recs), "err", case.relative_error(grad, fd))
    """)
    n = int(out.split("eigenrays")[-1].split()[0])
    err = float(out.split("err")[-1])
    assert n > 20
    assert err <= case.TOL, err


def test_adjoint_argument_checks(emulation_build):
    """Unsupported environments, a missing first checkpoint, a first
    checkpoint after stepping and records of another run are refused."""
    _run("""
        from underseamodelinglibrary_amd import _abi, _lib as L
        def status(fn):
            try:
                fn()
            except L.WQ3DError as e:
                return e.status
            return 0
        munk = scenarios.config1(de=seq_linear(-10.0, 5.0, 10.0), az=seq_linear(0.0, 90.0, 270.0),
                                 n_targets=(2, 2), time_max=1.0)
        q = wave_queue(*munk.args(), lib=lib)
        assert status(q.adjoint_checkpoint) == _abi.WQ3D_ERR_UNSUPPORTED
        sp = case.base_speeds()
        sc = case.make(sp)
        q = wave_queue(*sc.args(), lib=lib)
        g = np.zeros(len(sp))
        assert status(lambda: q.adjoint_gradient(q.drain_eigenrays()[0], np.zeros(0), np.zeros(0),
                                                 np.zeros(0), len(sp))) == _abi.WQ3D_ERR_STATE
        q._lib.check(q._lib.wq3d_step(q._fan, 3))
        assert status(q.adjoint_checkpoint) == _abi.WQ3D_ERR_STATE
        q = wave_queue(*sc.args(), lib=lib)
        recs, inten = case.run_forward(q, sc.steps)
        nf = sc.freq.size()
        wt, wi = case.weights(len(recs), nf)
        assert status(lambda: q.adjoint_gradient(recs, inten, wt, wi, len(sp) - 1)) == _abi.WQ3D_ERR_ARGUMENT
        bad = recs.copy()
        bad["travel_time"][0] += 1e-9
        assert status(lambda: q.adjoint_gradient(bad, inten, wt, wi, len(sp))) == _abi.WQ3D_ERR_STATE
        print("ok")
    """)


@pytest.mark.gpu
def test_adjoint_matches_central_differences_device(oracle_mod):
    """The device adjoint against central differences of the oracle, the
    same case and bar as the emulated check."""
    from underseamodelinglibrary_amd.usml.waveq3d import wave_queue
    sp = case.base_speeds()
    sc = case.make(sp)
    g = wave_queue(*sc.args())
    recs, inten = case.run_forward(g, sc.steps)
    nf = sc.freq.size()
    wt, wi = case.weights(len(recs), nf)
    grad = g.adjoint_gradient(recs, inten, wt, wi, len(sp))
    fd = case.central_differences(oracle_mod, sp, recs, wt, wi)
    err = case.relative_error(grad, fd)
    print(f"device adjoint: {len(recs)} eigenrays, error {err:.2e}", flush=True)
    assert len(recs) > 20
    assert err <= case.TOL, err
    # a second gradient on the same run (workspace reused) gives the same
    # result up to the order of atomic sums
    again = g.adjoint_gradient(recs, inten, wt, wi, len(sp))
    assert np.allclose(again, grad, rtol=1e-9, atol=1e-12 * np.abs(grad).max())


@pytest.mark.gpu
def test_autograd_function_batch(oracle_mod):
    """torch.autograd through eigenray_function over three perturbed
    profiles stepped as one fan_batch: each profile's gradient against
    central differences of the oracle."""
    import torch
    from underseamodelinglibrary_amd.usml.adjoint import eigenray_function, profile_runs
    base = case.base_speeds()
    rng = np.random.default_rng(7)
    speeds = np.stack([base + rng.normal(0.0, 0.5, size=base.size) for _ in range(3)])
    runs = profile_runs(case.make, checkpoint_every=case.CHECKPOINT, batched=True)
    x = torch.tensor(speeds, dtype=torch.float64, requires_grad=True)
    tt, ti = eigenray_function.apply(x, runs)
    counts = runs.counts()
    assert all(c > 20 for c in counts)
    nf = ti.shape[1]
    wt, wi = case.weights(len(tt), nf, seed=3)
    loss = (torch.from_numpy(wt) * tt).sum() + (torch.from_numpy(wi) * ti).sum()
    loss.backward()
    at = 0
    for p in range(3):
        n = counts[p]
        recs = runs.records[p][0]
        fd = case.central_differences(oracle_mod, speeds[p], recs, wt[at:at + n], wi[at:at + n])
        err = case.relative_error(x.grad[p].numpy(), fd)
        print(f"profile {p}: {n} eigenrays, error {err:.2e}", flush=True)
        assert err <= case.TOL, (p, err)
        at += n
    runs.close()
