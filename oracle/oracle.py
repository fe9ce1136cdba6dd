"""Python binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product package.  oracle_queue has the same interface as
underseamodelinglibrary_amd.usml.waveq3d.wave_queue so tests can run both on
one scenario and compare records and snapshots.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from underseamodelinglibrary_amd import _abi
from underseamodelinglibrary_amd.usml.types import wposition1
from underseamodelinglibrary_amd.usml.waveq3d import front_buffers, queue_base, wave_front

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libwq3d_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, i64, dp = C.c_void_p, C.c_int64, C.POINTER(C.c_double)
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [C.POINTER(_abi.EnvDesc), C.POINTER(_abi.FanDesc)]
    lib.orc_last_error.restype = C.c_char_p
    lib.orc_destroy.argtypes = [vp]
    lib.orc_step.argtypes = [vp, i64]
    lib.orc_set_thresholds.argtypes = [vp, C.POINTER(_abi.Thresholds)]
    lib.orc_set_listeners.argtypes = [vp, C.c_int32, C.c_int32]
    lib.orc_time.restype = C.c_double
    lib.orc_time.argtypes = [vp]
    lib.orc_source.argtypes = [vp, dp, dp, dp]
    for n in ("orc_num_eigenrays", "orc_num_eigenverbs", "orc_num_reflections"):
        getattr(lib, n).restype = i64
        getattr(lib, n).argtypes = [vp]
    lib.orc_take_eigenrays.argtypes = [vp, C.POINTER(_abi.EigenrayRec), dp, dp]
    lib.orc_sign_rule_counts.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    lib.orc_sign_rule_counts.restype = None
    lib.orc_trace_walks.argtypes = [vp, C.c_int32]
    lib.orc_trace_walks.restype = None
    lib.orc_walk_trace.argtypes = [vp, dp, i64]
    lib.orc_walk_trace.restype = i64
    lib.orc_take_eigenverbs.argtypes = [vp, C.POINTER(_abi.EigenverbRec), dp]
    lib.orc_take_reflections.argtypes = [vp, C.POINTER(_abi.ReflectionRec)]
    lib.orc_snapshot.argtypes = [vp, C.c_int32, C.POINTER(_abi.FrontView)]
    lib.orc_math.argtypes = [C.c_int32, dp, dp, dp, i64]
    lib.orc_sum_eigenrays.restype = C.c_int32
    lib.orc_sum_eigenrays.argtypes = [C.c_void_p, dp, dp, i64, dp, C.c_int32, i64, i64, C.c_int32,
                                      C.POINTER(_abi.EigenrayTotal), dp, dp]
    lib.orc_compute_offsets_core.restype = C.c_int32
    lib.orc_compute_offsets_core.argtypes = [dp, dp, C.c_int32, dp, dp]
    lib.orc_grid_interp.restype = C.c_double
    lib.orc_grid_interp.argtypes = [C.POINTER(_abi.Grid), dp, dp]
    lib.orc_sound_speed.argtypes = [C.POINTER(_abi.ProfileDesc), C.c_double, C.c_double,
                                    C.c_double, C.c_double, dp]
    lib.orc_attenuation.argtypes = [C.POINTER(_abi.ProfileDesc), C.POINTER(_abi.Seq),
                                    C.c_double, C.c_double, dp]
    lib.orc_reflect_loss.argtypes = [C.POINTER(_abi.BoundaryDesc), C.c_double, dp, dp]
    lib.orc_height.argtypes = [C.POINTER(_abi.BoundaryDesc), C.c_double, C.c_double,
                               C.c_double, dp]
    _lib = lib
    return lib


class oracle_queue(queue_base):
    """wave_queue semantics evaluated by the CPU oracle."""

    def __init__(self, ocean, freq, pos, de, az, time_step, target_pos=None,
                 spreading=_abi.SPREADING_HYBRID_GAUSSIAN):
        super().__init__(ocean, freq, pos, de, az, time_step, target_pos, spreading)
        self._lib = load()
        self._h = self._lib.orc_create(C.byref(self._env), C.byref(self._fan_desc))
        if not self._h:
            raise ValueError(self._lib.orc_last_error().decode())
        r, t, p = C.c_double(), C.c_double(), C.c_double()
        self._lib.orc_source(self._h, C.byref(r), C.byref(t), C.byref(p))
        self._source = wposition1.from_rtp(r.value, t.value, p.value)
        self._steps = 0

    def __del__(self):
        self.close_netcdf()
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def source_pos(self):
        return self._source

    def time(self):
        return self._lib.orc_time(self._h)

    def _push_thresholds(self):
        self._lib.orc_set_thresholds(self._h, C.byref(self._thr))

    def _push_listeners(self):
        self._lib.orc_set_listeners(self._h, 1 if self._verb_listeners else 0,
                                    1 if self._refl_listeners else 0)

    def step(self, n=1):
        for _ in range(int(n)):
            self._lib.orc_step(self._h, 1)
            self._steps += 1
            if self._ray_listeners or self._verb_listeners or self._refl_listeners:
                rays, inten, phase = self.drain_eigenrays()
                verbs, power = self.drain_eigenverbs()
                refls = self.drain_reflections()
                vm = self._eigenverb_models(verbs, power)
                for (typ, v), rec in zip(vm, verbs):
                    v._seq = int(rec["seq"])
                self._deliver(self.time(), self._eigenray_models(rays, inten, phase), vm, refls)

    def drain_eigenrays(self):
        n = self._lib.orc_num_eigenrays(self._h)
        nf = self._freq.size()
        recs = np.zeros(n, dtype=_abi.record_dtype(_abi.EigenrayRec))
        inten = np.zeros(n * nf)
        phase = np.zeros(n * nf)
        self._lib.orc_take_eigenrays(self._h, recs.ctypes.data_as(C.POINTER(_abi.EigenrayRec)),
                                     inten.ctypes.data_as(C.POINTER(C.c_double)),
                                     phase.ctypes.data_as(C.POINTER(C.c_double)))
        return recs, inten, phase

    def drain_eigenverbs(self):
        n = self._lib.orc_num_eigenverbs(self._h)
        nf = self._freq.size()
        recs = np.zeros(n, dtype=_abi.record_dtype(_abi.EigenverbRec))
        power = np.zeros(n * nf)
        self._lib.orc_take_eigenverbs(self._h, recs.ctypes.data_as(C.POINTER(_abi.EigenverbRec)),
                                      power.ctypes.data_as(C.POINTER(C.c_double)))
        return recs, power

    def drain_reflections(self):
        n = self._lib.orc_num_reflections(self._h)
        recs = np.zeros(n, dtype=_abi.record_dtype(_abi.ReflectionRec))
        self._lib.orc_take_reflections(self._h, recs.ctypes.data_as(C.POINTER(_abi.ReflectionRec)))
        return recs

    def snapshot_arrays(self, slot, fields=None):
        arrays, view = front_buffers(self.num_de() * self.num_az(), self._freq.size(), fields)
        self._lib.orc_snapshot(self._h, slot, C.byref(view))
        return arrays

    def _front(self, slot):
        return wave_front(self.snapshot_arrays(slot), self.num_de(), self.num_az(),
                          self._freq.size())

    def past(self):
        return self._front(_abi.PAST)

    def prev(self):
        return self._front(_abi.PREV)

    def curr(self):
        return self._front(_abi.CURR)

    def next(self):
        return self._front(_abi.NEXT)


def grid_interp(grid, loc):
    """Oracle gen_grid interpolation of a data_grid at one location."""
    lib = load()
    g = grid.to_abi()
    loc = np.ascontiguousarray(loc, dtype=np.float64)
    deriv = np.zeros(3)
    v = lib.orc_grid_interp(C.byref(g), loc.ctypes.data_as(C.POINTER(C.c_double)),
                            deriv.ctypes.data_as(C.POINTER(C.c_double)))
    return v, deriv[:grid.ndims()]


def sound_speed(profile, rho, theta, phi, R):
    lib = load()
    d = profile.to_abi()
    out = np.zeros(4)
    lib.orc_sound_speed(C.byref(d), rho, theta, phi, R, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def attenuation(profile, freq, alt, dist):
    lib = load()
    d = profile.to_abi()
    s = freq.to_abi()
    out = np.zeros(freq.size())
    lib.orc_attenuation(C.byref(d), C.byref(s), alt, dist, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def reflect_loss(boundary, grazing):
    lib = load()
    d = boundary.to_abi()
    amp, ph = C.c_double(), C.c_double()
    lib.orc_reflect_loss(C.byref(d), grazing, C.byref(amp), C.byref(ph))
    return amp.value, ph.value


def height(boundary, rho, theta, phi):
    lib = load()
    d = boundary.to_abi()
    out = np.zeros(4)
    lib.orc_height(C.byref(d), rho, theta, phi, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def math_fn(fn, x, y=None):
    """Shared elementary function `fn` (0 sin, 1 cos, 2 exp, 3 log, 4 log10,
    5 atan2, 6 asin, 7 10^x) evaluated on the host."""
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(x if y is None else y, dtype=np.float64)
    out = np.empty_like(x)
    dp = C.POINTER(C.c_double)
    lib.orc_math(int(fn), x.ctypes.data_as(dp), y.ctypes.data_as(dp), out.ctypes.data_as(dp),
                 x.size)
    return out
