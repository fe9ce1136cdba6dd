"""usml::waveq3d::wave_queue facade over libwq3d.so (the HIP library).

Mirrors the reference class (waveq3d/wave_queue.h:78-831): the constructor
takes ocean, frequencies, source position, D/E and AZ fans, time step and an
optional target grid; step() advances one time step; eigenray, eigenverb and
reflection listeners receive the products of each step in the reference's
call order; curr()/next()/prev()/past() return host snapshots of the four
queue slots.  Thresholds follow wave_thresholds (waveq3d/wave_thresholds.h).
"""
import ctypes as C

import numpy as np

from .. import _abi
from .types import earth_radius, seq_vector, to_degrees, to_latitude, wposition, wposition1


class eigenray_model:
    """eigenrays/eigenray_model.h:24-102."""

    __slots__ = ("travel_time", "frequencies", "intensity", "phase", "source_de", "source_az",
                 "target_de", "target_az", "surface", "bottom", "caustic", "upper", "lower")

    def __repr__(self):
        return (f"eigenray(t={self.travel_time:.9f}, de={self.source_de:.6f}, "
                f"az={self.source_az:.6f}, srf={self.surface}, btm={self.bottom}, "
                f"cst={self.caustic}, I0={self.intensity[0]:.4f})")


class eigenverb_model:
    """eigenverbs/eigenverb_model.h:34-160 (angles in radians)."""

    BOTTOM, SURFACE, VOLUME_UPPER, VOLUME_LOWER = 0, 1, 2, 3
    __slots__ = ("travel_time", "frequencies", "power", "length", "width", "position",
                 "direction", "grazing", "sound_speed", "de_index", "az_index", "source_de",
                 "source_az", "surface", "bottom", "caustic", "upper", "lower", "_seq")


class _component:
    """Callable view of one wvector/wposition component matrix."""

    def __init__(self, arr):
        self._a = arr

    def __call__(self, de=None, az=None):
        if de is None:
            return self._a
        return float(self._a[de, az])


class _wvec:
    def __init__(self, r, t, p):
        self.rho = _component(r)
        self.theta = _component(t)
        self.phi = _component(p)

    def altitude(self, de=None, az=None):
        a = self.rho() - earth_radius()
        return a if de is None else float(a[de, az])

    def latitude(self, de=None, az=None):
        a = 90.0 - self.theta() * (180.0 / np.pi)
        return a if de is None else float(a[de, az])

    def longitude(self, de=None, az=None):
        a = self.phi() * (180.0 / np.pi)
        return a if de is None else float(a[de, az])


class wave_front:
    """Host snapshot of one queue slot (waveq3d/wave_front.h:152-258)."""

    def __init__(self, arrays, nde, naz, nf):
        sh = (nde, naz)
        g = {k: v.reshape(sh) if v.size == nde * naz else v.reshape(sh + (nf,))
             for k, v in arrays.items()}
        self.position = _wvec(g["pos_rho"], g["pos_theta"], g["pos_phi"])
        self.pos_gradient = _wvec(g["pgrad_rho"], g["pgrad_theta"], g["pgrad_phi"])
        self.ndirection = _wvec(g["ndir_rho"], g["ndir_theta"], g["ndir_phi"])
        self.ndir_gradient = _wvec(g["dgrad_rho"], g["dgrad_theta"], g["dgrad_phi"])
        self.sound_gradient = _wvec(g["sgrad_rho"], g["sgrad_theta"], g["sgrad_phi"])
        self.sound_speed = _component(g["sound_speed"])
        self.distance = _component(g["distance"])
        self.path_length = _component(g["path_length"])
        self.attenuation = g["attenuation"]
        self.phase = g["phase"]
        self.surface = _component(g["surface"])
        self.bottom = _component(g["bottom"])
        self.caustic = _component(g["caustic"])
        self.upper = _component(g["upper"])
        self.lower = _component(g["lower"])
        self.on_edge = _component(g["on_edge"].astype(bool))
        self.arrays = g


def front_buffers(n, nf, fields=None):
    """Host arrays for the named fields of one queue slot (all of them when
    fields is None) and the C view over them; fields not named stay NULL in
    the view and are not copied."""
    want = None if fields is None else set(fields)
    arrays = {}
    view = _abi.FrontView()
    for name in _abi.FRONT_FIELDS:
        if want is None or name in want:
            arrays[name] = np.zeros(n * nf if name in ("attenuation", "phase") else n)
            setattr(view, name, arrays[name].ctypes.data_as(C.POINTER(C.c_double)))
    for name in _abi.FRONT_INT_FIELDS:
        if want is None or name in want:
            arrays[name] = np.zeros(n, dtype=np.int32)
            setattr(view, name, arrays[name].ctypes.data_as(C.POINTER(C.c_int32)))
    if want is None or "on_edge" in want:
        arrays["on_edge"] = np.zeros(n, dtype=np.uint8)
        view.on_edge = arrays["on_edge"].ctypes.data_as(C.POINTER(C.c_uint8))
    return arrays, view


def default_thresholds():
    t = _abi.Thresholds()
    t.intensity = 300.0
    t.eigenverb = 1e-30
    t.max_bottom = t.max_surface = t.max_caustic = t.max_upper = t.max_lower = 999
    return t


class queue_base:
    """Descriptor building, listener replay and thresholds shared by the
    device-backed wave_queue and any other backend with the same records."""

    CLASSIC_RAY = _abi.SPREADING_CLASSIC_RAY
    HYBRID_GAUSSIAN = _abi.SPREADING_HYBRID_GAUSSIAN

    def __init__(self, ocean, freq, pos, de, az, time_step, target_pos=None,
                 spreading=_abi.SPREADING_HYBRID_GAUSSIAN):
        for s in (freq, de, az):
            if not isinstance(s, seq_vector):
                raise TypeError("freq, de and az must be seq_vector instances")
        self._ocean = ocean
        self._freq = freq
        self._de = de
        self._az = az
        self._time_step = float(time_step)
        self._targets = target_pos
        self._run_id = 0
        self._thr = default_thresholds()
        self._ray_listeners = []
        self._verb_listeners = []
        self._refl_listeners = []
        self._env, keep = ocean.to_abi()
        self._keep = keep
        fd = _abi.FanDesc()
        fd.freq = freq.to_abi()
        fd.de = de.to_abi()
        fd.az = az.to_abi()
        fd.src_rho, fd.src_theta, fd.src_phi = pos.rho, pos.theta, pos.phi
        fd.time_step = self._time_step
        fd.earth_radius = earth_radius()
        if target_pos is not None:
            self._t_rho = np.ascontiguousarray(target_pos.rho, dtype=np.float64)
            self._t_theta = np.ascontiguousarray(target_pos.theta, dtype=np.float64)
            self._t_phi = np.ascontiguousarray(target_pos.phi, dtype=np.float64)
            fd.num_t1, fd.num_t2 = self._t_rho.shape
            dp = C.POINTER(C.c_double)
            fd.target_rho = self._t_rho.ctypes.data_as(dp)
            fd.target_theta = self._t_theta.ctypes.data_as(dp)
            fd.target_phi = self._t_phi.ctypes.data_as(dp)
        fd.spreading = int(spreading)
        fd.thresholds = self._thr
        self._fan_desc = fd

    # -- accessors (waveq3d/wave_queue.h:136-241)
    def frequencies(self):
        return self._freq

    def source_de(self, i):
        return self._de(i)

    def source_az(self, i):
        return self._az(i)

    def time_step(self):
        return self._time_step

    def targets(self):
        return self._targets

    def num_de(self):
        return self._de.size()

    def num_az(self):
        return self._az.size()

    def runID(self, run_id=None):
        if run_id is not None:
            self._run_id = int(run_id)
        return self._run_id

    # -- thresholds (waveq3d/wave_thresholds.h:54-192)
    def intensity_threshold(self, value=None):
        if value is None:
            return -self._thr.intensity
        self._thr.intensity = abs(float(value))
        self._push_thresholds()

    def eigenverb_threshold(self, value=None):
        if value is None:
            return 10.0 * np.log10(self._thr.eigenverb)
        self._thr.eigenverb = 10.0 ** (-0.1 * abs(float(value)))
        self._push_thresholds()

    def _set_max(self, name, value):
        if value is None:
            return getattr(self._thr, name)
        setattr(self._thr, name, int(value))
        self._push_thresholds()

    def max_bottom(self, value=None):
        return self._set_max("max_bottom", value)

    def max_surface(self, value=None):
        return self._set_max("max_surface", value)

    def max_caustic(self, value=None):
        return self._set_max("max_caustic", value)

    def max_upper(self, value=None):
        return self._set_max("max_upper", value)

    def max_lower(self, value=None):
        return self._set_max("max_lower", value)

    # -- listeners (eigenray_notifier.h, eigenverb_notifier.h, reflection_notifier.h)
    def add_eigenray_listener(self, listener):
        if listener not in self._ray_listeners:
            self._ray_listeners.append(listener)

    def remove_eigenray_listener(self, listener):
        self._ray_listeners.remove(listener)

    def add_eigenverb_listener(self, listener):
        if listener not in self._verb_listeners:
            self._verb_listeners.append(listener)
        self._push_listeners()

    def remove_eigenverb_listener(self, listener):
        self._verb_listeners.remove(listener)
        self._push_listeners()

    def has_eigenverb_listeners(self):
        return bool(self._verb_listeners)

    def add_reflection_listener(self, listener):
        if listener not in self._refl_listeners:
            self._refl_listeners.append(listener)
        self._push_listeners()

    def remove_reflection_listener(self, listener):
        self._refl_listeners.remove(listener)
        self._push_listeners()

    # -- record conversion
    def _eigenray_models(self, recs, inten, phase):
        nf = self._freq.size()
        out = []
        for k, r in enumerate(recs):
            m = eigenray_model()
            m.travel_time = float(r["travel_time"])
            m.frequencies = self._freq
            m.intensity = inten[k * nf:(k + 1) * nf].copy()
            m.phase = phase[k * nf:(k + 1) * nf].copy()
            m.source_de = float(r["source_de"])
            m.source_az = float(r["source_az"])
            m.target_de = float(r["target_de"])
            m.target_az = float(r["target_az"])
            for name in ("surface", "bottom", "caustic", "upper", "lower"):
                setattr(m, name, int(r[name]))
            out.append((int(r["t1"]), int(r["t2"]), m))
        return out

    def _eigenverb_models(self, recs, power):
        nf = self._freq.size()
        out = []
        for k, r in enumerate(recs):
            v = eigenverb_model()
            v.travel_time = float(r["travel_time"])
            v.frequencies = self._freq
            v.power = power[k * nf:(k + 1) * nf].copy()
            v.length = float(r["length"])
            v.width = float(r["width"])
            v.position = wposition1.from_rtp(r["pos_rho"], r["pos_theta"], r["pos_phi"])
            v.direction = float(r["direction"])
            v.grazing = float(r["grazing"])
            v.sound_speed = float(r["sound_speed"])
            v.de_index = int(r["de"])
            v.az_index = int(r["az"])
            v.source_de = float(r["source_de"])
            v.source_az = float(r["source_az"])
            for name in ("surface", "bottom", "caustic", "upper", "lower"):
                setattr(v, name, int(r[name]))
            out.append((int(r["type"]), v))
        return out

    def _deliver(self, step_time, rays, verbs, refls):
        """Replay one step's products in the reference's order: reflection and
        eigenverb callbacks happen inside detect_reflections (ordered by de, az,
        seq), eigenrays inside detect_eigenrays (t1, t2, de, az), then
        check_eigenrays (wave_queue.cc:155-192)."""
        events = []
        for r in refls:
            events.append(((int(r["de"]), int(r["az"]), int(r["seq"])), "refl", r))
        for typ, v in verbs:
            events.append(((v.de_index, v.az_index, v._seq), "verb", (typ, v)))
        events.sort(key=lambda e: e[0])
        for _, kind, item in events:
            if kind == "refl":
                pos = wposition1.from_rtp(item["pos_rho"], item["pos_theta"], item["pos_phi"])
                nd = wposition1.from_rtp(item["ndir_rho"], item["ndir_theta"], item["ndir_phi"])
                for lst in self._refl_listeners:
                    lst.reflect(float(item["time"]), int(item["de"]), int(item["az"]),
                                float(item["dt"]), float(item["grazing"]),
                                float(item["speed"]), pos, nd, int(item["type"]))
            else:
                typ, v = item
                for lst in self._verb_listeners:
                    lst.add_eigenverb(v, typ)
        for t1, t2, m in rays:
            for lst in self._ray_listeners:
                lst.add_eigenray(t1, t2, m, self._run_id)
        for lst in self._ray_listeners:
            if hasattr(lst, "check_eigenrays"):
                lst.check_eigenrays(step_time, self._run_id)

    def _push_thresholds(self):
        pass

    def _push_listeners(self):
        pass

    # a backend whose step() is synchronous completes the steps in step_end
    def step_begin(self, n=1):
        pass

    def step_end(self, n=1):
        self.step(n)

    # -- wavefront log (waveq3d/wave_queue_netcdf.cc:14-108, CDL in
    # wave_queue.h:760-814).  A backend provides time() and
    # snapshot_arrays(slot, fields).
    _LOG_FIELDS = ("pos_rho", "pos_theta", "pos_phi") + tuple(_abi.FRONT_INT_FIELDS) + ("on_edge",)
    _LOG_COUNTS = ("surface", "bottom", "caustic", "upper", "lower")

    def init_netcdf(self, filename, long_name=None):
        """Open the wavefront log (wave_queue_netcdf.cc:14-77): dimensions
        frequencies, source_de, source_az and the unlimited travel_time;
        the fan's coordinates; per record the time and the curr() positions
        as latitude / longitude / altitude, the five path counters (short)
        and on_edge (byte), indexed (travel_time, source_de, source_az).
        The file is netCDF classic (64-bit offset)."""
        from . import netcdf3 as nc
        self.close_netcdf()
        nf, nde, naz = self._freq.size(), self.num_de(), self.num_az()
        w = nc.Writer()
        if long_name is not None:
            w.atts["long_name"] = long_name
        w.atts["Conventions"] = "COARDS"
        fdim = w.add_dim("frequencies", nf)
        ddim = w.add_dim("source_de", nde)
        adim = w.add_dim("source_az", naz)
        tdim = w.add_dim("travel_time", nc.UNLIMITED)
        tda = [tdim, ddim, adim]
        w.add_var("frequencies", nc.NC_DOUBLE, [fdim], [self._freq(k) for k in range(nf)],
                  {"units": "hertz"})
        w.add_var("source_de", nc.NC_DOUBLE, [ddim], [self._de(k) for k in range(nde)],
                  {"units": "degrees", "positive": "up"})
        w.add_var("source_az", nc.NC_DOUBLE, [adim], [self._az(k) for k in range(naz)],
                  {"units": "degrees_true", "positive": "clockwise"})
        w.add_var("travel_time", nc.NC_DOUBLE, [tdim], atts={"units": "seconds"})
        w.add_var("latitude", nc.NC_DOUBLE, tda, atts={"units": "degrees_north"})
        w.add_var("longitude", nc.NC_DOUBLE, tda, atts={"units": "degrees_east"})
        w.add_var("altitude", nc.NC_DOUBLE, tda, atts={"units": "meters", "positive": "up"})
        for name in self._LOG_COUNTS:
            w.add_var(name, nc.NC_SHORT, tda, atts={"units": "count"})
        w.add_var("on_edge", nc.NC_BYTE, tda, atts={"units": "bool"})
        w.create(filename)
        self._nc_log = w

    def save_netcdf(self):
        """Append the current wavefront as one record (wave_queue_netcdf.cc:
        82-101): wposition's latitude = 90 - theta (180/pi), longitude =
        phi (180/pi), altitude = rho - earth radius (types/wposition.h)."""
        w = getattr(self, "_nc_log", None)
        if w is None:
            raise RuntimeError("save_netcdf() before init_netcdf()")
        a = self.snapshot_arrays(_abi.CURR, self._LOG_FIELDS)
        rec = {"travel_time": self.time(),
               "latitude": 90.0 - a["pos_theta"] * (180.0 / np.pi),
               "longitude": a["pos_phi"] * (180.0 / np.pi),
               "altitude": a["pos_rho"] - earth_radius(),
               "on_edge": a["on_edge"]}
        for name in self._LOG_COUNTS:
            rec[name] = a[name]
        w.put_record(rec)

    def close_netcdf(self):
        """wave_queue_netcdf.cc:106-108."""
        w = getattr(self, "_nc_log", None)
        if w is not None:
            w.close()
            self._nc_log = None


def _load_lib():
    from .. import _lib
    return _lib.load()


class wave_queue(queue_base):
    """Device-backed wave_queue; every step runs on the HIP library."""

    def __init__(self, ocean, freq, pos, de, az, time_step, target_pos=None,
                 spreading=_abi.SPREADING_HYBRID_GAUSSIAN, env_handle=None, lib=None):
        super().__init__(ocean, freq, pos, de, az, time_step, target_pos, spreading)
        self._lib = lib if lib is not None else _load_lib()
        lib = self._lib
        self._own_env = env_handle is None
        self._envh = C.c_void_p()
        self._fan = C.c_void_p()
        if env_handle is None:
            lib.check(lib.wq3d_env_create(C.byref(self._env), C.byref(self._envh)))
        else:
            self._envh = env_handle
        lib.check(lib.wq3d_fan_create(self._envh, C.byref(self._fan_desc), C.byref(self._fan)))
        self._time = 0.0
        self._steps = 0
        r, t, p = C.c_double(), C.c_double(), C.c_double()
        lib.check(lib.wq3d_source(self._fan, C.byref(r), C.byref(t), C.byref(p)))
        self._source = wposition1.from_rtp(r.value, t.value, p.value)

    def close(self):
        self.close_netcdf()
        if getattr(self, "_fan", None) and self._fan.value:
            self._lib.wq3d_fan_destroy(self._fan)
            self._fan = C.c_void_p()
        if getattr(self, "_own_env", False) and self._envh.value:
            self._lib.wq3d_env_destroy(self._envh)
            self._envh = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def source_pos(self):
        return self._source

    def time(self):
        return self._time

    def _push_thresholds(self):
        self._lib.check(self._lib.wq3d_set_thresholds(self._fan, C.byref(self._thr)))

    def _push_listeners(self):
        self._lib.check(self._lib.wq3d_set_listeners(
            self._fan, 1 if self._verb_listeners else 0, 1 if self._refl_listeners else 0))

    def step(self, n=1):
        """One wave_queue::step() per call (n > 1 batches steps on the device
        and replays listeners per step afterwards)."""
        self.step_begin(n)
        self.step_end(n)

    def step_begin(self, n=1):
        """Queue n steps on the fan's stream and return at once; step_end(n)
        completes them.  Fans on different queues overlap on the device."""
        self._lib.check(self._lib.wq3d_step(self._fan, int(n)))

    def step_end(self, n=1):
        """Wait for the n steps queued by step_begin(n), advance time() and
        replay their listener calls step by step."""
        if not (self._ray_listeners or self._verb_listeners or self._refl_listeners):
            for _ in range(int(n)):
                self._time += self._time_step
            self._steps += int(n)
            return
        rays, inten, phase = self.drain_eigenrays()
        verbs, power = self.drain_eigenverbs()
        refls = self.drain_reflections()
        for _ in range(int(n)):
            self._time += self._time_step
            self._steps += 1
            s = self._steps
            rs = rays[rays["step"] == s] if rays.size else rays
            idx = np.nonzero(rays["step"] == s)[0] if rays.size else []
            nf = self._freq.size()
            ri = np.concatenate([inten[k * nf:(k + 1) * nf] for k in idx]) if len(idx) else inten[:0]
            rp = np.concatenate([phase[k * nf:(k + 1) * nf] for k in idx]) if len(idx) else phase[:0]
            # reflections and eigenverbs are keyed by the step count before the
            # rotation (they come from detect_reflections), eigenrays after it
            vidx = np.nonzero(verbs["step"] == s - 1)[0] if verbs.size else []
            vs = verbs[vidx] if len(vidx) else verbs[:0]
            vp = np.concatenate([power[k * nf:(k + 1) * nf] for k in vidx]) if len(vidx) else power[:0]
            vm = self._eigenverb_models(vs, vp)
            for (typ, v), rec in zip(vm, vs):
                v._seq = int(rec["seq"])
            fs = refls[refls["step"] == s - 1] if refls.size else refls
            self._deliver(self._time, self._eigenray_models(rs, ri, rp), vm, fs)

    # -- differentiation (no counterpart in the reference; include/wq3d.h)
    def adjoint_checkpoint(self):
        """Keep a device copy of the whole fan state for the backward pass:
        once before the first step, then every K steps of the run."""
        self._lib.check(self._lib.wq3d_adjoint_checkpoint(self._fan))

    def adjoint_gradient(self, recs, intensity, grad_time, grad_intensity, num_nodes):
        """d loss / d node speed of the sound speed grid, for eigenray records
        of this fan (as drain_eigenrays returns them) and the loss gradients
        with respect to their travel times [n] and intensities [n, nf] (dB)."""
        nf = self._freq.size()
        n = len(recs)
        recs = np.ascontiguousarray(recs)
        inten = np.ascontiguousarray(intensity, dtype=np.float64).reshape(n * nf)
        gt = np.ascontiguousarray(grad_time, dtype=np.float64).reshape(n)
        gi = np.ascontiguousarray(grad_intensity, dtype=np.float64).reshape(n * nf)
        out = np.zeros(int(num_nodes))
        dp = C.POINTER(C.c_double)
        self._lib.check(self._lib.wq3d_adjoint_gradient(
            self._fan, recs.ctypes.data_as(C.POINTER(_abi.EigenrayRec)), inten.ctypes.data_as(dp),
            gt.ctypes.data_as(dp), gi.ctypes.data_as(dp), n, out.ctypes.data_as(dp), int(num_nodes)))
        return out

    def adjoint_release(self):
        self._lib.check(self._lib.wq3d_adjoint_release(self._fan))

    # -- raw record access (numpy structured arrays), used by tests and bench
    def drain_eigenrays(self):
        lib = self._lib
        n = C.c_int64()
        lib.check(lib.wq3d_drain_eigenrays(self._fan, None, None, None, 0, C.byref(n)))
        nf = self._freq.size()
        recs = np.zeros(n.value, dtype=_abi.record_dtype(_abi.EigenrayRec))
        inten = np.zeros(n.value * nf)
        phase = np.zeros(n.value * nf)
        lib.check(lib.wq3d_drain_eigenrays(
            self._fan, recs.ctypes.data_as(C.POINTER(_abi.EigenrayRec)),
            inten.ctypes.data_as(C.POINTER(C.c_double)),
            phase.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
        return recs[:n.value], inten[:n.value * nf], phase[:n.value * nf]

    def drain_eigenverbs(self):
        lib = self._lib
        n = C.c_int64()
        lib.check(lib.wq3d_drain_eigenverbs(self._fan, None, None, 0, C.byref(n)))
        nf = self._freq.size()
        recs = np.zeros(n.value, dtype=_abi.record_dtype(_abi.EigenverbRec))
        power = np.zeros(n.value * nf)
        lib.check(lib.wq3d_drain_eigenverbs(
            self._fan, recs.ctypes.data_as(C.POINTER(_abi.EigenverbRec)),
            power.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
        return recs[:n.value], power[:n.value * nf]

    def drain_reflections(self):
        lib = self._lib
        n = C.c_int64()
        lib.check(lib.wq3d_drain_reflections(self._fan, None, 0, C.byref(n)))
        recs = np.zeros(n.value, dtype=_abi.record_dtype(_abi.ReflectionRec))
        lib.check(lib.wq3d_drain_reflections(
            self._fan, recs.ctypes.data_as(C.POINTER(_abi.ReflectionRec)), n.value, C.byref(n)))
        return recs[:n.value]

    def snapshot_arrays(self, slot, fields=None):
        arrays, view = front_buffers(self.num_de() * self.num_az(), self._freq.size(), fields)
        self._lib.check(self._lib.wq3d_snapshot(self._fan, slot, C.byref(view)))
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

    def sync(self):
        self._lib.check(self._lib.wq3d_sync(self._fan))


class fan_batch:
    """Many device wave_queues stepped together through wq3d_batch_*: one
    fused launch per kernel of a step for every queue, 64-step loops replayed
    as HIP graphs.  step(n) is n calls of step() on every queue, with each
    queue's listener calls replayed in its own order afterwards, as
    wave_queue.step(n) replays them.  The queues must be fresh (or at the
    same step) and are stepped only through the batch while it is open."""

    def __init__(self, queues, lib=None):
        self._queues = list(queues)
        if not self._queues:
            raise ValueError("a batch needs at least one wave_queue")
        self._lib = lib if lib is not None else self._queues[0]._lib
        fans = (C.c_void_p * len(self._queues))(*[q._fan.value for q in self._queues])
        self._h = C.c_void_p()
        self._lib.check(self._lib.wq3d_batch_create(fans, len(self._queues), C.byref(self._h)))

    def step(self, n=1):
        self._lib.check(self._lib.wq3d_batch_step(self._h, int(n)))
        for q in self._queues:
            q.step_end(int(n))

    def shape(self):
        s = _abi.BatchShape()
        self._lib.check(self._lib.wq3d_batch_get_shape(self._h, C.byref(s)))
        return {"ray_threads": s.ray_threads, "build_blocks": s.build_blocks, "use_graph": bool(s.use_graph)}

    def set_shape(self, ray_threads=256, build_blocks=1024, use_graph=True):
        s = _abi.BatchShape(ray_threads, build_blocks, 1 if use_graph else 0, 0)
        self._lib.check(self._lib.wq3d_batch_set_shape(self._h, C.byref(s)))

    def tune(self, trial=8):
        """Run the autotuner over `trial` real steps per candidate shape (the
        queues advance by those steps, their listener calls replayed);
        returns the chosen shape and its time per step in ms."""
        ms = C.c_double()
        taken = C.c_int64()
        self._lib.check(self._lib.wq3d_batch_tune(self._h, int(trial), C.byref(ms), C.byref(taken)))
        for q in self._queues:
            q.step_end(taken.value)
        return self.shape(), ms.value

    def close(self):
        if self._h.value:
            self._lib.wq3d_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
