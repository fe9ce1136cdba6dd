"""usml::eigenrays::eigenray_collection over libwq3d.so.

Mirrors the reference class (eigenrays/eigenray_collection.h:34-262): an
eigenray listener that files each eigenray under its target (add_eigenray,
eigenray_collection.cc:90-101) and sums them per target (sum_eigenrays,
:105-191), coherently by default.  The sum runs on the GPU through
wq3d_sum_eigenrays; there is no host fallback.
"""
import ctypes as C

import numpy as np

from .. import _abi
from .waveq3d import eigenray_model

_REC = _abi.record_dtype(_abi.EigenrayRec)
_TOT = _abi.record_dtype(_abi.EigenrayTotal)


def pack_eigenrays(entries, num_freq):
    """(t1, t2, eigenray_model) entries, in arrival order, as the C ABI's
    record array plus intensity and phase arrays (num_freq per ray)."""
    n = len(entries)
    recs = np.zeros(n, dtype=_REC)
    inten = np.zeros(n * num_freq)
    phase = np.zeros(n * num_freq)
    for k, (t1, t2, ray) in enumerate(entries):
        r = recs[k]
        r["t1"], r["t2"] = t1, t2
        for name in ("travel_time", "source_de", "source_az", "target_de", "target_az",
                     "surface", "bottom", "caustic", "upper", "lower"):
            r[name] = getattr(ray, name)
        inten[k * num_freq:(k + 1) * num_freq] = ray.intensity
        phase[k * num_freq:(k + 1) * num_freq] = ray.phase
    return recs, inten, phase


def sum_eigenray_arrays(fn, recs, inten, phase, freq, num_t1, num_t2, coherent):
    """Call a wq3d_sum_eigenrays-shaped function `fn` on packed arrays;
    returns (totals, intensity, phase) arrays of num_t1 * num_t2 targets."""
    nf = len(freq)
    recs = np.ascontiguousarray(recs, dtype=_REC)
    inten = np.ascontiguousarray(inten, dtype=np.float64)
    phase = np.ascontiguousarray(phase, dtype=np.float64)
    freq = np.ascontiguousarray(freq, dtype=np.float64)
    nt = num_t1 * num_t2
    tot = np.zeros(max(1, nt), dtype=_TOT)
    ti = np.zeros(max(1, nt * nf))
    tp = np.zeros(max(1, nt * nf))
    dp = C.POINTER(C.c_double)
    status = fn(recs.ctypes.data_as(C.c_void_p) if len(recs) else None,
                inten.ctypes.data_as(dp), phase.ctypes.data_as(dp), len(recs),
                freq.ctypes.data_as(dp), nf, num_t1, num_t2, 1 if coherent else 0,
                tot.ctypes.data_as(C.POINTER(_abi.EigenrayTotal)), ti.ctypes.data_as(dp),
                tp.ctypes.data_as(dp))
    return status, tot[:nt], ti[:nt * nf], tp[:nt * nf]


class eigenray_collection:
    """eigenrays/eigenray_collection.h:34-262."""

    def __init__(self, frequencies, source_pos, target_pos, sourceID=0, targetIDs=None,
                 coherent=True, lib=None):
        from .. import _lib
        self._lib = lib if lib is not None else _lib.load()
        self._frequencies = frequencies
        self._source_pos = source_pos
        self._target_pos = target_pos
        self._sourceID = sourceID
        n1, n2 = target_pos.size1(), target_pos.size2()
        self._targetIDs = (np.zeros((n1, n2), dtype=np.uint64) if targetIDs is None
                           else np.asarray(targetIDs, dtype=np.uint64))
        self._coherent = bool(coherent)
        self._entries = []  # (t1, t2, eigenray_model) in arrival order
        self._lists = [[[] for _ in range(n2)] for _ in range(n1)]
        self._initial = np.zeros((n1, n2))
        self._total = [[self._empty_total() for _ in range(n2)] for _ in range(n1)]

    def _empty_total(self):
        m = eigenray_model()
        m.frequencies = self._frequencies
        m.intensity = np.zeros(self._frequencies.size())
        m.phase = np.zeros(self._frequencies.size())
        for name in ("travel_time", "source_de", "source_az", "target_de", "target_az"):
            setattr(m, name, 0.0)
        for name in ("surface", "bottom", "caustic", "upper", "lower"):
            setattr(m, name, 0)
        return m

    # eigenray_listener interface
    def add_eigenray(self, t1, t2, ray, runID=0):
        """eigenray_collection.cc:90-101."""
        self._entries.append((t1, t2, ray))
        self._lists[t1][t2].append(ray)
        old = self._initial[t1, t2]
        if old <= 0.0 or old > ray.travel_time:
            self._initial[t1, t2] = ray.travel_time

    def check_eigenrays(self, time, runID=0):
        pass

    # accessors
    def size1(self):
        return self._target_pos.size1()

    def size2(self):
        return self._target_pos.size2()

    def frequencies(self):
        return self._frequencies

    def source_pos(self):
        return self._source_pos

    def target_pos(self):
        return self._target_pos

    def sourceID(self):
        return self._sourceID

    def targetID(self, t1, t2):
        return int(self._targetIDs[t1, t2])

    def coherent(self):
        return self._coherent

    def num_eigenrays(self):
        return len(self._entries)

    def eigenrays(self, t1=0, t2=0):
        return list(self._lists[t1][t2])

    def initial_time(self, t1=0, t2=0):
        return float(self._initial[t1, t2])

    def total(self, t1=0, t2=0):
        return self._total[t1][t2]

    def write_netcdf(self, filename, long_name=None):
        """eigenray_collection.cc:196-356: the same dimensions, variables,
        attributes and record layout, in a netCDF classic (CDF-2) file.
        For each target the summed record comes first, then its eigenrays
        in arrival order; proploss_index / eigenray_index / eigenray_num
        locate them.  The reference writes sourceID and targetID as uint64,
        which the classic format lacks: here they are doubles (exact below
        2^53)."""
        from . import netcdf3 as nc
        n1, n2 = self.size1(), self.size2()
        nf = self._frequencies.size()
        w = nc.Writer()
        if long_name is not None:
            w.atts["long_name"] = long_name
        w.atts["Conventions"] = "COARDS"
        rows, cols = w.add_dim("rows", n1), w.add_dim("cols", n2)
        ray_dim = w.add_dim("eigenrays", self.num_eigenrays() + n1 * n2)
        freq_dim = w.add_dim("frequencies", nf)
        records = []  # eigenray_model per record, in file order
        prop_index = np.zeros((n1, n2), dtype=np.int32)
        ray_index = np.zeros((n1, n2), dtype=np.int32)
        ray_num = np.zeros((n1, n2), dtype=np.int32)
        for t1 in range(n1):
            for t2 in range(n2):
                prop_index[t1, t2] = len(records)
                ray_index[t1, t2] = len(records) + 1
                ray_num[t1, t2] = len(self._lists[t1][t2])
                records.append(self._total[t1][t2])
                records.extend(self._lists[t1][t2])
        tp = self._target_pos
        w.add_var("sourceID", nc.NC_DOUBLE, [], [float(self._sourceID)])
        w.add_var("source_latitude", nc.NC_DOUBLE, [], [self._source_pos.latitude()],
                  {"units": "degrees_north"})
        w.add_var("source_longitude", nc.NC_DOUBLE, [], [self._source_pos.longitude()],
                  {"units": "degrees_east"})
        w.add_var("source_altitude", nc.NC_DOUBLE, [], [self._source_pos.altitude()],
                  {"units": "meters", "positive": "up"})
        w.add_var("targetID", nc.NC_DOUBLE, [rows, cols], self._targetIDs.astype(np.float64))
        grid = [[(tp.latitude(t1, t2), tp.longitude(t1, t2), tp.altitude(t1, t2))
                 for t2 in range(n2)] for t1 in range(n1)]
        grid = np.array(grid, dtype=np.float64).reshape(n1, n2, 3)
        w.add_var("latitude", nc.NC_DOUBLE, [rows, cols], grid[:, :, 0], {"units": "degrees_north"})
        w.add_var("longitude", nc.NC_DOUBLE, [rows, cols], grid[:, :, 1], {"units": "degrees_east"})
        w.add_var("altitude", nc.NC_DOUBLE, [rows, cols], grid[:, :, 2],
                  {"units": "meters", "positive": "up"})
        w.add_var("initial_time", nc.NC_DOUBLE, [rows, cols], self._initial, {"units": "seconds"})
        w.add_var("frequencies", nc.NC_DOUBLE, [freq_dim],
                  [self._frequencies(k) for k in range(nf)], {"units": "hertz"})
        w.add_var("proploss_index", nc.NC_INT, [rows, cols], prop_index, {"units": "count"})
        w.add_var("eigenray_index", nc.NC_INT, [rows, cols], ray_index, {"units": "count"})
        w.add_var("eigenray_num", nc.NC_INT, [rows, cols], ray_num, {"units": "count"})
        w.add_var("intensity", nc.NC_DOUBLE, [ray_dim, freq_dim],
                  [np.asarray(r.intensity, dtype=np.float64) for r in records] or np.zeros((0, nf)),
                  {"units": "dB"})
        w.add_var("phase", nc.NC_DOUBLE, [ray_dim, freq_dim],
                  [np.asarray(r.phase, dtype=np.float64) for r in records] or np.zeros((0, nf)),
                  {"units": "radians"})
        w.add_var("travel_time", nc.NC_DOUBLE, [ray_dim], [r.travel_time for r in records],
                  {"units": "seconds"})
        angle_atts = {"source_de": {"units": "degrees", "positive": "up"},
                      "source_az": {"units": "degrees_true", "positive": "clockwise"},
                      "target_de": {"units": "degrees", "positive": "up"},
                      "target_az": {"units": "degrees_true", "positive": "clockwise"}}
        for name, atts in angle_atts.items():
            w.add_var(name, nc.NC_DOUBLE, [ray_dim], [getattr(r, name) for r in records], atts)
        for name in ("surface", "bottom", "caustic", "upper", "lower"):
            w.add_var(name, nc.NC_SHORT, [ray_dim], [getattr(r, name) for r in records],
                      {"units": "count"})
        w.write(filename)

    def sum_eigenrays(self):
        """eigenray_collection.cc:105-191, on the GPU."""
        nf = self._frequencies.size()
        freq = np.array([self._frequencies(k) for k in range(nf)])
        recs, inten, phase = pack_eigenrays(self._entries, nf)
        n1, n2 = self.size1(), self.size2()
        status, tot, ti, tp = sum_eigenray_arrays(self._lib.wq3d_sum_eigenrays, recs, inten, phase,
                                                  freq, n1, n2, self._coherent)
        self._lib.check(status)
        for t1 in range(n1):
            for t2 in range(n2):
                k = t1 * n2 + t2
                m = self._empty_total()
                rec = tot[k]
                for name in ("travel_time", "source_de", "source_az", "target_de", "target_az"):
                    setattr(m, name, float(rec[name]))
                for name in ("surface", "bottom", "caustic", "upper", "lower"):
                    setattr(m, name, int(rec[name]))
                m.intensity = ti[k * nf:(k + 1) * nf].copy()
                m.phase = tp[k * nf:(k + 1) * nf].copy()
                self._total[t1][t2] = m
