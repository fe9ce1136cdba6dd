"""usml::types subset used by the wave_queue path.

Sequences keep the values and increments exactly as the reference accumulates
them, because fan angles and axis spacings enter the arithmetic bit for bit:
  seq_linear  types/seq_linear.h:60-110 (v += increment)
  seq_log     types/seq_log.h:40-100   (v *= increment, increments as differences)
  seq_data    types/seq_data.h:86-120  (increments as forward differences)
  seq_rayfan  types/seq_rayfan.cc:19-47 (tangent spacing on top of seq_data)
Coordinates follow wposition/wposition1 (types/wposition.cc, types/wposition1.cc)
and the conversions of ublas/math_traits.h:156-170.
"""
import ctypes as C
import math

import numpy as np

from .. import _abi


def to_degrees(t):
    return t * (180.0 / math.pi)


def to_radians(t):
    return t * (math.pi / 180.0)


def to_latitude(t):
    return 90.0 - to_degrees(t)


def to_colatitude(t):
    return to_radians(90.0 - t)


class seq_vector:
    """Base: values, increments and the find_index rule of the subclass."""

    search = _abi.SEQ_DATA

    def __init__(self, values, increments):
        self._v = np.ascontiguousarray(values, dtype=np.float64)
        self._inc = np.ascontiguousarray(increments, dtype=np.float64)

    def size(self):
        return int(self._v.size)

    def __len__(self):
        return self.size()

    def _clamp(self, index):
        # the reference's index is a size_t: a negative index wraps to a huge
        # value and clamps to the last element (types/seq_vector.h:98-101)
        index = int(index)
        return self.size() - 1 if index < 0 or index > self.size() - 1 else index

    def __call__(self, index):
        return float(self._v[self._clamp(index)])

    def increment(self, index):
        return float(self._inc[self._clamp(index)])

    @property
    def values(self):
        return self._v

    @property
    def increments(self):
        return self._inc

    def to_abi(self):
        s = _abi.Seq()
        s.size = self.size()
        s.value = self._v.ctypes.data_as(C.POINTER(C.c_double))
        s.increment = self._inc.ctypes.data_as(C.POINTER(C.c_double))
        s.search = self.search
        return s


class seq_linear(seq_vector):
    """seq_linear(first, increment, size:int) or seq_linear(first, increment, last:float)."""

    search = _abi.SEQ_LINEAR

    def __init__(self, first, increment, size_or_last):
        first = float(first)
        increment = float(increment)
        if isinstance(size_or_last, (int, np.integer)):
            size = int(size_or_last)
        else:
            last = float(size_or_last)
            if increment == 0.0:
                size = 1
            else:
                size = max(1, int(math.floor(1.0 + (last + increment / 10 - first) / increment)))
        vals = np.empty(size)
        v = first
        for n in range(size):
            vals[n] = v
            v += increment
        super().__init__(vals, np.full(size, increment))


class seq_log(seq_vector):
    """seq_log(first, increment, size:int) or seq_log(first, increment, last:float)."""

    search = _abi.SEQ_LOG

    def __init__(self, first, increment, size_or_last):
        first = float(first)
        increment = float(increment)
        if isinstance(size_or_last, (int, np.integer)):
            size = int(size_or_last)
        else:
            last = float(size_or_last)
            if increment == 1.0:
                size = 1
            else:
                size = max(1, int(math.floor(
                    1.0 + (math.log((last + increment / 10) / first) / math.log(increment)))))
        vals = np.empty(size)
        inc = np.zeros(size)
        v = first
        for n in range(size):
            vals[n] = v
            v *= increment
            if n > 0:
                inc[n - 1] = vals[n] - vals[n - 1]
                inc[n] = inc[n - 1]
        super().__init__(vals, inc)


class seq_data(seq_vector):
    """Monotonic sequence from explicit values."""

    search = _abi.SEQ_DATA

    def __init__(self, data):
        data = np.asarray(data, dtype=np.float64).ravel()
        size = data.size
        inc = np.zeros(size)
        if size >= 2:
            left = data[1] - data[0]
            inc[0] = left
            for n in range(1, size):
                right = left
                if n < size - 1:
                    right = data[n + 1] - data[n]
                    if left * right <= 0:
                        raise ValueError("series not monotonic")
                left = right
                inc[n] = left
        super().__init__(data.copy(), inc)


class seq_rayfan(seq_data):
    """Tangent-spaced launch angles (types/seq_rayfan.cc:19-47)."""

    def __init__(self, first=-90.0, last=90.0, size=181, center=0.0, spread=6.0):
        if first > last:
            first, last = last, first
        first_ang = math.atan((first - center) / spread)
        last_ang = math.atan((last - center) / spread)
        scale = (last_ang - first_ang) / float(size - 1)
        data = [center + math.tan(first_ang + scale * float(n)) * spread for n in range(size)]
        super().__init__(data)


class _EarthRadius:
    """wposition::earth_radius, a process-wide mutable value (types/wposition.cc:13)."""

    value = 6371000.0


def earth_radius():
    return _EarthRadius.value


def set_earth_radius(r):
    _EarthRadius.value = float(r)


def compute_earth_radius(latitude):
    """WGS-84 radius of curvature (types/wposition.cc:33-42)."""
    a = 6378137.0
    f = 1.0 / 298.257223563
    e2 = f * (2.0 - f)
    sinT = math.sin(to_radians(latitude))
    w = math.sqrt(1.0 - e2 * sinT * sinT)
    rm = a * (1 - e2) / (w * w * w)
    rv = a / w
    _EarthRadius.value = math.sqrt(rm * rv)


class wposition1:
    """One geodetic position stored as (rho, theta, phi)."""

    def __init__(self, latitude=0.0, longitude=0.0, altitude=0.0):
        self.theta = to_colatitude(float(latitude))
        self.phi = to_radians(float(longitude))
        self.rho = float(altitude) + earth_radius()

    @classmethod
    def from_range_bearing(cls, location, range_m, bearing):
        """Point at a great-circle range (m) and bearing (rad) from location
        (types/wposition1.cc:23-37)."""
        lat1 = to_radians(location.latitude())
        lng1 = to_radians(location.longitude())
        R = earth_radius() + location.altitude()
        rng = range_m / R
        lat2 = math.asin(math.sin(lat1) * math.cos(rng)
                         + math.cos(lat1) * math.sin(rng) * math.cos(bearing))
        lng2 = lng1 + math.atan2(math.sin(bearing) * math.sin(rng) * math.cos(lat1),
                                 math.cos(rng) - math.sin(lat1) * math.sin(lat2))
        return cls(to_degrees(lat2), to_degrees(lng2), location.altitude())

    @classmethod
    def from_rtp(cls, rho, theta, phi):
        p = cls.__new__(cls)
        p.rho, p.theta, p.phi = float(rho), float(theta), float(phi)
        return p

    def latitude(self):
        return to_latitude(self.theta)

    def longitude(self):
        return to_degrees(self.phi)

    def altitude(self):
        return self.rho - earth_radius()


class wposition:
    """A rows x cols grid of positions (targets)."""

    def __init__(self, rows=1, cols=1, latitude=0.0, longitude=0.0, altitude=0.0):
        shape = (int(rows), int(cols))
        self.theta = np.empty(shape)
        self.phi = np.empty(shape)
        self.rho = np.empty(shape)
        self.theta[...] = to_colatitude(float(latitude))
        self.phi[...] = to_radians(float(longitude))
        self.rho[...] = float(altitude) + earth_radius()

    @classmethod
    def from_arrays(cls, latitude, longitude, altitude):
        """Element-wise latitude/longitude/altitude arrays of one shape."""
        lat = np.asarray(latitude, dtype=np.float64)
        p = cls(lat.shape[0], lat.shape[1])
        R = earth_radius()
        for idx in np.ndindex(lat.shape):
            p.theta[idx] = to_colatitude(float(lat[idx]))
            p.phi[idx] = to_radians(float(np.asarray(longitude)[idx]))
            p.rho[idx] = float(np.asarray(altitude)[idx]) + R
        return p

    def size1(self):
        return self.rho.shape[0]

    def size2(self):
        return self.rho.shape[1]

    def latitude(self, t1, t2):
        return to_latitude(self.theta[t1, t2])

    def longitude(self, t1, t2):
        return to_degrees(self.phi[t1, t2])

    def altitude(self, t1, t2):
        return self.rho[t1, t2] - earth_radius()
