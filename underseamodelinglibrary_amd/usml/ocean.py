"""usml::ocean models supported on the wave_queue path, lowered to wq3d_env_desc.

Each class keeps the reference constructor's argument meaning and defaults:
  profile_linear      ocean/profile_linear.h:53-92, profile_linear.cc:14-40
  profile_munk        ocean/profile_munk.h:87-94, profile_munk.cc:19-32
  profile_grid        ocean/profile_grid.h:68-130 (1-D or 3-D gen_grid)
  data_grid           types/gen_grid.h (axes, data, interp type, edge limit)
  data_grid_mackenzie ocean/data_grid_mackenzie.h:74-110
  attenuation_thorp / attenuation_constant   ocean/attenuation_*.cc
  boundary_flat       ocean/boundary_flat.cc:19-33 (height baked at construction)
  boundary_grid       ocean/boundary_grid.h:88-196
  reflect_loss_constant / reflect_loss_rayleigh   ocean/reflect_loss_*.{h,cc}
  volume_flat         ocean/volume_flat.h
  ocean_model         ocean/ocean_model.h:62-70
Models outside this set raise NotImplementedError, matching the library's
WQ3D_ERR_UNSUPPORTED (there is no host fallback).
"""
import ctypes as C
import math

import numpy as np

from .. import _abi
from .types import earth_radius, seq_vector

NEAREST, LINEAR, PCHIP = _abi.INTERP_NEAREST, _abi.INTERP_LINEAR, _abi.INTERP_PCHIP


class attenuation_thorp:
    kind = _abi.ATTEN_THORP
    coeff = 0.0


class attenuation_constant:
    kind = _abi.ATTEN_CONSTANT

    def __init__(self, coefficient):
        self.coeff = float(coefficient)


class data_grid:
    """N-D gridded data (N <= 3), row-major with axis 0 slowest."""

    def __init__(self, axes, data, interp=None, edge_limit=None):
        self.axes = list(axes)
        if not all(isinstance(a, seq_vector) for a in self.axes):
            raise TypeError("axes must be seq_vector instances")
        shape = tuple(a.size() for a in self.axes)
        self.data = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(shape))
        n = len(self.axes)
        self.interp = list(interp) if interp is not None else [LINEAR] * n
        self.edge_limit = list(edge_limit) if edge_limit is not None else [True] * n

    def ndims(self):
        return len(self.axes)

    def interp_type(self, dim, kind=None):
        if kind is not None:
            self.interp[dim] = kind
        return self.interp[dim]

    def to_abi(self):
        g = _abi.Grid()
        g.ndims = self.ndims()
        for d in range(self.ndims()):
            g.interp[d] = int(self.interp[d])
            g.edge_limit[d] = 1 if self.edge_limit[d] else 0
            g.axis[d] = self.axes[d].to_abi()
        g.data = self.data.ctypes.data_as(C.POINTER(C.c_double))
        return g


def data_grid_mackenzie(temperature, salinity):
    """Sound speed grid from T/S grids with Mackenzie's nine-term formula.

    Depth D = earth_radius - rho uses the radius in effect now, as the
    reference's constructor does.  Interp types and edge limits are copied
    from the temperature grid.
    """
    rho_axis = temperature.axes[0].values
    D = earth_radius() - rho_axis
    T = temperature.data
    S = salinity.data
    D = D.reshape((-1,) + (1,) * (T.ndim - 1))
    c = (1448.96 + 4.591 * T - 5.304e-2 * T * T + 2.374e-4 * T * T * T
         + (1.340 - 1.025e-2 * T) * (S - 35.0) + 1.630e-2 * D + 1.675e-7 * D * D
         - 7.139e-13 * T * D * D * D)
    return data_grid(temperature.axes, c, list(temperature.interp),
                     list(temperature.edge_limit))


class _profile:
    def __init__(self, attenuation=None):
        self.attenuation = attenuation if attenuation is not None else attenuation_thorp()
        self._flat_earth = False

    def flat_earth(self, flat):
        self._flat_earth = bool(flat)

    def _base(self):
        d = _abi.ProfileDesc()
        d.flat_earth = 1 if self._flat_earth else 0
        d.atten_kind = self.attenuation.kind
        d.atten_coeff = self.attenuation.coeff
        return d


class profile_linear(_profile):
    """profile_linear(c0=1500, attn) | (c0, g0, attn) | (c0, g0, z1, g1, attn)."""

    def __init__(self, *args, attenuation=None):
        args = list(args)
        if args and not isinstance(args[-1], (int, float)):
            attenuation = args.pop()
        super().__init__(attenuation)
        if len(args) == 0:
            c0, g0, z1, g1 = 1500.0, 0.0, 0.0, 0.0
        elif len(args) == 1:
            c0, g0, z1, g1 = args[0], 0.0, 0.0, 0.0
        elif len(args) == 2:
            c0, g0, z1, g1 = args[0], args[1], 0.0, args[1]
        else:
            c0, g0, z1, g1 = args[:4]
        self.p = [float(c0), float(g0), float(z1), float(g1)]

    def to_abi(self):
        d = self._base()
        d.kind = _abi.PROFILE_LINEAR
        for i in range(4):
            d.p[i] = self.p[i]
        return d


class profile_munk(_profile):
    def __init__(self, axis_depth=1300.0, scale=1300.0, axis_speed=1500.0, epsilon=7.37e-3,
                 attenuation=None):
        super().__init__(attenuation)
        self.p = [float(axis_depth), float(scale), float(axis_speed), float(epsilon)]

    def to_abi(self):
        d = self._base()
        d.kind = _abi.PROFILE_MUNK
        for i in range(4):
            d.p[i] = self.p[i]
        return d


class profile_grid(_profile):
    def __init__(self, grid, attenuation=None):
        super().__init__(attenuation)
        if grid.ndims() not in (1, 3):
            raise NotImplementedError("sound speed grid must be 1-D or 3-D")
        self.grid = grid

    def to_abi(self):
        d = self._base()
        d.kind = _abi.PROFILE_GRID
        d.grid = self.grid.to_abi()
        return d


class reflect_loss_constant:
    kind = _abi.LOSS_CONSTANT

    def __init__(self, amplitude, phase=0.0):
        self.params = [float(amplitude), float(phase), 0.0, 0.0, 0.0]


class reflect_loss_rayleigh:
    """Rayleigh loss from a bottom type name or explicit ratios
    (ocean/reflect_loss_rayleigh.cc:27-78)."""

    kind = _abi.LOSS_RAYLEIGH
    # type: density ratio, speed ratio, att (dB/lambda), shear speed ratio, shear att
    TABLE = {
        "clay": (1.5, 1.00, 0.2, 0.00, 1.0),
        "silt": (1.7, 1.05, 1.0, 0.00, 1.5),
        "sand": (1.9, 1.10, 0.8, 0.00, 2.5),
        "gravel": (2.0, 1.20, 0.6, 0.00, 1.5),
        "moraine": (2.1, 1.30, 0.4, 0.40, 1.0),
        "chalk": (2.2, 1.60, 0.2, 0.67, 0.5),
        "limestone": (2.4, 2.00, 0.1, 1.00, 0.2),
        "basalt": (2.7, 3.50, 0.1, 1.67, 0.2),
    }

    def __init__(self, bottom_type="sand", speed=None, att_bottom=0.0, speed_shear=0.0,
                 att_shear=0.0):
        if speed is None:
            self.params = list(self.TABLE[bottom_type])
        else:
            self.params = [float(bottom_type), float(speed), float(att_bottom),
                           float(speed_shear), float(att_shear)]


class boundary_flat:
    def __init__(self, depth=0.0, reflect_loss=None):
        self.rho = earth_radius() - abs(float(depth))
        if abs(float(depth)) < 1e-6:
            self.normal_rho = -1.0
            if reflect_loss is None:
                reflect_loss = reflect_loss_constant(0.0, math.pi)
        else:
            self.normal_rho = 1.0
        self.loss = reflect_loss if reflect_loss is not None else reflect_loss_constant(0.0, 0.0)

    def to_abi(self):
        b = _abi.BoundaryDesc()
        b.kind = _abi.BOUNDARY_FLAT
        b.rho = self.rho
        b.normal_rho = self.normal_rho
        b.loss_kind = self.loss.kind
        for i in range(5):
            b.loss[i] = self.loss.params[i]
        return b


class boundary_grid:
    """Bathymetry rho = f(theta) or f(theta, phi); axes in radians, data in
    meters from the earth's center (netcdf_bathy.cc:103-173 convention)."""

    def __init__(self, grid, reflect_loss=None):
        if grid.ndims() not in (1, 2):
            raise NotImplementedError("bathymetry must be 1-D or 2-D")
        self.grid = grid
        self.loss = reflect_loss if reflect_loss is not None else reflect_loss_constant(0.0, 0.0)

    def to_abi(self):
        b = _abi.BoundaryDesc()
        b.kind = _abi.BOUNDARY_GRID
        b.normal_rho = 1.0
        b.grid = self.grid.to_abi()
        b.loss_kind = self.loss.kind
        for i in range(5):
            b.loss[i] = self.loss.params[i]
        return b


class volume_flat:
    def __init__(self, depth=0.0, thickness=0.0, amplitude=-300.0):
        self.rho = earth_radius() - abs(float(depth))
        self.thickness = float(thickness)


class ocean_model:
    def __init__(self, surface, bottom, profile, volume=None):
        self.surface = surface
        self.bottom = bottom
        self.profile = profile
        self.volumes = list(volume) if volume else []

    def add_volume(self, layer):
        self.volumes.append(layer)

    def num_volume(self):
        return len(self.volumes)

    def to_abi(self):
        """Descriptor plus the ctypes objects that must outlive its use."""
        for part in (self.surface, self.bottom, self.profile):
            if not hasattr(part, "to_abi"):
                raise NotImplementedError(f"{type(part).__name__} is not supported")
        env = _abi.EnvDesc()
        env.profile = self.profile.to_abi()
        env.surface = self.surface.to_abi()
        env.bottom = self.bottom.to_abi()
        vols = (_abi.VolumeDesc * max(1, len(self.volumes)))()
        for i, v in enumerate(self.volumes):
            vols[i].rho = v.rho
            vols[i].thickness = v.thickness
        env.num_volume = len(self.volumes)
        env.volume = C.cast(vols, C.POINTER(_abi.VolumeDesc))
        return env, [vols]
