"""Seeded synthetic environments and fans of the benchmark workloads.

Each builder returns a `scenario` with the wave_queue constructor arguments
and the step count the reference's caller loop `while (time < t_max) step()`
produces with `time += time_step` accumulation (wave_queue.cc:167).

  config1  Munk profile (c0 1500 m/s, axis 1300 m, B 1300 m, eps 0.00737),
           flat 5000 m Rayleigh-sand bottom, flat surface, Thorp; source
           45N 45W 1000 m; DE -60:1:60, AZ 0:10:350, 1000 Hz; dt 0.1 s,
           t < 60 s; 10 x 10 targets within 0.5 deg, 100-2000 m deep.
  config5  range-independent: 64-node piecewise-linear c(z) (1-D
           profile_grid, linear in rho) from seeded perturbed Munk
           profiles, flat 4500 m bottom; DE -45:0.1:45, AZ 0:5:355,
           500-4000 Hz, dt 0.05 s, t < 120 s, 50 x 50 targets.
  config2  range-dependent: 3-D sound speed grid (1/4 deg over a 10 x 10 deg
           box, 40 depths to 5000 m) from Mackenzie on seeded T/S fields;
           1/60 deg bathymetry from 50 seeded Gaussian seamounts/troughs on
           a 4000 m base clipped to 1500-5500 m (bilinear gen_grid);
           DE -80:0.25:80 (641), AZ 0:2:358 (180), 8 log-spaced frequencies
           100-10000 Hz, dt 0.1 s, t < 120 s, 100 x 100 targets.
The T/S construction: T = 25 exp(-z/800) + 2 + dT, S = 35 + dS, where dT and
dS are sums of 64 Gaussian bumps (random centres in the box, widths 0.5-2 deg
horizontally and 200-1500 m vertically) rescaled to RMS 0.5 degC and 0.2 ppt.
"""
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .usml import ocean as oc
from .usml.types import (compute_earth_radius, earth_radius, seq_linear, seq_log, seq_rayfan,
                         set_earth_radius, to_colatitude, to_radians, wposition, wposition1)


@dataclass
class scenario:
    name: str
    ocean: object
    freq: object
    pos: object
    de: object
    az: object
    time_step: float
    targets: Optional[object]
    steps: int
    notes: dict = field(default_factory=dict)
    spreading: int = oc._abi.SPREADING_HYBRID_GAUSSIAN

    def num_rays(self):
        return self.de.size() * self.az.size()

    def num_targets(self):
        return 0 if self.targets is None else self.targets.size1() * self.targets.size2()

    def args(self):
        return (self.ocean, self.freq, self.pos, self.de, self.az, self.time_step, self.targets,
                self.spreading)


def loop_steps(time_step, time_max):
    """Number of step() calls made by `while (time < time_max)`."""
    t, n = 0.0, 0
    while t < time_max:
        t += time_step
        n += 1
    return n


def target_grid(lat0, lng0, half_width_deg, n1, n2, depth_min, depth_max):
    """n1 x n2 targets on a uniform lat/lon grid within +-half_width of
    (lat0, lng0); depth varies linearly with the row index."""
    lats = np.linspace(lat0 - half_width_deg, lat0 + half_width_deg, n2)
    lngs = np.linspace(lng0 - half_width_deg, lng0 + half_width_deg, n1)
    depths = np.linspace(depth_min, depth_max, n1)
    lat = np.empty((n1, n2))
    lng = np.empty((n1, n2))
    alt = np.empty((n1, n2))
    for a in range(n1):
        for b in range(n2):
            lat[a, b] = lats[b]
            lng[a, b] = lngs[a]
            alt[a, b] = -depths[a]
    return wposition.from_arrays(lat, lng, alt)


def config1(de=None, az=None, n_targets=(10, 10), time_max=60.0, targets=True):
    set_earth_radius(6371000.0)
    profile = oc.profile_munk(1300.0, 1300.0, 1500.0, 0.00737, oc.attenuation_thorp())
    bottom = oc.boundary_flat(5000.0, oc.reflect_loss_rayleigh("sand"))
    surface = oc.boundary_flat()
    ocean = oc.ocean_model(surface, bottom, profile)
    freq = seq_linear(1000.0, 0.0, 1)
    pos = wposition1(45.0, -45.0, -1000.0)
    de = de if de is not None else seq_linear(-60.0, 1.0, 60.0)
    az = az if az is not None else seq_linear(0.0, 10.0, 350.0)
    tg = target_grid(45.0, -45.0, 0.5, n_targets[0], n_targets[1], 100.0, 2000.0) if targets else None
    return scenario("config1", ocean, freq, pos, de, az, 0.1, tg, loop_steps(0.1, time_max))


def _bumps(rng, n, shape_fn):
    total = None
    for _ in range(n):
        term = shape_fn(rng)
        total = term if total is None else total + term
    return total


def config2_environment(seed=2024, box_deg=10.0, lat0=45.0, lng0=-45.0, bathy_step_deg=1.0 / 60.0):
    """Seeded gridded ocean of config 2 (see module docstring)."""
    set_earth_radius(6371000.0)
    R = earth_radius()
    rng = np.random.default_rng(seed)
    # sound speed grid: rho (R - depth) descending, colatitude, longitude
    n_z, step = 40, 0.25
    n_h = int(round(box_deg / step)) + 1
    depths = np.array([5000.0 * k / (n_z - 1) for k in range(n_z)])
    rho_axis = seq_linear(R, -5000.0 / (n_z - 1), n_z)
    lat_top = lat0 + box_deg / 2
    th_axis = seq_linear(to_colatitude(lat_top), to_radians(step), n_h)
    ph_axis = seq_linear(to_radians(lng0 - box_deg / 2), to_radians(step), n_h)
    lat = lat_top - step * np.arange(n_h)
    lng = (lng0 - box_deg / 2) + step * np.arange(n_h)
    Z, LA, LO = np.meshgrid(depths, lat, lng, indexing="ij")

    def bump(r):
        la = r.uniform(lat0 - box_deg / 2, lat0 + box_deg / 2)
        lo = r.uniform(lng0 - box_deg / 2, lng0 + box_deg / 2)
        z = r.uniform(0.0, 5000.0)
        sh = r.uniform(0.5, 2.0)
        sz = r.uniform(200.0, 1500.0)
        a = r.normal()
        return a * np.exp(-((LA - la) ** 2 + (LO - lo) ** 2) / (2 * sh * sh)
                          - (Z - z) ** 2 / (2 * sz * sz))

    dT = _bumps(rng, 64, bump)
    dT *= 0.5 / math.sqrt(float(np.mean(dT * dT)))
    dS = _bumps(rng, 64, bump)
    dS *= 0.2 / math.sqrt(float(np.mean(dS * dS)))
    T = 25.0 * np.exp(-Z / 800.0) + 2.0 + dT
    S = 35.0 + dS
    interp = [oc.PCHIP, oc.LINEAR, oc.LINEAR]
    tgrid = oc.data_grid([rho_axis, th_axis, ph_axis], T, interp, [True, True, True])
    sgrid = oc.data_grid([rho_axis, th_axis, ph_axis], S, interp, [True, True, True])
    ssp = oc.data_grid_mackenzie(tgrid, sgrid)
    profile = oc.profile_grid(ssp, oc.attenuation_thorp())
    # bathymetry: 1/60 deg, 50 seamounts / troughs on a 4000 m base
    n_b = int(round(box_deg / bathy_step_deg)) + 1
    blat = lat_top - bathy_step_deg * np.arange(n_b)
    blng = (lng0 - box_deg / 2) + bathy_step_deg * np.arange(n_b)
    BLA, BLO = np.meshgrid(blat, blng, indexing="ij")
    depth = np.full(BLA.shape, 4000.0)
    for _ in range(50):
        la = rng.uniform(lat0 - box_deg / 2, lat0 + box_deg / 2)
        lo = rng.uniform(lng0 - box_deg / 2, lng0 + box_deg / 2)
        w = rng.uniform(0.05, 0.5)
        amp = rng.uniform(500.0, 2500.0) * (1.0 if rng.uniform() < 0.5 else -1.0)
        depth -= amp * np.exp(-((BLA - la) ** 2 + (BLO - lo) ** 2) / (2 * w * w))
    depth = np.clip(depth, 1500.0, 5500.0)
    bth = seq_linear(to_colatitude(lat_top), to_radians(bathy_step_deg), n_b)
    bph = seq_linear(to_radians(lng0 - box_deg / 2), to_radians(bathy_step_deg), n_b)
    bgrid = oc.data_grid([bth, bph], R - depth, [oc.LINEAR, oc.LINEAR], [True, True])
    bottom = oc.boundary_grid(bgrid, oc.reflect_loss_rayleigh("sand"))
    surface = oc.boundary_flat()
    return oc.ocean_model(surface, bottom, profile)


def config2(de=None, az=None, n_targets=(100, 100), time_max=120.0, freq=None, seed=2024,
            ocean=None):
    ocean = ocean if ocean is not None else config2_environment(seed)
    freq = freq if freq is not None else seq_log(100.0, 10.0 ** (2.0 / 7.0), 8)
    pos = wposition1(45.0, -45.0, -1000.0)
    de = de if de is not None else seq_linear(-80.0, 0.25, 80.0)
    az = az if az is not None else seq_linear(0.0, 2.0, 358.0)
    tg = (target_grid(45.0, -45.0, 0.5, n_targets[0], n_targets[1], 100.0, 2000.0)
          if n_targets else None)
    return scenario("config2", ocean, freq, pos, de, az, 0.1, tg, loop_steps(0.1, time_max))


def config5_profile(profile=0, n_nodes=64, bottom_depth=4500.0, sigma=2.0):
    """Sound speed nodes of config 5's perturbed Munk profiles: n_nodes depths
    evenly spaced from the surface to the bottom, the Munk profile (axis
    1300 m, scale 1300 m, 1500 m/s, eps 0.00737) plus Gaussian noise of
    standard deviation `sigma` m/s smoothed over about three nodes and
    rescaled to that RMS; `profile` seeds the noise.  Returns (depths,
    speeds)."""
    depths = np.linspace(0.0, bottom_depth, n_nodes)
    z = 2.0 * (depths - 1300.0) / 1300.0
    munk = 1500.0 * (1.0 + 0.00737 * (z - 1.0 + np.exp(-z)))
    rng = np.random.default_rng(5000 + int(profile))
    noise = rng.normal(0.0, 1.0, n_nodes + 12)
    kernel = np.exp(-0.5 * (np.arange(-6, 7) / 2.0) ** 2)
    smooth = np.convolve(noise, kernel / kernel.sum(), mode="valid")[:n_nodes]
    smooth *= sigma / math.sqrt(float(np.mean(smooth * smooth)))
    return depths, munk + smooth


def config5_ocean(speeds, bottom_depth=4500.0):
    """Config 5's ocean for given node sound speeds: a 1-D profile_grid
    linear in rho (axis rho = R - depth, edge limits on), Thorp attenuation,
    flat bottom at bottom_depth with Rayleigh sand, flat surface."""
    R = earth_radius()
    n = len(speeds)
    rho_axis = seq_linear(R, -bottom_depth / (n - 1), n)
    grid = oc.data_grid([rho_axis], np.asarray(speeds, dtype=np.float64), [oc.LINEAR], [True])
    profile = oc.profile_grid(grid, oc.attenuation_thorp())
    bottom = oc.boundary_flat(bottom_depth, oc.reflect_loss_rayleigh("sand"))
    return oc.ocean_model(oc.boundary_flat(), bottom, profile)


def config5(profile=0, de=None, az=None, n_targets=(50, 50), time_max=120.0, freq=None, speeds=None,
            radius=6371000.0, half_width=0.5, source_depth=1000.0):
    """Config 5's forward run for one of its 32 profiles: 64-node piecewise
    linear c(z) (seeded Munk plus 2 m/s smoothed noise), range-independent,
    flat 4500 m bottom; DE -45:0.1:45 (901), AZ 0:5:355 (72), 500-4000 Hz
    (4 octave-spaced frequencies), dt 0.05 s, t < 120 s, 50 x 50 targets
    within 0.5 deg of the source (45N 45W, 1000 m) at 100-2000 m depth.
    `radius`, `half_width` (deg) and `source_depth` change the earth radius,
    the target box and the source depth of derivative checks."""
    set_earth_radius(radius)
    if speeds is None:
        speeds = config5_profile(profile)[1]
    ocean = config5_ocean(speeds)
    freq = freq if freq is not None else seq_log(500.0, 2.0, 4)
    pos = wposition1(45.0, -45.0, -source_depth)
    de = de if de is not None else seq_linear(-45.0, 0.1, 45.0)
    az = az if az is not None else seq_linear(0.0, 5.0, 355.0)
    tg = (target_grid(45.0, -45.0, half_width, n_targets[0], n_targets[1], 100.0, 2000.0)
          if n_targets else None)
    sc = scenario("config5", ocean, freq, pos, de, az, 0.05, tg, loop_steps(0.05, time_max))
    sc.notes["profile"] = profile
    return sc


def config2_analytic(n_targets=(100, 100), time_max=120.0):
    """Config 2's fan, frequencies, targets and run length (641 x 180 rays,
    8 log-spaced frequencies 100-10000 Hz, 100 x 100 targets, dt 0.1 s,
    t < 120 s) in config 1's analytic environment (Munk profile, flat 5000 m
    Rayleigh-sand bottom, Thorp attenuation)."""
    sc = config1(de=seq_linear(-80.0, 0.25, 80.0), az=seq_linear(0.0, 2.0, 358.0),
                 n_targets=n_targets, time_max=time_max)
    sc.freq = seq_log(100.0, 10.0 ** (2.0 / 7.0), 8)
    sc.name = "config2_analytic"
    return sc


def straight_ray_targets():
    """Straight rays for the sign of compute_offsets' DE offset
    (wave_queue.cc:863-912): an isovelocity ocean (1500 m/s, no absorption)
    with its boundaries 10 km above and below the source (45N 45W, 10 km
    deep), a coarse fan (DE -60:5:60, AZ 0:15:360 closing on itself), 1 kHz,
    dt 0.1 s for 6 s, and 12 x 12 seeded targets within 3 km of the source's
    depth and 0.05 deg of its position.  Direct paths only; each target's
    launch elevation is that of the straight chord from the source."""
    compute_earth_radius(45.0)
    profile = oc.profile_linear(1500.0, oc.attenuation_constant(0.0))
    ocean = oc.ocean_model(oc.boundary_flat(), oc.boundary_flat(20000.0), profile)
    rng = np.random.default_rng(3)
    n1 = n2 = 12
    lat = np.zeros((n1, n2))
    lng = np.zeros((n1, n2))
    alt = np.zeros((n1, n2))
    for a in range(n1):
        for b in range(n2):
            lat[a, b] = 45.0 + rng.uniform(-0.04, 0.04)
            lng[a, b] = -45.0 + rng.uniform(-0.05, 0.05)
            alt[a, b] = -10000.0 + rng.uniform(-3000.0, 3000.0)
    sc = scenario("straight_ray_targets", ocean, seq_log(1000.0, 1.0, 1), wposition1(45.0, -45.0, -10000.0),
                  seq_linear(-60.0, 5.0, 60.0), seq_linear(0.0, 15.0, 360.0), 0.1,
                  wposition.from_arrays(lat, lng, alt), loop_steps(0.1, 6.0))
    sc.notes["targets"] = (lat, lng, alt)
    return sc


def chord_elevation_deg(source, lat, lng, alt):
    """Elevation (deg, up positive) of the straight chord from `source`
    (a wposition1) to the point (lat, lng, alt) on the current earth
    sphere: the launch DE of a straight ray that reaches it."""
    R = earth_radius()

    def xyz(rho, lat_deg, lng_deg):
        th = math.radians(90.0 - lat_deg)
        ph = math.radians(lng_deg)
        return np.array([rho * math.sin(th) * math.cos(ph), rho * math.sin(th) * math.sin(ph), rho * math.cos(th)])

    s = xyz(R + source.altitude(), source.latitude(), source.longitude())
    d = xyz(R + alt, lat, lng) - s
    return math.degrees(math.asin(float(np.dot(d, s)) / (np.linalg.norm(d) * np.linalg.norm(s))))


def eigenray_branch_pt():
    """waveq3d/test/eigenray_test.cc:428-534 scenario: AZ fan closing on
    itself, DE -90:1:90, 12 targets on a 2226 m ring plus targets 500 m
    above and below the source (the source-branch rule)."""
    compute_earth_radius(0.0)
    profile = oc.profile_linear(1500.0, oc.attenuation_constant(0.0))
    ocean = oc.ocean_model(oc.boundary_flat(), oc.boundary_flat(3000.0), profile)
    freq = seq_log(1000.0, 1.0, 1)
    pos = wposition1(0.0, 0.0, -1000.0)
    de = seq_linear(-90.0, 1.0, 90.0)
    az = seq_linear(0.0, 15.0, 360.0)
    n = 12
    lat = np.zeros((n + 2, 1))
    lng = np.zeros((n + 2, 1))
    alt = np.full((n + 2, 1), -1000.0)
    alt[0, 0] = -1500.0
    alt[1, 0] = -500.0
    bearing = 0.0
    for k in range(2, n + 2):
        p = wposition1.from_range_bearing(pos, 2226.0, bearing)
        lat[k, 0], lng[k, 0], alt[k, 0] = p.latitude(), p.longitude(), p.altitude()
        bearing = bearing + 2.0 * math.pi / n
    target = wposition.from_arrays(lat, lng, alt)
    return scenario("eigenray_branch_pt", ocean, freq, pos, de, az, 0.1, target,
                    loop_steps(0.1, 3.5))


def proploss_classic():
    """waveq3d/test/proploss_test.cc:50-145 scenario: classic ray spreading
    (CLASSIC_RAY) in an isovelocity ocean (1500 m/s, no absorption) with the
    bottom 100 km down; source 45N 45W 15 km deep; DE -10.7:1:10, AZ
    -10.5:2:10; 10, 100, 1000 and 10000 Hz; dt 0.1 s for 8 s; ten targets due
    north at the source's depth, 0.02 to 0.11 deg away."""
    compute_earth_radius(45.0)
    profile = oc.profile_linear(1500.0, oc.attenuation_constant(0.0))
    ocean = oc.ocean_model(oc.boundary_flat(), oc.boundary_flat(1e5), profile)
    freq = seq_log(10.0, 10.0, 4)
    pos = wposition1(45.0, -45.0, -15000.0)
    de = seq_linear(-10.7, 1.0, 10.0)
    az = seq_linear(-10.5, 2.0, 10.0)
    lat = np.array([[45.0 + 0.01 * (n + 2.0)] for n in range(10)])
    target = wposition.from_arrays(lat, np.full((10, 1), -45.0), np.full((10, 1), -15000.0))
    sc = scenario("proploss_classic", ocean, freq, pos, de, az, 0.1, target, loop_steps(0.1, 8.0))
    sc.spreading = oc._abi.SPREADING_CLASSIC_RAY
    return sc


def bounce_threshold():
    """waveq3d/test/reflection_test.cc:597-656 scenario: isovelocity 1500 m/s,
    lossless flat 1000 m bottom, flat surface; source on the equator 500 m
    deep; DE -90:1:90, AZ 0:30:360 (a fan closing on itself), 900 Hz; one
    target 0.01 deg north at the source's depth; dt 0.1 s for 10 s.  The
    reference test caps surface bounces at 3 and bottom bounces at 5."""
    set_earth_radius(6371000.0)
    profile = oc.profile_linear(1500.0)
    ocean = oc.ocean_model(oc.boundary_flat(), oc.boundary_flat(1000.0, oc.reflect_loss_constant(0.0)),
                           profile)
    freq = seq_log(900.0, 1.0, 1)
    pos = wposition1(0.0, 0.0, -500.0)
    de = seq_linear(-90.0, 1.0, 90.0)
    az = seq_linear(0.0, 30.0, 360.0)
    target = wposition(1, 1, 0.01, 0.0, -500.0)
    return scenario("bounce_threshold", ocean, freq, pos, de, az, 0.1, target, loop_steps(0.1, 10.0))


def proploss_gaussian():
    """waveq3d/test/proploss_test.cc:153-245 scenario: hybrid Gaussian beams
    in an isovelocity ocean (1500 m/s, no absorption, flat 3000 m bottom);
    source 45N 45W 1000 m deep; DE -90:1:90, AZ -5:1:5; 100-1000 Hz in
    100 Hz steps; dt 0.1 s for 3.5 s; one target 0.02 deg north at the
    source's depth (direct, surface and bottom paths)."""
    compute_earth_radius(45.0)
    profile = oc.profile_linear(1500.0, oc.attenuation_constant(0.0))
    ocean = oc.ocean_model(oc.boundary_flat(), oc.boundary_flat(3000.0), profile)
    freq = seq_linear(100.0, 100.0, 1000.0)
    pos = wposition1(45.0, -45.0, -1000.0)
    target = wposition(1, 1, 45.02, -45.0, -1000.0)
    return scenario("proploss_gaussian", ocean, freq, pos, seq_linear(-90.0, 1.0, 90.0),
                    seq_linear(-5.0, 1.0, 5.0), 0.1, target, loop_steps(0.1, 3.5))


def _lloyds_ocean():
    profile = oc.profile_linear(1500.0, oc.attenuation_constant(0.0))
    profile.flat_earth(True)
    return oc.ocean_model(oc.boundary_flat(), oc.boundary_flat(1e5), profile)


def lloyds_range():
    """waveq3d/test/proploss_test.cc:357-410 scenario: Lloyd's mirror on a
    flat earth (isovelocity 1500 m/s, no absorption, bottom 100 km down);
    source 45N 45W 25 m deep; DE 181 tangent-spaced rays, AZ -20:5:20;
    500-2000 Hz in 500 Hz steps; dt 0.1 s for 8 s; 1961 targets 200 m deep
    due north at ranges 200 m to 10 km in 5 m steps (one nautical mile per
    arc minute of latitude).  The earth radius is the one the reference's
    test suite leaves behind (its first case computes it at 45N)."""
    compute_earth_radius(45.0)
    ocean = _lloyds_ocean()
    freq = seq_linear(500.0, 500.0, 4)
    pos = wposition1(45.0, -45.0, -25.0)
    ranges = seq_linear(200.0, 5.0, 10e3)
    lat = np.array([[45.0 + ranges(n) / (1852.0 * 60.0)] for n in range(ranges.size())])
    target = wposition.from_arrays(lat, np.full(lat.shape, -45.0), np.full(lat.shape, -200.0))
    sc = scenario("lloyds_range", ocean, freq, pos, seq_rayfan(), seq_linear(-20.0, 5.0, 20.0), 0.1,
                  target, loop_steps(0.1, 8.0))
    sc.notes = {"ranges": np.array([ranges(n) for n in range(ranges.size())]), "src_alt": -25.0,
                "trg_alt": -200.0}
    return sc


def lloyds_depth():
    """waveq3d/test/proploss_test.cc:562-620 scenario: the Lloyd's mirror
    ocean and fan of lloyds_range with 101 targets 10 km north of the
    source at depths 0.1 to 50.1 m in 0.5 m steps."""
    compute_earth_radius(45.0)
    ocean = _lloyds_ocean()
    freq = seq_linear(500.0, 500.0, 4)
    pos = wposition1(45.0, -45.0, -25.0)
    lat0 = 45.0 + math.degrees(10e3 / (earth_radius() - 25.0))
    depths = seq_linear(-0.1, -0.5, -50.1)
    alt = np.array([[depths(n)] for n in range(depths.size())])
    target = wposition.from_arrays(np.full(alt.shape, lat0), np.full(alt.shape, -45.0), alt)
    sc = scenario("lloyds_depth", ocean, freq, pos, seq_rayfan(), seq_linear(-20.0, 5.0, 20.0), 0.1,
                  target, loop_steps(0.1, 8.0))
    sc.notes = {"depths": alt[:, 0].copy(), "range": 10e3, "src_alt": -25.0}
    return sc


def eigenray_basic():
    """waveq3d/test/eigenray_test.cc:86-198 scenario."""
    compute_earth_radius(45.0)
    bottom = oc.boundary_flat(3000.0)
    surface = oc.boundary_flat()
    profile = oc.profile_linear(1500.0, oc.attenuation_constant(0.0))
    ocean = oc.ocean_model(surface, bottom, profile)
    freq = seq_log(10e3, 2.0, 3)
    pos = wposition1(45.0, -45.0, -1000.0)
    de = seq_linear(-60.0, 5.0, 60.0)
    az = seq_linear(-4.0, 1.0, 4.0)
    target = wposition(1, 1, 45.02, -45.0, -1000.0)
    return scenario("eigenray_basic", ocean, freq, pos, de, az, 0.1, target, loop_steps(0.1, 3.5))


def config4_sources(n_sources=256, seed=4096, time_max=90.0):
    """Config 4's independent sources: seeded positions in a 5 x 5 deg box
    around 45N 45W at 10-2000 m depth, uneven fans (DE count uniform in
    61-721 spread over -80..80 deg, AZ count uniform in 1-181 spread evenly
    round the circle), 4-16 log-spaced frequencies over 100-10000 Hz, dt
    0.1 s, t < time_max.  Returns a list of dicts with the source position,
    fans and frequencies; every source shares one ocean, which the caller
    picks (config 2's gridded one, or config 1's analytic one where grids
    are unavailable)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_sources):
        lat = 45.0 + rng.uniform(-2.5, 2.5)
        lng = -45.0 + rng.uniform(-2.5, 2.5)
        depth = rng.uniform(10.0, 2000.0)
        nde = int(rng.integers(61, 722))
        naz = int(rng.integers(1, 182))
        nf = int(rng.integers(4, 17))
        de = seq_linear(-80.0, 160.0 / (nde - 1), nde)
        az = seq_linear(0.0, 360.0 / naz, naz)
        freq = seq_log(100.0, 100.0 ** (1.0 / (nf - 1)), nf)
        out.append({"pos": wposition1(lat, lng, -depth), "de": de, "az": az, "freq": freq,
                    "time_step": 0.1, "time_max": time_max, "steps": loop_steps(0.1, time_max)})
    return out
