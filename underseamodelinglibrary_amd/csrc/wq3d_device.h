// wq3d_device.h - FP64 device arithmetic of the WaveQ3D hot path (gfx950).
//
// Every expression keeps the evaluation order of the reference's element-wise
// uBLAS expressions (the library is compiled with -ffp-contract=off), so the
// kernels reproduce the reference arithmetic operation for operation.  Each
// function cites the reference file:line it implements.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wq3d_math.h"

#if defined(WQ3D_EMULATE) && !defined(__HIP_DEVICE_COMPILE__)
#include <cstdio>
#include <cstdlib>
// CPU emulation build only: abort on values the test scenarios must never
// produce (non-finite grid locations, indices outside an array).  A test
// hook clears the flag to probe a function on such values directly.
inline bool& wq3d_emulation_strict() {
    static bool strict = true;
    return strict;
}
inline void wq3d_emulation_check(bool ok, const char* where, double value) {
    if (!ok && wq3d_emulation_strict()) {
        std::fprintf(stderr, "wq3d emulation check failed in %s: %.17g\n", where, value);
        std::abort();
    }
}
#endif

namespace wq3d {

constexpr int kMaxFreq = 32;     // frequencies per fan
constexpr int kMaxAz = 4096;     // AZ rays per fan (duplicate-ray bitset)
constexpr int kMaxVolume = 8;    // volume scattering layers

// ------------------------------------------------------------ environment
struct DSeq {                   // usml::types::seq_vector on the device
    const double* v;
    const double* inc;
    int32_t n;
    int32_t search;             // wq3d_seq_search
    double sign;                // seq_data: -1 for decreasing sequences
};

struct DGrid {                  // gen_grid<N>, N <= 3
    const double* data;
    int64_t size;               // elements of data (product of the axis sizes)
    DSeq ax[3];
    int32_t nd;
    int32_t interp[3];
    int32_t edge[3];
};

struct DProfile {
    int32_t kind;               // wq3d_profile_kind
    int32_t flat_earth;
    double p[4];
    DGrid grid;
    int32_t atten_kind;
    double atten_coeff;
};

struct DBoundary {
    int32_t kind;               // wq3d_boundary_kind
    int32_t loss_kind;          // wq3d_loss_kind
    double rho, normal_rho;
    double loss[5];
    // reflect_loss_rayleigh members (reflect_loss_rayleigh.cc:68-78)
    double den_w, spd_w, den_b, spd_b, att_b, spd_s, att_s;
    DGrid grid;
};

struct DEnv {
    DProfile profile;
    DBoundary surface, bottom;
    int32_t nvol;
    double vol_rho[kMaxVolume];
};

// std::max semantics (NaN in the first argument is returned)
__host__ __device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// ublas/math_traits.h:156-170
__host__ __device__ __forceinline__ double to_degrees(double t) { return t * (180.0 / M_PI); }
__host__ __device__ __forceinline__ double to_radians(double t) { return t * (M_PI / 180.0); }

// types/seq_vector.h:99-143: clamped accessors; negative indices arrive as
// huge size_t values in the reference and clamp to the last element.
__host__ __device__ __forceinline__ int64_t seq_clamp(const DSeq& s, int64_t i) {
    return (i < 0 || i > s.n - 1) ? (int64_t)(s.n - 1) : i;
}
__host__ __device__ __forceinline__ double seq_value(const DSeq& s, int64_t i) { return s.v[seq_clamp(s, i)]; }
__host__ __device__ __forceinline__ double seq_inc(const DSeq& s, int64_t i) { return s.inc[seq_clamp(s, i)]; }

// seq_linear::find_index (types/seq_linear.h:89-94), seq_log (seq_log.h:84),
// seq_data (seq_data.cc:15-36).  The reference converts floor(q) to a signed
// 64-bit integer and clamps it to [0, size-2].  Here q is clamped in double
// first, so the conversion is defined for every input: a NaN q, or one at
// or above 2^63, gives index 0, which is what the reference's conversion
// gives on x86-64 (the "integer indefinite" INT64_MIN, clamped to 0).  No
// input yields an index outside [0, size-2].
__host__ __device__ __forceinline__ int clamp_floor_index(double q, int n2) {
    if (!(q < 0x1p63)) return 0;
    if (q <= 0.0) return 0;
    if (q >= (double)n2) return n2;
    return (int)q;
}

__host__ __device__ inline int find_index(const DSeq& s, double value) {
#if defined(WQ3D_EMULATE) && !defined(__HIP_DEVICE_COMPILE__)
    wq3d_emulation_check(value == value && value > -1e15 && value < 1e15, "find_index", value);
#endif
    const int n2 = s.n - 2 > 0 ? s.n - 2 : 0;
    if (s.search == 0) return clamp_floor_index(floor((value - s.v[0]) / s.inc[0]), n2);
    if (s.search == 1)
        return clamp_floor_index(floor(wqm::log(value / s.v[0]) / wqm::log(s.inc[1] / s.inc[0])), n2);
    value *= s.sign;
    const int mi = s.n - 1;
    if (mi < 1 || !(value == value)) return 0;
    if (value <= s.sign * s.v[0]) return 0;
    if (value >= s.sign * s.v[mi - 1]) return mi - 1;
    int lo = 0, hi = mi;        // first index with sign*v > value
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (s.sign * s.v[mid] <= value) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

// ------------------------------------------------------------------ grids
// gen_grid::interp recursion (types/gen_grid.h:157-520) with compile-time
// dimension.  dv[] mirrors the reference's deriv_vec, including the
// overwrite order of its recursive calls; want == (deriv_vec != nullptr).
__host__ __device__ __forceinline__ int64_t grid_offset(const DGrid& g, const int idx[3]) {
    int64_t off = idx[0];
    for (int d = 1; d < g.nd; ++d) off = idx[d] + (int64_t)g.ax[d].n * off;
    return off;
}

template <int DIM>
__host__ __device__ double grid_interp(const DGrid& g, const int idx[3], const double loc[3],
                                       double& deriv, double dv[3], bool want);

__host__ __device__ __forceinline__ void pchip_harmonic(double d0, double d1, double dd0, double dd1,
                                               double w0, double w1, bool deriv, double& m,
                                               double& dm) {
    if (d0 * d1 > 0.0) m = (w0 + w1) / (w0 / d0 + w1 / d1);
    if (deriv && dd0 * dd1 > 0.0) dm = (w0 + w1) / (w0 / dd0 + w1 / dd1);
}

__host__ __device__ __forceinline__ void pchip_end_point(double d0, double d1, double dd0, double dd1,
                                                bool deriv, double& m, double& dm) {
    if (m * d0 < 0.0) {
        m = 0.0;
    } else if ((d0 * d1 < 0.0) && (fabs(m) > fabs(3.0 * d1))) {
        m = 3.0 * d1;
    }
    if (deriv) {
        if (dm * dd0 < 0.0) {
            dm = 0.0;
        } else if ((dd0 * dd1 < 0.0) && (fabs(dm) > fabs(3.0 * dd0))) {
            dm = 3.0 * dd0;
        }
    } else {
        dm = 0.0;
    }
}

template <int DIM>
__host__ __device__ __noinline__ double grid_level(const DGrid& g, const int idx[3], const double loc[3],
                                          double& deriv, double dv[3], bool want) {
    const DSeq& ax = g.ax[DIM];
    const int k = idx[DIM];
    const int type = g.interp[DIM];
    if (type == 0) {  // nearest (gen_grid.h:193-220)
        double da = 0.0, result;
        const double u = (loc[DIM] - seq_value(ax, k)) / seq_inc(ax, k);
        if (u < 0.5) {
            result = grid_interp<DIM - 1>(g, idx, loc, da, dv, want);
        } else {
            int nx[3] = {idx[0], idx[1], idx[2]};
            ++nx[DIM];
            result = grid_interp<DIM - 1>(g, nx, loc, da, dv, want);
        }
        if (want) {
            dv[DIM] = deriv;
            if (DIM > 0) dv[DIM > 0 ? DIM - 1 : 0] = da;
        }
        return result;
    }
    if (type == 1) {  // linear (gen_grid.h:248-283)
        double da = 0.0, db = 0.0;
        const double a = grid_interp<DIM - 1>(g, idx, loc, da, dv, want);
        int nx[3] = {idx[0], idx[1], idx[2]};
        ++nx[DIM];
        const double b = grid_interp<DIM - 1>(g, nx, loc, db, dv, want);
        const double h = seq_inc(ax, k);
        const double u = (loc[DIM] - seq_value(ax, k)) / h;
        const double result = a * (1.0 - u) + b * u;
        if (want) {
            deriv = (b - a) / h;
            dv[DIM] = deriv;
            if (DIM > 0) dv[DIM > 0 ? DIM - 1 : 0] = da * (1.0 - u) + db * u;
        }
        return result;
    }
    // pchip (gen_grid.h:350-520), end-point slopes as the reference writes them
    const int kmin = 1;
    const int kmax = ax.n - 3;
    double y0, y1, y2, y3, dy0 = 0.0, dy1 = 0.0, dy2 = 0.0, dy3 = 0.0;
    y1 = grid_interp<DIM - 1>(g, idx, loc, dy1, dv, want);
    if (k >= kmin) {
        int pv[3] = {idx[0], idx[1], idx[2]};
        --pv[DIM];
        y0 = grid_interp<DIM - 1>(g, pv, loc, dy0, dv, want);
    } else {
        y0 = y1;
        dy0 = dy1;
    }
    int nx[3] = {idx[0], idx[1], idx[2]};
    ++nx[DIM];
    y2 = grid_interp<DIM - 1>(g, nx, loc, dy2, dv, want);
    if (k <= kmax) {
        int ls[3] = {nx[0], nx[1], nx[2]};
        ++ls[DIM];
        y3 = grid_interp<DIM - 1>(g, ls, loc, dy3, dv, want);
    } else {
        y3 = y2;
        dy3 = dy2;
    }
    const double h0 = seq_inc(ax, k - 1);
    const double h1 = seq_inc(ax, k);
    const double h2 = seq_inc(ax, k + 1);
    const double h1_2 = h1 * h1;
    const double h1_3 = h1_2 * h1;
    const double s = loc[DIM] - seq_value(ax, k);
    const double s_2 = s * s, s_3 = s_2 * s;
    const double sh_minus = s - h1;
    const double sh_term = 3.0 * h1 * s_2 - 2.0 * s_3;
    const double deriv0 = (y1 - y0) / h0;
    const double deriv1 = (y2 - y1) / h1;
    const double deriv2 = (y3 - y2) / h2;
    double dderiv0 = 0.0, dderiv1 = 0.0, dderiv2 = 0.0;
    if (want) {
        dderiv0 = (dy1 - dy0) / h0;
        dderiv1 = (dy2 - dy1) / h1;
        dderiv2 = (dy3 - dy2) / h2;
    }
    double slope1 = 0.0, dslope1 = 0.0;
    if (k >= kmin) {
        const double w0 = 2.0 * h1 + h0;
        const double w1 = h1 + 2.0 * h0;
        pchip_harmonic(deriv0, deriv1, dderiv0, dderiv1, w0, w1, want, slope1, dslope1);
    } else {
        slope1 = ((2.0 + h1 + h2) * deriv1 - h1 * deriv2) / (h1 + h2);
        dslope1 = ((2.0 + h1 + h2) * dderiv1 - h1 * dderiv2) / (h1 + h2);
        pchip_end_point(deriv1, deriv2, dderiv1, dderiv2, want, slope1, dslope1);
    }
    double slope2 = 0.0, dslope2 = 0.0;
    if (k <= kmax) {
        const double w1 = 2.0 * h1 + h0;
        const double w2 = h1 + 2.0 * h0;
        pchip_harmonic(deriv1, deriv2, dderiv1, dderiv2, w1, w2, want, slope2, dslope2);
    } else {
        slope2 = ((2.0 + h1 + h2) * deriv1 - h1 * deriv0) / (h1 + h0);
        dslope2 = ((2.0 + h1 + h2) * dderiv1 - h1 * dderiv0) / (h1 + h0);
        pchip_end_point(deriv1, deriv0, dderiv1, dderiv0, want, slope2, dslope2);
    }
    const double result = y2 * sh_term / h1_3 + y1 * (h1_3 - sh_term) / h1_3 +
                          slope2 * s_2 * sh_minus / h1_2 + slope1 * s * sh_minus * sh_minus / h1_2;
    if (want) {
        const double v = s / h1;
        deriv = slope1 * (1.0 - v) + slope2 * v;
        dv[DIM] = deriv;
        if (DIM > 0) {
            dv[DIM > 0 ? DIM - 1 : 0] = dy2 * sh_term / h1_3 + dy1 * (h1_3 - sh_term) / h1_3 +
                                        dslope2 * s_2 * sh_minus / h1_2 +
                                        dslope1 * s * sh_minus * sh_minus / h1_2;
        }
    }
    return result;
}

template <int DIM>
__host__ __device__ __forceinline__ double grid_interp(const DGrid& g, const int idx[3],
                                                       const double loc[3], double& deriv, double dv[3], bool want) {
    return grid_level<DIM>(g, idx, loc, deriv, dv, want);
}

template <>
__host__ __device__ __forceinline__ double grid_interp<-1>(const DGrid& g, const int idx[3],
                                                  const double loc[3], double& deriv, double dv[3],
                                                  bool want) {
    const int64_t off = grid_offset(g, idx);
#if defined(WQ3D_EMULATE) && !defined(__HIP_DEVICE_COMPILE__)
    wq3d_emulation_check(off >= 0 && off < g.size, "grid_offset", (double)off);
#endif
    // every index is in range by construction (find_index and the edge
    // limits); the select keeps a read inside the grid's data regardless
    return g.data[(uint64_t)off < (uint64_t)g.size ? off : 0];
}

// gen_grid::interpolate (types/gen_grid.h:83-133): edge limiting and index
// search, then the recursion from the last dimension down.
template <int ND>
__host__ __device__ double grid_interpolate(const DGrid& g, const double location[], double dv[3],
                                   bool want) {
    int idx[3] = {0, 0, 0};
    double loc[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ND; ++d) loc[d] = location[d];
    for (int d = 0; d < ND; ++d) {
        const DSeq& a = g.ax[d];
        if (g.edge[d]) {
            const double lo = a.v[0];
            const double hi = a.v[a.n - 1];
            const double sgn = (a.inc[0] < 0) ? -1.0 : 1.0;
            const double x = loc[d] * sgn;
            if (x <= lo * sgn) {
                loc[d] = lo;
                idx[d] = 0;
            } else if (x >= hi * sgn) {
                loc[d] = hi;
                idx[d] = a.n - 2;
            } else {
                idx[d] = find_index(a, loc[d]);
            }
        } else {
            idx[d] = find_index(a, loc[d]);
        }
    }
    double dresult = 0.0;
    return grid_interp<ND - 1>(g, idx, loc, dresult, dv, want);
}

// profile_grid<1> with linear interpolation in rho: gen_grid::interpolate's
// edge limit and index search (types/gen_grid.h:83-133) and the linear
// level (:248-283), the same operations as grid_interpolate<1> with
// grid_level<0>, written inline.  The index is in [0, n-2] for every input.
__host__ __device__ __forceinline__ double grid1_linear(const DGrid& g, double rho, double& grho, bool want) {
    const DSeq& a = g.ax[0];
    double loc = rho;
    int k;
    if (g.edge[0]) {
        const double lo = a.v[0];
        const double hi = a.v[a.n - 1];
        const double sgn = (a.inc[0] < 0) ? -1.0 : 1.0;
        const double x = loc * sgn;
        if (x <= lo * sgn) {
            loc = lo;
            k = 0;
        } else if (x >= hi * sgn) {
            loc = hi;
            k = a.n - 2;
        } else {
            k = find_index(a, loc);
        }
    } else {
        k = find_index(a, loc);
    }
    const double va = g.data[k];
    const double vb = g.data[k + 1];
    const double h = seq_inc(a, k);
    const double u = (loc - seq_value(a, k)) / h;
    if (want) grho = (vb - va) / h;
    return va * (1.0 - u) + vb * u;
}

// profile kind of a kernel variant for a 1-D linear profile grid (not an
// ABI value: the descriptor says WQ3D_PROFILE_GRID)
constexpr int kProfileGrid1Linear = 3;

__host__ __device__ __forceinline__ bool is_grid1_linear(const DProfile& P) {
    return P.kind == 2 && P.grid.nd == 1 && P.grid.interp[0] == 1;
}

// Environment kinds known at compile time.  The functions on the
// propagation and reflection path take a policy K: with AnyKinds they read
// the kinds from the descriptor; with Kinds<PK, FLAT> the profile kind is PK
// and both boundaries are flat, so the other kinds' code (the grid
// interpolation above all) is not compiled in.
struct AnyKinds {
    static constexpr int kProfile = -1;
    static constexpr bool kFlat = false;
};
template <int PK, bool FLAT>
struct Kinds {
    static constexpr int kProfile = PK;
    static constexpr bool kFlat = FLAT;
};

// ------------------------------------------------------------- profiles
// profile_munk.cc:19-32, profile_linear.cc:14-40, profile_grid.h:68-130,
// profile_model::adjust_speed (profile_model.cc:14-26)
template <class K = AnyKinds>
__host__ __device__ __forceinline__ void sound_speed(const DProfile& P, double rho, double theta, double phi,
                                   double R, double& c, double g[3], bool want_gradient) {
    double grho = 0.0, gtheta = 0.0, gphi = 0.0, speed;
    const int kind = K::kProfile >= 0 ? K::kProfile : P.kind;
    if (kind == 1) {
        const double alt = rho - R;
        const double z = 2 * (-alt - P.p[0]) / P.p[1];
        speed = (((z - 1.0) + wqm::exp(-z)) * P.p[3] + 1.0) * P.p[2];
        grho = (1.0 - wqm::exp(-z)) * (-P.p[3] * P.p[2] * 2 / P.p[1]);
    } else if (kind == 0) {
        const double z = -(rho - R);
        if (z < P.p[2]) {
            speed = P.p[0] + P.p[1] * z;
            grho = -P.p[1];
        } else {
            speed = P.p[0] + P.p[1] * P.p[2] + P.p[3] * (z - P.p[2]);
            grho = -P.p[3];
        }
    } else if (kind == kProfileGrid1Linear || is_grid1_linear(P)) {
        speed = grid1_linear(P.grid, rho, grho, want_gradient);
    } else if (P.grid.nd == 1) {
        double loc[1] = {rho};
        double dv[3] = {0.0, 0.0, 0.0};
        speed = grid_interpolate<1>(P.grid, loc, dv, want_gradient);
        grho = dv[0];
    } else {
        double loc[3] = {rho, theta, phi};
        double dv[3] = {0.0, 0.0, 0.0};
        speed = grid_interpolate<3>(P.grid, loc, dv, want_gradient);
        grho = dv[0];
        gtheta = dv[1];
        gphi = dv[2];
    }
    if (P.flat_earth) {
        grho = (grho * rho + speed) / R;
        speed = speed * rho / R;
    }
    c = speed;
    g[0] = grho;
    g[1] = gtheta;
    g[2] = gphi;
}

// attenuation_thorp.cc:14-36 (coefficient recomputed as the reference does),
// attenuation_constant.cc:50-62
__host__ __device__ __forceinline__ double attenuation(const DProfile& P, double freq, double alt,
                                              double distance) {
    if (P.atten_kind == 1) return P.atten_coeff * distance * freq;
    double F2 = freq;
    F2 = 1e-6 * F2 * F2;
    const double alpha = 1e-3 * (3.3e-3 + F2 * (0.11 / (1.0 + F2) + 44.0 / (4100.0 + F2) + 3.0e-4)) /
                         (1.0 - 5.88264e-6 * 1000.0);
    return distance * alpha * (1.0 + 5.88264e-6 * alt);
}

// ------------------------------------------------------------ boundaries
// boundary_flat.cc:55-63, boundary_grid.h:145-196
template <class K = AnyKinds>
__host__ __device__ __forceinline__ void boundary_height(const DBoundary& B, double theta, double phi, double& h,
                                       double n[3], bool want_normal) {
    if (K::kFlat || B.kind == 0) {
        h = B.rho;
        n[0] = B.normal_rho;
        n[1] = 0.0;
        n[2] = 0.0;
        return;
    }
    if (B.grid.nd == 1) {
        double loc[1] = {theta};
        double g[3] = {0.0, 0.0, 0.0};
        h = grid_interpolate<1>(B.grid, loc, g, want_normal);
        if (want_normal) {
            const double t = -g[0] / h;
            const double N = sqrt(1 + t * t);
            n[0] = 1.0 / N;
            n[1] = t / N;
            n[2] = 0.0;
        }
    } else {
        double loc[2] = {theta, phi};
        double g[3] = {0.0, 0.0, 0.0};
        h = grid_interpolate<2>(B.grid, loc, g, want_normal);
        if (want_normal) {
            const double t = -g[0] / h;
            const double p = -g[1] / (h * wqm::sin(theta));
            const double N = sqrt(1 + t * t + p * p);
            n[0] = 1.0 / N;
            n[1] = t / N;
            n[2] = p / N;
        }
    }
}

using wqm::cadd;
using wqm::cdiv;
using wqm::cmul;
using wqm::cplx;
using wqm::cscale;
using wqm::csub;

// reflect_loss_rayleigh::impedance (reflect_loss_rayleigh.cc:129-144)
__host__ __device__ inline cplx rayleigh_impedance(const DBoundary& B, double density, double speed,
                                          double att, double angle, cplx& cosA, bool shear) {
    const cplx c = {speed, -att * speed};
    const cplx sinA = cscale(c, wqm::sin(angle));
    const cplx sA = {sinA.re / B.spd_w, sinA.im / B.spd_w};
    cosA = wqm::csqrt(csub({1.0, 0.0}, cmul(sA, sA)));
    if (shear) return cdiv(cscale(c, density), cosA);
    return cdiv(cosA, cscale(c, density));
}

// reflect_loss_constant.h, reflect_loss_rayleigh::reflect_loss (.cc:83-124);
// the same amplitude and phase for every frequency.
__host__ __device__ inline void reflect_loss(const DBoundary& B, double angle, double& amp, double& phase) {
    if (B.loss_kind == 0) {
        amp = B.loss[0];
        phase = B.loss[1];
        return;
    }
    if (angle >= M_PI_2) angle = M_PI_2 - 1e-10;
    angle = M_PI_2 - angle;
    const double Zw = wqm::cos(angle) / (B.spd_w * B.den_w);
    cplx cosAp, cosAs;
    cplx Zb = rayleigh_impedance(B, B.den_b, B.spd_b, B.att_b, angle, cosAp, false);
    if (B.spd_s != 0.0 || B.att_s != 0.0) {
        const cplx Zs = rayleigh_impedance(B, B.den_b, B.spd_s, B.att_s, angle, cosAs, true);
        const cplx one = {1.0, 0.0};
        const cplx sinAs = wqm::csqrt(csub(one, cmul(cosAs, cosAs)));
        const cplx cos2As = csub(cmul(cscale(cosAs, 2.0), cosAs), one);
        const cplx sin2As = cmul(cscale(sinAs, 2.0), cosAs);
        const cplx t1 = cmul(cmul(cdiv(one, Zb), cos2As), cos2As);
        const cplx t2 = cmul(cmul(Zs, sin2As), sin2As);
        Zb = cdiv(one, cadd(t1, t2));
    }
    const cplx Zwc = {Zw, 0.0};
    const cplx Rc = cdiv(csub(Zb, Zwc), cadd(Zb, Zwc));
    amp = -20.0 * wqm::log10(wqm::cabs(Rc)) * 1.0;
    phase = wqm::carg(Rc) * 1.0;
}

// --------------------------------------------------------- one ray's state
struct Ray {
    double pr, pt, pp;   // position (rho, theta, phi)
    double gr, gt, gp;   // pos_gradient
    double nr, nt, np;   // ndirection
    double hr, ht, hp;   // ndir_gradient
    double c, cr, ct, cp;  // sound speed and gradient
    double dist;         // step length (ode_integ.cc:150-155)
    double st;           // sin(theta) at the last update
};

// wave_front::update minus attenuation (wave_front.cc:125-178,241-250)
template <class K = AnyKinds>
__host__ __device__ __forceinline__ void update_ray(const DEnv& E, double R, Ray& w) {
    double g[3];
    sound_speed<K>(E.profile, w.pr, w.pt, w.pp, R, w.c, g, true);
    w.cr = g[0];
    w.ct = g[1];
    w.cp = g[2];
    const double dcc_r = w.cr / w.c;
    const double dcc_t = w.ct / w.c;
    const double dcc_p = w.cp / w.c;
    const double st = wqm::sin(w.pt);
    const double cot = wqm::cos(w.pt) / st;
    double c2r = w.c * w.c;
    w.gr = c2r * w.nr;
    c2r = c2r / w.pr;
    w.gt = c2r * w.nt;
    w.gp = c2r / st * w.np;
    w.hr = c2r * (w.nt * w.nt + w.np * w.np) - dcc_r;
    w.ht = -c2r * (w.nr * w.nt - w.np * w.np * cot) - dcc_t / w.pr;
    w.hp = -c2r * (w.np * (w.nr + w.nt * cot)) - dcc_p / (w.pr * st);
    w.st = st;
}

// ode_integ.cc:20-126 (RK3 stages)
__host__ __device__ __forceinline__ void rk1(double dt, const Ray& y0, Ray& y1) {
    y1.pr = y0.pr + 0.5 * dt * y0.gr;
    y1.pt = y0.pt + 0.5 * dt * y0.gt;
    y1.pp = y0.pp + 0.5 * dt * y0.gp;
    y1.nr = y0.nr + 0.5 * dt * y0.hr;
    y1.nt = y0.nt + 0.5 * dt * y0.ht;
    y1.np = y0.np + 0.5 * dt * y0.hp;
}
__host__ __device__ __forceinline__ void rk2(double dt, const Ray& y0, const Ray& y1, Ray& y2) {
    y2.pr = y0.pr + dt * (2.0 * y1.gr - y0.gr);
    y2.pt = y0.pt + dt * (2.0 * y1.gt - y0.gt);
    y2.pp = y0.pp + dt * (2.0 * y1.gp - y0.gp);
    y2.nr = y0.nr + dt * (2.0 * y1.hr - y0.hr);
    y2.nt = y0.nt + dt * (2.0 * y1.ht - y0.ht);
    y2.np = y0.np + dt * (2.0 * y1.hp - y0.hp);
}
__host__ __device__ __forceinline__ void rk3(double dt, const Ray& y0, const Ray& y1, const Ray& y2,
                                    Ray& y3) {
    const double pr = y0.pr + dt / 6.0 * (y0.gr + 4.0 * y1.gr + y2.gr);
    const double pt = y0.pt + dt / 6.0 * (y0.gt + 4.0 * y1.gt + y2.gt);
    const double pp = y0.pp + dt / 6.0 * (y0.gp + 4.0 * y1.gp + y2.gp);
    const double nr = y0.nr + dt / 6.0 * (y0.hr + 4.0 * y1.hr + y2.hr);
    const double nt = y0.nt + dt / 6.0 * (y0.ht + 4.0 * y1.ht + y2.ht);
    const double np = y0.np + dt / 6.0 * (y0.hp + 4.0 * y1.hp + y2.hp);
    y3.pr = pr;
    y3.pt = pt;
    y3.pp = pp;
    y3.nr = nr;
    y3.nt = nt;
    y3.np = np;
}

// ode_integ.cc:131-186 (AB3 position with step length, AB3 ndirection)
__host__ __device__ __forceinline__ void ab3(double dt, const Ray& y0, const Ray& y1, const Ray& y2,
                                    Ray& y3) {
    const double A2 = 23.0 / 12.0, A1 = 16.0 / 12.0, A0 = 5.0 / 12.0;
    const double dr = dt * (A2 * y2.gr - A1 * y1.gr + A0 * y0.gr);
    const double dth = dt * (A2 * y2.gt - A1 * y1.gt + A0 * y0.gt);
    const double dph = dt * (A2 * y2.gp - A1 * y1.gp + A0 * y0.gp);
    const double a = y2.pr * dth;
    const double b = y2.pr * (wqm::sin(y2.pt) * dph);
    y3.dist = sqrt(dr * dr + a * a + b * b);
    y3.pr = y2.pr + dr;
    y3.pt = y2.pt + dth;
    y3.pp = y2.pp + dph;
    y3.nr = y2.nr + dt * (A2 * y2.hr - A1 * y1.hr + A0 * y0.hr);
    y3.nt = y2.nt + dt * (A2 * y2.ht - A1 * y1.ht + A0 * y0.ht);
    y3.np = y2.np + dt * (A2 * y2.hp - A1 * y1.hp + A0 * y0.hp);
}

// wvector1::distance between two points (types/wvector1.cc:44-73)
__host__ __device__ __forceinline__ double point_distance(double r1, double t1, double p1, double r2,
                                                 double t2, double p2) {
    const double sin_theta = wqm::sin(0.5 * (t1 - t2));
    const double sin_phi = wqm::sin(0.5 * (p1 - p2));
    const double dotnorm = 1.0 - 2.0 * (sin_theta * sin_theta + wqm::sin(t1) * wqm::sin(t2) * sin_phi * sin_phi);
    const double dot = r1 * r2 * dotnorm;
    return sqrt(fabs(r1 * r1 + r2 * r2 - 2.0 * dot));
}

// wave_front::compute_target_distance (wave_front.cc:222-236) for one pair;
// hr = 0.5*theta_ray, ht2 = 0.5*theta_target etc. are not pre-scaled so the
// rounding matches the reference expression exactly.
__host__ __device__ __forceinline__ double target_distance2(double rho, double theta, double phi, double st,
                                                   double Rt, double Tt, double Pt, double St) {
    const double h = 0.5 * (theta - Tt);
    const double g = 0.5 * (phi - Pt);
    return fabs(rho * rho + Rt * Rt - 2.0 * Rt * (rho * (1.0 - 2.0 * (h * h + St * (st * (g * g))))));
}

}  // namespace wq3d
