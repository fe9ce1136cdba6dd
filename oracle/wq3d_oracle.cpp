/*
 * wq3d_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * Scalar FP64 CPU restatement of the reference's WaveQ3D wavefront hot path
 * (usml::waveq3d::wave_queue and what it calls), used as the parity oracle
 * for the HIP library and as the CPU baseline of bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product library never links or calls it.
 *
 * Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math), so every
 * expression below is evaluated in the order written, like the reference's
 * element-wise uBLAS expressions.  Each function cites the reference
 * file:line it follows.  Differences from the reference's *structure* (never
 * its arithmetic):
 *   - distance2 (waveq3d/wave_front.cc:222-236) is not stored as a targets x
 *     rays matrix per slot; each slot keeps the position it had at its last
 *     update() and distance2 is evaluated from it on demand with the same
 *     expression.  That is exactly the value the reference's stale matrix
 *     holds after reflection_copy (reflection_model.cc:335-360).
 *   - listeners are replaced by record lists with explicit keys.
 *
 * Pinning (what checks this file: tests/test_oracle_known_answers.py,
 * tests/test_oracle_pins.py, tests/test_oracle_offsets.py, tests/test_math.py,
 * tests/test_collection.py):
 *   the reference's known answers for eigenray_basic, eigenray_concave and
 *   eigenray_branch_pt (travel times, angles, intensities, counts),
 *   proploss_classic, proploss_gaussian, Lloyd's mirror in range and depth
 *   (path counts and types, bias / deviation / determination against the
 *   image-source field), bounce_threshold (the limited run's records are
 *   exactly the free run's within the limits), reflect_flat_test bounce
 *   times, refraction_munk_range cycle ranges and vertex counts,
 *   refraction_isovelocity straight lines, eigenverb_accuracy lengths /
 *   widths / grazing angles, the source clamp, Thorp against the Generic
 *   Sonar Model table, constant attenuation, gen_grid linear / quadratic /
 *   derivative accuracy and exact trilinear reproduction.  Closed forms:
 *   compute_offsets on neighbourhoods sampled from exact quadratics (the
 *   minimum, the determinant typo, the 2x2 and clipped branches); rays
 *   between flat boundaries following their image-source chains through
 *   the reflection restart (within 1 cm over 16 bounces); the phase of
 *   every ray and eigenray as -pi per surface bounce, the bottom loss phase
 *   per bottom bounce and -pi/2 per caustic; pchip interior slopes against
 *   Moler's weighted harmonic mean and its end slopes against the exact
 *   three-point formula with the reference's (2 + h1 + h2) weight; the
 *   Rayleigh loss against the textbook reflection coefficient; the shared
 *   elementary functions against glibc in ulps; the eigenray collection's
 *   191.4 dB total (eigenrays/test/eigenrays_test.cc:57-88) and closed
 *   forms of the coherent / incoherent sums; the hybrid Gaussian walks'
 *   stopping rule against the documented 0.01 dB (10^(0.001)) on traced walk
 *   steps; the caustic fold test against ray theory (caustics = turning
 *   points, or one fewer, for refracted rays in the Munk channel); the
 *   geometric DE sign on compute_offsets' unstable branch by straight rays
 *   (every direct eigenray's DE offset points toward the chord to its
 *   target).  No function is left unpinned.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../include/wq3d.h"
#include "../include/wq3d_math.h"

namespace orc {

static const double kPi = M_PI;
static const double kTwoPi = 2.0 * M_PI;  // ublas/math_traits.h:15

// std::max / std::min semantics, kept explicit because NaN handling matters
static inline double smax(double a, double b) { return (a < b) ? b : a; }

// ublas/math_traits.h:156-170
static inline double to_degrees(double t) { return t * (180.0 / M_PI); }
static inline double to_radians(double t) { return t * (M_PI / 180.0); }
static inline double to_latitude(double t) { return 90.0 - to_degrees(t); }

struct Error {
    std::string msg;
};

// ------------------------------------------------------------------ Seq
// types/seq_vector.h:99-143 (clamped accessors with size_t index wrap),
// types/seq_linear.h:89-94, types/seq_log.h:84, types/seq_data.cc:15-36.
struct Seq {
    std::vector<double> v, inc;
    int search = WQ3D_SEQ_LINEAR;
    size_t max_index = 0;
    double sign = 1.0;

    void load(const wq3d_seq& s) {
        if (s.size < 1 || !s.value || !s.increment) throw Error{"empty sequence"};
        v.assign(s.value, s.value + s.size);
        inc.assign(s.increment, s.increment + s.size);
        search = s.search;
        max_index = (size_t)s.size - 1;
        sign = (s.size > 1 && v[1] < v[0]) ? -1.0 : 1.0;
    }
    size_t size() const { return v.size(); }
    double operator()(size_t i) const { return v[std::min(max_index, i)]; }
    double increment(size_t i) const { return inc[std::min(max_index, i)]; }
    // The reference's (difference_type)floor(q) clamped to [0, size-2]; a
    // NaN q or one at or above 2^63 converts to INT64_MIN on x86-64 (the
    // "integer indefinite" value) and so clamps to 0.  The clamp is made in
    // double here so the conversion is defined for every input.
    static size_t clamp_floor(double q, long n2) {
        if (!(q < 0x1p63) || q <= 0.0) return 0;
        if (q >= (double)n2) return (size_t)std::max(0L, n2);
        return (size_t)q;
    }
    size_t find_index(double value) const {
        const long n2 = (long)size() - 2;
        if (search == WQ3D_SEQ_LINEAR) return clamp_floor(std::floor((value - v[0]) / inc[0]), n2);
        if (search == WQ3D_SEQ_LOG)
            return clamp_floor(std::floor(wqm::log(value / v[0]) / wqm::log(inc[1] / inc[0])), n2);
        value *= sign;
        if (max_index < 1 || !(value == value)) return 0;
        if (value <= sign * v[0]) return 0;
        if (value >= sign * v[max_index - 1]) return max_index - 1;
        size_t n = 0;  // first entry with key > value (std::map::upper_bound)
        while (n <= max_index && sign * v[n] <= value) ++n;
        return n - 1;
    }
};

// ------------------------------------------------------------------ Grid
// gen_grid<N>::interpolate and its recursion (types/gen_grid.h:83-520),
// derivative / end_point_derivative (types/gen_grid_utils.h:19-169).
struct Grid {
    int nd = 0;
    Seq ax[3];
    int interp_type[3] = {1, 1, 1};
    int edge[3] = {1, 1, 1};
    std::vector<double> data;

    void load(const wq3d_grid& g) {
        nd = g.ndims;
        if (nd < 1 || nd > 3 || !g.data) throw Error{"grid must have 1-3 dims and data"};
        size_t total = 1;
        for (int d = 0; d < nd; ++d) {
            ax[d].load(g.axis[d]);
            interp_type[d] = g.interp[d];
            edge[d] = g.edge_limit[d];
            if (ax[d].size() < 2) throw Error{"grid axis shorter than 2"};
            if (interp_type[d] == WQ3D_INTERP_PCHIP && ax[d].size() < 4)
                throw Error{"pchip axis shorter than 4"};
            total *= ax[d].size();
        }
        data.assign(g.data, g.data + total);
    }
    size_t offset(const size_t idx[]) const {
        size_t off = idx[0];
        for (int d = 1; d < nd; ++d) off = idx[d] + ax[d].size() * off;
        return off;
    }

    double interpolate(const double location[], double* deriv_vec) const {
        size_t index[3];
        double loc[3];
        for (int d = 0; d < nd; ++d) loc[d] = location[d];
        for (int dim = 0; dim < nd; ++dim) {
            const Seq& a = ax[dim];
            if (edge[dim]) {
                double lo = a(0);
                double hi = a(a.size() - 1);
                double sgn = (a.increment(0) < 0) ? -1.0 : 1.0;
                double d = loc[dim] * sgn;
                if (d <= lo * sgn) {
                    loc[dim] = lo;
                    index[dim] = 0;
                } else if (d >= hi * sgn) {
                    loc[dim] = hi;
                    index[dim] = a.size() - 2;
                } else {
                    index[dim] = a.find_index(loc[dim]);
                }
            } else {
                index[dim] = a.find_index(loc[dim]);
            }
        }
        double dresult = 0.0;
        return interp(nd - 1, index, loc, dresult, deriv_vec);
    }

    double interp(int dim, const size_t index[], const double loc[], double& deriv,
                  double* dv) const {
        if (dim < 0) return data[offset(index)];
        switch (interp_type[dim]) {
            case WQ3D_INTERP_NEAREST: return nearest(dim, index, loc, deriv, dv);
            case WQ3D_INTERP_LINEAR: return linear(dim, index, loc, deriv, dv);
            default: return pchip(dim, index, loc, deriv, dv);
        }
    }

    double nearest(int dim, const size_t index[], const double loc[], double& deriv,
                   double* dv) const {
        double result;
        double da = 0.0;
        const size_t k = index[dim];
        const double u = (loc[dim] - ax[dim](k)) / ax[dim].increment(k);
        if (u < 0.5) {
            result = interp(dim - 1, index, loc, da, dv);
        } else {
            size_t next[3] = {index[0], index[1], index[2]};
            ++next[dim];
            result = interp(dim - 1, next, loc, da, dv);
        }
        if (dv) {
            dv[dim] = deriv;
            if (dim > 0) dv[dim - 1] = da;
        }
        return result;
    }

    double linear(int dim, const size_t index[], const double loc[], double& deriv,
                  double* dv) const {
        double da = 0.0, db = 0.0;
        const double a = interp(dim - 1, index, loc, da, dv);
        size_t next[3] = {index[0], index[1], index[2]};
        ++next[dim];
        const double b = interp(dim - 1, next, loc, db, dv);
        const size_t k = index[dim];
        const double h = ax[dim].increment(k);
        const double u = (loc[dim] - ax[dim](k)) / h;
        double result = a * (1.0 - u) + b * u;
        if (dv) {
            deriv = (b - a) / h;
            dv[dim] = deriv;
            if (dim > 0) dv[dim - 1] = da * (1.0 - u) + db * u;
        }
        return result;
    }

    static void harmonic(double d0, double d1, double dd0, double dd1, double w0, double w1,
                         bool deriv, double& m, double& dm) {
        if (d0 * d1 > 0.0) m = (w0 + w1) / (w0 / d0 + w1 / d1);
        if (deriv && dd0 * dd1 > 0.0) dm = (w0 + w1) / (w0 / dd0 + w1 / dd1);
    }
    static void end_point(double d0, double d1, double dd0, double dd1, bool deriv, double& m,
                          double& dm) {
        if (m * d0 < 0.0) {
            m = 0.0;
        } else if ((d0 * d1 < 0.0) && (std::fabs(m) > std::fabs(3.0 * d1))) {
            m = 3.0 * d1;
        }
        if (deriv) {
            if (dm * dd0 < 0.0) {
                dm = 0.0;
            } else if ((dd0 * dd1 < 0.0) && (std::fabs(dm) > std::fabs(3.0 * dd0))) {
                dm = 3.0 * dd0;
            }
        } else {
            dm = 0.0;
        }
    }

    double pchip(int dim, const size_t index[], const double loc[], double& deriv,
                 double* dv) const {
        const Seq& a = ax[dim];
        const size_t kmin = 1u;
        const size_t kmax = a.size() - 3u;
        double y0, y1, y2, y3, dy0 = 0.0, dy1 = 0.0, dy2 = 0.0, dy3 = 0.0;
        const size_t k = index[dim];
        y1 = interp(dim - 1, index, loc, dy1, dv);
        if (k >= kmin) {
            size_t prev[3] = {index[0], index[1], index[2]};
            --prev[dim];
            y0 = interp(dim - 1, prev, loc, dy0, dv);
        } else {
            y0 = y1;
            dy0 = dy1;
        }
        size_t next[3] = {index[0], index[1], index[2]};
        ++next[dim];
        y2 = interp(dim - 1, next, loc, dy2, dv);
        if (k <= kmax) {
            size_t last[3] = {next[0], next[1], next[2]};
            ++last[dim];
            y3 = interp(dim - 1, last, loc, dy3, dv);
        } else {
            y3 = y2;
            dy3 = dy2;
        }
        const double h0 = a.increment(k - 1);
        const double h1 = a.increment(k);
        const double h2 = a.increment(k + 1);
        const double h1_2 = h1 * h1;
        const double h1_3 = h1_2 * h1;
        const double s = loc[dim] - a(k);
        const double s_2 = s * s, s_3 = s_2 * s;
        const double sh_minus = s - h1;
        const double sh_term = 3.0 * h1 * s_2 - 2.0 * s_3;
        const double deriv0 = (y1 - y0) / h0;
        const double deriv1 = (y2 - y1) / h1;
        const double deriv2 = (y3 - y2) / h2;
        double dderiv0 = 0.0, dderiv1 = 0.0, dderiv2 = 0.0;
        if (dv) {
            dderiv0 = (dy1 - dy0) / h0;
            dderiv1 = (dy2 - dy1) / h1;
            dderiv2 = (dy3 - dy2) / h2;
        }
        double slope1 = 0.0, dslope1 = 0.0;
        if (k >= kmin) {
            const double w0 = 2.0 * h1 + h0;
            const double w1 = h1 + 2.0 * h0;
            harmonic(deriv0, deriv1, dderiv0, dderiv1, w0, w1, dv != nullptr, slope1, dslope1);
        } else {
            slope1 = ((2.0 + h1 + h2) * deriv1 - h1 * deriv2) / (h1 + h2);
            dslope1 = ((2.0 + h1 + h2) * dderiv1 - h1 * dderiv2) / (h1 + h2);
            end_point(deriv1, deriv2, dderiv1, dderiv2, dv != nullptr, slope1, dslope1);
        }
        double slope2 = 0.0, dslope2 = 0.0;
        if (k <= kmax) {
            const double w1 = 2.0 * h1 + h0;
            const double w2 = h1 + 2.0 * h0;
            harmonic(deriv1, deriv2, dderiv1, dderiv2, w1, w2, dv != nullptr, slope2, dslope2);
        } else {
            slope2 = ((2.0 + h1 + h2) * deriv1 - h1 * deriv0) / (h1 + h0);
            dslope2 = ((2.0 + h1 + h2) * dderiv1 - h1 * dderiv0) / (h1 + h0);
            end_point(deriv1, deriv0, dderiv1, dderiv0, dv != nullptr, slope2, dslope2);
        }
        double result = y2 * sh_term / h1_3 + y1 * (h1_3 - sh_term) / h1_3 +
                        slope2 * s_2 * sh_minus / h1_2 + slope1 * s * sh_minus * sh_minus / h1_2;
        if (dv) {
            double v = s / h1;
            deriv = slope1 * (1.0 - v) + slope2 * v;
            dv[dim] = deriv;
            if (dim > 0) {
                dv[dim - 1] = dy2 * sh_term / h1_3 + dy1 * (h1_3 - sh_term) / h1_3 +
                              dslope2 * s_2 * sh_minus / h1_2 +
                              dslope1 * s * sh_minus * sh_minus / h1_2;
            }
        }
        return result;
    }
};

// --------------------------------------------------------------- Profile
struct Profile {
    int kind = WQ3D_PROFILE_LINEAR;
    bool flat_earth = false;
    double p[4] = {0, 0, 0, 0};
    Grid grid;
    int atten_kind = WQ3D_ATTEN_THORP;
    double atten_coeff = 0.0;

    void load(const wq3d_profile_desc& d) {
        kind = d.kind;
        flat_earth = d.flat_earth != 0;
        for (int i = 0; i < 4; ++i) p[i] = d.p[i];
        if (kind == WQ3D_PROFILE_GRID) {
            grid.load(d.grid);
            if (grid.nd == 2) throw Error{"2-D sound speed grids are not supported"};
        } else if (kind != WQ3D_PROFILE_LINEAR && kind != WQ3D_PROFILE_MUNK) {
            throw Error{"unknown profile kind"};
        }
        atten_kind = d.atten_kind;
        atten_coeff = d.atten_coeff;
    }

    // profile_munk.cc:19-32, profile_linear.cc:14-40, profile_grid.h:68-130,
    // then profile_model::adjust_speed (profile_model.cc:14-26).
    void sound_speed(double rho, double theta, double phi, double R, double* c, double* gr,
                     double* gt, double* gp) const {
        double grho = 0.0, gtheta = 0.0, gphi = 0.0, speed;
        if (kind == WQ3D_PROFILE_MUNK) {
            const double alt = rho - R;
            const double z = 2 * (-alt - p[0]) / p[1];
            speed = (((z - 1.0) + wqm::exp(-z)) * p[3] + 1.0) * p[2];
            grho = (1.0 - wqm::exp(-z)) * (-p[3] * p[2] * 2 / p[1]);
        } else if (kind == WQ3D_PROFILE_LINEAR) {
            const double z = -(rho - R);
            if (z < p[2]) {
                speed = p[0] + p[1] * z;
                grho = -p[1];
            } else {
                speed = p[0] + p[1] * p[2] + p[3] * (z - p[2]);
                grho = -p[3];
            }
        } else if (grid.nd == 1) {
            double loc[1] = {rho};
            double dv[1] = {0.0};
            speed = grid.interpolate(loc, dv);
            grho = dv[0];
        } else {
            double loc[3] = {rho, theta, phi};
            double dv[3] = {0.0, 0.0, 0.0};
            speed = grid.interpolate(loc, dv);
            grho = dv[0];
            gtheta = dv[1];
            gphi = dv[2];
        }
        if (flat_earth) {
            grho = (grho * rho + speed) / R;
            speed = speed * rho / R;
        }
        *c = speed;
        if (gr) {
            *gr = grho;
            *gt = gtheta;
            *gp = gphi;
        }
    }

    // attenuation_thorp.cc:14-36, attenuation_constant.cc:50-62
    void attenuation(double alt, double distance, const Seq& freq, double* out) const {
        for (size_t f = 0; f < freq.size(); ++f) {
            if (atten_kind == WQ3D_ATTEN_CONSTANT) {
                out[f] = atten_coeff * distance * freq(f);
            } else {
                double F2 = freq(f);
                F2 = 1e-6 * F2 * F2;
                const double alpha =
                    1e-3 * (3.3e-3 + F2 * (0.11 / (1.0 + F2) + 44.0 / (4100.0 + F2) + 3.0e-4)) /
                    (1.0 - 5.88264e-6 * 1000.0);
                out[f] = distance * alpha * (1.0 + 5.88264e-6 * alt);
            }
        }
    }
};

// -------------------------------------------------------------- Boundary
struct Boundary {
    int kind = WQ3D_BOUNDARY_FLAT;
    double rho = 0.0, normal_rho = 1.0;
    Grid grid;
    int loss_kind = WQ3D_LOSS_CONSTANT;
    double loss[5] = {0, 0, 0, 0, 0};
    // reflect_loss_rayleigh members (reflect_loss_rayleigh.cc:68-78)
    double den_w = 1000.0, spd_w = 1500.0, den_b = 0, spd_b = 0, att_b = 0, spd_s = 0, att_s = 0;

    void load(const wq3d_boundary_desc& d) {
        kind = d.kind;
        rho = d.rho;
        normal_rho = d.normal_rho;
        if (kind == WQ3D_BOUNDARY_GRID) {
            grid.load(d.grid);
            if (grid.nd > 2) throw Error{"bathymetry must be 1-D or 2-D"};
        } else if (kind != WQ3D_BOUNDARY_FLAT) {
            throw Error{"unknown boundary kind"};
        }
        loss_kind = d.loss_kind;
        for (int i = 0; i < 5; ++i) loss[i] = d.loss[i];
        if (loss_kind == WQ3D_LOSS_RAYLEIGH) {
            const double att_convert = 1.0 / (20.0 * M_LOG10E * kTwoPi);
            den_b = den_w * loss[0];
            spd_b = spd_w * loss[1];
            att_b = loss[2] * att_convert;
            spd_s = spd_w * loss[3];
            att_s = loss[4] * att_convert;
        } else if (loss_kind != WQ3D_LOSS_CONSTANT) {
            throw Error{"unknown reflection loss kind"};
        }
    }

    // boundary_flat.cc:55-63, boundary_grid.h:145-196
    void height(double r, double theta, double phi, double* h, double* n) const {
        if (kind == WQ3D_BOUNDARY_FLAT) {
            *h = rho;
            if (n) {
                n[0] = normal_rho;
                n[1] = 0.0;
                n[2] = 0.0;
            }
            return;
        }
        (void)r;
        if (grid.nd == 1) {
            double loc[1] = {theta};
            double g[1] = {0.0};
            *h = grid.interpolate(loc, n ? g : nullptr);
            if (n) {
                const double t = -g[0] / (*h);
                const double N = std::sqrt(1 + t * t);
                n[0] = 1.0 / N;
                n[1] = t / N;
                n[2] = 0.0;
            }
        } else {
            double loc[2] = {theta, phi};
            double g[2] = {0.0, 0.0};
            *h = grid.interpolate(loc, n ? g : nullptr);
            if (n) {
                const double t = -g[0] / (*h);
                const double p = -g[1] / ((*h) * wqm::sin(theta));
                const double N = std::sqrt(1 + t * t + p * p);
                n[0] = 1.0 / N;
                n[1] = t / N;
                n[2] = p / N;
            }
        }
    }

    wqm::cplx impedance(double density, double speed, double attenuation, double angle,
                        wqm::cplx* cosA, bool shear) const {
        const wqm::cplx c = {speed, -attenuation * speed};
        const wqm::cplx sinA = wqm::cdivs(wqm::cscale(c, wqm::sin(angle)), spd_w);
        *cosA = wqm::csqrt(wqm::csub({1.0, 0.0}, wqm::cmul(sinA, sinA)));
        if (shear) return wqm::cdiv(wqm::cscale(c, density), *cosA);
        return wqm::cdiv(*cosA, wqm::cscale(c, density));
    }

    // reflect_loss_constant.h, reflect_loss_rayleigh.cc:83-124
    void reflect_loss(double angle, size_t nf, double* amp, double* phase) const {
        if (loss_kind == WQ3D_LOSS_CONSTANT) {
            for (size_t f = 0; f < nf; ++f) {
                amp[f] = loss[0];
                if (phase) phase[f] = loss[1];
            }
            return;
        }
        if (angle >= M_PI_2) angle = M_PI_2 - 1e-10;
        angle = M_PI_2 - angle;
        const double Zw = wqm::cos(angle) / (spd_w * den_w);
        wqm::cplx cosAp, cosAs;
        wqm::cplx Zb = impedance(den_b, spd_b, att_b, angle, &cosAp, false);
        if (spd_s != 0.0 || att_s != 0.0) {
            const wqm::cplx one = {1.0, 0.0};
            const wqm::cplx Zs = impedance(den_b, spd_s, att_s, angle, &cosAs, true);
            const wqm::cplx sinAs = wqm::csqrt(wqm::csub(one, wqm::cmul(cosAs, cosAs)));
            const wqm::cplx cos2As = wqm::csub(wqm::cmul(wqm::cscale(cosAs, 2.0), cosAs), one);
            const wqm::cplx sin2As = wqm::cmul(wqm::cscale(sinAs, 2.0), cosAs);
            Zb = wqm::cdiv(one, wqm::cadd(wqm::cmul(wqm::cmul(wqm::cdiv(one, Zb), cos2As), cos2As),
                                          wqm::cmul(wqm::cmul(Zs, sin2As), sin2As)));
        }
        const wqm::cplx Zwc = {Zw, 0.0};
        const wqm::cplx R = wqm::cdiv(wqm::csub(Zb, Zwc), wqm::cadd(Zb, Zwc));
        const double a = -20.0 * wqm::log10(wqm::cabs(R)) * 1.0;
        const double ph = wqm::carg(R) * 1.0;
        for (size_t f = 0; f < nf; ++f) {
            amp[f] = a;
            if (phase) phase[f] = ph;
        }
    }
};

// ------------------------------------------------------------ WaveFront
// waveq3d/wave_front.h:152-258 as parallel arrays, de-major.
struct Front {
    size_t nde = 0, naz = 0, nf = 0;
    std::vector<double> pr, pt, pp, gr, gt, gp, nr, nt, np, hr, ht, hp;
    std::vector<double> c, cr, ct, cp, dist, path, atten, phase;
    std::vector<int32_t> srf, btm, cst, upr, lwr;
    std::vector<uint8_t> edge;
    std::vector<double> tr, tt, tp, ts;  // position and sin(theta) at last update()

    void init(size_t d, size_t a, size_t f) {
        nde = d;
        naz = a;
        nf = f;
        const size_t n = d * a;
        for (auto* v : {&pr, &pt, &pp, &gr, &gt, &gp, &nr, &nt, &np, &hr, &ht, &hp, &c, &cr,
                        &ct, &cp, &dist, &path, &tr, &tt, &tp, &ts})
            v->assign(n, 0.0);
        atten.assign(n * f, 0.0);
        phase.assign(n * f, 0.0);
        for (auto* v : {&srf, &btm, &cst, &upr, &lwr}) v->assign(n, 0);
        edge.assign(n, 0);
    }
    size_t id(size_t de, size_t az) const { return de * naz + az; }
};

struct Queue;

// ------------------------------------------------------------- Spreading
// waveq3d/spreading_hybrid_gaussian.cc:20-571, .h:148-153
struct Spreading {
    Queue* w = nullptr;
    std::vector<double> norm_de, norm_az, init_area, spread, I_de, I_az, bw;
    std::vector<uint8_t> dup;
    bool classic = false;     // spreading_ray (CLASSIC_RAY) instead of the hybrid Gaussian model
    double init_c = 0.0;      // spreading_ray::_init_sound_speed
    static constexpr double OVERLAP = 2.0;
    static constexpr double THRESHOLD = 1.002305238;
    // test hook: when on, every walk step appends (walk 0-3, sum before,
    // sum after, 0 going on / 1 stopped by THRESHOLD / 2 stopped at a
    // virtual ray) for the lowest frequency
    bool trace_walks = false;
    std::vector<double> walk_trace;
    void trace(int walk, double before, double after, int how) {
        if (!trace_walks) return;
        walk_trace.insert(walk_trace.end(), {(double)walk, before, after, (double)how});
    }
    void init(Queue* q);
    void gaussian(double d, double wdt, double A, double* out);
    void intensity(double trho, double ttheta, double tphi, size_t de, size_t az,
                   const double offset[3], const double distance[3], double* result);
    void intensity_classic(double trho, double ttheta, double tphi, size_t de, size_t az,
                           const double offset[3], double* result);
    void intensity_de(size_t de, size_t az, const double offset[3], const double distance[3]);
    void intensity_az(size_t de, size_t az, const double offset[3], const double distance[3]);
    double width_de(size_t de, size_t az, const double offset[3]) const;
    double width_az(size_t de, size_t az, const double offset[3]) const;
};

struct Queue {
    Profile profile;
    Boundary surface, bottom;
    std::vector<wq3d_volume_desc> volume;
    Seq freq, de_seq, az_seq;
    double src_r = 0, src_t = 0, src_p = 0;  // clamped source (_source_pos)
    double time_step = 0, time = 0, R = 6371000.0;
    size_t max_de = 0, max_az = 0;
    bool az_boundary = false, de_branch = false;
    bool want_verbs = false, want_refl = false, has_spreading = false;
    wq3d_thresholds thr{};
    size_t nt1 = 0, nt2 = 0;
    std::vector<double> tgt_r, tgt_t, tgt_p, tgt_s;
    Front slots[4];
    Front* past = &slots[0];
    Front* prev = &slots[1];
    Front* curr = &slots[2];
    Front* next = &slots[3];
    Spreading spreading;
    int64_t step_count = 0;
    // per-ray emission counter within one step (record ordering key)
    int32_t emit_seq = 0;

    std::vector<wq3d_eigenray> rays;
    std::vector<double> ray_int, ray_phase;
    std::vector<wq3d_eigenverb> verbs;
    std::vector<double> verb_power;
    std::vector<wq3d_reflection> refl;

    size_t nde() const { return de_seq.size(); }
    size_t naz() const { return az_seq.size(); }
    size_t nf() const { return freq.size(); }

    // ---------------------------------------------------------- update
    // wave_front::update + compute_profile (wave_front.cc:125-178,241-250)
    void update_ray(Front& w, size_t i, bool atten = true) {
        double c, gr, gt, gp;
        profile.sound_speed(w.pr[i], w.pt[i], w.pp[i], R, &c, &gr, &gt, &gp);
        w.c[i] = c;
        w.cr[i] = gr;
        w.ct[i] = gt;
        w.cp[i] = gp;
        if (atten) profile.attenuation(w.pr[i] - R, w.dist[i], freq, &w.atten[i * w.nf]);
        for (size_t f = 0; f < w.nf; ++f) w.phase[i * w.nf + f] = 0.0;
        const double dcc_r = w.cr[i] / w.c[i];
        const double dcc_t = w.ct[i] / w.c[i];
        const double dcc_p = w.cp[i] / w.c[i];
        const double st = wqm::sin(w.pt[i]);
        const double cot = wqm::cos(w.pt[i]) / st;
        double c2r = w.c[i] * w.c[i];
        w.gr[i] = c2r * w.nr[i];
        c2r = c2r / w.pr[i];
        w.gt[i] = c2r * w.nt[i];
        w.gp[i] = c2r / st * w.np[i];
        w.hr[i] = c2r * (w.nt[i] * w.nt[i] + w.np[i] * w.np[i]) - dcc_r;
        w.ht[i] = -c2r * (w.nr[i] * w.nt[i] - w.np[i] * w.np[i] * cot) - dcc_t / w.pr[i];
        w.hp[i] = -c2r * (w.np[i] * (w.nr[i] + w.nt[i] * cot)) - dcc_p / (w.pr[i] * st);
        w.tr[i] = w.pr[i];
        w.tt[i] = w.pt[i];
        w.tp[i] = w.pp[i];
        w.ts[i] = st;
    }
    void update(Front& w) {
        for (size_t i = 0; i < w.pr.size(); ++i) update_ray(w, i);
    }

    // ode_integ.cc:20-126 (RK3 stages), 131-186 (AB3), one ray at a time
    static void rk1(double dt, const Front& y0, size_t i0, Front& y1, size_t i1) {
        y1.pr[i1] = y0.pr[i0] + 0.5 * dt * y0.gr[i0];
        y1.pt[i1] = y0.pt[i0] + 0.5 * dt * y0.gt[i0];
        y1.pp[i1] = y0.pp[i0] + 0.5 * dt * y0.gp[i0];
        y1.nr[i1] = y0.nr[i0] + 0.5 * dt * y0.hr[i0];
        y1.nt[i1] = y0.nt[i0] + 0.5 * dt * y0.ht[i0];
        y1.np[i1] = y0.np[i0] + 0.5 * dt * y0.hp[i0];
    }
    static void rk2(double dt, const Front& y0, size_t i0, const Front& y1, size_t i1, Front& y2,
                    size_t i2) {
        y2.pr[i2] = y0.pr[i0] + dt * (2.0 * y1.gr[i1] - y0.gr[i0]);
        y2.pt[i2] = y0.pt[i0] + dt * (2.0 * y1.gt[i1] - y0.gt[i0]);
        y2.pp[i2] = y0.pp[i0] + dt * (2.0 * y1.gp[i1] - y0.gp[i0]);
        y2.nr[i2] = y0.nr[i0] + dt * (2.0 * y1.hr[i1] - y0.hr[i0]);
        y2.nt[i2] = y0.nt[i0] + dt * (2.0 * y1.ht[i1] - y0.ht[i0]);
        y2.np[i2] = y0.np[i0] + dt * (2.0 * y1.hp[i1] - y0.hp[i0]);
    }
    static void rk3(double dt, const Front& y0, size_t i0, const Front& y1, size_t i1,
                    const Front& y2, size_t i2, Front& y3, size_t i3) {
        const double pr = y0.pr[i0] + dt / 6.0 * (y0.gr[i0] + 4.0 * y1.gr[i1] + y2.gr[i2]);
        const double pt = y0.pt[i0] + dt / 6.0 * (y0.gt[i0] + 4.0 * y1.gt[i1] + y2.gt[i2]);
        const double pp = y0.pp[i0] + dt / 6.0 * (y0.gp[i0] + 4.0 * y1.gp[i1] + y2.gp[i2]);
        const double nr = y0.nr[i0] + dt / 6.0 * (y0.hr[i0] + 4.0 * y1.hr[i1] + y2.hr[i2]);
        const double nt = y0.nt[i0] + dt / 6.0 * (y0.ht[i0] + 4.0 * y1.ht[i1] + y2.ht[i2]);
        const double np = y0.np[i0] + dt / 6.0 * (y0.hp[i0] + 4.0 * y1.hp[i1] + y2.hp[i2]);
        y3.pr[i3] = pr;
        y3.pt[i3] = pt;
        y3.pp[i3] = pp;
        y3.nr[i3] = nr;
        y3.nt[i3] = nt;
        y3.np[i3] = np;
    }
    static void ab3(double dt, const Front& y0, size_t i0, const Front& y1, size_t i1,
                    const Front& y2, size_t i2, Front& y3, size_t i3) {
        const double A2 = 23.0 / 12.0, A1 = 16.0 / 12.0, A0 = 5.0 / 12.0;
        const double dr = dt * (A2 * y2.gr[i2] - A1 * y1.gr[i1] + A0 * y0.gr[i0]);
        const double dth = dt * (A2 * y2.gt[i2] - A1 * y1.gt[i1] + A0 * y0.gt[i0]);
        const double dph = dt * (A2 * y2.gp[i2] - A1 * y1.gp[i1] + A0 * y0.gp[i0]);
        const double a = y2.pr[i2] * dth;
        const double b = y2.pr[i2] * (wqm::sin(y2.pt[i2]) * dph);
        y3.dist[i3] = std::sqrt(dr * dr + a * a + b * b);
        y3.pr[i3] = y2.pr[i2] + dr;
        y3.pt[i3] = y2.pt[i2] + dth;
        y3.pp[i3] = y2.pp[i2] + dph;
        y3.nr[i3] = y2.nr[i2] + dt * (A2 * y2.hr[i2] - A1 * y1.hr[i1] + A0 * y0.hr[i0]);
        y3.nt[i3] = y2.nt[i2] + dt * (A2 * y2.ht[i2] - A1 * y1.ht[i1] + A0 * y0.ht[i0]);
        y3.np[i3] = y2.np[i2] + dt * (A2 * y2.hp[i2] - A1 * y1.hp[i1] + A0 * y0.hp[i0]);
    }

    // --------------------------------------------------- construction
    // wave_queue::wave_queue (wave_queue.cc:30-98)
    void create(const wq3d_env_desc& env, const wq3d_fan_desc& fd) {
        profile.load(env.profile);
        surface.load(env.surface);
        bottom.load(env.bottom);
        if (env.num_volume > 0) volume.assign(env.volume, env.volume + env.num_volume);
        freq.load(fd.freq);
        de_seq.load(fd.de);
        az_seq.load(fd.az);
        time_step = fd.time_step;
        R = fd.earth_radius;
        want_verbs = fd.want_eigenverbs != 0;
        want_refl = fd.want_reflections != 0;
        thr = fd.thresholds;
        max_de = nde() - 1;
        max_az = naz() - 1;
        az_boundary = false;
        if (naz() > 1) {
            const double az_first = std::fabs(az_seq(0));
            const double az_last = std::fabs(az_seq(naz() - 1));
            az_boundary = std::fmod(az_first + 360.0, 360.0) == std::fmod(az_last + 360.0, 360.0);
        }
        nt1 = (size_t)std::max<int64_t>(0, fd.num_t1);
        nt2 = (size_t)std::max<int64_t>(0, fd.num_t2);
        if (nt1 * nt2 > 0) {
            const size_t nt = nt1 * nt2;
            tgt_r.assign(fd.target_rho, fd.target_rho + nt);
            tgt_t.assign(fd.target_theta, fd.target_theta + nt);
            tgt_p.assign(fd.target_phi, fd.target_phi + nt);
            tgt_s.resize(nt);
            for (size_t n = 0; n < nt; ++n) tgt_s[n] = wqm::sin(tgt_t[n]);
        }
        // source clamp: affects _source_pos only; init_wave gets the
        // constructor's argument (wave_queue.cc:58-68, 83)
        src_r = fd.src_rho;
        src_t = fd.src_theta;
        src_p = fd.src_phi;
        const double offset = 0.1;
        double bottom_rho = 0.0;
        bottom.height(src_r, src_t, src_p, &bottom_rho, nullptr);
        bottom_rho += offset;
        if (src_r - R > -offset) {
            src_r = -offset + R;
        } else if (src_r < bottom_rho) {
            src_r = bottom_rho;
        }
        for (auto& s : slots) s.init(nde(), naz(), nf());
        init_wave(fd.src_rho, fd.src_theta, fd.src_phi);
        update(*curr);
        init_wavefronts();
        has_spreading = nde() >= 3 && naz() >= 3;
        if (nt1 * nt2 > 0 && !has_spreading && naz() >= 2 && az_boundary && nde() >= 3)
            throw Error{"targets need a fan of at least 3 x 3 rays"};
        if (has_spreading) {
            if (fd.spreading != WQ3D_SPREADING_HYBRID_GAUSSIAN && fd.spreading != WQ3D_SPREADING_CLASSIC_RAY)
                throw Error{"unknown spreading model"};
            spreading.init(this);
            spreading.classic = fd.spreading == WQ3D_SPREADING_CLASSIC_RAY;
            // spreading_ray.cc:41: the curr front's sound speed of ray (0, 0)
            // once the queue is built
            spreading.init_c = curr->c[curr->id(0, 0)];
        }
    }

    // wave_front::init_wave (wave_front.cc:84-121)
    void init_wave(double r, double t, double p) {
        Front& w = *curr;
        for (size_t d = 0; d < nde(); ++d) {
            const double cos_de = wqm::cos(to_radians(de_seq(d)));
            const double sin_de = wqm::sin(to_radians(de_seq(d)));
            for (size_t a = 0; a < naz(); ++a) {
                const double cos_az = wqm::cos(to_radians(az_seq(a)));
                const double sin_az = wqm::sin(to_radians(az_seq(a)));
                const size_t i = w.id(d, a);
                w.nr[i] = sin_de;
                w.nt[i] = -cos_de * cos_az;
                w.np[i] = cos_de * sin_az;
                w.pr[i] = r;
                w.pt[i] = t;
                w.pp[i] = p;
            }
        }
        double c0;
        profile.sound_speed(r, t, p, R, &c0, nullptr, nullptr, nullptr);
        for (size_t i = 0; i < w.nr.size(); ++i) {
            w.nr[i] = w.nr[i] / c0;
            w.nt[i] = w.nt[i] / c0;
            w.np[i] = w.np[i] / c0;
        }
    }

    // wave_queue::init_wavefronts (wave_queue.cc:114-150)
    void init_wavefronts() {
        const size_t n = curr->pr.size();
        const double dt = time_step;
        for (size_t i = 0; i < n; ++i) rk1(-dt, *curr, i, *next, i);
        update(*next);
        for (size_t i = 0; i < n; ++i) rk2(-dt, *curr, i, *next, i, *past, i);
        update(*past);
        for (size_t i = 0; i < n; ++i) rk3(-dt, *curr, i, *next, i, *past, i, *prev, i);
        update(*prev);
        for (size_t i = 0; i < n; ++i) rk1(-dt, *prev, i, *next, i);
        update(*next);
        for (size_t i = 0; i < n; ++i) rk2(-dt, *prev, i, *next, i, *past, i);
        update(*past);
        for (size_t i = 0; i < n; ++i) rk3(-dt, *prev, i, *next, i, *past, i, *past, i);
        update(*past);
        for (size_t i = 0; i < n; ++i) ab3(dt, *past, i, *prev, i, *curr, i, *next, i);
        update(*next);
        for (size_t i = 0; i < n; ++i) next->path[i] = next->dist[i] + curr->path[i];
    }

    // ------------------------------------------------------------ step
    // wave_queue::step (wave_queue.cc:155-192)
    void step() {
        detect_reflections();
        Front* save = past;
        past = prev;
        prev = curr;
        curr = next;
        next = save;
        time += time_step;
        ++step_count;
        const size_t n = curr->pr.size();
        for (size_t i = 0; i < n; ++i) ab3(time_step, *past, i, *prev, i, *curr, i, *next, i);
        update(*next);
        for (size_t i = 0; i < n; ++i) {
            next->path[i] = next->dist[i] + curr->path[i];
            for (size_t f = 0; f < nf(); ++f) {
                next->atten[i * nf() + f] += curr->atten[i * nf() + f];
                next->phase[i * nf() + f] += curr->phase[i * nf() + f];
            }
            next->srf[i] = curr->srf[i];
            next->btm[i] = curr->btm[i];
            next->upr[i] = curr->upr[i];
            next->lwr[i] = curr->lwr[i];
            next->cst[i] = curr->cst[i];
        }
        detect_eigenrays();
    }

    // wave_queue::detect_reflections (wave_queue.cc:197-216)
    void detect_reflections() {
        for (size_t de = 0; de < nde(); ++de) {
            for (size_t az = 0; az < naz(); ++az) {
                emit_seq = 0;
                detect_volume_scattering(de, az);
                if (!detect_reflections_surface(de, az)) {
                    if (!detect_reflections_bottom(de, az)) {
                        detect_vertices(de, az);
                        detect_caustics(de, az);
                    }
                }
            }
        }
        find_edges(*next);
    }

    bool above_bounce(const Front& w, size_t i) const {
        return w.btm[i] > thr.max_bottom || w.srf[i] > thr.max_surface ||
               w.cst[i] > thr.max_caustic || w.upr[i] > thr.max_upper ||
               w.lwr[i] > thr.max_lower;
    }

    // wave_queue.cc:222-234
    bool detect_reflections_surface(size_t de, size_t az) {
        const size_t i = next->id(de, az);
        if (next->pr[i] - R > 0.0) {
            if (surface_reflection(de, az)) {
                next->srf[i] += 1;
                curr->srf[i] = prev->srf[i] = past->srf[i] = next->srf[i];
                detect_volume_scattering(de, az);
                detect_reflections_bottom(de, az);
                return true;
            }
        }
        return false;
    }

    // wave_queue.cc:240-256
    bool detect_reflections_bottom(size_t de, size_t az) {
        const size_t i = next->id(de, az);
        double height;
        bottom.height(next->pr[i], next->pt[i], next->pp[i], &height, nullptr);
        const double depth = height - next->pr[i];
        if (depth > 0.0) {
            if (bottom_reflection(de, az, depth)) {
                next->btm[i] += 1;
                curr->btm[i] = prev->btm[i] = past->btm[i] = next->btm[i];
                detect_volume_scattering(de, az);
                detect_reflections_surface(de, az);
                return true;
            }
        }
        return false;
    }

    // wave_queue.cc:261-269
    void detect_vertices(size_t de, size_t az) {
        const size_t i = next->id(de, az);
        const double L = curr->nr[i];
        const double Rr = next->nr[i];
        if (L * Rr < 0.0 && Rr < 0.0) {
            next->upr[i]++;
        } else if (L * Rr < 0.0 && 0.0 < Rr) {
            next->lwr[i]++;
        }
    }

    // wave_queue.cc:274-292
    void detect_caustics(size_t de, size_t az) {
        if (de < max_de) {
            const size_t i = next->id(de, az);
            const size_t j = next->id(de + 1, az);
            const double A = curr->pr[j];
            const double B = curr->pr[i];
            const double C = next->pr[j];
            const double D = next->pr[i];
            const bool fold = next->srf[j] == next->srf[i] && next->btm[j] == next->btm[i];
            if ((C - D) * (A - B) < 0 && fold) {
                next->cst[j]++;
                for (size_t f = 0; f < nf(); ++f) next->phase[j * nf() + f] -= M_PI_2;
            }
        }
    }

    // wave_front::find_edges (wave_front.cc:184-216)
    static void find_edges(Front& w) {
        std::fill(w.edge.begin(), w.edge.end(), 0);
        const size_t md = w.nde - 1;
        for (size_t az = 0; az < w.naz; ++az) {
            w.edge[w.id(0, az)] = 1;
            w.edge[w.id(md, az)] = 1;
        }
        for (size_t az = 0; az < w.naz; ++az) {
            for (size_t de = 1; de < md; ++de) {
                const double r = w.pr[w.id(de, az)];
                const double rp = w.pr[w.id(de + 1, az)];
                const double rm = w.pr[w.id(de - 1, az)];
                if ((r < rp && r < rm) || (r > rp && r > rm)) {
                    w.edge[w.id(de, az)] = 1;
                    const double x = w.nr[w.id(de, az)];
                    if (std::fabs(x - w.nr[w.id(de - 1, az)]) > std::fabs(x - w.nr[w.id(de + 1, az)])) {
                        w.edge[w.id(de - 1, az)] = 1;
                    } else {
                        w.edge[w.id(de + 1, az)] = 1;
                    }
                }
            }
        }
    }

    // wave_queue::collision_location (wave_queue.cc:982-1057)
    void collision_location(size_t de, size_t az, double tw, double pos[3], double nd[3],
                            double* speed) const {
        const size_t i = curr->id(de, az);
        const double time1 = 2.0 * time_step;
        const double time2 = time_step * time_step;
        const double dtime2 = tw * tw;
        double drho = (next->c[i] - prev->c[i]) / time1;
        double d2rho = (next->c[i] + prev->c[i] - 2.0 * curr->c[i]) / time2;
        *speed = curr->c[i] + drho * tw + 0.5 * d2rho * dtime2;
        const std::vector<double>* P[3][3] = {{&prev->pr, &curr->pr, &next->pr},
                                              {&prev->pt, &curr->pt, &next->pt},
                                              {&prev->pp, &curr->pp, &next->pp}};
        const std::vector<double>* N[3][3] = {{&prev->nr, &curr->nr, &next->nr},
                                              {&prev->nt, &curr->nt, &next->nt},
                                              {&prev->np, &curr->np, &next->np}};
        for (int k = 0; k < 3; ++k) {
            const double d1 = ((*P[k][2])[i] - (*P[k][0])[i]) / time1;
            const double d2 = ((*P[k][2])[i] + (*P[k][0])[i] - 2.0 * (*P[k][1])[i]) / time2;
            pos[k] = (*P[k][1])[i] + d1 * tw + 0.5 * d2 * dtime2;
        }
        for (int k = 0; k < 3; ++k) {
            const double d1 = ((*N[k][2])[i] - (*N[k][0])[i]) / time1;
            const double d2 = ((*N[k][2])[i] + (*N[k][0])[i] - 2.0 * (*N[k][1])[i]) / time2;
            nd[k] = (*N[k][1])[i] + d1 * tw + 0.5 * d2 * dtime2;
        }
    }

    void notify_reflection(size_t de, size_t az, double tw, double grazing, double c,
                           const double pos[3], const double nd[3], int type) {
        wq3d_reflection r{};
        r.step = step_count;
        r.de = (int32_t)de;
        r.az = (int32_t)az;
        r.seq = emit_seq++;
        r.type = type;
        r.time = time + tw;
        r.dt = tw;
        r.grazing = grazing;
        r.speed = c;
        r.pos_rho = pos[0];
        r.pos_theta = pos[1];
        r.pos_phi = pos[2];
        r.ndir_rho = nd[0];
        r.ndir_theta = nd[1];
        r.ndir_phi = nd[2];
        refl.push_back(r);
    }

    // reflection_model::surface_reflection (reflection_model.cc:187-243)
    bool surface_reflection(size_t de, size_t az) {
        const size_t i = curr->id(de, az);
        double c = curr->c[i];
        const double d = c * c * curr->nr[i];
        const double time_water = (d == 0.0) ? 0.0 : -(curr->pr[i] - R) / d;
        double pos[3], nd[3];
        collision_location(de, az, time_water, pos, nd, &c);
        const double grazing =
            wqm::atan2(curr->nr[i], std::sqrt(curr->nt[i] * curr->nt[i] + curr->np[i] * curr->np[i]));
        if (grazing <= 0.0) return false;
        if (want_refl) notify_reflection(de, az, time_water, grazing, c, pos, nd, 1);
        if (want_verbs) build_eigenverb(de, az, time_water, grazing, c, pos, nd, 1);
        std::vector<double> amp(nf());
        surface.reflect_loss(grazing, nf(), amp.data(), nullptr);
        for (size_t f = 0; f < nf(); ++f) {
            next->atten[i * nf() + f] += amp[f];
            next->phase[i * nf() + f] -= M_PI;
        }
        nd[0] = -nd[0];
        reflection_reinit(de, az, time_water, pos, nd);
        return true;
    }

    // reflection_model::bottom_reflection (reflection_model.cc:33-182)
    bool bottom_reflection(size_t de, size_t az, double depth) {
        const double MIN_REFLECT = 6.0;
        const double TOO_SHALLOW = 300.0 * time_step;
        const size_t i = curr->id(de, az);
        double pos[3] = {curr->pr[i], curr->pt[i], curr->pp[i]};
        double nd[3] = {curr->nr[i], curr->nt[i], curr->np[i]};
        double c = curr->c[i];
        double c2 = c * c;
        double bottom_rho, nrm[3];
        bottom.height(pos[0], pos[1], pos[2], &bottom_rho, nrm);
        double height_water = pos[0] - bottom_rho;
        const bool shallow = (R - bottom_rho) < TOO_SHALLOW;
        if (shallow) {
            const double N = std::sqrt(nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            nrm[0] = 0.0;
            nrm[1] = nrm[1] / N;
            nrm[2] = nrm[2] / N;
        }
        nd[0] = c2 * nd[0];
        nd[1] = c2 * nd[1];
        nd[2] = c2 * nd[2];
        double dot_full = nrm[0] * nd[0] + nrm[1] * nd[1] + nrm[2] * nd[2];
        double max_dot = -smax(MIN_REFLECT, (height_water + depth) * nrm[0]);
        if (dot_full >= max_dot) dot_full = max_dot;
        const double dot_water = -height_water * nrm[0];
        const double time_water = smax(0.0, dot_water / dot_full);
        collision_location(de, az, time_water, pos, nd, &c);
        bottom.height(pos[0], pos[1], pos[2], &bottom_rho, nrm);
        if (shallow) {
            const double N = std::sqrt(nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            nrm[0] = 0.0;
            nrm[1] = nrm[1] / N;
            nrm[2] = nrm[2] / N;
        }
        c2 = c * c;
        height_water = pos[0] - bottom_rho;
        nd[0] = c2 * nd[0];
        nd[1] = c2 * nd[1];
        nd[2] = c2 * nd[2];
        dot_full = nrm[0] * nd[0] + nrm[1] * nd[1] + nrm[2] * nd[2];
        max_dot = -smax(MIN_REFLECT, (height_water + depth) * nrm[0]);
        if (dot_full >= max_dot) dot_full = max_dot;
        double grazing;
        if (dot_full / c >= 1.0) {
            grazing = -M_PI_2;
        } else if (dot_full / c <= -1.0) {
            grazing = M_PI_2;
        } else {
            grazing = wqm::asin(-dot_full / c);
        }
        if (want_refl) notify_reflection(de, az, time_water, grazing, c, pos, nd, 0);
        if (want_verbs) build_eigenverb(de, az, time_water, grazing, c, pos, nd, 0);
        std::vector<double> amp(nf()), ph(nf());
        bottom.reflect_loss(grazing, nf(), amp.data(), ph.data());
        for (size_t f = 0; f < nf(); ++f) {
            next->atten[i * nf() + f] += amp[f];
            next->phase[i * nf() + f] += ph[f];
        }
        dot_full *= 2.0;
        nd[0] = nd[0] - dot_full * nrm[0];
        nd[1] = nd[1] - dot_full * nrm[1];
        nd[2] = nd[2] - dot_full * nrm[2];
        const double N = std::sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]) * c;
        nd[0] = nd[0] / N;
        nd[1] = nd[1] / N;
        nd[2] = nd[2] / N;
        reflection_reinit(de, az, time_water, pos, nd);
        return true;
    }

    // reflection_model::reflection_reinit / reflection_copy
    // (reflection_model.cc:248-360), five 1x1 temporary fronts
    void reflection_reinit(size_t de, size_t az, double time_water, const double pos[3],
                           const double nd[3]) {
        Front tpast, tprev, tcurr, tnext, temp;
        for (Front* f : {&tpast, &tprev, &tcurr, &tnext, &temp}) f->init(1, 1, nf());
        temp.pr[0] = pos[0];
        temp.pt[0] = pos[1];
        temp.pp[0] = pos[2];
        temp.nr[0] = nd[0];
        temp.nt[0] = nd[1];
        temp.np[0] = nd[2];
        update(temp);
        rk1(-time_water, temp, 0, tnext, 0);
        update(tnext);
        rk2(-time_water, temp, 0, tnext, 0, tpast, 0);
        update(tpast);
        rk3(-time_water, temp, 0, tnext, 0, tpast, 0, tcurr, 0);
        update(tcurr);
        reflection_copy(*curr, de, az, tcurr);
        const double dt = time_step;
        rk1(-dt, tcurr, 0, tnext, 0);
        update(tnext);
        rk2(-dt, tcurr, 0, tnext, 0, tpast, 0);
        update(tpast);
        rk3(-dt, tcurr, 0, tnext, 0, tpast, 0, tprev, 0);
        update(tprev);
        reflection_copy(*prev, de, az, tprev);
        rk1(-dt, tprev, 0, tnext, 0);
        update(tnext);
        rk2(-dt, tprev, 0, tnext, 0, temp, 0);  // writes temp, then updates past
        update(tpast);
        rk3(-dt, tprev, 0, tnext, 0, temp, 0, tpast, 0);
        update(tpast);
        reflection_copy(*past, de, az, tpast);
        ab3(dt, tpast, 0, tprev, 0, tcurr, 0, tnext, 0);
        update(tnext);
        reflection_copy(*next, de, az, tnext);
    }

    static void reflection_copy(Front& e, size_t de, size_t az, const Front& r) {
        const size_t i = e.id(de, az);
        e.pr[i] = r.pr[0];
        e.pt[i] = r.pt[0];
        e.pp[i] = r.pp[0];
        e.gr[i] = r.gr[0];
        e.gt[i] = r.gt[0];
        e.gp[i] = r.gp[0];
        e.nr[i] = r.nr[0];
        e.nt[i] = r.nt[0];
        e.np[i] = r.np[0];
        e.hr[i] = r.hr[0];
        e.ht[i] = r.ht[0];
        e.hp[i] = r.hp[0];
        e.cr[i] = r.cr[0];
        e.ct[i] = r.ct[0];
        e.cp[i] = r.cp[0];
        e.c[i] = r.c[0];
        e.dist[i] = r.dist[0];
        e.path[i] += r.path[0];
    }

    // wave_queue::detect_volume_scattering (wave_queue.cc:297-345)
    void detect_volume_scattering(size_t de, size_t az) {
        if (!want_verbs) return;
        const size_t i = curr->id(de, az);
        if (above_bounce(*curr, i)) return;
        for (size_t l = 0; l < volume.size(); ++l) {
            const double height = volume[l].rho;
            const double d1 = height - next->pr[i];
            const double d2 = height - curr->pr[i];
            if (d1 * d2 >= 0) continue;
            double c = curr->c[i];
            const double d = c * c * curr->nr[i];
            const double time_water = (d == 0.0) ? 0.0 : -(curr->pr[i] - R) / d;
            double pos[3], nd[3];
            collision_location(de, az, time_water, pos, nd, &c);
            const double grazing = wqm::atan2(
                curr->nr[i], std::sqrt(curr->nt[i] * curr->nt[i] + curr->np[i] * curr->np[i]));
            const int type = (d1 > 0) ? 3 + (int)l * 2 : 2 + (int)l * 2;
            build_eigenverb(de, az, time_water, grazing, c, pos, nd, type);
        }
    }

    // wave_queue::build_eigenverb (wave_queue.cc:1063-1164)
    void build_eigenverb(size_t de, size_t az, double dt, double grazing, double speed,
                         const double pos[3], const double nd[3], int type) {
        grazing = std::fabs(grazing);
        const size_t i = curr->id(de, az);
        if (!want_verbs || above_bounce(*curr, i) || time <= 0.0 || grazing < 1e-6 ||
            (az_boundary && az == max_az) || std::fabs(de_seq(de)) > 89.9)
            return;
        const double de_angle = to_radians(de_seq(de));
        const double de_plus = de_angle + 0.5 * to_radians(de_seq.increment(de));
        const double de_minus = de_angle - 0.5 * to_radians(de_seq.increment(de - 1));
        const double az_angle = to_radians(az_seq(az));
        const double az_plus = az_angle + 0.5 * to_radians(az_seq.increment(az));
        const size_t az_index = (az == 0 && az_boundary) ? max_az : az;
        const double az_minus = az_angle - 0.5 * to_radians(az_seq.increment(az_index - 1));
        const double area = (wqm::sin(de_plus) - wqm::sin(de_minus)) * (az_plus - az_minus);
        const double de_delta = de_plus - de_minus;
        const double az_delta = area / de_delta;
        const double path_length = curr->path[i] + curr->c[i] * dt;
        if (path_length <= 0) return;
        const double sin_grazing = wqm::sin(grazing);
        wq3d_eigenverb v{};
        v.length = 0.5 * path_length * de_delta / sin_grazing;
        v.width = 0.5 * path_length * az_delta;
        std::vector<double> power(nf());
        bool above = false;
        for (size_t f = 0; f < nf(); ++f) {
            power[f] = wqm::pow10(-0.1 * curr->atten[i * nf() + f]) * area / sin_grazing;
            if (power[f] >= thr.eigenverb) above = true;
        }
        if (!above) return;
        const double a = to_degrees(wqm::atan2(nd[2], -nd[1]));
        v.direction = to_radians(a);
        v.step = step_count;
        v.de = (int32_t)de;
        v.az = (int32_t)az;
        v.seq = emit_seq++;
        v.type = type;
        v.travel_time = time + dt;
        v.grazing = grazing;
        v.sound_speed = speed;
        v.pos_rho = pos[0];
        v.pos_theta = pos[1];
        v.pos_phi = pos[2];
        v.source_de = to_radians(de_seq(de));
        v.source_az = to_radians(az_seq(az));
        v.surface = curr->srf[i];
        v.bottom = curr->btm[i];
        v.caustic = curr->cst[i];
        v.upper = curr->upr[i];
        v.lower = curr->lwr[i];
        verbs.push_back(v);
        verb_power.insert(verb_power.end(), power.begin(), power.end());
    }

    // ------------------------------------------------------- eigenrays
    // wave_front::compute_target_distance (wave_front.cc:222-236)
    double distance2(const Front& w, size_t t, size_t i) const {
        const double rho = w.tr[i];
        const double Rt = tgt_r[t];
        const double h = 0.5 * (w.tt[i] - tgt_t[t]);
        const double g = 0.5 * (w.tp[i] - tgt_p[t]);
        return std::fabs(rho * rho + Rt * Rt -
                         2.0 * Rt * (rho * (1.0 - 2.0 * (h * h + tgt_s[t] * (w.ts[i] * (g * g))))));
    }

    // wave_queue::detect_eigenrays (wave_queue.cc:351-404)
    void detect_eigenrays() {
        if (nt1 * nt2 == 0) return;
        double D[3][3][3];
        const size_t az_start = az_boundary ? 0 : 1;
        const double src_lat = to_latitude(src_t), src_lng = to_degrees(src_p);
        for (size_t t1 = 0; t1 < nt1; ++t1) {
            for (size_t t2 = 0; t2 < nt2; ++t2) {
                const size_t t = t1 * nt2 + t2;
                de_branch = std::fabs(src_lat - to_latitude(tgt_t[t])) < 1e-4 &&
                            std::fabs(src_lng - to_degrees(tgt_p[t])) < 1e-4;
                for (size_t de = 1; de < max_de; ++de) {
                    for (size_t az = az_start; az < max_az; ++az) {
                        const size_t i = curr->id(de, az);
                        if (curr->edge[i]) continue;
                        const double center = distance2(*curr, t, i);
                        D[1][1][1] = center;
                        D[2][1][1] = distance2(*next, t, i);
                        if (D[2][1][1] <= center) continue;
                        D[0][1][1] = distance2(*prev, t, i);
                        if (D[0][1][1] < center) continue;
                        if (is_closest_ray(t, de, az, center, D)) build_eigenray(t1, t2, de, az, D);
                    }
                }
            }
        }
    }

    // wave_queue::is_closest_ray (wave_queue.cc:411-488)
    bool is_closest_ray(size_t t, size_t de, size_t az, double center, double D[3][3][3]) {
        for (size_t nde = 0; nde < 3; ++nde) {
            for (size_t naz = 0; naz < 3; ++naz) {
                if (nde == 1 && naz == 1) continue;
                const size_t d = de + nde - 1;
                size_t a = az + naz - 1;
                if (az_boundary) {
                    if (az + naz == 0) {
                        a = naz_() - 2;
                    } else if (a >= max_az) {
                        a = 0;
                    }
                }
                const size_t i = curr->id(d, a);
                D[0][nde][naz] = distance2(*prev, t, i);
                D[1][nde][naz] = distance2(*curr, t, i);
                D[2][nde][naz] = distance2(*next, t, i);
                if (a == 0 && !az_boundary) continue;
                if (a == max_az) continue;
                if (de_branch) {
                    if (curr->edge[i]) continue;
                } else if (nde != 1) {
                    if (curr->edge[i]) continue;
                }
                if (nde == 2 || naz == 2) {
                    if (de_branch) {
                        if (az == 0) {
                            if (D[1][nde][naz] < center) return false;
                        } else {
                            return false;
                        }
                    } else if (D[1][nde][naz] <= center) {
                        return false;
                    }
                } else if (D[1][nde][naz] < center) {
                    return false;
                }
                if (D[2][nde][naz] <= center) return false;
                if (D[0][nde][naz] < center) return false;
            }
        }
        return true;
    }
    size_t naz_() const { return naz(); }

    // wave_queue::make_taylor_coeff (wave_queue.cc:929-977)
    static void make_taylor_coeff(const double v[3][3][3], const double delta[3], double& center,
                                  double g[3], double H[3][3]) {
        const double d0 = 2.0 * delta[0];
        const double d1 = 2.0 * delta[1];
        const double d2 = 2.0 * delta[2];
        center = v[1][1][1];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) H[a][b] = 0.0;
        H[0][0] = (v[2][1][1] + v[0][1][1] - 2.0 * center) / (delta[0] * delta[0]);
        H[1][1] = (v[1][2][1] + v[1][0][1] - 2.0 * center) / (delta[1] * delta[1]);
        H[2][2] = (v[1][1][2] + v[1][1][0] - 2.0 * center) / (delta[2] * delta[2]);
        g[0] = (v[2][0][1] - v[0][0][1]) / d0;
        g[2] = (v[2][2][1] - v[0][2][1]) / d0;
        H[0][1] = (g[2] - g[0]) / d1;
        H[1][0] = H[0][1];
        g[0] = (v[2][1][0] - v[0][1][0]) / d0;
        g[2] = (v[2][1][2] - v[0][1][2]) / d0;
        H[0][2] = (g[2] - g[0]) / d2;
        H[2][0] = H[0][2];
        g[0] = (v[1][2][0] - v[1][0][0]) / d1;
        g[2] = (v[1][2][2] - v[1][0][2]) / d1;
        H[1][2] = (g[2] - g[0]) / d2;
        H[2][1] = H[1][2];
        g[0] = (v[2][1][1] - v[0][1][1]) / d0;
        g[1] = (v[1][2][1] - v[1][0][1]) / d1;
        g[2] = (v[1][1][2] - v[1][1][0]) / d2;
    }

    static double inner3(const double a[3], const double b[3]) {
        double t = 0.0;
        for (int k = 0; k < 3; ++k) t += a[k] * b[k];
        return t;
    }
    static void prod3(const double M[3][3], const double x[3], double y[3]) {
        for (int r = 0; r < 3; ++r) {
            double t = 0.0;
            for (int k = 0; k < 3; ++k) t += M[r][k] * x[k];
            y[r] = t;
        }
    }
    static double norm3(const double g[3]) {
        double t = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double u = std::fabs(g[k]);
            t += u * u;
        }
        return std::sqrt(t);
    }

    // wave_queue::compute_offsets (wave_queue.cc:708-924)
    void compute_offsets(size_t t, size_t de, size_t az, const double D[3][3][3],
                         const double delta[3], double offset[3], double distance[3]) {
        bool unstable = false;
        const size_t ic = curr->id(de, az);
        const int s0 = curr->srf[ic], b0 = curr->btm[ic], c0 = curr->cst[ic];
        for (size_t nde = 0; nde < 3 && !unstable; ++nde) {
            const size_t d = de + nde - 1;
            for (size_t naz = 0; naz < 3 && !unstable; ++naz) {
                size_t a = az + naz - 1;
                if (az_boundary) {
                    if (az + naz == 0) {
                        a = naz_() - 2;
                    } else if (a >= naz_() - 1) {
                        a = 0;
                    }
                }
                const size_t i = curr->id(d, a);
                if (prev->srf[i] != s0 || curr->srf[i] != s0 || next->srf[i] != s0 ||
                    prev->btm[i] != b0 || curr->btm[i] != b0 || next->btm[i] != b0 ||
                    prev->cst[i] != c0 || curr->cst[i] != c0 || next->cst[i] != c0 ||
                    prev->edge[i] || prev->edge[i] || prev->edge[i]) {
                    unstable = true;
                }
            }
        }
        if (offsets_core(D, delta, unstable, offset, distance) && distance[1] > 0.0) {
            // the reference's else-branch (a positive DE remainder) takes its
            // sign from the direction toward the next DE ray and the target
            ++sign_rule_count;
            double cpa[3], nxt[3], prv[3], dir[3], speed;
            collision_location(de, az, offset[0], cpa, dir, &speed);
            collision_location(de + 1, az, offset[0], nxt, dir, &speed);
            collision_location(de - 1, az, offset[0], prv, dir, &speed);
            double nxt_dir[3], tgt_dir[3];
            nxt_dir[0] = nxt[0] - cpa[0];
            nxt_dir[1] = (nxt[1] - prv[1]) * cpa[0];
            nxt_dir[2] = (nxt[2] - prv[2]) * cpa[0] * wqm::sin(cpa[1]);
            if (std::fabs(nxt_dir[0]) < 0.01) nxt_dir[0] = 0.0;
            tgt_dir[0] = tgt_r[t] - cpa[0];
            tgt_dir[1] = (tgt_t[t] - cpa[1]) * cpa[0];
            tgt_dir[2] = (tgt_p[t] - cpa[2]) * cpa[0] * wqm::sin(cpa[1]);
            const double dot = inner3(nxt_dir, tgt_dir);
            if (dot < 0.0) {
                distance[1] *= -1.0;
                offset[1] *= -1.0;
                ++sign_flip_count;
            }
        }
    }
    // how often compute_offsets' unstable branch took the geometric DE sign,
    // and how often that sign was negative (a test hook)
    int64_t sign_rule_count = 0, sign_flip_count = 0;

    // compute_offsets (wave_queue.cc:708-924) from the neighbourhood's
    // distances alone: make_taylor_coeff, the stability tests (determinant
    // typo kept, wave_queue.cc:757-763), the 3 x 3 or 2 x 2 (time, AZ)
    // solve, the per-axis distances and, when unstable, the DE distance as
    // the remainder (its sign set by the caller).  Returns `unstable`.
    static bool offsets_core(const double D[3][3][3], const double delta[3], bool unstable, double offset[3],
                             double distance[3]) {
        double center, g[3], H[3][3], inv[3][3];
        make_taylor_coeff(D, delta, center, g, H);
        double determinant = smax(1e-6, H[0][0] * (H[1][1] * H[2][2] - H[1][2] * H[2][1]) +
                                            H[0][1] * (H[1][2] * H[2][0] - H[1][0] * H[2][2]) +
                                            H[0][2] * (H[1][0] * H[2][1] - H[2][1] * H[2][0]));
        unstable = unstable || determinant < norm3(g);
        for (int n = 0; n < 3; ++n) {
            const double h = smax(1e-10, H[n][n]);
            offset[n] = -g[n] / h;
            if (std::fabs(offset[n] / delta[n]) > 0.5) {
                unstable = true;
                offset[n] = std::copysign(0.5 * delta[n], offset[n]);
            }
        }
        if (!unstable) {
            inv[0][0] = H[1][1] * H[2][2] - H[1][2] * H[2][1];
            inv[1][0] = H[1][2] * H[2][0] - H[1][0] * H[2][2];
            inv[2][0] = H[1][0] * H[2][1] - H[1][1] * H[2][0];
            inv[0][1] = inv[1][0];
            inv[1][1] = H[0][0] * H[2][2] - H[0][2] * H[2][0];
            inv[2][1] = H[2][0] * H[0][1] - H[0][0] * H[2][1];
            inv[0][2] = inv[2][0];
            inv[1][2] = inv[2][1];
            inv[2][2] = H[0][0] * H[1][1] - H[0][1] * H[1][0];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) inv[a][b] /= determinant;
            const double ng[3] = {-g[0], -g[1], -g[2]};
            prod3(inv, ng, offset);
            unstable = unstable || std::fabs(offset[1]) >= 0.5 * std::fabs(delta[1]);
        } else {
            const double save_offset = offset[1];
            g[1] = 0.0;
            determinant = H[0][0] * H[2][2] - H[0][2] * H[2][0];
            if (determinant >= norm3(g)) {
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) inv[a][b] = 0.0;
                inv[0][0] = H[2][2];
                inv[0][2] = -H[0][2];
                inv[1][1] = 1.0;
                inv[2][0] = -H[2][0];
                inv[2][2] = H[0][0];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) inv[a][b] /= determinant;
                const double ng[3] = {-g[0], -g[1], -g[2]};
                prod3(inv, ng, offset);
                offset[1] = save_offset;
            }
        }
        for (int n = 0; n < 3; ++n) {
            distance[n] = -g[n] * offset[n] - 0.5 * H[n][n] * offset[n] * offset[n];
            distance[n] = std::sqrt(smax(0.0, distance[n]));
            if (offset[n] < 0.0) distance[n] *= -1.0;
        }
        if (unstable) {
            distance[1] = center - distance[0] * distance[0] - distance[2] * distance[2];
            if (distance[1] <= 0.0) {
                distance[1] = 0.0;
                offset[1] = 0.0;
            } else {
                distance[1] = std::sqrt(distance[1]);
                offset[1] = std::fabs(offset[1]);
            }
        }
        return unstable;
    }

    // wvector1::direction(double*, double*) (types/wvector1.cc:79-82)
    static void direction(double nr, double nt, double np, double* de, double* az) {
        *de = to_degrees(wqm::atan2(nr, std::sqrt(nt * nt + np * np)));
        *az = to_degrees(wqm::atan2(np, -nt));
    }

    // wave_queue::build_eigenray (wave_queue.cc:495-702)
    void build_eigenray(size_t t1, size_t t2, size_t de, size_t az, double D[3][3][3]) {
        const size_t ic = curr->id(de, az);
        if (above_bounce(*curr, ic)) return;
        const size_t t = t1 * nt2 + t2;
        const double delta[3] = {time_step, de_seq.increment(de), az_seq.increment(az)};
        double offset[3], distance[3];
        compute_offsets(t, de, az, D, delta, offset, distance);
        wq3d_eigenray r{};
        r.step = step_count;
        r.t1 = (int32_t)t1;
        r.t2 = (int32_t)t2;
        r.de = (int32_t)de;
        r.az = (int32_t)az;
        r.travel_time = time + offset[0];
        r.source_de = de_seq(de) + offset[1];
        r.source_az = az_seq(az) + offset[2];
        r.surface = curr->srf[ic];
        r.bottom = curr->btm[ic];
        r.caustic = curr->cst[ic];
        r.upper = curr->upr[ic];
        r.lower = curr->lwr[ic];
        std::vector<double> inten(nf()), spread(nf());
        spreading.intensity(tgt_r[t], tgt_t[t], tgt_p[t], de, az, offset, distance, spread.data());
        for (size_t f = 0; f < nf(); ++f) inten[f] = -10.0 * wqm::log10(smax(spread[f], 1e-30));
        double dt = offset[0] / time_step;
        if (dt >= 0.0) {
            for (size_t f = 0; f < nf(); ++f)
                inten[f] = inten[f] + curr->atten[ic * nf() + f] * (1.0 - dt) +
                           next->atten[ic * nf() + f] * dt;
        } else {
            dt = 1.0 + dt;
            for (size_t f = 0; f < nf(); ++f)
                inten[f] = inten[f] + prev->atten[ic * nf() + f] * (1.0 - dt) +
                           curr->atten[ic * nf() + f] * dt;
        }
        bool above = false;
        for (size_t f = 0; f < nf(); ++f)
            if (inten[f] < thr.intensity) above = true;
        if (!above) return;
        double V[3][3][3], W[3][3][3];
        const Front* S[3] = {prev, curr, next};
        for (size_t nde = 0; nde < 3; ++nde) {
            for (size_t naz = 0; naz < 3; ++naz) {
                const size_t d = de + nde - 1;
                size_t a = az + naz - 1;
                if (az_boundary) {
                    if (az + naz == 0) {
                        a = max_az - 1;
                    } else if (a >= max_az) {
                        a = 0;
                    }
                }
                const size_t i = curr->id(d, a);
                for (int s = 0; s < 3; ++s)
                    direction(S[s]->nr[i], S[s]->nt[i], S[s]->np[i], &V[s][nde][naz], &W[s][nde][naz]);
            }
        }
        double center, g[3], H[3][3], Ho[3];
        make_taylor_coeff(V, delta, center, g, H);
        prod3(H, offset, Ho);
        r.target_de = center + inner3(g, offset) + 0.5 * inner3(offset, Ho);
        make_taylor_coeff(W, delta, center, g, H);
        prod3(H, offset, Ho);
        r.target_az = center + inner3(g, offset) + 0.5 * inner3(offset, Ho);
        rays.push_back(r);
        ray_int.insert(ray_int.end(), inten.begin(), inten.end());
        for (size_t f = 0; f < nf(); ++f) ray_phase.push_back(curr->phase[ic * nf() + f]);
    }
};

// ------------------------------------------------------------- Spreading
void Spreading::init(Queue* q) {
    w = q;
    const size_t nde = q->nde(), naz = q->naz();
    norm_de.assign(nde, 0.0);
    norm_az.assign(nde * naz, 0.0);
    init_area.assign(nde * naz, 0.0);
    spread.assign(q->nf(), 0.0);
    I_de.assign(q->nf(), 0.0);
    I_az.assign(q->nf(), 0.0);
    bw.assign(q->nf(), 0.0);
    dup.assign(naz, 0);
    for (size_t d = 0; d < nde - 1; ++d) {
        const double de1 = to_radians(q->de_seq(d));
        const double de2 = to_radians(q->de_seq(d + 1));
        norm_de[d] = de2 - de1;
        for (size_t a = 0; a < naz - 1; ++a) {
            const double az1 = to_radians(q->az_seq(a));
            const double az2 = to_radians(q->az_seq(a + 1));
            init_area[d * naz + a] = (wqm::sin(de2) - wqm::sin(de1)) * (az2 - az1);
            norm_az[d * naz + a] = init_area[d * naz + a] / norm_de[d];
        }
        norm_az[d * naz + naz - 1] = norm_az[d * naz + naz - 2];
        init_area[d * naz + naz - 1] = init_area[d * naz + naz - 2];  // spreading_ray.cc:35
    }
    norm_de[nde - 1] = norm_de[nde - 2];
    for (size_t a = 0; a < naz - 1; ++a) {
        init_area[(nde - 1) * naz + a] = init_area[(nde - 2) * naz + a];
        norm_az[(nde - 1) * naz + a] = norm_az[(nde - 2) * naz + a];
    }
    const double s = std::sqrt(kTwoPi);
    for (auto& x : norm_de) x = x / s;
    for (auto& x : norm_az) x = x / s;
}

void Spreading::gaussian(double d, double wdt, double A, double* out) {
    for (size_t f = 0; f < spread.size(); ++f) {
        bw[f] = spread[f] + OVERLAP * OVERLAP * wdt * wdt;
        out[f] = wqm::exp((-0.5 * d * d) / bw[f]) / std::sqrt(bw[f]) * A;
    }
}

static double ray_distance(const Front& w, size_t i, size_t j);

// wvector1::area of four points (types/wvector1.cc:151-169): triangles
// (p1, p2, p3) and (p1, p3, p4) by Heron's formula
static double quad_area(const Front& f, size_t i1, size_t i2, size_t i3, size_t i4) {
    double a = ray_distance(f, i1, i2);
    const double b = ray_distance(f, i1, i3);
    double c = ray_distance(f, i2, i3);
    const double area1 = 0.25 * std::sqrt((a + b + c) * (b + c - a) * (c + a - b) * (a + b - c));
    a = ray_distance(f, i1, i4);
    c = ray_distance(f, i4, i3);
    const double area2 = 0.25 * std::sqrt((a + b + c) * (b + c - a) * (c + a - b) * (a + b - c));
    return area1 + area2;
}

// spreading_ray::intensity (waveq3d/spreading_ray.cc:47-130): the ratio of
// the ray tube's initial area to its area now (two triangles per cell,
// interpolated in time toward prev or next), times c(target) / c(source);
// zero more than four beam widths outside the fan.  The reference steps to
// the lower cell with a size_t `--az`, which at az = 0 of a fan closing on
// itself indexes outside the front; here that cell is the fan's wrap,
// naz - 2, as the hybrid model's intensity takes it.
void Spreading::intensity_classic(double trho, double ttheta, double tphi, size_t de, size_t az,
                                  const double offset[3], double* result) {
    if (offset[1] < 0.0) --de;
    if (offset[2] < 0.0) az = (az > 0) ? az - 1 : w->naz() - 2;
    double c;
    w->profile.sound_speed(trho, ttheta, tphi, w->R, &c, nullptr, nullptr, nullptr);
    const Front& p1 = *w->curr;
    const double area1 = quad_area(p1, p1.id(de, az), p1.id(de + 1, az), p1.id(de + 1, az + 1), p1.id(de, az + 1));
    const Front& p2 = (offset[0] < 0.0) ? *w->prev : *w->next;
    const double area2 = quad_area(p2, p2.id(de, az), p2.id(de + 1, az), p2.id(de + 1, az + 1), p2.id(de, az + 1));
    double u = std::fabs(offset[0]) / w->time_step;
    const double area = (1.0 - u) * area1 + u * area2;
    double loss = init_area[de * w->naz() + az] * c / (area * init_c);
    u = std::fabs(offset[1]) / w->de_seq.increment(de) - 1.0;
    if (u > 4.0) loss *= 0.0;
    u = std::fabs(offset[2]) / w->az_seq.increment(az) - 1.0;
    if (u > 4.0) loss *= 0.0;
    for (size_t f = 0; f < spread.size(); ++f) result[f] = loss;
}

// spreading_hybrid_gaussian::intensity (spreading_hybrid_gaussian.cc:65-119)
void Spreading::intensity(double trho, double ttheta, double tphi, size_t de, size_t az,
                          const double offset[3], const double distance[3], double* result) {
    if (classic) {
        intensity_classic(trho, ttheta, tphi, de, az, offset, result);
        return;
    }
    double c;
    w->profile.sound_speed(trho, ttheta, tphi, w->R, &c, nullptr, nullptr, nullptr);
    for (size_t f = 0; f < spread.size(); ++f) spread[f] = kTwoPi * c / w->freq(f);
    for (size_t f = 0; f < spread.size(); ++f) spread[f] = spread[f] * spread[f];
    size_t d = de, a = az;
    double no[3] = {offset[0], offset[1], offset[2]};
    if (offset[2] < 0.0) {
        if (az > 0) {
            a = az - 1;
        } else if (w->az_boundary) {
            a = w->naz() - 2;
        }
        no[2] = offset[2] + w->az_seq.increment(a);
    }
    intensity_de(de, a, no, distance);
    no[0] = offset[0];
    no[1] = offset[1];
    no[2] = offset[2];
    if (offset[1] < 0.0 && !w->curr->edge[w->curr->id(de - 1, az)]) {
        d = de - 1;
        no[1] = offset[1] + w->de_seq.increment(d);
    }
    intensity_az(d, az, no, distance);
    for (size_t f = 0; f < spread.size(); ++f) result[f] = I_de[f] * I_az[f];
}

// spreading_hybrid_gaussian::intensity_de (spreading_hybrid_gaussian.cc:124-258)
void Spreading::intensity_de(size_t de, size_t az, const double offset[3],
                             const double distance[3]) {
    const size_t nf = spread.size();
    std::vector<double> g(nf);
    const Front& cf = *w->curr;
    int d = (int)de;
    double cell_width = width_de((size_t)d, az, offset);
    const double initial_width = cell_width;
    const double L = distance[1];
    double cell_dist = L - cell_width;
    gaussian(cell_dist, cell_width, norm_de[d], I_de.data());
    d = (int)de - 1;
    cell_width = width_de((size_t)d, az, offset);
    cell_dist = L + cell_width;
    gaussian(cell_dist, cell_width, norm_de[d], g.data());
    for (size_t f = 0; f < nf; ++f) I_de[f] += g[f];
    if (I_de[0] < 1e-10) return;
    for (d = (int)de - 2; d >= 0; --d) {
        const bool virtual_ray = cf.edge[cf.id(d + 1, az)] && cf.edge[cf.id(d, az)];
        double nn;
        cell_dist += cell_width;
        if (virtual_ray) {
            cell_dist += cell_width;
            nn = norm_de[d + 1];
        } else {
            cell_width = width_de((size_t)d, az, offset);
            cell_dist += cell_width;
            nn = norm_de[d];
        }
        const double old_tl = I_de[0];
        gaussian(cell_dist, cell_width, nn, g.data());
        for (size_t f = 0; f < nf; ++f) I_de[f] += g[f];
        if (I_de[0] / old_tl < THRESHOLD) {
            trace(0, old_tl, I_de[0], 1);
            break;
        }
        trace(0, old_tl, I_de[0], virtual_ray ? 2 : 0);
        if (virtual_ray) break;
    }
    cell_width = initial_width;
    cell_dist = L - cell_width;
    const size_t size = w->nde() - 1;
    for (d = (int)de + 1; (size_t)d < size; ++d) {
        const bool virtual_ray = cf.edge[cf.id(d + 1, az)] && cf.edge[cf.id(d, az)];
        double nn;
        cell_dist -= cell_width;
        if (virtual_ray) {
            cell_dist -= cell_width;
            nn = norm_de[d - 1];
        } else {
            cell_width = width_de((size_t)d, az, offset);
            cell_dist -= cell_width;
            nn = norm_de[d];
        }
        const double old_tl = I_de[0];
        gaussian(cell_dist, cell_width, nn, g.data());
        for (size_t f = 0; f < nf; ++f) I_de[f] += g[f];
        if (I_de[0] / old_tl < THRESHOLD) {
            trace(1, old_tl, I_de[0], 1);
            break;
        }
        trace(1, old_tl, I_de[0], virtual_ray ? 2 : 0);
        if (virtual_ray) break;
    }
}

// spreading_hybrid_gaussian::intensity_az (spreading_hybrid_gaussian.cc:264-425)
void Spreading::intensity_az(size_t de, size_t az, const double offset[3],
                             const double distance[3]) {
    const size_t nf = spread.size();
    const size_t naz = w->naz();
    std::vector<double> g(nf);
    const Front& cf = *w->curr;
    const size_t max_de = w->nde() - 2;
    const size_t max_az = naz - 1;
    size_t az_lower, az_upper;
    if (w->az_boundary) {
        az_lower = az_upper = az;
    } else {
        az_lower = 0;
        az_upper = max_az - 1;
    }
    std::fill(dup.begin(), dup.end(), 0);
    size_t a = az;
    dup[a] = 1;
    double cell_width = width_az(de, a, offset);
    const double initial_width = cell_width;
    const double L = distance[2];
    double cell_dist = L - cell_width;
    auto norm = [&](size_t dd, size_t aa) {
        return (dd >= max_de) ? norm_az[1 * naz + aa] : norm_az[dd * naz + aa];
    };
    gaussian(cell_dist, cell_width, norm(de, a), I_az.data());
    a = (az < 1) ? max_az - 1 : az - 1;
    dup[a] = 1;
    cell_width = width_az(de, a, offset);
    cell_dist = L + cell_width;
    gaussian(cell_dist, cell_width, norm(de, a), g.data());
    for (size_t f = 0; f < nf; ++f) I_az[f] += g[f];
    if (I_az[0] < 1e-10) return;
    if (a < 1) {
        a = max_az - 1;
    } else {
        --a;
    }
    while ((a % max_az) != az_lower) {
        if (dup[a]) break;
        dup[a] = 1;
        if (cf.edge[cf.id(de, a)]) break;
        cell_dist += cell_width;
        cell_width = width_az(de, a, offset);
        cell_dist += cell_width;
        const double old_tl = I_az[0];
        gaussian(cell_dist, cell_width, norm(de, a), g.data());
        for (size_t f = 0; f < nf; ++f) I_az[f] += g[f];
        if (I_az[0] / old_tl < THRESHOLD) {
            trace(2, old_tl, I_az[0], 1);
            break;
        }
        trace(2, old_tl, I_az[0], 0);
        if (a == 0) {
            a += max_az - 1;
        } else {
            --a;
        }
    }
    cell_width = initial_width;
    cell_dist = L - cell_width;
    a = az + 1;
    while ((a % max_az) != az_upper) {
        if (a == max_az) a = 0;
        if (dup[a]) break;
        dup[a] = 1;
        if (cf.edge[cf.id(de, a)]) break;
        cell_dist -= cell_width;
        cell_width = width_az(de, a, offset);
        cell_dist -= cell_width;
        const double old_tl = I_az[0];
        gaussian(cell_dist, cell_width, norm(de, a), g.data());
        for (size_t f = 0; f < nf; ++f) I_az[f] += g[f];
        if (I_az[0] / old_tl < THRESHOLD) {
            trace(3, old_tl, I_az[0], 1);
            break;
        }
        trace(3, old_tl, I_az[0], 0);
        ++a;
    }
}

// wvector1::distance (types/wvector1.cc:44-73) between two rays of a front
static double ray_distance(const Front& w, size_t i, size_t j) {
    const double r1 = w.pr[i], t1 = w.pt[i], p1 = w.pp[i];
    const double r2 = w.pr[j], t2 = w.pt[j], p2 = w.pp[j];
    const double sin_theta = wqm::sin(0.5 * (t1 - t2));
    const double sin_phi = wqm::sin(0.5 * (p1 - p2));
    const double dotnorm =
        1.0 - 2.0 * (sin_theta * sin_theta + wqm::sin(t1) * wqm::sin(t2) * sin_phi * sin_phi);
    const double dot = r1 * r2 * dotnorm;
    return std::sqrt(std::fabs(r1 * r1 + r2 * r2 - 2.0 * dot));
}

// spreading_hybrid_gaussian::width_de (spreading_hybrid_gaussian.cc:430-495)
double Spreading::width_de(size_t de, size_t az, const double offset[3]) const {
    const double u = std::fabs(offset[0]) / w->time_step;
    const double v = std::fabs(offset[2]) / w->az_seq.increment(az);
    const size_t max_az = w->naz() - 1;
    const size_t az_wrap = (az + 1 >= max_az) ? 0 : az + 1;
    const Front& p1 = *w->curr;
    double L1 = ray_distance(p1, p1.id(de, az), p1.id(de + 1, az));
    double length1, length2;
    if (v < 1e-10) {
        length1 = L1;
    } else {
        const double L2 = ray_distance(p1, p1.id(de, az_wrap), p1.id(de + 1, az_wrap));
        length1 = (1.0 - v) * L1 + v * L2;
    }
    if (u < 1e-10) return 0.5 * length1;
    const Front& p2 = (offset[0] < 0.0) ? *w->prev : *w->next;
    L1 = ray_distance(p2, p2.id(de, az), p2.id(de + 1, az));
    if (v < 1e-10) {
        length2 = L1;
    } else {
        const double L2 = ray_distance(p2, p2.id(de, az_wrap + 1), p2.id(de + 1, az_wrap + 1));
        length2 = (1.0 - v) * L1 + v * L2;
    }
    return 0.5 * ((1.0 - u) * length1 + u * length2);
}

// spreading_hybrid_gaussian::width_az (spreading_hybrid_gaussian.cc:500-571)
double Spreading::width_az(size_t de, size_t az, const double offset[3]) const {
    const double u = std::fabs(offset[0]) / w->time_step;
    const double v = std::fabs(offset[1]) / w->de_seq.increment(de);
    const size_t max_az = w->naz() - 1;
    const size_t max_de = w->nde() - 1;
    const size_t de_upper = (de + 1 >= max_de) ? max_de - 2 : de;
    const size_t az_wrap = (az + 1 > max_az) ? 0 : az + 1;
    const Front& p1 = *w->curr;
    double L1 = ray_distance(p1, p1.id(de, az), p1.id(de, az_wrap));
    double length1, length2;
    if (v < 1e-10 || std::fabs(v - 1.0) < 1e-10) {
        length1 = L1;
    } else {
        const double L2 = ray_distance(p1, p1.id(de_upper + 1, az), p1.id(de_upper + 1, az_wrap));
        length1 = (1.0 - v) * L1 + v * L2;
    }
    if (u < 1e-10) return 0.5 * length1;
    const Front& p2 = (offset[0] < 0.0) ? *w->prev : *w->next;
    L1 = ray_distance(p2, p2.id(de, az), p2.id(de, az_wrap));
    if (v < 1e-10 || std::fabs(v - 1.0) < 1e-10) {
        length2 = L1;
    } else {
        const double L2 = ray_distance(p2, p2.id(de_upper + 1, az), p2.id(de_upper + 1, az_wrap));
        length2 = (1.0 - v) * L1 + v * L2;
    }
    return 0.5 * ((1.0 - u) * length1 + u * length2);
}

}  // namespace orc

// ============================================================== C API
extern "C" {

struct orc_handle {
    orc::Queue q;
};

static thread_local std::string g_orc_error;

const char* orc_last_error(void) { return g_orc_error.c_str(); }

void* orc_create(const wq3d_env_desc* env, const wq3d_fan_desc* fan) {
    g_orc_error.clear();
    auto* h = new orc_handle();
    try {
        h->q.create(*env, *fan);
    } catch (const orc::Error& e) {
        g_orc_error = e.msg;
        delete h;
        return nullptr;
    }
    return h;
}

void orc_destroy(void* h) { delete static_cast<orc_handle*>(h); }

void orc_set_thresholds(void* h, const wq3d_thresholds* t) { static_cast<orc_handle*>(h)->q.thr = *t; }

void orc_set_listeners(void* h, int32_t verbs, int32_t refl) {
    auto& q = static_cast<orc_handle*>(h)->q;
    q.want_verbs = verbs != 0;
    q.want_refl = refl != 0;
}

void orc_step(void* h, int64_t n) {
    auto& q = static_cast<orc_handle*>(h)->q;
    for (int64_t k = 0; k < n; ++k) q.step();
}

double orc_time(void* h) { return static_cast<orc_handle*>(h)->q.time; }

void orc_source(void* h, double* r, double* t, double* p) {
    auto& q = static_cast<orc_handle*>(h)->q;
    *r = q.src_r;
    *t = q.src_t;
    *p = q.src_p;
}

int64_t orc_num_eigenrays(void* h) { return (int64_t) static_cast<orc_handle*>(h)->q.rays.size(); }

/* Test hook: trace the hybrid Gaussian walks (on / off), and copy out the
 * trace (4 doubles per walk step; returns the number of doubles, copying at
 * most cap of them). */
void orc_trace_walks(void* h, int32_t on) {
    auto& sp = static_cast<orc_handle*>(h)->q.spreading;
    sp.trace_walks = on != 0;
    sp.walk_trace.clear();
}
int64_t orc_walk_trace(void* h, double* out, int64_t cap) {
    const auto& t = static_cast<orc_handle*>(h)->q.spreading.walk_trace;
    const int64_t n = (int64_t)t.size();
    for (int64_t k = 0; k < n && k < cap; ++k) out[k] = t[(size_t)k];
    return n;
}

/* Test hook: the number of eigenrays whose DE offset took its sign from the
 * geometry (compute_offsets' unstable branch) and how many of those signs
 * were negative. */
void orc_sign_rule_counts(void* h, int64_t* taken, int64_t* negative) {
    *taken = static_cast<orc_handle*>(h)->q.sign_rule_count;
    *negative = static_cast<orc_handle*>(h)->q.sign_flip_count;
}
int64_t orc_num_eigenverbs(void* h) { return (int64_t) static_cast<orc_handle*>(h)->q.verbs.size(); }
int64_t orc_num_reflections(void* h) { return (int64_t) static_cast<orc_handle*>(h)->q.refl.size(); }

/* Copy out and clear all pending records (caller sizes buffers with orc_num_*). */
void orc_take_eigenrays(void* h, wq3d_eigenray* out, double* inten, double* phase) {
    auto& q = static_cast<orc_handle*>(h)->q;
    std::copy(q.rays.begin(), q.rays.end(), out);
    std::copy(q.ray_int.begin(), q.ray_int.end(), inten);
    std::copy(q.ray_phase.begin(), q.ray_phase.end(), phase);
    q.rays.clear();
    q.ray_int.clear();
    q.ray_phase.clear();
}
void orc_take_eigenverbs(void* h, wq3d_eigenverb* out, double* power) {
    auto& q = static_cast<orc_handle*>(h)->q;
    std::copy(q.verbs.begin(), q.verbs.end(), out);
    std::copy(q.verb_power.begin(), q.verb_power.end(), power);
    q.verbs.clear();
    q.verb_power.clear();
}
void orc_take_reflections(void* h, wq3d_reflection* out) {
    auto& q = static_cast<orc_handle*>(h)->q;
    std::copy(q.refl.begin(), q.refl.end(), out);
    q.refl.clear();
}

static void copy_vec(const std::vector<double>& v, double* dst) {
    if (dst) std::copy(v.begin(), v.end(), dst);
}
static void copy_ivec(const std::vector<int32_t>& v, int32_t* dst) {
    if (dst) std::copy(v.begin(), v.end(), dst);
}

void orc_snapshot(void* h, int32_t slot, wq3d_front_view* view) {
    auto& q = static_cast<orc_handle*>(h)->q;
    const orc::Front* w = slot == WQ3D_PAST ? q.past
                          : slot == WQ3D_PREV ? q.prev
                          : slot == WQ3D_CURR ? q.curr
                                              : q.next;
    copy_vec(w->pr, view->pos_rho);
    copy_vec(w->pt, view->pos_theta);
    copy_vec(w->pp, view->pos_phi);
    copy_vec(w->gr, view->pgrad_rho);
    copy_vec(w->gt, view->pgrad_theta);
    copy_vec(w->gp, view->pgrad_phi);
    copy_vec(w->nr, view->ndir_rho);
    copy_vec(w->nt, view->ndir_theta);
    copy_vec(w->np, view->ndir_phi);
    copy_vec(w->hr, view->dgrad_rho);
    copy_vec(w->ht, view->dgrad_theta);
    copy_vec(w->hp, view->dgrad_phi);
    copy_vec(w->c, view->sound_speed);
    copy_vec(w->cr, view->sgrad_rho);
    copy_vec(w->ct, view->sgrad_theta);
    copy_vec(w->cp, view->sgrad_phi);
    copy_vec(w->dist, view->distance);
    copy_vec(w->path, view->path_length);
    copy_vec(w->atten, view->attenuation);
    copy_vec(w->phase, view->phase);
    copy_ivec(w->srf, view->surface);
    copy_ivec(w->btm, view->bottom);
    copy_ivec(w->cst, view->caustic);
    copy_ivec(w->upr, view->upper);
    copy_ivec(w->lwr, view->lower);
    if (view->on_edge) std::copy(w->edge.begin(), w->edge.end(), view->on_edge);
}

/* Shared elementary functions applied element-wise (accuracy tests):
 * fn 0 sin, 1 cos, 2 exp, 3 log, 4 log10, 5 atan2(x, y), 6 asin, 7 10^x. */
/* eigenray_collection (eigenrays/eigenray_collection.cc): add_eigenray
 * (:90-101) appends each ray to its target's list in arrival order and keeps
 * the initial time; sum_eigenrays (:105-191) then sums each list.  Same
 * arguments and outputs as wq3d_sum_eigenrays. */
int32_t orc_sum_eigenrays(const wq3d_eigenray* rays, const double* intensity, const double* phase,
                          int64_t num_rays, const double* freq, int32_t nf, int64_t num_t1, int64_t num_t2,
                          int32_t coherent, wq3d_eigenray_total* totals, double* total_intensity,
                          double* total_phase) {
    using namespace orc;
    std::vector<std::vector<int64_t>> lists((size_t)(num_t1 * num_t2));
    std::vector<double> initial((size_t)(num_t1 * num_t2), 0.0);
    for (int64_t k = 0; k < num_rays; ++k) {  // add_eigenray
        const int64_t t1 = rays[k].t1, t2 = rays[k].t2;
        if (t1 < 0 || t1 >= num_t1 || t2 < 0 || t2 >= num_t2) return 1;
        const size_t t = (size_t)(t1 * num_t2 + t2);
        lists[t].push_back(k);
        const double old_initial = initial[t];
        const double new_initial = rays[k].travel_time;
        if (old_initial <= 0.0 || old_initial > new_initial) initial[t] = new_initial;
    }
    for (int64_t t1 = 0; t1 < num_t1; ++t1) {  // sum_eigenrays
        for (int64_t t2 = 0; t2 < num_t2; ++t2) {
            const size_t t = (size_t)(t1 * num_t2 + t2);
            const std::vector<int64_t>& ray_list = lists[t];
            double time = 0.0, source_de = 0.0, source_az_x = 0.0, source_az_y = 0.0;
            double target_de = 0.0, target_az_x = 0.0, target_az_y = 0.0;
            int surface = -1, bottom = -1, caustic = -1, upper = -1, lower = -1;
            double wgt = 0.0, max_a = 0.0;
            for (int f = 0; f < nf; ++f) {
                double ph_re = 0.0, ph_im = 0.0;
                for (int64_t r : ray_list) {
                    const wq3d_eigenray& ray = rays[r];
                    double a = wqm::pow10(intensity[r * nf + f] / -20.0);
                    if (coherent) {
                        double p = kTwoPi * freq[f] * ray.travel_time + phase[r * nf + f];
                        p = std::fmod(p, kTwoPi);
                        const double vre = a * wqm::cos(p), vim = a * wqm::sin(p);
                        ph_re += vre;
                        ph_im += vim;
                    } else {
                        ph_re += a;
                    }
                    a *= a;
                    wgt += a;
                    time += a * ray.travel_time;
                    source_de += a * ray.source_de;
                    source_az_x += a * wqm::sin(to_radians(ray.source_az));
                    source_az_y += a * wqm::cos(to_radians(ray.source_az));
                    target_de += a * ray.target_de;
                    target_az_x += a * wqm::sin(to_radians(ray.target_az));
                    target_az_y += a * wqm::cos(to_radians(ray.target_az));
                    if (a > max_a) {
                        max_a = a;
                        surface = ray.surface;
                        bottom = ray.bottom;
                        caustic = ray.caustic;
                        upper = ray.upper;
                        lower = ray.lower;
                    }
                }
                const wqm::cplx phasor = {ph_re, ph_im};
                total_intensity[t * nf + f] = -20.0 * wqm::log10(smax(1e-15, wqm::cabs(phasor)));
                total_phase[t * nf + f] = wqm::carg(phasor);
            }
            wq3d_eigenray_total& T = totals[t];
            T.travel_time = time / wgt;
            T.source_de = source_de / wgt;
            T.source_az = 90.0 - to_degrees(wqm::atan2(source_az_y, source_az_x));
            T.target_de = target_de / wgt;
            T.target_az = 90.0 - to_degrees(wqm::atan2(target_az_y, target_az_x));
            T.initial_time = initial[t];
            T.surface = surface;
            T.bottom = bottom;
            T.caustic = caustic;
            T.upper = upper;
            T.lower = lower;
            T.count = (int32_t)ray_list.size();
        }
    }
    return 0;
}

void orc_math(int32_t fn, const double* x, const double* y, double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = wqm::sin(x[i]); break;
            case 1: out[i] = wqm::cos(x[i]); break;
            case 2: out[i] = wqm::exp(x[i]); break;
            case 3: out[i] = wqm::log(x[i]); break;
            case 4: out[i] = wqm::log10(x[i]); break;
            case 5: out[i] = wqm::atan2(x[i], y[i]); break;
            case 6: out[i] = wqm::asin(x[i]); break;
            default: out[i] = wqm::pow10(x[i]); break;
        }
    }
}

/* Interpolate a grid descriptor at one location (pins gen_grid). */
double orc_grid_interp(const wq3d_grid* g, const double* loc, double* deriv) {
    orc::Grid grid;
    try {
        grid.load(*g);
    } catch (const orc::Error& e) {
        g_orc_error = e.msg;
        return NAN;
    }
    return grid.interpolate(loc, deriv);
}

/* Sound speed of a profile descriptor at one location (pins profiles). */
void orc_sound_speed(const wq3d_profile_desc* d, double rho, double theta, double phi, double R,
                     double* out4) {
    orc::Profile p;
    p.load(*d);
    p.sound_speed(rho, theta, phi, R, &out4[0], &out4[1], &out4[2], &out4[3]);
}

/* Attenuation of a profile descriptor (pins Thorp / constant). */
void orc_attenuation(const wq3d_profile_desc* d, const wq3d_seq* freq, double alt, double dist,
                     double* out) {
    orc::Profile p;
    p.load(*d);
    orc::Seq f;
    f.load(*freq);
    p.attenuation(alt, dist, f, out);
}

/* Reflection loss of a boundary descriptor (pins Rayleigh). */
void orc_reflect_loss(const wq3d_boundary_desc* d, double grazing, double* amp, double* phase) {
    orc::Boundary b;
    b.load(*d);
    b.reflect_loss(grazing, 1, amp, phase);
}

/* The distance-only part of compute_offsets on a 3 x 3 x 3 neighbourhood
 * (d27 in [time][DE][AZ] order): offsets, distances, and the returned
 * unstable flag (pinned by exact quadratic neighbourhoods). */
int32_t orc_compute_offsets_core(const double* d27, const double* delta, int32_t counters_differ,
                                 double* offset, double* distance) {
    double D[3][3][3];
    for (int k = 0; k < 27; ++k) D[k / 9][(k / 3) % 3][k % 3] = d27[k];
    return orc::Queue::offsets_core(D, delta, counters_differ != 0, offset, distance) ? 1 : 0;
}

/* Height and normal of a boundary descriptor (pins boundary_grid). */
void orc_height(const wq3d_boundary_desc* d, double rho, double theta, double phi, double* out4) {
    orc::Boundary b;
    b.load(*d);
    b.height(rho, theta, phi, &out4[0], &out4[1]);
}

}  // extern "C"
