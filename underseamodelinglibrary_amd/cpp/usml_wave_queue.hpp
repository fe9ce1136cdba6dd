// usml_wave_queue.hpp - C++ host side of the WaveQ3D path over libwq3d.so.
//
// Header-only C++17 mirror of the reference interface a caller of the hot
// path uses, with the same names, argument meaning and error behaviour:
//   usml::types  seq_linear / seq_log / seq_data (types/seq_linear.h:60-110,
//                types/seq_log.h:40-100, types/seq_data.h:86-120), wposition1
//                and wposition (types/wposition1.cc, types/wposition.cc)
//   usml::ocean  profile_linear / profile_munk / profile_grid, attenuation
//                models, boundary_flat / boundary_grid, reflect_loss_constant /
//                reflect_loss_rayleigh, volume_flat, ocean_model
//                (ocean/ocean_model.h:62-70)
//   usml::waveq3d::wave_queue (waveq3d/wave_queue.h:78-831): constructor,
//                step(), time(), source_pos(), frequencies(), source_de/az(),
//                num_de/az(), runID(), curr()/next()/prev()/past(), listener
//                registration and wave_thresholds.
// Listeners are kept in std::set<listener*> and called in pointer order, as
// eigenray_notifier.h:70 and eigenverb_notifier.h:56 do.  Construction errors
// throw std::invalid_argument (bad arguments, unsupported models), other
// library failures std::runtime_error.  Every step runs on the GPU library;
// there is no host fallback.
#ifndef USML_AMD_WAVE_QUEUE_HPP
#define USML_AMD_WAVE_QUEUE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/wq3d.h"

namespace usml_amd {

inline void check(wq3d_status s) {
    if (s == WQ3D_OK) return;
    const std::string msg = wq3d_last_error();
    if (s == WQ3D_ERR_ARGUMENT || s == WQ3D_ERR_UNSUPPORTED) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// ------------------------------------------------------------------ types
namespace types {

inline double to_degrees(double t) { return t * (180.0 / M_PI); }
inline double to_radians(double t) { return t * (M_PI / 180.0); }
inline double to_latitude(double t) { return 90.0 - to_degrees(t); }
inline double to_colatitude(double t) { return to_radians(90.0 - t); }

// wposition::earth_radius, process-wide and mutable (types/wposition.cc:13)
inline double& earth_radius() {
    static double r = 6371000.0;
    return r;
}

// WGS-84 radius of curvature (types/wposition.cc:33-42)
inline void compute_earth_radius(double latitude) {
    const double a = 6378137.0;
    const double f = 1.0 / 298.257223563;
    const double e2 = f * (2.0 - f);
    const double s = std::sin(to_radians(latitude));
    const double w = std::sqrt(1.0 - e2 * s * s);
    const double rm = a * (1 - e2) / (w * w * w);
    const double rv = a / w;
    earth_radius() = std::sqrt(rm * rv);
}

class seq_vector {
public:
    virtual ~seq_vector() = default;
    size_t size() const { return _v.size(); }
    // clamped accessors (types/seq_vector.h:99-143)
    double operator()(long i) const { return _v[clamp(i)]; }
    double increment(long i) const { return _inc[clamp(i)]; }
    const std::vector<double>& values() const { return _v; }
    wq3d_seq abi() const { return wq3d_seq{(int64_t)_v.size(), _v.data(), _inc.data(), _search, 0}; }

protected:
    // the reference's index is a size_t: a negative index wraps to a huge
    // value and clamps to the last element, like one past the end
    size_t clamp(long i) const {
        return (i < 0 || (size_t)i >= _v.size()) ? _v.size() - 1 : (size_t)i;
    }
    std::vector<double> _v, _inc;
    int32_t _search = WQ3D_SEQ_DATA;
};

// values accumulate v += increment (types/seq_linear.h:104-110)
class seq_linear : public seq_vector {
public:
    seq_linear(double first, double increment, size_t size) { fill(first, increment, size); }
    seq_linear(double first, double increment, double last) {
        size_t n = 1;
        if (increment != 0.0)
            n = (size_t)std::max(1.0, std::floor(1.0 + (last + increment / 10 - first) / increment));
        fill(first, increment, n);
    }

private:
    void fill(double first, double increment, size_t size) {
        _search = WQ3D_SEQ_LINEAR;
        _v.resize(size);
        _inc.assign(size, increment);
        double v = first;
        for (size_t n = 0; n < size; ++n) {
            _v[n] = v;
            v += increment;
        }
    }
};

// values accumulate v *= increment; increments are differences (types/seq_log.h:84)
class seq_log : public seq_vector {
public:
    seq_log(double first, double increment, size_t size) {
        _search = WQ3D_SEQ_LOG;
        _v.resize(size);
        _inc.assign(size, 0.0);
        double v = first;
        for (size_t n = 0; n < size; ++n) {
            _v[n] = v;
            v *= increment;
            if (n > 0) {
                _inc[n - 1] = _v[n] - _v[n - 1];
                _inc[n] = _inc[n - 1];
            }
        }
    }
};

// explicit monotonic data; increments are forward differences (types/seq_data.h:86-120)
class seq_data : public seq_vector {
public:
    explicit seq_data(const std::vector<double>& data) {
        _search = WQ3D_SEQ_DATA;
        _v = data;
        _inc.assign(data.size(), 0.0);
        if (data.size() >= 2) {
            double left = data[1] - data[0];
            _inc[0] = left;
            for (size_t n = 1; n < data.size(); ++n) {
                double right = left;
                if (n < data.size() - 1) {
                    right = data[n + 1] - data[n];
                    if (left * right <= 0) throw std::invalid_argument("series not monotonic");
                }
                left = right;
                _inc[n] = left;
            }
        }
    }
};

struct wposition1 {
    double rho, theta, phi;
    wposition1(double latitude = 0.0, double longitude = 0.0, double altitude = 0.0)
        : rho(altitude + earth_radius()), theta(to_colatitude(latitude)), phi(to_radians(longitude)) {}
    static wposition1 rtp(double r, double t, double p) {
        wposition1 w;
        w.rho = r;
        w.theta = t;
        w.phi = p;
        return w;
    }
    double latitude() const { return to_latitude(theta); }
    double longitude() const { return to_degrees(phi); }
    double altitude() const { return rho - earth_radius(); }
};

// rows x cols target grid, row-major
struct wposition {
    size_t rows, cols;
    std::vector<double> rho, theta, phi;
    wposition(size_t r, size_t c, double latitude = 0.0, double longitude = 0.0, double altitude = 0.0)
        : rows(r), cols(c), rho(r * c, altitude + earth_radius()), theta(r * c, to_colatitude(latitude)),
          phi(r * c, to_radians(longitude)) {}
    size_t size1() const { return rows; }
    size_t size2() const { return cols; }
    void set(size_t t1, size_t t2, double latitude, double longitude, double altitude) {
        const size_t k = t1 * cols + t2;
        rho[k] = altitude + earth_radius();
        theta[k] = to_colatitude(latitude);
        phi[k] = to_radians(longitude);
    }
};

}  // namespace types

// ------------------------------------------------------------------ ocean
namespace ocean {

// gen_grid<N> data, N <= 3, row-major with axis 0 slowest
struct data_grid {
    std::vector<std::shared_ptr<const types::seq_vector>> axes;
    std::vector<double> data;
    int interp[3] = {WQ3D_INTERP_LINEAR, WQ3D_INTERP_LINEAR, WQ3D_INTERP_LINEAR};
    bool edge_limit[3] = {true, true, true};
};

struct attenuation_model {
    int32_t kind = WQ3D_ATTEN_THORP;
    double coeff = 0.0;
};
inline attenuation_model attenuation_thorp() { return {}; }
inline attenuation_model attenuation_constant(double coefficient) { return {WQ3D_ATTEN_CONSTANT, coefficient}; }

struct profile_model {
    int32_t kind = WQ3D_PROFILE_LINEAR;
    double p[4] = {1500.0, 0.0, 0.0, 0.0};
    bool flat_earth = false;
    attenuation_model attenuation;
    std::shared_ptr<const data_grid> grid;
};
// ocean/profile_linear.h:53-92
inline profile_model profile_linear(double c0 = 1500.0, attenuation_model a = {}) {
    profile_model m;
    m.p[0] = c0;
    m.attenuation = a;
    return m;
}
inline profile_model profile_linear(double c0, double g0, double z1, double g1, attenuation_model a = {}) {
    profile_model m;
    m.p[0] = c0;
    m.p[1] = g0;
    m.p[2] = z1;
    m.p[3] = g1;
    m.attenuation = a;
    return m;
}
// ocean/profile_munk.h:87-94
inline profile_model profile_munk(double axis_depth = 1300.0, double scale = 1300.0, double axis_speed = 1500.0,
                                  double epsilon = 7.37e-3, attenuation_model a = {}) {
    profile_model m;
    m.kind = WQ3D_PROFILE_MUNK;
    m.p[0] = axis_depth;
    m.p[1] = scale;
    m.p[2] = axis_speed;
    m.p[3] = epsilon;
    m.attenuation = a;
    return m;
}
// ocean/profile_grid.h:68-130
inline profile_model profile_grid(std::shared_ptr<const data_grid> g, attenuation_model a = {}) {
    if (!g || (g->axes.size() != 1 && g->axes.size() != 3))
        throw std::invalid_argument("sound speed grid must be 1-D or 3-D");
    profile_model m;
    m.kind = WQ3D_PROFILE_GRID;
    m.grid = std::move(g);
    m.attenuation = a;
    return m;
}

struct reflect_loss_model {
    int32_t kind = WQ3D_LOSS_CONSTANT;
    double params[5] = {0, 0, 0, 0, 0};
};
inline reflect_loss_model reflect_loss_constant(double amplitude, double phase = 0.0) {
    reflect_loss_model m;
    m.params[0] = amplitude;
    m.params[1] = phase;
    return m;
}
// density ratio, speed ratio, attenuation (dB/wavelength), shear speed ratio,
// shear attenuation (ocean/reflect_loss_rayleigh.cc:27-39)
inline reflect_loss_model reflect_loss_rayleigh(double density, double speed, double att_bottom = 0.0,
                                                double speed_shear = 0.0, double att_shear = 0.0) {
    reflect_loss_model m;
    m.kind = WQ3D_LOSS_RAYLEIGH;
    m.params[0] = density;
    m.params[1] = speed;
    m.params[2] = att_bottom;
    m.params[3] = speed_shear;
    m.params[4] = att_shear;
    return m;
}
enum bottom_type_enum { CLAY, SILT, SAND, GRAVEL, MORAINE, CHALK, LIMESTONE, BASALT };
inline reflect_loss_model reflect_loss_rayleigh(bottom_type_enum type) {
    static const double table[8][5] = {
        {1.5, 1.00, 0.2, 0.00, 1.0}, {1.7, 1.05, 1.0, 0.00, 1.5}, {1.9, 1.10, 0.8, 0.00, 2.5},
        {2.0, 1.20, 0.6, 0.00, 1.5}, {2.1, 1.30, 0.4, 0.40, 1.0}, {2.2, 1.60, 0.2, 0.67, 0.5},
        {2.4, 2.00, 0.1, 1.00, 0.2}, {2.7, 3.50, 0.1, 1.67, 0.2}};
    const double* r = table[(int)type];
    return reflect_loss_rayleigh(r[0], r[1], r[2], r[3], r[4]);
}

struct boundary_model {
    int32_t kind = WQ3D_BOUNDARY_FLAT;
    double rho = 0.0;
    double normal_rho = -1.0;
    reflect_loss_model loss;
    std::shared_ptr<const data_grid> grid;
};
// ocean/boundary_flat.cc:19-33; the ocean surface (depth 0) defaults to a
// pressure-release loss of (0 dB, pi)
inline boundary_model boundary_flat(double depth = 0.0) {
    boundary_model b;
    b.rho = types::earth_radius() - std::fabs(depth);
    if (std::fabs(depth) < 1e-6) {
        b.normal_rho = -1.0;
        b.loss = reflect_loss_constant(0.0, M_PI);
    } else {
        b.normal_rho = 1.0;
    }
    return b;
}
inline boundary_model boundary_flat(double depth, reflect_loss_model loss) {
    boundary_model b = boundary_flat(depth);
    b.loss = loss;
    return b;
}
// ocean/boundary_grid.h:88-196; rho = f(theta[, phi])
inline boundary_model boundary_grid(std::shared_ptr<const data_grid> g, reflect_loss_model loss = {}) {
    if (!g || (g->axes.size() != 1 && g->axes.size() != 2))
        throw std::invalid_argument("bathymetry must be 1-D or 2-D");
    boundary_model b;
    b.kind = WQ3D_BOUNDARY_GRID;
    b.normal_rho = 1.0;
    b.grid = std::move(g);
    b.loss = loss;
    return b;
}

struct volume_flat {
    double depth = 0.0, thickness = 0.0;
};

struct ocean_model {
    boundary_model surface, bottom;
    profile_model profile;
    std::vector<volume_flat> volumes;
    ocean_model(boundary_model s, boundary_model b, profile_model p) : surface(s), bottom(b), profile(p) {}
    void add_volume(volume_flat v) { volumes.push_back(v); }
    using csptr = std::shared_ptr<const ocean_model>;  // ocean/ocean_model.h:38
};

}  // namespace ocean

// ---------------------------------------------------------------- waveq3d
namespace waveq3d {

// eigenrays/eigenray_model.h:24-102
struct eigenray_model {
    double time = 0.0;
    std::vector<double> intensity, phase;
    double source_de = 0, source_az = 0, target_de = 0, target_az = 0;
    int surface = 0, bottom = 0, caustic = 0, upper = 0, lower = 0;
    using csptr = std::shared_ptr<const eigenray_model>;
};

// eigenverbs/eigenverb_model.h:34-160 (angles in radians)
struct eigenverb_model {
    double time = 0.0;
    std::vector<double> power;
    double length = 0, width = 0;
    types::wposition1 position;
    double direction = 0, grazing = 0, sound_speed = 0;
    size_t de_index = 0, az_index = 0;
    double source_de = 0, source_az = 0;
    int surface = 0, bottom = 0, caustic = 0, upper = 0, lower = 0;
    using csptr = std::shared_ptr<const eigenverb_model>;
};

// eigenrays/eigenray_listener.h:41,53
class eigenray_listener {
public:
    virtual ~eigenray_listener() = default;
    virtual void add_eigenray(size_t t1, size_t t2, eigenray_model::csptr ray, size_t runID) = 0;
    virtual void check_eigenrays(double /*time*/, size_t /*runID*/) {}
};

// eigenverbs/eigenverb_listener.h:40
class eigenverb_listener {
public:
    virtual ~eigenverb_listener() = default;
    virtual void add_eigenverb(eigenverb_model::csptr verb, int interface_type) = 0;
};

// waveq3d/reflection_listener.h
class reflection_listener {
public:
    virtual ~reflection_listener() = default;
    virtual bool reflect(double time, int de, int az, double dt, double grazing, double speed,
                         const types::wposition1& position, const types::wposition1& ndirection,
                         int type) = 0;
};

enum spreading_type { CLASSIC_RAY = WQ3D_SPREADING_CLASSIC_RAY, HYBRID_GAUSSIAN = WQ3D_SPREADING_HYBRID_GAUSSIAN };

// host snapshot of one queue slot (waveq3d/wave_front.h:152-258), de-major
struct wave_front {
    size_t num_de = 0, num_az = 0, num_freq = 0;
    std::vector<double> pos_rho, pos_theta, pos_phi, ndir_rho, ndir_theta, ndir_phi;
    std::vector<double> sound_speed, distance, path_length, attenuation, phase;
    std::vector<int32_t> surface, bottom, caustic, upper, lower;
    std::vector<uint8_t> on_edge;
    size_t index(size_t de, size_t az) const { return de * num_az + az; }
};

class wave_queue {
public:
    // waveq3d/wave_queue.h:122-126: the ocean is shared and immutable
    wave_queue(const ocean::ocean_model::csptr& ocean, std::shared_ptr<const types::seq_vector> freq,
               const types::wposition1& pos, std::shared_ptr<const types::seq_vector> de,
               std::shared_ptr<const types::seq_vector> az, double time_step,
               const types::wposition* target_pos = nullptr, spreading_type type = HYBRID_GAUSSIAN)
        : _ocean(ocean), _freq(std::move(freq)), _de(std::move(de)), _az(std::move(az)), _time_step(time_step),
          _targets(target_pos) {
        if (!_ocean || !_freq || !_de || !_az) throw std::invalid_argument("null ocean or sequence");
        build_env(*_ocean);
        wq3d_fan_desc d{};
        d.freq = _freq->abi();
        d.de = _de->abi();
        d.az = _az->abi();
        d.src_rho = pos.rho;
        d.src_theta = pos.theta;
        d.src_phi = pos.phi;
        d.time_step = time_step;
        d.earth_radius = types::earth_radius();
        if (target_pos) {
            d.num_t1 = (int64_t)target_pos->rows;
            d.num_t2 = (int64_t)target_pos->cols;
            d.target_rho = target_pos->rho.data();
            d.target_theta = target_pos->theta.data();
            d.target_phi = target_pos->phi.data();
        }
        d.spreading = (int32_t)type;
        _thr.intensity = 300.0;
        _thr.eigenverb = 1e-30;
        _thr.max_bottom = _thr.max_surface = _thr.max_caustic = _thr.max_upper = _thr.max_lower = 999;
        d.thresholds = _thr;
        wq3d_status s = wq3d_fan_create(_env, &d, &_fan);
        if (s != WQ3D_OK) {
            wq3d_env_destroy(_env);
            check(s);
        }
        double r, t, p;
        check(wq3d_source(_fan, &r, &t, &p));
        _source = types::wposition1::rtp(r, t, p);
    }
    wave_queue(const wave_queue&) = delete;
    wave_queue& operator=(const wave_queue&) = delete;
    ~wave_queue() {
        if (_fan) wq3d_fan_destroy(_fan);
        if (_env) wq3d_env_destroy(_env);
    }

    // -- accessors (waveq3d/wave_queue.h:136-241)
    double time() const { return _time; }
    double time_step() const { return _time_step; }
    const types::wposition1& source_pos() const { return _source; }
    std::shared_ptr<const types::seq_vector> frequencies() const { return _freq; }
    double source_de(size_t i) const { return (*_de)((long)i); }
    double source_az(size_t i) const { return (*_az)((long)i); }
    size_t num_de() const { return _de->size(); }
    size_t num_az() const { return _az->size(); }
    const types::wposition* targets() const { return _targets; }
    size_t runID() const { return _run_id; }
    void runID(size_t id) { _run_id = id; }

    // -- wave_thresholds (waveq3d/wave_thresholds.h:54-192)
    double intensity_threshold() const { return -_thr.intensity; }
    void intensity_threshold(double dB) {
        _thr.intensity = std::fabs(dB);
        check(wq3d_set_thresholds(_fan, &_thr));
    }
    void max_bottom(int n) {
        _thr.max_bottom = n;
        check(wq3d_set_thresholds(_fan, &_thr));
    }
    void max_surface(int n) {
        _thr.max_surface = n;
        check(wq3d_set_thresholds(_fan, &_thr));
    }

    // -- listeners (eigenray_notifier.h:28-33, eigenverb_notifier.h:28-33)
    void add_eigenray_listener(eigenray_listener* l) { _ray_listeners.insert(l); }
    void remove_eigenray_listener(eigenray_listener* l) { _ray_listeners.erase(l); }
    void add_eigenverb_listener(eigenverb_listener* l) {
        _verb_listeners.insert(l);
        push_listeners();
    }
    void remove_eigenverb_listener(eigenverb_listener* l) {
        _verb_listeners.erase(l);
        push_listeners();
    }
    bool has_eigenverb_listeners() const { return !_verb_listeners.empty(); }
    void add_reflection_listener(reflection_listener* l) {
        _refl_listeners.insert(l);
        push_listeners();
    }
    void remove_reflection_listener(reflection_listener* l) {
        _refl_listeners.erase(l);
        push_listeners();
    }

    // wave_queue::step() (waveq3d/wave_queue.cc:155-192)
    void step() { step_many(1); }

    // n steps queued on the device at once; listener calls are replayed per
    // step afterwards in the order step() would have made them
    void step_many(int64_t n) {
        check(wq3d_step(_fan, n));
        if (_ray_listeners.empty() && _verb_listeners.empty() && _refl_listeners.empty()) {
            for (int64_t k = 0; k < n; ++k) _time += _time_step;
            _steps += n;
            return;
        }
        deliver(n);
    }

    const wave_front& past() { return snapshot(WQ3D_PAST); }
    const wave_front& prev() { return snapshot(WQ3D_PREV); }
    const wave_front& curr() { return snapshot(WQ3D_CURR); }
    const wave_front& next() { return snapshot(WQ3D_NEXT); }

private:
    void build_env(const ocean::ocean_model& o) {
        wq3d_env_desc e{};
        auto grid = [&](const std::shared_ptr<const ocean::data_grid>& g, wq3d_grid& out) {
            out.ndims = (int32_t)g->axes.size();
            for (size_t k = 0; k < g->axes.size(); ++k) {
                out.axis[k] = g->axes[k]->abi();
                out.interp[k] = g->interp[k];
                out.edge_limit[k] = g->edge_limit[k] ? 1 : 0;
            }
            out.data = g->data.data();
        };
        e.profile.kind = o.profile.kind;
        e.profile.flat_earth = o.profile.flat_earth ? 1 : 0;
        for (int k = 0; k < 4; ++k) e.profile.p[k] = o.profile.p[k];
        e.profile.atten_kind = o.profile.attenuation.kind;
        e.profile.atten_coeff = o.profile.attenuation.coeff;
        if (o.profile.grid) grid(o.profile.grid, e.profile.grid);
        auto boundary = [&](const ocean::boundary_model& b, wq3d_boundary_desc& out) {
            out.kind = b.kind;
            out.rho = b.rho;
            out.normal_rho = b.normal_rho;
            out.loss_kind = b.loss.kind;
            for (int k = 0; k < 5; ++k) out.loss[k] = b.loss.params[k];
            if (b.grid) grid(b.grid, out.grid);
        };
        boundary(o.surface, e.surface);
        boundary(o.bottom, e.bottom);
        std::vector<wq3d_volume_desc> vols;
        for (const auto& v : o.volumes)
            vols.push_back(wq3d_volume_desc{types::earth_radius() - std::fabs(v.depth), v.thickness});
        e.num_volume = (int32_t)vols.size();
        e.volume = vols.empty() ? nullptr : vols.data();
        check(wq3d_env_create(&e, &_env));
    }

    void push_listeners() {
        check(wq3d_set_listeners(_fan, _verb_listeners.empty() ? 0 : 1, _refl_listeners.empty() ? 0 : 1));
    }

    void deliver(int64_t n) {
        const size_t nf = _freq->size();
        int64_t cnt = 0;
        check(wq3d_drain_eigenrays(_fan, nullptr, nullptr, nullptr, 0, &cnt));
        std::vector<wq3d_eigenray> rays((size_t)cnt);
        std::vector<double> inten((size_t)cnt * nf), phase((size_t)cnt * nf);
        check(wq3d_drain_eigenrays(_fan, rays.data(), inten.data(), phase.data(), cnt, &cnt));
        check(wq3d_drain_eigenverbs(_fan, nullptr, nullptr, 0, &cnt));
        std::vector<wq3d_eigenverb> verbs((size_t)cnt);
        std::vector<double> power((size_t)cnt * nf);
        check(wq3d_drain_eigenverbs(_fan, verbs.data(), power.data(), cnt, &cnt));
        check(wq3d_drain_reflections(_fan, nullptr, 0, &cnt));
        std::vector<wq3d_reflection> refl((size_t)cnt);
        check(wq3d_drain_reflections(_fan, refl.data(), cnt, &cnt));
        size_t ir = 0, iv = 0, il = 0;
        for (int64_t k = 0; k < n; ++k) {
            _time += _time_step;
            ++_steps;
            // reflection and eigenverb records carry the step count before the
            // queue rotation (detect_reflections), eigenrays the count after it;
            // both record lists arrive sorted by step first
            std::vector<std::tuple<int, int, int, int, size_t>> events;  // de, az, seq, kind, index
            for (; il < refl.size() && refl[il].step == _steps - 1; ++il)
                events.emplace_back(refl[il].de, refl[il].az, refl[il].seq, 0, il);
            for (; iv < verbs.size() && verbs[iv].step == _steps - 1; ++iv)
                events.emplace_back(verbs[iv].de, verbs[iv].az, verbs[iv].seq, 1, iv);
            std::sort(events.begin(), events.end());
            for (const auto& ev : events) {
                const size_t j = std::get<4>(ev);
                if (std::get<3>(ev) == 0) {
                    const wq3d_reflection& r = refl[j];
                    const auto p = types::wposition1::rtp(r.pos_rho, r.pos_theta, r.pos_phi);
                    const auto d = types::wposition1::rtp(r.ndir_rho, r.ndir_theta, r.ndir_phi);
                    for (auto* l : _refl_listeners)
                        l->reflect(r.time, r.de, r.az, r.dt, r.grazing, r.speed, p, d, r.type);
                } else {
                    const wq3d_eigenverb& v = verbs[j];
                    auto m = std::make_shared<eigenverb_model>();
                    m->time = v.travel_time;
                    m->power.assign(power.begin() + j * nf, power.begin() + (j + 1) * nf);
                    m->length = v.length;
                    m->width = v.width;
                    m->position = types::wposition1::rtp(v.pos_rho, v.pos_theta, v.pos_phi);
                    m->direction = v.direction;
                    m->grazing = v.grazing;
                    m->sound_speed = v.sound_speed;
                    m->de_index = (size_t)v.de;
                    m->az_index = (size_t)v.az;
                    m->source_de = v.source_de;
                    m->source_az = v.source_az;
                    m->surface = v.surface;
                    m->bottom = v.bottom;
                    m->caustic = v.caustic;
                    m->upper = v.upper;
                    m->lower = v.lower;
                    for (auto* l : _verb_listeners) l->add_eigenverb(m, v.type);
                }
            }
            for (; ir < rays.size() && rays[ir].step == _steps; ++ir) {
                const wq3d_eigenray& r = rays[ir];
                auto m = std::make_shared<eigenray_model>();
                m->time = r.travel_time;
                m->intensity.assign(inten.begin() + ir * nf, inten.begin() + (ir + 1) * nf);
                m->phase.assign(phase.begin() + ir * nf, phase.begin() + (ir + 1) * nf);
                m->source_de = r.source_de;
                m->source_az = r.source_az;
                m->target_de = r.target_de;
                m->target_az = r.target_az;
                m->surface = r.surface;
                m->bottom = r.bottom;
                m->caustic = r.caustic;
                m->upper = r.upper;
                m->lower = r.lower;
                for (auto* l : _ray_listeners) l->add_eigenray((size_t)r.t1, (size_t)r.t2, m, _run_id);
            }
            for (auto* l : _ray_listeners) l->check_eigenrays(_time, _run_id);
        }
    }

    const wave_front& snapshot(int slot) {
        wave_front& w = _front;
        w.num_de = num_de();
        w.num_az = num_az();
        w.num_freq = _freq->size();
        const size_t n = w.num_de * w.num_az;
        for (auto* v : {&w.pos_rho, &w.pos_theta, &w.pos_phi, &w.ndir_rho, &w.ndir_theta, &w.ndir_phi,
                        &w.sound_speed, &w.distance, &w.path_length})
            v->assign(n, 0.0);
        w.attenuation.assign(n * w.num_freq, 0.0);
        w.phase.assign(n * w.num_freq, 0.0);
        for (auto* v : {&w.surface, &w.bottom, &w.caustic, &w.upper, &w.lower}) v->assign(n, 0);
        w.on_edge.assign(n, 0);
        wq3d_front_view view{};
        view.pos_rho = w.pos_rho.data();
        view.pos_theta = w.pos_theta.data();
        view.pos_phi = w.pos_phi.data();
        view.ndir_rho = w.ndir_rho.data();
        view.ndir_theta = w.ndir_theta.data();
        view.ndir_phi = w.ndir_phi.data();
        view.sound_speed = w.sound_speed.data();
        view.distance = w.distance.data();
        view.path_length = w.path_length.data();
        view.attenuation = w.attenuation.data();
        view.phase = w.phase.data();
        view.surface = w.surface.data();
        view.bottom = w.bottom.data();
        view.caustic = w.caustic.data();
        view.upper = w.upper.data();
        view.lower = w.lower.data();
        view.on_edge = w.on_edge.data();
        check(wq3d_snapshot(_fan, slot, &view));
        return w;
    }

    ocean::ocean_model::csptr _ocean;
    std::shared_ptr<const types::seq_vector> _freq, _de, _az;
    double _time_step;
    const types::wposition* _targets;
    types::wposition1 _source;
    wq3d_env* _env = nullptr;
    wq3d_fan* _fan = nullptr;
    wq3d_thresholds _thr{};
    double _time = 0.0;
    int64_t _steps = 0;
    size_t _run_id = 0;
    std::set<eigenray_listener*> _ray_listeners;
    std::set<eigenverb_listener*> _verb_listeners;
    std::set<reflection_listener*> _refl_listeners;
    wave_front _front;
};

}  // namespace waveq3d

// -------------------------------------------------------------- eigenrays
namespace eigenrays {

// eigenrays/eigenray_collection.h:34-262: an eigenray listener that files
// each eigenray under its target in arrival order (add_eigenray,
// eigenray_collection.cc:90-101) and sums each target's list on the GPU
// (sum_eigenrays, :105-191, through wq3d_sum_eigenrays).
class eigenray_collection : public waveq3d::eigenray_listener {
public:
    eigenray_collection(std::shared_ptr<const types::seq_vector> frequencies, const types::wposition1& source_pos,
                        const types::wposition& target_pos, uint64_t sourceID = 0,
                        std::vector<uint64_t> targetIDs = {}, bool coherent = true)
        : _freq(std::move(frequencies)), _source(source_pos), _n1(target_pos.size1()),
          _n2(target_pos.size2()), _sourceID(sourceID), _targetIDs(std::move(targetIDs)), _coherent(coherent),
          _lists(_n1 * _n2), _initial(_n1 * _n2, 0.0), _total(_n1 * _n2) {
        if (_targetIDs.empty()) _targetIDs.assign(_n1 * _n2, 0);
        for (auto& t : _total) t = empty_total();
    }

    void add_eigenray(size_t t1, size_t t2, waveq3d::eigenray_model::csptr ray, size_t) override {
        const size_t k = t1 * _n2 + t2;
        _arrivals.emplace_back(t1, t2, ray);
        _lists[k].push_back(ray);
        if (_initial[k] <= 0.0 || _initial[k] > ray->time) _initial[k] = ray->time;
    }

    void sum_eigenrays() {
        const size_t nf = _freq->size();
        std::vector<wq3d_eigenray> recs(_arrivals.size());
        std::vector<double> inten(recs.size() * nf), phase(recs.size() * nf);
        for (size_t k = 0; k < _arrivals.size(); ++k) {
            const auto& r = *std::get<2>(_arrivals[k]);
            wq3d_eigenray& e = recs[k];
            e = wq3d_eigenray{};
            e.t1 = (int32_t)std::get<0>(_arrivals[k]);
            e.t2 = (int32_t)std::get<1>(_arrivals[k]);
            e.travel_time = r.time;
            e.source_de = r.source_de;
            e.source_az = r.source_az;
            e.target_de = r.target_de;
            e.target_az = r.target_az;
            e.surface = r.surface;
            e.bottom = r.bottom;
            e.caustic = r.caustic;
            e.upper = r.upper;
            e.lower = r.lower;
            std::copy(r.intensity.begin(), r.intensity.end(), inten.begin() + k * nf);
            std::copy(r.phase.begin(), r.phase.end(), phase.begin() + k * nf);
        }
        const size_t nt = _n1 * _n2;
        std::vector<wq3d_eigenray_total> tot(nt);
        std::vector<double> ti(nt * nf), tp(nt * nf);
        check(wq3d_sum_eigenrays(recs.data(), inten.data(), phase.data(), (int64_t)recs.size(),
                                 _freq->values().data(), (int32_t)nf, (int64_t)_n1, (int64_t)_n2,
                                 _coherent ? 1 : 0, tot.data(), ti.data(), tp.data()));
        for (size_t k = 0; k < nt; ++k) {
            auto m = std::make_shared<waveq3d::eigenray_model>();
            m->time = tot[k].travel_time;
            m->source_de = tot[k].source_de;
            m->source_az = tot[k].source_az;
            m->target_de = tot[k].target_de;
            m->target_az = tot[k].target_az;
            m->surface = tot[k].surface;
            m->bottom = tot[k].bottom;
            m->caustic = tot[k].caustic;
            m->upper = tot[k].upper;
            m->lower = tot[k].lower;
            m->intensity.assign(ti.begin() + k * nf, ti.begin() + (k + 1) * nf);
            m->phase.assign(tp.begin() + k * nf, tp.begin() + (k + 1) * nf);
            _total[k] = m;
        }
    }

    size_t size1() const { return _n1; }
    size_t size2() const { return _n2; }
    uint64_t sourceID() const { return _sourceID; }
    uint64_t targetID(size_t t1, size_t t2) const { return _targetIDs[t1 * _n2 + t2]; }
    bool coherent() const { return _coherent; }
    size_t num_eigenrays() const { return _arrivals.size(); }
    const std::vector<waveq3d::eigenray_model::csptr>& eigenrays(size_t t1, size_t t2) const {
        return _lists[t1 * _n2 + t2];
    }
    double initial_time(size_t t1, size_t t2) const { return _initial[t1 * _n2 + t2]; }
    waveq3d::eigenray_model::csptr total(size_t t1, size_t t2) const { return _total[t1 * _n2 + t2]; }

private:
    waveq3d::eigenray_model::csptr empty_total() const {
        auto m = std::make_shared<waveq3d::eigenray_model>();
        m->intensity.assign(_freq->size(), 0.0);
        m->phase.assign(_freq->size(), 0.0);
        return m;
    }

    std::shared_ptr<const types::seq_vector> _freq;
    types::wposition1 _source;
    size_t _n1, _n2;
    uint64_t _sourceID;
    std::vector<uint64_t> _targetIDs;
    bool _coherent;
    std::vector<std::tuple<size_t, size_t, waveq3d::eigenray_model::csptr>> _arrivals;
    std::vector<std::vector<waveq3d::eigenray_model::csptr>> _lists;
    std::vector<double> _initial;
    std::vector<waveq3d::eigenray_model::csptr> _total;
};

}  // namespace eigenrays
}  // namespace usml_amd

#endif  // USML_AMD_WAVE_QUEUE_HPP
