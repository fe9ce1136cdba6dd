// The reference's eigenray_basic scenario (waveq3d/test/eigenray_test.cc:86-198)
// driven through the C++ wave_queue mirror over libwq3d.so.  Prints one line
// per listener call and the eigenray collection's total for the target;
// tests/test_cpp_facade.py checks them.
#include <cstdio>
#include <exception>

#include "../../underseamodelinglibrary_amd/cpp/usml_wave_queue.hpp"

using namespace usml_amd;

struct printer : waveq3d::eigenray_listener {
    int checks = 0;
    void add_eigenray(size_t t1, size_t t2, waveq3d::eigenray_model::csptr r, size_t runID) override {
        std::printf("eigenray %zu %zu %zu %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d %d %d\n", t1, t2, runID,
                    r->time, r->intensity[0], r->phase[0], r->source_de, r->target_de, r->source_az,
                    r->target_az, r->surface, r->bottom, r->caustic);
    }
    void check_eigenrays(double, size_t) override { ++checks; }
};

int main() {
    try {
        types::compute_earth_radius(45.0);
        auto ocean = std::make_shared<const ocean::ocean_model>(
            ocean::boundary_flat(), ocean::boundary_flat(3000.0),
            ocean::profile_linear(1500.0, ocean::attenuation_constant(0.0)));
        auto freq = std::make_shared<types::seq_log>(10e3, 2.0, (size_t)3);
        types::wposition1 pos(45.0, -45.0, -1000.0);
        auto de = std::make_shared<types::seq_linear>(-60.0, 5.0, 60.0);
        auto az = std::make_shared<types::seq_linear>(-4.0, 1.0, 4.0);
        types::wposition target(1, 1, 45.02, -45.0, -1000.0);
        waveq3d::wave_queue wave(ocean, freq, pos, de, az, 0.1, &target);
        wave.runID(7);
        printer p;
        wave.add_eigenray_listener(&p);
        eigenrays::eigenray_collection collection(freq, pos, target, 11, {42});
        wave.add_eigenray_listener(&collection);
        while (wave.time() < 3.5) wave.step();
        std::printf("checks %d time %.17g\n", p.checks, wave.time());
        // clamped sequence accessors: a negative index is a huge size_t in the
        // reference and clamps to the last element (types/seq_vector.h:98-101)
        std::printf("seq %.17g %.17g %.17g %.17g\n", (*de)(-1), de->increment(-1), (*de)(1000), (*de)(0));
        collection.sum_eigenrays();
        const auto total = collection.total(0, 0);
        std::printf("total %zu %.17g %.17g %.17g %.17g %.17g %d %d %.17g %llu\n", collection.num_eigenrays(),
                    total->time, total->intensity[0], total->intensity[2], total->phase[0], total->source_de,
                    total->surface, total->bottom, collection.initial_time(0, 0),
                    (unsigned long long)collection.targetID(0, 0));
        const auto& curr = wave.curr();
        std::printf("curr %zu %zu %.17g\n", curr.num_de, curr.num_az, curr.pos_rho[curr.index(12, 4)]);
        // the classic ray model (spreading_ray) on the same scenario
        {
            waveq3d::wave_queue classic(ocean, freq, pos, de, az, 0.1, &target, waveq3d::CLASSIC_RAY);
            struct counter : waveq3d::eigenray_listener {
                int n = 0;
                void add_eigenray(size_t, size_t, waveq3d::eigenray_model::csptr, size_t) override { ++n; }
            } c;
            classic.add_eigenray_listener(&c);
            while (classic.time() < 3.5) classic.step();
            std::printf("classic %d\n", c.n);
        }
        // an unsupported model is rejected with std::invalid_argument
        try {
            waveq3d::wave_queue bad(ocean, freq, pos, de, az, 0.1, &target, (waveq3d::spreading_type)7);
            std::printf("unknown accepted\n");
        } catch (const std::invalid_argument& e) {
            std::printf("unknown rejected\n");
        }
    } catch (const std::exception& e) {
        std::printf("error %s\n", e.what());
        return 1;
    }
    return 0;
}
