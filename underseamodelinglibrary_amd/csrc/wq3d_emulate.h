// wq3d_emulate.h - checks of the CPU emulation build (WQ3D_EMULATE) only.
//
// The emulation build runs the kernel bodies on the host so that tests can
// check them with the host's sanitizers.  Two things it adds here:
//
//  * An allocation registry.  Every "device" allocation of the emulated
//    runtime is recorded, and before each launch the driver checks that every
//    pointer a kernel will dereference lies inside one of them with the extent
//    the kernel assumes.  A host pointer (a std::vector's buffer, a caller's
//    array) handed to a kernel by mistake reads fine on the host but faults
//    on the GPU; here it aborts with the field's name.
//
//  * Lane groups run in lockstep.  The GPU builds each eigenray candidate on
//    a group of G lanes that exchange values by cross-lane shuffles and an LDS
//    area.  The emulation runs the G lanes as G host threads; every group
//    operation (any, broadcast, shuffle, the LDS ordering point) is a barrier
//    at which all G lanes must arrive from the same call site of the same
//    candidate.  Lanes that diverge - one lane skipping a group operation
//    that another makes - abort with both call sites instead of silently
//    exchanging garbage, and an LDS read that is not ordered after the write
//    it needs shows up as a data race under ThreadSanitizer.
#pragma once

#if defined(WQ3D_EMULATE) && !defined(__HIP_DEVICE_COMPILE__)

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace wq3d {
namespace emu {

// ------------------------------------------------------------ registry
// A plain array rather than a std::map: the map's tree operations live in
// the (uninstrumented) C++ runtime library, which MemorySanitizer builds of
// the emulation cannot see into.
struct Block {
    uintptr_t start;
    size_t bytes;
};

struct Registry {
    std::mutex m;
    std::vector<Block> blocks;
};

inline Registry& registry() {
    static Registry r;
    return r;
}

inline void note_alloc(const void* p, size_t bytes) {
    std::lock_guard<std::mutex> g(registry().m);
    registry().blocks.push_back(Block{(uintptr_t)p, bytes});
}

inline void note_free(const void* p) {
    std::lock_guard<std::mutex> g(registry().m);
    auto& b = registry().blocks;
    for (size_t k = 0; k < b.size(); ++k) {
        if (b[k].start == (uintptr_t)p) {
            b[k] = b.back();
            b.pop_back();
            return;
        }
    }
}

// [p, p + bytes) must lie inside one registered allocation
inline void check_range(const void* p, size_t bytes, const char* what) {
    if (bytes == 0) return;
    std::lock_guard<std::mutex> g(registry().m);
    const uintptr_t a = (uintptr_t)p;
    for (const Block& k : registry().blocks)
        if (p != nullptr && a >= k.start && a + bytes <= k.start + k.bytes) return;
    std::fprintf(stderr, "wq3d emulation check failed: %s (%p, %zu bytes) is not inside a device allocation\n",
                 what, p, bytes);
    std::abort();
}

// -------------------------------------------------------- lane groups
struct Group {
    int G = 1;
    std::mutex m;
    std::condition_variable cv;
    int arrived = 0;
    uint64_t generation = 0;
    long long site[64] = {};
    long long cand[64] = {};
    unsigned long long u[64] = {};
    double d[64] = {};

    // every lane stores its value, then waits for the others; all lanes
    // must come from the same call site of the same candidate
    void barrier(int lane, long long where, long long tag) {
        std::unique_lock<std::mutex> lk(m);
        site[lane] = where;
        cand[lane] = tag;
        const uint64_t gen = generation;
        if (++arrived == G) {
            for (int k = 1; k < G; ++k) {
                if (site[k] != site[0] || cand[k] != cand[0]) {
                    std::fprintf(stderr,
                                 "wq3d emulation check failed: lanes of a group diverged "
                                 "(lane 0 at line %lld of candidate %lld, lane %d at line %lld of "
                                 "candidate %lld)\n", site[0], cand[0], k, site[k], cand[k]);
                    std::abort();
                }
            }
            arrived = 0;
            ++generation;
            cv.notify_all();
            return;
        }
        if (!cv.wait_for(lk, std::chrono::seconds(60), [&] { return generation != gen; })) {
            std::fprintf(stderr, "wq3d emulation check failed: lane %d waits alone at line %lld "
                         "(candidate %lld): the group's lanes diverged\n", lane, where, tag);
            std::abort();
        }
    }
};

struct LaneState {
    Group* group = nullptr;
    int lane = 0;
    long long tag = 0;  // the candidate this lane is building
};

inline LaneState& lane_state() {
    static thread_local LaneState s;
    return s;
}

inline bool in_group() { return lane_state().group != nullptr; }

inline bool group_any(bool p, long long where) {
    LaneState& s = lane_state();
    Group& g = *s.group;
    g.u[s.lane] = p ? 1ull : 0ull;
    g.barrier(s.lane, where, s.tag);
    bool r = false;
    for (int k = 0; k < g.G; ++k) r = r || g.u[k] != 0ull;
    g.barrier(s.lane, where, s.tag);
    return r;
}

inline double group_shfl(double v, int src, long long where) {
    LaneState& s = lane_state();
    Group& g = *s.group;
    if (src < 0 || src >= g.G) {
        std::fprintf(stderr, "wq3d emulation check failed: shuffle from lane %d of a group of %d\n", src, g.G);
        std::abort();
    }
    g.d[s.lane] = v;
    g.barrier(s.lane, where, s.tag);
    const double r = g.d[src];
    g.barrier(s.lane, where, s.tag);
    return r;
}

inline unsigned long long group_bcast(unsigned long long v, long long where) {
    LaneState& s = lane_state();
    Group& g = *s.group;
    g.u[s.lane] = v;
    g.barrier(s.lane, where, s.tag);
    const unsigned long long r = g.u[0];
    g.barrier(s.lane, where, s.tag);
    return r;
}

inline void wave_sync(long long where) {
    LaneState& s = lane_state();
    s.group->barrier(s.lane, where, s.tag);
}

}  // namespace emu
}  // namespace wq3d

#endif
