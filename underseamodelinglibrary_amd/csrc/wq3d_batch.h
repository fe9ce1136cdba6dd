// wq3d_batch.h - many fans stepped together (part of wq3d.hip's translation
// unit; included at its end).
//
// The reference runs independent fans as wavefront_generator tasks on a CPU
// thread pool (wavegen/wavefront_generator.cc:52-127, threads/
// thread_pool.cc:21-75), each task stepping its own wave_queue.  A batch
// steps every fan of a list with one launch per kernel of the step instead
// of one per fan: the fused launches (k_*_multi) give each fan a contiguous
// range of blocks, so small fans fill the device together.  Every 64 steps
// the record sets of all fans rotate, exactly as each fan's own wq3d_step
// would rotate them; between rotations the 64 steps' launches are built
// once per record set into a HIP graph and replayed.  Launch shapes (threads
// per block of the per-ray launches, blocks of the construction launch) are
// picked by wq3d_batch_tune, which times real steps of the batch: the shape
// changes only how work is spread over the device, never a result.
#include <hip/hip_runtime.h>

#include <tuple>

namespace {

constexpr int kBatchGraphSteps = 64;  // the record sets' rotation period

// a device copy of n elements that stays writable (DeviceBuffers::upload
// hands out const pointers)
template <class T>
hipError_t put(DeviceBuffers& mem, T** out, const T* src, size_t n) {
    hipError_t e = mem.alloc(out, n);
    if (e == hipSuccess && n) e = rt::memcpy(*out, src, n * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

bool identity_slots(const wq3d_fan* f) {
    return f->phys[0] == 0 && f->phys[1] == 1 && f->phys[2] == 2 && f->phys[3] == 3;
}

#ifndef WQ3D_EMULATE
// The launches of a step go straight onto the batch's stream or, while a
// graph is being built, into that graph as a chain of kernel nodes.  The
// graphs are built node by node instead of by capturing the stream: while
// a stream of the device is capturing, ROCm refuses legacy-stream work from
// every other host thread (a record drain's copy, say) and the capture is
// lost, so capturing would break batches stepped from several threads.
struct Launcher {
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;     // set: add nodes instead of launching
    hipGraphNode_t last = nullptr;  // the chain's last node
    hipError_t err = hipSuccess;

    template <class... P, class... A>
    void operator()(void (*kernel)(P...), dim3 grid, dim3 block, A... args) {
        if (err != hipSuccess) return;
        if (!graph) {
            hipLaunchKernelGGL(kernel, grid, block, 0, stream, args...);
            return;
        }
        std::tuple<P...> vals(args...);
        std::apply(
            [&](auto&... v) {
                void* ptrs[] = {static_cast<void*>(&v)...};
                hipKernelNodeParams np{};
                np.func = reinterpret_cast<void*>(kernel);
                np.gridDim = grid;
                np.blockDim = block;
                np.sharedMemBytes = 0;
                np.kernelParams = ptrs;
                np.extra = nullptr;
                hipGraphNode_t node = nullptr;
                err = hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &np);
                if (err == hipSuccess) last = node;
            },
            vals);
    }
};
#endif

}  // namespace

struct wq3d_batch {
    std::vector<wq3d_fan*> fans;
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceBuffers mem;
    KParams* kps = nullptr;         // device [2][n]: copy of fan f's parameters for record set r at r * n + f
    FanClock* clk = nullptr;        // device [n]
    struct Group {                  // fans of one environment variant (per-ray launches)
        int variant = 0;
        std::vector<int32_t> ids;
        int32_t* d_ids = nullptr;
        int64_t* d_first = nullptr;
        int64_t blocks = 0;
        int64_t rays = 0;                       // over the group's fans
        unsigned long long* d_blist = nullptr;  // [slots] this step's listed rays, (fan of the group << 32 | ray),
                                                //   in lists per die as k_reflect_light's
        unsigned long long* d_bcnt = nullptr;   // [2][dies][2] the lists' lengths by step parity
        int64_t slots = 0;                      // rays + 256 per fan: the grid's threads at any ray_threads
        std::vector<int64_t> first;             // host copy of d_first
    };
    std::vector<Group> groups;
    std::vector<int32_t> tids;  // fans with targets (the eigenray search)
    int32_t* d_tids = nullptr;
    int64_t *d_sup_first = nullptr, *d_det_first = nullptr;
    int64_t sup_blocks = 0, det_blocks = 0;
    wq3d_batch_shape shape{kBlock, 1024, 1, 0};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    // timing of the search stage (fused supmask + detect launches)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> stage_events, event_pool;
    double stage_ms = 0.0;
    int64_t stage_steps = 0;
};

namespace {

void batch_drop_graphs(wq3d_batch* b) {
#ifndef WQ3D_EMULATE
    for (auto& e : b->exec) {
        if (e) (void)hipGraphExecDestroy(e);
        e = nullptr;
    }
#endif
}

// block ranges of the per-ray launches for the current threads per block
wq3d_status batch_tables(wq3d_batch* b) {
    const int64_t tpb = b->shape.ray_threads;
    for (auto& g : b->groups) {
        std::vector<int64_t> first(g.ids.size() + 1, 0);
        for (size_t k = 0; k < g.ids.size(); ++k)
            first[k + 1] = first[k] + (b->fans[g.ids[k]]->F.N + tpb - 1) / tpb;
        g.blocks = first.back();
        HIP_TRY(rt::memcpy(g.d_first, first.data(), first.size() * sizeof(int64_t), hipMemcpyHostToDevice));
        g.first = first;
    }
    return WQ3D_OK;
}

// one step's launches for record set `rset`, onto the batch's stream or into
// the graph being built (L); the host state (slots, times, record sets) is
// advanced by the caller
#ifndef WQ3D_EMULATE
wq3d_status batch_launch_step(wq3d_batch* b, Launcher& L, const Slots& S, int rset) {
    const int n = (int)b->fans.size();
    const KParams* kps = b->kps + (size_t)rset * n;
    const dim3 tpb((unsigned)b->shape.ray_threads);
    const FanClock* clk = b->clk;
    for (const auto& g : b->groups)
        L(&k_reflect_light_multi, dim3((unsigned)g.blocks), tpb, kps, g.d_ids, g.d_first, (int)g.ids.size(), S, clk,
          g.d_blist, g.d_bcnt);
    for (const auto& g : b->groups) {
        const unsigned long long* blist = g.d_blist;
        const int64_t lb = g.blocks, lt = b->shape.ray_threads;
        const dim3 grid(bounce_grid(lb, lt)), block(64);
        unsigned long long* cnt = g.d_bcnt;
        switch (g.variant) {
        case 1:
            L(&k_bounce_multi<WQ3D_PROFILE_MUNK, true>, grid, block, kps, g.d_ids, S, clk, blist, cnt, lb, lt);
            break;
        case 2:
            L(&k_bounce_multi<WQ3D_PROFILE_LINEAR, true>, grid, block, kps, g.d_ids, S, clk, blist, cnt, lb, lt);
            break;
        case 3:
            L(&k_bounce_multi<kProfileGrid1Linear, true>, grid, block, kps, g.d_ids, S, clk, blist, cnt, lb, lt);
            break;
        default:
            return fail(WQ3D_ERR_UNSUPPORTED, "environment kind without a batch variant");
        }
    }
    // the rotation (past <- prev <- curr <- next <- past)
    const Slots R{S.prev, S.curr, S.next, S.past};
    for (const auto& g : b->groups) {
        const int ng = (int)g.ids.size();
        const dim3 grid((unsigned)g.blocks);
        switch (g.variant) {
        case 1:
            L(&k_advance_multi<WQ3D_PROFILE_MUNK, true>, grid, tpb, kps, g.d_ids, g.d_first, ng, R, b->clk);
            break;
        case 2:
            L(&k_advance_multi<WQ3D_PROFILE_LINEAR, true>, grid, tpb, kps, g.d_ids, g.d_first, ng, R, b->clk);
            break;
        default:
            L(&k_advance_multi<kProfileGrid1Linear, true>, grid, tpb, kps, g.d_ids, g.d_first, ng, R, b->clk);
            break;
        }
    }
    if (!b->tids.empty()) {
        const int nt = (int)b->tids.size();
        std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
        if (b->timing) {  // eager steps only (batch_run builds no graph while timing)
            if (!b->event_pool.empty()) {
                ev = b->event_pool.back();
                b->event_pool.pop_back();
            } else {
                HIP_TRY(rt::event_create(&ev.first));
                HIP_TRY(rt::event_create(&ev.second));
            }
            HIP_TRY(rt::event_record(ev.first, b->stream));
        }
        L(&k_supmask_multi, dim3((unsigned)b->sup_blocks), dim3(kBlock), kps, (const int32_t*)b->d_tids,
          (const int64_t*)b->d_sup_first, nt);
        L(&k_detect_multi, dim3((unsigned)b->det_blocks), dim3(kBlock), kps, (const int32_t*)b->d_tids,
          (const int64_t*)b->d_det_first, nt, R);
        if (b->timing) {
            HIP_TRY(rt::event_record(ev.second, b->stream));
            b->stage_events.push_back(ev);
        }
        L(&k_build_multi, dim3((unsigned)b->shape.build_blocks), dim3(kBlock), kps, (const int32_t*)b->d_tids, nt,
          R, (const FanClock*)b->clk, kBuildLanes);
    }
    if (L.err != hipSuccess) return fail(WQ3D_ERR_DEVICE, std::string("batch graph node: ") + hipGetErrorString(L.err));
    if (!L.graph) HIP_TRY(rt::last_error());
    return WQ3D_OK;
}
#else
// The reflection launches of one group as the device runs them: every
// thread of k_reflect_light_multi's grid, then every thread of
// k_bounce_multi's, with the group's die lists in its device list array
// (their slots checked against the grid) and the fans' host clocks.
void emu_group_reflections(wq3d_batch* b, const wq3d_batch::Group& g, const Slots& S, int rset) {
    const int64_t tpb = b->shape.ray_threads, blocks = g.blocks;
    emu_range(g.d_blist, g.slots, "group reflection list");
    unsigned long long cnt[2 * kXcds] = {};
    auto cap_of = [&](int x) { return die_blocks(blocks, x) * tpb; };
    auto slot = [&](int x, int64_t k) {
        const int64_t at = die_slot(x, k, tpb);
        if (k < 0 || k >= cap_of(x) || at < 0 || at >= blocks * tpb || at >= g.slots) {
            std::fprintf(stderr, "wq3d emulation check failed: reflection list slot out of range\n");
            std::abort();
        }
        return at;
    };
    for (int64_t w = 0; w < blocks; ++w) {
        const int gi = (int)(std::upper_bound(g.first.begin(), g.first.end(), w) - g.first.begin()) - 1;
        const KParams* P = b->fans[g.ids[gi]]->kp + rset;
        emu::check_range(P, sizeof(KParams), "KParams");
        const int x = (int)(w % kXcds);
        for (int64_t t = 0; t < tpb; ++t) {
            const int64_t i = (w - g.first[gi]) * tpb + t;
            if (i == 0) {
                emu_check_launch(P->E, P->F, P->A, &P->R);
                *P->R.cand_count = 0ull;
                *P->R.hit_count = 0ull;
                *P->R.pass_count = 0ull;
            }
            if (i >= P->F.N) continue;
            body_reflect_light(i, P->F, P->A, S);
            if (!bounce_wanted(P->E, P->F, P->A, i)) continue;
            const unsigned long long e = ((unsigned long long)gi << 32) | (unsigned long long)i;
            if (bounce_side(P->A, i) == 0)
                g.d_blist[slot(x, (int64_t)cnt[2 * x]++)] = e;
            else
                g.d_blist[slot(x, cap_of(x) - 1 - (int64_t)cnt[2 * x + 1]++)] = e;
        }
    }
    const int64_t bthreads = (int64_t)bounce_grid(blocks, tpb) * 64;
    for (int64_t gid = 0; gid < bthreads; ++gid) {
        const int x = (int)((gid / 64) % kXcds);
        const unsigned long long* c = cnt + 2 * x;
        const int64_t j = (gid / 64 / kXcds) * 64 + gid % 64;
        if (j >= (int64_t)(c[0] + c[1])) continue;
        const unsigned long long e = g.d_blist[slot(x, die_entry(j, c, cap_of(x)))];
        const int gi = (int)(e >> 32);
        if (gi < 0 || gi >= (int)g.ids.size()) std::abort();
        wq3d_fan* f = b->fans[g.ids[gi]];
        const KParams* P = f->kp + rset;
        body_bounce((int64_t)(e & 0xffffffffull), P->E, P->F, P->A, S, P->R, f->time, f->steps);
    }
}

// the emulation build runs each group's reflections as above, then steps
// the fans one after another through the single-fan launch paths, with the
// same slots, times and record sets
wq3d_status batch_launch_step(wq3d_batch* b, const Slots& S, int rset) {
    const Slots R{S.prev, S.curr, S.next, S.past};
    for (const auto& g : b->groups) emu_group_reflections(b, g, S, rset);
    for (wq3d_fan* f : b->fans) {
        const unsigned grid = (unsigned)((f->F.N + kBlock - 1) / kBlock);
        launch_advance(b->stream, grid, f->env->dev, f->F, f->A, R);
        if (f->F.nt > 0 && f->F.has_spreading) {
            const int64_t tiles = ((f->F.nde + kTile - 1) / kTile) * ((f->F.naz + kTile - 1) / kTile);
            launch_detect(b->stream, (tiles + 3) / 4, f->F, f->A, R, f->R);
            double t = f->time + f->F.dt;
            launch_build(b->stream, f->kp + rset, f->env->dev, f->F, f->A, R, f->R, t, f->steps + 1);
        }
    }
    return WQ3D_OK;
}
#endif

// the host side of one step for every fan: the rotation, time and step
// count as wq3d_step keeps them, and the record sets every 64 steps
wq3d_status batch_host_step(wq3d_batch* b) {
    for (wq3d_fan* f : b->fans) {
        const int save = f->phys[0];
        f->phys[0] = f->phys[1];
        f->phys[1] = f->phys[2];
        f->phys[2] = f->phys[3];
        f->phys[3] = save;
        f->time += f->F.dt;
        f->steps += 1;
        f->pending_steps += 1;
    }
    if (b->fans[0]->pending_steps >= kBatchGraphSteps) {
        for (wq3d_fan* f : b->fans) {
            wq3d_status st = rotate_record_sets(f, b->stream);
            if (st) return st;
        }
    }
    return WQ3D_OK;
}

wq3d_status batch_eager_step(wq3d_batch* b) {
    wq3d_fan* f0 = b->fans[0];
#ifndef WQ3D_EMULATE
    Launcher L;
    L.stream = b->stream;
    wq3d_status st = batch_launch_step(b, L, slots_of(f0), f0->rset);
#else
    wq3d_status st = batch_launch_step(b, slots_of(f0), f0->rset);
#endif
    if (st) return st;
#ifdef WQ3D_EMULATE
    // the emulated launches read the host clocks directly
#endif
    return batch_host_step(b);
}

#ifndef WQ3D_EMULATE
// 64 steps from the identity slots for record set `rset`, built once
wq3d_status batch_graph(wq3d_batch* b, int rset, hipGraphExec_t* out) {
    if (b->exec[rset]) {
        *out = b->exec[rset];
        return WQ3D_OK;
    }
    hipGraph_t graph = nullptr;
    HIP_TRY(hipGraphCreate(&graph, 0));
    Launcher L;
    L.stream = b->stream;
    L.graph = graph;
    int phys[4] = {0, 1, 2, 3};
    wq3d_status st = WQ3D_OK;
    for (int k = 0; k < kBatchGraphSteps && !st; ++k) {
        st = batch_launch_step(b, L, Slots{phys[0], phys[1], phys[2], phys[3]}, rset);
        const int save = phys[0];
        phys[0] = phys[1];
        phys[1] = phys[2];
        phys[2] = phys[3];
        phys[3] = save;
    }
    if (st) {
        (void)hipGraphDestroy(graph);
        return st;
    }
    const hipError_t ei = hipGraphInstantiate(&b->exec[rset], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(WQ3D_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ei));
    *out = b->exec[rset];
    return WQ3D_OK;
}
#endif

wq3d_status batch_check_aligned(const wq3d_batch* b) {
    const wq3d_fan* f0 = b->fans[0];
    for (const wq3d_fan* f : b->fans) {
        if (f->steps != f0->steps || f->pending_steps != f0->pending_steps || f->rset != f0->rset ||
            f->phys[0] != f0->phys[0] || f->phys[1] != f0->phys[1] || f->phys[2] != f0->phys[2])
            return fail(WQ3D_ERR_STATE, "the batch's fans are no longer at the same step (stepped alone?)");
    }
    return WQ3D_OK;
}

// the fans' parameter copies (each fan's own and the batch's) and clocks as
// the host holds them, before a batch call's first launch (the stream is
// idle between calls)
wq3d_status batch_prepare(wq3d_batch* b) {
    wq3d_status st = batch_check_aligned(b);
    if (st) return st;
    const size_t n = b->fans.size();
    std::vector<FanClock> clk(n);
    std::vector<KParams> hk(2 * n);
    for (size_t k = 0; k < n; ++k) {
        wq3d_fan* f = b->fans[k];
        if (f->kp_dirty) {
            HIP_TRY(rt::sync(f->stream));
            const int active = f->rset;
            for (int r = 0; r < 2; ++r) {
                use_record_set(f, r);
                const KParams kp{f->env->dev, f->F, f->A, f->R};
                HIP_TRY(rt::memcpy(f->kp + r, &kp, sizeof(KParams), hipMemcpyHostToDevice));
            }
            use_record_set(f, active);
            f->kp_dirty = false;
        }
        clk[k] = FanClock{f->time, f->steps, f->F.dt, 0};
        const int active = f->rset;
        for (int r = 0; r < 2; ++r) {
            use_record_set(f, r);
            hk[(size_t)r * n + k] = KParams{f->env->dev, f->F, f->A, f->R};
        }
        use_record_set(f, active);
    }
    HIP_TRY(rt::memcpy(b->kps, hk.data(), hk.size() * sizeof(KParams), hipMemcpyHostToDevice));
    HIP_TRY(rt::memcpy(b->clk, clk.data(), n * sizeof(FanClock), hipMemcpyHostToDevice));
    // the reflection lists' counts: the clocks' parity restarts from the
    // fans' step counts, which may have moved on without the batch
    for (auto& g : b->groups)
        if (g.d_bcnt) HIP_TRY(rt::memset_async(g.d_bcnt, 0, kListCounts * sizeof(unsigned long long), b->stream));
    return WQ3D_OK;
}

wq3d_status batch_run(wq3d_batch* b, int64_t n) {
    int64_t done = 0;
    while (done < n) {
        wq3d_fan* f0 = b->fans[0];
#ifndef WQ3D_EMULATE
        if (b->shape.use_graph && !b->timing && n - done >= kBatchGraphSteps && f0->pending_steps == 0 && identity_slots(f0)) {
            hipGraphExec_t exec = nullptr;
            wq3d_status st = batch_graph(b, f0->rset, &exec);
            if (st) return st;
            HIP_TRY(hipGraphLaunch(exec, b->stream));
            for (int k = 0; k < kBatchGraphSteps; ++k) {
                st = batch_host_step(b);
                if (st) return st;
            }
            done += kBatchGraphSteps;
            continue;
        }
#endif
        wq3d_status st = batch_eager_step(b);
        if (st) return st;
        done += 1;
    }
    return WQ3D_OK;
}

}  // namespace

extern "C" {

wq3d_status wq3d_batch_create(wq3d_fan* const* fans, int32_t num_fans, wq3d_batch** out) {
    if (!fans || !out || num_fans < 1) return fail(WQ3D_ERR_ARGUMENT, "no fans");
    *out = nullptr;
    if (num_fans > kMaxBatchFans) return fail(WQ3D_ERR_UNSUPPORTED, "more than 1024 fans in a batch");
    auto* b = new wq3d_batch();
    for (int32_t k = 0; k < num_fans; ++k) {
        wq3d_fan* f = fans[k];
        if (!f) {
            delete b;
            return fail(WQ3D_ERR_ARGUMENT, "null fan");
        }
        for (wq3d_fan* g : b->fans)
            if (g == f) {
                delete b;
                return fail(WQ3D_ERR_ARGUMENT, "a fan appears twice in the batch");
            }
        if (f->device != fans[0]->device) {
            delete b;
            return fail(WQ3D_ERR_ARGUMENT, "the batch's fans are on different devices");
        }
        if (f->timing) {
            delete b;
            return fail(WQ3D_ERR_UNSUPPORTED, "per-kernel timing is not available in a batch");
        }
        b->fans.push_back(f);
    }
    wq3d_status st = batch_check_aligned(b);
    if (st) {
        delete b;
        return st;
    }
    b->device = fans[0]->device;
    auto bail = [&](wq3d_status s) {
        if (b->stream) (void)rt::stream_destroy(b->stream);
        delete b;
        return s;
    };
    if (rt::set_device(b->device) != hipSuccess || rt::stream_create(&b->stream) != hipSuccess)
        return bail(fail(WQ3D_ERR_DEVICE, "batch stream creation failed"));
    const int n = num_fans;
    if (b->mem.alloc(&b->kps, (size_t)2 * n) != hipSuccess || b->mem.alloc(&b->clk, (size_t)n) != hipSuccess)
        return bail(fail(WQ3D_ERR_DEVICE, "batch table upload failed"));
    // per-ray launches: fans grouped by environment variant
    for (int k = 0; k < n; ++k) {
        int v = 1;
#ifndef WQ3D_EMULATE
        v = env_variant(b->fans[k]->env->dev);
        if (v == 0) return bail(fail(WQ3D_ERR_UNSUPPORTED, "environment kind without a batch variant"));
#endif
        auto it = std::find_if(b->groups.begin(), b->groups.end(), [&](const wq3d_batch::Group& g) { return g.variant == v; });
        if (it == b->groups.end()) {
            b->groups.push_back(wq3d_batch::Group{});
            b->groups.back().variant = v;
            it = b->groups.end() - 1;
        }
        it->ids.push_back(k);
    }
    for (auto& g : b->groups) {
        for (int32_t k : g.ids) g.rays += b->fans[k]->F.N;
        g.slots = g.rays + 256 * (int64_t)g.ids.size();
        if (put(b->mem, &g.d_ids, g.ids.data(), g.ids.size()) != hipSuccess ||
            b->mem.alloc(&g.d_first, g.ids.size() + 1) != hipSuccess)
            return bail(fail(WQ3D_ERR_DEVICE, "batch table upload failed"));
        if (b->mem.alloc(&g.d_blist, (size_t)g.slots) != hipSuccess ||
            b->mem.alloc(&g.d_bcnt, kListCounts) != hipSuccess)
            return bail(fail(WQ3D_ERR_DEVICE, "batch reflection list allocation failed"));
    }
    // the eigenray search: the fans with targets, each a whole grid of its own
    std::vector<int64_t> sup{0}, det{0};
    for (int k = 0; k < n; ++k) {
        const FanConst& F = b->fans[k]->F;
        if (!(F.nt > 0 && F.has_spreading)) continue;
        b->tids.push_back(k);
        const int64_t tiles = ((F.nde + kTile - 1) / kTile) * ((F.naz + kTile - 1) / kTile);
        const int64_t quads = (tiles + 3) / 4;
        sup.push_back(sup.back() + (tiles + kBlock / 64 - 1) / (kBlock / 64));
        det.push_back(det.back() + kXcds * ((quads + kXcds - 1) / kXcds) * F.nsup);
    }
    b->sup_blocks = sup.back();
    b->det_blocks = det.back();
    if (!b->tids.empty() &&
        (put(b->mem, &b->d_tids, b->tids.data(), b->tids.size()) != hipSuccess ||
         put(b->mem, &b->d_sup_first, sup.data(), sup.size()) != hipSuccess ||
         put(b->mem, &b->d_det_first, det.data(), det.size()) != hipSuccess))
        return bail(fail(WQ3D_ERR_DEVICE, "batch table upload failed"));
    st = batch_tables(b);
    if (st) return bail(st);
    for (auto& e : b->ev)
        if (rt::event_create(&e) != hipSuccess) return bail(fail(WQ3D_ERR_DEVICE, "event creation failed"));
    *out = b;
    return WQ3D_OK;
}

wq3d_status wq3d_batch_destroy(wq3d_batch* b) {
    if (!b) return WQ3D_OK;
    (void)rt::set_device(b->device);
    (void)rt::sync(b->stream);
    batch_drop_graphs(b);
    for (auto* v : {&b->stage_events, &b->event_pool})
        for (auto& e : *v) {
            (void)rt::event_destroy(e.first);
            (void)rt::event_destroy(e.second);
        }
    for (auto& e : b->ev)
        if (e) (void)rt::event_destroy(e);
    (void)rt::stream_destroy(b->stream);
    delete b;
    return WQ3D_OK;
}

wq3d_status wq3d_batch_step(wq3d_batch* b, int64_t n) {
    if (!b) return fail(WQ3D_ERR_ARGUMENT, "null batch");
    if (n < 0) return fail(WQ3D_ERR_ARGUMENT, "negative step count");
    HIP_TRY(rt::set_device(b->device));
    wq3d_status st = batch_prepare(b);
    if (st) return st;
    st = batch_run(b, n);
    if (st) return st;
    HIP_TRY(rt::sync(b->stream));
    return WQ3D_OK;
}

wq3d_status wq3d_batch_set_timing(wq3d_batch* b, int32_t enable) {
    if (!b) return fail(WQ3D_ERR_ARGUMENT, "null batch");
    b->timing = enable != 0;
    return WQ3D_OK;
}

wq3d_status wq3d_batch_stage_times(wq3d_batch* b, double* ms_per_step, int64_t* steps) {
    if (!b || !ms_per_step || !steps) return fail(WQ3D_ERR_ARGUMENT, "null argument");
    HIP_TRY(rt::set_device(b->device));
    HIP_TRY(rt::sync(b->stream));
    for (auto& e : b->stage_events) {
        float t = 0.0f;
        HIP_TRY(rt::event_elapsed(&t, e.first, e.second));
        b->stage_ms += t;
        b->stage_steps += 1;
        b->event_pool.push_back(e);
    }
    b->stage_events.clear();
    *steps = b->stage_steps;
    *ms_per_step = b->stage_steps ? b->stage_ms / (double)b->stage_steps : 0.0;
    b->stage_ms = 0.0;
    b->stage_steps = 0;
    return WQ3D_OK;
}

wq3d_status wq3d_batch_get_shape(const wq3d_batch* b, wq3d_batch_shape* shape) {
    if (!b || !shape) return fail(WQ3D_ERR_ARGUMENT, "null argument");
    *shape = b->shape;
    return WQ3D_OK;
}

wq3d_status wq3d_batch_set_shape(wq3d_batch* b, const wq3d_batch_shape* shape) {
    if (!b || !shape) return fail(WQ3D_ERR_ARGUMENT, "null argument");
    const int t = shape->ray_threads;
    if (!(t == 64 || t == 128 || t == 256)) return fail(WQ3D_ERR_ARGUMENT, "ray_threads must be 64, 128 or 256");
    if (shape->build_blocks < 1 || shape->build_blocks > (1 << 16))
        return fail(WQ3D_ERR_ARGUMENT, "build_blocks must be in 1 .. 65536");
    HIP_TRY(rt::set_device(b->device));
    HIP_TRY(rt::sync(b->stream));
    batch_drop_graphs(b);
    b->shape = *shape;
    b->shape.use_graph = shape->use_graph ? 1 : 0;
    b->shape.pad_ = 0;
    return batch_tables(b);
}

// The autotuner: the time per step of each launch shape over `trial` real
// steps of the batch (eager launches, HIP events on the batch's stream),
// first over threads per block of the per-ray launches, then over blocks of
// the construction launch; the fastest shape is kept.  Each trial advances
// the batch by `trial` steps - the results of every step are the same
// whatever the shape.
wq3d_status wq3d_batch_tune(wq3d_batch* b, int32_t trial, double* best_ms_per_step, int64_t* steps_taken) {
    if (!b || trial < 1) return fail(WQ3D_ERR_ARGUMENT, "null batch or trial length below 1");
    int64_t taken = 0;
    if (steps_taken) *steps_taken = 0;
    HIP_TRY(rt::set_device(b->device));
    wq3d_status st = batch_prepare(b);
    if (st) return st;
    const int use_graph = b->shape.use_graph;
    auto time_shape = [&](int threads, int blocks, double& ms) -> wq3d_status {
        wq3d_batch_shape s{threads, blocks, 0, 0};
        wq3d_status e = wq3d_batch_set_shape(b, &s);
        if (e) return e;
        // one untimed step settles the new shape
        e = batch_eager_step(b);
        if (e) return e;
        taken += 1;
        if (steps_taken) *steps_taken = taken;
        HIP_TRY(rt::event_record(b->ev[0], b->stream));
        for (int k = 0; k < trial; ++k) {
            e = batch_eager_step(b);
            if (e) return e;
            taken += 1;
            if (steps_taken) *steps_taken = taken;
        }
        HIP_TRY(rt::event_record(b->ev[1], b->stream));
        HIP_TRY(rt::event_sync(b->ev[1]));
        float t = 0.0f;
        HIP_TRY(rt::event_elapsed(&t, b->ev[0], b->ev[1]));
        ms = (double)t / trial;
        return WQ3D_OK;
    };
    int best_threads = b->shape.ray_threads, best_blocks = b->shape.build_blocks;
    double best = 1e300;
    for (int threads : {64, 128, 256}) {
        double ms = 0.0;
        st = time_shape(threads, best_blocks, ms);
        if (st) return st;
        if (ms < best) {
            best = ms;
            best_threads = threads;
        }
    }
    if (!b->tids.empty()) {
        for (int blocks : {256, 512, 1024, 2048, 4096}) {
            double ms = 0.0;
            st = time_shape(best_threads, blocks, ms);
            if (st) return st;
            if (ms < best) {
                best = ms;
                best_blocks = blocks;
            }
        }
    }
    wq3d_batch_shape s{best_threads, best_blocks, use_graph, 0};
    st = wq3d_batch_set_shape(b, &s);
    if (st) return st;
    HIP_TRY(rt::sync(b->stream));
    if (best_ms_per_step) *best_ms_per_step = best;
    return WQ3D_OK;
}

}  // extern "C"
