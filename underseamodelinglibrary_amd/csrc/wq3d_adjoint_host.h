// wq3d_adjoint_host.h - checkpoints and the backward pass of
// wq3d_adjoint_* (part of wq3d.hip's translation unit; included at its end).
//
// A fan being differentiated keeps a copy of its whole state every time the
// caller asks (wq3d_adjoint_checkpoint, every K steps of the forward run).
// wq3d_adjoint_gradient then walks the segments between checkpoints from
// the last to the first.  For each segment it
//   1. restores the checkpoint into a replay copy of the fan's arrays and
//      steps it forward with the ordinary reflection and advance kernels (no
//      eigenray search: the caller's eigenrays say where the seeds are),
//      keeping every step's four slots in a history;
//   2. builds the seeds of all the segment's eigenrays at once, each from
//      its step's history, as entries (ray, step, field, value);
//   3. sorts the entries by ray and step (radix sort);
//   4. sweeps every ray backward through the segment in one launch, one
//      thread per ray: entries, step adjoint, reflection adjoint
//      (wq3d_adjoint.h) - the rays are independent once the seeds exist.
// The fan's own state is not touched, so it may go on stepping afterwards.
#ifndef WQ3D_EMULATE
#include <hipcub/hipcub.hpp>
#endif

#include <memory>

#include "wq3d_adjoint.h"

namespace {

using wq3d::adj::AdjDev;
using wq3d::adj::AdjSeed;

constexpr int32_t kTapeSmall = 2048;    // operations of a reflection (chain) or start-up tape
constexpr int32_t kTapeSeed = 16384;    // operations of an eigenray's tape
constexpr int kAdjBlock = 256;

#ifndef WQ3D_EMULATE
constexpr size_t kTapePoolBytes = size_t(4) << 30;
constexpr int64_t kEntriesPerSegment = int64_t(1) << 25;
#else
constexpr size_t kTapePoolBytes = size_t(16) << 20;
constexpr int64_t kEntriesPerSegment = int64_t(1) << 22;
#endif

struct AdjCheckpoint {
    int64_t steps = 0;
    double time = 0.0;
    int phys[4] = {0, 1, 2, 3};
    Arrays A{};
};

}  // namespace

// A backward pass's device workspace (replay arrays, history, tapes,
// entries).  Workspaces go back to a per-process pool when a fan releases
// its adjoint and are reused by the next fan of the same shape, so the
// backward passes of many fans allocate and free nothing after the first
// (hipFree would synchronise the device and stop passes from overlapping).
struct AdjWorkspace {
    DeviceBuffers mem;
    int device = 0;
    int64_t N = 0, nf = 0, n_nodes = 0, calls = 0;
    Arrays W{};                 // replay arrays
    KParams* kpw = nullptr;     // {E, F without records, W, Rw}
    Records Rw{};
    AdjDev D{};
    Slots* slots = nullptr;     // [calls] slots after each replayed step
    double* times = nullptr;    // [calls] time after each replayed step
    unsigned long long* skey = nullptr;  // sorted entries
    double* sval = nullptr;
    int64_t *rbeg = nullptr, *rend = nullptr;  // [N] each ray's range of sorted entries
    void* sort_tmp = nullptr;
    size_t sort_bytes = 0;
};

struct FanAdjoint {
    DeviceBuffers cp_mem;  // checkpoints
    std::vector<AdjCheckpoint> cps;
    std::unique_ptr<AdjWorkspace> ws;
    AdjSeed* seeds = nullptr;
    double *seed_int = nullptr, *seed_grad = nullptr;
    int64_t seed_cap = 0;
    DeviceBuffers seed_mem;
    int64_t n_nodes = 0;
};

namespace {

constexpr size_t kPooledWorkspaces = 8;
std::mutex& ws_pool_mutex() {
    static std::mutex m;
    return m;
}
std::vector<std::unique_ptr<AdjWorkspace>>& ws_pool() {
    static std::vector<std::unique_ptr<AdjWorkspace>> p;
    return p;
}

}  // namespace

namespace {

int64_t tapes_of(int32_t cap) {
    return (int64_t)(kTapePoolBytes / (sizeof(wq3d::ad::Node) + sizeof(double)) / (size_t)cap) & ~int64_t(63);
}

#ifndef WQ3D_EMULATE
__global__ void __launch_bounds__(kAdjBlock) k_adj_capture(FanConst F, Arrays A, AdjDev D, int64_t h) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F.N) wq3d::adj::body_adj_capture(i, F, A, D, h);
}

__global__ void __launch_bounds__(64) k_adj_seed(DEnv E, FanConst F, const Slots* slots, const double* times,
                                                 const AdjSeed* seeds, const double* seed_int,
                                                 const double* seed_grad, int64_t n, AdjDev D) {
    const int64_t tk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = tk; e < n; e += nth)
        wq3d::adj::body_adj_seed(e, tk, E, F, slots, times, seeds, seed_int, seed_grad, D);
}

// each ray's range in the sorted entries
__global__ void __launch_bounds__(kAdjBlock) k_adj_ranges(const unsigned long long* key, int64_t n, int64_t* beg,
                                                          int64_t* end) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int shift = 8 + wq3d::adj::kEntryStepBits;
    const int64_t r = (int64_t)(key[k] >> shift);
    if (k == 0 || (int64_t)(key[k - 1] >> shift) != r) beg[r] = k;
    if (k == n - 1 || (int64_t)(key[k + 1] >> shift) != r) end[r] = k + 1;
}

// one ray per thread through the segment; the node adjoints are summed per
// block in LDS first
__global__ void __launch_bounds__(kAdjBlock) k_adj_sweep(DEnv E, FanConst F, const Slots* slots, int64_t K,
                                                         const unsigned long long* key, const double* val,
                                                         const int64_t* beg, const int64_t* end, AdjDev D,
                                                         int32_t n_nodes) {
    extern __shared__ double s_node[];
    for (int k = threadIdx.x; k < n_nodes; k += blockDim.x) s_node[k] = 0.0;
    __syncthreads();
    const int64_t tk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    auto add = [&](int k0, double d0, double d1) {
        if (d0 != 0.0) atomicAdd(&s_node[k0], d0);
        if (d1 != 0.0) atomicAdd(&s_node[k0 + 1], d1);
    };
    for (int64_t i = tk; i < F.N; i += nth)
        wq3d::adj::body_adj_sweep(i, tk, E, F, slots, K, key, val, beg[i], end[i], D, add);
    __syncthreads();
    for (int k = threadIdx.x; k < n_nodes; k += blockDim.x)
        if (s_node[k] != 0.0) atomicAdd(D.node_grad + k, s_node[k]);
}

__global__ void __launch_bounds__(64) k_adj_init(DEnv E, FanConst F, Slots S, AdjDev D) {
    const int64_t tk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tk; i < F.N; i += nth) wq3d::adj::body_adj_init(i, tk, E, F, S, D);
}
#endif

bool adjoint_supported(const DEnv& E) {
    return is_grid1_linear(E.profile) && E.surface.kind == WQ3D_BOUNDARY_FLAT &&
           E.bottom.kind == WQ3D_BOUNDARY_FLAT;
}

// every state field of a fan's four slots, for checkpoint copies
template <class Fn>
void for_each_field(const FanConst& F, Arrays& dst, const Arrays& src, Fn fn) {
    const int64_t n4 = 4 * F.N;
    double* Arrays::*df[] = {&Arrays::pr, &Arrays::pt, &Arrays::pp, &Arrays::gr, &Arrays::gt, &Arrays::gp,
                             &Arrays::nr, &Arrays::nt, &Arrays::np, &Arrays::hr, &Arrays::ht, &Arrays::hp,
                             &Arrays::c,  &Arrays::cr, &Arrays::ct, &Arrays::cp, &Arrays::dist, &Arrays::path,
                             &Arrays::tr, &Arrays::tt, &Arrays::tp, &Arrays::ts};
    for (auto m : df) fn((void**)&(dst.*m), (const void*)(src.*m), n4 * sizeof(double));
    fn((void**)&dst.atten, (const void*)src.atten, n4 * F.nf * sizeof(double));
    fn((void**)&dst.phase, (const void*)src.phase, n4 * F.nf * sizeof(double));
    int32_t* Arrays::*iff[] = {&Arrays::srf, &Arrays::btm, &Arrays::cst, &Arrays::upr, &Arrays::lwr};
    for (auto m : iff) fn((void**)&(dst.*m), (const void*)(src.*m), n4 * sizeof(int32_t));
    fn((void**)&dst.edge, (const void*)src.edge, n4 * sizeof(uint8_t));
    fn((void**)&dst.flag, (const void*)src.flag, F.N * sizeof(uint8_t));
}

// allocate dst's fields in mem (when alloc: one allocation, 256-byte aligned
// fields, so that freeing a copy is one hipFree) and copy src into them on
// the stream
wq3d_status copy_state(DeviceBuffers& mem, const FanConst& F, Arrays& dst, const Arrays& src, hipStream_t st,
                       bool alloc) {
    Arrays s = src;
    bool ok = true;
    uint8_t* base = nullptr;
    if (alloc) {
        size_t total = 0;
        Arrays probe{};
        for_each_field(F, probe, s, [&](void**, const void*, size_t bytes) { total += (bytes + 255) & ~size_t(255); });
        if (mem.alloc(&base, total) != hipSuccess) return fail(WQ3D_ERR_DEVICE, "adjoint state allocation failed");
    }
    size_t at = 0;
    for_each_field(F, dst, s, [&](void** d, const void* p, size_t bytes) {
        if (!ok) return;
        if (alloc) {
            *d = base + at;
            at += (bytes + 255) & ~size_t(255);
        }
        if (rt::memcpy_async(*d, p, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) ok = false;
    });
    if (!ok) return fail(WQ3D_ERR_DEVICE, "adjoint state copy failed");
    return WQ3D_OK;
}

// a new workspace for fans of f's shape with a history of `calls` steps
wq3d_status new_workspace(wq3d_fan* f, int64_t calls, std::unique_ptr<AdjWorkspace>& out) {
    const FanConst& F = f->F;
    auto w = std::make_unique<AdjWorkspace>();
    w->device = f->device;
    w->N = F.N;
    w->nf = F.nf;
    w->n_nodes = f->adjoint->n_nodes;
    w->calls = calls;
    DeviceBuffers& m = w->mem;
    const int64_t N = F.N;
    auto nomem = [&](const char* what) {
        return fail(WQ3D_ERR_DEVICE, std::string("adjoint workspace allocation failed (") + what + ", history of " +
                                         std::to_string(calls) + " steps)");
    };
    if (copy_state(m, F, w->W, f->A, f->stream, true)) return WQ3D_ERR_DEVICE;
    {
        unsigned long long* c = nullptr;
        int32_t* e = nullptr;
        if (m.alloc(&c, 6 + 2 * kScanSlots) != hipSuccess || m.alloc(&e, 1) != hipSuccess ||
            rt::memset(c, 0, (6 + 2 * kScanSlots) * sizeof(unsigned long long)) != hipSuccess ||
            rt::memset(e, 0, sizeof(int32_t)) != hipSuccess)
            return nomem("records");
        Records& R = w->Rw;
        R = Records{};
        R.ray_count = c;
        R.verb_count = c + 1;
        R.refl_count = c + 2;
        R.cand_count = c + 3;
        R.hit_count = c + 4;
        R.pass_count = c + 5;
        R.scan = c + 6;
        R.error = e;
        R.cand_cap = 0;
    }
    if (m.alloc(&w->kpw, 1) != hipSuccess) return nomem("parameters");
    AdjDev& D = w->D;
    D = AdjDev{};
    D.N = N;
    D.nf = F.nf;
    const int64_t n4 = 4 * N;
    const int64_t nadj = (int64_t)(wq3d::adj::kAtt + F.nf) * n4;
    const size_t tape_nodes = kTapePoolBytes / (sizeof(wq3d::ad::Node) + sizeof(double));
    double *hist = nullptr, *hatt = nullptr;
    int32_t* hint = nullptr;
    uint8_t *hedge = nullptr, *hflag = nullptr;
    if (m.alloc(&D.adj, nadj) != hipSuccess || m.alloc(&D.node_grad, std::max<int64_t>(1, w->n_nodes)) != hipSuccess)
        return nomem("adjoint state");
    if (m.alloc(&hist, (size_t)calls * wq3d::adj::kAtt * n4) != hipSuccess ||
        m.alloc(&hatt, (size_t)calls * F.nf * n4) != hipSuccess || m.alloc(&hint, (size_t)calls * 5 * n4) != hipSuccess ||
        m.alloc(&hedge, (size_t)calls * n4) != hipSuccess || m.alloc(&hflag, (size_t)calls * N) != hipSuccess)
        return nomem("history");
    if (m.alloc(&D.tape_node, tape_nodes) != hipSuccess || m.alloc(&D.tape_adj, tape_nodes) != hipSuccess)
        return nomem("tapes");
    if (m.alloc(&D.ent_key, kEntriesPerSegment) != hipSuccess || m.alloc(&D.ent_val, kEntriesPerSegment) != hipSuccess ||
        m.alloc(&w->skey, kEntriesPerSegment) != hipSuccess || m.alloc(&w->sval, kEntriesPerSegment) != hipSuccess ||
        m.alloc(&w->rbeg, N) != hipSuccess || m.alloc(&w->rend, N) != hipSuccess ||
        m.alloc(&D.ent_count, 1) != hipSuccess || m.alloc(&D.err, 1) != hipSuccess)
        return nomem("seed entries");
    if (m.alloc(&w->slots, calls) != hipSuccess || m.alloc(&w->times, calls) != hipSuccess) return nomem("steps");
#ifndef WQ3D_EMULATE
    w->sort_bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, w->sort_bytes, D.ent_key, w->skey, D.ent_val, w->sval,
                                           (int)kEntriesPerSegment, 0, 64, f->stream) != hipSuccess)
        return fail(WQ3D_ERR_DEVICE, "radix sort set-up failed");
    uint8_t* tmp = nullptr;
    if (m.alloc(&tmp, w->sort_bytes) != hipSuccess) return nomem("sort");
    w->sort_tmp = tmp;
#endif
    D.hist = hist;
    D.hatt = hatt;
    D.hint = hint;
    D.hedge = hedge;
    D.hflag = hflag;
    D.ent_cap = kEntriesPerSegment;
    out = std::move(w);
    return WQ3D_OK;
}

// give f a workspace with a history of at least `calls` steps (its own, one
// from the pool, or a new one) and point it at f's parameters
wq3d_status adjoint_workspace(wq3d_fan* f, int64_t calls) {
    FanAdjoint& a = *f->adjoint;
    const FanConst& F = f->F;
    auto fits = [&](const AdjWorkspace& w) {
        return w.device == f->device && w.N == F.N && w.nf == F.nf && w.n_nodes == a.n_nodes && w.calls >= calls;
    };
    if (a.ws && !fits(*a.ws)) a.ws.reset();
    if (!a.ws) {
        std::lock_guard<std::mutex> lock(ws_pool_mutex());
        auto& pool = ws_pool();
        for (size_t k = 0; k < pool.size(); ++k)
            if (fits(*pool[k])) {
                a.ws = std::move(pool[k]);
                pool.erase(pool.begin() + (long)k);
                break;
            }
    }
    if (!a.ws) {
        wq3d_status st = new_workspace(f, calls, a.ws);
        if (st) return st;
    }
    AdjWorkspace& w = *a.ws;
    // the search record stays the fan's: k_advance rewrites it every step
    // before k_detect reads it
    w.W.win = f->A.win;
    w.W.tsup = f->A.tsup;
    FanConst Fw = F;
    Fw.want_verbs = 0;
    Fw.want_refl = 0;
    const KParams kp{f->env->dev, Fw, w.W, w.Rw};
    if (rt::memcpy(w.kpw, &kp, sizeof(KParams), hipMemcpyHostToDevice) != hipSuccess)
        return fail(WQ3D_ERR_DEVICE, "adjoint parameter upload failed");
    return WQ3D_OK;
}

// a fan's workspace goes back to the pool (or is freed when the pool is full)
void give_workspace(FanAdjoint& a) {
    if (!a.ws) return;
    std::lock_guard<std::mutex> lock(ws_pool_mutex());
    if (ws_pool().size() < kPooledWorkspaces) ws_pool().push_back(std::move(a.ws));
    a.ws.reset();
}

// the replayed fan's slot order, time and step count
struct Replay {
    int phys[4];
    double time;
    int64_t steps;
    Slots slots() const { return Slots{phys[0], phys[1], phys[2], phys[3]}; }
    void rotate() {
        const int save = phys[0];
        phys[0] = phys[1];
        phys[1] = phys[2];
        phys[2] = phys[3];
        phys[3] = save;
    }
};

void launch_adj_capture(wq3d_fan* f, int64_t h) {
    FanAdjoint& a = *f->adjoint;
    const FanConst& F = f->F;
#ifndef WQ3D_EMULATE
    const unsigned grid = (unsigned)((F.N + kAdjBlock - 1) / kAdjBlock);
    hipLaunchKernelGGL(k_adj_capture, dim3(grid), dim3(kAdjBlock), 0, f->stream, F, a.ws->W, a.ws->D, h);
#else
    emu_check_arrays(F, a.ws->W);
    for (int64_t i = 0; i < F.N; ++i) wq3d::adj::body_adj_capture(i, F, a.ws->W, a.ws->D, h);
#endif
}

void launch_adj_seed(wq3d_fan* f, int64_t n) {
    FanAdjoint& a = *f->adjoint;
    AdjDev D = a.ws->D;
    D.tape_cap = kTapeSeed;
#ifndef WQ3D_EMULATE
    const int64_t threads = std::min<int64_t>(tapes_of(kTapeSeed), (n + 63) & ~int64_t(63));
    const unsigned grid = (unsigned)(threads / 64);
    hipLaunchKernelGGL(k_adj_seed, dim3(grid), dim3(64), 0, f->stream, f->env->dev, f->F, (const Slots*)a.ws->slots,
                       (const double*)a.ws->times, (const AdjSeed*)a.seeds, (const double*)a.seed_int,
                       (const double*)a.seed_grad, n, D);
#else
    emu_check_launch(f->env->dev, f->F, a.ws->W, nullptr);
    for (int64_t e = 0; e < n; ++e)
        wq3d::adj::body_adj_seed(e, 0, f->env->dev, f->F, a.ws->slots, a.ws->times, a.seeds, a.seed_int, a.seed_grad, D);
#endif
}

// sort the n entries by key into skey / sval and find each ray's range
wq3d_status sort_entries(wq3d_fan* f, int64_t n) {
    FanAdjoint& a = *f->adjoint;
    const int64_t N = f->F.N;
    HIP_TRY(rt::memset_async(a.ws->rbeg, 0, (size_t)N * sizeof(int64_t), f->stream));
    HIP_TRY(rt::memset_async(a.ws->rend, 0, (size_t)N * sizeof(int64_t), f->stream));
    if (n == 0) return WQ3D_OK;
#ifndef WQ3D_EMULATE
    size_t bytes = a.ws->sort_bytes;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(a.ws->sort_tmp, bytes, a.ws->D.ent_key, a.ws->skey, a.ws->D.ent_val, a.ws->sval, (int)n,
                                               0, 64, f->stream));
    const unsigned grid = (unsigned)((n + kAdjBlock - 1) / kAdjBlock);
    hipLaunchKernelGGL(k_adj_ranges, dim3(grid), dim3(kAdjBlock), 0, f->stream, (const unsigned long long*)a.ws->skey, n,
                       a.ws->rbeg, a.ws->rend);
#else
    std::vector<std::pair<unsigned long long, double>> e((size_t)n);
    for (int64_t k = 0; k < n; ++k) e[(size_t)k] = {a.ws->D.ent_key[k], a.ws->D.ent_val[k]};
    std::stable_sort(e.begin(), e.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    const int shift = 8 + wq3d::adj::kEntryStepBits;
    for (int64_t k = 0; k < n; ++k) {
        a.ws->skey[k] = e[(size_t)k].first;
        a.ws->sval[k] = e[(size_t)k].second;
    }
    for (int64_t k = 0; k < n; ++k) {
        const int64_t r = (int64_t)(a.ws->skey[k] >> shift);
        emu_range(a.ws->rbeg + r, 1, "entry ray");
        if (k == 0 || (int64_t)(a.ws->skey[k - 1] >> shift) != r) a.ws->rbeg[r] = k;
        if (k == n - 1 || (int64_t)(a.ws->skey[k + 1] >> shift) != r) a.ws->rend[r] = k + 1;
    }
#endif
    return WQ3D_OK;
}

void launch_adj_sweep(wq3d_fan* f, int64_t K) {
    FanAdjoint& a = *f->adjoint;
    const FanConst& F = f->F;
    AdjDev D = a.ws->D;
    D.tape_cap = kTapeSmall;
#ifndef WQ3D_EMULATE
    const int64_t threads = std::min<int64_t>(tapes_of(kTapeSmall), (F.N + kAdjBlock - 1) / kAdjBlock * kAdjBlock);
    const unsigned grid = (unsigned)((threads + kAdjBlock - 1) / kAdjBlock);
    hipLaunchKernelGGL(k_adj_sweep, dim3(grid), dim3(kAdjBlock), a.n_nodes * sizeof(double), f->stream, f->env->dev,
                       F, (const Slots*)a.ws->slots, K, (const unsigned long long*)a.ws->skey, (const double*)a.ws->sval,
                       (const int64_t*)a.ws->rbeg, (const int64_t*)a.ws->rend, D, (int32_t)a.n_nodes);
#else
    auto add = [&](int k0, double d0, double d1) {
        emu_range(a.ws->D.node_grad + k0, 2, "node gradient");
        a.ws->D.node_grad[k0] += d0;
        a.ws->D.node_grad[k0 + 1] += d1;
    };
    for (int64_t k = 0; k < F.N; ++k) {
        const int64_t i = emu_ray(k, F.N);
        wq3d::adj::body_adj_sweep(i, 0, f->env->dev, F, a.ws->slots, K, a.ws->skey, a.ws->sval, a.ws->rbeg[i], a.ws->rend[i], D, add);
    }
#endif
}

void launch_adj_init(wq3d_fan* f, const Slots& S) {
    FanAdjoint& a = *f->adjoint;
    const FanConst& F = f->F;
    AdjDev D = a.ws->D;
    D.tape_cap = kTapeSmall;
#ifndef WQ3D_EMULATE
    const int64_t threads = std::min<int64_t>(tapes_of(kTapeSmall), (F.N + 63) & ~int64_t(63));
    const unsigned grid = (unsigned)(threads / 64);
    hipLaunchKernelGGL(k_adj_init, dim3(grid), dim3(64), 0, f->stream, f->env->dev, F, S, D);
#else
    for (int64_t k = 0; k < F.N; ++k) wq3d::adj::body_adj_init(emu_ray(k, F.N), 0, f->env->dev, F, S, D);
#endif
}

}  // namespace

void adjoint_free(wq3d_fan* f) {
    if (f->adjoint) give_workspace(*f->adjoint);
    delete f->adjoint;
    f->adjoint = nullptr;
}

extern "C" {

wq3d_status wq3d_adjoint_checkpoint(wq3d_fan* f) {
    if (!f) return fail(WQ3D_ERR_ARGUMENT, "null fan");
    if (!adjoint_supported(f->env->dev))
        return fail(WQ3D_ERR_UNSUPPORTED,
                    "the adjoint covers a 1-D linear sound speed grid between flat boundaries");
    if (f->F.spreading != WQ3D_SPREADING_HYBRID_GAUSSIAN)
        return fail(WQ3D_ERR_UNSUPPORTED, "the adjoint covers hybrid Gaussian spreading");
    if (f->env->dev.profile.grid.ax[0].n > 4096)
        return fail(WQ3D_ERR_UNSUPPORTED, "the adjoint covers sound speed grids of at most 4096 nodes");
    HIP_TRY(rt::set_device(f->device));
    if (!f->adjoint) {
        if (f->steps != 0) return fail(WQ3D_ERR_STATE, "the first checkpoint must be taken before the first step");
        f->adjoint = new FanAdjoint();
        f->adjoint->n_nodes = f->env->dev.profile.grid.ax[0].n;
    }
    FanAdjoint& a = *f->adjoint;
    if (!a.cps.empty() && a.cps.back().steps == f->steps) return WQ3D_OK;
    if (!a.cps.empty() && f->steps - a.cps.back().steps >= (int64_t(1) << wq3d::adj::kEntryStepBits))
        return fail(WQ3D_ERR_ARGUMENT, "checkpoints must be less than 65536 steps apart");
    AdjCheckpoint cp;
    cp.steps = f->steps;
    cp.time = f->time;
    for (int k = 0; k < 4; ++k) cp.phys[k] = f->phys[k];
    wq3d_status st = copy_state(a.cp_mem, f->F, cp.A, f->A, f->stream, true);
    if (st) return st;
    a.cps.push_back(cp);
    HIP_TRY(rt::sync(f->stream));
    return WQ3D_OK;
}

wq3d_status wq3d_adjoint_release(wq3d_fan* f) {
    if (!f) return fail(WQ3D_ERR_ARGUMENT, "null fan");
    HIP_TRY(rt::set_device(f->device));
    HIP_TRY(rt::sync(f->stream));
    adjoint_free(f);
    return WQ3D_OK;
}

wq3d_status wq3d_adjoint_gradient(wq3d_fan* f, const wq3d_eigenray* rays, const double* intensity,
                                  const double* grad_time, const double* grad_intensity, int64_t num_rays,
                                  double* grad_nodes, int64_t num_nodes) {
    if (!f) return fail(WQ3D_ERR_ARGUMENT, "null fan");
    if (num_rays < 0 || (num_rays > 0 && (!rays || !intensity || !grad_time || !grad_intensity)) || !grad_nodes)
        return fail(WQ3D_ERR_ARGUMENT, "null argument");
    if (!f->adjoint || f->adjoint->cps.empty() || f->adjoint->cps[0].steps != 0)
        return fail(WQ3D_ERR_STATE, "no checkpoint at step 0 (wq3d_adjoint_checkpoint before stepping)");
    FanAdjoint& a = *f->adjoint;
    if (num_nodes != a.n_nodes) return fail(WQ3D_ERR_ARGUMENT, "num_nodes is not the sound speed grid's size");
    if (f->steps - a.cps.back().steps >= (int64_t(1) << wq3d::adj::kEntryStepBits))
        return fail(WQ3D_ERR_ARGUMENT, "checkpoints must be less than 65536 steps apart");
    HIP_TRY(rt::set_device(f->device));
    const FanConst& F = f->F;
    const int nf = F.nf;
    // the eigenrays by step
    std::vector<int64_t> order((size_t)num_rays);
    for (int64_t k = 0; k < num_rays; ++k) {
        const wq3d_eigenray& r = rays[k];
        // an eigenray's ray is one the search visits (detect_interior): its
        // neighbourhood lies inside the fan
        const int az_start = F.az_boundary ? 0 : 1;
        if (r.step < 1 || r.step > f->steps || r.de < 1 || r.de >= F.nde - 1 || r.az < az_start ||
            r.az >= F.naz - 1 || r.t1 < 0 || r.t2 < 0 || r.t2 >= F.nt2 || (int64_t)r.t1 * F.nt2 + r.t2 >= F.nt)
            return fail(WQ3D_ERR_ARGUMENT, "eigenray " + std::to_string(k) + " is not one of this fan's");
        order[(size_t)k] = k;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return rays[x].step < rays[y].step; });
    // segments [cps[j].steps, end_j)
    const size_t ns = a.cps.size();
    std::vector<int64_t> seg_end(ns);
    int64_t longest = 0;
    for (size_t j = 0; j < ns; ++j) {
        seg_end[j] = (j + 1 < ns) ? a.cps[j + 1].steps : f->steps;
        longest = std::max<int64_t>(longest, seg_end[j] - a.cps[j].steps);
    }
    wq3d_status st = adjoint_workspace(f, longest + 1);
    if (st) return st;
    AdjDev& D = a.ws->D;
    HIP_TRY(rt::sync(f->stream));
    HIP_TRY(rt::memset(D.adj, 0, (size_t)(wq3d::adj::kAtt + nf) * 4 * F.N * sizeof(double)));
    HIP_TRY(rt::memset(D.node_grad, 0, (size_t)std::max<int64_t>(1, a.n_nodes) * sizeof(double)));
    HIP_TRY(rt::memset(D.err, 0, sizeof(int32_t)));
    const unsigned grid = (unsigned)((F.N + kBlock - 1) / kBlock);
    const DEnv& E = f->env->dev;
    FanConst Fw = F;
    Fw.want_verbs = 0;
    Fw.want_refl = 0;
    size_t next_ray = order.size();  // eigenrays are consumed from the back, segment by segment
    for (size_t jj = ns; jj-- > 0;) {
        const AdjCheckpoint& cp = a.cps[jj];
        const int64_t a0 = cp.steps, e0 = seg_end[jj], K = e0 - a0;
        if (K <= 0) continue;
        // this segment's eigenrays
        size_t first = next_ray;
        while (first > 0 && rays[order[first - 1]].step > a0) --first;
        const int64_t nseg = (int64_t)(next_ray - first);
        if (nseg > a.seed_cap) {
            a.seed_mem.reset();
            a.seed_cap = std::max<int64_t>(nseg, 1024);
            if (a.seed_mem.alloc(&a.seeds, a.seed_cap) != hipSuccess ||
                a.seed_mem.alloc(&a.seed_int, a.seed_cap * nf) != hipSuccess ||
                a.seed_mem.alloc(&a.seed_grad, a.seed_cap * nf) != hipSuccess) {
                a.seed_cap = 0;
                return fail(WQ3D_ERR_DEVICE, "adjoint seed allocation failed");
            }
        }
        if (nseg) {
            std::vector<AdjSeed> hs((size_t)nseg);
            std::vector<double> hi((size_t)(nseg * nf)), hg((size_t)(nseg * nf));
            for (int64_t k = 0; k < nseg; ++k) {
                const int64_t r = order[first + (size_t)k];
                const wq3d_eigenray& x = rays[r];
                hs[(size_t)k] = AdjSeed{(int64_t)x.t1 * F.nt2 + x.t2, (int64_t)x.de * F.naz + x.az, x.travel_time,
                                        grad_time[r], x.step - a0};
                for (int q = 0; q < nf; ++q) {
                    hi[(size_t)(k * nf + q)] = intensity[r * nf + q];
                    hg[(size_t)(k * nf + q)] = grad_intensity[r * nf + q];
                }
            }
            HIP_TRY(rt::memcpy(a.seeds, hs.data(), hs.size() * sizeof(AdjSeed), hipMemcpyHostToDevice));
            HIP_TRY(rt::memcpy(a.seed_int, hi.data(), hi.size() * sizeof(double), hipMemcpyHostToDevice));
            HIP_TRY(rt::memcpy(a.seed_grad, hg.data(), hg.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        next_ray = first;
        // 1. replay the segment into the history
        st = copy_state(a.ws->mem, F, a.ws->W, cp.A, f->stream, false);
        if (st) return st;
        Replay rp{{cp.phys[0], cp.phys[1], cp.phys[2], cp.phys[3]}, cp.time, cp.steps};
        std::vector<Slots> slots_at((size_t)(K + 1));
        std::vector<double> times_at((size_t)(K + 1));
        slots_at[0] = rp.slots();
        times_at[0] = rp.time;
        launch_adj_capture(f, 0);
        for (int64_t h = 1; h <= K; ++h) {
            launch_bounce(f->stream, grid, a.ws->kpw, E, Fw, a.ws->W, rp.slots(), a.ws->Rw, rp.time, rp.steps);
            rp.rotate();
            rp.time += F.dt;
            rp.steps += 1;
            launch_advance(f->stream, grid, E, Fw, a.ws->W, rp.slots());
            launch_adj_capture(f, h);
            slots_at[(size_t)h] = rp.slots();
            times_at[(size_t)h] = rp.time;
        }
        HIP_TRY(rt::last_error());
        HIP_TRY(rt::memcpy_async(a.ws->slots, slots_at.data(), slots_at.size() * sizeof(Slots), hipMemcpyHostToDevice,
                                 f->stream));
        HIP_TRY(rt::memcpy_async(a.ws->times, times_at.data(), times_at.size() * sizeof(double), hipMemcpyHostToDevice,
                                 f->stream));
        // 2. the seeds of the segment's eigenrays, 3. sorted by ray and step
        HIP_TRY(rt::memset_async(D.ent_count, 0, sizeof(unsigned long long), f->stream));
        if (nseg) launch_adj_seed(f, nseg);
        HIP_TRY(rt::last_error());
        unsigned long long n_ent = 0;
        HIP_TRY(rt::sync(f->stream));  // the fan's stream does not wait for the legacy stream's copies
        HIP_TRY(rt::memcpy(&n_ent, D.ent_count, sizeof(n_ent), hipMemcpyDeviceToHost));
        n_ent = std::min<unsigned long long>(n_ent, (unsigned long long)kEntriesPerSegment);
        st = sort_entries(f, (int64_t)n_ent);
        if (st) return st;
        // 4. every ray backward through the segment
        launch_adj_sweep(f, K);
        HIP_TRY(rt::last_error());
        HIP_TRY(rt::sync(f->stream));
    }
    launch_adj_init(f, Slots{a.cps[0].phys[0], a.cps[0].phys[1], a.cps[0].phys[2], a.cps[0].phys[3]});
    HIP_TRY(rt::last_error());
    HIP_TRY(rt::sync(f->stream));
    int32_t err = 0;
    HIP_TRY(rt::memcpy(&err, D.err, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (err & (1 | 4))
        return fail(WQ3D_ERR_CAPACITY, "adjoint tape or seed-entry capacity exceeded (code " + std::to_string(err) + ")");
    if (err & 16) return fail(WQ3D_ERR_CAPACITY, "reflection chain longer than the adjoint replays");
    if (err & (2 | 8))
        return fail(WQ3D_ERR_STATE, "the replay does not reproduce the forward run or its eigenrays (code " +
                                        std::to_string(err) + ")");
    HIP_TRY(rt::memcpy(grad_nodes, D.node_grad, (size_t)a.n_nodes * sizeof(double), hipMemcpyDeviceToHost));
    give_workspace(a);  // the next pass, of this fan or another, takes it from the pool
    return WQ3D_OK;
}

}  // extern "C"

#ifdef WQ3D_EMULATE
extern "C" {
// Test hook of the CPU emulation build only: the hand-written adjoint of
// one ray's advance (step_vjp) against the same step recorded on a tape,
// at ray i of the fan's current state with pseudo-random output adjoints.
// Returns the largest difference over the input and node adjoints, relative
// to the largest adjoint.
double wq3d_emu_check_step_vjp(wq3d_fan* f, int64_t i, uint64_t seed) {
    using namespace wq3d;
    using namespace wq3d::adj;
    const FanConst& F = f->F;
    const DEnv& E = f->env->dev;
    const Slots S = slots_of(f);
    const Arrays& A = f->A;
    const int64_t N = F.N;
    auto rnd = [&]() {
        seed = seed * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(seed >> 11) / 9007199254740992.0 * 2.0 - 1.0;
    };
    StepIn in;
    const int sp[3] = {S.past, S.prev, S.curr};
    double* g3[3] = {in.g0, in.g1, in.g2};
    double* h3[3] = {in.h0, in.h1, in.h2};
    const double* G[3] = {A.gr, A.gt, A.gp};
    const double* H[3] = {A.hr, A.ht, A.hp};
    const double* P[3] = {A.pr, A.pt, A.pp};
    const double* Q[3] = {A.nr, A.nt, A.np};
    for (int s = 0; s < 3; ++s)
        for (int q = 0; q < 3; ++q) {
            g3[s][q] = G[q][at(sp[s], N, i)];
            h3[s][q] = H[q][at(sp[s], N, i)];
        }
    for (int q = 0; q < 3; ++q) {
        in.p2[q] = P[q][at(S.curr, N, i)];
        in.n2[q] = Q[q][at(S.curr, N, i)];
    }
    StepBar bar;
    for (int q = 0; q < 3; ++q) {
        bar.p[q] = rnd() * 1e-3;
        bar.n[q] = rnd() * 1e3;
        bar.g[q] = rnd() * 1e-1;
        bar.h[q] = rnd() * 1e3;
    }
    bar.c = rnd();
    for (int q = 0; q < 4; ++q) bar.s[q] = rnd() * (q == 3 ? 1e3 : 1e-3);
    for (int k = 0; k < F.nf; ++k) bar.att[k] = rnd();
    StepGrad out;
    step_vjp(E, F, in, bar, out);
    // the tape
    std::vector<ad::Node> nodes(1 << 14);
    std::vector<double> adjv(1 << 14);
    ad::Tape tp;
    ad::tape_init(tp, nodes.data(), adjv.data(), (int32_t)nodes.size());
    const NodesV nv{E.profile.grid.data, &tp};
    RayT<V> y[4];
    for (int s = 0; s < 4; ++s) y[s] = fresh_t<V>();
    int tag = 0;
    V* gi[3][3] = {{&y[0].gr, &y[0].gt, &y[0].gp}, {&y[1].gr, &y[1].gt, &y[1].gp}, {&y[2].gr, &y[2].gt, &y[2].gp}};
    V* hi[3][3] = {{&y[0].hr, &y[0].ht, &y[0].hp}, {&y[1].hr, &y[1].ht, &y[1].hp}, {&y[2].hr, &y[2].ht, &y[2].hp}};
    for (int s = 0; s < 3; ++s)
        for (int q = 0; q < 3; ++q) {
            *gi[s][q] = ad::input_plain(&tp, g3[s][q], tag++);
            *hi[s][q] = ad::input_plain(&tp, h3[s][q], tag++);
        }
    V* pi[3] = {&y[2].pr, &y[2].pt, &y[2].pp};
    V* ni[3] = {&y[2].nr, &y[2].nt, &y[2].np};
    for (int q = 0; q < 3; ++q) {
        *pi[q] = ad::input_plain(&tp, in.p2[q], tag++);
        *ni[q] = ad::input_plain(&tp, in.n2[q], tag++);
    }
    ab3_t(F.dt, y[0], y[1], y[2], y[3]);
    update_t(E, nv, F.R, y[3]);
    V att[kMaxFreq];
    for (int k = 0; k < F.nf; ++k) att[k] = attenuation_t(E.profile, seq_value(F.freq, k), y[3].pr - F.R, y[3].dist);
    ad::clear_adjoints(&tp);
    const V* po[3] = {&y[3].pr, &y[3].pt, &y[3].pp};
    const V* no[3] = {&y[3].nr, &y[3].nt, &y[3].np};
    const V* go[3] = {&y[3].gr, &y[3].gt, &y[3].gp};
    const V* ho[3] = {&y[3].hr, &y[3].ht, &y[3].hp};
    for (int q = 0; q < 3; ++q) {
        ad::seed(&tp, *po[q], bar.p[q] + bar.s[q]);
        ad::seed(&tp, *no[q], bar.n[q]);
        ad::seed(&tp, *go[q], bar.g[q]);
        ad::seed(&tp, *ho[q], bar.h[q]);
    }
    ad::seed(&tp, y[3].c, bar.c);
    ad::seed(&tp, y[3].st, bar.s[3]);
    for (int k = 0; k < F.nf; ++k) ad::seed(&tp, att[k], bar.att[k]);
    ad::reverse(&tp);
    std::vector<double> ref(tag, 0.0), node_ref(E.profile.grid.ax[0].n, 0.0);
    for (int32_t k = 0; k < tp.n; ++k) {
        if (tp.node[k].a == ad::kInputPlain) ref[tp.node[k].b] += tp.adj[k];
        if (tp.node[k].a == ad::kInputNode) node_ref[tp.node[k].b] += tp.adj[k];
    }
    std::vector<double> mine(tag, 0.0);
    int t2 = 0;
    const double* og[3] = {out.g0, out.g1, out.g2};
    const double* oh[3] = {out.h0, out.h1, out.h2};
    for (int s = 0; s < 3; ++s)
        for (int q = 0; q < 3; ++q) {
            mine[t2++] = og[s][q];
            mine[t2++] = oh[s][q];
        }
    for (int q = 0; q < 3; ++q) {
        mine[t2++] = out.p2[q];
        mine[t2++] = out.n2[q];
    }
    double worst = 0.0, scale = 0.0;
    for (int k = 0; k < tag; ++k) scale = std::max(scale, std::fabs(ref[k]));
    for (int k = 0; k < tag; ++k) {
        const double d = std::fabs(mine[k] - ref[k]) / std::max(std::fabs(ref[k]), 1e-300);
        if (d > 1e-9) std::fprintf(stderr, "step_vjp input %d: %.17g tape %.17g\n", k, mine[k], ref[k]);
        worst = std::max(worst, std::fabs(mine[k] - ref[k]) / scale);
    }
    std::vector<double> node_mine(node_ref.size(), 0.0);
    node_mine[out.node] += out.dnode[0];
    node_mine[out.node + 1] += out.dnode[1];
    for (size_t k = 0; k < node_ref.size(); ++k) {
        const double d = std::fabs(node_mine[k] - node_ref[k]);
        if (d > 1e-9 * std::max(std::fabs(node_ref[k]), 1e-300))
            std::fprintf(stderr, "step_vjp node %zu: %.17g tape %.17g\n", k, node_mine[k], node_ref[k]);
        worst = std::max(worst, d / std::max(scale, 1e-300));
    }
    return worst;
}
}  // extern "C"
#endif
