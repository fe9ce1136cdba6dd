// wq3d_adjoint_host.h - checkpoints and the backward pass of
// wq3d_adjoint_* (part of wq3d.hip's translation unit; included at its end).
//
// A fan being differentiated keeps a copy of its whole state every time the
// caller asks (wq3d_adjoint_checkpoint, every K steps of the forward run).
// wq3d_adjoint_gradient then walks the segments between checkpoints from
// the last to the first: it restores the segment's checkpoint into a
// replay copy of the fan's arrays, steps it forward with the ordinary
// reflection and advance kernels (no eigenray search: the caller's
// eigenrays say where the seeds are), keeping every step's window and
// computing the seeds of that step's eigenrays, and then runs the step
// adjoints of the segment in reverse order (wq3d_adjoint.h).  The fan's own
// state is not touched, so it may go on stepping afterwards.
#include "wq3d_adjoint.h"

namespace {

using wq3d::adj::AdjDev;
using wq3d::adj::AdjSeed;

constexpr int32_t kTapeSmall = 4096;    // operations of a reflection or start-up tape
constexpr int32_t kTapeSeed = 65536;    // operations of an eigenray's tape
constexpr int kAdjBlock = 256;

#ifndef WQ3D_EMULATE
constexpr size_t kTapePoolBytes = size_t(512) << 20;
constexpr int64_t kEntriesPerSegment = int64_t(1) << 24;
#else
constexpr size_t kTapePoolBytes = size_t(16) << 20;
constexpr int64_t kEntriesPerSegment = int64_t(1) << 20;
#endif

struct AdjCheckpoint {
    int64_t steps = 0;
    double time = 0.0;
    int phys[4] = {0, 1, 2, 3};
    Arrays A{};
};

}  // namespace

struct FanAdjoint {
    DeviceBuffers cp_mem;  // checkpoints
    std::vector<AdjCheckpoint> cps;
    // backward workspace
    DeviceBuffers ws_mem;
    bool have_ws = false;
    int64_t hist_calls = 0;
    Arrays W{};                 // replay arrays
    KParams* kpw = nullptr;     // {E, F without records, W, Rw}
    Records Rw{};
    AdjDev D{};
    double* hist = nullptr;
    uint8_t* hflag = nullptr;
    unsigned long long* eoff = nullptr;  // [hist_calls] entry counts after each replayed step
    AdjSeed* seeds = nullptr;
    double *seed_int = nullptr, *seed_grad = nullptr;
    int64_t seed_cap = 0;
    DeviceBuffers seed_mem;
    int64_t n_nodes = 0;
};

namespace {

#ifndef WQ3D_EMULATE
__global__ void __launch_bounds__(kAdjBlock) k_adj_capture(FanConst F, Arrays A, double* hist, uint8_t* hflag,
                                                           int64_t h) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F.N) wq3d::adj::body_adj_capture(i, F, A, hist, hflag, h);
}

__global__ void k_adj_mark(const unsigned long long* count, unsigned long long* eoff, int64_t h) {
    if (threadIdx.x == 0 && blockIdx.x == 0) eoff[h] = *count;
}

// the seed entries of replayed step h, added to the adjoint state
__global__ void __launch_bounds__(kAdjBlock) k_adj_scatter(AdjDev D, const unsigned long long* eoff, int64_t h) {
    const unsigned long long b = eoff[h - 1], e = eoff[h];
    for (unsigned long long k = b + (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; k < e;
         k += (unsigned long long)gridDim.x * blockDim.x)
        atomicAdd(D.adj + D.ent_idx[k], D.ent_val[k]);
}

__global__ void __launch_bounds__(64) k_adj_seed(const KParams* __restrict__ P, Slots S, double time,
                                                 const AdjSeed* seeds, const double* seed_int,
                                                 const double* seed_grad, int64_t lo, int64_t hi, AdjDev D) {
    const int64_t tk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = lo + tk; e < hi; e += nth)
        wq3d::adj::body_adj_seed(e, tk, P->E, P->F, P->A, S, time, seeds, seed_int, seed_grad, D);
}

// one ray per thread; the node adjoints are summed per block in LDS first
__global__ void __launch_bounds__(kAdjBlock) k_adj_advance(DEnv E, FanConst F, Slots S, int64_t h, AdjDev D,
                                                           int32_t n_nodes) {
    extern __shared__ double s_node[];
    for (int k = threadIdx.x; k < n_nodes; k += blockDim.x) s_node[k] = 0.0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F.N) {
        int k0;
        double d0, d1;
        wq3d::adj::body_adj_advance(i, E, F, S, h, D, k0, d0, d1);
        if (d0 != 0.0) atomicAdd(&s_node[k0], d0);
        if (d1 != 0.0) atomicAdd(&s_node[k0 + 1], d1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n_nodes; k += blockDim.x)
        if (s_node[k] != 0.0) atomicAdd(D.node_grad + k, s_node[k]);
}

__global__ void __launch_bounds__(64) k_adj_reflect(DEnv E, FanConst F, Slots Sp, Slots Sq, int64_t hp,
                                                    int64_t hq, AdjDev D) {
    const int64_t tk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tk; i < F.N; i += nth)
        if (D.hflag[hp * F.N + i] != 0) wq3d::adj::body_adj_reflect(i, tk, E, F, Sp, Sq, hp, hq, D);
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

// allocate dst's fields in mem (when alloc) and copy src into them on the stream
wq3d_status copy_state(DeviceBuffers& mem, const FanConst& F, Arrays& dst, const Arrays& src, hipStream_t st,
                       bool alloc) {
    Arrays s = src;
    bool ok = true;
    for_each_field(F, dst, s, [&](void** d, const void* p, size_t bytes) {
        if (!ok) return;
        if (alloc) {
            uint8_t* q = nullptr;
            if (mem.alloc(&q, bytes) != hipSuccess) {
                ok = false;
                return;
            }
            *d = q;
        }
        if (rt::memcpy_async(*d, p, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) ok = false;
    });
    if (!ok) return fail(WQ3D_ERR_DEVICE, "adjoint state copy failed");
    return WQ3D_OK;
}

wq3d_status adjoint_workspace(wq3d_fan* f, int64_t calls) {
    FanAdjoint& a = *f->adjoint;
    const FanConst& F = f->F;
    if (a.have_ws && a.hist_calls >= calls) return WQ3D_OK;
    a.ws_mem.reset();
    a.have_ws = false;
    DeviceBuffers& m = a.ws_mem;
    const int64_t N = F.N;
    // replay arrays (the search record stays the fan's: k_advance rewrites it
    // every step before k_detect reads it)
    if (copy_state(m, F, a.W, f->A, f->stream, true)) return WQ3D_ERR_DEVICE;
    a.W.win = f->A.win;
    a.W.tsup = f->A.tsup;
    {
        unsigned long long* c = nullptr;
        int32_t* e = nullptr;
        if (m.alloc(&c, 6 + 2 * kScanSlots) != hipSuccess || m.alloc(&e, 1) != hipSuccess ||
            rt::memset(c, 0, (6 + 2 * kScanSlots) * sizeof(unsigned long long)) != hipSuccess ||
            rt::memset(e, 0, sizeof(int32_t)) != hipSuccess)
            return fail(WQ3D_ERR_DEVICE, "adjoint workspace allocation failed");
        Records& R = a.Rw;
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
    FanConst Fw = F;
    Fw.want_verbs = 0;
    Fw.want_refl = 0;
    if (m.alloc(&a.kpw, 1) != hipSuccess) return fail(WQ3D_ERR_DEVICE, "adjoint workspace allocation failed");
    {
        const KParams kp{f->env->dev, Fw, a.W, a.Rw};
        if (rt::memcpy(a.kpw, &kp, sizeof(KParams), hipMemcpyHostToDevice) != hipSuccess)
            return fail(WQ3D_ERR_DEVICE, "adjoint parameter upload failed");
    }
    AdjDev& D = a.D;
    D = AdjDev{};
    D.N = N;
    const int64_t nadj = (int64_t)(wq3d::adj::kAtt + F.nf) * 4 * N;
    const size_t tape_nodes = kTapePoolBytes / (sizeof(wq3d::ad::Node) + sizeof(double));
    if (m.alloc(&D.adj, nadj) != hipSuccess || m.alloc(&D.node_grad, std::max<int64_t>(1, a.n_nodes)) != hipSuccess ||
        m.alloc(&a.hist, (size_t)calls * wq3d::adj::kHistFields * 4 * N) != hipSuccess ||
        m.alloc(&a.hflag, (size_t)calls * N) != hipSuccess || m.alloc(&a.eoff, calls) != hipSuccess ||
        m.alloc(&D.tape_node, tape_nodes) != hipSuccess || m.alloc(&D.tape_adj, tape_nodes) != hipSuccess ||
        m.alloc(&D.ent_idx, kEntriesPerSegment) != hipSuccess || m.alloc(&D.ent_val, kEntriesPerSegment) != hipSuccess ||
        m.alloc(&D.ent_count, 1) != hipSuccess || m.alloc(&D.err, 1) != hipSuccess)
        return fail(WQ3D_ERR_DEVICE, "adjoint workspace allocation failed (history of " + std::to_string(calls) +
                                         " steps)");
    D.hist = a.hist;
    D.hflag = a.hflag;
    D.ent_cap = kEntriesPerSegment;
    a.hist_calls = calls;
    a.have_ws = true;
    return WQ3D_OK;
}

// one replayed step of the fan's replay arrays (the reflections and the
// advance of wq3d_step, no search)
struct Replay {
    wq3d_fan* f;
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
    void unrotate() {
        const int save = phys[3];
        phys[3] = phys[2];
        phys[2] = phys[1];
        phys[1] = phys[0];
        phys[0] = save;
    }
};

void launch_adj_capture(wq3d_fan* f, int64_t h) {
    FanAdjoint& a = *f->adjoint;
    const FanConst& F = f->F;
#ifndef WQ3D_EMULATE
    const unsigned grid = (unsigned)((F.N + kAdjBlock - 1) / kAdjBlock);
    hipLaunchKernelGGL(k_adj_capture, dim3(grid), dim3(kAdjBlock), 0, f->stream, F, a.W, a.hist, a.hflag, h);
#else
    emu_check_arrays(F, a.W);
    for (int64_t i = 0; i < F.N; ++i) wq3d::adj::body_adj_capture(i, F, a.W, a.hist, a.hflag, h);
#endif
}

void launch_adj_mark(wq3d_fan* f, int64_t h) {
    FanAdjoint& a = *f->adjoint;
#ifndef WQ3D_EMULATE
    hipLaunchKernelGGL(k_adj_mark, dim3(1), dim3(64), 0, f->stream, a.D.ent_count, a.eoff, h);
#else
    a.eoff[h] = *a.D.ent_count;
#endif
}

void launch_adj_scatter(wq3d_fan* f, int64_t h) {
    FanAdjoint& a = *f->adjoint;
#ifndef WQ3D_EMULATE
    hipLaunchKernelGGL(k_adj_scatter, dim3(64), dim3(kAdjBlock), 0, f->stream, a.D, (const unsigned long long*)a.eoff, h);
#else
    for (unsigned long long k = a.eoff[h - 1]; k < a.eoff[h]; ++k) {
        emu_range(a.D.adj + a.D.ent_idx[k], 1, "adjoint entry");
        a.D.adj[a.D.ent_idx[k]] += a.D.ent_val[k];
    }
#endif
}

int64_t seed_tapes() { return (int64_t)(kTapePoolBytes / (sizeof(wq3d::ad::Node) + sizeof(double)) / kTapeSeed); }
int64_t small_tapes() { return (int64_t)(kTapePoolBytes / (sizeof(wq3d::ad::Node) + sizeof(double)) / kTapeSmall); }

void launch_adj_seed(wq3d_fan* f, const Slots& S, double time, int64_t lo, int64_t hi) {
    FanAdjoint& a = *f->adjoint;
    AdjDev D = a.D;
    D.tape_cap = kTapeSeed;
#ifndef WQ3D_EMULATE
    const int64_t threads = std::min<int64_t>(seed_tapes(), hi - lo);
    const unsigned grid = (unsigned)((threads + 63) / 64);
    hipLaunchKernelGGL(k_adj_seed, dim3(grid), dim3(64), 0, f->stream, (const KParams*)a.kpw, S, time,
                       (const AdjSeed*)a.seeds, (const double*)a.seed_int, (const double*)a.seed_grad, lo, hi, D);
#else
    const KParams* P = a.kpw;
    emu_check_launch(P->E, P->F, P->A, nullptr);
    for (int64_t e = lo; e < hi; ++e)
        wq3d::adj::body_adj_seed(e, 0, P->E, P->F, P->A, S, time, a.seeds, a.seed_int, a.seed_grad, D);
#endif
}

void launch_adj_advance(wq3d_fan* f, const Slots& S, int64_t h) {
    FanAdjoint& a = *f->adjoint;
    const FanConst& F = f->F;
#ifndef WQ3D_EMULATE
    const unsigned grid = (unsigned)((F.N + kAdjBlock - 1) / kAdjBlock);
    hipLaunchKernelGGL(k_adj_advance, dim3(grid), dim3(kAdjBlock), a.n_nodes * sizeof(double), f->stream,
                       f->env->dev, F, S, h, a.D, (int32_t)a.n_nodes);
#else
    for (int64_t k = 0; k < F.N; ++k) {
        const int64_t i = emu_ray(k, F.N);
        int k0;
        double d0, d1;
        wq3d::adj::body_adj_advance(i, f->env->dev, F, S, h, a.D, k0, d0, d1);
        emu_range(a.D.node_grad + k0, 2, "node gradient");
        a.D.node_grad[k0] += d0;
        a.D.node_grad[k0 + 1] += d1;
    }
#endif
}

void launch_adj_reflect(wq3d_fan* f, const Slots& Sp, const Slots& Sq, int64_t hp, int64_t hq) {
    FanAdjoint& a = *f->adjoint;
    const FanConst& F = f->F;
    AdjDev D = a.D;
    D.tape_cap = kTapeSmall;
#ifndef WQ3D_EMULATE
    const int64_t threads = std::min<int64_t>(small_tapes(), F.N);
    const unsigned grid = (unsigned)((threads + 63) / 64);
    hipLaunchKernelGGL(k_adj_reflect, dim3(grid), dim3(64), 0, f->stream, f->env->dev, F, Sp, Sq, hp, hq, D);
#else
    for (int64_t k = 0; k < F.N; ++k) {
        const int64_t i = emu_ray(k, F.N);
        if (a.hflag[hp * F.N + i] != 0) wq3d::adj::body_adj_reflect(i, 0, f->env->dev, F, Sp, Sq, hp, hq, D);
    }
#endif
}

void launch_adj_init(wq3d_fan* f, const Slots& S) {
    FanAdjoint& a = *f->adjoint;
    const FanConst& F = f->F;
    AdjDev D = a.D;
    D.tape_cap = kTapeSmall;
#ifndef WQ3D_EMULATE
    const int64_t threads = std::min<int64_t>(small_tapes(), F.N);
    const unsigned grid = (unsigned)((threads + 63) / 64);
    hipLaunchKernelGGL(k_adj_init, dim3(grid), dim3(64), 0, f->stream, f->env->dev, F, S, D);
#else
    for (int64_t k = 0; k < F.N; ++k) wq3d::adj::body_adj_init(emu_ray(k, F.N), 0, f->env->dev, F, S, D);
#endif
}

}  // namespace

void adjoint_free(wq3d_fan* f) {
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
    HIP_TRY(rt::set_device(f->device));
    if (f->env->dev.profile.grid.ax[0].n > 4096)
        return fail(WQ3D_ERR_UNSUPPORTED, "the adjoint covers sound speed grids of at most 4096 nodes");
    if (!f->adjoint) {
        if (f->steps != 0) return fail(WQ3D_ERR_STATE, "the first checkpoint must be taken before the first step");
        f->adjoint = new FanAdjoint();
        f->adjoint->n_nodes = f->env->dev.profile.grid.ax[0].n;
    }
    FanAdjoint& a = *f->adjoint;
    if (!a.cps.empty() && a.cps.back().steps == f->steps) return WQ3D_OK;
    if (!a.cps.empty() && a.cps.back().steps > f->steps) return fail(WQ3D_ERR_STATE, "checkpoints must advance");
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
    AdjDev& D = a.D;
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
        const int64_t a0 = cp.steps, e0 = seg_end[jj];
        if (e0 <= a0) continue;
        // this segment's eigenrays, uploaded by step
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
        std::vector<int64_t> call_lo((size_t)(e0 - a0 + 2), nseg);
        {
            std::vector<AdjSeed> hs((size_t)nseg);
            std::vector<double> hi((size_t)(nseg * nf)), hg((size_t)(nseg * nf));
            for (int64_t k = nseg - 1; k >= 0; --k) {
                const int64_t r = order[first + (size_t)k];
                const wq3d_eigenray& x = rays[r];
                hs[(size_t)k] = AdjSeed{(int64_t)x.t1 * F.nt2 + x.t2, (int64_t)x.de * F.naz + x.az, x.travel_time,
                                        grad_time[r]};
                for (int q = 0; q < nf; ++q) {
                    hi[(size_t)(k * nf + q)] = intensity[r * nf + q];
                    hg[(size_t)(k * nf + q)] = grad_intensity[r * nf + q];
                }
                call_lo[(size_t)(x.step - a0)] = k;
            }
            for (int64_t h = e0 - a0; h >= 1; --h) call_lo[(size_t)h] = std::min(call_lo[(size_t)h], call_lo[(size_t)h + 1]);
            if (nseg) {
                HIP_TRY(rt::memcpy(a.seeds, hs.data(), hs.size() * sizeof(AdjSeed), hipMemcpyHostToDevice));
                HIP_TRY(rt::memcpy(a.seed_int, hi.data(), hi.size() * sizeof(double), hipMemcpyHostToDevice));
                HIP_TRY(rt::memcpy(a.seed_grad, hg.data(), hg.size() * sizeof(double), hipMemcpyHostToDevice));
            }
        }
        next_ray = first;
        // replay the segment, keeping its windows and the seeds of its steps
        st = copy_state(a.ws_mem, F, a.W, cp.A, f->stream, false);
        if (st) return st;
        HIP_TRY(rt::memset_async(D.ent_count, 0, sizeof(unsigned long long), f->stream));
        HIP_TRY(rt::memset_async(a.eoff, 0, sizeof(unsigned long long), f->stream));
        Replay rp{f, {cp.phys[0], cp.phys[1], cp.phys[2], cp.phys[3]}, cp.time, cp.steps};
        std::vector<Slots> slots_at((size_t)(e0 - a0 + 1));
        slots_at[0] = rp.slots();
        launch_adj_capture(f, 0);
        for (int64_t h = 1; h <= e0 - a0; ++h) {
            launch_bounce(f->stream, grid, a.kpw, E, Fw, a.W, rp.slots(), a.Rw, rp.time, rp.steps);
            rp.rotate();
            rp.time += F.dt;
            rp.steps += 1;
            launch_advance(f->stream, grid, E, Fw, a.W, rp.slots());
            const int64_t lo = call_lo[(size_t)h], hi_ = call_lo[(size_t)h + 1];
            if (hi_ > lo) launch_adj_seed(f, rp.slots(), rp.time, lo, hi_);
            launch_adj_mark(f, h);
            launch_adj_capture(f, h);
            slots_at[(size_t)h] = rp.slots();
        }
        HIP_TRY(rt::last_error());
        // the step adjoints, last step first
        for (int64_t h = e0 - a0; h >= 1; --h) {
            launch_adj_scatter(f, h);
            launch_adj_advance(f, slots_at[(size_t)h], h);
            launch_adj_reflect(f, slots_at[(size_t)h - 1], slots_at[(size_t)h], h - 1, h);
        }
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
