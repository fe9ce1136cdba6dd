// wq3d_kernels.h - per-ray / per-pair bodies of the WaveQ3D kernels.
//
// Every function here is __host__ __device__: wq3d.hip launches them from
// thin __global__ wrappers on gfx950, and the CPU emulation build of the
// same library (WQ3D_EMULATE, used only by tests to run the kernels under
// host bounds checking) calls them in serial loops.  See wq3d.hip for the
// step pipeline.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/wq3d.h"
#include "wq3d_device.h"
#include "wq3d_emulate.h"

#define WQ_HD __host__ __device__
#define WQ_HD_INLINE __host__ __device__ __forceinline__

namespace wq3d {

// atomics: device atomics on the GPU, plain read-modify-write in the serial
// emulation
WQ_HD_INLINE unsigned long long atom_add(unsigned long long* p, unsigned long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return atomicAdd(p, v);
#else
    const unsigned long long old = *p;
    *p = old + v;
    return old;
#endif
}

WQ_HD_INLINE void atom_or(int32_t* p, int32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr(p, v);
#else
    *p |= v;
#endif
}


// --------------------------------------------------------------- layout
struct Arrays {
    // [4*N] per field, slot s at offset s*N
    double *pr, *pt, *pp, *gr, *gt, *gp, *nr, *nt, *np, *hr, *ht, *hp;
    double *c, *cr, *ct, *cp, *dist, *path;
    double *tr, *tt, *tp, *ts;  // position and wqm::sin(theta) at the slot's last update()
    double *atten, *phase;      // [4][nf][N]
    int32_t *srf, *btm, *cst, *upr, *lwr;
    uint8_t* edge;
    uint8_t* flag;              // [N] reflection kind for the coming step
    double* win;                // eigenray pre-test window of the step, tile by tile (store_window)
    uint32_t* tsup;             // [tiles][sup_words(F)] bit u: some ray of the tile passes super-block u's window bound
};

struct FanConst {
    int64_t N;
    int32_t nde, naz, nf;
    int32_t az_boundary;
    double dt, R;
    DSeq freq, de, az;
    const double* cos_de;       // launch direction tables (host libm)
    const double* sin_de;
    const double* cos_az;
    const double* sin_az;
    double c0;                  // sound speed at the (unclamped) source
    double src_r, src_t, src_p; // ray start (constructor argument)
    // targets
    int64_t nt;
    int32_t nt2;
    const double *tgt_r, *tgt_t, *tgt_p, *tgt_s;
    const uint8_t* tgt_branch;  // _de_branch per target (wave_queue.cc:363-368)
    // per-target terms of the distance expression, each computed once:
    // theta/2, phi/2, rho^2, 2 rho
    const double *tgt_th, *tgt_ph, *tgt_r2, *tgt_2r;
    // [nt][6] target features of the distance expression written as a
    // linear form (see lin_coeffs) and the rounding margin of that form
    // Targets in scan order: spatially sorted and cut into blocks of
    // kTargetBlock.  tgt_lin holds their features, tgt_perm the target index
    // of each scan position, tgt_box one TargetBox per block.
    const double* tgt_lin;
    const uint32_t* tgt_perm;
    // tgt_box: nblk block boxes, then nsup boxes of kSuperBlock consecutive
    // blocks; blk_branch: the same boxes' flag "holds a source-branch target"
    const double* tgt_box;
    const uint8_t* blk_branch;
    const uint8_t* scan_branch; // source-branch flag per scan position
    int64_t nblk, nsup;
    double lin_margin;
    double lin_fmin[6], lin_fmax[6];  // feature ranges over all targets
    // spreading_hybrid_gaussian norms (spreading_hybrid_gaussian.cc:28-59)
    const double* norm_de;
    const double* norm_az;
    int32_t has_spreading;
    // spreading_ray (CLASSIC_RAY, spreading_ray.cc:26-130): the model in use,
    // the initial ray-tube areas and the sound speed at the source
    int32_t spreading;          // wq3d_spreading
    const double* init_area;    // [nde][naz]
    double init_c;
    // thresholds
    double thr_intensity, thr_eigenverb;
    int32_t max_bottom, max_surface, max_caustic, max_upper, max_lower;
    int32_t want_verbs, want_refl;
};

struct Slots {
    int past, prev, curr, next;
};

struct Records {
    wq3d_eigenray* rays;
    double* ray_int;
    double* ray_phase;
    unsigned long long* ray_count;
    int64_t ray_cap;
    wq3d_eigenverb* verbs;
    double* verb_power;
    unsigned long long* verb_count;
    int64_t verb_cap;
    wq3d_reflection* refl;
    unsigned long long* refl_count;
    int64_t refl_cap;
    unsigned long long* hit_count;  // pairs the first-stage filter kept, this step
    unsigned long long* pass_count; // pairs that passed is_closest_ray, this step
    unsigned long long* cand;   // pairs the first-stage filter kept: (target << 32) | ray
    unsigned long long* cand_count;
    int64_t cand_cap;
    unsigned long long* scan;   // [kScanSlots][2] (target, ray) pairs put to the
                                // window test and (ray, box) bounds tested,
                                // summed over steps
    int32_t* error;             // bit 0: record overflow, bit 1: candidate overflow,
                                // bit 2: reflection chain exceeded its bound
};

constexpr int kMaxBounceChain = 64;
// counters of the window test's pair count, spread to keep atomics apart
constexpr int kScanSlots = 64;

WQ_HD_INLINE int64_t at(int s, int64_t N, int64_t i) { return (int64_t)s * N + i; }

WQ_HD_INLINE Ray load_ray(const Arrays& A, int64_t k) {
    Ray w;
    w.pr = A.pr[k]; w.pt = A.pt[k]; w.pp = A.pp[k];
    w.gr = A.gr[k]; w.gt = A.gt[k]; w.gp = A.gp[k];
    w.nr = A.nr[k]; w.nt = A.nt[k]; w.np = A.np[k];
    w.hr = A.hr[k]; w.ht = A.ht[k]; w.hp = A.hp[k];
    w.c = A.c[k]; w.cr = A.cr[k]; w.ct = A.ct[k]; w.cp = A.cp[k];
    w.dist = A.dist[k];
    w.st = 0.0;
    return w;
}

// reflection_model::reflection_copy (reflection_model.cc:335-360): the
// target-distance snapshot (tr/tt/tp/ts), attenuation and phase stay as they
// were; path_length gains the temporary front's zero path length.
WQ_HD_INLINE void copy_ray(const Arrays& A, int64_t k, const Ray& w) {
    A.pr[k] = w.pr; A.pt[k] = w.pt; A.pp[k] = w.pp;
    A.gr[k] = w.gr; A.gt[k] = w.gt; A.gp[k] = w.gp;
    A.nr[k] = w.nr; A.nt[k] = w.nt; A.np[k] = w.np;
    A.hr[k] = w.hr; A.ht[k] = w.ht; A.hp[k] = w.hp;
    A.cr[k] = w.cr; A.ct[k] = w.ct; A.cp[k] = w.cp;
    A.c[k] = w.c;
    A.dist[k] = w.dist;
}

// full slot write after update(): state plus target-distance snapshot
WQ_HD_INLINE void store_updated(const Arrays& A, int64_t k, const Ray& w) {
    copy_ray(A, k, w);
    A.tr[k] = w.pr; A.tt[k] = w.pt; A.tp[k] = w.pp; A.ts[k] = w.st;
}

WQ_HD_INLINE Ray fresh_ray() {
    Ray w;
    w.pr = w.pt = w.pp = w.gr = w.gt = w.gp = w.nr = w.nt = w.np = 0.0;
    w.hr = w.ht = w.hp = w.c = w.cr = w.ct = w.cp = w.dist = w.st = 0.0;
    return w;
}

// reflection kind decided from the state detect_reflections() will see
// (wave_queue.cc:222-256): 1 surface, 2 bottom, 0 none.
template <class K = AnyKinds>
WQ_HD_INLINE uint8_t reflect_kind(const DEnv& E, const FanConst& F, double next_r,
                                                double next_t, double next_p, double cnr, double cnt,
                                                double cnp) {
    if (next_r - F.R > 0.0) {
        const double grazing = wqm::atan2(cnr, sqrt(cnt * cnt + cnp * cnp));
        if (grazing > 0.0) return 1;
    }
    double h, n[3];
    boundary_height<K>(E.bottom, next_t, next_p, h, n, false);
    return (h - next_r > 0.0) ? 2 : 0;
}

// ------------------------------------------------------------------ init
// wave_front::init_wave + update + wave_queue::init_wavefronts
// (wave_front.cc:84-121, wave_queue.cc:83-85, 114-150), one thread per ray.
template <class K = AnyKinds>
WQ_HD void body_init(int64_t i, const DEnv& E, const FanConst& F, const Arrays& A, const Slots& S) {
    const int de = (int)(i / F.naz), az = (int)(i % F.naz);
    const double dt = F.dt;
    Ray cur = fresh_ray();
    cur.nr = F.sin_de[de];
    cur.nt = -F.cos_de[de] * F.cos_az[az];
    cur.np = F.cos_de[de] * F.sin_az[az];
    cur.pr = F.src_r;
    cur.pt = F.src_t;
    cur.pp = F.src_p;
    cur.nr = cur.nr / F.c0;
    cur.nt = cur.nt / F.c0;
    cur.np = cur.np / F.c0;
    update_ray<K>(E, F.R, cur);
    Ray nxt = fresh_ray(), pst = fresh_ray(), prv = fresh_ray();
    rk1(-dt, cur, nxt);
    update_ray<K>(E, F.R, nxt);
    rk2(-dt, cur, nxt, pst);
    update_ray<K>(E, F.R, pst);
    rk3(-dt, cur, nxt, pst, prv);
    update_ray<K>(E, F.R, prv);
    rk1(-dt, prv, nxt);
    update_ray<K>(E, F.R, nxt);
    rk2(-dt, prv, nxt, pst);
    update_ray<K>(E, F.R, pst);
    rk3(-dt, prv, nxt, pst, pst);
    update_ray<K>(E, F.R, pst);
    ab3(dt, pst, prv, cur, nxt);
    update_ray<K>(E, F.R, nxt);
    const int sl[4] = {S.past, S.prev, S.curr, S.next};
    const Ray* st[4] = {&pst, &prv, &cur, &nxt};
    const int64_t N = F.N;
    for (int s = 0; s < 4; ++s) {
        const int64_t k = at(sl[s], N, i);
        store_updated(A, k, *st[s]);
        A.path[k] = 0.0;
        A.srf[k] = A.btm[k] = A.cst[k] = A.upr[k] = A.lwr[k] = 0;
        A.edge[k] = 0;
        for (int f = 0; f < F.nf; ++f) {
            const int64_t kf = ((int64_t)sl[s] * F.nf + f) * N + i;
            A.atten[kf] = attenuation(E.profile, seq_value(F.freq, f), st[s]->pr - F.R, st[s]->dist);
            A.phase[kf] = 0.0;
        }
    }
    A.path[at(S.next, N, i)] = nxt.dist + 0.0;
    A.flag[i] = reflect_kind<K>(E, F, nxt.pr, nxt.pt, nxt.pp, cur.nr, cur.nt, cur.np);
}

// ------------------------------------------------------------- reflect
struct RayCtx {
    const DEnv& E;
    const FanConst& F;
    const Arrays& A;
    const Slots& S;
    const Records& R;
    int64_t i;
    int de, az;
    double time;      // wave_queue::time() before the rotation
    int64_t step_key;
    int seq;
};

// wave_queue::collision_location (wave_queue.cc:982-1057)
WQ_HD void collision_location(const RayCtx& x, int64_t i, double tw, double pos[3],
                                   double nd[3], double& speed) {
    const Arrays& A = x.A;
    const int64_t N = x.F.N;
    const int64_t kp = at(x.S.prev, N, i), kc = at(x.S.curr, N, i), kn = at(x.S.next, N, i);
    const double time1 = 2.0 * x.F.dt;
    const double time2 = x.F.dt * x.F.dt;
    const double dtime2 = tw * tw;
    {
        const double d1 = (A.c[kn] - A.c[kp]) / time1;
        const double d2 = (A.c[kn] + A.c[kp] - 2.0 * A.c[kc]) / time2;
        speed = A.c[kc] + d1 * tw + 0.5 * d2 * dtime2;
    }
    const double* P[3] = {A.pr, A.pt, A.pp};
    const double* Q[3] = {A.nr, A.nt, A.np};
    for (int k = 0; k < 3; ++k) {
        const double d1 = (P[k][kn] - P[k][kp]) / time1;
        const double d2 = (P[k][kn] + P[k][kp] - 2.0 * P[k][kc]) / time2;
        pos[k] = P[k][kc] + d1 * tw + 0.5 * d2 * dtime2;
    }
    for (int k = 0; k < 3; ++k) {
        const double d1 = (Q[k][kn] - Q[k][kp]) / time1;
        const double d2 = (Q[k][kn] + Q[k][kp] - 2.0 * Q[k][kc]) / time2;
        nd[k] = Q[k][kc] + d1 * tw + 0.5 * d2 * dtime2;
    }
}

WQ_HD_INLINE bool above_bounce(const FanConst& F, const Arrays& A, int64_t k) {
    return A.btm[k] > F.max_bottom || A.srf[k] > F.max_surface || A.cst[k] > F.max_caustic ||
           A.upr[k] > F.max_upper || A.lwr[k] > F.max_lower;
}

WQ_HD void emit_reflection(RayCtx& x, double tw, double grazing, double c, const double pos[3],
                                const double nd[3], int type) {
    const unsigned long long slot = atom_add(x.R.refl_count, 1ull);
    const int seq = x.seq++;
    if ((int64_t)slot >= x.R.refl_cap) {
        atom_or(x.R.error, 1);
        return;
    }
    wq3d_reflection& r = x.R.refl[slot];
    r.step = x.step_key;
    r.de = x.de;
    r.az = x.az;
    r.seq = seq;
    r.type = type;
    r.time = x.time + tw;
    r.dt = tw;
    r.grazing = grazing;
    r.speed = c;
    r.pos_rho = pos[0];
    r.pos_theta = pos[1];
    r.pos_phi = pos[2];
    r.ndir_rho = nd[0];
    r.ndir_theta = nd[1];
    r.ndir_phi = nd[2];
}

// wave_queue::build_eigenverb (wave_queue.cc:1063-1164)
WQ_HD void build_eigenverb(RayCtx& x, double dt, double grazing, double speed,
                                const double pos[3], const double nd[3], int type) {
    const FanConst& F = x.F;
    const Arrays& A = x.A;
    grazing = fabs(grazing);
    const int64_t kc = at(x.S.curr, F.N, x.i);
    const int max_az = F.naz - 1;
    if (!F.want_verbs || above_bounce(F, A, kc) || x.time <= 0.0 || grazing < 1e-6 ||
        (F.az_boundary && x.az == max_az) || fabs(seq_value(F.de, x.de)) > 89.9)
        return;
    const double de_angle = to_radians(seq_value(F.de, x.de));
    const double de_plus = de_angle + 0.5 * to_radians(seq_inc(F.de, x.de));
    const double de_minus = de_angle - 0.5 * to_radians(seq_inc(F.de, (int64_t)x.de - 1));
    const double az_angle = to_radians(seq_value(F.az, x.az));
    const double az_plus = az_angle + 0.5 * to_radians(seq_inc(F.az, x.az));
    const int az_index = (x.az == 0 && F.az_boundary) ? max_az : x.az;
    const double az_minus = az_angle - 0.5 * to_radians(seq_inc(F.az, (int64_t)az_index - 1));
    const double area = (wqm::sin(de_plus) - wqm::sin(de_minus)) * (az_plus - az_minus);
    const double de_delta = de_plus - de_minus;
    const double az_delta = area / de_delta;
    const double path_length = A.path[kc] + A.c[kc] * dt;
    if (path_length <= 0) return;
    const double sin_grazing = wqm::sin(grazing);
    double power[kMaxFreq];
    bool above = false;
    for (int f = 0; f < F.nf; ++f) {
        const double att = A.atten[((int64_t)x.S.curr * F.nf + f) * F.N + x.i];
        power[f] = wqm::pow10(-0.1 * att) * area / sin_grazing;
        if (power[f] >= F.thr_eigenverb) above = true;
    }
    if (!above) return;
    const unsigned long long slot = atom_add(x.R.verb_count, 1ull);
    const int seq = x.seq++;
    if ((int64_t)slot >= x.R.verb_cap) {
        atom_or(x.R.error, 1);
        return;
    }
    wq3d_eigenverb& v = x.R.verbs[slot];
    v.step = x.step_key;
    v.de = x.de;
    v.az = x.az;
    v.seq = seq;
    v.type = type;
    v.length = 0.5 * path_length * de_delta / sin_grazing;
    v.width = 0.5 * path_length * az_delta;
    v.direction = to_radians(to_degrees(wqm::atan2(nd[2], -nd[1])));
    v.travel_time = x.time + dt;
    v.grazing = grazing;
    v.sound_speed = speed;
    v.pos_rho = pos[0];
    v.pos_theta = pos[1];
    v.pos_phi = pos[2];
    v.source_de = to_radians(seq_value(F.de, x.de));
    v.source_az = to_radians(seq_value(F.az, x.az));
    v.surface = A.srf[kc];
    v.bottom = A.btm[kc];
    v.caustic = A.cst[kc];
    v.upper = A.upr[kc];
    v.lower = A.lwr[kc];
    for (int f = 0; f < F.nf; ++f) x.R.verb_power[slot * F.nf + f] = power[f];
}

// wave_queue::detect_volume_scattering (wave_queue.cc:297-345)
WQ_HD void volume_scattering(RayCtx& x) {
    const FanConst& F = x.F;
    if (!F.want_verbs) return;
    const Arrays& A = x.A;
    const int64_t kc = at(x.S.curr, F.N, x.i), kn = at(x.S.next, F.N, x.i);
    if (above_bounce(F, A, kc)) return;
    for (int l = 0; l < x.E.nvol; ++l) {
        const double height = x.E.vol_rho[l];
        const double d1 = height - A.pr[kn];
        const double d2 = height - A.pr[kc];
        if (d1 * d2 >= 0) continue;
        double c = A.c[kc];
        const double d = c * c * A.nr[kc];
        const double time_water = (d == 0.0) ? 0.0 : -(A.pr[kc] - F.R) / d;
        double pos[3], nd[3];
        collision_location(x, x.i, time_water, pos, nd, c);
        const double grazing = wqm::atan2(A.nr[kc], sqrt(A.nt[kc] * A.nt[kc] + A.np[kc] * A.np[kc]));
        const int type = (d1 > 0) ? 3 + l * 2 : 2 + l * 2;
        build_eigenverb(x, time_water, grazing, c, pos, nd, type);
    }
}

// reflection_model::reflection_reinit (reflection_model.cc:248-330), all
// temporaries in registers; the reference's third block advances a
// temporary it never updates, so its last RK stage reuses the gradient at
// the collision point (reflection_model.cc:313-318).
template <class K = AnyKinds>
WQ_HD void reflection_reinit(RayCtx& x, double time_water, const double pos[3],
                                  const double nd[3]) {
    const DEnv& E = x.E;
    const double R = x.F.R;
    const double dt = x.F.dt;
    const int64_t N = x.F.N;
    Ray temp = fresh_ray();
    temp.pr = pos[0];
    temp.pt = pos[1];
    temp.pp = pos[2];
    temp.nr = nd[0];
    temp.nt = nd[1];
    temp.np = nd[2];
    update_ray<K>(E, R, temp);
    Ray tn = fresh_ray(), tp = fresh_ray(), tc = fresh_ray(), tq = fresh_ray();
    rk1(-time_water, temp, tn);
    update_ray<K>(E, R, tn);
    rk2(-time_water, temp, tn, tp);
    update_ray<K>(E, R, tp);
    rk3(-time_water, temp, tn, tp, tc);
    update_ray<K>(E, R, tc);
    copy_ray(x.A, at(x.S.curr, N, x.i), tc);
    rk1(-dt, tc, tn);
    update_ray<K>(E, R, tn);
    rk2(-dt, tc, tn, tp);
    update_ray<K>(E, R, tp);
    rk3(-dt, tc, tn, tp, tq);
    update_ray<K>(E, R, tq);
    copy_ray(x.A, at(x.S.prev, N, x.i), tq);
    rk1(-dt, tq, tn);
    update_ray<K>(E, R, tn);
    rk2(-dt, tq, tn, temp);
    rk3(-dt, tq, tn, temp, tp);
    update_ray<K>(E, R, tp);
    copy_ray(x.A, at(x.S.past, N, x.i), tp);
    ab3(dt, tp, tq, tc, tn);
    update_ray<K>(E, R, tn);
    copy_ray(x.A, at(x.S.next, N, x.i), tn);
}

// reflection_model::surface_reflection (reflection_model.cc:187-243)
template <class K = AnyKinds>
WQ_HD bool surface_reflection(RayCtx& x) {
    const FanConst& F = x.F;
    const Arrays& A = x.A;
    const int64_t kc = at(x.S.curr, F.N, x.i);
    double c = A.c[kc];
    const double d = c * c * A.nr[kc];
    const double time_water = (d == 0.0) ? 0.0 : -(A.pr[kc] - F.R) / d;
    double pos[3], nd[3];
    collision_location(x, x.i, time_water, pos, nd, c);
    const double grazing = wqm::atan2(A.nr[kc], sqrt(A.nt[kc] * A.nt[kc] + A.np[kc] * A.np[kc]));
    if (grazing <= 0.0) return false;
    if (F.want_refl) emit_reflection(x, time_water, grazing, c, pos, nd, 1);
    if (F.want_verbs) build_eigenverb(x, time_water, grazing, c, pos, nd, 1);
    double amp, ph;
    reflect_loss(x.E.surface, grazing, amp, ph);
    for (int f = 0; f < F.nf; ++f) {
        const int64_t kf = ((int64_t)x.S.next * F.nf + f) * F.N + x.i;
        A.atten[kf] += amp;
        A.phase[kf] -= M_PI;
    }
    nd[0] = -nd[0];
    reflection_reinit<K>(x, time_water, pos, nd);
    return true;
}

// reflection_model::bottom_reflection (reflection_model.cc:33-182)
template <class K = AnyKinds>
WQ_HD bool bottom_reflection(RayCtx& x, double depth) {
    const double MIN_REFLECT = 6.0;
    const FanConst& F = x.F;
    const Arrays& A = x.A;
    const double TOO_SHALLOW = 300.0 * F.dt;
    const int64_t kc = at(x.S.curr, F.N, x.i);
    double pos[3] = {A.pr[kc], A.pt[kc], A.pp[kc]};
    double nd[3] = {A.nr[kc], A.nt[kc], A.np[kc]};
    double c = A.c[kc];
    double c2 = c * c;
    double bottom_rho, nrm[3];
    boundary_height<K>(x.E.bottom, pos[1], pos[2], bottom_rho, nrm, true);
    double height_water = pos[0] - bottom_rho;
    const bool shallow = (F.R - bottom_rho) < TOO_SHALLOW;
    if (shallow) {
        const double Nn = sqrt(nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        nrm[0] = 0.0;
        nrm[1] = nrm[1] / Nn;
        nrm[2] = nrm[2] / Nn;
    }
    nd[0] = c2 * nd[0];
    nd[1] = c2 * nd[1];
    nd[2] = c2 * nd[2];
    double dot_full = nrm[0] * nd[0] + nrm[1] * nd[1] + nrm[2] * nd[2];
    double max_dot = -smax(MIN_REFLECT, (height_water + depth) * nrm[0]);
    if (dot_full >= max_dot) dot_full = max_dot;
    const double dot_water = -height_water * nrm[0];
    const double time_water = smax(0.0, dot_water / dot_full);
    collision_location(x, x.i, time_water, pos, nd, c);
    boundary_height<K>(x.E.bottom, pos[1], pos[2], bottom_rho, nrm, true);
    if (shallow) {
        const double Nn = sqrt(nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        nrm[0] = 0.0;
        nrm[1] = nrm[1] / Nn;
        nrm[2] = nrm[2] / Nn;
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
    if (F.want_refl) emit_reflection(x, time_water, grazing, c, pos, nd, 0);
    if (F.want_verbs) build_eigenverb(x, time_water, grazing, c, pos, nd, 0);
    double amp, ph;
    reflect_loss(x.E.bottom, grazing, amp, ph);
    for (int f = 0; f < F.nf; ++f) {
        const int64_t kf = ((int64_t)x.S.next * F.nf + f) * F.N + x.i;
        A.atten[kf] += amp;
        A.phase[kf] += ph;
    }
    dot_full *= 2.0;
    nd[0] = nd[0] - dot_full * nrm[0];
    nd[1] = nd[1] - dot_full * nrm[1];
    nd[2] = nd[2] - dot_full * nrm[2];
    const double Nn = sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]) * c;
    nd[0] = nd[0] / Nn;
    nd[1] = nd[1] / Nn;
    nd[2] = nd[2] / Nn;
    reflection_reinit<K>(x, time_water, pos, nd);
    return true;
}

// detect_reflections_surface / _bottom bodies (wave_queue.cc:222-256)
template <class K = AnyKinds>
WQ_HD bool try_surface(RayCtx& x) {
    const int64_t N = x.F.N;
    const int64_t kn = at(x.S.next, N, x.i);
    if (!(x.A.pr[kn] - x.F.R > 0.0)) return false;
    if (!surface_reflection<K>(x)) return false;
    const int v = x.A.srf[kn] + 1;
    x.A.srf[kn] = v;
    x.A.srf[at(x.S.curr, N, x.i)] = v;
    x.A.srf[at(x.S.prev, N, x.i)] = v;
    x.A.srf[at(x.S.past, N, x.i)] = v;
    volume_scattering(x);
    return true;
}

template <class K = AnyKinds>
WQ_HD bool try_bottom(RayCtx& x) {
    const int64_t N = x.F.N;
    const int64_t kn = at(x.S.next, N, x.i);
    double height, n[3];
    boundary_height<K>(x.E.bottom, x.A.pt[kn], x.A.pp[kn], height, n, false);
    const double depth = height - x.A.pr[kn];
    if (!(depth > 0.0)) return false;
    if (!bottom_reflection<K>(x, depth)) return false;
    const int v = x.A.btm[kn] + 1;
    x.A.btm[kn] = v;
    x.A.btm[at(x.S.curr, N, x.i)] = v;
    x.A.btm[at(x.S.prev, N, x.i)] = v;
    x.A.btm[at(x.S.past, N, x.i)] = v;
    volume_scattering(x);
    return true;
}

// the rays the reflection part of detect_reflections must visit: those that
// reflect this step, or every ray when volume layers emit eigenverbs
WQ_HD_INLINE bool bounce_wanted(const DEnv& E, const FanConst& F, const Arrays& A, int64_t i) {
    return A.flag[i] != 0 || (F.want_verbs && E.nvol > 0);
}

// The reflection list a wanted ray goes to in a fused batch step: 1 for a
// bottom reflection, 0 otherwise.
WQ_HD_INLINE int bounce_side(const Arrays& A, int64_t i) { return A.flag[i] == 2 ? 1 : 0; }

// detect_reflections (wave_queue.cc:197-292) in two parts.  The first runs
// for every ray: the caustic test against the lower DE neighbour and, for
// rays that do not reflect this step, the vertex counters.  The second
// runs volume scattering and then, for the rays flag[] marks, the
// reflections.  A ray's caustic test only reads a neighbour that does not
// reflect, which the second part never writes; the caustic and vertex
// updates touch the next slot's counters and phase, which volume
// scattering does not read, so the split keeps the reference's results.
WQ_HD void body_reflect_light(int64_t i, const FanConst& F, const Arrays& A, const Slots& S) {
    const int64_t N = F.N;
    const int de = (int)(i / F.naz);
    const int64_t kc = at(S.curr, N, i), kn = at(S.next, N, i);
    // caustic found while processing the lower DE neighbour, which precedes
    // this ray in the reference's de-major loop (wave_queue.cc:274-292)
    if (de >= 1 && A.flag[i - F.naz] == 0) {
        const int64_t j = i - F.naz;
        const double Av = A.pr[kc];
        const double Bv = A.pr[at(S.curr, N, j)];
        const double Cv = A.pr[kn];
        const double Dv = A.pr[at(S.next, N, j)];
        const bool fold = A.srf[kn] == A.srf[at(S.next, N, j)] && A.btm[kn] == A.btm[at(S.next, N, j)];
        if ((Cv - Dv) * (Av - Bv) < 0 && fold) {
            A.cst[kn]++;
            for (int f = 0; f < F.nf; ++f) A.phase[((int64_t)S.next * F.nf + f) * N + i] -= M_PI_2;
        }
    }
    if (A.flag[i] != 0) return;
    // wave_queue::detect_vertices (wave_queue.cc:261-269)
    const double L = A.nr[kc];
    const double Rr = A.nr[kn];
    if (L * Rr < 0.0 && Rr < 0.0) {
        A.upr[kn]++;
    } else if (L * Rr < 0.0 && 0.0 < Rr) {
        A.lwr[kn]++;
    }
}

template <class K = AnyKinds>
WQ_HD void body_bounce(int64_t i, const DEnv& E, const FanConst& F, const Arrays& A, const Slots& S,
                       const Records& Rc, double time, int64_t step_key) {
    const bool reflects = A.flag[i] != 0;
    if (!reflects && !(F.want_verbs && E.nvol > 0)) return;
    const int de = (int)(i / F.naz), az = (int)(i % F.naz);
    RayCtx x{E, F, A, S, Rc, i, de, az, time, step_key, 0};
    volume_scattering(x);
    if (!reflects) return;
    // surface <-> bottom recursion as a bounded loop
    bool bottom_next;
    if (try_surface<K>(x)) {
        bottom_next = true;
    } else {
        try_bottom<K>(x);
        bottom_next = false;
    }
    for (int n = 0; n < kMaxBounceChain; ++n) {
        const bool hit = bottom_next ? try_bottom<K>(x) : try_surface<K>(x);
        if (!hit) return;
        bottom_next = !bottom_next;
    }
    atom_or(Rc.error, 4);
}

// ------------------------------------------------------ eigenray totals
// eigenray_collection::add_eigenray (eigenrays/eigenray_collection.cc:90-101)
// and sum_eigenrays (:105-191) for one target, over its eigenrays in the
// order the collection received them (order[begin .. end)): frequency-major
// loops, sums in that order, so the result is the reference's bit for bit
// given the same elementary functions.
struct SumIn {
    const wq3d_eigenray* rays;
    const double *inten, *phase;  // [ray][nf]
    const int64_t* order;         // ray indices grouped by target, in order
    const int64_t* offset;        // [targets + 1] into order
    const double* freq;
    int32_t nf, coherent;
};

struct SumOut {
    wq3d_eigenray_total* tot;
    double *inten, *phase;  // [target][nf]
};

WQ_HD void body_sum_eigenrays(int64_t target, const SumIn& in, const SumOut& out) {
    const double TWO_PI = 2.0 * M_PI;
    const int64_t b = in.offset[target], e = in.offset[target + 1];
    const int nf = in.nf;
    double initial = 0.0;
    for (int64_t k = b; k < e; ++k) {
        const double t = in.rays[in.order[k]].travel_time;
        if (initial <= 0.0 || initial > t) initial = t;
    }
    double time = 0.0, source_de = 0.0, source_az_x = 0.0, source_az_y = 0.0;
    double target_de = 0.0, target_az_x = 0.0, target_az_y = 0.0;
    int32_t surface = -1, bottom = -1, caustic = -1, upper = -1, lower = -1;
    double wgt = 0.0, max_a = 0.0;
    for (int f = 0; f < nf; ++f) {
        wqm::cplx phasor = {0.0, 0.0};
        for (int64_t k = b; k < e; ++k) {
            const int64_t r = in.order[k];
            const wq3d_eigenray& ray = in.rays[r];
            double a = wqm::pow10(in.inten[r * nf + f] / -20.0);
            if (in.coherent) {
                double p = TWO_PI * in.freq[f] * ray.travel_time + in.phase[r * nf + f];
                p = ::fmod(p, TWO_PI);
                phasor.re += a * wqm::cos(p);
                phasor.im += a * wqm::sin(p);
            } else {
                phasor.re += a;
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
        out.inten[target * nf + f] = -20.0 * wqm::log10(smax(1e-15, wqm::cabs(phasor)));
        out.phase[target * nf + f] = wqm::carg(phasor);
    }
    wq3d_eigenray_total& T = out.tot[target];
    T.travel_time = time / wgt;
    T.source_de = source_de / wgt;
    T.source_az = 90.0 - to_degrees(wqm::atan2(source_az_y, source_az_x));
    T.target_de = target_de / wgt;
    T.target_az = 90.0 - to_degrees(wqm::atan2(target_az_y, target_az_x));
    T.initial_time = initial;
    T.surface = surface;
    T.bottom = bottom;
    T.caustic = caustic;
    T.upper = upper;
    T.lower = lower;
    T.count = (int32_t)(e - b);
}

// The per-target lists body_sum_eigenrays reads (order / offset), built on
// the device: count per target into offset[t + 1], an inclusive scan turns
// the counts into offsets, each ray takes a position in its target's list
// with an atomic cursor, and each list is then sorted by ray index, which is
// arrival order - the order a stable counting sort on the host gives.
WQ_HD_INLINE void body_count_target(int64_t k, const wq3d_eigenray* rays, int64_t nt2, int64_t* offset) {
    const int64_t t = (int64_t)rays[k].t1 * nt2 + rays[k].t2;
    atom_add(reinterpret_cast<unsigned long long*>(offset + t + 1), 1ull);
}

WQ_HD_INLINE void body_scatter_ray(int64_t k, const wq3d_eigenray* rays, int64_t nt2, const int64_t* offset,
                                   int64_t* cursor, int64_t* order) {
    const int64_t t = (int64_t)rays[k].t1 * nt2 + rays[k].t2;
    const unsigned long long at_ = atom_add(reinterpret_cast<unsigned long long*>(cursor + t), 1ull);
    order[offset[t] + (int64_t)at_] = k;
}

// insertion sort of one target's list (lists hold tens of rays)
WQ_HD_INLINE void body_sort_list(int64_t t, const int64_t* offset, int64_t* order) {
    const int64_t b = offset[t], e = offset[t + 1];
    for (int64_t k = b + 1; k < e; ++k) {
        const int64_t v = order[k];
        int64_t j = k;
        while (j > b && order[j - 1] > v) {
            order[j] = order[j - 1];
            --j;
        }
        order[j] = v;
    }
}

// --------------------------------------------------------------- advance
// find_edges (gather form of wave_front.cc:184-216) on the old next slot,
// then AB3 + update() of the new next slot and the carries of step()
// (wave_queue.cc:171-183).
WQ_HD_INLINE bool is_extremum(const double* r, int64_t base, int naz, int d, int md) {
    if (d < 1 || d > md - 1) return false;
    const double x = r[base + (int64_t)d * naz];
    const double xp = r[base + (int64_t)(d + 1) * naz];
    const double xm = r[base + (int64_t)(d - 1) * naz];
    return (x < xp && x < xm) || (x > xp && x > xm);
}

template <class K = AnyKinds>
WQ_HD void body_advance(int64_t i, const DEnv& E, const FanConst& F, const Arrays& A, const Slots& S) {
    const int64_t N = F.N;
    const int de = (int)(i / F.naz), az = (int)(i % F.naz);
    const int md = F.nde - 1;
    // on_edge of the slot that just became curr
    {
        const int64_t base = (int64_t)S.curr * N + az;
        const double* r = A.pr;
        const double* x = A.nr;
        bool e = (de == 0 || de == md) || is_extremum(r, base, F.naz, de, md);
        if (!e && is_extremum(r, base, F.naz, de + 1, md)) {
            const double c = x[base + (int64_t)(de + 1) * F.naz];
            if (fabs(c - x[base + (int64_t)de * F.naz]) > fabs(c - x[base + (int64_t)(de + 2) * F.naz]))
                e = true;
        }
        if (!e && is_extremum(r, base, F.naz, de - 1, md)) {
            const double c = x[base + (int64_t)(de - 1) * F.naz];
            if (!(fabs(c - x[base + (int64_t)(de - 2) * F.naz]) > fabs(c - x[base + (int64_t)de * F.naz])))
                e = true;
        }
        A.edge[at(S.curr, N, i)] = e ? 1 : 0;
    }
    const int64_t k0 = at(S.past, N, i), k1 = at(S.prev, N, i), k2 = at(S.curr, N, i);
    const int64_t k3 = at(S.next, N, i);
    Ray y0, y1, y2, y3 = fresh_ray();
    y0.gr = A.gr[k0]; y0.gt = A.gt[k0]; y0.gp = A.gp[k0];
    y0.hr = A.hr[k0]; y0.ht = A.ht[k0]; y0.hp = A.hp[k0];
    y1.gr = A.gr[k1]; y1.gt = A.gt[k1]; y1.gp = A.gp[k1];
    y1.hr = A.hr[k1]; y1.ht = A.ht[k1]; y1.hp = A.hp[k1];
    y2 = load_ray(A, k2);
    ab3(F.dt, y0, y1, y2, y3);
    update_ray<K>(E, F.R, y3);
    store_updated(A, k3, y3);
    A.path[k3] = y3.dist + A.path[k2];
    for (int f = 0; f < F.nf; ++f) {
        const int64_t f3 = ((int64_t)S.next * F.nf + f) * N + i;
        const int64_t f2 = ((int64_t)S.curr * F.nf + f) * N + i;
        A.atten[f3] = attenuation(E.profile, seq_value(F.freq, f), y3.pr - F.R, y3.dist) + A.atten[f2];
        A.phase[f3] = 0.0 + A.phase[f2];
    }
    A.srf[k3] = A.srf[k2];
    A.btm[k3] = A.btm[k2];
    A.upr[k3] = A.upr[k2];
    A.lwr[k3] = A.lwr[k2];
    A.cst[k3] = A.cst[k2];
    A.flag[i] = reflect_kind<K>(E, F, y3.pr, y3.pt, y3.pp, y2.nr, y2.nt, y2.np);
}

// ---------------------------------------------------------------- detect
// The eigenray search of detect_eigenrays (wave_queue.cc:372-398) is split
// in two: k_detect filters every (target, interior ray) pair with linear
// forms in the target's features - a window equivalent to the time-CPA
// pre-test's outcome (see detect_window), the previous-time comparison and
// the current-time comparison with the four in-plane neighbours, each with
// a rounding margin so that it never drops a pair the exact test keeps -
// and k_build runs the exact pre-test and is_closest_ray in full on the
// pairs it keeps.  A pair is only ever dropped on a condition for which
// is_closest_ray would return false.  detect_pair below is the exact test
// of one pair, which the emulation build's filter check compares against.
//
// Distance: the expression of wave_front.cc:227-232 with per-ray and
// per-target terms hoisted.  It is the same IEEE value: 0.5*(a-b) equals
// a/2-b/2 (power-of-two scaling commutes with rounding, no underflow at
// these magnitudes) and 1-2Y is fma(-2,Y,1) because 2Y is exact.
struct DPoint {
    double th, ph, r, r2, s;  // theta/2, phi/2, rho, rho^2, sin(theta)
};

struct DTarget {
    double th, ph, s, R, R2, twoR;
};

WQ_HD_INLINE DPoint dpoint(const Arrays& A, int64_t k) {
    DPoint p;
    p.th = 0.5 * A.tt[k];
    p.ph = 0.5 * A.tp[k];
    p.r = A.tr[k];
    p.r2 = p.r * p.r;
    p.s = A.ts[k];
    return p;
}

WQ_HD_INLINE double dist2_fast(const DPoint& p, const DTarget& t) {
    const double h = p.th - t.th;
    const double g = p.ph - t.ph;
    return fabs((p.r2 + t.R2) - t.twoR * (p.r * ::fma(-2.0, h * h + t.s * (p.s * (g * g)), 1.0)));
}

// The time-CPA pre-test compares distances of one target to the same ray at
// three times.  Expanding the distance expression above,
//   d2 = (r - R)^2 + 4 r R ((a - at)^2 + s st (b - bt)^2)
// (a = theta/2, b = phi/2, s = sin theta of the ray; at, bt, st, R of the
// target), a difference d2(q) - d2(p) of two ray points is linear in six
// target features
//   F = (R, R at, R at^2, R st, R st bt, R st bt^2)
// with coefficients c(q) - c(p), c = (-2r + 4r a^2, -8r a, 4r, 4r s b^2,
// -8r s b, 4r s), plus (r_q - r_p)(r_q + r_p).  Six fused multiply-adds give
// the sign of the difference whenever its magnitude exceeds lin_margin, a
// bound on the rounding of the estimate and of the two exact distances
// (both well under 1 m^2 for positions near the earth's radius, against a
// margin of a few hundred m^2).  Pairs inside the margin, and every pair
// that passes, are decided with the exact expression, so the outcome is
// that of the reference's comparisons bit for bit.
WQ_HD_INLINE void lin_coeffs(const DPoint& p, double c[6]) {
    const double r = p.r, a = p.th, b = p.ph;
    const double rs = r * p.s;
    c[0] = -2.0 * r + 4.0 * r * (a * a);
    c[1] = -8.0 * r * a;
    c[2] = 4.0 * r;
    c[3] = 4.0 * rs * (b * b);
    c[4] = -8.0 * rs * b;
    c[5] = 4.0 * rs;
}

struct LinForm {
    double k, a[6];
};

WQ_HD_INLINE LinForm lin_diff(const DPoint& q, const DPoint& p) {
    double cq[6], cp[6];
    lin_coeffs(q, cq);
    lin_coeffs(p, cp);
    LinForm L;
    L.k = (q.r - p.r) * (q.r + p.r);
    for (int j = 0; j < 6; ++j) L.a[j] = cq[j] - cp[j];
    return L;
}

WQ_HD_INLINE double lin_eval(const LinForm& L, const double f[6]) {
    double d = ::fma(L.a[0], f[0], L.k);
    for (int j = 1; j < 6; ++j) d = ::fma(L.a[j], f[j], d);
    return d;
}

struct DetectRay {
    int64_t i;           // ray index
    DPoint cur;
    int32_t nbj[4];      // neighbours de-1, de+1, az-1, az+1 (AZ wrap of is_closest_ray)
    uint32_t use;        // bit n: neighbour n not skipped by is_closest_ray (non-branch targets)
    uint32_t upper;      // bit n: nde == 2 || naz == 2, rejection on <= instead of <
};

// The pre-test window of one ray: d2(next) - d2(cur) shifted to the centre
// of the interval it must lie in, and the half width of that interval.
struct DetectWindow {
    LinForm form;
    double w;
    LinForm prev;  // d2(prev) - d2(cur)
    double m;      // rounding margin
};

// The pre-test needs d2(next) - d2(cur) > 0 and d2(prev) - d2(cur) >= 0.
// Their sum, the second difference of d2 along the ray, has small
// coefficients, so an interval bound of it over the box of target features
// (S_up) is tight; any target whose next-difference lies outside
// [-margin, S_up + 2 margin] fails the pre-test.  One linear form and one
// compare against that window reject almost every pair; the prev form then
// settles most of the rest.
WQ_HD_INLINE DetectWindow detect_window(const FanConst& F, const LinForm& next, const LinForm& prev) {
    const double m = F.lin_margin;
    double up = next.k + prev.k;
    for (int j = 0; j < 6; ++j) {
        const double a = next.a[j] + prev.a[j];
        up += a >= 0.0 ? a * F.lin_fmax[j] : a * F.lin_fmin[j];
    }
    const double lo = -m, hi = up + 2.0 * m;
    const double c = 0.5 * (lo + hi);
    DetectWindow W;
    W.form = next;
    W.form.k = next.k - c;
    W.w = 0.5 * (hi - lo) + 1.0;
    if (!(W.w < 1e300)) W.w = __builtin_inf();  // NaN / overflow: no rejection
    W.prev = prev;
    W.m = m;
    return W;
}

// false when the window rejects the pair, or, for the pairs inside it, when
// d2(prev) - d2(cur) is below -margin (prev surely closer than center).
// The second form matters for rays whose window is wide, such as rays whose
// three snapshots straddle a reflection.  NaN passes through.
WQ_HD_INLINE bool detect_window_in(const DetectWindow& W, const double f[6]) {
    return !(fabs(lin_eval(W.form, f)) > W.w);
}
WQ_HD_INLINE bool detect_window_prev_ok(const DetectWindow& W, const double f[6]) {
    return !(lin_eval(W.prev, f) < -W.m);
}
WQ_HD_INLINE bool detect_window_ok(const DetectWindow& W, const double f[6]) {
    return detect_window_in(W, f) && detect_window_prev_ok(W, f);
}

// Box of one block of targets, in the variables of the features: theta/2
// and phi/2 as centre and half width, rho and sin(theta) as ranges.
constexpr int kTargetBlock = 64;
constexpr int kSuperBlock = 8;  // blocks per super-block (one k_detect work unit)
struct TargetBox {
    double at0, hx, bt0, hy, r_lo, r_hi, s_lo, s_hi;
};

// Interval of the window form over every target of a box.  With the
// features written as R (P(at) + st Q(bt)), P and Q quadratics in theta/2
// and phi/2, their Taylor expansions about the box centre bound P and Q
// exactly; st and R are positive ranges.  false when no target of the box
// can be inside the window (one more margin absorbs the rounding of this
// bound and of the features).
WQ_HD_INLINE void block_form_range(const LinForm& L, const TargetBox& B, double& lo, double& hi) {
    const double* a = L.a;
    const double P0 = a[0] + a[1] * B.at0 + a[2] * (B.at0 * B.at0);
    const double P1 = fabs(a[1] + 2.0 * a[2] * B.at0) * B.hx;
    const double P2 = a[2] * (B.hx * B.hx);
    const double p_lo = P0 - P1 + (P2 < 0.0 ? P2 : 0.0);
    const double p_hi = P0 + P1 + (P2 > 0.0 ? P2 : 0.0);
    const double Q0 = a[3] + a[4] * B.bt0 + a[5] * (B.bt0 * B.bt0);
    const double Q1 = fabs(a[4] + 2.0 * a[5] * B.bt0) * B.hy;
    const double Q2 = a[5] * (B.hy * B.hy);
    const double q_lo = Q0 - Q1 + (Q2 < 0.0 ? Q2 : 0.0);
    const double q_hi = Q0 + Q1 + (Q2 > 0.0 ? Q2 : 0.0);
    const double sq_lo = fmin(fmin(B.s_lo * q_lo, B.s_hi * q_lo), fmin(B.s_lo * q_hi, B.s_hi * q_hi));
    const double sq_hi = fmax(fmax(B.s_lo * q_lo, B.s_hi * q_lo), fmax(B.s_lo * q_hi, B.s_hi * q_hi));
    const double g_lo = p_lo + sq_lo, g_hi = p_hi + sq_hi;
    lo = L.k + fmin(B.r_lo * g_lo, B.r_hi * g_lo);
    hi = L.k + fmax(B.r_lo * g_hi, B.r_hi * g_hi);
}

WQ_HD_INLINE bool block_window_ok(const DetectWindow& W, const TargetBox& B) {
    double lo, hi;
    block_form_range(W.form, B, lo, hi);
    const double w = W.w + W.m;
    return !(lo > w || hi < -w);
}

// In-plane neighbours of ray i that is_closest_ray compares a target
// against (wave_queue.cc:423-478): DE-1, DE+1, AZ-1, AZ+1 with the AZ wrap
// of a fan that closes on itself.  stat bit n: neighbour n is not excluded
// by the AZ rule (the DE neighbours' edge flags are checked separately);
// upper bit n: rejection on <= instead of <.
WQ_HD_INLINE void ray_neighbours(const FanConst& F, int64_t i, int32_t nbj[4], uint32_t& stat, uint32_t& upper) {
    const int de = (int)(i / F.naz), az = (int)(i % F.naz);
    const int max_az = F.naz - 1;
    const int dd[4] = {de - 1, de + 1, de, de};
    const int nz[4] = {1, 1, 0, 2};
    const int nd[4] = {0, 2, 1, 1};
    stat = 0;
    upper = 0;
    for (int n = 0; n < 4; ++n) {
        int a = az + nz[n] - 1;
        if (F.az_boundary) {
            if (az + nz[n] == 0) {
                a = F.naz - 2;
            } else if (a >= max_az) {
                a = 0;
            }
        }
        nbj[n] = (int32_t)((int64_t)dd[n] * F.naz + a);
        stat |= (((a == 0 && !F.az_boundary) || a == max_az) ? 0u : 1u) << n;
        upper |= ((nd[n] == 2 || nz[n] == 2) ? 1u : 0u) << n;
    }
}

// The candidate region of one ray: the pre-test window and, for each
// neighbour is_closest_ray compares against, d2(neighbour) - d2(cur) as a
// linear form.  A target for which one of these is below -margin is closer
// to that neighbour than to this ray (the rounding of the form is within
// the margin), so is_closest_ray rejects it, unless the target has the
// source-branch rule, which skips the neighbour comparison.
struct DetectCell {
    DetectWindow W;
    LinForm nb[4];
    uint32_t use;  // bit n: neighbour n is compared
};

WQ_HD_INLINE bool detect_cell_nb_ok(const DetectCell& C, const double f[6]) {
    for (int n = 0; n < 4; ++n)
        if (((C.use >> n) & 1u) && lin_eval(C.nb[n], f) < -C.W.m) return false;
    return true;
}

WQ_HD_INLINE bool detect_cell_ok(const DetectCell& C, const double f[6], bool branch) {
    return detect_window_ok(C.W, f) && (branch || detect_cell_nb_ok(C, f));
}

// Whole-block form of detect_cell_ok: false when no target of the box can
// pass (blocks holding a source-branch target skip the neighbour part).
// The neighbour bound gets one more margin for its own rounding, as the
// window bound does.
// the neighbour part of block_cell_ok
WQ_HD_INLINE bool block_cell_nb_ok(const DetectCell& C, const TargetBox& B) {
    for (int n = 0; n < 4; ++n) {
        if (!((C.use >> n) & 1u)) continue;
        double lo, hi;
        block_form_range(C.nb[n], B, lo, hi);
        if (hi < -2.0 * C.W.m) return false;
    }
    return true;
}

WQ_HD_INLINE bool block_cell_ok(const DetectCell& C, const TargetBox& B, bool branch) {
    if (!block_window_ok(C.W, B)) return false;
    if (branch) return true;
    return block_cell_nb_ok(C, B);
}

// The rays detect_eigenrays visits: interior and not on an edge.
WQ_HD_INLINE bool detect_interior(const FanConst& F, const Arrays& A, const Slots& S, int64_t i) {
    if (i < 0 || i >= F.N) return false;
    const int de = (int)(i / F.naz), az = (int)(i % F.naz);
    const int az_start = F.az_boundary ? 0 : 1;
    if (!(de >= 1 && de < F.nde - 1 && az >= az_start && az < F.naz - 1)) return false;
    return A.edge[at(S.curr, F.N, i)] == 0;
}

// Window of one ray for the first stage.
WQ_HD_INLINE bool detect_ray_window(const FanConst& F, const Arrays& A, const Slots& S, int64_t i,
                                    DetectWindow& W) {
    if (!detect_interior(F, A, S, i)) return false;
    const int64_t N = F.N;
    const DPoint cur = dpoint(A, at(S.curr, N, i));
    const DPoint nxt = dpoint(A, at(S.next, N, i));
    const DPoint prv = dpoint(A, at(S.prev, N, i));
    W = detect_window(F, lin_diff(nxt, cur), lin_diff(prv, cur));
    return true;
}

// The pre-test window of ray i for the current step, computed once by
// k_advance (after the new next slot's update()) and read by k_supmask and
// by every target group of k_detect.  Record of kWinStride doubles:
// {form.k, form.a[6], w}; w = -inf marks a ray that is not searched.  The
// rest of the candidate region - the previous-time form and the neighbour
// forms - k_detect computes from the slots' distance snapshots when a
// wavefront needs it (load_cell_rest): the same points through the same
// lin_diff give the same forms, and the record stays a quarter of the size
// (k_advance writes 64 B per ray instead of 352).
constexpr int kWinStride = 8;
// Records are stored tile by tile (kWinTile x kWinTile rays, AZ fastest,
// the k_detect wavefront's tile) and field-major within a tile, so that a
// wavefront reads each field of its 64 rays as one contiguous 512-byte row.
constexpr int kWinTile = 8;

WQ_HD_INLINE int64_t win_tiles(const FanConst& F) {
    return (int64_t)((F.nde + kWinTile - 1) / kWinTile) * ((F.naz + kWinTile - 1) / kWinTile);
}

// words of one tile's super-block bit mask
WQ_HD_INLINE int64_t sup_words(const FanConst& F) { return (F.nsup + 31) / 32; }

// offset of field 0 of ray i's record; field q is at q * kWinTile^2 past it
WQ_HD_INLINE int64_t win_slot(const FanConst& F, int64_t i) {
    const int64_t de = i / F.naz, az = i % F.naz;
    const int64_t tiles_az = (F.naz + kWinTile - 1) / kWinTile;
    const int64_t tile = (de / kWinTile) * tiles_az + az / kWinTile;
    return tile * (kWinTile * kWinTile) * kWinStride + (de % kWinTile) * kWinTile + az % kWinTile;
}
constexpr int kWinField = kWinTile * kWinTile;

WQ_HD_INLINE void store_window(const FanConst& F, const Arrays& A, const Slots& S, int64_t i) {
    double* out = A.win + win_slot(F, i);
    DetectWindow W;
    if (F.nt <= 0 || !detect_ray_window(F, A, S, i, W)) {
        out[7 * kWinField] = -HUGE_VAL;  // the width field marks the ray as not searched
        return;
    }
    out[0] = W.form.k;
    for (int j = 0; j < 6; ++j) out[(1 + j) * kWinField] = W.form.a[j];
    // a searched ray's width is never -inf (no window test tells -inf from
    // the most negative finite width)
    out[7 * kWinField] = W.w == -HUGE_VAL ? -0x1.fffffffffffffp+1023 : W.w;
}

// The candidate region of ray i in two parts for k_detect: the window form
// and width from the record (false for rays not searched, whose width field
// is -inf), which the super-block's window bound needs, and the rest - the
// previous-time form, the neighbour forms and the neighbours
// is_closest_ray compares - read only by wavefronts some of whose rays pass
// it.
WQ_HD_INLINE bool load_cell_window(const FanConst& F, const Arrays& A, int64_t i, DetectCell& C) {
    if (i < 0 || i >= F.N) return false;
    const double* in = A.win + win_slot(F, i);
    C.W.form.k = in[0];
    for (int j = 0; j < 6; ++j) C.W.form.a[j] = in[(1 + j) * kWinField];
    C.W.w = in[7 * kWinField];
    C.W.m = F.lin_margin;
    return C.W.w != -HUGE_VAL;
}

// The rest of the region of a searched ray (detect_interior, as the record
// was written for it): d2(prev) - d2(cur) and d2(neighbour) - d2(cur) from
// the prev and curr slots' snapshots, and which neighbours are compared -
// the AZ rule's (ray_neighbours) less the DE neighbours on an edge.  The
// edge flags are the curr slot's, which k_advance wrote in the launch
// before.  Rays not searched get no forms and no neighbours.
WQ_HD_INLINE void load_cell_rest(const FanConst& F, const Arrays& A, const Slots& S, int64_t i, DetectCell& C) {
    C.use = 0u;
    if (!detect_interior(F, A, S, i)) {
        // defined forms for every lane: k_detect reads another lane's forms
        // with a shuffle that every lane takes part in
        C.W.prev.k = 0.0;
        for (int j = 0; j < 6; ++j) C.W.prev.a[j] = 0.0;
        for (int n = 0; n < 4; ++n) {
            C.nb[n].k = 0.0;
            for (int j = 0; j < 6; ++j) C.nb[n].a[j] = 0.0;
        }
        return;
    }
    const int64_t N = F.N;
    const DPoint cur = dpoint(A, at(S.curr, N, i));
    C.W.prev = lin_diff(dpoint(A, at(S.prev, N, i)), cur);
    int32_t nbj[4];
    uint32_t stat, upper;
    ray_neighbours(F, i, nbj, stat, upper);
    for (int n = 0; n < 4; ++n) {
        if ((stat >> n) & 1u) {
            C.nb[n] = lin_diff(dpoint(A, at(S.curr, N, nbj[n])), cur);
        } else {
            C.nb[n].k = 0.0;
            for (int j = 0; j < 6; ++j) C.nb[n].a[j] = 0.0;
        }
    }
    uint32_t use = stat;
    if ((use & 1u) && A.edge[at(S.curr, N, i - F.naz)] != 0) use &= ~1u;
    if ((use & 2u) && A.edge[at(S.curr, N, i + F.naz)] != 0) use &= ~2u;
    C.use = use;
}

// the whole region of ray i as a DetectCell; false for rays not searched
WQ_HD_INLINE bool load_cell(const FanConst& F, const Arrays& A, const Slots& S, int64_t i, DetectCell& C) {
    if (!load_cell_window(F, A, i, C)) return false;
    load_cell_rest(F, A, S, i, C);
    return true;
}

// Exact test of one (target, ray) pair: the time-CPA pre-test (next >
// center, prev >= center; wave_queue.cc:384-394) and, for targets without
// the source-branch rule, the current-time part of is_closest_ray for the
// four in-plane neighbours (wave_queue.cc:458-478).
template <class Points>
WQ_HD_INLINE bool detect_pair(const DetectRay& r, const Points& pts, const DTarget& t, bool branch) {
    const double center = dist2_fast(r.cur, t);
    if (!(dist2_fast(pts.nxt(r.i), t) > center)) return false;
    if (dist2_fast(pts.prv(r.i), t) < center) return false;
    if (branch) return true;
#pragma unroll 1
    for (int n = 0; n < 4; ++n) {
        if (!((r.use >> n) & 1u)) continue;
        const double d = dist2_fast(pts.nb(r.nbj[n]), t);
        if (((r.upper >> n) & 1u) ? (d <= center) : (d < center)) return false;
    }
    return true;
}

WQ_HD_INLINE DTarget dtarget(const FanConst& F, int64_t t) {
    DTarget x;
    x.th = F.tgt_th[t];
    x.ph = F.tgt_ph[t];
    x.s = F.tgt_s[t];
    x.R = F.tgt_r[t];
    x.R2 = F.tgt_r2[t];
    x.twoR = F.tgt_2r[t];
    return x;
}

// ----------------------------------------------------------------- build
struct BuildCtx {
    const DEnv& E;
    const FanConst& F;
    const Arrays& A;
    const Slots& S;
};

WQ_HD_INLINE double dist2_slot(const BuildCtx& b, int slot, int64_t i, int64_t t) {
    const int64_t k = at(slot, b.F.N, i);
    return target_distance2(b.A.tr[k], b.A.tt[k], b.A.tp[k], b.A.ts[k], b.F.tgt_r[t], b.F.tgt_t[t],
                            b.F.tgt_p[t], b.F.tgt_s[t]);
}

// Ray of the 3 x 3 neighbourhood point (nde, naz) of ray (de, az), with the
// AZ wrap of is_closest_ray and build_eigenray (wave_queue.cc:423-431,
// :544-548).
WQ_HD_INLINE int64_t cube_ray(const FanConst& F, int de, int az, int nde, int naz) {
    const int max_az = F.naz - 1;
    const int d = de + nde - 1;
    int a = az + naz - 1;
    if (F.az_boundary) {
        if (az + naz == 0) {
            a = F.naz - 2;
        } else if (a >= max_az) {
            a = 0;
        }
    }
    return (int64_t)d * F.naz + a;
}

// wave_queue::is_closest_ray (wave_queue.cc:411-488) on the distances D of
// the whole (time, DE, AZ) neighbourhood, checked in the reference's order.
WQ_HD_INLINE bool is_closest_ray(const BuildCtx& b, int64_t t, int de, int az, const double D[3][3][3],
                          const double edge_cur[3][3]) {
    const FanConst& F = b.F;
    const int max_az = F.naz - 1;
    const bool branch = F.tgt_branch[t] != 0;
    const double center = D[1][1][1];
    for (int nde = 0; nde < 3; ++nde) {
        for (int naz = 0; naz < 3; ++naz) {
            if (nde == 1 && naz == 1) continue;
            const int64_t i = cube_ray(F, de, az, nde, naz);
            const int a = (int)(i % F.naz);
            if (a == 0 && !F.az_boundary) continue;
            if (a == max_az) continue;
            const bool edge = edge_cur[nde][naz] != 0.0;  // edge of ray i in the current slot
            if (branch) {
                if (edge) continue;
            } else if (nde != 1) {
                if (edge) continue;
            }
            if (nde == 2 || naz == 2) {
                if (branch) {
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

// wave_queue::make_taylor_coeff (wave_queue.cc:929-977)
WQ_HD_INLINE void make_taylor_coeff(const double v[3][3][3], const double delta[3], double& center,
                                  double g[3], double H[3][3]) {
    const double d0 = 2.0 * delta[0];
    const double d1 = 2.0 * delta[1];
    const double d2 = 2.0 * delta[2];
    center = v[1][1][1];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) H[a][c] = 0.0;
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

WQ_HD_INLINE double inner3(const double a[3], const double b[3]) {
    double t = 0.0;
    for (int k = 0; k < 3; ++k) t += a[k] * b[k];
    return t;
}
WQ_HD_INLINE void prod3(const double M[3][3], const double x[3], double y[3]) {
    for (int r = 0; r < 3; ++r) {
        double t = 0.0;
        for (int k = 0; k < 3; ++k) t += M[r][k] * x[k];
        y[r] = t;
    }
}
WQ_HD_INLINE double norm3(const double g[3]) {
    double t = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double u = fabs(g[k]);
        t += u * u;
    }
    return sqrt(t);
}

// collision_location of another ray (used by compute_offsets)
WQ_HD_INLINE void collision_at(const BuildCtx& b, int64_t i, double tw, double pos[3]) {
    const Arrays& A = b.A;
    const int64_t N = b.F.N;
    const int64_t kp = at(b.S.prev, N, i), kc = at(b.S.curr, N, i), kn = at(b.S.next, N, i);
    const double time1 = 2.0 * b.F.dt;
    const double time2 = b.F.dt * b.F.dt;
    const double dtime2 = tw * tw;
    const double* P[3] = {A.pr, A.pt, A.pp};
    for (int k = 0; k < 3; ++k) {
        const double d1 = (P[k][kn] - P[k][kp]) / time1;
        const double d2 = (P[k][kn] + P[k][kp] - 2.0 * P[k][kc]) / time2;
        pos[k] = P[k][kc] + d1 * tw + 0.5 * d2 * dtime2;
    }
}

// wave_queue::compute_offsets (wave_queue.cc:708-924)
// The first stability test of compute_offsets at one neighbourhood point:
// ray j in slot s (0 prev, 1 curr, 2 next) against the centre's counts.
WQ_HD_INLINE bool counters_differ_at(const BuildCtx& b, int s, int64_t j, int s0, int b0, int c0) {
    const int slot = s == 0 ? b.S.prev : (s == 1 ? b.S.curr : b.S.next);
    const int64_t k = at(slot, b.F.N, j);
    return b.A.srf[k] != s0 || b.A.btm[k] != b0 || b.A.cst[k] != c0 || (s == 0 && b.A.edge[k] != 0);
}

// counters_differ: some ray of the neighbourhood, in some of the three
// slots, has surface/bottom/caustic counts other than the centre's, or is on
// an edge in the prev slot (the first stability test, wave_queue.cc:720-745;
// see counters_differ_at)
// The part of compute_offsets that depends only on the neighbourhood's
// distances: the Taylor coefficients, the stability tests, the 3 x 3 or
// 2 x 2 (time, AZ) solve, the per-axis distances and, on the unstable
// branch, the DE distance as the remainder.  Returns whether the branch was
// unstable; the caller then gives a positive DE distance its sign from the
// rays' geometry.
WQ_HD_INLINE bool offsets_core(const double D[3][3][3], const double delta[3], bool counters_differ,
                               double offset[3], double distance[3]) {
    bool unstable = counters_differ;
    double center, g[3], H[3][3], inv[3][3];
    make_taylor_coeff(D, delta, center, g, H);
    double determinant = smax(1e-6, H[0][0] * (H[1][1] * H[2][2] - H[1][2] * H[2][1]) +
                                        H[0][1] * (H[1][2] * H[2][0] - H[1][0] * H[2][2]) +
                                        H[0][2] * (H[1][0] * H[2][1] - H[2][1] * H[2][0]));
    unstable = unstable || determinant < norm3(g);
    for (int n = 0; n < 3; ++n) {
        const double h = smax(1e-10, H[n][n]);
        offset[n] = -g[n] / h;
        if (fabs(offset[n] / delta[n]) > 0.5) {
            unstable = true;
            offset[n] = copysign(0.5 * delta[n], offset[n]);
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
            for (int c = 0; c < 3; ++c) inv[a][c] /= determinant;
        const double ng[3] = {-g[0], -g[1], -g[2]};
        prod3(inv, ng, offset);
        unstable = unstable || fabs(offset[1]) >= 0.5 * fabs(delta[1]);
    } else {
        const double save_offset = offset[1];
        g[1] = 0.0;
        determinant = H[0][0] * H[2][2] - H[0][2] * H[2][0];
        if (determinant >= norm3(g)) {
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) inv[a][c] = 0.0;
            inv[0][0] = H[2][2];
            inv[0][2] = -H[0][2];
            inv[1][1] = 1.0;
            inv[2][0] = -H[2][0];
            inv[2][2] = H[0][0];
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) inv[a][c] /= determinant;
            const double ng[3] = {-g[0], -g[1], -g[2]};
            prod3(inv, ng, offset);
            offset[1] = save_offset;
        }
    }
    for (int n = 0; n < 3; ++n) {
        distance[n] = -g[n] * offset[n] - 0.5 * H[n][n] * offset[n] * offset[n];
        distance[n] = sqrt(smax(0.0, distance[n]));
        if (offset[n] < 0.0) distance[n] *= -1.0;
    }
    if (unstable) {
        distance[1] = center - distance[0] * distance[0] - distance[2] * distance[2];
        if (distance[1] <= 0.0) {
            distance[1] = 0.0;
            offset[1] = 0.0;
        } else {
            distance[1] = sqrt(distance[1]);
            offset[1] = fabs(offset[1]);
        }
    }
    return unstable;
}

WQ_HD_INLINE void compute_offsets(const BuildCtx& b, int64_t t, int de, int az, const double D[3][3][3],
                                const double delta[3], bool counters_differ, double offset[3],
                                double distance[3]) {
    const FanConst& F = b.F;
    const bool unstable = offsets_core(D, delta, counters_differ, offset, distance);
    // a positive remainder (the reference's else-branch) takes its sign
    // from the direction toward the next DE ray and toward the target
    if (unstable && distance[1] > 0.0) {
        double cpa[3], nxt[3], prv[3];
        collision_at(b, (int64_t)de * F.naz + az, offset[0], cpa);
        collision_at(b, (int64_t)(de + 1) * F.naz + az, offset[0], nxt);
        collision_at(b, (int64_t)(de - 1) * F.naz + az, offset[0], prv);
        double nxt_dir[3], tgt_dir[3];
        nxt_dir[0] = nxt[0] - cpa[0];
        nxt_dir[1] = (nxt[1] - prv[1]) * cpa[0];
        nxt_dir[2] = (nxt[2] - prv[2]) * cpa[0] * wqm::sin(cpa[1]);
        if (fabs(nxt_dir[0]) < 0.01) nxt_dir[0] = 0.0;
        tgt_dir[0] = F.tgt_r[t] - cpa[0];
        tgt_dir[1] = (F.tgt_t[t] - cpa[1]) * cpa[0];
        tgt_dir[2] = (F.tgt_p[t] - cpa[2]) * cpa[0] * wqm::sin(cpa[1]);
        const double dot = inner3(nxt_dir, tgt_dir);
        if (dot < 0.0) {
            distance[1] *= -1.0;
            offset[1] *= -1.0;
        }
    }
}

// spreading_hybrid_gaussian (spreading_hybrid_gaussian.cc:20-571)
// Lane-group helpers of the split build: a candidate's lanes are
// consecutive lanes of one wavefront, starting at a multiple of G.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ bool group_any(bool p, int G) {
    const unsigned long long m = __ballot(p);
    const int base = (int)(threadIdx.x & 63) & ~(G - 1);
    const unsigned long long sel = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << base);
    return (m & sel) != 0ull;
}
__device__ __forceinline__ double group_shfl(double v, int src, int G) {
    const int lane = (int)(threadIdx.x & 63) & ~(G - 1);
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned int lo = (unsigned int)__shfl((int)(unsigned int)(u & 0xffffffffull), lane + src);
    const unsigned int hi = (unsigned int)__shfl((int)(unsigned int)(u >> 32), lane + src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ unsigned long long group_bcast(unsigned long long v, int G) {
    const int base = (int)(threadIdx.x & 63) & ~(G - 1);
    const unsigned int lo = (unsigned int)__shfl((int)(unsigned int)(v & 0xffffffffull), base);
    const unsigned int hi = (unsigned int)__shfl((int)(unsigned int)(v >> 32), base);
    return ((unsigned long long)hi << 32) | lo;
}
#elif defined(WQ3D_EMULATE)
// emulation build: the lanes of a group are host threads in lockstep
// (wq3d_emulate.h); a single thread (the serial build) is its own group
inline bool group_any(bool p, int, int where = __builtin_LINE()) {
    return emu::in_group() ? emu::group_any(p, where) : p;
}
inline unsigned long long group_bcast(unsigned long long v, int, int where = __builtin_LINE()) {
    return emu::in_group() ? emu::group_bcast(v, where) : v;
}
inline double group_shfl(double v, int src, int, int where = __builtin_LINE()) {
    return emu::in_group() ? emu::group_shfl(v, src, where) : v;
}
#else
inline bool group_any(bool p, int) { return p; }
inline unsigned long long group_bcast(unsigned long long v, int) { return v; }
inline double group_shfl(double v, int, int) { return v; }
#endif

// Frequencies one thread evaluates in body_build.  Split: the candidate is
// built by a group of lanes, lane g evaluating frequency 0 (which steers the
// spreading walks) and frequency g, in two register slots.  Not split: one
// thread evaluates every frequency.
template <bool Split>
struct FreqSel {
    static constexpr int kSlots = Split ? 2 : kMaxFreq;
    int nf;
    int fsel;  // Split only: this lane's index in the group; its frequency if below nf
    int G;     // Split only: lanes in the group (all share the widths and neighbourhoods)
    WQ_HD_INLINE int lanes() const { return G; }
    WQ_HD_INLINE bool writes(int k) const { return !Split || (k == 0 ? fsel == 0 : fsel < nf); }
    WQ_HD_INLINE int count() const { return Split ? (fsel > 0 && fsel < nf ? 2 : 1) : nf; }
    WQ_HD_INLINE int freq(int k) const { return Split ? (k == 0 ? 0 : fsel) : k; }
};

template <bool Split>
struct Spread {
    const BuildCtx& b;
    FreqSel<Split> fs;
    double spread[FreqSel<Split>::kSlots];
    double I_de[FreqSel<Split>::kSlots];
    double I_az[FreqSel<Split>::kSlots];
#if defined(WQ3D_EMULATE) && !defined(__HIP_DEVICE_COMPILE__)
    uint32_t dup[kMaxAz / 32];  // the plain set of visited AZ cells, checks AzMarks
#endif
};

// AZ cells an intensity_az walk has visited (the reference's duplicate test,
// spreading_hybrid_gaussian.cc:264-425).  The walk visits az, then az-1,
// az-2, ... and afterwards az+1, az+2, ... (indices modulo max_az), so the
// visited cells are always two cyclic runs, az down to az-(dn-1) and az+1
// up to az+up, and a test is two modular compares instead of a bitset.
struct AzMarks {
    int az, max_az, dn = 0, up = 0;
    WQ_HD_INLINE int cyc(int x) const { return ((x % max_az) + max_az) % max_az; }
    // true if cell a was visited; marks it otherwise (it is the next cell of
    // the down run while up == 0, of the up run after)
    WQ_HD_INLINE bool test_set(int a) {
        if (cyc(az - a) < dn || cyc(a - az - 1) < up) return true;
        if (up == 0 && cyc(az - a) == dn) {
            ++dn;
        } else {
            ++up;
        }
        return false;
    }
};

WQ_HD_INLINE double ray_dist(const BuildCtx& b, int slot, int d1, int a1, int d2,
                                           int a2) {
    const int64_t k1 = at(slot, b.F.N, (int64_t)d1 * b.F.naz + a1);
    const int64_t k2 = at(slot, b.F.N, (int64_t)d2 * b.F.naz + a2);
    return point_distance(b.A.pr[k1], b.A.pt[k1], b.A.pp[k1], b.A.pr[k2], b.A.pt[k2], b.A.pp[k2]);
}

// width_de (spreading_hybrid_gaussian.cc:430-495)
WQ_HD_INLINE double width_de(const BuildCtx& b, int de, int az, const double offset[3]) {
    const FanConst& F = b.F;
    const double u = fabs(offset[0]) / F.dt;
    const double v = fabs(offset[2]) / seq_inc(F.az, az);
    const int max_az = F.naz - 1;
    const int az_wrap = (az + 1 >= max_az) ? 0 : az + 1;
    double L1 = ray_dist(b, b.S.curr, de, az, de + 1, az);
    double length1, length2;
    if (v < 1e-10) {
        length1 = L1;
    } else {
        const double L2 = ray_dist(b, b.S.curr, de, az_wrap, de + 1, az_wrap);
        length1 = (1.0 - v) * L1 + v * L2;
    }
    if (u < 1e-10) return 0.5 * length1;
    const int s2 = (offset[0] < 0.0) ? b.S.prev : b.S.next;
    L1 = ray_dist(b, s2, de, az, de + 1, az);
    if (v < 1e-10) {
        length2 = L1;
    } else {
        const double L2 = ray_dist(b, s2, de, az_wrap + 1, de + 1, az_wrap + 1);
        length2 = (1.0 - v) * L1 + v * L2;
    }
    return 0.5 * ((1.0 - u) * length1 + u * length2);
}

// width_az (spreading_hybrid_gaussian.cc:500-571)
WQ_HD_INLINE double width_az(const BuildCtx& b, int de, int az, const double offset[3]) {
    const FanConst& F = b.F;
    const double u = fabs(offset[0]) / F.dt;
    const double v = fabs(offset[1]) / seq_inc(F.de, de);
    const int max_az = F.naz - 1;
    const int max_de = F.nde - 1;
    const int de_upper = (de + 1 >= max_de) ? max_de - 2 : de;
    const int az_wrap = (az + 1 > max_az) ? 0 : az + 1;
    double L1 = ray_dist(b, b.S.curr, de, az, de, az_wrap);
    double length1, length2;
    if (v < 1e-10 || fabs(v - 1.0) < 1e-10) {
        length1 = L1;
    } else {
        const double L2 = ray_dist(b, b.S.curr, de_upper + 1, az, de_upper + 1, az_wrap);
        length1 = (1.0 - v) * L1 + v * L2;
    }
    if (u < 1e-10) return 0.5 * length1;
    const int s2 = (offset[0] < 0.0) ? b.S.prev : b.S.next;
    L1 = ray_dist(b, s2, de, az, de, az_wrap);
    if (v < 1e-10 || fabs(v - 1.0) < 1e-10) {
        length2 = L1;
    } else {
        const double L2 = ray_dist(b, s2, de_upper + 1, az, de_upper + 1, az_wrap);
        length2 = (1.0 - v) * L1 + v * L2;
    }
    return 0.5 * ((1.0 - u) * length1 + u * length2);
}

// What a spreading walk reads of one cell: its width, the two norms it may
// use and one flag (a virtual ray in DE, an edge in AZ).
struct CellData {
    double w, n0, n1;
    bool flag;
};

// The cells one spreading walk visits, cell k of the walk being cell(k).  A
// split build fetches them L at a time (L = the lanes of the group), lane g
// fetching cell base + g, and hands each out by shuffles, so the walk's
// sequential loop waits on no memory; the data do not depend on the walk's
// state, so fetching a few cells past the walk's end changes nothing.  A
// plain build fetches each cell when it is asked for.
template <bool Split, class CellOf>
struct CellStream {
    CellOf cell_of;  // data of cell k of the walk; any k >= 0 must be safe
    int G, L, g;
    int base = 0;
    CellData c{0.0, 0.0, 0.0, false};
    bool loaded = false;
    WQ_HD_INLINE CellData get(int k) {
        if (!Split) return cell_of(k);
        if (!loaded || k >= base + L) {
            base = k;
            c = cell_of(k + g);
            loaded = true;
        }
        const int src = k - base;
        CellData out;
        out.w = group_shfl(c.w, src, G);
        out.n0 = group_shfl(c.n0, src, G);
        out.n1 = group_shfl(c.n1, src, G);
        out.flag = group_shfl(c.flag ? 1.0 : 0.0, src, G) != 0.0;
        return out;
    }
};

template <bool Split, class CellOf>
WQ_HD_INLINE CellStream<Split, CellOf> cell_stream(const FreqSel<Split>& fs, CellOf fn) {
    return CellStream<Split, CellOf>{fn, fs.G, fs.lanes(), fs.fsel};
}

// Orders the LDS writes of a wavefront's lanes before their reads by other
// lanes of the same wavefront.
WQ_HD_INLINE void wave_sync(int where = __builtin_LINE()) {
#if defined(__HIP_DEVICE_COMPILE__)
    (void)where;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#elif defined(WQ3D_EMULATE)
    if (emu::in_group()) emu::wave_sync(where);
#else
    (void)where;
#endif
}

// Values at the 27 points k = s*9 + nde*3 + naz of a (time, DE, AZ)
// neighbourhood, NVal per point (value v of point k at [v * 27 + k]).  A
// split build with an LDS area computes them a lane at a time (lane g takes
// k = g, g + L, ...) into that area, and every lane then reads them there:
// the returned pointer is the LDS area, which holds them until the group's
// next fill_cube on it (the group's lanes run in one wavefront, so program
// order and the wave barriers order one candidate's reads before the next
// candidate's writes).  Otherwise one thread computes them all into `out`,
// which is returned.
template <int NVal, bool Split, class PointFn>
WQ_HD_INLINE const double* fill_cube(const FreqSel<Split>& fs, double* lds, double* out, PointFn fn) {
    if (Split) {  // a split build always has its group's LDS area
        wave_sync();
        for (int k = fs.fsel; k < 27; k += fs.lanes()) fn(k, lds + k, 27);
        wave_sync();
        return lds;
    }
    for (int k = 0; k < 27; ++k) fn(k, out + k, 27);
    return out;
}

// spreading_hybrid_gaussian::gaussian (.h:148-153), accumulated into out
// (slot k holds frequency s.fs.freq(k))
template <bool Split>
WQ_HD_INLINE void gaussian_add(const Spread<Split>& s, double d, double w, double A, double* out, bool assign) {
    for (int k = 0; k < FreqSel<Split>::kSlots; ++k) {
        if (k >= s.fs.count()) break;
        const double bw = s.spread[k] + 2.0 * 2.0 * w * w;
        const double g = wqm::exp((-0.5 * d * d) / bw) / sqrt(bw) * A;
        out[k] = assign ? g : out[k] + g;
    }
}

WQ_HD_INLINE bool edge_at(const BuildCtx& b, int de, int az) {
    return b.A.edge[at(b.S.curr, b.F.N, (int64_t)de * b.F.naz + az)] != 0;
}

// intensity_de (spreading_hybrid_gaussian.cc:124-258)
template <bool Split>
WQ_HD_INLINE void intensity_de(Spread<Split>& s, int de, int az, const double offset[3], const double distance[3]) {
    const BuildCtx& b = s.b;
    const double THRESHOLD = 1.002305238;
    const double* norm_de = b.F.norm_de;
    int d = de;
    double cell_width = width_de(b, d, az, offset);
    const double initial_width = cell_width;
    const double L = distance[1];
    double cell_dist = L - cell_width;
    gaussian_add(s, cell_dist, cell_width, norm_de[d], s.I_de, true);
    d = de - 1;
    cell_width = width_de(b, d, az, offset);
    cell_dist = L + cell_width;
    gaussian_add(s, cell_dist, cell_width, norm_de[d], s.I_de, false);
    if (s.I_de[0] < 1e-10) return;
    // cells de-2, de-3, ... 0: width, norm_de[d], norm_de[d+1], virtual ray
    auto down = cell_stream(s.fs, [&](int k) {
        const int dd = de - 2 - k;
        if (dd < 0) return CellData{0.0, 0.0, 0.0, false};
        return CellData{width_de(b, dd, az, offset), norm_de[dd], norm_de[dd + 1],
                        edge_at(b, dd + 1, az) && edge_at(b, dd, az)};
    });
    for (d = de - 2; d >= 0; --d) {
        const CellData cd = down.get(de - 2 - d);
        const bool virtual_ray = cd.flag;
        double nn;
        cell_dist += cell_width;
        if (virtual_ray) {
            cell_dist += cell_width;
            nn = cd.n1;
        } else {
            cell_width = cd.w;
            cell_dist += cell_width;
            nn = cd.n0;
        }
        const double old_tl = s.I_de[0];
        gaussian_add(s, cell_dist, cell_width, nn, s.I_de, false);
        if (s.I_de[0] / old_tl < THRESHOLD) break;
        if (virtual_ray) break;
    }
    cell_width = initial_width;
    cell_dist = L - cell_width;
    const int size = b.F.nde - 1;
    // cells de+1, de+2, ... size-1: width, norm_de[d], norm_de[d-1], virtual ray
    auto up = cell_stream(s.fs, [&](int k) {
        const int dd = de + 1 + k;
        if (dd >= size) return CellData{0.0, 0.0, 0.0, false};
        return CellData{width_de(b, dd, az, offset), norm_de[dd], norm_de[dd - 1],
                        edge_at(b, dd + 1, az) && edge_at(b, dd, az)};
    });
    for (d = de + 1; d < size; ++d) {
        const CellData cd = up.get(d - de - 1);
        const bool virtual_ray = cd.flag;
        double nn;
        cell_dist -= cell_width;
        if (virtual_ray) {
            cell_dist -= cell_width;
            nn = cd.n1;
        } else {
            cell_width = cd.w;
            cell_dist -= cell_width;
            nn = cd.n0;
        }
        const double old_tl = s.I_de[0];
        gaussian_add(s, cell_dist, cell_width, nn, s.I_de, false);
        if (s.I_de[0] / old_tl < THRESHOLD) break;
        if (virtual_ray) break;
    }
}

template <bool Split>
WQ_HD_INLINE bool dup_test_set(Spread<Split>& s, AzMarks& m, int a) {
    const bool was = m.test_set(a);
#if defined(WQ3D_EMULATE) && !defined(__HIP_DEVICE_COMPILE__)
    const uint32_t bit = 1u << (a & 31);
    const bool set_was = (s.dup[a >> 5] & bit) != 0;
    s.dup[a >> 5] |= bit;
    wq3d_emulation_check(set_was == was, "AZ visited-cell runs", (double)a);
#endif
    return was;
}

// intensity_az (spreading_hybrid_gaussian.cc:264-425)
template <bool Split>
WQ_HD_INLINE void intensity_az(Spread<Split>& s, int de, int az, const double offset[3], const double distance[3]) {
    const BuildCtx& b = s.b;
    const FanConst& F = b.F;
    const double THRESHOLD = 1.002305238;
    const int naz = F.naz;
    const int max_de = F.nde - 2;
    const int max_az = naz - 1;
    int az_lower, az_upper;
    if (F.az_boundary) {
        az_lower = az_upper = az;
    } else {
        az_lower = 0;
        az_upper = max_az - 1;
    }
#if defined(WQ3D_EMULATE) && !defined(__HIP_DEVICE_COMPILE__)
    for (int w = 0; w < (naz + 31) / 32; ++w) s.dup[w] = 0u;
#endif
    AzMarks marks{az, max_az};
    auto norm = [&](int dd, int aa) {
        return (dd >= max_de) ? F.norm_az[1 * (int64_t)naz + aa] : F.norm_az[(int64_t)dd * naz + aa];
    };
    int a = az;
    dup_test_set(s, marks, a);
    double cell_width = width_az(b, de, a, offset);
    const double initial_width = cell_width;
    const double L = distance[2];
    double cell_dist = L - cell_width;
    gaussian_add(s, cell_dist, cell_width, norm(de, a), s.I_az, true);
    a = (az < 1) ? max_az - 1 : az - 1;
    dup_test_set(s, marks, a);
    cell_width = width_az(b, de, a, offset);
    cell_dist = L + cell_width;
    gaussian_add(s, cell_dist, cell_width, norm(de, a), s.I_az, false);
    if (s.I_az[0] < 1e-10) return;
    if (a < 1) {
        a = max_az - 1;
    } else {
        --a;
    }
    // the walk visits a, a-1, ... wrapping from 0 to max_az-1: width, norm,
    // edge of each
    const int a_down = a;
    auto down = cell_stream(s.fs, [&](int k) {
        const int c = ((a_down - k) % max_az + max_az) % max_az;
        const double n = norm(de, c);
        return CellData{width_az(b, de, c, offset), n, n, edge_at(b, de, c)};
    });
    int kd = 0;
    while ((a % max_az) != az_lower) {
        const CellData cd = down.get(kd++);
        if (dup_test_set(s, marks, a)) break;
        if (cd.flag) break;
        cell_dist += cell_width;
        cell_width = cd.w;
        cell_dist += cell_width;
        const double old_tl = s.I_az[0];
        gaussian_add(s, cell_dist, cell_width, cd.n0, s.I_az, false);
        if (s.I_az[0] / old_tl < THRESHOLD) break;
        if (a == 0) {
            a += max_az - 1;
        } else {
            --a;
        }
    }
    cell_width = initial_width;
    cell_dist = L - cell_width;
    a = az + 1;
    // the walk visits az+1, az+2, ... modulo max_az: width, norm, edge
    auto up = cell_stream(s.fs, [&](int k) {
        const int c = (az + 1 + k) % max_az;
        const double n = norm(de, c);
        return CellData{width_az(b, de, c, offset), n, n, edge_at(b, de, c)};
    });
    int ku = 0;
    while ((a % max_az) != az_upper) {
        if (a == max_az) a = 0;
        const CellData cd = up.get(ku++);
        if (dup_test_set(s, marks, a)) break;
        if (cd.flag) break;
        cell_dist -= cell_width;
        cell_width = cd.w;
        cell_dist -= cell_width;
        const double old_tl = s.I_az[0];
        gaussian_add(s, cell_dist, cell_width, cd.n0, s.I_az, false);
        if (s.I_az[0] / old_tl < THRESHOLD) break;
        ++a;
    }
}

// wvector1::direction(double*, double*) (types/wvector1.cc:79-82)
WQ_HD_INLINE void ndir_angles(const Arrays& A, int64_t k, double& de, double& az) {
    const double nr = A.nr[k], nt = A.nt[k], np = A.np[k];
    de = to_degrees(wqm::atan2(nr, sqrt(nt * nt + np * np)));
    az = to_degrees(wqm::atan2(np, -nt));
}

// wvector1::area of four points of one slot (types/wvector1.cc:151-169):
// the triangles (p1, p2, p3) and (p1, p3, p4) by Heron's formula, for the
// cell p1 = (de, az), p2 = (de+1, az), p3 = (de+1, az+1), p4 = (de, az+1)
WQ_HD_INLINE double cell_area(const BuildCtx& b, int slot, int de, int az) {
    double a = ray_dist(b, slot, de, az, de + 1, az);
    const double bb = ray_dist(b, slot, de, az, de + 1, az + 1);
    double c = ray_dist(b, slot, de + 1, az, de + 1, az + 1);
    const double area1 = 0.25 * sqrt((a + bb + c) * (bb + c - a) * (c + a - bb) * (a + bb - c));
    a = ray_dist(b, slot, de, az, de, az + 1);
    c = ray_dist(b, slot, de, az + 1, de + 1, az + 1);
    const double area2 = 0.25 * sqrt((a + bb + c) * (bb + c - a) * (c + a - bb) * (a + bb - c));
    return area1 + area2;
}

// spreading_ray::intensity (spreading_ray.cc:47-130): the ray tube's
// initial area over its area now (interpolated in time toward prev or
// next), times c(target) / c(source), zero more than four beam widths
// outside the fan; the same for every frequency.  The reference steps to
// the lower cell with a size_t `--az`, which at az = 0 of a fan closing on
// itself indexes outside the front; here that cell is the fan's wrap,
// naz - 2, as the hybrid model takes it.
WQ_HD_INLINE double classic_spread(const BuildCtx& b, int de, int az, const double offset[3], double c_target) {
    const FanConst& F = b.F;
    if (offset[1] < 0.0) --de;
    if (offset[2] < 0.0) az = (az > 0) ? az - 1 : F.naz - 2;
    const double area1 = cell_area(b, b.S.curr, de, az);
    const double area2 = cell_area(b, (offset[0] < 0.0) ? b.S.prev : b.S.next, de, az);
    double u = fabs(offset[0]) / F.dt;
    const double area = (1.0 - u) * area1 + u * area2;
    double loss = F.init_area[(int64_t)de * F.naz + az] * c_target / (area * F.init_c);
    u = fabs(offset[1]) / seq_inc(F.de, de) - 1.0;
    if (u > 4.0) loss *= 0.0;
    u = fabs(offset[2]) / seq_inc(F.az, az) - 1.0;
    if (u > 4.0) loss *= 0.0;
    return loss;
}

// is_closest_ray + build_eigenray (wave_queue.cc:411-702) for candidate c.
// Split: called by the G lanes of a group with fsel = the lane's frequency;
// every lane runs the same control flow (the walks stop on frequency 0,
// which every lane evaluates), so the group stays converged.
template <bool Split>
WQ_HD_INLINE void body_build(int64_t c, int fsel, int G, double* lds, const DEnv& E, const FanConst& F,
                      const Arrays& A, const Slots& S, const Records& Rc, double time, int64_t step_key) {
    const BuildCtx b{E, F, A, S};
    const FreqSel<Split> fs{F.nf, fsel, G};
    constexpr int kSlots = FreqSel<Split>::kSlots;
    {
        const unsigned long long key = Rc.cand[c];
        const int64_t t = (int64_t)(key >> 32);
        const int64_t i = (int64_t)(key & 0xffffffffull);
        const int de = (int)(i / F.naz), az = (int)(i % F.naz);
        const int sl[3] = {S.prev, S.curr, S.next};
        // distances of the neighbourhood, and the edge flags of its rays in
        // the current slot (the second value of the s = 1 points)
        // (and, for compute_offsets, whether any point's counts differ from
        // the centre's)
        const int64_t ic0 = at(S.curr, F.N, i);
        const int s0 = A.srf[ic0], b0 = A.btm[ic0], c0 = A.cst[ic0];
        bool differ = false;
        // the split build keeps both cubes in its group's LDS area (2 x 54
        // values), the serial build in these private arrays
        double DE_own[Split ? 1 : 2 * 27];
        const double(*DE)[3][3][3] = reinterpret_cast<const double(*)[3][3][3]>(
            fill_cube<2>(fs, lds, DE_own, [&](int k, double* out, int stride) {
                const int64_t j = cube_ray(F, de, az, (k / 3) % 3, k % 3);
                out[0] = dist2_slot(b, sl[k / 9], j, t);
                out[stride] = (k / 9 == 1 && A.edge[at(S.curr, F.N, j)] != 0) ? 1.0 : 0.0;
                differ = differ || counters_differ_at(b, k / 9, j, s0, b0, c0);
            }));
        if (Split) differ = group_any(differ, G);
        const double(&D)[3][3][3] = DE[0];
        // the time-CPA pre-test (wave_queue.cc:384-394), then is_closest_ray
        if (!(D[2][1][1] > D[1][1][1]) || D[0][1][1] < D[1][1][1]) return;
        if (!is_closest_ray(b, t, de, az, D, DE[1][1])) return;
        if (!Split || fsel == 0) atom_add(Rc.pass_count, 1ull);
        // build_eigenray (wave_queue.cc:495-702)
        const int64_t ic = at(S.curr, F.N, i);
        if (above_bounce(F, A, ic)) return;
        const double delta[3] = {F.dt, seq_inc(F.de, de), seq_inc(F.az, az)};
        double offset[3], distance[3];
        compute_offsets(b, t, de, az, D, delta, differ, offset, distance);
        Spread<Split> sp{b, fs};
        double ts_c, ts_g[3];
        sound_speed(E.profile, F.tgt_r[t], F.tgt_t[t], F.tgt_p[t], F.R, ts_c, ts_g, false);
        for (int k = 0; k < kSlots; ++k) {
            if (k >= fs.count()) break;
            const double w = (2.0 * M_PI) * ts_c / seq_value(F.freq, fs.freq(k));
            sp.spread[k] = w * w;
        }
        if (F.spreading == WQ3D_SPREADING_CLASSIC_RAY) {
            const double loss = classic_spread(b, de, az, offset, ts_c);
            for (int k = 0; k < kSlots; ++k) {
                sp.I_de[k] = loss;
                sp.I_az[k] = 1.0;
            }
        } else {
            int a = az;
            double no[3] = {offset[0], offset[1], offset[2]};
            if (offset[2] < 0.0) {
                if (az > 0) {
                    a = az - 1;
                } else if (F.az_boundary) {
                    a = F.naz - 2;
                }
                no[2] = offset[2] + seq_inc(F.az, a);
            }
            intensity_de(sp, de, a, no, distance);
            no[2] = offset[2];
            int d = de;
            if (offset[1] < 0.0 && !edge_at(b, de - 1, az)) {
                d = de - 1;
                no[1] = offset[1] + seq_inc(F.de, d);
            }
            intensity_az(sp, d, az, no, distance);
        }
        double inten[kSlots];
        double dt = offset[0] / F.dt;
        const bool fwd = dt >= 0.0;
        if (!fwd) dt = 1.0 + dt;
        bool above = false;
        for (int k = 0; k < kSlots; ++k) {
            if (k >= fs.count()) break;
            const int f = fs.freq(k);
            inten[k] = -10.0 * wqm::log10(smax(sp.I_de[k] * sp.I_az[k], 1e-30));
            if (fwd) {
                const double ac = A.atten[((int64_t)S.curr * F.nf + f) * F.N + i];
                const double an = A.atten[((int64_t)S.next * F.nf + f) * F.N + i];
                inten[k] = inten[k] + ac * (1.0 - dt) + an * dt;
            } else {
                const double ap = A.atten[((int64_t)S.prev * F.nf + f) * F.N + i];
                const double ac = A.atten[((int64_t)S.curr * F.nf + f) * F.N + i];
                inten[k] = inten[k] + ap * (1.0 - dt) + ac * dt;
            }
            if (inten[k] < F.thr_intensity) above = true;
        }
        if (Split) above = group_any(above, G);
        if (!above) return;
        // launch-angle DE and AZ of the neighbourhood's directions
        double VW_own[Split ? 1 : 2 * 27];
        const double(*VW)[3][3][3] = reinterpret_cast<const double(*)[3][3][3]>(
            fill_cube<2>(fs, Split ? lds + 2 * 27 : nullptr, VW_own, [&](int k, double* out, int stride) {
                ndir_angles(A, at(sl[k / 9], F.N, cube_ray(F, de, az, (k / 3) % 3, k % 3)), out[0], out[stride]);
            }));
        const double(&V)[3][3][3] = VW[0];
        const double(&W)[3][3][3] = VW[1];
        double cen, g[3], H[3][3], Ho[3];
        make_taylor_coeff(V, delta, cen, g, H);
        prod3(H, offset, Ho);
        const double target_de = cen + inner3(g, offset) + 0.5 * inner3(offset, Ho);
        make_taylor_coeff(W, delta, cen, g, H);
        prod3(H, offset, Ho);
        const double target_az = cen + inner3(g, offset) + 0.5 * inner3(offset, Ho);
        unsigned long long slot = 0;
        if (!Split || fsel == 0) slot = atom_add(Rc.ray_count, 1ull);
        if (Split) slot = group_bcast(slot, G);
        if ((int64_t)slot >= Rc.ray_cap) {
            if (!Split || fsel == 0) atom_or(Rc.error, 1);
            return;
        }
        if (!Split || fsel == 0) {
            wq3d_eigenray& r = Rc.rays[slot];
            r.step = step_key;
            r.t1 = (int32_t)(t / F.nt2);
            r.t2 = (int32_t)(t % F.nt2);
            r.de = de;
            r.az = az;
            r.travel_time = time + offset[0];
            r.source_de = seq_value(F.de, de) + offset[1];
            r.source_az = seq_value(F.az, az) + offset[2];
            r.target_de = target_de;
            r.target_az = target_az;
            r.surface = A.srf[ic];
            r.bottom = A.btm[ic];
            r.caustic = A.cst[ic];
            r.upper = A.upr[ic];
            r.lower = A.lwr[ic];
        }
        for (int k = 0; k < kSlots; ++k) {
            if (k >= fs.count()) break;
            if (!fs.writes(k)) continue;  // each frequency is written by its own lane
            const int f = fs.freq(k);
            Rc.ray_int[slot * F.nf + f] = inten[k];
            Rc.ray_phase[slot * F.nf + f] = A.phase[((int64_t)S.curr * F.nf + f) * F.N + i];
        }
    }
}

}  // namespace wq3d
