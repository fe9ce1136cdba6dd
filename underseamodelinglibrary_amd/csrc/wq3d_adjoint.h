// wq3d_adjoint.h - discrete adjoint of eigenray travel time and
// transmission loss with respect to the node speeds of a 1-D linear
// sound-speed grid (profile_grid<1>, types/gen_grid.h:248-283), for fans in
// that profile between flat boundaries.
//
// The forward run is the ordinary step pipeline (wave_queue.cc:155-192).
// The backward pass replays it segment by segment from checkpoints of the
// whole fan state, keeps every step's wavefront window of the segment, and
// then runs the adjoint of each step in reverse order, per ray:
//   seeds     for each eigenray of the step, the derivatives of its travel
//             time and its per-frequency intensity (dB) with respect to
//             everything its construction read (the 27-point neighbourhood's
//             target-distance snapshots, the walked rays' positions, the
//             ray's attenuation, the target's sound speed), weighted by the
//             caller's gradients - a tape replay of build_eigenray
//             (wave_queue.cc:495-702) and spreading_hybrid_gaussian
//             (spreading_hybrid_gaussian.cc:124-571), computed during the
//             segment's replay and applied at the step's turn;
//   advance   the adjoint of AB3 + update() + Thorp attenuation
//             (ode_integ.cc:131-186, wave_front.cc:125-178,
//             attenuation_thorp.cc:14-36), hand-written;
//   reflect   for the rays that reflected in the step, a tape replay of the
//             surface/bottom chain (reflection_model.cc:33-360,
//             wave_queue.cc:982-1057);
// and at the end the adjoint of the fan's start (wave_queue.cc:114-150).
// Every tape replay recomputes the forward values with the same
// operations, so it follows the forward run's branches (the eigenray set,
// the reflection steps, compute_offsets' branch, the walks' lengths); each
// replayed travel time, intensity and reflected state is compared with the
// forward run's, bit for bit.
#pragma once

#include "wq3d_ad.h"
#include "wq3d_kernels.h"

namespace wq3d {
namespace adj {

using ad::V;
using ad::val;

// adjoint state fields, each [4 slots][N]; attenuation per frequency last
enum Field {
    kPr, kPt, kPp, kNr, kNt, kNp, kGr, kGt, kGp, kHr, kHt, kHp, kC,
    kSr, kSt, kSp, kSs,  // target-distance snapshot (tr, tt, tp, ts)
    kAtt                 // + frequency
};
constexpr int kStateFields = 13;  // kPr .. kC: what update() and a reflection write

WQ_HD_INLINE int32_t code(int field, int slot) { return field * 4 + slot; }

// ------------------------------------------------------------ number-generic physics
// One ray's state over either number type: the forward functions of
// wq3d_device.h / wq3d_kernels.h written once more as templates, operation
// for operation (tests check the double instantiation against them).
template <class T>
struct RayT {
    T pr, pt, pp, gr, gt, gp, nr, nt, np, hr, ht, hp, c, cr, ct, cp, dist, st;
};

template <class T>
WQ_HD_INLINE RayT<T> fresh_t() {
    RayT<T> w;
    w.pr = w.pt = w.pp = w.gr = w.gt = w.gp = w.nr = w.nt = w.np = T(0.0);
    w.hr = w.ht = w.hp = w.c = w.cr = w.ct = w.cp = w.dist = w.st = T(0.0);
    return w;
}

template <class T>
WQ_HD_INLINE T smax_t(const T& a, const T& b) { return (a < b) ? b : a; }

// node speeds read as plain doubles or as tape inputs
struct NodesD {
    const double* d;
    WQ_HD_INLINE double at(int k) const { return d[k]; }
};
struct NodesV {
    const double* d;
    ad::Tape* t;
    WQ_HD_INLINE V at(int k) const { return ad::input_node(t, d[k], k); }
};

// grid1_linear (wq3d_device.h) over node values of type T
template <class T, class Nodes>
WQ_HD_INLINE T grid1_linear_t(const DGrid& g, const Nodes& nd, const T& rho, T& grho) {
    const DSeq& a = g.ax[0];
    T loc = rho;
    int k;
    if (g.edge[0]) {
        const double lo = a.v[0];
        const double hi = a.v[a.n - 1];
        const double sgn = (a.inc[0] < 0) ? -1.0 : 1.0;
        const double x = val(loc) * sgn;
        if (x <= lo * sgn) {
            loc = T(lo);
            k = 0;
        } else if (x >= hi * sgn) {
            loc = T(hi);
            k = a.n - 2;
        } else {
            k = find_index(a, val(loc));
        }
    } else {
        k = find_index(a, val(loc));
    }
    const T va = nd.at(k);
    const T vb = nd.at(k + 1);
    const double h = seq_inc(a, k);
    const T u = (loc - seq_value(a, k)) / h;
    grho = (vb - va) / h;
    return va * (1.0 - u) + vb * u;
}

// sound_speed for the 1-D linear grid, with profile_model::adjust_speed
template <class T, class Nodes>
WQ_HD_INLINE void sound_speed_t(const DProfile& P, const Nodes& nd, const T& rho, double R, T& c, T& grho) {
    T speed = grid1_linear_t(P.grid, nd, rho, grho);
    if (P.flat_earth) {
        grho = (grho * rho + speed) / R;
        speed = speed * rho / R;
    }
    c = speed;
}

// update_ray (wq3d_device.h)
template <class T, class Nodes>
WQ_HD_INLINE void update_t(const DEnv& E, const Nodes& nd, double R, RayT<T>& w) {
    T g0;
    sound_speed_t(E.profile, nd, w.pr, R, w.c, g0);
    w.cr = g0;
    w.ct = T(0.0);
    w.cp = T(0.0);
    const T dcc_r = w.cr / w.c;
    const T dcc_t = w.ct / w.c;
    const T dcc_p = w.cp / w.c;
    const T st = wqm::sin(w.pt);
    const T cot = wqm::cos(w.pt) / st;
    T c2r = w.c * w.c;
    w.gr = c2r * w.nr;
    c2r = c2r / w.pr;
    w.gt = c2r * w.nt;
    w.gp = c2r / st * w.np;
    w.hr = c2r * (w.nt * w.nt + w.np * w.np) - dcc_r;
    w.ht = -c2r * (w.nr * w.nt - w.np * w.np * cot) - dcc_t / w.pr;
    w.hp = -c2r * (w.np * (w.nr + w.nt * cot)) - dcc_p / (w.pr * st);
    w.st = st;
}

// rk1 / rk2 / rk3 / ab3 (wq3d_device.h)
template <class T, class D>
WQ_HD_INLINE void rk1_t(const D& dt, const RayT<T>& y0, RayT<T>& y1) {
    y1.pr = y0.pr + 0.5 * dt * y0.gr;
    y1.pt = y0.pt + 0.5 * dt * y0.gt;
    y1.pp = y0.pp + 0.5 * dt * y0.gp;
    y1.nr = y0.nr + 0.5 * dt * y0.hr;
    y1.nt = y0.nt + 0.5 * dt * y0.ht;
    y1.np = y0.np + 0.5 * dt * y0.hp;
}
template <class T, class D>
WQ_HD_INLINE void rk2_t(const D& dt, const RayT<T>& y0, const RayT<T>& y1, RayT<T>& y2) {
    y2.pr = y0.pr + dt * (2.0 * y1.gr - y0.gr);
    y2.pt = y0.pt + dt * (2.0 * y1.gt - y0.gt);
    y2.pp = y0.pp + dt * (2.0 * y1.gp - y0.gp);
    y2.nr = y0.nr + dt * (2.0 * y1.hr - y0.hr);
    y2.nt = y0.nt + dt * (2.0 * y1.ht - y0.ht);
    y2.np = y0.np + dt * (2.0 * y1.hp - y0.hp);
}
template <class T, class D>
WQ_HD_INLINE void rk3_t(const D& dt, const RayT<T>& y0, const RayT<T>& y1, const RayT<T>& y2, RayT<T>& y3) {
    const T pr = y0.pr + dt / 6.0 * (y0.gr + 4.0 * y1.gr + y2.gr);
    const T pt = y0.pt + dt / 6.0 * (y0.gt + 4.0 * y1.gt + y2.gt);
    const T pp = y0.pp + dt / 6.0 * (y0.gp + 4.0 * y1.gp + y2.gp);
    const T nr = y0.nr + dt / 6.0 * (y0.hr + 4.0 * y1.hr + y2.hr);
    const T nt = y0.nt + dt / 6.0 * (y0.ht + 4.0 * y1.ht + y2.ht);
    const T np = y0.np + dt / 6.0 * (y0.hp + 4.0 * y1.hp + y2.hp);
    y3.pr = pr;
    y3.pt = pt;
    y3.pp = pp;
    y3.nr = nr;
    y3.nt = nt;
    y3.np = np;
}
template <class T>
WQ_HD_INLINE void ab3_t(double dt, const RayT<T>& y0, const RayT<T>& y1, const RayT<T>& y2, RayT<T>& y3) {
    using ::sqrt;
    const double A2 = 23.0 / 12.0, A1 = 16.0 / 12.0, A0 = 5.0 / 12.0;
    const T dr = dt * (A2 * y2.gr - A1 * y1.gr + A0 * y0.gr);
    const T dth = dt * (A2 * y2.gt - A1 * y1.gt + A0 * y0.gt);
    const T dph = dt * (A2 * y2.gp - A1 * y1.gp + A0 * y0.gp);
    const T a = y2.pr * dth;
    const T b = y2.pr * (wqm::sin(y2.pt) * dph);
    y3.dist = sqrt(dr * dr + a * a + b * b);
    y3.pr = y2.pr + dr;
    y3.pt = y2.pt + dth;
    y3.pp = y2.pp + dph;
    y3.nr = y2.nr + dt * (A2 * y2.hr - A1 * y1.hr + A0 * y0.hr);
    y3.nt = y2.nt + dt * (A2 * y2.ht - A1 * y1.ht + A0 * y0.ht);
    y3.np = y2.np + dt * (A2 * y2.hp - A1 * y1.hp + A0 * y0.hp);
}

// attenuation (Thorp or constant, wq3d_device.h)
template <class T>
WQ_HD_INLINE T attenuation_t(const DProfile& P, double freq, const T& alt, const T& distance) {
    if (P.atten_kind == 1) return P.atten_coeff * distance * freq;
    double F2 = freq;
    F2 = 1e-6 * F2 * F2;
    const double alpha = 1e-3 * (3.3e-3 + F2 * (0.11 / (1.0 + F2) + 44.0 / (4100.0 + F2) + 3.0e-4)) /
                         (1.0 - 5.88264e-6 * 1000.0);
    return distance * alpha * (1.0 + 5.88264e-6 * alt);
}

// reflect_loss and rayleigh_impedance (wq3d_device.h)
template <class T>
WQ_HD_INLINE ad::cx<T> rayleigh_impedance_t(const DBoundary& B, double density, double speed, double att,
                                           const T& angle, ad::cx<T>& cosA, bool shear) {
    const ad::cx<T> c = {T(speed), T(-att * speed)};
    const ad::cx<T> sinA = ad::cscale(c, T(wqm::sin(angle)));
    const ad::cx<T> sA = {sinA.re / B.spd_w, sinA.im / B.spd_w};
    cosA = ad::csqrt(ad::csub(ad::cx<T>{T(1.0), T(0.0)}, ad::cmul(sA, sA)));
    if (shear) return ad::cdiv(ad::cscale(c, T(density)), cosA);
    return ad::cdiv(cosA, ad::cscale(c, T(density)));
}

template <class T>
WQ_HD_INLINE void reflect_loss_t(const DBoundary& B, T angle, T& amp, T& phase) {
    if (B.loss_kind == 0) {
        amp = T(B.loss[0]);
        phase = T(B.loss[1]);
        return;
    }
    if (val(angle) >= M_PI_2) angle = T(M_PI_2 - 1e-10);
    angle = M_PI_2 - angle;
    const T Zw = wqm::cos(angle) / (B.spd_w * B.den_w);
    ad::cx<T> cosAp, cosAs;
    ad::cx<T> Zb = rayleigh_impedance_t(B, B.den_b, B.spd_b, B.att_b, angle, cosAp, false);
    if (B.spd_s != 0.0 || B.att_s != 0.0) {
        const ad::cx<T> Zs = rayleigh_impedance_t(B, B.den_b, B.spd_s, B.att_s, angle, cosAs, true);
        const ad::cx<T> one = {T(1.0), T(0.0)};
        const ad::cx<T> sinAs = ad::csqrt(ad::csub(one, ad::cmul(cosAs, cosAs)));
        const ad::cx<T> cos2As = ad::csub(ad::cmul(ad::cscale(cosAs, T(2.0)), cosAs), one);
        const ad::cx<T> sin2As = ad::cmul(ad::cscale(sinAs, T(2.0)), cosAs);
        const ad::cx<T> t1 = ad::cmul(ad::cmul(ad::cdiv(one, Zb), cos2As), cos2As);
        const ad::cx<T> t2 = ad::cmul(ad::cmul(Zs, sin2As), sin2As);
        Zb = ad::cdiv(one, ad::cadd(t1, t2));
    }
    const ad::cx<T> Zwc = {Zw, T(0.0)};
    const ad::cx<T> Rc = ad::cdiv(ad::csub(Zb, Zwc), ad::cadd(Zb, Zwc));
    amp = -20.0 * wqm::log10(ad::cabs(Rc)) * 1.0;
    phase = ad::carg(Rc) * 1.0;
}

// ------------------------------------------------------------ reflections
// One ray's four slots during detect_reflections (before the rotation):
// s[0] past, s[1] prev, s[2] curr, s[3] next.  amp accumulates the
// reflection losses added to the next slot's attenuation.
template <class T>
struct WinT {
    RayT<T> s[4];
    T amp;
};

// collision_location (wave_queue.cc:982-1057) on the window
template <class T>
WQ_HD_INLINE void collision_t(const FanConst& F, const WinT<T>& w, const T& tw, T pos[3], T nd[3], T& speed) {
    const RayT<T>& p = w.s[1];
    const RayT<T>& c = w.s[2];
    const RayT<T>& n = w.s[3];
    const double time1 = 2.0 * F.dt;
    const double time2 = F.dt * F.dt;
    const T dtime2 = tw * tw;
    {
        const T d1 = (n.c - p.c) / time1;
        const T d2 = (n.c + p.c - 2.0 * c.c) / time2;
        speed = c.c + d1 * tw + 0.5 * d2 * dtime2;
    }
    const T* P3[3] = {&p.pr, &p.pt, &p.pp};
    const T* C3[3] = {&c.pr, &c.pt, &c.pp};
    const T* N3[3] = {&n.pr, &n.pt, &n.pp};
    for (int k = 0; k < 3; ++k) {
        const T d1 = (*N3[k] - *P3[k]) / time1;
        const T d2 = (*N3[k] + *P3[k] - 2.0 * *C3[k]) / time2;
        pos[k] = *C3[k] + d1 * tw + 0.5 * d2 * dtime2;
    }
    const T* Pn[3] = {&p.nr, &p.nt, &p.np};
    const T* Cn[3] = {&c.nr, &c.nt, &c.np};
    const T* Nn[3] = {&n.nr, &n.nt, &n.np};
    for (int k = 0; k < 3; ++k) {
        const T d1 = (*Nn[k] - *Pn[k]) / time1;
        const T d2 = (*Nn[k] + *Pn[k] - 2.0 * *Cn[k]) / time2;
        nd[k] = *Cn[k] + d1 * tw + 0.5 * d2 * dtime2;
    }
}

// reflection_copy: the fields copy_ray writes
template <class T>
WQ_HD_INLINE void copy_t(RayT<T>& dst, const RayT<T>& w) {
    dst.pr = w.pr; dst.pt = w.pt; dst.pp = w.pp;
    dst.gr = w.gr; dst.gt = w.gt; dst.gp = w.gp;
    dst.nr = w.nr; dst.nt = w.nt; dst.np = w.np;
    dst.hr = w.hr; dst.ht = w.ht; dst.hp = w.hp;
    dst.cr = w.cr; dst.ct = w.ct; dst.cp = w.cp;
    dst.c = w.c;
    dst.dist = w.dist;
}

// reflection_reinit (wq3d_kernels.h)
template <class T, class Nodes>
WQ_HD void reinit_t(const DEnv& E, const FanConst& F, const Nodes& nd, WinT<T>& w, const T& time_water,
                    const T pos[3], const T dir[3]) {
    const double R = F.R;
    const double dt = F.dt;
    RayT<T> temp = fresh_t<T>();
    temp.pr = pos[0];
    temp.pt = pos[1];
    temp.pp = pos[2];
    temp.nr = dir[0];
    temp.nt = dir[1];
    temp.np = dir[2];
    update_t(E, nd, R, temp);
    RayT<T> tn = fresh_t<T>(), tp = fresh_t<T>(), tc = fresh_t<T>(), tq = fresh_t<T>();
    const T mtw = -time_water;
    rk1_t(mtw, temp, tn);
    update_t(E, nd, R, tn);
    rk2_t(mtw, temp, tn, tp);
    update_t(E, nd, R, tp);
    rk3_t(mtw, temp, tn, tp, tc);
    update_t(E, nd, R, tc);
    copy_t(w.s[2], tc);
    rk1_t(-dt, tc, tn);
    update_t(E, nd, R, tn);
    rk2_t(-dt, tc, tn, tp);
    update_t(E, nd, R, tp);
    rk3_t(-dt, tc, tn, tp, tq);
    update_t(E, nd, R, tq);
    copy_t(w.s[1], tq);
    rk1_t(-dt, tq, tn);
    update_t(E, nd, R, tn);
    rk2_t(-dt, tq, tn, temp);
    rk3_t(-dt, tq, tn, temp, tp);
    update_t(E, nd, R, tp);
    copy_t(w.s[0], tp);
    ab3_t(dt, tp, tq, tc, tn);
    update_t(E, nd, R, tn);
    copy_t(w.s[3], tn);
}

// surface_reflection (wq3d_kernels.h) on the window
template <class T, class Nodes>
WQ_HD bool surface_t(const DEnv& E, const FanConst& F, const Nodes& nd, WinT<T>& w) {
    using ::sqrt;
    const RayT<T>& cur = w.s[2];
    T c = cur.c;
    const T d = c * c * cur.nr;
    const T time_water = (val(d) == 0.0) ? T(0.0) : -(cur.pr - F.R) / d;
    T pos[3], dir[3];
    collision_t(F, w, time_water, pos, dir, c);
    const T grazing = wqm::atan2(cur.nr, sqrt(cur.nt * cur.nt + cur.np * cur.np));
    if (val(grazing) <= 0.0) return false;
    T amp, ph;
    reflect_loss_t(E.surface, grazing, amp, ph);
    w.amp = w.amp + amp;
    dir[0] = -dir[0];
    reinit_t(E, F, nd, w, time_water, pos, dir);
    return true;
}

// bottom_reflection (wq3d_kernels.h) on the window, flat bottom
template <class T, class Nodes>
WQ_HD bool bottom_t(const DEnv& E, const FanConst& F, const Nodes& nd, WinT<T>& w, const T& depth) {
    using ::sqrt;
    const double MIN_REFLECT = 6.0;
    const double TOO_SHALLOW = 300.0 * F.dt;
    const RayT<T>& cur = w.s[2];
    T pos[3] = {cur.pr, cur.pt, cur.pp};
    T dir[3] = {cur.nr, cur.nt, cur.np};
    T c = cur.c;
    T c2 = c * c;
    double bottom_rho = E.bottom.rho;
    double nrm[3] = {E.bottom.normal_rho, 0.0, 0.0};
    T height_water = pos[0] - bottom_rho;
    const bool shallow = (F.R - bottom_rho) < TOO_SHALLOW;
    if (shallow) {
        const double Nn = ::sqrt(nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        nrm[0] = 0.0;
        nrm[1] = nrm[1] / Nn;
        nrm[2] = nrm[2] / Nn;
    }
    dir[0] = c2 * dir[0];
    dir[1] = c2 * dir[1];
    dir[2] = c2 * dir[2];
    T dot_full = nrm[0] * dir[0] + nrm[1] * dir[1] + nrm[2] * dir[2];
    T max_dot = -smax_t(T(MIN_REFLECT), (height_water + depth) * nrm[0]);
    if (dot_full >= max_dot) dot_full = max_dot;
    const T dot_water = -height_water * nrm[0];
    const T time_water = smax_t(T(0.0), dot_water / dot_full);
    collision_t(F, w, time_water, pos, dir, c);
    bottom_rho = E.bottom.rho;
    nrm[0] = E.bottom.normal_rho;
    nrm[1] = 0.0;
    nrm[2] = 0.0;
    if (shallow) {
        const double Nn = ::sqrt(nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        nrm[0] = 0.0;
        nrm[1] = nrm[1] / Nn;
        nrm[2] = nrm[2] / Nn;
    }
    c2 = c * c;
    height_water = pos[0] - bottom_rho;
    dir[0] = c2 * dir[0];
    dir[1] = c2 * dir[1];
    dir[2] = c2 * dir[2];
    dot_full = nrm[0] * dir[0] + nrm[1] * dir[1] + nrm[2] * dir[2];
    max_dot = -smax_t(T(MIN_REFLECT), (height_water + depth) * nrm[0]);
    if (dot_full >= max_dot) dot_full = max_dot;
    T grazing;
    if (val(dot_full / c) >= 1.0) {
        grazing = T(-M_PI_2);
    } else if (val(dot_full / c) <= -1.0) {
        grazing = T(M_PI_2);
    } else {
        grazing = wqm::asin(-dot_full / c);
    }
    T amp, ph;
    reflect_loss_t(E.bottom, grazing, amp, ph);
    w.amp = w.amp + amp;
    dot_full *= 2.0;
    dir[0] = dir[0] - dot_full * nrm[0];
    dir[1] = dir[1] - dot_full * nrm[1];
    dir[2] = dir[2] - dot_full * nrm[2];
    const T Nn = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]) * c;
    dir[0] = dir[0] / Nn;
    dir[1] = dir[1] / Nn;
    dir[2] = dir[2] / Nn;
    reinit_t(E, F, nd, w, time_water, pos, dir);
    return true;
}

// the reflection part of body_bounce (the surface <-> bottom chain)
template <class T, class Nodes>
WQ_HD bool bounce_chain_t(const DEnv& E, const FanConst& F, const Nodes& nd, WinT<T>& w) {
    auto try_surface = [&]() -> bool {
        if (!(w.s[3].pr - F.R > 0.0)) return false;
        return surface_t(E, F, nd, w);
    };
    auto try_bottom = [&]() -> bool {
        const T depth = E.bottom.rho - w.s[3].pr;
        if (!(depth > 0.0)) return false;
        return bottom_t(E, F, nd, w, depth);
    };
    bool bottom_next;
    if (try_surface()) {
        bottom_next = true;
    } else {
        try_bottom();
        bottom_next = false;
    }
    for (int n = 0; n < kMaxBounceChain; ++n) {
        const bool hit = bottom_next ? try_bottom() : try_surface();
        if (!hit) return true;
        bottom_next = !bottom_next;
    }
    return false;
}

// ------------------------------------------------------------ eigenray construction
// target_distance2, point_distance, make_taylor_coeff, offsets_core,
// compute_offsets, the Gaussian walks (wq3d_device.h, wq3d_kernels.h), one
// thread per eigenray.
template <class T>
WQ_HD_INLINE T target_distance2_t(const T& rho, const T& theta, const T& phi, const T& st, double Rt, double Tt,
                                  double Pt, double St) {
    using ::fabs;
    const T h = 0.5 * (theta - Tt);
    const T g = 0.5 * (phi - Pt);
    return fabs(rho * rho + Rt * Rt - 2.0 * Rt * (rho * (1.0 - 2.0 * (h * h + St * (st * (g * g))))));
}

template <class T>
WQ_HD_INLINE T point_distance_t(const T& r1, const T& t1, const T& p1, const T& r2, const T& t2, const T& p2) {
    using ::fabs;
    using ::sqrt;
    const T sin_theta = wqm::sin(0.5 * (t1 - t2));
    const T sin_phi = wqm::sin(0.5 * (p1 - p2));
    const T dotnorm = 1.0 - 2.0 * (sin_theta * sin_theta + wqm::sin(t1) * wqm::sin(t2) * sin_phi * sin_phi);
    const T dot = r1 * r2 * dotnorm;
    return sqrt(fabs(r1 * r1 + r2 * r2 - 2.0 * dot));
}

template <class T>
WQ_HD_INLINE void make_taylor_coeff_t(const T v[3][3][3], const double delta[3], T& center, T g[3], T H[3][3]) {
    const double d0 = 2.0 * delta[0];
    const double d1 = 2.0 * delta[1];
    const double d2 = 2.0 * delta[2];
    center = v[1][1][1];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) H[a][c] = T(0.0);
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

template <class T>
WQ_HD_INLINE void prod3_t(const T M[3][3], const T x[3], T y[3]) {
    for (int r = 0; r < 3; ++r) {
        T t = T(0.0);
        for (int k = 0; k < 3; ++k) t += M[r][k] * x[k];
        y[r] = t;
    }
}
template <class T>
WQ_HD_INLINE T norm3_t(const T g[3]) {
    using ::fabs;
    using ::sqrt;
    T t = T(0.0);
    for (int k = 0; k < 3; ++k) {
        const T u = fabs(g[k]);
        t += u * u;
    }
    return sqrt(t);
}

template <class T>
WQ_HD_INLINE bool offsets_core_t(const T D[3][3][3], const double delta[3], bool counters_differ, T offset[3],
                                 T distance[3]) {
    using ::copysign;
    using ::fabs;
    using ::sqrt;
    bool unstable = counters_differ;
    T center, g[3], H[3][3], inv[3][3];
    make_taylor_coeff_t(D, delta, center, g, H);
    T determinant = smax_t(T(1e-6), H[0][0] * (H[1][1] * H[2][2] - H[1][2] * H[2][1]) +
                                        H[0][1] * (H[1][2] * H[2][0] - H[1][0] * H[2][2]) +
                                        H[0][2] * (H[1][0] * H[2][1] - H[2][1] * H[2][0]));
    unstable = unstable || determinant < norm3_t(g);
    for (int n = 0; n < 3; ++n) {
        const T h = smax_t(T(1e-10), H[n][n]);
        offset[n] = -g[n] / h;
        if (val(fabs(offset[n] / delta[n])) > 0.5) {
            unstable = true;
            offset[n] = copysign(T(0.5 * delta[n]), offset[n]);
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
        const T ng[3] = {-g[0], -g[1], -g[2]};
        prod3_t(inv, ng, offset);
        unstable = unstable || val(fabs(offset[1])) >= 0.5 * ::fabs(delta[1]);
    } else {
        const T save_offset = offset[1];
        g[1] = T(0.0);
        determinant = H[0][0] * H[2][2] - H[0][2] * H[2][0];
        if (determinant >= norm3_t(g)) {
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) inv[a][c] = T(0.0);
            inv[0][0] = H[2][2];
            inv[0][2] = -H[0][2];
            inv[1][1] = T(1.0);
            inv[2][0] = -H[2][0];
            inv[2][2] = H[0][0];
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) inv[a][c] /= determinant;
            const T ng[3] = {-g[0], -g[1], -g[2]};
            prod3_t(inv, ng, offset);
            offset[1] = save_offset;
        }
    }
    for (int n = 0; n < 3; ++n) {
        distance[n] = -g[n] * offset[n] - 0.5 * H[n][n] * offset[n] * offset[n];
        distance[n] = sqrt(smax_t(T(0.0), distance[n]));
        if (val(offset[n]) < 0.0) distance[n] *= -1.0;
    }
    if (unstable) {
        distance[1] = center - distance[0] * distance[0] - distance[2] * distance[2];
        if (val(distance[1]) <= 0.0) {
            distance[1] = T(0.0);
            offset[1] = T(0.0);
        } else {
            distance[1] = sqrt(distance[1]);
            offset[1] = fabs(offset[1]);
        }
    }
    return unstable;
}

// point_distance and target_distance2 on tape values, each recorded as one
// combined value of its inputs with partials derived from its formula (the
// value itself is the double function's, bit for bit)
WQ_HD_INLINE V point_distance_v(const V& r1, const V& t1, const V& p1, const V& r2, const V& t2, const V& p2) {
    const double L = point_distance(r1.v, t1.v, p1.v, r2.v, t2.v, p2.v);
    const double ht = 0.5 * (t1.v - t2.v), hp = 0.5 * (p1.v - p2.v);
    const double st = wqm::sin(ht), sp = wqm::sin(hp);
    const double s1 = wqm::sin(t1.v), s2 = wqm::sin(t2.v);
    const double dn = 1.0 - 2.0 * (st * st + s1 * s2 * sp * sp);
    const double q = r1.v * r1.v + r2.v * r2.v - 2.0 * (r1.v * r2.v * dn);
    const double dq = L > 0.0 ? (q < 0.0 ? -0.5 : 0.5) / L : 0.0;   // dL / dq
    const double ddn = -2.0 * r1.v * r2.v * dq;                      // dL / d dotnorm
    const double dst = -4.0 * st * ddn, dsp = -4.0 * s1 * s2 * sp * ddn;
    const double cht = 0.5 * wqm::cos(ht), chp = 0.5 * wqm::cos(hp);
    const double x[6] = {(2.0 * r1.v - 2.0 * r2.v * dn) * dq,
                         dst * cht - 2.0 * wqm::cos(t1.v) * s2 * sp * sp * ddn,
                         dsp * chp,
                         (2.0 * r2.v - 2.0 * r1.v * dn) * dq,
                         -dst * cht - 2.0 * s1 * wqm::cos(t2.v) * sp * sp * ddn,
                         -dsp * chp};
    const V in[6] = {r1, t1, p1, r2, t2, p2};
    return ad::combine<6>(L, in, x);
}

WQ_HD_INLINE V target_distance2_v(const V& rho, const V& theta, const V& phi, const V& st, double Rt, double Tt,
                                  double Pt, double St) {
    const double D = target_distance2(rho.v, theta.v, phi.v, st.v, Rt, Tt, Pt, St);
    const double h = 0.5 * (theta.v - Tt);
    const double g = 0.5 * (phi.v - Pt);
    const double inner = 1.0 - 2.0 * (h * h + St * (st.v * (g * g)));
    const double q = rho.v * rho.v + Rt * Rt - 2.0 * Rt * (rho.v * inner);
    const double sq = q < 0.0 ? -1.0 : 1.0;                          // dD / dq
    const double dinner = -2.0 * Rt * rho.v * sq;
    const double x[4] = {(2.0 * rho.v - 2.0 * Rt * inner) * sq, dinner * (-4.0 * h) * 0.5,
                         dinner * (-4.0 * St * st.v * g) * 0.5, dinner * (-2.0 * St * g * g)};
    const V in[4] = {rho, theta, phi, st};
    return ad::combine<4>(D, in, x);
}

// What an eigenray's construction reads from the fan, as tape inputs (or
// plain values): target-distance snapshots, positions and attenuations of
// (slot, ray), each tagged with its adjoint field.
struct Reader {
    const FanConst& F;
    const Arrays& A;
    ad::Tape* t;
    WQ_HD_INLINE V get(const double* arr, int field, int slot, int64_t ray) const {
        return ad::input_state(t, arr[at(slot, F.N, ray)], code(field, slot), ray);
    }
    WQ_HD_INLINE V atten(int slot, int f, int64_t ray) const {
        return ad::input_state(t, A.atten[((int64_t)slot * F.nf + f) * F.N + ray], code(kAtt + f, slot), ray);
    }
};

WQ_HD_INLINE V ray_dist_v(const Reader& rd, int slot, int d1, int a1, int d2, int a2) {
    const int64_t i1 = (int64_t)d1 * rd.F.naz + a1;
    const int64_t i2 = (int64_t)d2 * rd.F.naz + a2;
    return point_distance_v(rd.get(rd.A.pr, kPr, slot, i1), rd.get(rd.A.pt, kPt, slot, i1),
                            rd.get(rd.A.pp, kPp, slot, i1), rd.get(rd.A.pr, kPr, slot, i2),
                            rd.get(rd.A.pt, kPt, slot, i2), rd.get(rd.A.pp, kPp, slot, i2));
}

// width_de / width_az (wq3d_kernels.h)
WQ_HD_INLINE V width_de_v(const Reader& rd, const Slots& S, int de, int az, const V offset[3]) {
    using ad::fabs;
    const FanConst& F = rd.F;
    const V u = fabs(offset[0]) / F.dt;
    const V v = fabs(offset[2]) / seq_inc(F.az, az);
    const int max_az = F.naz - 1;
    const int az_wrap = (az + 1 >= max_az) ? 0 : az + 1;
    V L1 = ray_dist_v(rd, S.curr, de, az, de + 1, az);
    V length1, length2;
    if (val(v) < 1e-10) {
        length1 = L1;
    } else {
        const V L2 = ray_dist_v(rd, S.curr, de, az_wrap, de + 1, az_wrap);
        length1 = (1.0 - v) * L1 + v * L2;
    }
    if (val(u) < 1e-10) return 0.5 * length1;
    const int s2 = (val(offset[0]) < 0.0) ? S.prev : S.next;
    L1 = ray_dist_v(rd, s2, de, az, de + 1, az);
    if (val(v) < 1e-10) {
        length2 = L1;
    } else {
        const V L2 = ray_dist_v(rd, s2, de, az_wrap + 1, de + 1, az_wrap + 1);
        length2 = (1.0 - v) * L1 + v * L2;
    }
    return 0.5 * ((1.0 - u) * length1 + u * length2);
}

WQ_HD_INLINE V width_az_v(const Reader& rd, const Slots& S, int de, int az, const V offset[3]) {
    using ad::fabs;
    const FanConst& F = rd.F;
    const V u = fabs(offset[0]) / F.dt;
    const V v = fabs(offset[1]) / seq_inc(F.de, de);
    const int max_az = F.naz - 1;
    const int max_de = F.nde - 1;
    const int de_upper = (de + 1 >= max_de) ? max_de - 2 : de;
    const int az_wrap = (az + 1 > max_az) ? 0 : az + 1;
    V L1 = ray_dist_v(rd, S.curr, de, az, de, az_wrap);
    V length1, length2;
    if (val(v) < 1e-10 || ::fabs(val(v) - 1.0) < 1e-10) {
        length1 = L1;
    } else {
        const V L2 = ray_dist_v(rd, S.curr, de_upper + 1, az, de_upper + 1, az_wrap);
        length1 = (1.0 - v) * L1 + v * L2;
    }
    if (val(u) < 1e-10) return 0.5 * length1;
    const int s2 = (val(offset[0]) < 0.0) ? S.prev : S.next;
    L1 = ray_dist_v(rd, s2, de, az, de, az_wrap);
    if (val(v) < 1e-10 || ::fabs(val(v) - 1.0) < 1e-10) {
        length2 = L1;
    } else {
        const V L2 = ray_dist_v(rd, s2, de_upper + 1, az, de_upper + 1, az_wrap);
        length2 = (1.0 - v) * L1 + v * L2;
    }
    return 0.5 * ((1.0 - u) * length1 + u * length2);
}

struct SpreadV {
    int nf;
    V spread[kMaxFreq];
    V I_de[kMaxFreq];
    V I_az[kMaxFreq];
};

WQ_HD_INLINE void gaussian_add_v(const SpreadV& s, const V& d, const V& w, double A, V* out, bool assign) {
    for (int k = 0; k < s.nf; ++k) {
        // gaussian_add's value, and its partials in d, w and the spread
        const double bw = s.spread[k].v + 2.0 * 2.0 * w.v * w.v;
        const double gv = wqm::exp((-0.5 * d.v * d.v) / bw) / ::sqrt(bw) * A;
        const double dbw = gv * (0.5 * d.v * d.v / (bw * bw) - 0.5 / bw);
        const double x[3] = {-gv * d.v / bw, dbw * 8.0 * w.v, dbw};
        const V in[3] = {d, w, s.spread[k]};
        const V g = ad::combine<3>(gv, in, x);
        out[k] = assign ? g : out[k] + g;
    }
}

// intensity_de (wq3d_kernels.h), the serial walk
WQ_HD_INLINE void intensity_de_v(const Reader& rd, const Slots& S, SpreadV& s, int de, int az, const V offset[3],
                                 const V distance[3]) {
    const FanConst& F = rd.F;
    const double THRESHOLD = 1.002305238;
    const double* norm_de = F.norm_de;
    auto edge = [&](int d, int a) { return rd.A.edge[at(S.curr, F.N, (int64_t)d * F.naz + a)] != 0; };
    int d = de;
    V cell_width = width_de_v(rd, S, d, az, offset);
    const V initial_width = cell_width;
    const V L = distance[1];
    V cell_dist = L - cell_width;
    gaussian_add_v(s, cell_dist, cell_width, norm_de[d], s.I_de, true);
    d = de - 1;
    cell_width = width_de_v(rd, S, d, az, offset);
    cell_dist = L + cell_width;
    gaussian_add_v(s, cell_dist, cell_width, norm_de[d], s.I_de, false);
    if (val(s.I_de[0]) < 1e-10) return;
    for (d = de - 2; d >= 0; --d) {
        const bool virtual_ray = edge(d + 1, az) && edge(d, az);
        double nn;
        cell_dist += cell_width;
        if (virtual_ray) {
            cell_dist += cell_width;
            nn = norm_de[d + 1];
        } else {
            cell_width = width_de_v(rd, S, d, az, offset);
            cell_dist += cell_width;
            nn = norm_de[d];
        }
        const double old_tl = val(s.I_de[0]);
        gaussian_add_v(s, cell_dist, cell_width, nn, s.I_de, false);
        if (val(s.I_de[0]) / old_tl < THRESHOLD) break;
        if (virtual_ray) break;
    }
    cell_width = initial_width;
    cell_dist = L - cell_width;
    const int size = F.nde - 1;
    for (d = de + 1; d < size; ++d) {
        const bool virtual_ray = edge(d + 1, az) && edge(d, az);
        double nn;
        cell_dist -= cell_width;
        if (virtual_ray) {
            cell_dist -= cell_width;
            nn = norm_de[d - 1];
        } else {
            cell_width = width_de_v(rd, S, d, az, offset);
            cell_dist -= cell_width;
            nn = norm_de[d];
        }
        const double old_tl = val(s.I_de[0]);
        gaussian_add_v(s, cell_dist, cell_width, nn, s.I_de, false);
        if (val(s.I_de[0]) / old_tl < THRESHOLD) break;
        if (virtual_ray) break;
    }
}

// intensity_az (wq3d_kernels.h), the serial walk
WQ_HD_INLINE void intensity_az_v(const Reader& rd, const Slots& S, SpreadV& s, int de, int az, const V offset[3],
                                 const V distance[3]) {
    const FanConst& F = rd.F;
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
    AzMarks marks{az, max_az};
    auto norm = [&](int dd, int aa) {
        return (dd >= max_de) ? F.norm_az[1 * (int64_t)naz + aa] : F.norm_az[(int64_t)dd * naz + aa];
    };
    auto edge = [&](int d, int a) { return rd.A.edge[at(S.curr, F.N, (int64_t)d * F.naz + a)] != 0; };
    int a = az;
    marks.test_set(a);
    V cell_width = width_az_v(rd, S, de, a, offset);
    const V initial_width = cell_width;
    const V L = distance[2];
    V cell_dist = L - cell_width;
    gaussian_add_v(s, cell_dist, cell_width, norm(de, a), s.I_az, true);
    a = (az < 1) ? max_az - 1 : az - 1;
    marks.test_set(a);
    cell_width = width_az_v(rd, S, de, a, offset);
    cell_dist = L + cell_width;
    gaussian_add_v(s, cell_dist, cell_width, norm(de, a), s.I_az, false);
    if (val(s.I_az[0]) < 1e-10) return;
    if (a < 1) {
        a = max_az - 1;
    } else {
        --a;
    }
    while ((a % max_az) != az_lower) {
        const int c = ((a % max_az) + max_az) % max_az;
        if (marks.test_set(a)) break;
        if (edge(de, c)) break;
        cell_dist += cell_width;
        cell_width = width_az_v(rd, S, de, c, offset);
        cell_dist += cell_width;
        const double old_tl = val(s.I_az[0]);
        gaussian_add_v(s, cell_dist, cell_width, norm(de, c), s.I_az, false);
        if (val(s.I_az[0]) / old_tl < THRESHOLD) break;
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
        const int c = a % max_az;
        if (marks.test_set(a)) break;
        if (edge(de, c)) break;
        cell_dist -= cell_width;
        cell_width = width_az_v(rd, S, de, c, offset);
        cell_dist -= cell_width;
        const double old_tl = val(s.I_az[0]);
        gaussian_add_v(s, cell_dist, cell_width, norm(de, c), s.I_az, false);
        if (val(s.I_az[0]) / old_tl < THRESHOLD) break;
        ++a;
    }
}

// is_closest_ray + build_eigenray (body_build) for one eigenray of the
// current step, on tape values: travel time T and intensities I[f] (dB).
// Returns false if the pair does not make an eigenray any more.
WQ_HD bool eigenray_v(const DEnv& E, const FanConst& F, const Arrays& A, const Slots& S, ad::Tape* tp,
                      int64_t t, int64_t i, double time, V& T, V* I) {
    const BuildCtx b{E, F, A, S};
    const Reader rd{F, A, tp};
    const int de = (int)(i / F.naz), az = (int)(i % F.naz);
    const int sl[3] = {S.prev, S.curr, S.next};
    const int64_t ic0 = at(S.curr, F.N, i);
    const int s0 = A.srf[ic0], b0 = A.btm[ic0], c0 = A.cst[ic0];
    bool differ = false;
    V D[3][3][3];
    double Dv[3][3][3];
    double edge_cur[3][3];
    for (int k = 0; k < 27; ++k) {
        const int s = k / 9, nde = (k / 3) % 3, naz = k % 3;
        const int64_t j = cube_ray(F, de, az, nde, naz);
        const int slot = sl[s];
        D[s][nde][naz] = target_distance2_v(rd.get(A.tr, kSr, slot, j), rd.get(A.tt, kSt, slot, j),
                                            rd.get(A.tp, kSp, slot, j), rd.get(A.ts, kSs, slot, j), F.tgt_r[t],
                                            F.tgt_t[t], F.tgt_p[t], F.tgt_s[t]);
        Dv[s][nde][naz] = D[s][nde][naz].v;
        if (s == 1) edge_cur[nde][naz] = A.edge[at(S.curr, F.N, j)] != 0 ? 1.0 : 0.0;
        differ = differ || counters_differ_at(b, s, j, s0, b0, c0);
    }
    if (!(Dv[2][1][1] > Dv[1][1][1]) || Dv[0][1][1] < Dv[1][1][1]) return false;
    if (!is_closest_ray(b, t, de, az, Dv, edge_cur)) return false;
    if (above_bounce(F, A, ic0)) return false;
    const double delta[3] = {F.dt, seq_inc(F.de, de), seq_inc(F.az, az)};
    V offset[3], distance[3];
    const bool unstable = offsets_core_t(D, delta, differ, offset, distance);
    if (unstable && distance[1].v > 0.0) {
        double cpa[3], nxt[3], prv[3];
        collision_at(b, (int64_t)de * F.naz + az, offset[0].v, cpa);
        collision_at(b, (int64_t)(de + 1) * F.naz + az, offset[0].v, nxt);
        collision_at(b, (int64_t)(de - 1) * F.naz + az, offset[0].v, prv);
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
    // spreading_hybrid_gaussian at the target's sound speed
    SpreadV sp;
    sp.nf = F.nf;
    {
        V ts_c, g0;
        const NodesV nodes{E.profile.grid.data, tp};
        sound_speed_t(E.profile, nodes, V(F.tgt_r[t]), F.R, ts_c, g0);
        for (int f = 0; f < F.nf; ++f) {
            const V w = (2.0 * M_PI) * ts_c / seq_value(F.freq, f);
            sp.spread[f] = w * w;
        }
    }
    {
        int a = az;
        V no[3] = {offset[0], offset[1], offset[2]};
        if (offset[2].v < 0.0) {
            if (az > 0) {
                a = az - 1;
            } else if (F.az_boundary) {
                a = F.naz - 2;
            }
            no[2] = offset[2] + seq_inc(F.az, a);
        }
        intensity_de_v(rd, S, sp, de, a, no, distance);
        no[2] = offset[2];
        int d = de;
        if (offset[1].v < 0.0 && !edge_at(b, de - 1, az)) {
            d = de - 1;
            no[1] = offset[1] + seq_inc(F.de, d);
        }
        intensity_az_v(rd, S, sp, d, az, no, distance);
    }
    V dt = offset[0] / F.dt;
    const bool fwd = dt.v >= 0.0;
    if (!fwd) dt = 1.0 + dt;
    for (int f = 0; f < F.nf; ++f) {
        I[f] = -10.0 * wqm::log10(smax_t(sp.I_de[f] * sp.I_az[f], V(1e-30)));
        if (fwd) {
            I[f] = I[f] + rd.atten(S.curr, f, i) * (1.0 - dt) + rd.atten(S.next, f, i) * dt;
        } else {
            I[f] = I[f] + rd.atten(S.prev, f, i) * (1.0 - dt) + rd.atten(S.curr, f, i) * dt;
        }
    }
    T = time + offset[0];
    return true;
}

// ------------------------------------------------------------ hand-written step adjoint
// The forward step of one ray, AB3 (ode_integ.cc:131-186) + update()
// (wave_front.cc:125-178) on the 1-D linear grid + attenuation
// (attenuation_thorp.cc:14-36, attenuation_constant.cc:50-62), and its
// adjoint: given the adjoints of the new slot's outputs, it adds the
// adjoints of the inputs (the three older slots' gradients, the current
// slot's position and direction, the current attenuation) and the node
// speeds'.
struct StepIn {
    double g0[3], h0[3], g1[3], h1[3], g2[3], h2[3];  // past, prev, curr gradients
    double p2[3], n2[3];                                 // curr position, direction
};

struct StepBar {
    // adjoints of the new slot: position, direction, gradients, sound speed,
    // snapshot (r, theta, phi, sin theta); atten per frequency
    double p[3], n[3], g[3], h[3], c, s[4];
    double att[kMaxFreq];
};

struct StepGrad {
    double g0[3], h0[3], g1[3], h1[3], g2[3], h2[3], p2[3], n2[3];
    int node;          // first of the two node speeds read (k, k + 1)
    double dnode[2];
};

// p3 / n3 / dist of AB3, as ab3() computes them
WQ_HD_INLINE void step_ab3(double dt, const StepIn& in, double p3[3], double n3[3], double& dr, double& dth,
                           double& dph, double& a, double& b, double& sn2, double& dist) {
    const double A2 = 23.0 / 12.0, A1 = 16.0 / 12.0, A0 = 5.0 / 12.0;
    dr = dt * (A2 * in.g2[0] - A1 * in.g1[0] + A0 * in.g0[0]);
    dth = dt * (A2 * in.g2[1] - A1 * in.g1[1] + A0 * in.g0[1]);
    dph = dt * (A2 * in.g2[2] - A1 * in.g1[2] + A0 * in.g0[2]);
    sn2 = wqm::sin(in.p2[1]);
    a = in.p2[0] * dth;
    b = in.p2[0] * (sn2 * dph);
    dist = sqrt(dr * dr + a * a + b * b);
    p3[0] = in.p2[0] + dr;
    p3[1] = in.p2[1] + dth;
    p3[2] = in.p2[2] + dph;
    for (int k = 0; k < 3; ++k) n3[k] = in.n2[k] + dt * (A2 * in.h2[k] - A1 * in.h1[k] + A0 * in.h0[k]);
}

WQ_HD_INLINE void step_vjp(const DEnv& E, const FanConst& F, const StepIn& in, const StepBar& bar, StepGrad& out) {
    const double dt = F.dt;
    const double A2 = 23.0 / 12.0, A1 = 16.0 / 12.0, A0 = 5.0 / 12.0;
    double p3[3], n3[3], dr, dth, dph, a, b, sn2, dist;
    step_ab3(dt, in, p3, n3, dr, dth, dph, a, b, sn2, dist);
    // update() at (p3, n3): the grid lookup and the ray equations
    const DGrid& g = E.profile.grid;
    const DSeq& ax = g.ax[0];
    double loc = p3[0];
    bool clamped = false;
    int k;
    if (g.edge[0]) {
        const double lo = ax.v[0];
        const double hi = ax.v[ax.n - 1];
        const double sgn = (ax.inc[0] < 0) ? -1.0 : 1.0;
        const double x = loc * sgn;
        if (x <= lo * sgn) {
            loc = lo;
            k = 0;
            clamped = true;
        } else if (x >= hi * sgn) {
            loc = hi;
            k = ax.n - 2;
            clamped = true;
        } else {
            k = find_index(ax, loc);
        }
    } else {
        k = find_index(ax, loc);
    }
    const double va = g.data[k], vb = g.data[k + 1];
    const double hk = seq_inc(ax, k);
    const double u = (loc - seq_value(ax, k)) / hk;
    double c = va * (1.0 - u) + vb * u;
    double cr = (vb - va) / hk;
    // flat earth: cr' = (cr * rho + c) / R, c' = c * rho / R
    const bool flat = E.profile.flat_earth != 0;
    const double c_grid = c, cr_grid = cr;
    if (flat) {
        cr = (cr_grid * p3[0] + c_grid) / F.R;
        c = c_grid * p3[0] / F.R;
    }
    const double pr = p3[0];
    const double st = wqm::sin(p3[1]);
    const double ct = wqm::cos(p3[1]);
    const double cot = ct / st;
    const double c2 = c * c;
    const double c2r = c2 / pr;
    const double nr = n3[0], nt = n3[1], np = n3[2];
    // adjoints of the new slot's outputs
    double pb[3] = {bar.p[0] + bar.s[0], bar.p[1] + bar.s[1], bar.p[2] + bar.s[2]};
    double nb[3] = {bar.n[0], bar.n[1], bar.n[2]};
    double stb = bar.s[3];
    double cb = bar.c;
    // attenuation: att3 = att2 + distance * alpha * (1 + 5.88264e-6 * alt)
    double distb = 0.0;
    for (int f = 0; f < F.nf; ++f) {
        const double ab = bar.att[f];
        if (ab == 0.0) continue;
        const double freq = seq_value(F.freq, f);
        if (E.profile.atten_kind == 1) {
            distb += ab * E.profile.atten_coeff * freq;
        } else {
            double F2 = freq;
            F2 = 1e-6 * F2 * F2;
            const double alpha = 1e-3 * (3.3e-3 + F2 * (0.11 / (1.0 + F2) + 44.0 / (4100.0 + F2) + 3.0e-4)) /
                                 (1.0 - 5.88264e-6 * 1000.0);
            distb += ab * alpha * (1.0 + 5.88264e-6 * (pr - F.R));
            pb[0] += ab * dist * alpha * 5.88264e-6;
        }
    }
    // gr = c2 nr; gt = c2r nt; gp = c2r / st np;
    // hr = c2r (nt^2 + np^2) - cr / c;
    // ht = -c2r (nr nt - np^2 cot) - 0 / pr; hp = -c2r np (nr + nt cot) - 0 / (pr st)
    const double* gb = bar.g;
    const double* hb = bar.h;
    double c2b = gb[0] * nr;
    double c2rb = gb[1] * nt + gb[2] * np / st + hb[0] * (nt * nt + np * np) - hb[1] * (nr * nt - np * np * cot) -
                  hb[2] * (np * (nr + nt * cot));
    nb[0] += gb[0] * c2 - hb[1] * c2r * nt - hb[2] * c2r * np;
    nb[1] += gb[1] * c2r + hb[0] * c2r * 2.0 * nt - hb[1] * c2r * nr - hb[2] * c2r * np * cot;
    nb[2] += gb[2] * c2r / st + hb[0] * c2r * 2.0 * np + hb[1] * c2r * 2.0 * np * cot - hb[2] * c2r * (nr + nt * cot);
    const double cotb = hb[1] * c2r * np * np - hb[2] * c2r * np * nt;
    stb += -gb[2] * c2r * np / (st * st);
    // cot = ct / st
    const double ctb = cotb / st;
    stb += -cotb * ct / (st * st);
    pb[1] += stb * ct - ctb * st;
    // c2r = c2 / pr
    c2b += c2rb / pr;
    pb[0] += -c2rb * c2 / (pr * pr);
    // hr's - cr / c
    double crb = -hb[0] / c;
    cb += hb[0] * cr / (c * c);
    cb += c2b * 2.0 * c;
    // flat earth and the grid lookup
    double cgb = cb, crgb = crb;
    if (flat) {
        cgb = cb * pr / F.R + crb / F.R;
        crgb = crb * pr / F.R;
        pb[0] += crb * cr_grid / F.R + cb * c_grid / F.R;
    }
    // c = va (1 - u) + vb u, u = (loc - v_k) / h; cr = (vb - va) / h
    if (!clamped) pb[0] += cgb * (vb - va) / hk;
    out.node = k;
    out.dnode[0] = cgb * (1.0 - u) - crgb / hk;
    out.dnode[1] = cgb * u + crgb / hk;
    // AB3: p3 = p2 + d, n3 = n2 + dt (A2 h2 - A1 h1 + A0 h0),
    // dist = |(dr, p2r dth, p2r sin(p2t) dph)|
    double drb = pb[0], dthb = pb[1], dphb = pb[2];
    double p2b[3] = {pb[0], pb[1], pb[2]};
    if (dist > 0.0) {
        drb += distb * dr / dist;
        const double ab_ = distb * a / dist;
        const double bb_ = distb * b / dist;
        dthb += ab_ * in.p2[0];
        dphb += bb_ * in.p2[0] * sn2;
        p2b[0] += ab_ * dth + bb_ * sn2 * dph;
        p2b[1] += bb_ * in.p2[0] * dph * wqm::cos(in.p2[1]);
    }
    const double db[3] = {drb, dthb, dphb};
    for (int q = 0; q < 3; ++q) {
        out.g2[q] = dt * A2 * db[q];
        out.g1[q] = -dt * A1 * db[q];
        out.g0[q] = dt * A0 * db[q];
        out.h2[q] = dt * A2 * nb[q];
        out.h1[q] = -dt * A1 * nb[q];
        out.h0[q] = dt * A0 * nb[q];
        out.p2[q] = p2b[q];
        out.n2[q] = nb[q];
    }
}


// ------------------------------------------------------------ backward-pass bodies
// Device state of a fan's backward pass.  The history keeps, for each step
// of the segment being differentiated, the fan's four slots as they were
// after the step: kPr .. kSs of every slot ([step][kAtt][4][N]), the
// attenuations in the fan's own layout ([step][4][nf][N]), the counters
// ([step][5][4][N]: surface, bottom, caustic, upper, lower), edge flags
// ([step][4][N]) and flag[] ([step][N]).
struct AdjDev {
    double* adj;              // [kAtt + nf][4][N]: adjoint of every slot field
    double* node_grad;        // [nodes]
    const double* hist;
    const double* hatt;
    const int32_t* hint;
    const uint8_t* hedge;
    const uint8_t* hflag;
    int64_t N;
    int32_t nf;
    ad::Node* tape_node;      // tapes of tape_cap operations each
    double* tape_adj;
    int32_t tape_cap;
    unsigned long long* ent_key;  // seed entries: (ray << 24) | (step << 8) | field code
    double* ent_val;
    unsigned long long* ent_count;
    int64_t ent_cap;
    int32_t* err;             // bit 0 tape full, 1 replay differs, 2 entries full,
                              // 3 eigenray not rebuilt, 4 reflection chain too long
};

// the eigenrays the caller differentiates, with their step in the segment
struct AdjSeed {
    int64_t target;
    int64_t ray;
    double travel_time;       // the forward run's value, checked by the replay
    double grad_time;         // dL / d travel_time
    int64_t h;                // step of the segment (1 .. segment length)
};

constexpr int kEntryStepBits = 16;  // segments of at most 65535 steps

WQ_HD_INLINE double* adj_at(const AdjDev& D, int field, int slot, int64_t i) {
    return D.adj + (int64_t)code(field, slot) * D.N + i;
}
WQ_HD_INLINE double hist_at(const AdjDev& D, int64_t h, int field, int slot, int64_t i) {
    return D.hist[((h * kAtt + field) * 4 + slot) * D.N + i];
}

// the fan's arrays as they were after step h of the segment (read only)
WQ_HD_INLINE Arrays hist_view(const AdjDev& D, int64_t h) {
    const int64_t n4 = 4 * D.N;
    double* b = const_cast<double*>(D.hist) + h * kAtt * n4;
    Arrays A{};
    double** f[kAtt] = {&A.pr, &A.pt, &A.pp, &A.nr, &A.nt, &A.np, &A.gr, &A.gt, &A.gp,
                        &A.hr, &A.ht, &A.hp, &A.c, &A.tr, &A.tt, &A.tp, &A.ts};
    for (int k = 0; k < kAtt; ++k) *f[k] = b + k * n4;
    A.atten = const_cast<double*>(D.hatt) + h * n4 * D.nf;
    int32_t* ib = const_cast<int32_t*>(D.hint) + h * 5 * n4;
    A.srf = ib;
    A.btm = ib + n4;
    A.cst = ib + 2 * n4;
    A.upr = ib + 3 * n4;
    A.lwr = ib + 4 * n4;
    A.edge = const_cast<uint8_t*>(D.hedge) + h * n4;
    A.flag = const_cast<uint8_t*>(D.hflag) + h * D.N;
    return A;
}

WQ_HD_INLINE void atom_add_d(double* p, double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicAdd(p, v);
#else
    *p += v;
#endif
}

WQ_HD_INLINE void tape_of(const AdjDev& D, int64_t k, ad::Tape& t) {
    ad::tape_init(t, D.tape_node + k * D.tape_cap, D.tape_adj + k * D.tape_cap, D.tape_cap);
}

// the node-speed adjoints a tape's inputs received
WQ_HD_INLINE void flush_nodes(const AdjDev& D, const ad::Tape& t) {
    for (int32_t k = 0; k < t.n; ++k)
        if (t.node[k].a == ad::kInputNode && t.adj[k] != 0.0) atom_add_d(D.node_grad + t.node[k].b, t.adj[k]);
}

// history capture of ray i after step h of the segment
WQ_HD_INLINE void body_adj_capture(int64_t i, const FanConst& F, const Arrays& A, const AdjDev& D, int64_t h) {
    const Arrays H = hist_view(D, h);
    const int64_t N = F.N;
    const double* src[kAtt] = {A.pr, A.pt, A.pp, A.nr, A.nt, A.np, A.gr, A.gt, A.gp,
                               A.hr, A.ht, A.hp, A.c, A.tr, A.tt, A.tp, A.ts};
    double* dst[kAtt] = {H.pr, H.pt, H.pp, H.nr, H.nt, H.np, H.gr, H.gt, H.gp,
                         H.hr, H.ht, H.hp, H.c, H.tr, H.tt, H.tp, H.ts};
    for (int f = 0; f < kAtt; ++f)
        for (int s = 0; s < 4; ++s) dst[f][s * N + i] = src[f][s * N + i];
    for (int k = 0; k < 4 * F.nf; ++k) H.atten[k * N + i] = A.atten[k * N + i];
    const int32_t* is[5] = {A.srf, A.btm, A.cst, A.upr, A.lwr};
    int32_t* id[5] = {H.srf, H.btm, H.cst, H.upr, H.lwr};
    for (int f = 0; f < 5; ++f)
        for (int s = 0; s < 4; ++s) id[f][s * N + i] = is[f][s * N + i];
    for (int s = 0; s < 4; ++s) H.edge[s * N + i] = A.edge[s * N + i];
    H.flag[i] = A.flag[i];
}

// Seeds of one eigenray: its construction replayed on a tape from the
// history of its step, the caller's gradients pushed back to what it read;
// the adjoints of state fields become entries keyed by (ray, step, field),
// applied at the step's turn in the backward sweep; the node speeds' go
// straight to node_grad.
WQ_HD void body_adj_seed(int64_t e, int64_t tape_k, const DEnv& E, const FanConst& F, const Slots* slots,
                         const double* times, const AdjSeed* seeds, const double* seed_int,
                         const double* seed_grad, const AdjDev& D) {
    ad::Tape tp;
    tape_of(D, tape_k, tp);
    const AdjSeed sd = seeds[e];
    const Arrays A = hist_view(D, sd.h);
    V T;
    V I[kMaxFreq];
    if (!eigenray_v(E, F, A, slots[sd.h], &tp, sd.target, sd.ray, times[sd.h], T, I)) {
        atom_or(D.err, 8);
        return;
    }
    if (tp.overflow) {
        atom_or(D.err, 1);
        return;
    }
    bool same = T.v == sd.travel_time;
    for (int f = 0; f < F.nf; ++f) same = same && I[f].v == seed_int[e * F.nf + f];
    if (!same) atom_or(D.err, 2);
    ad::clear_adjoints(&tp);
    ad::seed(&tp, T, sd.grad_time);
    for (int f = 0; f < F.nf; ++f) ad::seed(&tp, I[f], seed_grad[e * F.nf + f]);
    ad::reverse(&tp);
    unsigned long long n = 0;
    for (int32_t k = 0; k < tp.n; ++k)
        if (tp.node[k].a == ad::kInputState && tp.adj[k] != 0.0) ++n;
    const unsigned long long base = atom_add(D.ent_count, n);
    if ((int64_t)(base + n) > D.ent_cap) {
        atom_or(D.err, 4);
        return;
    }
    unsigned long long w = base;
    for (int32_t k = 0; k < tp.n; ++k) {
        const ad::Node& nd = tp.node[k];
        if (nd.a != ad::kInputState || tp.adj[k] == 0.0) continue;
        D.ent_key[w] = ((unsigned long long)nd.da << (8 + kEntryStepBits)) |
                       ((unsigned long long)sd.h << 8) | (unsigned long long)nd.b;
        D.ent_val[w] = tp.adj[k];
        ++w;
    }
    flush_nodes(D, tp);
}

// The adjoint of one ray's advance in the step (post-rotation slots S,
// window h of the history): the new slot's adjoints go back to the older
// slots and the node speeds, and the new slot's are cleared.  Returns the
// node adjoints through k0 / d0 / d1 for a block-level sum.
WQ_HD_INLINE void body_adj_advance(int64_t i, const DEnv& E, const FanConst& F, const Slots& S, int64_t h,
                                   const AdjDev& D, int& k0, double& d0, double& d1) {
    StepIn in;
    const int sp[3] = {S.past, S.prev, S.curr};
    double* g3[3] = {in.g0, in.g1, in.g2};
    double* h3[3] = {in.h0, in.h1, in.h2};
    for (int s = 0; s < 3; ++s)
        for (int q = 0; q < 3; ++q) {
            g3[s][q] = hist_at(D, h, kGr + q, sp[s], i);
            h3[s][q] = hist_at(D, h, kHr + q, sp[s], i);
        }
    for (int q = 0; q < 3; ++q) {
        in.p2[q] = hist_at(D, h, kPr + q, S.curr, i);
        in.n2[q] = hist_at(D, h, kNr + q, S.curr, i);
    }
    {   // the replayed AB3 must give the forward run's new slot
        double p3[3], n3[3], dr, dth, dph, a, b, sn2, dist;
        step_ab3(F.dt, in, p3, n3, dr, dth, dph, a, b, sn2, dist);
        bool same = true;
        for (int q = 0; q < 3; ++q)
            same = same && p3[q] == hist_at(D, h, kPr + q, S.next, i) && n3[q] == hist_at(D, h, kNr + q, S.next, i);
        if (!same) atom_or(D.err, 2);
    }
    StepBar bar;
    const int sn = S.next;
    for (int q = 0; q < 3; ++q) {
        bar.p[q] = *adj_at(D, kPr + q, sn, i);
        bar.n[q] = *adj_at(D, kNr + q, sn, i);
        bar.g[q] = *adj_at(D, kGr + q, sn, i);
        bar.h[q] = *adj_at(D, kHr + q, sn, i);
    }
    bar.c = *adj_at(D, kC, sn, i);
    for (int q = 0; q < 4; ++q) bar.s[q] = *adj_at(D, kSr + q, sn, i);
    for (int f = 0; f < F.nf; ++f) bar.att[f] = *adj_at(D, kAtt + f, sn, i);
    StepGrad out;
    step_vjp(E, F, in, bar, out);
    const double* og[3] = {out.g0, out.g1, out.g2};
    const double* oh[3] = {out.h0, out.h1, out.h2};
    for (int s = 0; s < 3; ++s)
        for (int q = 0; q < 3; ++q) {
            *adj_at(D, kGr + q, sp[s], i) += og[s][q];
            *adj_at(D, kHr + q, sp[s], i) += oh[s][q];
        }
    for (int q = 0; q < 3; ++q) {
        *adj_at(D, kPr + q, S.curr, i) += out.p2[q];
        *adj_at(D, kNr + q, S.curr, i) += out.n2[q];
    }
    // attenuation carries: att(next) = att(curr) + the step's absorption
    for (int f = 0; f < F.nf; ++f) *adj_at(D, kAtt + f, S.curr, i) += bar.att[f];
    for (int fld = 0; fld < kAtt + F.nf; ++fld) *adj_at(D, fld, sn, i) = 0.0;
    k0 = out.node;
    d0 = out.dnode[0];
    d1 = out.dnode[1];
}

// The adjoint of one ray's reflections in the step: the surface/bottom
// chain replayed on a tape from the pre-reflection window (history h_pre,
// slots Sp before the rotation), checked against the post-step window
// (history h_post, slots Sq), then the adjoints of what it wrote pushed
// back to what it read.
WQ_HD void body_adj_reflect(int64_t i, int64_t tape_k, const DEnv& E, const FanConst& F, const Slots& Sp,
                            const Slots& Sq, int64_t h_pre, int64_t h_post, const AdjDev& D) {
    ad::Tape tp;
    tape_of(D, tape_k, tp);
    const NodesV nodes{E.profile.grid.data, &tp};
    const int sp[4] = {Sp.past, Sp.prev, Sp.curr, Sp.next};
    WinT<V> w;
    for (int q = 0; q < 4; ++q) {
        w.s[q] = fresh_t<V>();
        if (q == 0) continue;  // the reflection reads prev, curr and next
        RayT<V>& r = w.s[q];
        V* p[3] = {&r.pr, &r.pt, &r.pp};
        V* n[3] = {&r.nr, &r.nt, &r.np};
        for (int k = 0; k < 3; ++k) {
            *p[k] = ad::input_state(&tp, hist_at(D, h_pre, kPr + k, sp[q], i), code(kPr + k, sp[q]), i);
            *n[k] = ad::input_state(&tp, hist_at(D, h_pre, kNr + k, sp[q], i), code(kNr + k, sp[q]), i);
        }
        r.c = ad::input_state(&tp, hist_at(D, h_pre, kC, sp[q], i), code(kC, sp[q]), i);
    }
    w.amp = V(0.0);
    if (!bounce_chain_t(E, F, nodes, w)) {
        atom_or(D.err, 16);
        return;
    }
    if (tp.overflow) {
        atom_or(D.err, 1);
        return;
    }
    // after the rotation, the reflected prev / curr / next are past / prev / curr
    {
        const int sq[3] = {Sq.past, Sq.prev, Sq.curr};
        bool same = true;
        for (int q = 1; q < 4; ++q) {
            const RayT<V>& r = w.s[q];
            const V* f13[kStateFields] = {&r.pr, &r.pt, &r.pp, &r.nr, &r.nt, &r.np, &r.gr,
                                         &r.gt, &r.gp, &r.hr, &r.ht, &r.hp, &r.c};
            for (int fld = 0; fld < kStateFields; ++fld)
                same = same && f13[fld]->v == hist_at(D, h_post, fld, sq[q - 1], i);
        }
        if (!same) atom_or(D.err, 2);
    }
    ad::clear_adjoints(&tp);
    for (int q = 0; q < 4; ++q) {
        const RayT<V>& r = w.s[q];
        const V* f13[kStateFields] = {&r.pr, &r.pt, &r.pp, &r.nr, &r.nt, &r.np, &r.gr,
                                     &r.gt, &r.gp, &r.hr, &r.ht, &r.hp, &r.c};
        for (int fld = 0; fld < kStateFields; ++fld) {
            double* a = adj_at(D, fld, sp[q], i);
            ad::seed(&tp, *f13[fld], *a);
            *a = 0.0;
        }
    }
    {
        double ab = 0.0;
        for (int f = 0; f < F.nf; ++f) ab += *adj_at(D, kAtt + f, Sp.next, i);
        ad::seed(&tp, w.amp, ab);
    }
    ad::reverse(&tp);
    for (int32_t k = 0; k < tp.n; ++k) {
        const ad::Node& nd = tp.node[k];
        if (nd.a == ad::kInputState && tp.adj[k] != 0.0) D.adj[(int64_t)nd.b * D.N + (int64_t)nd.da] += tp.adj[k];
    }
    flush_nodes(D, tp);
}

// The backward sweep of ray i over the segment's steps K .. 1: the step's
// seed entries for the ray (sorted by ray, then step: the ray's range
// [beg, end) is read from its end), the adjoint of its advance, the
// adjoint of its reflections.  Rays are independent here - the seeds carry
// every dependence between rays - so one thread takes a ray through the
// whole segment.  The advance's node adjoints go to add_node(k, d0, d1).
template <class AddNode>
WQ_HD void body_adj_sweep(int64_t i, int64_t tape_k, const DEnv& E, const FanConst& F, const Slots* slots,
                          int64_t K, const unsigned long long* key, const double* val, int64_t beg, int64_t end,
                          const AdjDev& D, AddNode add_node) {
    int64_t p = end;
    for (int64_t h = K; h >= 1; --h) {
        while (p > beg) {
            const unsigned long long k = key[p - 1];
            if ((int64_t)((k >> 8) & ((1ull << kEntryStepBits) - 1)) != h) break;
            D.adj[(int64_t)(k & 0xffull) * D.N + i] += val[p - 1];
            --p;
        }
        int k0;
        double d0, d1;
        body_adj_advance(i, E, F, slots[h], h, D, k0, d0, d1);
        add_node(k0, d0, d1);
        if (D.hflag[(h - 1) * D.N + i] != 0) body_adj_reflect(i, tape_k, E, F, slots[h - 1], slots[h], h - 1, h, D);
    }
}

// The adjoint of the fan's start (body_init): the node speeds' share of
// the adjoints of the four initial slots, through the launch direction's
// 1 / c at the source and the start-up's update() calls.
WQ_HD void body_adj_init(int64_t i, int64_t tape_k, const DEnv& E, const FanConst& F, const Slots& S,
                         const AdjDev& D) {
    ad::Tape tp;
    tape_of(D, tape_k, tp);
    const NodesV nodes{E.profile.grid.data, &tp};
    const int de = (int)(i / F.naz), az = (int)(i % F.naz);
    const double dt = F.dt;
    V c0, g0;
    sound_speed_t(E.profile, nodes, V(F.src_r), F.R, c0, g0);
    if (c0.v != F.c0) atom_or(D.err, 2);
    RayT<V> cur = fresh_t<V>();
    cur.nr = V(F.sin_de[de]);
    cur.nt = V(-F.cos_de[de] * F.cos_az[az]);
    cur.np = V(F.cos_de[de] * F.sin_az[az]);
    cur.pr = V(F.src_r);
    cur.pt = V(F.src_t);
    cur.pp = V(F.src_p);
    cur.nr = cur.nr / c0;
    cur.nt = cur.nt / c0;
    cur.np = cur.np / c0;
    update_t(E, nodes, F.R, cur);
    RayT<V> nxt = fresh_t<V>(), pst = fresh_t<V>(), prv = fresh_t<V>();
    rk1_t(-dt, cur, nxt);
    update_t(E, nodes, F.R, nxt);
    rk2_t(-dt, cur, nxt, pst);
    update_t(E, nodes, F.R, pst);
    rk3_t(-dt, cur, nxt, pst, prv);
    update_t(E, nodes, F.R, prv);
    rk1_t(-dt, prv, nxt);
    update_t(E, nodes, F.R, nxt);
    rk2_t(-dt, prv, nxt, pst);
    update_t(E, nodes, F.R, pst);
    rk3_t(-dt, prv, nxt, pst, pst);
    update_t(E, nodes, F.R, pst);
    ab3_t(dt, pst, prv, cur, nxt);
    update_t(E, nodes, F.R, nxt);
    const int sl[4] = {S.past, S.prev, S.curr, S.next};
    const RayT<V>* st[4] = {&pst, &prv, &cur, &nxt};
    V att[4][kMaxFreq];
    for (int s = 0; s < 4; ++s)
        for (int f = 0; f < F.nf; ++f)
            att[s][f] = attenuation_t(E.profile, seq_value(F.freq, f), st[s]->pr - F.R, st[s]->dist);
    if (tp.overflow) {
        atom_or(D.err, 1);
        return;
    }
    ad::clear_adjoints(&tp);
    for (int s = 0; s < 4; ++s) {
        const RayT<V>& r = *st[s];
        const V* f13[kStateFields] = {&r.pr, &r.pt, &r.pp, &r.nr, &r.nt, &r.np, &r.gr,
                                     &r.gt, &r.gp, &r.hr, &r.ht, &r.hp, &r.c};
        for (int fld = 0; fld < kStateFields; ++fld) ad::seed(&tp, *f13[fld], *adj_at(D, fld, sl[s], i));
        // the snapshot of update(): position and sin(theta)
        ad::seed(&tp, r.pr, *adj_at(D, kSr, sl[s], i));
        ad::seed(&tp, r.pt, *adj_at(D, kSt, sl[s], i));
        ad::seed(&tp, r.pp, *adj_at(D, kSp, sl[s], i));
        ad::seed(&tp, r.st, *adj_at(D, kSs, sl[s], i));
        for (int f = 0; f < F.nf; ++f) ad::seed(&tp, att[s][f], *adj_at(D, kAtt + f, sl[s], i));
    }
    ad::reverse(&tp);
    flush_nodes(D, tp);
}

}  // namespace adj
}  // namespace wq3d
