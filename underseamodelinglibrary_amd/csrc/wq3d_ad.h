// wq3d_ad.h - reverse-mode differentiation on a tape (adjoint of the
// eigenray travel time and transmission loss, wq3d_adjoint_*).
//
// A value V carries its double value and, when it depends on a registered
// input, the index of the operation that made it on its tape.  Every
// operation on V values computes the value exactly as the same expression
// on doubles does (same operations, same order, the shared elementary
// functions of wq3d_math.h), so code written once as a template over the
// number type takes the forward path's branches bit for bit, and records
// the partial derivatives of each operation at that point.  reverse() then
// accumulates the adjoints from the outputs back to the inputs.
//
// Tapes are used where the work per ray is rare or irregular: the
// reflections of one step (about ten update() calls per bounce), the start
// of the fan, and the construction of each eigenray (the 27-point
// neighbourhood, compute_offsets and the Gaussian spreading walks).  The
// propagation step every ray takes has a hand-written adjoint
// (wq3d_adjoint.h), checked against the tape.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wq3d_math.h"

namespace wq3d {
namespace ad {

// value = f(a, b) with partials da = d value / d a, db = d value / d b at
// the recorded point.  Inputs have a < 0: kInputState for a state field of
// one ray (b = field code, da = ray index), kInputNode for a node of the
// sound speed grid (b = node index).
struct Node {
    int32_t a, b;
    double da, db;
};

constexpr int32_t kNone = -1;
constexpr int32_t kInputState = -2;
constexpr int32_t kInputNode = -3;
constexpr int32_t kInputPlain = -4;

struct Tape {
    Node* node;
    double* adj;
    int32_t n, cap;
    int32_t overflow;
};

struct V {
    double v;
    int32_t i;  // operation index on t, or < 0 for a constant
    Tape* t;
    __host__ __device__ V() : v(0.0), i(-1), t(nullptr) {}
    __host__ __device__ V(double x) : v(x), i(-1), t(nullptr) {}
    __host__ __device__ V(double x, int32_t k, Tape* tp) : v(x), i(k), t(tp) {}
};

__host__ __device__ __forceinline__ void tape_init(Tape& t, Node* node, double* adj, int32_t cap) {
    t.node = node;
    t.adj = adj;
    t.n = 0;
    t.cap = cap;
    t.overflow = 0;
}

__host__ __device__ __forceinline__ int32_t push(Tape* t, int32_t a, double da, int32_t b, double db) {
    if (t->n >= t->cap) {
        t->overflow = 1;
        return -1;
    }
    const int32_t k = t->n++;
    t->node[k] = Node{a, b, da, db};
    return k;
}

__host__ __device__ __forceinline__ V rec1(const V& x, double v, double dx) {
    if (x.i < 0) return V(v);
    return V(v, push(x.t, x.i, dx, kNone, 0.0), x.t);
}

__host__ __device__ __forceinline__ V rec2(const V& x, const V& y, double v, double dx, double dy) {
    if (x.i < 0 && y.i < 0) return V(v);
    Tape* t = x.i >= 0 ? x.t : y.t;
    return V(v, push(t, x.i, dx, y.i, dy), t);
}

// inputs: a state field of a ray, a grid node, or a plain variable
__host__ __device__ __forceinline__ V input_state(Tape* t, double v, int32_t code, int64_t ray) {
    return V(v, push(t, kInputState, (double)ray, code, 0.0), t);
}
__host__ __device__ __forceinline__ V input_node(Tape* t, double v, int32_t node) {
    return V(v, push(t, kInputNode, 0.0, node, 0.0), t);
}
__host__ __device__ __forceinline__ V input_plain(Tape* t, double v, int32_t tag) {
    return V(v, push(t, kInputPlain, 0.0, tag, 0.0), t);
}

__host__ __device__ __forceinline__ void clear_adjoints(Tape* t) {
    for (int32_t k = 0; k < t->n; ++k) t->adj[k] = 0.0;
}

__host__ __device__ __forceinline__ void seed(Tape* t, const V& y, double g) {
    if (y.i >= 0 && y.t == t) t->adj[y.i] += g;
}

__host__ __device__ inline void reverse(Tape* t) {
    for (int32_t k = t->n - 1; k >= 0; --k) {
        const double g = t->adj[k];
        if (g == 0.0) continue;
        const Node nd = t->node[k];
        if (nd.a < kNone) continue;  // an input
        if (nd.a >= 0) t->adj[nd.a] += g * nd.da;
        if (nd.b >= 0) t->adj[nd.b] += g * nd.db;
    }
}

// A value computed outside the tape from n inputs, with its partial
// derivatives d value / d x[k] given: recorded as a chain of n - 1 two-input
// operations instead of the operations of its formula.
template <int NI>
__host__ __device__ __forceinline__ V combine(double v, const V (&x)[NI], const double (&d)[NI]) {
    Tape* t = nullptr;
    for (int k = 0; k < NI; ++k)
        if (x[k].i >= 0) t = x[k].t;
    if (!t) return V(v);
    int32_t acc = -1;
    double dacc = 0.0;
    for (int k = 0; k < NI; ++k) {
        if (x[k].i < 0 || d[k] == 0.0) continue;
        if (acc < 0) {
            acc = x[k].i;
            dacc = d[k];
            continue;
        }
        acc = push(t, acc, dacc, x[k].i, d[k]);
        dacc = 1.0;
        if (acc < 0) return V(v);
    }
    if (acc < 0) return V(v);
    return V(v, push(t, acc, dacc, kNone, 0.0), t);
}

// ------------------------------------------------------------ arithmetic
__host__ __device__ __forceinline__ V operator+(const V& x, const V& y) { return rec2(x, y, x.v + y.v, 1.0, 1.0); }
__host__ __device__ __forceinline__ V operator-(const V& x, const V& y) { return rec2(x, y, x.v - y.v, 1.0, -1.0); }
__host__ __device__ __forceinline__ V operator*(const V& x, const V& y) { return rec2(x, y, x.v * y.v, y.v, x.v); }
__host__ __device__ __forceinline__ V operator/(const V& x, const V& y) {
    const double q = x.v / y.v;
    return rec2(x, y, q, 1.0 / y.v, -q / y.v);
}
__host__ __device__ __forceinline__ V operator-(const V& x) { return rec1(x, -x.v, -1.0); }
__host__ __device__ __forceinline__ V& operator+=(V& x, const V& y) { return x = x + y; }
__host__ __device__ __forceinline__ V& operator-=(V& x, const V& y) { return x = x - y; }
__host__ __device__ __forceinline__ V& operator*=(V& x, const V& y) { return x = x * y; }
__host__ __device__ __forceinline__ V& operator/=(V& x, const V& y) { return x = x / y; }
__host__ __device__ __forceinline__ bool operator<(const V& x, const V& y) { return x.v < y.v; }
__host__ __device__ __forceinline__ bool operator<=(const V& x, const V& y) { return x.v <= y.v; }
__host__ __device__ __forceinline__ bool operator>(const V& x, const V& y) { return x.v > y.v; }
__host__ __device__ __forceinline__ bool operator>=(const V& x, const V& y) { return x.v >= y.v; }
__host__ __device__ __forceinline__ bool operator==(const V& x, const V& y) { return x.v == y.v; }
__host__ __device__ __forceinline__ bool operator!=(const V& x, const V& y) { return x.v != y.v; }

__host__ __device__ __forceinline__ V sqrt(const V& x) {
    const double s = ::sqrt(x.v);
    return rec1(x, s, s > 0.0 ? 0.5 / s : 0.0);
}
__host__ __device__ __forceinline__ V fabs(const V& x) { return rec1(x, ::fabs(x.v), x.v < 0.0 ? -1.0 : 1.0); }
__host__ __device__ __forceinline__ V copysign(const V& x, const V& y) {
    const double r = ::copysign(x.v, y.v);
    return rec1(x, r, (r == x.v) ? 1.0 : -1.0);
}

// plain values of either number type
__host__ __device__ __forceinline__ double val(double x) { return x; }
__host__ __device__ __forceinline__ double val(const V& x) { return x.v; }

}  // namespace ad
}  // namespace wq3d

// the shared elementary functions on tape values
namespace wqm {
using wq3d::ad::rec1;
using wq3d::ad::rec2;
using wq3d::ad::V;
constexpr double kLn10 = 2.302585092994045684;
__host__ __device__ __forceinline__ V sin(const V& x) { return rec1(x, wqm::sin(x.v), wqm::cos(x.v)); }
__host__ __device__ __forceinline__ V cos(const V& x) { return rec1(x, wqm::cos(x.v), -wqm::sin(x.v)); }
__host__ __device__ __forceinline__ V exp(const V& x) {
    const double e = wqm::exp(x.v);
    return rec1(x, e, e);
}
__host__ __device__ __forceinline__ V log(const V& x) { return rec1(x, wqm::log(x.v), 1.0 / x.v); }
__host__ __device__ __forceinline__ V log10(const V& x) { return rec1(x, wqm::log10(x.v), 1.0 / (x.v * kLn10)); }
__host__ __device__ __forceinline__ V pow10(const V& x) {
    const double p = wqm::pow10(x.v);
    return rec1(x, p, p * kLn10);
}
__host__ __device__ __forceinline__ V atan2(const V& y, const V& x) {
    const double r2 = x.v * x.v + y.v * y.v;
    return rec2(y, x, wqm::atan2(y.v, x.v), r2 > 0.0 ? x.v / r2 : 0.0, r2 > 0.0 ? -y.v / r2 : 0.0);
}
__host__ __device__ __forceinline__ V asin(const V& x) {
    const double q = ::sqrt((1.0 - x.v) * (1.0 + x.v));
    return rec1(x, wqm::asin(x.v), q > 0.0 ? 1.0 / q : 0.0);
}
}  // namespace wqm

namespace wq3d {
namespace ad {

// complex arithmetic over either number type, operation for operation as
// wqm::cplx (the Rayleigh reflection coefficient)
template <class T>
struct cx {
    T re, im;
};
template <class T>
__host__ __device__ __forceinline__ cx<T> cadd(const cx<T>& a, const cx<T>& b) { return {a.re + b.re, a.im + b.im}; }
template <class T>
__host__ __device__ __forceinline__ cx<T> csub(const cx<T>& a, const cx<T>& b) { return {a.re - b.re, a.im - b.im}; }
template <class T>
__host__ __device__ __forceinline__ cx<T> cmul(const cx<T>& a, const cx<T>& b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T>
__host__ __device__ __forceinline__ cx<T> cscale(const cx<T>& a, const T& s) { return {a.re * s, a.im * s}; }
template <class T>
__host__ __device__ __forceinline__ cx<T> cdiv(const cx<T>& a, const cx<T>& b) {
    if (::fabs(val(b.re)) >= ::fabs(val(b.im))) {
        const T r = b.im / b.re;
        const T d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const T r = b.re / b.im;
    const T d = b.im + b.re * r;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
template <class T>
__host__ __device__ __forceinline__ T chypot(const T& x, const T& y) {
    using ::sqrt;
    return sqrt(x * x + y * y);
}
template <class T>
__host__ __device__ __forceinline__ cx<T> csqrt(const cx<T>& z) {
    using ::copysign;
    using ::fabs;
    using ::sqrt;
    if (val(z.re) == 0.0 && val(z.im) == 0.0) return {T(0.0), z.im};
    const T t = sqrt((fabs(z.re) + chypot(z.re, z.im)) * 0.5);
    if (val(z.re) >= 0.0) return {t, z.im / (2.0 * t)};
    return {fabs(z.im) / (2.0 * t), copysign(t, z.im)};
}
template <class T>
__host__ __device__ __forceinline__ T cabs(const cx<T>& z) { return chypot(z.re, z.im); }
template <class T>
__host__ __device__ __forceinline__ T carg(const cx<T>& z) { return wqm::atan2(z.im, z.re); }

}  // namespace ad
}  // namespace wq3d
