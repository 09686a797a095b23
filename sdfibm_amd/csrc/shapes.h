// shapes.h -- device shape predicates and signed distances.
//
// Two predicates per shape, as in the reference (src/libshape/ishape.h:51-52):
//   inside(p)  strict comparisons without square roots (isInside), FP64 always,
//              so candidate-cell classification is bit-identical to an FP64 CPU
//              evaluation in the same operation order;
//   sdf<R>(p)  signed distance through sdf::filter (signedDistance), in R =
//              float or double, used only for volume fractions.
// Both act on local coordinates conj(q) (p - t) q (ishape.h:43-46).
//
// This file is compiled with FMA contraction off (see the build line), so
// a*b + c stays two rounded operations, as on the CPU.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/sdfibm_hip.h"

namespace sdfibm_dev {

constexpr double kTOL = 1e-8;    // sdf.h:9
constexpr double kSMALL = 1e-6;  // types.h:19

struct d3 {
    double x, y, z;
};
__host__ __device__ inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline d3 cross(d3 a, d3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Derived constants of a CSG shape (sdfibm_csg_shape), computed once per
// set_csg_shapes: per primitive the kind, the operation combining it with the
// primitives before it, the cylinder axis, the centre in the solid's frame, the
// half extents and the squares the strict predicates compare against.
struct CsgPrimK {
    int kind, op, axis, pad;
    double c[3];
    double h[3];  // half extents: box half sides, ellipsoid semi-axes, sphere (r, r, r), cylinder (r, half length, r)
    double k[3];  // ellipsoid 1/(a*a), 1/(b*b), 1/(c*c); sphere and cylinder r*r in k[0]
};
struct CsgK {
    int n, pad;
    CsgPrimK p[SDFIBM_CSG_MAX_PRIMS];
};

__host__ inline CsgK make_csg_k(const sdfibm_csg_shape& g) {
    CsgK K{};
    K.n = g.n_prims;
    for (int t = 0; t < g.n_prims && t < SDFIBM_CSG_MAX_PRIMS; ++t) {
        const sdfibm_csg_prim& a = g.prim[t];
        CsgPrimK& b = K.p[t];
        b.kind = a.kind;
        b.op = a.op;
        b.axis = a.axis;
        for (int d = 0; d < 3; ++d) b.c[d] = a.center[d];
        switch (a.kind) {
        case SDFIBM_CSG_BOX:
            for (int d = 0; d < 3; ++d) b.h[d] = a.size[d];
            break;
        case SDFIBM_CSG_ELLIPSOID:  // the ellipsoid constructor's inverse squares (ellipsoid.h:18-20)
            for (int d = 0; d < 3; ++d) {
                b.h[d] = a.size[d];
                b.k[d] = 1.0 / (a.size[d] * a.size[d]);
            }
            break;
        case SDFIBM_CSG_SPHERE:
            b.h[0] = b.h[1] = b.h[2] = a.size[0];
            b.k[0] = a.size[0] * a.size[0];
            break;
        case SDFIBM_CSG_CYLINDER:
            for (int d = 0; d < 3; ++d) b.h[d] = d == a.axis ? a.size[1] : a.size[0];
            b.k[0] = a.size[0] * a.size[0];
            break;
        default: break;
        }
    }
    return K;
}

// Per-solid constants derived once per step from the solid record, with the
// same expressions the reference's shape constructors use (e.g. 1.0/(a*a)).
struct SolidK {
    int shape;
    int pad;
    const CsgK* csg;  // CSG solids: their shape's constants (device or host table, as make_solid_k was given)
    double c[3];    // centre
    double q[4];    // (w, x, y, z)
    double com[3];  // shape com offset (zero where the shape has none)
    double k[6];    // shape constants, see make_solid_k
    double v[3];    // velocity
    double w[3];    // angular velocity
    double rt[9];   // R(q)^T row-major: local = rt (p - c), for the signed distance
    double qc[9];   // columns of rotate_to_local's linear map (the inside test's rotation), column-major
};

__host__ __device__ inline d3 rotate_to_local(const SolidK& K, d3 u);

// csg_table: the CSG shape constants (device or host copy, whichever side
// evaluates K); a CSG solid's prm[0] is its index there.
__host__ __device__ inline SolidK make_solid_k(const sdfibm_solid& s, const CsgK* csg_table = nullptr) {
    SolidK K{};
    K.shape = s.shape;
    K.csg = (s.shape == SDFIBM_SHAPE_CSG && csg_table) ? csg_table + (long long)s.prm[0] : nullptr;
    for (int d = 0; d < 3; ++d) {
        K.c[d] = s.center[d];
        K.v[d] = s.vel[d];
        K.w[d] = s.omega[d];
    }
    for (int d = 0; d < 4; ++d) K.q[d] = s.q[d];
    const double* p = s.prm;
    switch (s.shape) {
    case SDFIBM_SHAPE_SPHERE:
    case SDFIBM_SHAPE_CIRCLE:
        K.k[0] = p[0];
        K.k[1] = p[0] * p[0];  // spheres: k[2..4] = the world centre (set below, after rt)
        K.com[0] = p[1];
        K.com[1] = p[2];
        K.com[2] = p[3];
        break;
    case SDFIBM_SHAPE_ELLIPSOID:
        K.k[0] = 1.0 / (p[0] * p[0]);
        K.k[1] = 1.0 / (p[1] * p[1]);
        K.k[2] = 1.0 / (p[2] * p[2]);
        break;
    case SDFIBM_SHAPE_ELLIPSE:
        K.k[0] = 1.0 / (p[0] * p[0]);
        K.k[1] = 1.0 / (p[1] * p[1]);
        K.com[0] = p[3];
        K.com[1] = p[4];
        K.com[2] = p[5];
        break;
    case SDFIBM_SHAPE_BOX:
        K.k[0] = p[0];
        K.k[1] = p[1];
        K.k[2] = p[2];
        K.com[0] = p[3];
        K.com[1] = p[4];
        K.com[2] = p[5];
        break;
    case SDFIBM_SHAPE_RECTANGLE:
        K.k[0] = p[0];
        K.k[1] = p[1];
        K.com[0] = p[3];
        K.com[1] = p[4];
        K.com[2] = p[5];
        break;
    case SDFIBM_SHAPE_CIRCLE_TAIL:
    case SDFIBM_SHAPE_CIRCLE_TWOTAIL:
        K.k[0] = p[0];
        K.k[1] = p[0] * p[0];
        K.k[2] = (p[1] + 1) * 0.5 * p[0];  // tail half-length radiusa
        K.k[3] = (s.shape == SDFIBM_SHAPE_CIRCLE_TAIL) ? p[2] : p[2] * 0.5;  // half-thickness
        K.com[0] = p[3];
        K.com[1] = p[4];
        K.com[2] = p[5];
        break;
    default: break;
    }
    {
        const double w = s.q[0], x = s.q[1], y = s.q[2], z = s.q[3];
        // rows of R^T = columns of R(q)
        K.rt[0] = 1 - 2 * (y * y + z * z);
        K.rt[1] = 2 * (x * y + w * z);
        K.rt[2] = 2 * (x * z - w * y);
        K.rt[3] = 2 * (x * y - w * z);
        K.rt[4] = 1 - 2 * (x * x + z * z);
        K.rt[5] = 2 * (y * z + w * x);
        K.rt[6] = 2 * (x * z + w * y);
        K.rt[7] = 2 * (y * z - w * x);
        K.rt[8] = 1 - 2 * (x * x + y * y);
    }
    if (s.shape == SDFIBM_SHAPE_SPHERE)  // rt (p - c) + com = 0  <=>  p = c - R com
        for (int d = 0; d < 3; ++d)
            K.k[2 + d] = K.c[d] - (K.rt[d] * K.com[0] + K.rt[3 + d] * K.com[1] + K.rt[6 + d] * K.com[2]);
    for (int a = 0; a < 3; ++a) {
        const d3 col = rotate_to_local(K, d3{a == 0 ? 1.0 : 0.0, a == 1 ? 1.0 : 0.0, a == 2 ? 1.0 : 0.0});
        K.qc[3 * a + 0] = col.x;
        K.qc[3 * a + 1] = col.y;
        K.qc[3 * a + 2] = col.z;
    }
    return K;
}

// The rotation part of world_to_local: conj(q) u q in the same operation order.
__host__ __device__ inline d3 rotate_to_local(const SolidK& K, d3 u) {
    const double aw = K.q[0];
    const d3 av{-K.q[1], -K.q[2], -K.q[3]};
    const double mw = -dot(av, u);
    const d3 mv = (aw * u) + cross(av, u);
    const d3 cav{K.q[1], K.q[2], K.q[3]};
    return (mw * cav + aw * mv) + cross(mv, cav);
}

// world2local = conjugate(q).transform(p - t), with OpenFOAM's quaternion
// product order: m = a.mulq0v(u), result = (m * conjugate(a)).v, a = conj(q).
__host__ __device__ inline d3 world_to_local(const SolidK& K, d3 p) {
    return rotate_to_local(K, p - d3{K.c[0], K.c[1], K.c[2]});
}

__host__ __device__ inline bool rect_bool(double x, double y, double ra, double rb) {
    return fabs(x) < ra && fabs(y) < rb;
}

// CSG: the strict predicates of the primitives combined with the reference's
// boolean operators (sdf.h:133-142: U = or, D(a, b) = a and not b, I = and),
// left to right.
__host__ __device__ inline bool csg_prim_inside(const CsgPrimK& g, d3 p) {
    const double x = p.x - g.c[0], y = p.y - g.c[1], z = p.z - g.c[2];
    switch (g.kind) {
    case SDFIBM_CSG_BOX: return fabs(x) < g.h[0] && fabs(y) < g.h[1] && fabs(z) < g.h[2];
    case SDFIBM_CSG_ELLIPSOID: return x * x * g.k[0] + y * y * g.k[1] + z * z * g.k[2] < 1.0;
    case SDFIBM_CSG_SPHERE: return x * x + y * y + z * z < g.k[0];
    case SDFIBM_CSG_CYLINDER: {
        const double u = g.axis == 0 ? y : x, v = g.axis == 2 ? y : z, w = g.axis == 0 ? x : (g.axis == 1 ? y : z);
        return u * u + v * v < g.k[0] && fabs(w) < g.h[g.axis];
    }
    default: return false;
    }
}

__host__ __device__ inline bool csg_inside(const CsgK& G, d3 p) {
    bool in = csg_prim_inside(G.p[0], p);
    for (int t = 1; t < G.n; ++t) {
        const bool b = csg_prim_inside(G.p[t], p);
        in = G.p[t].op == SDFIBM_CSG_UNION ? (in || b) : (G.p[t].op == SDFIBM_CSG_DIFFERENCE ? (in && !b) : (in && b));
    }
    return in;
}

// Strict inside test in local coordinates (each shape's isInside).
__host__ __device__ inline bool inside_local(const SolidK& K, d3 p) {
    switch (K.shape) {
    case SDFIBM_SHAPE_SPHERE: {
        const d3 q = d3{K.com[0], K.com[1], K.com[2]} + p;
        return q.x * q.x + q.y * q.y + q.z * q.z < K.k[1];
    }
    case SDFIBM_SHAPE_CIRCLE: {
        const d3 q = d3{K.com[0], K.com[1], K.com[2]} + d3{p.x, p.y, 0.0};
        return q.x * q.x + q.y * q.y + q.z * q.z < K.k[1];
    }
    case SDFIBM_SHAPE_ELLIPSOID:
        return p.x * p.x * K.k[0] + p.y * p.y * K.k[1] + p.z * p.z * K.k[2] < 1.0;
    case SDFIBM_SHAPE_ELLIPSE: {
        const d3 q = d3{K.com[0], K.com[1], K.com[2]} + p;
        return q.x * q.x * K.k[0] + q.y * q.y * K.k[1] < 1.0;
    }
    case SDFIBM_SHAPE_BOX: {
        const d3 q = d3{K.com[0], K.com[1], K.com[2]} + p;
        return fabs(q.x) < K.k[0] && fabs(q.y) < K.k[1] && fabs(q.z) < K.k[2];
    }
    case SDFIBM_SHAPE_RECTANGLE: {
        const d3 q = d3{K.com[0], K.com[1], K.com[2]} + p;
        return rect_bool(q.x, q.y, K.k[0], K.k[1]);
    }
    case SDFIBM_SHAPE_PLANE: return p.y < 0;
    case SDFIBM_SHAPE_CIRCLE_TAIL: {
        const d3 q = d3{K.com[0], K.com[1], K.com[2]} + p;
        const bool b1 = rect_bool(q.x - K.k[2], q.y - 0.0, K.k[2], K.k[3]);
        const bool b2 = q.x * q.x + q.y * q.y + 0.0 * 0.0 < K.k[1];
        return b1 || b2;
    }
    case SDFIBM_SHAPE_CIRCLE_TWOTAIL: {
        const d3 q = d3{K.com[0], K.com[1], K.com[2]} + p;
        const bool dc = q.x * q.x + q.y * q.y + 0.0 * 0.0 < K.k[1];
        // rot30 (sdf.h:155-158), then offset by (radiusa, 0, 0)
        const double r1x = 0.866025404 * q.x + 0.5 * q.y, r1y = 0.866025404 * q.y - 0.5 * q.x;
        const double r2x = 0.866025404 * q.x + 0.5 * (-q.y), r2y = 0.866025404 * (-q.y) - 0.5 * q.x;
        const bool d1 = rect_bool(r1x - K.k[2], r1y - 0.0, K.k[2], K.k[3]);
        const bool d2 = rect_bool(r2x - K.k[2], r2y - 0.0, K.k[2], K.k[3]);
        return dc || d1 || d2;
    }
    case SDFIBM_SHAPE_CSG: return csg_inside(*K.csg, p);
    default: return false;
    }
}

__host__ __device__ inline bool inside_world(const SolidK& K, d3 p) { return inside_local(K, world_to_local(K, p)); }

// ---- inside bits of a row of vertices (quadric solids) ----------------------
// Along a row of vertices x_0 < x_1 < ... at fixed (y, z), the local point is
// affine in t = x - c.x (rotate_to_local is linear), so the strict predicate of
// a sphere, circle, ellipsoid or ellipse reads f(t) = A t^2 + 2 B t + C < 0,
// with A > 0: its inside set is the interval between the roots t1 < t2.  A
// vertex farther than a margin from both roots takes the interval's answer;
// one within it, and every vertex of a row near tangency (|B^2 - AC| below
// 1e-6 of B^2 + |AC|), takes the exact test inside_world.
//
// The margin (1e-6 of the row's largest |t1|, |t2|, |y - c.y|, |z - c.z|)
// keeps |f| at a decided vertex above 1e-9 A T^2, where the rounding of the
// exact test's operation order moves f by about 1e-14 w T^2 (w the largest
// weight below): the decided bits are the exact test's bits for solids whose
// squared axis ratio stays below 1e4.  The roots' own rounding (some 1e-13 T
// after the 1e-3 relative separation the tangency test leaves) is far inside
// the margin.

__host__ __device__ inline bool row_quadric(int shape) {
    return shape == SDFIBM_SHAPE_SPHERE || shape == SDFIBM_SHAPE_ELLIPSOID || shape == SDFIBM_SHAPE_CIRCLE ||
           shape == SDFIBM_SHAPE_ELLIPSE;
}

// Inside bits (bit v for vertex v) of the row of n <= 64 vertices at x = xs[0..n), y, z.
// The predicate reads f = sum_d w_d (off_d + l_d)^2 + e < 0 with the local point
// l = m t + c1 (y - c.y) + c2 (z - c.z), (m, c1, c2) = K.qc.  The exact tests of
// the undecided vertices run after the interval pass.
__host__ __device__ inline unsigned long long row_inside_bits(const SolidK& K, const double* xs, int n, double y,
                                                              double z) {
    double w[3] = {1.0, 1.0, 1.0}, off[3] = {K.com[0], K.com[1], K.com[2]}, e;
    switch (K.shape) {
    case SDFIBM_SHAPE_SPHERE: e = -K.k[1]; break;
    case SDFIBM_SHAPE_CIRCLE:  // q = com + (l.x, l.y, 0)
        w[2] = 0.0;
        e = K.com[2] * K.com[2] - K.k[1];
        break;
    case SDFIBM_SHAPE_ELLIPSOID:
        for (int d = 0; d < 3; ++d) {
            w[d] = K.k[d];
            off[d] = 0.0;
        }
        e = -1.0;
        break;
    default:  // ellipse: q = com + l, the z term absent
        w[0] = K.k[0];
        w[1] = K.k[1];
        w[2] = 0.0;
        e = -1.0;
        break;
    }
    const double uy = y - K.c[1], uz = z - K.c[2];
    double A = 0.0, B = 0.0, C = e;
    for (int d = 0; d < 3; ++d) {
        const double md = K.qc[d], a = off[d] + (K.qc[3 + d] * uy + K.qc[6 + d] * uz);
        A += w[d] * md * md;
        B += w[d] * md * a;
        C += w[d] * a * a;
    }
    const double disc = B * B - A * C, S = B * B + fabs(A * C);
    const unsigned long long all = n >= 64 ? ~0ull : (1ull << n) - 1ull;
    unsigned long long bits = 0ull, undecided = all;
    if (A > 0.0 && fabs(disc) <= 1e300 && S <= 1e300) {
        if (disc <= -1e-6 * S) return 0ull;  // the row misses the solid
        if (disc >= 1e-6 * S) {
            const double sq = sqrt(disc);
            const double t1 = (-B - sq) / A, t2 = (-B + sq) / A;
            const double T = fmax(fmax(fabs(t1), fabs(t2)), fmax(fabs(uy), fabs(uz)));
            const double mg = 1e-6 * T;
            const double in_lo = t1 + mg, in_hi = t2 - mg, out_lo = t1 - mg, out_hi = t2 + mg;
            undecided = 0ull;
            for (int v = 0; v < n; ++v) {
                const double t = xs[v] - K.c[0];
                const bool in = t > in_lo && t < in_hi;
                const bool out = t < out_lo || t > out_hi;
                bits |= (unsigned long long)in << v;
                undecided |= (unsigned long long)(!in && !out) << v;
            }
        }
    }
    while (undecided) {
        const int v = __builtin_ctzll(undecided);
        undecided &= undecided - 1ull;
        if (inside_world(K, d3{xs[v], y, z})) bits |= 1ull << v;
    }
    return bits;
}

// ---- signed distance in R ---------------------------------------------------
template <class R>
__host__ __device__ inline R filter_sd(R phi) {
    return (fabs(phi) < R(kTOL)) ? R(-kTOL) : phi;
}

template <class R>
__host__ __device__ inline R rect_sd(R x, R y, R ra, R rb) {
    const R dx = fabs(x) - ra;
    const R dy = fabs(y) - rb;
    const R dxp = fmax(R(0), dx);
    const R dyp = fmax(R(0), dy);
    return sqrt(dxp * dxp + dyp * dyp) + fmin(R(0), fmax(dx, dy));
}

// CSG signed distance in R before the filter: the primitives' distances (box
// and ellipsoid: sdf.h:43-57, 72-83; sphere: |x| - r; finite cylinder: the
// rectangle formula of sdf.h:29-37 in (radius, axial) coordinates) combined
// with the reference's operators (sdf.h:133-142: U = min, D(a, b) = max(a, -b),
// I = max), left to right.
template <class R>
__host__ __device__ inline R csg_prim_sd(const CsgPrimK& g, R px, R py, R pz) {
    const R x = px - R(g.c[0]), y = py - R(g.c[1]), z = pz - R(g.c[2]);
    switch (g.kind) {
    case SDFIBM_CSG_BOX: {
        const R dx = fabs(x) - R(g.h[0]), dy = fabs(y) - R(g.h[1]), dz = fabs(z) - R(g.h[2]);
        const R dxp = fmax(R(0), dx), dyp = fmax(R(0), dy), dzp = fmax(R(0), dz);
        return sqrt(dxp * dxp + dyp * dyp + dzp * dzp) + fmin(R(0), fmax(dz, fmax(dx, dy)));
    }
    case SDFIBM_CSG_ELLIPSOID: {
        const R ai = R(g.k[0]), bi = R(g.k[1]), ci = R(g.k[2]);
        const R xa = x * x * ai, yb = y * y * bi, zc = z * z * ci;
        return R(0.5) * (xa + yb + zc - R(1)) / sqrt(xa * ai + yb * bi + zc * ci);
    }
    case SDFIBM_CSG_SPHERE: return sqrt(x * x + y * y + z * z) - R(g.h[0]);
    case SDFIBM_CSG_CYLINDER: {
        const R u = g.axis == 0 ? y : x, v = g.axis == 2 ? y : z, w = g.axis == 0 ? x : (g.axis == 1 ? y : z);
        return rect_sd<R>(sqrt(u * u + v * v), w, R(g.h[0 == g.axis ? 1 : 0]), R(g.h[g.axis]));
    }
    default: return R(1);
    }
}

template <class R>
__host__ __device__ inline R csg_sd(const CsgK& G, R px, R py, R pz) {
    R d = csg_prim_sd<R>(G.p[0], px, py, pz);
    for (int t = 1; t < G.n; ++t) {
        const R e = csg_prim_sd<R>(G.p[t], px, py, pz);
        d = G.p[t].op == SDFIBM_CSG_UNION ? fmin(d, e) : (G.p[t].op == SDFIBM_CSG_DIFFERENCE ? fmax(d, -e) : fmax(d, e));
    }
    return d;
}

// signedDistance of each shape, local coordinates given in R.
template <class R>
__host__ __device__ inline R sdf_local(const SolidK& K, R px, R py, R pz) {
    const R cx = R(K.com[0]), cy = R(K.com[1]), cz = R(K.com[2]);
    switch (K.shape) {
    case SDFIBM_SHAPE_SPHERE: {
        const R x = cx + px, y = cy + py, z = cz + pz;
        return filter_sd<R>(sqrt(x * x + y * y + z * z) - R(K.k[0]));
    }
    case SDFIBM_SHAPE_CIRCLE: {
        const R x = cx + px, y = cy + py, z = cz;
        return filter_sd<R>(sqrt(x * x + y * y + z * z) - R(K.k[0]));
    }
    case SDFIBM_SHAPE_ELLIPSOID: {
        const R ai = R(K.k[0]), bi = R(K.k[1]), ci = R(K.k[2]);
        const R xa = px * px * ai, yb = py * py * bi, zc = pz * pz * ci;
        return filter_sd<R>(R(0.5) * (xa + yb + zc - R(1)) / sqrt(xa * ai + yb * bi + zc * ci));
    }
    case SDFIBM_SHAPE_ELLIPSE: {
        const R ai = R(K.k[0]), bi = R(K.k[1]);
        const R x = cx + px, y = cy + py;
        const R xa = x * x * ai, yb = y * y * bi;
        return filter_sd<R>(R(0.5) * (xa + yb - R(1)) / sqrt(xa * ai + yb * bi));
    }
    case SDFIBM_SHAPE_BOX: {
        const R x = cx + px, y = cy + py, z = cz + pz;
        const R dx = fabs(x) - R(K.k[0]), dy = fabs(y) - R(K.k[1]), dz = fabs(z) - R(K.k[2]);
        const R dxp = fmax(R(0), dx), dyp = fmax(R(0), dy), dzp = fmax(R(0), dz);
        return filter_sd<R>(sqrt(dxp * dxp + dyp * dyp + dzp * dzp) + fmin(R(0), fmax(dz, fmax(dx, dy))));
    }
    case SDFIBM_SHAPE_RECTANGLE: {
        const R x = cx + px, y = cy + py;
        return filter_sd<R>(rect_sd<R>(x, y, R(K.k[0]), R(K.k[1])));
    }
    case SDFIBM_SHAPE_PLANE: return py;
    case SDFIBM_SHAPE_CIRCLE_TAIL: {
        const R x = cx + px, y = cy + py;
        const R d1 = sqrt(x * x + y * y) - R(K.k[0]);
        const R d2 = rect_sd<R>(x - R(K.k[2]), y, R(K.k[2]), R(K.k[3]));
        return filter_sd<R>(fmin(d1, d2));
    }
    case SDFIBM_SHAPE_CIRCLE_TWOTAIL: {
        const R x = cx + px, y = cy + py;
        const R dc = sqrt(x * x + y * y) - R(K.k[0]);
        const R c30 = R(0.866025404), h = R(0.5);
        const R r1x = c30 * x + h * y, r1y = c30 * y - h * x;
        const R r2x = c30 * x + h * (-y), r2y = c30 * (-y) - h * x;
        const R d1 = rect_sd<R>(r1x - R(K.k[2]), r1y, R(K.k[2]), R(K.k[3]));
        const R d2 = rect_sd<R>(r2x - R(K.k[2]), r2y, R(K.k[2]), R(K.k[3]));
        return filter_sd<R>(fmin(fmin(dc, d1), d2));
    }
    case SDFIBM_SHAPE_CSG: return filter_sd<R>(csg_sd<R>(*K.csg, px, py, pz));
    default: return R(1);
    }
}

// FP32 signed distance with the cancelling part in FP64: the value near the
// surface is a small difference of O(1) terms (|p|^2 - r^2, sum x_i^2/a_i^2 - 1,
// |x| - a), which is formed in FP64 and rounded once; only square roots and
// divides of well-conditioned quantities run in FP32, on the hardware's 1-ulp
// square root, reciprocal and reciprocal square root.  Same function as the
// reference formula up to rounding.
// A sphere's distance from its squared centre distance d2 (radius r, r2 = r^2):
// (d2 - r2) / (sqrt(d2) + r), the difference formed in FP64.
__device__ inline float sphere_sd_mixed(double d2, double r, double r2) {
    const float num = (float)(d2 - r2);
    return num * __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf((float)d2) + (float)r);
}

__device__ inline float sdf_local_mixed(const SolidK& K, d3 p) {
    const double cx = K.com[0], cy = K.com[1], cz = K.com[2];
    auto sph = [](double d2, double r, double r2) -> float { return sphere_sd_mixed(d2, r, r2); };
    auto rect = [](double x, double y, double ra, double rb) -> float {
        const float dx = (float)(fabs(x) - ra), dy = (float)(fabs(y) - rb);
        const float dxp = fmaxf(0.f, dx), dyp = fmaxf(0.f, dy);
        return __builtin_amdgcn_sqrtf(dxp * dxp + dyp * dyp) + fminf(0.f, fmaxf(dx, dy));
    };
    switch (K.shape) {
    case SDFIBM_SHAPE_SPHERE: {
        const double x = cx + p.x, y = cy + p.y, z = cz + p.z;
        return filter_sd<float>(sph(x * x + y * y + z * z, K.k[0], K.k[1]));
    }
    case SDFIBM_SHAPE_CIRCLE: {
        const double x = cx + p.x, y = cy + p.y, z = cz;
        return filter_sd<float>(sph(x * x + y * y + z * z, K.k[0], K.k[1]));
    }
    case SDFIBM_SHAPE_ELLIPSOID: {
        const double xa = p.x * p.x * K.k[0], yb = p.y * p.y * K.k[1], zc = p.z * p.z * K.k[2];
        const float qm1 = (float)(xa + yb + zc - 1.0);
        const float g = (float)(xa * K.k[0] + yb * K.k[1] + zc * K.k[2]);
        return filter_sd<float>(0.5f * qm1 * __builtin_amdgcn_rsqf(g));
    }
    case SDFIBM_SHAPE_ELLIPSE: {
        const double x = cx + p.x, y = cy + p.y;
        const double xa = x * x * K.k[0], yb = y * y * K.k[1];
        const float qm1 = (float)(xa + yb - 1.0);
        const float g = (float)(xa * K.k[0] + yb * K.k[1]);
        return filter_sd<float>(0.5f * qm1 * __builtin_amdgcn_rsqf(g));
    }
    case SDFIBM_SHAPE_BOX: {
        const double x = cx + p.x, y = cy + p.y, z = cz + p.z;
        const float dx = (float)(fabs(x) - K.k[0]), dy = (float)(fabs(y) - K.k[1]), dz = (float)(fabs(z) - K.k[2]);
        const float dxp = fmaxf(0.f, dx), dyp = fmaxf(0.f, dy), dzp = fmaxf(0.f, dz);
        return filter_sd<float>(__builtin_amdgcn_sqrtf(dxp * dxp + dyp * dyp + dzp * dzp) +
                                fminf(0.f, fmaxf(dz, fmaxf(dx, dy))));
    }
    case SDFIBM_SHAPE_RECTANGLE: return filter_sd<float>(rect(cx + p.x, cy + p.y, K.k[0], K.k[1]));
    case SDFIBM_SHAPE_PLANE: return (float)p.y;
    case SDFIBM_SHAPE_CIRCLE_TAIL: {
        const double x = cx + p.x, y = cy + p.y;
        const float d1 = sph(x * x + y * y, K.k[0], K.k[1]);
        const float d2 = rect(x - K.k[2], y, K.k[2], K.k[3]);
        return filter_sd<float>(fminf(d1, d2));
    }
    case SDFIBM_SHAPE_CIRCLE_TWOTAIL: {
        const double x = cx + p.x, y = cy + p.y;
        const float dc = sph(x * x + y * y, K.k[0], K.k[1]);
        const double r1x = 0.866025404 * x + 0.5 * y, r1y = 0.866025404 * y - 0.5 * x;
        const double r2x = 0.866025404 * x + 0.5 * (-y), r2y = 0.866025404 * (-y) - 0.5 * x;
        const float d1 = rect(r1x - K.k[2], r1y, K.k[2], K.k[3]);
        const float d2 = rect(r2x - K.k[2], r2y, K.k[2], K.k[3]);
        return filter_sd<float>(fminf(fminf(dc, d1), d2));
    }
    case SDFIBM_SHAPE_CSG: {
        const CsgK& G = *K.csg;
        auto prim = [&](const CsgPrimK& g) -> float {
            const double x = p.x - g.c[0], y = p.y - g.c[1], z = p.z - g.c[2];
            switch (g.kind) {
            case SDFIBM_CSG_BOX: {
                const float dx = (float)(fabs(x) - g.h[0]), dy = (float)(fabs(y) - g.h[1]), dz = (float)(fabs(z) - g.h[2]);
                const float dxp = fmaxf(0.f, dx), dyp = fmaxf(0.f, dy), dzp = fmaxf(0.f, dz);
                return __builtin_amdgcn_sqrtf(dxp * dxp + dyp * dyp + dzp * dzp) + fminf(0.f, fmaxf(dz, fmaxf(dx, dy)));
            }
            case SDFIBM_CSG_ELLIPSOID: {
                const double xa = x * x * g.k[0], yb = y * y * g.k[1], zc = z * z * g.k[2];
                const float qm1 = (float)(xa + yb + zc - 1.0);
                const float gg = (float)(xa * g.k[0] + yb * g.k[1] + zc * g.k[2]);
                return 0.5f * qm1 * __builtin_amdgcn_rsqf(gg);
            }
            case SDFIBM_CSG_SPHERE: return sph(x * x + y * y + z * z, g.h[0], g.k[0]);
            case SDFIBM_CSG_CYLINDER: {
                const double u = g.axis == 0 ? y : x, v = g.axis == 2 ? y : z, w = g.axis == 0 ? x : (g.axis == 1 ? y : z);
                const float dr = sph(u * u + v * v, g.h[g.axis == 0 ? 1 : 0], g.k[0]);
                const float dw = (float)(fabs(w) - g.h[g.axis]);
                const float drp = fmaxf(0.f, dr), dwp = fmaxf(0.f, dw);
                return __builtin_amdgcn_sqrtf(drp * drp + dwp * dwp) + fminf(0.f, fmaxf(dr, dw));
            }
            default: return 1.f;
            }
        };
        float d = prim(G.p[0]);
        for (int t = 1; t < G.n; ++t) {
            const float e = prim(G.p[t]);
            d = G.p[t].op == SDFIBM_CSG_UNION ? fminf(d, e) : (G.p[t].op == SDFIBM_CSG_DIFFERENCE ? fmaxf(d, -e) : fmaxf(d, e));
        }
        return filter_sd<float>(d);
    }
    default: return 1.f;
    }
}

template <class R>
__device__ inline R sdf_world(const SolidK& K, d3 p) {
    const d3 l = world_to_local(K, p);
    if constexpr (sizeof(R) == sizeof(float))
        return sdf_local_mixed(K, l);
    else
        return sdf_local<R>(K, R(l.x), R(l.y), R(l.z));
}

// Local-frame axis-aligned bounds of the inside region: centre lc and half
// extents lh (lh = +inf along z for the z-extruded 2-D shapes).  Returns false
// for unbounded shapes (plane).
__host__ __device__ inline bool local_bounds(const SolidK& K, d3& lc, d3& lh) {
    const double INF = __builtin_huge_val();
    const d3 mcom{-K.com[0], -K.com[1], -K.com[2]};
    switch (K.shape) {
    case SDFIBM_SHAPE_SPHERE: lc = mcom; lh = {K.k[0], K.k[0], K.k[0]}; return true;
    case SDFIBM_SHAPE_CIRCLE: lc = mcom; lh = {K.k[0], K.k[0], INF}; return true;
    case SDFIBM_SHAPE_ELLIPSOID: {
        lc = {0, 0, 0};
        lh = {1.0 / sqrt(K.k[0]), 1.0 / sqrt(K.k[1]), 1.0 / sqrt(K.k[2])};
        return true;
    }
    case SDFIBM_SHAPE_ELLIPSE: lc = mcom; lh = {1.0 / sqrt(K.k[0]), 1.0 / sqrt(K.k[1]), INF}; return true;
    case SDFIBM_SHAPE_BOX: lc = mcom; lh = {K.k[0], K.k[1], K.k[2]}; return true;
    case SDFIBM_SHAPE_RECTANGLE: lc = mcom; lh = {K.k[0], K.k[1], INF}; return true;
    case SDFIBM_SHAPE_CIRCLE_TAIL:
    case SDFIBM_SHAPE_CIRCLE_TWOTAIL: {
        // circle of radius r and tail rectangle(s) [0, 2 ra] x [-rb, rb] (rotated by
        // about +-30 degrees for the two-tail shape): bound by the disc of radius
        // max(r, |(2 ra, rb)|) around -com.
        const double r = K.k[0], L = 2.0 * K.k[2], rb = K.k[3];
        const double R = fmax(r, sqrt(L * L + rb * rb));
        lc = mcom;
        lh = {R, R, INF};
        return true;
    }
    case SDFIBM_SHAPE_CSG: {
        // the box of the first primitive grown by the united ones and cut by the intersected ones
        // (a difference removes inside points only)
        const CsgK& G = *K.csg;
        double lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = G.p[0].c[d] - G.p[0].h[d];
            hi[d] = G.p[0].c[d] + G.p[0].h[d];
        }
        for (int t = 1; t < G.n; ++t)
            for (int d = 0; d < 3; ++d) {
                const double a = G.p[t].c[d] - G.p[t].h[d], b = G.p[t].c[d] + G.p[t].h[d];
                if (G.p[t].op == SDFIBM_CSG_UNION) {
                    lo[d] = fmin(lo[d], a);
                    hi[d] = fmax(hi[d], b);
                } else if (G.p[t].op == SDFIBM_CSG_INTERSECTION) {
                    lo[d] = fmax(lo[d], a);
                    hi[d] = fmin(hi[d], b);
                }
            }
        for (int d = 0; d < 3; ++d)
            if (!(hi[d] >= lo[d])) hi[d] = lo[d];  // empty intersection: a point, no vertex inside
        lc = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
        lh = {0.5 * (hi[0] - lo[0]), 0.5 * (hi[1] - lo[1]), 0.5 * (hi[2] - lo[2])};
        return true;
    }
    default: return false;
    }
}

}  // namespace sdfibm_dev
