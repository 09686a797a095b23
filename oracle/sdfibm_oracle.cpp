// sdfibm_oracle.cpp -- CPU FP64 oracle for the solid-fluid coupling step.
//
// TEST INFRASTRUCTURE ONLY (see sdfibm_oracle.h): the checker that the HIP
// product path is compared against, and the timed CPU baseline.  Never linked
// into the product library.
//
// Restates, in FP64 and in the reference's operation order:
//   * src/libshape/sdf/sdf.h:9-155           SDF primitives, filter, TOL
//   * src/libshape/{sphere,ellipsoid,circle,ellipse,box,rectangle,plane,
//     circle_tail,circle_twotail}.h          isInside / signedDistance
//   * src/libshape/ishape.h:43-52            world2local = conj(q).transform(p - t)
//   * src/solid.h:110-124                    phi01 / phi / evalPointVelocity
//   * src/cellenumerator.{h,cpp}             BFS candidate enumeration + seed fallback
//   * src/geometrictools.cpp:6-116           apex / face-area / cell-volume fraction
//   * src/solidcloud.cpp:361-464, 564-570    per-solid forcing, reduction, clip
//   * src/types.h:19                         SMALL = 1e-6
// and, for the third-party parts the reference calls into (OpenFOAM-12,
// not vendored): quaternion product/transform, primitiveMesh cell/face
// addressing and face/cell centre/area/volume, meshSearch::findNearestCell.
//
// Built with -ffp-contract=off so that no multiply-add is fused.

#include "sdfibm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <queue>
#include <set>
#include <unordered_map>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// vector algebra with OpenFOAM's evaluation order
// ---------------------------------------------------------------------------
struct V3 {
    double x, y, z;
};
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline double magSqr(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double mag(V3 a) { return std::sqrt(magSqr(a)); }

// ---------------------------------------------------------------------------
// quaternion: OpenFOAM quaternion (w, v); transform(u) = (mulq0v(u) * conj(q)).v
// ---------------------------------------------------------------------------
// world2local(p) = conjugate(q).transform(p - t)   (ishape.h:43-46)
static inline V3 world_to_local(const double* q, const double* t, V3 p) {
    const V3 u = p - V3{t[0], t[1], t[2]};
    const double aw = q[0];
    const V3 av{-q[1], -q[2], -q[3]};  // conjugate(q)
    // m = a.mulq0v(u) = ( -(a.v & u), a.w*u + (a.v ^ u) )
    const double mw = -dot(av, u);
    const V3 mv = (aw * u) + cross(av, u);
    // (m * conjugate(a)).v = m.w*ca.v + ca.w*m.v + (m.v ^ ca.v), ca = (aw, -av)
    const V3 cav{q[1], q[2], q[3]};
    return (mw * cav + aw * mv) + cross(mv, cav);
}

// ---------------------------------------------------------------------------
// shapes
// ---------------------------------------------------------------------------
static const double TOL = 1e-8;   // sdf.h:9
static const double SMALL = 1e-6; // types.h:19

static inline double filter(double phi) { return (std::fabs(phi) < TOL) ? -TOL : phi; }

static inline bool rect_bool(V3 p, double ra, double rb) {
    return std::fabs(p.x) < ra && std::fabs(p.y) < rb;
}
static inline double rect_sd(V3 p, double ra, double rb) {
    const double dx = std::fabs(p.x) - ra;
    const double dy = std::fabs(p.y) - rb;
    const double dxp = std::max(0.0, dx);
    const double dyp = std::max(0.0, dy);
    return std::sqrt(dxp * dxp + dyp * dyp) + std::min(0.0, std::max(dx, dy));
}
static inline V3 rot30(V3 p) {
    return {0.866025404 * p.x + 0.5 * p.y, 0.866025404 * p.y - 0.5 * p.x, 0.0};
}
static inline V3 flipy(V3 p) { return {p.x, -p.y, p.z}; }

// CSG shapes: the table set by orc_set_csg.  Primitive predicates are strict
// comparisons without square roots; distances are the exact box (sdf.h:43-57),
// the reference's first-order ellipsoid (sdf.h:72-83), the sphere and a finite
// cylinder (sdf.h:29-37's rectangle in radius/axial coordinates); the operators
// are sdf.h:133-142 (U: or / min; D(a, b): a and not b / max(a, -b); I: and / max).
static std::vector<orc_csg_shape> g_csg;

static inline void csg_frame(const orc_csg_prim& g, V3 p, double& x, double& y, double& z) {
    x = p.x - g.center[0];
    y = p.y - g.center[1];
    z = p.z - g.center[2];
}
// radial pair and axial coordinate of a point for a cylinder along `axis`
static inline void cyl_coords(int axis, double x, double y, double z, double& u, double& v, double& w) {
    if (axis == 0) { u = y; v = z; w = x; }
    else if (axis == 1) { u = x; v = z; w = y; }
    else { u = x; v = y; w = z; }
}

static bool csg_prim_in(const orc_csg_prim& g, V3 p) {
    double x, y, z;
    csg_frame(g, p, x, y, z);
    switch (g.kind) {
    case ORC_CSG_BOX: return std::fabs(x) < g.size[0] && std::fabs(y) < g.size[1] && std::fabs(z) < g.size[2];
    case ORC_CSG_ELLIPSOID: {
        const double ai = 1.0 / (g.size[0] * g.size[0]), bi = 1.0 / (g.size[1] * g.size[1]),
                     ci = 1.0 / (g.size[2] * g.size[2]);
        return x * x * ai + y * y * bi + z * z * ci < 1.0;
    }
    case ORC_CSG_SPHERE: return x * x + y * y + z * z < g.size[0] * g.size[0];
    case ORC_CSG_CYLINDER: {
        double u, v, w;
        cyl_coords(g.axis, x, y, z, u, v, w);
        return u * u + v * v < g.size[0] * g.size[0] && std::fabs(w) < g.size[1];
    }
    default: return false;
    }
}

static double csg_prim_dist(const orc_csg_prim& g, V3 p) {
    double x, y, z;
    csg_frame(g, p, x, y, z);
    switch (g.kind) {
    case ORC_CSG_BOX: {
        const double dx = std::fabs(x) - g.size[0], dy = std::fabs(y) - g.size[1], dz = std::fabs(z) - g.size[2];
        const double dxp = std::max(0.0, dx), dyp = std::max(0.0, dy), dzp = std::max(0.0, dz);
        return std::sqrt(dxp * dxp + dyp * dyp + dzp * dzp) + std::min(0.0, std::max(dz, std::max(dx, dy)));
    }
    case ORC_CSG_ELLIPSOID: {
        const double ai = 1.0 / (g.size[0] * g.size[0]), bi = 1.0 / (g.size[1] * g.size[1]),
                     ci = 1.0 / (g.size[2] * g.size[2]);
        const double xa = x * x * ai, yb = y * y * bi, zc = z * z * ci;
        return 0.5 * (xa + yb + zc - 1.0) / (std::sqrt(xa * ai + yb * bi + zc * ci));
    }
    case ORC_CSG_SPHERE: return std::sqrt(x * x + y * y + z * z) - g.size[0];
    case ORC_CSG_CYLINDER: {
        double u, v, w;
        cyl_coords(g.axis, x, y, z, u, v, w);
        return rect_sd(V3{std::sqrt(u * u + v * v), w, 0.0}, g.size[0], g.size[1]);
    }
    default: return 1.0;
    }
}

static bool csg_in(const orc_csg_shape& G, V3 p) {
    bool in = csg_prim_in(G.prim[0], p);
    for (int t = 1; t < G.n_prims; ++t) {
        const bool b = csg_prim_in(G.prim[t], p);
        switch (G.prim[t].op) {
        case ORC_CSG_UNION: in = in || b; break;
        case ORC_CSG_DIFFERENCE: in = in && !b; break;
        default: in = in && b; break;
        }
    }
    return in;
}

static double csg_dist(const orc_csg_shape& G, V3 p) {
    double d = csg_prim_dist(G.prim[0], p);
    for (int t = 1; t < G.n_prims; ++t) {
        const double e = csg_prim_dist(G.prim[t], p);
        switch (G.prim[t].op) {
        case ORC_CSG_UNION: d = std::min(d, e); break;
        case ORC_CSG_DIFFERENCE: d = std::max(d, -e); break;
        default: d = std::max(d, e); break;
        }
    }
    return d;
}

static const orc_csg_shape* csg_of(const orc_solid& s) {
    const long long k = (long long)s.prm[0];
    return (k >= 0 && k < (long long)g_csg.size()) ? &g_csg[k] : nullptr;
}

static inline V3 com_of(const orc_solid& s) {
    switch (s.shape) {
    case ORC_SPHERE:
    case ORC_CIRCLE: return {s.prm[1], s.prm[2], s.prm[3]};
    case ORC_ELLIPSE:
    case ORC_RECTANGLE:
    case ORC_BOX:
    case ORC_CIRCLE_TAIL:
    case ORC_CIRCLE_TWOTAIL: return {s.prm[3], s.prm[4], s.prm[5]};
    default: return {0.0, 0.0, 0.0};
    }
}

// isInside in local coordinates (strict comparisons, no sqrt)
static bool is_inside_local(const orc_solid& s, V3 p) {
    switch (s.shape) {
    case ORC_SPHERE: {  // sphere.h:37-40
        const double r = s.prm[0];
        return magSqr(com_of(s) + p) < r * r;
    }
    case ORC_CIRCLE: {  // circle.h:38-41
        const double r = s.prm[0];
        return magSqr(com_of(s) + V3{p.x, p.y, 0.0}) < r * r;
    }
    case ORC_ELLIPSOID: {  // ellipsoid.h:43-46
        const double ai = 1.0 / (s.prm[0] * s.prm[0]);
        const double bi = 1.0 / (s.prm[1] * s.prm[1]);
        const double ci = 1.0 / (s.prm[2] * s.prm[2]);
        return p.x * p.x * ai + p.y * p.y * bi + p.z * p.z * ci < 1.0;
    }
    case ORC_ELLIPSE: {  // ellipse.h:40-45
        V3 q = com_of(s) + p;
        q.z = 0.0;
        const double ai = 1.0 / (s.prm[0] * s.prm[0]);
        const double bi = 1.0 / (s.prm[1] * s.prm[1]);
        return q.x * q.x * ai + q.y * q.y * bi < 1.0;
    }
    case ORC_BOX: {  // box.h:40-48
        const V3 q = com_of(s) + p;
        return std::fabs(q.x) < s.prm[0] && std::fabs(q.y) < s.prm[1] && std::fabs(q.z) < s.prm[2];
    }
    case ORC_RECTANGLE: {  // rectangle.h:41-48
        V3 q = com_of(s) + p;
        q.z = 0.0;
        return rect_bool(q, s.prm[0], s.prm[1]);
    }
    case ORC_CSG: {
        const orc_csg_shape* G = csg_of(s);
        return G ? csg_in(*G, p) : false;
    }
    case ORC_PLANE: return p.y < 0;  // plane.h:20-23
    case ORC_CIRCLE_TAIL: {          // circle_tail.h:42-50
        V3 q = com_of(s) + p;
        q.z = 0.0;
        const double r = s.prm[0];
        const double ra = (s.prm[1] + 1) * 0.5 * r;
        const double rb = s.prm[2];
        const bool b1 = rect_bool(q - V3{ra, 0.0, 0.0}, ra, rb);
        const bool b2 = magSqr(V3{q.x, q.y, 0.0}) < r * r;
        return b1 || b2;
    }
    case ORC_CIRCLE_TWOTAIL: {  // circle_twotail.h:42-53
        V3 q = com_of(s) + p;
        q.z = 0.0;
        const double r = s.prm[0];
        const double ra = (s.prm[1] + 1) * 0.5 * r;
        const double rb = s.prm[2] * 0.5;
        const bool dc = magSqr(q) < r * r;
        const bool d1 = rect_bool(rot30(q) - V3{ra, 0, 0}, ra, rb);
        const bool d2 = rect_bool(rot30(flipy(q)) - V3{ra, 0, 0}, ra, rb);
        return dc || d1 || d2;
    }
    default: return false;
    }
}

// signedDistance in local coordinates (negative inside)
static double signed_distance_local(const orc_solid& s, V3 p) {
    switch (s.shape) {
    case ORC_SPHERE: return filter(mag(com_of(s) + p) - s.prm[0]);  // sphere.h:42-45
    case ORC_CIRCLE:
        return filter(mag(com_of(s) + V3{p.x, p.y, 0.0}) - s.prm[0]);  // circle.h:43-46
    case ORC_ELLIPSOID: {                                                // ellipsoid.h:48-53
        const double ai = 1.0 / (s.prm[0] * s.prm[0]);
        const double bi = 1.0 / (s.prm[1] * s.prm[1]);
        const double ci = 1.0 / (s.prm[2] * s.prm[2]);
        const double xa = p.x * p.x * ai;
        const double yb = p.y * p.y * bi;
        const double zc = p.z * p.z * ci;
        return filter(0.5 * (xa + yb + zc - 1.0) / (std::sqrt(xa * ai + yb * bi + zc * ci)));
    }
    case ORC_ELLIPSE: {  // ellipse.h:47-54
        V3 q = com_of(s) + p;
        q.z = 0.0;
        const double ai = 1.0 / (s.prm[0] * s.prm[0]);
        const double bi = 1.0 / (s.prm[1] * s.prm[1]);
        const double xa = q.x * q.x * ai;
        const double yb = q.y * q.y * bi;
        return filter(0.5 * (xa + yb - 1.0) / (std::sqrt(xa * ai + yb * bi)));
    }
    case ORC_BOX: {  // box.h:50-57
        const V3 q = com_of(s) + p;
        const double dx = std::fabs(q.x) - s.prm[0];
        const double dy = std::fabs(q.y) - s.prm[1];
        const double dz = std::fabs(q.z) - s.prm[2];
        const double dxp = std::max(0.0, dx);
        const double dyp = std::max(0.0, dy);
        const double dzp = std::max(0.0, dz);
        return filter(std::sqrt(dxp * dxp + dyp * dyp + dzp * dzp) +
                      std::min(0.0, std::max(dz, std::max(dx, dy))));
    }
    case ORC_RECTANGLE: {  // rectangle.h:50-58
        V3 q = com_of(s) + p;
        q.z = 0.0;
        return filter(rect_sd(q, s.prm[0], s.prm[1]));
    }
    case ORC_CSG: {
        const orc_csg_shape* G = csg_of(s);
        return G ? filter(csg_dist(*G, p)) : 1.0;
    }
    case ORC_PLANE: return p.y;  // plane.h:25-28 (no filter)
    case ORC_CIRCLE_TAIL: {      // circle_tail.h:52-61
        V3 q = com_of(s) + p;
        q.z = 0.0;
        const double r = s.prm[0];
        const double ra = (s.prm[1] + 1) * 0.5 * r;
        const double rb = s.prm[2];
        const double d1 = mag(q) - r;
        const double d2 = rect_sd(q - V3{ra, 0.0, 0.0}, ra, rb);
        return filter(std::min(d1, d2));
    }
    case ORC_CIRCLE_TWOTAIL: {  // circle_twotail.h:55-66
        V3 q = com_of(s) + p;
        q.z = 0.0;
        const double r = s.prm[0];
        const double ra = (s.prm[1] + 1) * 0.5 * r;
        const double rb = s.prm[2] * 0.5;
        const double dc = mag(q) - r;
        const double d1 = rect_sd(rot30(q) - V3{ra, 0, 0}, ra, rb);
        const double d2 = rect_sd(rot30(flipy(q)) - V3{ra, 0, 0}, ra, rb);
        return filter(std::min(std::min(dc, d1), d2));
    }
    default: return 1.0;
    }
}

static inline bool phi01(const orc_solid& s, V3 p) {
    return is_inside_local(s, world_to_local(s.q, s.center, p));
}
static inline double phi(const orc_solid& s, V3 p) {
    return signed_distance_local(s, world_to_local(s.q, s.center, p));
}

// ---------------------------------------------------------------------------
// mesh: polyMesh + primitiveMesh-style derived addressing and geometry
// ---------------------------------------------------------------------------
struct Mesh {
    int64_t nP = 0, nF = 0, nIF = 0, nC = 0;
    std::vector<V3> P;
    std::vector<int64_t> fOff;
    std::vector<int32_t> fPts;
    std::vector<int32_t> own, nei;
    // cell -> faces (OpenFOAM calcCells order: owned faces ascending, then neighbour faces ascending)
    std::vector<int64_t> cfOff;
    std::vector<int32_t> cf;
    // cell -> points, ascending labels
    std::vector<int64_t> cpOff;
    std::vector<int32_t> cp;
    // cell -> cells across internal faces, in face order
    std::vector<int64_t> ccOff;
    std::vector<int32_t> cc;
    // geometry
    std::vector<V3> Cf, Sf, C;
    std::vector<double> Vol;
    // bucket grid over cell centres for nearest-cell queries
    V3 gLo{}, gInv{};
    int gn[3] = {1, 1, 1};
    std::vector<int64_t> gOff;
    std::vector<int32_t> gCells;
};

// Face centre and area vector of a polygon (OpenFOAM primitiveMesh face
// geometry: triangle fan about the point average, centre weighted by the
// triangle areas projected on the face normal).
static void face_geometry(const Mesh& m, int64_t f, V3& fc, V3& sf) {
    const int64_t b = m.fOff[f], n = m.fOff[f + 1] - b;
    const int32_t* lab = &m.fPts[b];
    if (n == 3) {
        const V3 p0 = m.P[lab[0]], p1 = m.P[lab[1]], p2 = m.P[lab[2]];
        fc = (1.0 / 3.0) * ((p0 + p1) + p2);
        sf = 0.5 * cross(p1 - p0, p2 - p0);
        return;
    }
    V3 fCentre = m.P[lab[0]];
    for (int64_t i = 1; i < n; ++i) fCentre = fCentre + m.P[lab[i]];
    fCentre = fCentre / double(n);
    V3 sumN{0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        const V3 a = m.P[lab[i]], b2 = m.P[lab[(i + 1) % n]];
        sumN = sumN + cross(b2 - a, fCentre - a);
    }
    const double sumNmag = mag(sumN);
    if (!(sumNmag > 1e-300)) {
        fc = fCentre;
        sf = V3{0, 0, 0};
        return;
    }
    const V3 nHat = sumN * (1.0 / sumNmag);
    double sumA = 0.0;
    V3 sumAc{0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        const V3 a = m.P[lab[i]], b2 = m.P[lab[(i + 1) % n]];
        const V3 c = (a + b2) + fCentre;
        const double an = dot(cross(b2 - a, fCentre - a), nHat);
        sumA += an;
        sumAc = sumAc + an * c;
    }
    fc = ((1.0 / 3.0) * sumAc) / sumA;
    sf = 0.5 * sumN;
}

static void build_mesh(Mesh& m) {
    // faces geometry
    m.Cf.resize(m.nF);
    m.Sf.resize(m.nF);
    for (int64_t f = 0; f < m.nF; ++f) face_geometry(m, f, m.Cf[f], m.Sf[f]);

    // cells: owned faces ascending then neighbour faces ascending
    int64_t nC = 0;
    for (int64_t f = 0; f < m.nF; ++f) nC = std::max<int64_t>(nC, m.own[f] + 1);
    for (int64_t f = 0; f < m.nIF; ++f) nC = std::max<int64_t>(nC, m.nei[f] + 1);
    m.nC = nC;
    std::vector<int64_t> cnt(nC + 1, 0);
    for (int64_t f = 0; f < m.nF; ++f) cnt[m.own[f] + 1]++;
    for (int64_t f = 0; f < m.nIF; ++f) cnt[m.nei[f] + 1]++;
    m.cfOff.assign(nC + 1, 0);
    for (int64_t c = 0; c < nC; ++c) m.cfOff[c + 1] = m.cfOff[c] + cnt[c + 1];
    m.cf.resize(m.cfOff[nC]);
    std::vector<int64_t> fill(m.cfOff.begin(), m.cfOff.end() - 1);
    for (int64_t f = 0; f < m.nF; ++f) m.cf[fill[m.own[f]]++] = (int32_t)f;
    for (int64_t f = 0; f < m.nIF; ++f) m.cf[fill[m.nei[f]]++] = (int32_t)f;

    // cell points ascending (invert point->cells built in cell order)
    m.cpOff.assign(nC + 1, 0);
    std::vector<int32_t> tmp;
    std::vector<std::vector<int32_t>> per(nC);
    for (int64_t c = 0; c < nC; ++c) {
        tmp.clear();
        for (int64_t k = m.cfOff[c]; k < m.cfOff[c + 1]; ++k) {
            const int32_t f = m.cf[k];
            for (int64_t j = m.fOff[f]; j < m.fOff[f + 1]; ++j) tmp.push_back(m.fPts[j]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        m.cpOff[c + 1] = m.cpOff[c] + (int64_t)tmp.size();
        per[c] = tmp;
    }
    m.cp.resize(m.cpOff[nC]);
    for (int64_t c = 0; c < nC; ++c) std::copy(per[c].begin(), per[c].end(), m.cp.begin() + m.cpOff[c]);
    per.clear();
    per.shrink_to_fit();

    // cell cells over internal faces, face order
    std::vector<int64_t> ncc(nC + 1, 0);
    for (int64_t f = 0; f < m.nIF; ++f) {
        ncc[m.own[f] + 1]++;
        ncc[m.nei[f] + 1]++;
    }
    m.ccOff.assign(nC + 1, 0);
    for (int64_t c = 0; c < nC; ++c) m.ccOff[c + 1] = m.ccOff[c] + ncc[c + 1];
    m.cc.resize(m.ccOff[nC]);
    std::vector<int64_t> f2(m.ccOff.begin(), m.ccOff.end() - 1);
    for (int64_t f = 0; f < m.nIF; ++f) {
        m.cc[f2[m.own[f]]++] = m.nei[f];
        m.cc[f2[m.nei[f]]++] = m.own[f];
    }

    // cell centres and volumes: face-pyramid decomposition about the mean of
    // the face centres (OpenFOAM primitiveMesh cell geometry)
    std::vector<V3> cEst(nC, V3{0, 0, 0});
    std::vector<int> nCF(nC, 0);
    for (int64_t f = 0; f < m.nF; ++f) {
        cEst[m.own[f]] = cEst[m.own[f]] + m.Cf[f];
        nCF[m.own[f]]++;
    }
    for (int64_t f = 0; f < m.nIF; ++f) {
        cEst[m.nei[f]] = cEst[m.nei[f]] + m.Cf[f];
        nCF[m.nei[f]]++;
    }
    for (int64_t c = 0; c < nC; ++c) cEst[c] = cEst[c] / double(nCF[c]);
    m.C.assign(nC, V3{0, 0, 0});
    m.Vol.assign(nC, 0.0);
    for (int64_t f = 0; f < m.nF; ++f) {
        const int32_t c = m.own[f];
        const double pyr3 = dot(m.Sf[f], m.Cf[f] - cEst[c]);
        const V3 pc = (3.0 / 4.0) * m.Cf[f] + (1.0 / 4.0) * cEst[c];
        m.C[c] = m.C[c] + pyr3 * pc;
        m.Vol[c] += pyr3;
    }
    for (int64_t f = 0; f < m.nIF; ++f) {
        const int32_t c = m.nei[f];
        const double pyr3 = dot(m.Sf[f], cEst[c] - m.Cf[f]);
        const V3 pc = (3.0 / 4.0) * m.Cf[f] + (1.0 / 4.0) * cEst[c];
        m.C[c] = m.C[c] + pyr3 * pc;
        m.Vol[c] += pyr3;
    }
    for (int64_t c = 0; c < nC; ++c) {
        if (std::fabs(m.Vol[c]) > 1e-300)
            m.C[c] = m.C[c] / m.Vol[c];
        else
            m.C[c] = cEst[c];
        m.Vol[c] *= (1.0 / 3.0);
    }

    // bucket grid for nearest-cell queries (~2 cells per bucket)
    V3 lo = m.C[0], hi = m.C[0];
    for (const V3& c : m.C) {
        lo = {std::min(lo.x, c.x), std::min(lo.y, c.y), std::min(lo.z, c.z)};
        hi = {std::max(hi.x, c.x), std::max(hi.y, c.y), std::max(hi.z, c.z)};
    }
    const double ext[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
    const double target = std::max<double>(1.0, double(nC) / 2.0);
    double vol = 1.0;
    int ndim = 0;
    for (int d = 0; d < 3; ++d)
        if (ext[d] > 0) {
            vol *= ext[d];
            ndim++;
        }
    const double side = ndim ? std::pow(vol / target, 1.0 / ndim) : 1.0;
    for (int d = 0; d < 3; ++d)
        m.gn[d] = ext[d] > 0 ? std::max(1, std::min(4096, (int)std::ceil(ext[d] / side))) : 1;
    m.gLo = lo;
    m.gInv = {ext[0] > 0 ? m.gn[0] / ext[0] : 0.0, ext[1] > 0 ? m.gn[1] / ext[1] : 0.0,
              ext[2] > 0 ? m.gn[2] / ext[2] : 0.0};
    const int64_t nb = (int64_t)m.gn[0] * m.gn[1] * m.gn[2];
    std::vector<int64_t> bc(nb + 1, 0);
    auto bucket_of = [&](V3 p, int ijk[3]) {
        const double v[3] = {(p.x - m.gLo.x) * m.gInv.x, (p.y - m.gLo.y) * m.gInv.y,
                             (p.z - m.gLo.z) * m.gInv.z};
        for (int d = 0; d < 3; ++d) {
            int i = (int)std::floor(v[d]);
            ijk[d] = std::min(std::max(i, 0), m.gn[d] - 1);
        }
    };
    std::vector<int64_t> bid(nC);
    for (int64_t c = 0; c < nC; ++c) {
        int ijk[3];
        bucket_of(m.C[c], ijk);
        bid[c] = ijk[0] + (int64_t)m.gn[0] * (ijk[1] + (int64_t)m.gn[1] * ijk[2]);
        bc[bid[c] + 1]++;
    }
    for (int64_t b = 0; b < nb; ++b) bc[b + 1] += bc[b];
    m.gOff = bc;
    m.gCells.resize(nC);
    std::vector<int64_t> gf(bc.begin(), bc.end() - 1);
    for (int64_t c = 0; c < nC; ++c) m.gCells[gf[bid[c]]++] = (int32_t)c;
}

// meshSearch::findNearestCell: the cell whose centre is nearest.  Ties --
// centres whose squared distance is within a relative 1e-10 of the minimum --
// go to the lowest cell index, so that rounding differences in the centres do
// not decide between equidistant cells.  Ring search over the bucket grid.
static const double kNearTie = 1e-10;
static int64_t find_nearest_cell(const Mesh& m, V3 p) {
    if (m.nC == 0) return -1;
    const double v[3] = {(p.x - m.gLo.x) * m.gInv.x, (p.y - m.gLo.y) * m.gInv.y,
                         (p.z - m.gLo.z) * m.gInv.z};
    int b0[3];
    for (int d = 0; d < 3; ++d) {
        double f = std::floor(v[d]);
        f = std::min(std::max(f, 0.0), double(m.gn[d] - 1));
        b0[d] = (int)f;
    }
    // bucket widths (for the ring stopping rule)
    double bw[3];
    bw[0] = m.gInv.x > 0 ? 1.0 / m.gInv.x : 0.0;
    bw[1] = m.gInv.y > 0 ? 1.0 / m.gInv.y : 0.0;
    bw[2] = m.gInv.z > 0 ? 1.0 / m.gInv.z : 0.0;
    double best = INFINITY;
    int64_t bestc = -1;
    const int maxr = std::max(m.gn[0], std::max(m.gn[1], m.gn[2]));
    for (int r = 0; r <= maxr; ++r) {
        for (int k = b0[2] - r; k <= b0[2] + r; ++k) {
            if (k < 0 || k >= m.gn[2]) continue;
            for (int j = b0[1] - r; j <= b0[1] + r; ++j) {
                if (j < 0 || j >= m.gn[1]) continue;
                for (int i = b0[0] - r; i <= b0[0] + r; ++i) {
                    if (i < 0 || i >= m.gn[0]) continue;
                    if (std::max(std::abs(i - b0[0]), std::max(std::abs(j - b0[1]), std::abs(k - b0[2]))) != r)
                        continue;
                    const int64_t b = i + (int64_t)m.gn[0] * (j + (int64_t)m.gn[1] * k);
                    for (int64_t q = m.gOff[b]; q < m.gOff[b + 1]; ++q) {
                        const int32_t c = m.gCells[q];
                        const double d2 = magSqr(m.C[c] - p);
                        if (d2 < best) best = d2;
                    }
                }
            }
        }
        // every unvisited bucket is at least r*min(width) away from p's bucket boundary
        double wmin = INFINITY;
        for (int d = 0; d < 3; ++d)
            if (m.gn[d] > 1) wmin = std::min(wmin, bw[d]);
        if (!std::isfinite(wmin)) wmin = 0.0;
        // distance from p to the outside of the (2r+1)^3 block around b0, lower bound
        double lb = INFINITY;
        const double lo3[3] = {m.gLo.x, m.gLo.y, m.gLo.z};
        const double pp[3] = {p.x, p.y, p.z};
        bool any = false;
        for (int d = 0; d < 3; ++d) {
            if (m.gn[d] <= 1) continue;
            any = true;
            const double lo_edge = lo3[d] + (b0[d] - r) * bw[d];
            const double hi_edge = lo3[d] + (b0[d] + r + 1) * bw[d];
            double dd = INFINITY;
            if (b0[d] - r > 0) dd = std::min(dd, pp[d] - lo_edge);
            if (b0[d] + r + 1 < m.gn[d]) dd = std::min(dd, hi_edge - pp[d]);
            lb = std::min(lb, dd);
        }
        const double lim = best * (1.0 + kNearTie) + 1e-300;
        const bool done = !any || !std::isfinite(lb) || (std::isfinite(best) && lb > 0 && lb * lb > lim);
        if (done) {
            // lowest index among the near-ties, over the buckets visited so far
            for (int k = b0[2] - r; k <= b0[2] + r; ++k) {
                if (k < 0 || k >= m.gn[2]) continue;
                for (int j = b0[1] - r; j <= b0[1] + r; ++j) {
                    if (j < 0 || j >= m.gn[1]) continue;
                    for (int i = b0[0] - r; i <= b0[0] + r; ++i) {
                        if (i < 0 || i >= m.gn[0]) continue;
                        const int64_t b = i + (int64_t)m.gn[0] * (j + (int64_t)m.gn[1] * k);
                        for (int64_t q = m.gOff[b]; q < m.gOff[b + 1]; ++q) {
                            const int32_t c = m.gCells[q];
                            if (magSqr(m.C[c] - p) <= lim && (bestc < 0 || c < bestc)) bestc = c;
                        }
                    }
                }
            }
            return bestc;
        }
    }
    return bestc;
}

// ---------------------------------------------------------------------------
// geometric tools (geometrictools.cpp:13-116), phi values passed in by the caller
// ---------------------------------------------------------------------------
static inline double line_fraction(double pa, double pb) {  // :13-23
    if (pa > 0 && pb > 0) return 0;
    if (pa <= 0 && pb <= 0) return 1;
    if (pa > 0) return -pb / (pa - pb);
    return -pa / (pb - pa);
}

// calcApex (:25-45): A = first vertex; B = first later vertex whose phi sign
// differs from A's (phiA*phiB <= 0), else the last vertex.
template <class PhiOf>
static V3 apex(const Mesh& m, const int32_t* lab, int64_t n, PhiOf phiOf) {
    const V3 A = m.P[lab[0]];
    const double pA = phiOf(lab[0]);
    V3 B{0, 0, 0};
    double pB = 0.0;
    for (int64_t i = 1; i < n; ++i) {
        B = m.P[lab[i]];
        pB = phiOf(lab[i]);
        if (pA * pB <= 0) break;
    }
    return A - (std::fabs(pA) / (SMALL + std::fabs(pA) + std::fabs(pB))) * (A - B);
}

template <class PhiOf>
static double face_area_fraction(const Mesh& m, int64_t f, PhiOf phiOf) {  // :98-116, :74-96
    const int64_t b = m.fOff[f], n = m.fOff[f + 1] - b;
    const int32_t* lab = &m.fPts[b];
    int sign_sum = 0;
    for (int64_t i = 0; i < n; ++i) sign_sum += (phiOf(lab[i]) > 0) ? 1 : -1;
    if (sign_sum == n) return 0.0;
    if (sign_sum == -n) return 1.0;
    const V3 ap = apex(m, lab, n, phiOf);
    double area = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double pO = phiOf(lab[i]);
        const double pA = phiOf(lab[(i + 1) % n]);
        const V3 O = m.P[lab[i]];
        const V3 A = m.P[lab[(i + 1) % n]];
        area += std::fabs(0.5 * mag(cross(A - O, ap - O))) * line_fraction(pO, pA);
    }
    return area / mag(m.Sf[f]);
}

template <class PhiOf>
static double cell_volume(const Mesh& m, int64_t c, bool twoD, PhiOf phiOf) {  // :47-72
    const int32_t* lab = &m.cp[m.cpOff[c]];
    const int64_t n = m.cpOff[c + 1] - m.cpOff[c];
    V3 ap = apex(m, lab, n, phiOf);
    if (twoD) ap.z = 0.0;
    double vol = 0.0;
    for (int64_t k = m.cfOff[c]; k < m.cfOff[c + 1]; ++k) {
        const int32_t f = m.cf[k];
        const double eps = face_area_fraction(m, f, phiOf);
        vol += (1.0 / 3.0) * eps * std::fabs(dot(ap - m.Cf[f], m.Sf[f]));
    }
    return vol;
}

// ---------------------------------------------------------------------------
// per-solid interaction (solidcloud.cpp:361-433)
// ---------------------------------------------------------------------------
enum CellType { UNVISITED = 0, ALL_INSIDE = 1, CENTER_INSIDE = 2, CENTER_OUTSIDE = 3, ALL_OUTSIDE = 4 };

struct Workspace {
    // fast mode: stamp-based visited marks and vertex phi cache reused across solids
    std::vector<uint32_t> stamp;
    std::vector<uint8_t> type;
    std::vector<uint32_t> vstamp;
    std::vector<double> vphi;
    uint32_t gen = 0;
};

static int count_inside(const Mesh& m, const orc_solid& s, int64_t c) {  // cellenumerator.cpp:36-45
    int n = 0;
    for (int64_t k = m.cpOff[c]; k < m.cpOff[c + 1]; ++k) n += phi01(s, m.P[m.cp[k]]) ? 1 : 0;
    return n;
}

struct SolidResult {
    std::vector<int64_t> lists[3];  // ALL_INSIDE, CENTER_INSIDE, CENTER_OUTSIDE ascending
    int64_t seed = -1;
};

// CellEnumerator (cellenumerator.{h,cpp}) in the reference's bookkeeping:
// an N-sized type array per solid, a queue, std::set per class.
static void enumerate_faithful(const Mesh& m, const orc_solid& s, SolidResult& out) {
    std::function<bool(const V3&)> pred = [&](const V3& v) { return phi01(s, v); };
    auto countv = [&](int64_t c) {
        int n = 0;
        for (int64_t k = m.cpOff[c]; k < m.cpOff[c + 1]; ++k) n += pred(m.P[m.cp[k]]);
        return n;
    };
    std::vector<int> ct(m.nC, UNVISITED);
    std::unordered_map<int, std::set<size_t>> is;
    std::queue<int> q;
    int64_t seed = find_nearest_cell(m, V3{s.center[0], s.center[1], s.center[2]});
    if (seed < 0 || countv(seed) == 0) {
        seed = -1;
        for (int64_t c = 0; c < m.nC; ++c)
            if (countv(c) > 0) {
                seed = c;
                break;
            }
    }
    if (seed >= 0) {
        q.push((int)seed);
        const int nin = countv(seed);
        if (nin == (int)(m.cpOff[seed + 1] - m.cpOff[seed]))
            ct[seed] = ALL_INSIDE;
        else if (pred(m.C[seed]))
            ct[seed] = CENTER_INSIDE;
        else
            ct[seed] = CENTER_OUTSIDE;
        is[ct[seed]].insert(seed);
    }
    while (!q.empty()) {
        const int icur = q.front();
        for (int64_t k = m.ccOff[icur]; k < m.ccOff[icur + 1]; ++k) {
            const int inb = m.cc[k];
            if (ct[inb] != UNVISITED) continue;
            const int nin = countv(inb);
            if (nin == 0) {
                ct[inb] = ALL_OUTSIDE;
                is[ct[inb]].insert(inb);
                continue;
            }
            q.push(inb);
            if (nin == (int)(m.cpOff[icur + 1] - m.cpOff[icur]))  // current cell's count (cellenumerator.cpp:25)
                ct[inb] = ALL_INSIDE;
            else if (pred(m.C[inb]))
                ct[inb] = CENTER_INSIDE;
            else
                ct[inb] = CENTER_OUTSIDE;
            is[ct[inb]].insert(inb);
        }
        q.pop();
    }
    for (int t = 0; t < 3; ++t) {
        auto it = is.find((CellType)(t + 1));
        out.lists[t].clear();
        if (it != is.end()) out.lists[t].assign(it->second.begin(), it->second.end());
    }
    out.seed = seed;
}

// Same sets with stamp-based visited marks (no N-sized fills per solid).
static void enumerate_fast(const Mesh& m, const orc_solid& s, Workspace& w, SolidResult& out) {
    if (w.stamp.size() != (size_t)m.nC) {
        w.stamp.assign(m.nC, 0);
        w.type.assign(m.nC, 0);
        w.gen = 0;
    }
    if (++w.gen == 0) {
        std::fill(w.stamp.begin(), w.stamp.end(), 0);
        w.gen = 1;
    }
    const uint32_t g = w.gen;
    for (auto& l : out.lists) l.clear();
    int64_t seed = find_nearest_cell(m, V3{s.center[0], s.center[1], s.center[2]});
    if (seed < 0 || count_inside(m, s, seed) == 0) {
        seed = -1;
        for (int64_t c = 0; c < m.nC; ++c)
            if (count_inside(m, s, c) > 0) {
                seed = c;
                break;
            }
    }
    out.seed = seed;
    if (seed < 0) return;
    std::vector<int64_t> queue;
    queue.push_back(seed);
    w.stamp[seed] = g;
    {
        const int nin = count_inside(m, s, seed);
        uint8_t t;
        if (nin == (int)(m.cpOff[seed + 1] - m.cpOff[seed]))
            t = ALL_INSIDE;
        else if (phi01(s, m.C[seed]))
            t = CENTER_INSIDE;
        else
            t = CENTER_OUTSIDE;
        w.type[seed] = t;
        out.lists[t - 1].push_back(seed);
    }
    for (size_t h = 0; h < queue.size(); ++h) {
        const int64_t icur = queue[h];
        for (int64_t k = m.ccOff[icur]; k < m.ccOff[icur + 1]; ++k) {
            const int32_t inb = m.cc[k];
            if (w.stamp[inb] == g) continue;
            w.stamp[inb] = g;
            const int nin = count_inside(m, s, inb);
            if (nin == 0) {
                w.type[inb] = ALL_OUTSIDE;
                continue;
            }
            queue.push_back(inb);
            uint8_t t;
            if (nin == (int)(m.cpOff[icur + 1] - m.cpOff[icur]))
                t = ALL_INSIDE;
            else if (phi01(s, m.C[inb]))
                t = CENTER_INSIDE;
            else
                t = CENTER_OUTSIDE;
            w.type[inb] = t;
            out.lists[t - 1].push_back(inb);
        }
    }
    for (auto& l : out.lists) std::sort(l.begin(), l.end());
}

struct Fields {
    double *As, *Fs, *Ts, *Ct;
};

static void solid_fluid_interact(const Mesh& m, bool twoD, double rho_f, double dt, const orc_solid& s,
                                 int64_t sid, const double* U, Fields F, double* ft, int mode, Workspace& w,
                                 SolidResult& r, std::vector<double>* alpha_out) {
    if (mode == 1)
        enumerate_faithful(m, s, r);
    else
        enumerate_fast(m, s, w, r);

    // Ct marks (solidcloud.cpp:376-382): last solid in index order wins
    const double insideType = double(sid + 4);
    for (int64_t c : r.lists[0]) F.Ct[c] = insideType;
    for (int64_t c : r.lists[1]) F.Ct[c] = CENTER_INSIDE;
    for (int64_t c : r.lists[2]) F.Ct[c] = CENTER_OUTSIDE;

    const V3 ctr{s.center[0], s.center[1], s.center[2]};
    const V3 vel{s.vel[0], s.vel[1], s.vel[2]};
    const V3 om{s.omega[0], s.omega[1], s.omega[2]};
    const double dtINV = 1.0 / dt;
    V3 force{0, 0, 0}, torque{0, 0, 0};

    // vertex phi cache (geometrictools.cpp:6-11), cleared per solid
    std::unordered_map<int, double> cache;
    if (mode != 1) {
        if (w.vstamp.size() != (size_t)m.nP) {
            w.vstamp.assign(m.nP, 0);
            w.vphi.assign(m.nP, 0.0);
        }
    }
    const uint32_t g = w.gen;
    auto phiOf = [&](int32_t v) -> double {
        if (mode == 1) {
            auto it = cache.find(v);
            if (it == cache.end()) it = cache.emplace(v, phi(s, m.P[v])).first;
            return it->second;
        }
        if (w.vstamp[v] != g) {
            w.vstamp[v] = g;
            w.vphi[v] = phi(s, m.P[v]);
        }
        return w.vphi[v];
    };

    const int64_t nIn = (int64_t)r.lists[0].size();
    int64_t counter = 0;
    for (int t = 0; t < 3; ++t) {
        for (int64_t c : r.lists[t]) {
            double alpha = nIn > 0 ? 1.0 : 0.0;
            if (counter >= nIn) alpha = cell_volume(m, c, twoD, phiOf) / m.Vol[c];
            if (alpha_out) alpha_out->push_back(alpha);
            F.As[c] += alpha;
            const V3 cc = m.C[c];
            const V3 uf{U[3 * c + 0], U[3 * c + 1], U[3 * c + 2]};
            const V3 us = vel + cross(om, cc - ctr);
            const V3 f = alpha * (uf - us);
            const V3 tq = cross(cc - ctr, f);
            force = force + (f * m.Vol[c]) * dtINV;
            torque = torque + (tq * m.Vol[c]) * dtINV;
            const V3 fs = f * dtINV;
            F.Fs[3 * c + 0] += fs.x;
            F.Fs[3 * c + 1] += fs.y;
            F.Fs[3 * c + 2] += fs.z;
            F.Ts[c] += alpha;
            ++counter;
        }
    }
    force = force * rho_f;
    torque = torque * rho_f;
    ft[0] = force.x;
    ft[1] = force.y;
    ft[2] = force.z;
    ft[3] = torque.x;
    ft[4] = torque.y;
    ft[5] = torque.z;
}

}  // namespace orc

using namespace orc;

struct orc_mesh {
    Mesh m;
};

extern "C" {

orc_mesh* orc_mesh_create(int64_t n_points, const double* points, int64_t n_faces, const int64_t* face_offsets,
                          const int32_t* face_points, const int32_t* owner, int64_t n_internal,
                          const int32_t* neighbour) {
    if (n_points <= 0 || n_faces <= 0 || !points || !face_offsets || !face_points || !owner) return nullptr;
    if (n_internal < 0 || n_internal > n_faces || (n_internal > 0 && !neighbour)) return nullptr;
    orc_mesh* om = new orc_mesh;
    Mesh& m = om->m;
    m.nP = n_points;
    m.nF = n_faces;
    m.nIF = n_internal;
    m.P.resize(n_points);
    for (int64_t i = 0; i < n_points; ++i) m.P[i] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};
    m.fOff.assign(face_offsets, face_offsets + n_faces + 1);
    m.fPts.assign(face_points, face_points + m.fOff[n_faces]);
    for (int32_t v : m.fPts)
        if (v < 0 || v >= n_points) {
            delete om;
            return nullptr;
        }
    m.own.assign(owner, owner + n_faces);
    m.nei.assign(neighbour, neighbour + n_internal);
    build_mesh(m);
    return om;
}

// Single-block hex mesh (points i-fastest) exported the way blockMesh writes it:
// internal faces by owner then neighbour (+x, +y, +z), then the boundary
// patches xmin, xmax, ymin, ymax, zmin, zmax; face vertex order from the hex
// cell model.  points: [(nx+1)(ny+1)(nz+1) * 3].
orc_mesh* orc_mesh_create_structured(int64_t nx, int64_t ny, int64_t nz, const double* points) {
    if (nx < 1 || ny < 1 || nz < 1 || !points) return nullptr;
    const int64_t nC = nx * ny * nz, nP = (nx + 1) * (ny + 1) * (nz + 1);
    auto P = [&](int64_t i, int64_t j, int64_t k) { return (int32_t)(i + (nx + 1) * (j + (ny + 1) * k)); };
    // hex model vertex offsets and faces (xmin xmax ymin ymax zmin zmax)
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    static const int mf[6][4] = {{0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4}, {3, 7, 6, 2}, {0, 3, 2, 1}, {4, 5, 6, 7}};
    const int64_t nIF = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1);
    const int64_t nBF = 2 * (ny * nz + nx * nz + nx * ny);
    const int64_t nF = nIF + nBF;
    orc_mesh* om = new orc_mesh;
    Mesh& m = om->m;
    m.nP = nP;
    m.nF = nF;
    m.nIF = nIF;
    m.P.resize(nP);
    for (int64_t q = 0; q < nP; ++q) m.P[q] = {points[3 * q], points[3 * q + 1], points[3 * q + 2]};
    m.fOff.resize(nF + 1);
    for (int64_t f = 0; f <= nF; ++f) m.fOff[f] = 4 * f;
    m.fPts.resize(4 * nF);
    m.own.resize(nF);
    m.nei.resize(nIF);
    int64_t f = 0;
    auto put = [&](int64_t c, int64_t i, int64_t j, int64_t k, int face) {
        for (int t = 0; t < 4; ++t) {
            const int* o = off[mf[face][t]];
            m.fPts[4 * f + t] = P(i + o[0], j + o[1], k + o[2]);
        }
        m.own[f] = (int32_t)c;
    };
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                const int64_t c = i + nx * (j + ny * k);
                if (i < nx - 1) { put(c, i, j, k, 1); m.nei[f] = (int32_t)(c + 1); ++f; }
                if (j < ny - 1) { put(c, i, j, k, 3); m.nei[f] = (int32_t)(c + nx); ++f; }
                if (k < nz - 1) { put(c, i, j, k, 5); m.nei[f] = (int32_t)(c + nx * ny); ++f; }
            }
    for (int face = 0; face < 6; ++face)
        for (int64_t k = 0; k < nz; ++k)
            for (int64_t j = 0; j < ny; ++j)
                for (int64_t i = 0; i < nx; ++i) {
                    const bool on = (face == 0 && i == 0) || (face == 1 && i == nx - 1) || (face == 2 && j == 0) ||
                                    (face == 3 && j == ny - 1) || (face == 4 && k == 0) || (face == 5 && k == nz - 1);
                    if (!on) continue;
                    put(i + nx * (j + ny * k), i, j, k, face);
                    ++f;
                }
    (void)nC;
    build_mesh(m);
    return om;
}

void orc_mesh_destroy(orc_mesh* m) { delete m; }

int64_t orc_mesh_ncells(const orc_mesh* m) { return m ? m->m.nC : 0; }

void orc_mesh_geometry(const orc_mesh* om, double* cc, double* V, double* fc, double* Sf) {
    const Mesh& m = om->m;
    for (int64_t c = 0; c < m.nC; ++c) {
        if (cc) {
            cc[3 * c] = m.C[c].x;
            cc[3 * c + 1] = m.C[c].y;
            cc[3 * c + 2] = m.C[c].z;
        }
        if (V) V[c] = m.Vol[c];
    }
    for (int64_t f = 0; f < m.nF; ++f) {
        if (fc) {
            fc[3 * f] = m.Cf[f].x;
            fc[3 * f + 1] = m.Cf[f].y;
            fc[3 * f + 2] = m.Cf[f].z;
        }
        if (Sf) {
            Sf[3 * f] = m.Sf[f].x;
            Sf[3 * f + 1] = m.Sf[f].y;
            Sf[3 * f + 2] = m.Sf[f].z;
        }
    }
}

int64_t orc_mesh_cell_points(const orc_mesh* om, int64_t* offsets, int32_t* labels) {
    const Mesh& m = om->m;
    if (offsets) std::copy(m.cpOff.begin(), m.cpOff.end(), offsets);
    if (labels) std::copy(m.cp.begin(), m.cp.end(), labels);
    return (int64_t)m.cp.size();
}

void orc_set_csg(const orc_csg_shape* shapes, int64_t n) {
    g_csg.assign(shapes, shapes + (n > 0 ? n : 0));
}

int orc_is_inside(const orc_solid* s, const double* p) { return phi01(*s, V3{p[0], p[1], p[2]}) ? 1 : 0; }

double orc_signed_distance(const orc_solid* s, const double* p) { return phi(*s, V3{p[0], p[1], p[2]}); }

void orc_world_to_local(const orc_solid* s, const double* p, double* out) {
    const V3 l = world_to_local(s->q, s->center, V3{p[0], p[1], p[2]});
    out[0] = l.x;
    out[1] = l.y;
    out[2] = l.z;
}

double orc_cell_fraction(const orc_mesh* om, const orc_solid* s, int64_t cell, int two_d) {
    const Mesh& m = om->m;
    if (cell < 0 || cell >= m.nC) return NAN;
    std::unordered_map<int, double> cache;
    auto phiOf = [&](int32_t v) -> double {
        auto it = cache.find(v);
        if (it == cache.end()) it = cache.emplace(v, phi(*s, m.P[v])).first;
        return it->second;
    };
    return cell_volume(m, cell, two_d != 0, phiOf) / m.Vol[cell];
}

int64_t orc_find_nearest_cell(const orc_mesh* om, const double* p) {
    return find_nearest_cell(om->m, V3{p[0], p[1], p[2]});
}

int orc_interact_range(const orc_mesh* om, int two_d, double rho_f, double dt, int64_t ns,
                       const orc_solid* solids, int64_t s0, int64_t s1, const double* U, double* As, double* Fs,
                       double* Ts, double* Ct, double* ft, int mode) {
    if (!om || !solids || !U || !As || !Fs || !Ts || !Ct || !ft || s0 < 0 || s1 > ns || s0 > s1) return -1;
    const Mesh& m = om->m;
    Workspace w;
    SolidResult r;
    for (int64_t s = s0; s < s1; ++s)
        solid_fluid_interact(m, two_d != 0, rho_f, dt, solids[s], s, U, Fields{As, Fs, Ts, Ct}, ft + 6 * s, mode,
                             w, r, nullptr);
    return 0;
}

int orc_interact(const orc_mesh* om, int two_d, double rho_f, double dt, int64_t ns, const orc_solid* solids,
                 const double* U, double* As, double* Fs, double* Ts, double* Ct, double* ft, int mode,
                 orc_capture* cap) {
    if (!om || (!solids && ns > 0) || !U || !As || !Fs || !Ts || !Ct || (!ft && ns > 0)) return -1;
    const Mesh& m = om->m;
    // zero-fill (solidcloud.cpp:438-441)
    std::fill(As, As + m.nC, 0.0);
    std::fill(Fs, Fs + 3 * m.nC, 0.0);
    std::fill(Ts, Ts + m.nC, 0.0);
    std::fill(Ct, Ct + m.nC, 0.0);
    Workspace w;
    SolidResult r;
    std::vector<double> alphas;
    if (cap) cap->used = 0;
    int rc = 0;
    for (int64_t s = 0; s < ns; ++s) {
        alphas.clear();
        solid_fluid_interact(m, two_d != 0, rho_f, dt, solids[s], s, U, Fields{As, Fs, Ts, Ct}, ft + 6 * s, mode, w,
                             r, cap ? &alphas : nullptr);
        if (cap) {
            const int64_t n = (int64_t)(r.lists[0].size() + r.lists[1].size() + r.lists[2].size());
            if (cap->counts) {
                cap->counts[4 * s + 0] = (int64_t)r.lists[0].size();
                cap->counts[4 * s + 1] = (int64_t)r.lists[1].size();
                cap->counts[4 * s + 2] = (int64_t)r.lists[2].size();
                cap->counts[4 * s + 3] = r.seed;
            }
            if (cap->used + n > cap->capacity) {
                rc = -2;
            } else {
                int64_t k = cap->used;
                for (int t = 0; t < 3; ++t)
                    for (int64_t c : r.lists[t]) {
                        if (cap->cells) cap->cells[k] = c;
                        if (cap->alpha) cap->alpha[k] = alphas[k - cap->used];
                        ++k;
                    }
                cap->used = k;
            }
        }
    }
    // checkAlpha (solidcloud.cpp:564-570)
    for (int64_t c = 0; c < m.nC; ++c) As[c] = std::min(As[c], 1.0);
    return rc;
}

// U = U - Fs*dt (src/main.cpp:70) on the internal field.
void orc_update_velocity(int64_t n, double* U, const double* Fs, double dt) {
    for (int64_t c = 0; c < 3 * n; ++c) U[c] = U[c] - Fs[c] * dt;
}

// fixInternal (src/solidcloud.cpp:288-301): cells fully inside a solid (Ct >= 4)
// take that solid's rigid velocity v + omega x (cc - centre).
void orc_fix_internal(const orc_mesh* om, double* U, const double* Ct, int64_t ns, const orc_solid* solids) {
    const Mesh& m = om->m;
    for (int64_t c = 0; c < m.nC; ++c) {
        if (Ct[c] >= 4) {
            const int64_t id = (int64_t)Ct[c] - 4;
            if (id < 0 || id >= ns) continue;
            const orc_solid& s = solids[id];
            const V3 v = V3{s.vel[0], s.vel[1], s.vel[2]} +
                         cross(V3{s.omega[0], s.omega[1], s.omega[2]}, m.C[c] - V3{s.center[0], s.center[1], s.center[2]});
            U[3 * c] = v.x;
            U[3 * c + 1] = v.y;
            U[3 * c + 2] = v.z;
        }
    }
}

// ---------------------------------------------------------------------------
// rigid-body update
// ---------------------------------------------------------------------------
void orc_mass_properties(const orc_solid* sp, double* volume, double* moi3) {
    const orc_solid& s = *sp;
    const double PI = 3.14159265358979323846;
    double V = 0.0, m0 = 1.0, m1 = 1.0, m2 = 1.0;  // plane: m_volume 0, m_moi = I (ishape.h:33-36)
    switch (s.shape) {
    case ORC_SPHERE: {  // sphere.h:21-25
        const double r = s.prm[0], r2 = r * r;
        V = 4.0 / 3.0 * PI * r2 * r;
        m0 = m1 = m2 = 0.4 * V * r2;
        break;
    }
    case ORC_ELLIPSOID: {  // ellipsoid.h:25-30
        const double a = s.prm[0], b = s.prm[1], c = s.prm[2];
        V = 4.0 / 3.0 * PI * a * b * c;
        m0 = 0.2 * V * (b * b + c * c);
        m1 = 0.2 * V * (a * a + c * c);
        m2 = 0.2 * V * (a * a + b * b);
        break;
    }
    case ORC_CIRCLE:
    case ORC_CIRCLE_TAIL:
    case ORC_CIRCLE_TWOTAIL: {  // circle.h:22-25, circle_tail.h:28-31
        const double r = s.prm[0], r2 = r * r;
        V = PI * r2;
        m0 = m1 = m2 = 0.5 * V * r2;
        break;
    }
    case ORC_ELLIPSE: {  // ellipse.h:24-28
        const double a = s.prm[0], b = s.prm[1];
        V = PI * a * b;
        m0 = m1 = m2 = 0.25 * V * (a * a + b * b);
        break;
    }
    case ORC_BOX: {  // box.h:22-27
        const double a = s.prm[0], b = s.prm[1], c = s.prm[2];
        V = 8.0 * a * b * c;
        m0 = 1.0 / 3.0 * V * (b * b + c * c);
        m1 = 1.0 / 3.0 * V * (a * a + c * c);
        m2 = 1.0 / 3.0 * V * (b * b + a * a);
        break;
    }
    case ORC_RECTANGLE: {  // rectangle.h:25-29
        const double a = s.prm[0], b = s.prm[1];
        V = 4.0 * a * b;
        m0 = m1 = m2 = 1.0 / 3.0 * V * (a * a + b * b);
        break;
    }
    case ORC_CSG: {  // the values the CSG shape record declares (IShape::m_volume, m_moi)
        const orc_csg_shape* G = csg_of(s);
        if (G) {
            V = G->volume;
            m0 = G->moi[0];
            m1 = G->moi[1];
            m2 = G->moi[2];
        }
        break;
    }
    default: break;
    }
    *volume = V;
    moi3[0] = m0;
    moi3[1] = m1;
    moi3[2] = m2;
}

namespace {
struct Quat {
    double w;
    V3 v;
};
// Hamilton product, OpenFOAM's quaternion operator*: (w1 w2 - v1.v2, w1 v2 + w2 v1 + v1 x v2)
Quat qmul(const Quat& a, const Quat& b) {
    return {a.w * b.w - dot(a.v, b.v), (a.w * b.v + b.w * a.v) + cross(a.v, b.v)};
}
// quaternion::R(): rotation tensor, row-major
void qR(const Quat& q, double R[9]) {
    const double w2 = q.w * q.w, x2 = q.v.x * q.v.x, y2 = q.v.y * q.v.y, z2 = q.v.z * q.v.z;
    const double txy = 2 * q.v.x * q.v.y, twz = 2 * q.w * q.v.z, txz = 2 * q.v.x * q.v.z;
    const double twy = 2 * q.w * q.v.y, tyz = 2 * q.v.y * q.v.z, twx = 2 * q.w * q.v.x;
    R[0] = w2 + x2 - y2 - z2;
    R[1] = txy - twz;
    R[2] = txz + twy;
    R[3] = txy + twz;
    R[4] = w2 - x2 + y2 - z2;
    R[5] = tyz - twx;
    R[6] = txz - twy;
    R[7] = tyz + twx;
    R[8] = w2 - x2 - y2 + z2;
}
void tmul(const double* A, const double* B, double* C) {  // C = A & B
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
}  // namespace

}  // extern "C"

namespace {
using namespace orc;

struct Body {
    double mass, mass_inv, Minv[9];
};

// Solid::setMaterial (solid.h:94-103): mass = V rho, mass_inv = (1/V)/rho, moi_inv = inv(moi)/rho
Body body_of(const orc_solid& s, const orc_dyn& d) {
    double vol, moi[3];
    orc_mass_properties(&s, &vol, moi);
    const double volinv = vol > 0 ? 1.0 / vol : 0.0;
    Body b;
    b.mass = vol * d.rho;
    b.mass_inv = volinv / d.rho;
    for (int a = 0; a < 9; ++a) b.Minv[a] = 0.0;
    for (int a = 0; a < 3; ++a) b.Minv[4 * a] = (1.0 / moi[a]) / d.rho;
    return b;
}

// One sub-step of one solid: clear, AB2 fluid force (first step: old = current),
// buoyant gravity, the contact force fc (if any), Solid::move.
void substep(orc_solid& s, orc_dyn& d, const Body& b, V3 ff, V3 ft, V3 g, double rho_f, const V3* fc, double dts) {
    if (d.first_fluid) {
        for (int a = 0; a < 3; ++a) {
            d.ff_old[a] = (&ff.x)[a];
            d.ft_old[a] = (&ft.x)[a];
        }
        d.first_fluid = 0;
    }
    const V3 ffo{d.ff_old[0], d.ff_old[1], d.ff_old[2]}, fto{d.ft_old[0], d.ft_old[1], d.ft_old[2]};
    V3 force = V3{0, 0, 0} + (1.5 * ff - 0.5 * ffo);
    const V3 torque = V3{0, 0, 0} + (1.5 * ft - 0.5 * fto);
    const V3 gprime = ((d.rho - rho_f) / d.rho) * g;
    force = force + b.mass * gprime;
    if (fc) force = force + *fc;
    // Solid::move
    V3 v{s.vel[0], s.vel[1], s.vel[2]}, w{s.omega[0], s.omega[1], s.omega[2]};
    const V3 v_old = v, w_old = w;
    v = v + (force * b.mass_inv) * dts;
    Quat q{s.q[0], {s.q[1], s.q[2], s.q[3]}};
    double R[9], RT[9], T1[9], Mw[9];
    qR(q, R);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) RT[3 * i + j] = R[3 * j + i];
    tmul(R, b.Minv, T1);
    tmul(T1, RT, Mw);
    const V3 aw{Mw[0] * torque.x + Mw[1] * torque.y + Mw[2] * torque.z,
                Mw[3] * torque.x + Mw[4] * torque.y + Mw[5] * torque.z,
                Mw[6] * torque.x + Mw[7] * torque.y + Mw[8] * torque.z};
    w = w + aw * dts;
    if (d.mask >= 0) {  // Motion01Mask::constraint: cmptMultiply with 0/1 masks
        v = {v.x * ((d.mask >> 0) & 1), v.y * ((d.mask >> 1) & 1), v.z * ((d.mask >> 2) & 1)};
        w = {w.x * ((d.mask >> 3) & 1), w.y * ((d.mask >> 4) & 1), w.z * ((d.mask >> 5) & 1)};
    }
    const V3 c = V3{s.center[0], s.center[1], s.center[2]} + (0.5 * (v + v_old)) * dts;
    const Quat hw{0.5 * 0.0, 0.5 * (0.5 * (w + w_old))};  // 0.5 * quaternion(0.5 (w + w_old))
    const Quat dq = qmul(hw, q);
    q = {q.w + dq.w * dts, q.v + dq.v * dts};
    const double qm = std::sqrt(q.w * q.w + magSqr(q.v));
    q = {q.w / qm, q.v / qm};
    for (int a = 0; a < 3; ++a) {
        s.vel[a] = (&v.x)[a];
        s.omega[a] = (&w.x)[a];
        s.center[a] = (&c.x)[a];
    }
    s.q[0] = q.w;
    s.q[1] = q.v.x;
    s.q[2] = q.v.y;
    s.q[3] = q.v.z;
}

// storeOldForce (solid.h:143-149)
void store_old(orc_dyn& d, V3 ff, V3 ft) {
    for (int a = 0; a < 3; ++a) {
        d.ff_old[a] = (&ff.x)[a];
        d.ft_old[a] = (&ft.x)[a];
    }
}

V3 cvec(const double* p) { return {p[0], p[1], p[2]}; }

// Contact force on a from finite solid b (sphereSphereCollision geometry,
// collision.cpp:7-17: normal from a to b, contact point 0.5 (ca + cb + (ra - rb) n)).
V3 pair_force(const orc_solid& a, double ra, const orc_solid& b, double rb, double k, double cdamp) {
    const V3 ca = cvec(a.center), cb = cvec(b.center);
    const V3 ab = cb - ca;
    const double dist = mag(ab);
    const double delta = ra + rb - dist;
    const V3 n = ab / (dist + 1e-300);  // Foam::normalised
    const V3 cp = 0.5 * ((ca + cb) + (ra - rb) * n);
    const V3 va = cvec(a.vel) + cross(cvec(a.omega), cp - ca);
    const V3 vb = cvec(b.vel) + cross(cvec(b.omega), cp - cb);
    const double vn = dot(vb - va, n);
    const double fm = k * delta - cdamp * vn;
    if (!(delta > 0) || !(fm > 0)) return {0, 0, 0};
    return (-fm) * n;
}

// Contact force on finite solid s from plane p (planeSphereCollision geometry,
// collision.cpp:19-28: normal = the plane's local +y axis, overlap r - local y
// of the centre, contact point centre - r n).
V3 plane_force(const orc_solid& s, double r, const orc_solid& p, double k, double cdamp) {
    double R[9];
    qR(Quat{p.q[0], {p.q[1], p.q[2], p.q[3]}}, R);
    const V3 n{R[1], R[4], R[7]};
    const V3 cs = cvec(s.center), cpl = cvec(p.center);
    const double delta = r - dot(n, cs - cpl);
    const V3 cp = cs - r * n;
    const V3 vs = cvec(s.vel) + cross(cvec(s.omega), cp - cs);
    const V3 vp = cvec(p.vel) + cross(cvec(p.omega), cp - cpl);
    const double vn = dot(vs - vp, n);
    const double fm = k * delta - cdamp * vn;
    if (!(delta > 0) || !(fm > 0)) return {0, 0, 0};
    return fm * n;
}

}  // namespace

extern "C" {

void orc_evolve(int64_t ns, orc_solid* solids, orc_dyn* dyn, const double* fluid_ft, double time, double dt,
                const double* gravity, double rho_f) {
    (void)time;
    const int n_sub = ns == 1 ? 1 : 20;  // solidcloud.cpp:524-526
    const double dts = dt / n_sub;
    const V3 g{gravity[0], gravity[1], gravity[2]};
    for (int64_t id = 0; id < ns; ++id) {
        const Body b = body_of(solids[id], dyn[id]);
        const V3 ff = cvec(fluid_ft + 6 * id), ft = cvec(fluid_ft + 6 * id + 3);
        for (int it = 0; it < n_sub; ++it) substep(solids[id], dyn[id], b, ff, ft, g, rho_f, nullptr, dts);
        store_old(dyn[id], ff, ft);
    }
}

double orc_bounding_radius(const orc_solid* s) {
    switch (s->shape) {
    case 0:  // Sphere, sphere.h:28
    case 2:  // Circle, circle.h:27
        return s->prm[0];
    case 1:  // Ellipsoid, ellipsoid.h:32
    case 4:  // Box, box.h:29
        return std::max(std::max(s->prm[0], s->prm[1]), s->prm[2]);
    case 3:  // Ellipse, ellipse.h:30
    case 5:  // Rectangle, rectangle.h:31
        return std::max(s->prm[0], s->prm[1]);
    case 7:  // Circle_Tail, circle_tail.h:28,33: m_radiusB = 2 m_radiusa, m_radiusa = (ratio + 1) r / 2
    case 8:  // Circle_TwoTail, circle_twotail.h:26,31
        return 2 * ((s->prm[1] + 1) * 0.5 * s->prm[0]);
    case ORC_CSG: {  // declared by the CSG shape record (IShape::m_radiusB)
        const orc_csg_shape* G = csg_of(*s);
        return G ? G->radius_b : 0.0;
    }
    default:
        return 0.0;
    }
}

void orc_contact_forces(int64_t ns, const orc_solid* solids, double k, double cdamp, double* out) {
    std::vector<double> rb(ns);
    for (int64_t s = 0; s < ns; ++s) rb[s] = orc_bounding_radius(&solids[s]);
    for (int64_t s = 0; s < ns; ++s) {
        V3 fc{0, 0, 0};
        if (rb[s] > 0) {
            for (int64_t j = 0; j < ns; ++j) {  // every finite partner, ascending
                if (j == s || !(rb[j] > 0)) continue;
                const V3 ab = cvec(solids[j].center) - cvec(solids[s].center);
                if (!(rb[s] + rb[j] - mag(ab) > 0)) continue;
                fc = fc + pair_force(solids[s], rb[s], solids[j], rb[j], k, cdamp);
            }
            for (int64_t p = 0; p < ns; ++p)  // then the planes, ascending
                if (solids[p].shape == 6) fc = fc + plane_force(solids[s], rb[s], solids[p], k, cdamp);
        }
        out[3 * s] = fc.x;
        out[3 * s + 1] = fc.y;
        out[3 * s + 2] = fc.z;
    }
}

void orc_evolve_contact(int64_t ns, orc_solid* solids, orc_dyn* dyn, const double* fluid_ft, double time, double dt,
                        const double* gravity, double rho_f, double k, double cdamp) {
    (void)time;
    const int n_sub = ns == 1 ? 1 : 20;
    const double dts = dt / n_sub;
    const V3 g{gravity[0], gravity[1], gravity[2]};
    std::vector<double> fc(3 * ns);
    std::vector<Body> body(ns);
    for (int64_t id = 0; id < ns; ++id) body[id] = body_of(solids[id], dyn[id]);
    for (int it = 0; it < n_sub; ++it) {
        // contact forces from the state at the start of the sub-step, then every solid moves
        orc_contact_forces(ns, solids, k, cdamp, fc.data());
        for (int64_t id = 0; id < ns; ++id) {
            const V3 f = cvec(&fc[3 * id]);
            substep(solids[id], dyn[id], body[id], cvec(fluid_ft + 6 * id), cvec(fluid_ft + 6 * id + 3), g, rho_f,
                    orc_bounding_radius(&solids[id]) > 0 ? &f : nullptr, dts);
        }
    }
    for (int64_t id = 0; id < ns; ++id) store_old(dyn[id], cvec(fluid_ft + 6 * id), cvec(fluid_ft + 6 * id + 3));
}

}  // extern "C"
