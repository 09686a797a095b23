// fraction.h -- solid volume fraction of one hex cell from its corner signed
// distances (the reference's GeometricTools, src/geometrictools.cpp:13-116).
//
// Cell vertices are indexed in ascending point-label order for an i-fastest
// numbered block (offsets (0,0,0) (1,0,0) (0,1,0) (1,1,0) (0,0,1) (1,0,1)
// (0,1,1) (1,1,1)); that is OpenFOAM's cellPoints order.  Face vertex lists
// follow OpenFOAM's hex model (owner-outward winding): an internal face takes
// the winding and starting vertex of its owner, the lower-index cell, so a
// cell's -x/-y/-z faces are listed differently when internal and when on the
// boundary.  The starting vertex matters: it is vertex A of the apex rule.
//
// Arithmetic is in R (float or double) on coordinates relative to the cell's
// vertex 0, which keeps FP32 results accurate for small cells far from the
// origin.
#pragma once

#include "shapes.h"

namespace sdfibm_dev {

template <class R>
struct r3 {
    R x, y, z;
};
template <class R>
__device__ inline r3<R> sub(r3<R> a, r3<R> b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class R>
__device__ inline R rdot(r3<R> a, r3<R> b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
__device__ inline r3<R> rcross(r3<R> a, r3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R>
__device__ inline R rmag(r3<R> a) {
    return sqrt(rdot(a, a));
}

// Quotient of the fraction arithmetic: exact division in FP64; in FP32 the
// hardware reciprocal (1 ulp) times the numerator -- two operations instead of
// the ten of a correctly rounded divide.  Every divisor here is positive.
template <class R>
__device__ inline R qdiv(R a, R b) {
    return a / b;
}
template <>
__device__ inline float qdiv<float>(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
}

// Line fraction (geometrictools.cpp:13-23).
template <class R>
__device__ inline R line_fraction(R pa, R pb) {
    if (pa > 0 && pb > 0) return R(0);
    if (pa <= 0 && pb <= 0) return R(1);
    if (pa > 0) return qdiv<R>(-pb, pa - pb);
    return qdiv<R>(-pa, pb - pa);
}

template <class R>
__device__ inline r3<R> apex_point(r3<R> A, R pA, r3<R> B, R pB) {
    const R s = qdiv<R>(fabs(pA), R(kSMALL) + fabs(pA) + fabs(pB));
    return {A.x - s * (A.x - B.x), A.y - s * (A.y - B.y), A.z - s * (A.z - B.z)};
}

// The 8 vertices of a hex cell (relative to vertex 0) and their signed
// distances, in cellPoints order: offsets (0,0,0) (1,0,0) (0,1,0) (1,1,0)
// (0,0,1) (1,0,1) (0,1,1) (1,1,1).  Named members with compile-time access keep
// them in registers (an array would be selected through memory).
template <class R>
struct Hex8 {
    r3<R> p0, p1, p2, p3, p4, p5, p6, p7;
    R f0, f1, f2, f3, f4, f5, f6, f7;
};
template <int I, class R>
__device__ inline r3<R> hp(const Hex8<R>& h) {
    if constexpr (I == 0) return h.p0;
    else if constexpr (I == 1) return h.p1;
    else if constexpr (I == 2) return h.p2;
    else if constexpr (I == 3) return h.p3;
    else if constexpr (I == 4) return h.p4;
    else if constexpr (I == 5) return h.p5;
    else if constexpr (I == 6) return h.p6;
    else return h.p7;
}
template <int I, class R>
__device__ inline R hf(const Hex8<R>& h) {
    if constexpr (I == 0) return h.f0;
    else if constexpr (I == 1) return h.f1;
    else if constexpr (I == 2) return h.f2;
    else if constexpr (I == 3) return h.f3;
    else if constexpr (I == 4) return h.f4;
    else if constexpr (I == 5) return h.f5;
    else if constexpr (I == 6) return h.f6;
    else return h.f7;
}

// Area fraction of the quad face (V0 V1 V2 V3) (geometrictools.cpp:74-116).
// The apex rule (geometrictools.cpp:25-45): A = first listed vertex, B = first
// later vertex with phi_A*phi_B <= 0, else the last one.
template <class R, int V0, int V1, int V2, int V3>
__device__ inline R face_fraction(const Hex8<R>& h, R sf_mag, R l01, R l12, R l23, R l30) {
    // l01 .. l30: line fractions of the listed edges (V0 V1) (V1 V2) (V2 V3) (V3 V0)
    const R p0 = hf<V0>(h), p1 = hf<V1>(h), p2 = hf<V2>(h), p3 = hf<V3>(h);
    const int ss = (p0 > 0 ? 1 : -1) + (p1 > 0 ? 1 : -1) + (p2 > 0 ? 1 : -1) + (p3 > 0 ? 1 : -1);
    if (ss == 4) return R(0);
    if (ss == -4) return R(1);
    const r3<R> q0 = hp<V0>(h), q1 = hp<V1>(h), q2 = hp<V2>(h), q3 = hp<V3>(h);
    // later assignments win: the first vertex in list order with a sign change
    r3<R> B = q3;
    R pB = p3;
    if (p0 * p2 <= 0) {
        B = q2;
        pB = p2;
    }
    if (p0 * p1 <= 0) {
        B = q1;
        pB = p1;
    }
    const r3<R> ap = apex_point(q0, p0, B, pB);
    auto tri = [&](r3<R> O, r3<R> A, R lf) { return fabs(R(0.5) * rmag(rcross(sub(A, O), sub(ap, O)))) * lf; };
    const R area = tri(q0, q1, l01) + tri(q1, q2, l12) + tri(q2, q3, l23) + tri(q3, q0, l30);
    return area / sf_mag;
}

// Solid volume fraction of a hex cell (geometrictools.cpp:47-72): sum over faces
// of (1/3) eps_f |(apex - fc).Sf|, divided by V.  two_d: the apex z is forced to
// apex_z_rel (world z = 0, geometrictools.cpp:57-58).  face(slot, fc, sf) gives
// the centre (relative to vertex 0) and area vector of face slot in the order
// -x, +x, -y, +y, -z, +z; bnd[a]: the minus face along axis a is on the boundary
// (its vertex list then starts/winds differently: an internal face takes its
// owner's -- the lower-index neighbour's -- listing).  Each face is fetched,
// used and dropped in turn, so no more than one face's geometry is live.
template <class R, class FaceAt>
__device__ inline R hex_fraction(const Hex8<R>& h, FaceAt face, const bool* bnd, R V, bool two_d, R apex_z_rel) {
    // cell apex: A = vertex 0, B = first later vertex with a sign change, else vertex 7
    // B is blended as sum_t w_t P_t with exactly one weight 1 and the rest 0, which is
    // exact and keeps the vertices in registers (a select chain is lowered to an
    // indexed load from memory).
    const R f0 = h.f0;
    const bool c1 = f0 * h.f1 <= 0, c2 = f0 * h.f2 <= 0, c3 = f0 * h.f3 <= 0, c4 = f0 * h.f4 <= 0,
               c5 = f0 * h.f5 <= 0, c6 = f0 * h.f6 <= 0;
    const R w1 = c1 ? R(1) : R(0);
    const R w2 = (!c1 && c2) ? R(1) : R(0);
    const R w3 = (!c1 && !c2 && c3) ? R(1) : R(0);
    const R w4 = (!c1 && !c2 && !c3 && c4) ? R(1) : R(0);
    const R w5 = (!c1 && !c2 && !c3 && !c4 && c5) ? R(1) : R(0);
    const R w6 = (!c1 && !c2 && !c3 && !c4 && !c5 && c6) ? R(1) : R(0);
    const R w7 = R(1) - (w1 + w2 + w3 + w4 + w5 + w6);
    auto blend = [&](R a1, R a2, R a3, R a4, R a5, R a6, R a7) {
        return w1 * a1 + w2 * a2 + w3 * a3 + w4 * a4 + w5 * a5 + w6 * a6 + w7 * a7;
    };
    const r3<R> B{blend(h.p1.x, h.p2.x, h.p3.x, h.p4.x, h.p5.x, h.p6.x, h.p7.x),
                  blend(h.p1.y, h.p2.y, h.p3.y, h.p4.y, h.p5.y, h.p6.y, h.p7.y),
                  blend(h.p1.z, h.p2.z, h.p3.z, h.p4.z, h.p5.z, h.p6.z, h.p7.z)};
    const R pB = blend(h.f1, h.f2, h.f3, h.f4, h.f5, h.f6, h.f7);
    r3<R> ap = apex_point(h.p0, f0, B, pB);
    if (two_d) ap.z = apex_z_rel;
    // the 12 edges' line fractions, each formed once (calcLineFraction is symmetric in its
    // two arguments, bit for bit)
    const R x01 = line_fraction<R>(h.f0, h.f1), x23 = line_fraction<R>(h.f2, h.f3),
            x45 = line_fraction<R>(h.f4, h.f5), x67 = line_fraction<R>(h.f6, h.f7);
    const R y02 = line_fraction<R>(h.f0, h.f2), y13 = line_fraction<R>(h.f1, h.f3),
            y46 = line_fraction<R>(h.f4, h.f6), y57 = line_fraction<R>(h.f5, h.f7);
    const R z04 = line_fraction<R>(h.f0, h.f4), z15 = line_fraction<R>(h.f1, h.f5),
            z26 = line_fraction<R>(h.f2, h.f6), z37 = line_fraction<R>(h.f3, h.f7);
    R vol = R(0);
    // one face: its fraction eps (sf_mag = |Sf|) and, where eps != 0, its pyramid term
    auto add = [&](const r3<R>& fc, const r3<R>& sf, R eps) {
        if (eps != R(0)) vol += (R(1) / R(3)) * eps * fabs(rdot(sub(ap, fc), sf));
    };
    r3<R> fc, sf;
    face(0, fc, sf);
    add(fc, sf, bnd[0] ? face_fraction<R, 0, 4, 6, 2>(h, rmag(sf), z04, y46, z26, y02)
                       : face_fraction<R, 0, 2, 6, 4>(h, rmag(sf), y02, z26, y46, z04));
    face(1, fc, sf);
    add(fc, sf, face_fraction<R, 1, 3, 7, 5>(h, rmag(sf), y13, z37, y57, z15));
    face(2, fc, sf);
    add(fc, sf, bnd[1] ? face_fraction<R, 0, 1, 5, 4>(h, rmag(sf), x01, z15, x45, z04)
                       : face_fraction<R, 0, 4, 5, 1>(h, rmag(sf), z04, x45, z15, x01));
    face(3, fc, sf);
    add(fc, sf, face_fraction<R, 2, 6, 7, 3>(h, rmag(sf), z26, x67, z37, x23));
    face(4, fc, sf);
    add(fc, sf, bnd[2] ? face_fraction<R, 0, 2, 3, 1>(h, rmag(sf), y02, x23, y13, x01)
                       : face_fraction<R, 0, 1, 3, 2>(h, rmag(sf), x01, y13, x23, y02));
    face(5, fc, sf);
    add(fc, sf, face_fraction<R, 4, 5, 7, 6>(h, rmag(sf), x45, y57, x67, y46));
    return vol / V;
}

// ---------------------------------------------------------------------------
// Axis-aligned box cells (rectilinear meshes): hex_fraction in coordinates
// normalised by the cell size.  Every face is a unit square whose apex lies in
// its plane, so the triangle (O, A, apex) on an edge has area dist/2, dist = the
// apex's distance from the edge's line, and |(apex - fc).Sf| / V is the cell
// apex's distance from the face's plane.  Same apex rules, face listings and
// sums as above; no square roots, and each of the 12 edge line fractions is
// formed once (calcLineFraction is symmetric in its two arguments).
// ---------------------------------------------------------------------------
template <class R>
struct Phi8 {
    R f0, f1, f2, f3, f4, f5, f6, f7;
};
template <int I, class R>
__device__ inline R pf(const Phi8<R>& h) {
    if constexpr (I == 0) return h.f0;
    else if constexpr (I == 1) return h.f1;
    else if constexpr (I == 2) return h.f2;
    else if constexpr (I == 3) return h.f3;
    else if constexpr (I == 4) return h.f4;
    else if constexpr (I == 5) return h.f5;
    else if constexpr (I == 6) return h.f6;
    else return h.f7;
}

// Area fraction of the unit face listed (V0 V1 V2 V3); AU, AV are its in-plane
// axes (0 x, 1 y, 2 z; vertex bit a of the cellPoints index is the coordinate
// along axis a); lu0, lu1, lv0, lv1 are the line fractions of its edges on the
// lines u = 0, u = 1, v = 0, v = 1.
template <class R, int V0, int V1, int V2, int V3, int AU, int AV>
__device__ inline R unit_face_fraction(const Phi8<R>& h, R lu0, R lu1, R lv0, R lv1) {
    const R p0 = pf<V0>(h), p1 = pf<V1>(h), p2 = pf<V2>(h), p3 = pf<V3>(h);
    const int ss = (p0 > 0 ? 1 : -1) + (p1 > 0 ? 1 : -1) + (p2 > 0 ? 1 : -1) + (p3 > 0 ? 1 : -1);
    if (ss == 4) return R(0);
    if (ss == -4) return R(1);
    const bool c1 = p0 * p1 <= 0, c2 = p0 * p2 <= 0;
    const R pB = c1 ? p1 : (c2 ? p2 : p3);
    constexpr R u0 = R((V0 >> AU) & 1), v0 = R((V0 >> AV) & 1);
    const R bu = c1 ? R((V1 >> AU) & 1) : (c2 ? R((V2 >> AU) & 1) : R((V3 >> AU) & 1));
    const R bv = c1 ? R((V1 >> AV) & 1) : (c2 ? R((V2 >> AV) & 1) : R((V3 >> AV) & 1));
    const R s = qdiv<R>(fabs(p0), R(kSMALL) + fabs(p0) + fabs(pB));
    const R au = u0 - s * (u0 - bu), av = v0 - s * (v0 - bv);
    return R(0.5) * (lv0 * av + lu1 * (R(1) - au) + lv1 * (R(1) - av) + lu0 * au);
}

// Volume fraction of an axis-aligned box cell; bx, by, bz: the cell's -x, -y,
// -z faces lie on the mesh boundary; two_d: the apex z is az_2d (world z = 0 in
// the cell's normalised z coordinate).
template <class R>
__device__ inline R box_fraction(const Phi8<R>& h, bool bx, bool by, bool bz, bool two_d, R az_2d) {
    const R x01 = line_fraction<R>(h.f0, h.f1), x23 = line_fraction<R>(h.f2, h.f3),
            x45 = line_fraction<R>(h.f4, h.f5), x67 = line_fraction<R>(h.f6, h.f7);
    const R y02 = line_fraction<R>(h.f0, h.f2), y13 = line_fraction<R>(h.f1, h.f3),
            y46 = line_fraction<R>(h.f4, h.f6), y57 = line_fraction<R>(h.f5, h.f7);
    const R z04 = line_fraction<R>(h.f0, h.f4), z15 = line_fraction<R>(h.f1, h.f5),
            z26 = line_fraction<R>(h.f2, h.f6), z37 = line_fraction<R>(h.f3, h.f7);
    const R exm = bx ? unit_face_fraction<R, 0, 4, 6, 2, 1, 2>(h, z04, z26, y02, y46)
                     : unit_face_fraction<R, 0, 2, 6, 4, 1, 2>(h, z04, z26, y02, y46);
    const R exp_ = unit_face_fraction<R, 1, 3, 7, 5, 1, 2>(h, z15, z37, y13, y57);
    const R eym = by ? unit_face_fraction<R, 0, 1, 5, 4, 0, 2>(h, z04, z15, x01, x45)
                     : unit_face_fraction<R, 0, 4, 5, 1, 0, 2>(h, z04, z15, x01, x45);
    const R eyp = unit_face_fraction<R, 2, 6, 7, 3, 0, 2>(h, z26, z37, x23, x67);
    const R ezm = bz ? unit_face_fraction<R, 0, 2, 3, 1, 0, 1>(h, y02, y13, x01, x23)
                     : unit_face_fraction<R, 0, 1, 3, 2, 0, 1>(h, y02, y13, x01, x23);
    const R ezp = unit_face_fraction<R, 4, 5, 7, 6, 0, 1>(h, y46, y57, x45, x67);
    // cell apex: A = vertex 0 (the origin), B = first later vertex with a sign change, else 7
    const R f0 = h.f0;
    int b = 7;
    R pB = h.f7;
    if (f0 * h.f6 <= 0) { b = 6; pB = h.f6; }
    if (f0 * h.f5 <= 0) { b = 5; pB = h.f5; }
    if (f0 * h.f4 <= 0) { b = 4; pB = h.f4; }
    if (f0 * h.f3 <= 0) { b = 3; pB = h.f3; }
    if (f0 * h.f2 <= 0) { b = 2; pB = h.f2; }
    if (f0 * h.f1 <= 0) { b = 1; pB = h.f1; }
    const R s = qdiv<R>(fabs(f0), R(kSMALL) + fabs(f0) + fabs(pB));
    const R ax = (b & 1) ? s : R(0), ay = (b & 2) ? s : R(0);
    const R az = two_d ? az_2d : ((b & 4) ? s : R(0));
    const R vol = exm * fabs(ax) + exp_ * fabs(ax - R(1)) + eym * fabs(ay) + eyp * fabs(ay - R(1)) +
                  ezm * fabs(az) + ezp * fabs(az - R(1));
    return (R(1) / R(3)) * vol;
}

}  // namespace sdfibm_dev
