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

// Face vertex lists (indices into the 8 cell vertices).
// +x, +y, +z (internal or boundary), -x/-y/-z internal, -x/-y/-z boundary.
__constant__ const unsigned char kFacePX[4] = {1, 3, 7, 5};
__constant__ const unsigned char kFacePY[4] = {2, 6, 7, 3};
__constant__ const unsigned char kFacePZ[4] = {4, 5, 7, 6};
__constant__ const unsigned char kFaceMXi[4] = {0, 2, 6, 4};
__constant__ const unsigned char kFaceMYi[4] = {0, 4, 5, 1};
__constant__ const unsigned char kFaceMZi[4] = {0, 1, 3, 2};
__constant__ const unsigned char kFaceMXb[4] = {0, 4, 6, 2};
__constant__ const unsigned char kFaceMYb[4] = {0, 1, 5, 4};
__constant__ const unsigned char kFaceMZb[4] = {0, 2, 3, 1};

// Line fraction (geometrictools.cpp:13-23).
template <class R>
__device__ inline R line_fraction(R pa, R pb) {
    if (pa > 0 && pb > 0) return R(0);
    if (pa <= 0 && pb <= 0) return R(1);
    if (pa > 0) return -pb / (pa - pb);
    return -pa / (pb - pa);
}

// Apex rule (geometrictools.cpp:25-45): A = first listed vertex, B = first
// later vertex with phi_A*phi_B <= 0 (else the last one).
template <class R, int N>
__device__ inline r3<R> apex_of(const r3<R>* P, const R* phi, const unsigned char* lst) {
    const r3<R> A = P[lst[0]];
    const R pA = phi[lst[0]];
    r3<R> B = P[lst[1]];
    R pB = phi[lst[1]];
#pragma unroll
    for (int t = 1; t < N; ++t) {
        B = P[lst[t]];
        pB = phi[lst[t]];
        if (pA * pB <= R(0)) break;
    }
    const R s = fabs(pA) / (R(kSMALL) + fabs(pA) + fabs(pB));
    return {A.x - s * (A.x - B.x), A.y - s * (A.y - B.y), A.z - s * (A.z - B.z)};
}

// Area fraction of a quad face (geometrictools.cpp:74-116); sf = face area vector.
template <class R>
__device__ inline R face_area_fraction(const r3<R>* P, const R* phi, const unsigned char* f, r3<R> sf) {
    int ss = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) ss += (phi[f[t]] > R(0)) ? 1 : -1;
    if (ss == 4) return R(0);
    if (ss == -4) return R(1);
    const r3<R> ap = apex_of<R, 4>(P, phi, f);
    R area = R(0);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int a = f[t], b = f[(t + 1) & 3];
        const r3<R> O = P[a], A = P[b];
        const R tri = fabs(R(0.5) * rmag(rcross(sub(A, O), sub(ap, O))));
        area += tri * line_fraction<R>(phi[a], phi[b]);
    }
    return area / rmag(sf);
}

// Geometry of the 6 faces of a cell, in the cell's OpenFOAM face order.
template <class R>
struct CellFaces {
    r3<R> fc[6];  // centres relative to vertex 0
    r3<R> sf[6];  // area vectors
    const unsigned char* lst[6];
    int n;
};

// Volume fraction: sum over faces of (1/3) eps_f |(apex - fc) . Sf|, divided by V
// (geometrictools.cpp:47-72).  apex_z_rel: if two_d, the apex z (relative) is
// forced to this value, i.e. world z = 0 (geometrictools.cpp:57-58).
template <class R>
__device__ inline R cell_fraction(const r3<R>* P, const R* phi, const CellFaces<R>& F, R V, bool two_d,
                                  R apex_z_rel) {
    const unsigned char kCell[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    r3<R> ap = apex_of<R, 8>(P, phi, kCell);
    if (two_d) ap.z = apex_z_rel;
    R vol = R(0);
    for (int f = 0; f < F.n; ++f) {
        const R eps = face_area_fraction<R>(P, phi, F.lst[f], F.sf[f]);
        if (eps != R(0)) vol += (R(1) / R(3)) * eps * fabs(rdot(sub(ap, F.fc[f]), F.sf[f]));
    }
    return vol / V;
}

}  // namespace sdfibm_dev
