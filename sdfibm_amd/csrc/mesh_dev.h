// mesh_dev.h -- device view of a logically structured hex mesh and the
// per-solid cell tiles (shared by host set-up code and kernels).
#pragma once

#include "fraction.h"

namespace sdfibm_dev {

enum MeshKind { MK_RECT = 0, MK_STRUCT = 1 };

struct MeshDev {
    int nx, ny, nz, two_d;
    // rectilinear coordinates
    const double* xs;
    const double* ys;
    const double* zs;
    // structured: SoA points, SoA cell centres, volumes, AoS face centres/areas per family
    const double* pxyz;  // structured: points AoS [(nx+1)(ny+1)(nz+1)*3]
    const double* ccx;
    const double* ccy;
    const double* ccz;
    const double* vol;
    const double* fc[3];
    const double* sf[3];
    // structured: nominal uniform grid the points scatter around (origin, spacing) and
    // the largest point displacement from it, in cells
    double o[3], h[3], jit;
};

__host__ __device__ inline long long cell_id(const MeshDev& m, int i, int j, int k) {
    return (long long)i + (long long)m.nx * ((long long)j + (long long)m.ny * k);
}

// a = q d + r (0 <= r < d) for 0 <= a < 2^52 through the FP64 product a * (1/d),
// which is within one of the quotient: one correction step (a 64-bit integer
// division would cost far more instructions).
__device__ inline void divmod_i64(long long a, int d, double inv_d, long long& q, int& r) {
    long long t = (long long)((double)a * inv_d);
    long long rr = a - t * d;
    if (rr < 0) {
        --t;
        rr += d;
    } else if (rr >= d) {
        ++t;
        rr -= d;
    }
    q = t;
    r = (int)rr;
}

// (i, j, k) of cell c
__device__ inline void cell_ijk(const MeshDev& m, long long c, int& i, int& j, int& k) {
    long long q, kk;
    divmod_i64(c, m.nx, 1.0 / m.nx, q, i);
    divmod_i64(q, m.ny, 1.0 / m.ny, kk, j);
    k = (int)kk;
}

template <int MK>
__device__ inline d3 vertex_pos(const MeshDev& m, int i, int j, int k) {
    if constexpr (MK == MK_RECT) {
        return {m.xs[i], m.ys[j], m.zs[k]};
    } else {
        const long long p = (long long)i + (long long)(m.nx + 1) * ((long long)j + (long long)(m.ny + 1) * k);
        const double* q = m.pxyz + 3 * p;  // AoS: one cache line for a point's coordinates
        return {q[0], q[1], q[2]};
    }
}

template <int MK>
__device__ inline d3 cell_centre(const MeshDev& m, int i, int j, int k) {
    if constexpr (MK == MK_RECT) {
        return {(m.xs[i] + m.xs[i + 1]) * 0.5, (m.ys[j] + m.ys[j + 1]) * 0.5, (m.zs[k] + m.zs[k + 1]) * 0.5};
    } else {
        const long long c = cell_id(m, i, j, k);
        return {m.ccx[c], m.ccy[c], m.ccz[c]};
    }
}

template <int MK>
__device__ inline double cell_volume(const MeshDev& m, int i, int j, int k) {
    if constexpr (MK == MK_RECT) {
        return ((m.xs[i + 1] - m.xs[i]) * (m.ys[j + 1] - m.ys[j])) * (m.zs[k + 1] - m.zs[k]);
    } else {
        return m.vol[cell_id(m, i, j, k)];
    }
}

// The 8 vertex positions of cell (i,j,k) in cellPoints order.
template <int MK>
__device__ inline void cell_vertices(const MeshDev& m, int i, int j, int k, d3* P) {
#pragma unroll
    for (int v = 0; v < 8; ++v) P[v] = vertex_pos<MK>(m, i + (v & 1), j + ((v >> 1) & 1), k + ((v >> 2) & 1));
}

// Centre (relative to v0) and area vector of face slot of cell (i,j,k), slots
// -x, +x, -y, +y, -z, +z.
template <class R, int MK>
__device__ inline void hex_face(const MeshDev& m, int i, int j, int k, d3 v0, int slot, r3<R>& fco, r3<R>& sfo) {
    const int fam = slot >> 1, a = i + (slot == 1), b = j + (slot == 3), c = k + (slot == 5);
    {
        d3 fc, sf;
        if constexpr (MK == MK_RECT) {
            if (fam == 0) {
                fc = {m.xs[a], (m.ys[b] + m.ys[b + 1]) * 0.5, (m.zs[c] + m.zs[c + 1]) * 0.5};
                sf = {(m.ys[b + 1] - m.ys[b]) * (m.zs[c + 1] - m.zs[c]), 0.0, 0.0};
            } else if (fam == 1) {
                fc = {(m.xs[a] + m.xs[a + 1]) * 0.5, m.ys[b], (m.zs[c] + m.zs[c + 1]) * 0.5};
                sf = {0.0, (m.xs[a + 1] - m.xs[a]) * (m.zs[c + 1] - m.zs[c]), 0.0};
            } else {
                fc = {(m.xs[a] + m.xs[a + 1]) * 0.5, (m.ys[b] + m.ys[b + 1]) * 0.5, m.zs[c]};
                sf = {0.0, 0.0, (m.xs[a + 1] - m.xs[a]) * (m.ys[b + 1] - m.ys[b])};
            }
        } else {
            long long f;
            if (fam == 0)
                f = (long long)a + (long long)(m.nx + 1) * ((long long)b + (long long)m.ny * c);
            else if (fam == 1)
                f = (long long)a + (long long)m.nx * ((long long)b + (long long)(m.ny + 1) * c);
            else
                f = (long long)a + (long long)m.nx * ((long long)b + (long long)m.ny * c);
            const double* pc = m.fc[fam] + 3 * f;
            const double* ps = m.sf[fam] + 3 * f;
            fc = {pc[0], pc[1], pc[2]};
            sf = {ps[0], ps[1], ps[2]};
        }
        fco = {R(fc.x - v0.x), R(fc.y - v0.y), R(fc.z - v0.z)};
        sfo = {R(sf.x), R(sf.y), R(sf.z)};
    }
}

// Signed distance at world point p for the fraction: local = R^T (p - c) with
// the solid's rotation matrix (the inside tests keep the quaternion form).  A
// plane keeps the quaternion form: its distance is the local y itself, and a
// wall lying in a mesh face gives its points distances of a few 1e-17 (the
// rotation's rounding) whose ratios the fraction's edge crossings take, so the
// bits of those distances are the reference's only in its own transform.
template <class R>
__device__ inline R sdf_fast(const SolidK& K, d3 p) {
    if (K.shape == SDFIBM_SHAPE_PLANE) return R(world_to_local(K, p).y);
    const double dx = p.x - K.c[0], dy = p.y - K.c[1], dz = p.z - K.c[2];
    const d3 l{K.rt[0] * dx + K.rt[1] * dy + K.rt[2] * dz, K.rt[3] * dx + K.rt[4] * dy + K.rt[5] * dz,
               K.rt[6] * dx + K.rt[7] * dy + K.rt[8] * dz};
    if constexpr (sizeof(R) == sizeof(float))
        return sdf_local_mixed(K, l);
    else
        return sdf_local<R>(K, R(l.x), R(l.y), R(l.z));
}

// Volume fraction of cell (i,j,k) (calcCellVolume / V) from its 8 corner signed
// distances; phi_at(di, dj, dk) returns the distance at vertex (i+di, j+dj, k+dk).
template <class R, int MK, class PhiAt>
__device__ inline R fraction_with_phi(const MeshDev& m, int i, int j, int k, PhiAt phi_at) {
    const d3 w0 = vertex_pos<MK>(m, i, j, k);
    Hex8<R> h;
    auto vtx = [&](int di, int dj, int dk, r3<R>& p, R& f) {
        const d3 w = vertex_pos<MK>(m, i + di, j + dj, k + dk);
        p = {R(w.x - w0.x), R(w.y - w0.y), R(w.z - w0.z)};
        f = phi_at(di, dj, dk, w);
    };
    vtx(0, 0, 0, h.p0, h.f0);
    vtx(1, 0, 0, h.p1, h.f1);
    vtx(0, 1, 0, h.p2, h.f2);
    vtx(1, 1, 0, h.p3, h.f3);
    vtx(0, 0, 1, h.p4, h.f4);
    vtx(1, 0, 1, h.p5, h.f5);
    vtx(0, 1, 1, h.p6, h.f6);
    vtx(1, 1, 1, h.p7, h.f7);
    const bool bnd[3] = {i == 0, j == 0, k == 0};
    const R V = R(cell_volume<MK>(m, i, j, k));
    return hex_fraction<R>(h, [&](int slot, r3<R>& fc, r3<R>& sf) { hex_face<R, MK>(m, i, j, k, w0, slot, fc, sf); },
                           bnd, V, m.two_d != 0, R(0.0 - w0.z));
}

// Volume fraction of cell (i,j,k) inside solid K, corner distances evaluated here.
template <class R, int MK>
__device__ inline R fraction_of_cell(const MeshDev& m, const SolidK& K, int i, int j, int k) {
    return fraction_with_phi<R, MK>(m, i, j, k, [&](int, int, int, d3 w) { return sdf_fast<R>(K, w); });
}

// Corner signed distances of box cell (i,j,k) of a rectilinear mesh in FP32
// and the local coordinates of the cell centre.  Corner 0's local coordinates
// rt (p - c) in FP64, the others by FP64 steps along the cell's edges (the
// rotated edge vectors) in a Gray-code walk 0 1 3 2 6 7 5 4, one step per
// corner: rounding of order 1e-16 of |p - c|, far below the FP32 distances,
// with 12 FP64 values live instead of the 18 products of all corners.
__device__ inline void box_corner_phi(const MeshDev& m, const SolidK& K, int i, int j, int k, Phi8<float>& h, d3& lc) {
    const double dx = m.xs[i] - K.c[0], dy = m.ys[j] - K.c[1], dz = m.zs[k] - K.c[2];
    const double hx = m.xs[i + 1] - m.xs[i], hy = m.ys[j + 1] - m.ys[j], hz = m.zs[k + 1] - m.zs[k];
    d3 l{K.rt[0] * dx + K.rt[1] * dy + K.rt[2] * dz, K.rt[3] * dx + K.rt[4] * dy + K.rt[5] * dz,
         K.rt[6] * dx + K.rt[7] * dy + K.rt[8] * dz};
    const d3 ex{K.rt[0] * hx, K.rt[3] * hx, K.rt[6] * hx}, ey{K.rt[1] * hy, K.rt[4] * hy, K.rt[7] * hy},
        ez{K.rt[2] * hz, K.rt[5] * hz, K.rt[8] * hz};
    const d3 l0 = l;
    h.f0 = sdf_local_mixed(K, l);
    l = l + ex;
    h.f1 = sdf_local_mixed(K, l);
    l = l + ey;
    h.f3 = sdf_local_mixed(K, l);
    l = l - ex;
    h.f2 = sdf_local_mixed(K, l);
    l = l + ez;
    h.f6 = sdf_local_mixed(K, l);
    l = l + ex;
    h.f7 = sdf_local_mixed(K, l);
    lc = 0.5 * (l0 + l);
    l = l - ey;
    h.f5 = sdf_local_mixed(K, l);
    l = l - ex;
    h.f4 = sdf_local_mixed(K, l);
}

// Corner signed distances of box cell (i,j,k) for a sphere in FP32: its
// distance is invariant under the solid's rotation, |rt (p - c) + com| = |p - cw|
// with cw = c - R com its world centre (K.k[2..4]), so the 8 squared distances
// are sums of per-axis squares (6 products, 16 sums).
__device__ inline void sphere_corner_phi(const MeshDev& m, const SolidK& K, int i, int j, int k, Phi8<float>& h) {
    double sx[2], sy[2], sz[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const double dx = m.xs[i + b] - K.k[2], dy = m.ys[j + b] - K.k[3], dz = m.zs[k + b] - K.k[4];
        sx[b] = dx * dx;
        sy[b] = dy * dy;
        sz[b] = dz * dz;
    }
    const double r = K.k[0], r2 = K.k[1];
    auto f = [&](int bx, int by, int bz) { return filter_sd<float>(sphere_sd_mixed(sx[bx] + sy[by] + sz[bz], r, r2)); };
    h.f0 = f(0, 0, 0);
    h.f1 = f(1, 0, 0);
    h.f2 = f(0, 1, 0);
    h.f3 = f(1, 1, 0);
    h.f4 = f(0, 0, 1);
    h.f5 = f(1, 0, 1);
    h.f6 = f(0, 1, 1);
    h.f7 = f(1, 1, 1);
}

// Corner signed distances of box cell (i,j,k) for a plane in FP32: the local y of
// each corner by the quaternion transform, as sdf_fast forms it.  A plane's
// distance is not filtered, so for a wall lying in a mesh face the rounding-level
// distances of that face's vertices decide the cut; they keep the reference's
// operation order.
__device__ inline void plane_corner_phi(const MeshDev& m, const SolidK& K, int i, int j, int k, Phi8<float>& h) {
    auto f = [&](int bx, int by, int bz) {
        return (float)world_to_local(K, d3{m.xs[i + bx], m.ys[j + by], m.zs[k + bz]}).y;
    };
    h.f0 = f(0, 0, 0);
    h.f1 = f(1, 0, 0);
    h.f2 = f(0, 1, 0);
    h.f3 = f(1, 1, 0);
    h.f4 = f(0, 0, 1);
    h.f5 = f(1, 0, 1);
    h.f6 = f(0, 1, 1);
    h.f7 = f(1, 1, 1);
}

// Volume fraction of cell (i,j,k) inside solid K as the cut-cell kernels form
// it: in FP32 on a rectilinear mesh the box form (spheres without the
// rotation, planes by the quaternion transform); otherwise the general hex
// form (FP64 keeps the reference's operation order).  PL = false for the
// kernels of finite solids, which never see a plane.
template <class R, int MK, bool PL = true>
__device__ inline R cell_fraction(const MeshDev& m, const SolidK& K, int i, int j, int k) {
    if constexpr (MK == MK_RECT && sizeof(R) == sizeof(float)) {
        Phi8<float> h;
        d3 lc;
        if (K.shape == SDFIBM_SHAPE_SPHERE)
            sphere_corner_phi(m, K, i, j, k, h);
        else if (PL && K.shape == SDFIBM_SHAPE_PLANE)
            plane_corner_phi(m, K, i, j, k, h);
        else
            box_corner_phi(m, K, i, j, k, h, lc);
        const double z0 = m.zs[k], z1 = m.zs[k + 1];
        const float az = (float)((0.0 - z0) / (z1 - z0));
        return box_fraction<float>(h, i == 0, j == 0, k == 0, m.two_d != 0, az);
    } else {
        return fraction_of_cell<R, MK>(m, K, i, j, k);
    }
}

// ---------------------------------------------------------------------------
// per-solid tiles
// ---------------------------------------------------------------------------
// TC_BIG: a tile too big for the on-chip tile pass (> kMaxTileX cells along x or
// > kRowsMedium vertex rows, where one workgroup per solid would be the step's
// long pole) but with fewer than kBigVerts vertices: the big-tile pass (flat
// index spaces over the whole chip, global-memory labelling, chunked cut cells)
// with the same outputs.
// TC_TOOBIG: a tile beyond that; set_solids sends such a solid the planes' way
// (TC_HUGE: cells over the mesh, union-find component of the seed, its tile box
// as the cull).
enum TileClass {
    TC_EMPTY = 0, TC_SMALL = 1, TC_PLANE = 3, TC_TOOBIG = 4, TC_MEDIUM = 5, TC_HUGE = 6, TC_BIG = 7
};

__host__ __device__ inline bool tile_is_finite(int cls) {
    return cls == TC_SMALL || cls == TC_MEDIUM || cls == TC_BIG;
}

struct TileInfo {
    int lo[3];
    int n[3];
    long long off;  // offset of the tile's first code
    int seed;       // nearest cell to the solid centre (global index), -1 if none
    int cls;        // TileClass
};

constexpr int kMaxTileX = 63;      // cells along x per tile (vertex row fits one 64-bit word)
constexpr int kRowsTiny = 64;      // vertex rows of a tiny tile (classified by one wave)
constexpr int kRowsSmall = 400;    // vertex rows (ny+1)(nz+1) of a small tile
constexpr int kRowsMedium = 1024;  // ... of a medium tile (beyond: the big-tile pass)
constexpr long long kBigVerts = 1LL << 22;  // vertices of a big tile (the big pass's 22-bit index arithmetic)

// Searches in an increasing array a[0..n): start from a linear-interpolation
// guess and step to the exact answer, which takes one or two probes on a
// uniform axis (a binary search would need a chain of log2(n) dependent loads).
__host__ __device__ inline int guess_index(const double* a, int n, double v) {
    const double a0 = a[0], a1 = a[n - 1];
    double g = (a1 > a0) ? (v - a0) / (a1 - a0) * (double)(n - 1) : 0.0;
    if (!(g > 0.0)) g = 0.0;
    if (g > (double)(n - 1)) g = (double)(n - 1);
    return (int)g;
}
// first index with a[idx] >= v (n if none)
__host__ __device__ inline int lower_bound_d(const double* a, int n, double v) {
    int g = guess_index(a, n, v);
    while (g > 0 && a[g - 1] >= v) --g;
    while (g < n && a[g] < v) ++g;
    return g;
}
// last index with a[idx] <= v (-1 if none)
__host__ __device__ inline int last_le_d(const double* a, int n, double v) {
    int g = guess_index(a, n, v);
    while (g < n && a[g] <= v) ++g;
    while (g > 0 && a[g - 1] > v) --g;
    return g - 1;
}

// Rotation matrix of q (local -> world).
__host__ __device__ inline void rot_of(const double* q, double Rm[3][3]) {
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    Rm[0][0] = 1 - 2 * (y * y + z * z);
    Rm[0][1] = 2 * (x * y - w * z);
    Rm[0][2] = 2 * (x * z + w * y);
    Rm[1][0] = 2 * (x * y + w * z);
    Rm[1][1] = 1 - 2 * (x * x + z * z);
    Rm[1][2] = 2 * (y * z - w * x);
    Rm[2][0] = 2 * (x * z - w * y);
    Rm[2][1] = 2 * (y * z + w * x);
    Rm[2][2] = 1 - 2 * (x * x + y * y);
}

// World-space axis-aligned bounds (centre wc, half extents wh, padded) of the
// region where a solid's inside test can hold; false for unbounded shapes.
// Bodies are bounded through their local box; the tailed discs through the
// union of the disc's box and the exact corners of the tail rectangle(s).
__host__ __device__ inline bool world_bounds(const SolidK& K, double* wc, double* wh) {
    const double INF = __builtin_huge_val();
    double Rm[3][3];
    rot_of(K.q, Rm);
    double lo[3] = {INF, INF, INF}, hi[3] = {-INF, -INF, -INF};
    auto add_point = [&](double px, double py, double pz) {  // local point
        for (int d = 0; d < 3; ++d) {
            const double w = K.c[d] + (Rm[d][0] * px + Rm[d][1] * py + Rm[d][2] * pz);
            lo[d] = fmin(lo[d], w);
            hi[d] = fmax(hi[d], w);
        }
    };
    // local z extrusion of the 2-D shapes: unbounded along any world axis local z touches
    auto extrude = [&]() {
        for (int d = 0; d < 3; ++d)
            if (fabs(Rm[d][2]) > 1e-12) {
                lo[d] = -INF;
                hi[d] = INF;
            }
    };
    if (K.shape == SDFIBM_SHAPE_CIRCLE_TAIL || K.shape == SDFIBM_SHAPE_CIRCLE_TWOTAIL) {
        const double r = K.k[0], ra = K.k[2], rb = K.k[3];
        const double cx = -K.com[0], cy = -K.com[1];
        for (int d = 0; d < 3; ++d) {  // the disc about -com in the local xy plane
            const double w = K.c[d] + (Rm[d][0] * cx + Rm[d][1] * cy);
            const double e = r * sqrt(Rm[d][0] * Rm[d][0] + Rm[d][1] * Rm[d][1]);
            lo[d] = fmin(lo[d], w - e);
            hi[d] = fmax(hi[d], w + e);
        }
        const double ux[4] = {0.0, 2 * ra, 2 * ra, 0.0}, uy[4] = {-rb, -rb, rb, rb};
        if (K.shape == SDFIBM_SHAPE_CIRCLE_TAIL) {
            for (int t = 0; t < 4; ++t) add_point(ux[t] + cx, uy[t] + cy, 0.0);
        } else {
            // tails: rot30(q) or rot30(flipy(q)) in [0, 2ra] x [-rb, rb]; invert rot30
            const double c30 = 0.866025404, det = c30 * c30 + 0.25;
            for (int t = 0; t < 4; ++t) {
                const double qx = (c30 * ux[t] - 0.5 * uy[t]) / det, qy = (0.5 * ux[t] + c30 * uy[t]) / det;
                add_point(qx + cx, qy + cy, 0.0);
                add_point(qx + cx, -qy + cy, 0.0);
            }
        }
        extrude();
    } else {
        d3 lc, lh;
        if (!local_bounds(K, lc, lh)) return false;
        // ellipsoids and ellipses: exact support function sqrt(sum (R_de l_e)^2);
        // boxes and rectangles: sum |R_de| l_e (the rotated corners)
        const bool quadric = K.shape == SDFIBM_SHAPE_ELLIPSOID || K.shape == SDFIBM_SHAPE_ELLIPSE ||
                             K.shape == SDFIBM_SHAPE_SPHERE || K.shape == SDFIBM_SHAPE_CIRCLE;
        const double lcv[3] = {lc.x, lc.y, lc.z}, lhv[3] = {lh.x, lh.y, lh.z};
        for (int d = 0; d < 3; ++d) {
            const double w = K.c[d] + (Rm[d][0] * lcv[0] + Rm[d][1] * lcv[1] + Rm[d][2] * lcv[2]);
            double e = 0.0;
            for (int f = 0; f < 3; ++f) {
                const double rr = fabs(Rm[d][f]);
                if (lhv[f] > 1e300) {
                    if (rr > 1e-12) e = INF;
                } else {
                    e += quadric ? (rr * lhv[f]) * (rr * lhv[f]) : rr * lhv[f];
                }
            }
            if (quadric && e < INF) e = sqrt(e);
            lo[d] = w - e;
            hi[d] = w + e;
        }
    }
    for (int d = 0; d < 3; ++d) {
        if (!(hi[d] - lo[d] < 1e300)) {
            wc[d] = 0.0;
            wh[d] = INF;
            continue;
        }
        wc[d] = 0.5 * (lo[d] + hi[d]);
        wh[d] = 0.5 * (hi[d] - lo[d]) * (1.0 + 1e-6) + 1e-12 * (fabs(wc[d]) + 1.0);
    }
    return true;
}

// Cell tile of a solid: the index box containing every cell that has a vertex
// strictly inside the solid, and the nearest cell to the solid centre.
// `cx`,`cy`,`cz` are the 1-D coordinates (rectilinear) used for the seed.
template <int MK>
__host__ __device__ inline TileInfo make_tile(const MeshDev& m, const SolidK& K) {
    TileInfo T{};
    T.seed = -1;
    T.off = 0;
    double wc[3], wh[3];
    if (!world_bounds(K, wc, wh)) {
        T.cls = TC_PLANE;
        return T;
    }
    const int nn[3] = {m.nx, m.ny, m.nz};
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {  // unrolled: the per-axis arrays stay in registers
        const double a = wc[d] - wh[d], b = wc[d] + wh[d];
        int vlo, vhi;  // vertex index range that may hold inside vertices
        if (!(wh[d] < 1e300)) {
            vlo = 0;
            vhi = nn[d];
        } else if constexpr (MK == MK_RECT) {
            const double* xs = d == 0 ? m.xs : (d == 1 ? m.ys : m.zs);
            vlo = lower_bound_d(xs, nn[d] + 1, a);
            vhi = last_le_d(xs, nn[d] + 1, b);
        } else {
            const double fa = (a - m.o[d]) / m.h[d] - m.jit, fb = (b - m.o[d]) / m.h[d] + m.jit;
            vlo = fa < -1.0 ? 0 : (int)ceil(fa);
            vhi = fb > (double)nn[d] + 1.0 ? nn[d] : (int)floor(fb);
            if (vlo < 0) vlo = 0;
            if (vhi > nn[d]) vhi = nn[d];
        }
        if (vlo > vhi) {
            T.cls = TC_EMPTY;
            return T;
        }
        lo[d] = vlo - 1 < 0 ? 0 : vlo - 1;
        hi[d] = vhi > nn[d] - 1 ? nn[d] - 1 : vhi;
        if (lo[d] > hi[d]) {
            T.cls = TC_EMPTY;
            return T;
        }
    }
    for (int d = 0; d < 3; ++d) {
        T.lo[d] = lo[d];
        T.n[d] = hi[d] - lo[d] + 1;
    }
    const long long vrows = (long long)(T.n[1] + 1) * (T.n[2] + 1);
    if (T.n[0] > kMaxTileX || vrows > kRowsMedium)
        T.cls = (vrows * (T.n[0] + 1) < kBigVerts && T.n[0] < 4095 && T.n[1] < 4095) ? TC_BIG : TC_TOOBIG;
    else
        T.cls = vrows > kRowsSmall ? TC_MEDIUM : TC_SMALL;
    return T;
}

// Nearest cell centre to p (meshSearch::findNearestCell).  Ties -- squared
// distances within a relative 1e-10 of the minimum -- go to the lowest cell
// index, so that rounding of the centres does not decide between equidistant
// cells.
constexpr double kNearTie = 1e-10;

template <int MK>
__host__ __device__ inline int nearest_cell(const MeshDev& m, const double* p) {
    int c0[3], r;
    const int nn[3] = {m.nx, m.ny, m.nz};
    if constexpr (MK == MK_RECT) {
        // per axis the nearest centre is one of two neighbours; search that 2x2x2 block
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double* xs = d == 0 ? m.xs : (d == 1 ? m.ys : m.zs);
            // first cell whose centre is >= p (centres increase), by guess and step
            int lo = guess_index(xs, nn[d] + 1, p[d]);
            if (lo > nn[d] - 1) lo = nn[d] - 1;
            while (lo > 0 && (xs[lo - 1] + xs[lo]) * 0.5 >= p[d]) --lo;
            while (lo < nn[d] - 1 && (xs[lo] + xs[lo + 1]) * 0.5 < p[d]) ++lo;
            c0[d] = lo > 0 ? lo - 1 : 0;
        }
        r = 1;
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double f = floor((p[d] - m.o[d]) / m.h[d]);
            if (f < 0) f = 0;
            if (f > nn[d] - 1) f = nn[d] - 1;
            c0[d] = (int)f;
        }
        r = (int)ceil(2.0 * m.jit) + 1;
    }
    const int lo0 = MK == MK_RECT ? 0 : -r;
    double best = __builtin_huge_val();
    for (int pass = 0; pass < 2; ++pass) {
        long long bc = -1;
        const double lim = best * (1.0 + kNearTie) + 1e-300;
        for (int k = c0[2] + lo0; k <= c0[2] + r; ++k) {
            if (k < 0 || k >= m.nz) continue;
            for (int j = c0[1] + lo0; j <= c0[1] + r; ++j) {
                if (j < 0 || j >= m.ny) continue;
                for (int i = c0[0] + lo0; i <= c0[0] + r; ++i) {
                    if (i < 0 || i >= m.nx) continue;
                    const long long c = cell_id(m, i, j, k);
                    double cx, cy, cz;
                    if constexpr (MK == MK_RECT) {
                        cx = (m.xs[i] + m.xs[i + 1]) * 0.5;
                        cy = (m.ys[j] + m.ys[j + 1]) * 0.5;
                        cz = (m.zs[k] + m.zs[k + 1]) * 0.5;
                    } else {
                        cx = m.ccx[c];
                        cy = m.ccy[c];
                        cz = m.ccz[c];
                    }
                    const double dx = cx - p[0], dy = cy - p[1], dz = cz - p[2];
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    if (pass == 0) {
                        if (d2 < best) best = d2;
                    } else if (d2 <= lim && (bc < 0 || c < bc)) {
                        bc = c;
                    }
                }
            }
        }
        if (pass == 1) return (int)bc;
    }
    return -1;
}

}  // namespace sdfibm_dev
