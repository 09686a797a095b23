// poly_mesh.h -- the coupling step on a general polyhedral mesh (OpenFOAM's
// polyMesh: points, faces, owner, neighbour), for meshes that are not one
// structured block: several blockMesh blocks, O-grids, snappyHexMesh output.
//
// Included by interact.hip inside its kernel namespace, after the helpers it
// uses (scan, EF_* flags, code_none).  The results are the reference's
// (src/solidcloud.cpp:361-433 over src/cellenumerator.cpp:6-78 and
// src/geometrictools.cpp:6-116): per solid the cells with >= 1 strictly inside
// point, connected through faces to the seed (the nearest cell if it has an
// inside point, else the lowest such cell), ALL_INSIDE when every point is
// inside; fractions from the cell's faces with the apex rule over cellPoints
// (ascending labels) and each face's own point order.
//
// The mesh carries no (i, j, k), so the pass works on (solid, cell) pairs:
//   k_poly_setup     per solid (a wave each): shape constants, world box, the
//                    nearest cell (ring search over the cell bins), the number
//                    of chunks of 512 bins of its box's bin index box
//   k_poly_visit     per chunk (a wave each): the cells of its bins, counted,
//                    then listed bin by bin (a fixed order: solid, bin, cell)
//   k_poly_test      per visited pair: strict inside test of the cell's points
//                    (FP64) -> active / ALL_INSIDE flags; lowest active cell
//   k_poly_compact   active pairs -> pair arrays (solid-major: a solid's pairs
//                    are contiguous and in solid order) and a hash table
//                    (solid, cell) -> pair
//   k_poly_hook      union over the faces between active pairs of one solid
//   k_poly_compress  labels -> roots
//   k_poly_force     per pair: membership in the seed's component, code (the
//                    fraction, signed by the centre test; +inf ALL_INSIDE; NaN
//                    not a candidate), FP64 force/torque contribution
//   k_poly_pair_part per chunk of 4,096 of a solid's pairs (a wave each): the
//                    partial sums of their contributions and their counts
//   k_poly_reduce    per solid (a wave each): its chunks' partial sums in chunk
//                    order (x rho_f), its candidate counts
//   k_poly_cells     per cell: the member pairs that touch it (counted, then
//                    listed); k_poly_fields writes every cell once, its solids
//                    in index order (= pair order), As clipped.
// The host reads the number of visits and of pairs back (two synchronisations
// per step) and sizes the buffers to them.
#pragma once

struct PolyDev {
    int n_cells, n_faces, n_points, two_d;
    const double* pxyz;  // points AoS [n_points*3]
    const int* f_off;    // faces: point lists (CSR)
    const int* f_pts;
    const double* fc;    // face centres [n_faces*3]
    const double* sf;    // face area vectors [n_faces*3]
    const int* c_foff;   // cells: faces, owned faces ascending then neighbour faces ascending
    const int* c_faces;
    const int* c_poff;   // cells: points, ascending labels
    const int* c_pts;
    const int* c_noff;   // cells: neighbour cells across internal faces
    const int* c_nbs;
    const double* cc;    // cell centres [n_cells*3]
    const double* vol;   // cell volumes
    double mlo[3], mhi[3];  // the box of the mesh points
    // bins of cells by centre: a uniform grid over the cell centres' box
    double blo[3], binv[3], bw[3];
    int bn[3];
    const int* b_off;     // [nbins+1]
    const int* b_cells;   // cells of each bin, ascending
    const double* b_box;  // [nbins*6]: the union of the bin's cells' point boxes (lo xyz, hi xyz)
};

constexpr unsigned long long kPolyEmpty = ~0ull;

__device__ inline d3 poly_point(const PolyDev& P, int p) {
    const double* q = P.pxyz + 3LL * p;
    return {q[0], q[1], q[2]};
}

__device__ inline unsigned long long poly_key(int s, int c) {
    return ((unsigned long long)(unsigned)s << 32) | (unsigned)c;
}
__device__ inline unsigned long long poly_hash(unsigned long long key, int bits) {
    return (key * 0x9E3779B97F4A7C15ull) >> (64 - bits);
}
// Pair index of (solid s, cell c), -1 if it is not an active pair.
__device__ inline int poly_lookup(const unsigned long long* hk, const int* hv, int bits, int s, int c) {
    const unsigned long long key = poly_key(s, c), mask = (1ull << bits) - 1ull;
    for (unsigned long long h = poly_hash(key, bits);; h = (h + 1) & mask) {
        const unsigned long long k = hk[h];
        if (k == key) return hv[h];
        if (k == kPolyEmpty) return -1;
    }
}

// The bins a world box [lo, hi] overlaps (index box, clamped), false if none.
__device__ inline bool poly_bin_box(const PolyDev& P, const double* lo, const double* hi, int* b0, int* b1) {
    for (int d = 0; d < 3; ++d)
        if (lo[d] > hi[d]) {  // an empty box: no bins
            b0[0] = b0[1] = b0[2] = 0;
            b1[0] = b1[1] = b1[2] = -1;
            return false;
        }
    for (int d = 0; d < 3; ++d) {
        double a = (lo[d] - P.blo[d]) * P.binv[d], b = (hi[d] - P.blo[d]) * P.binv[d];
        if (!(a > -1e9)) a = -1e9;  // unbounded axes
        if (!(b < 1e9)) b = 1e9;
        int ia = (int)floor(a), ib = (int)floor(b);
        // a cell's points may reach beyond its bin: the bin boxes decide; here the index box
        // is only widened to every bin (the bin boxes are tested one by one)
        ia = ia < 0 ? 0 : (ia >= P.bn[d] ? P.bn[d] - 1 : ia);
        ib = ib < 0 ? 0 : (ib >= P.bn[d] ? P.bn[d] - 1 : ib);
        b0[d] = ia;
        b1[d] = ib;
    }
    return true;
}

__device__ inline bool poly_box_hit(const double* bb, const double* lo, const double* hi) {
    return bb[0] <= hi[0] && bb[3] >= lo[0] && bb[1] <= hi[1] && bb[4] >= lo[1] && bb[2] <= hi[2] && bb[5] >= lo[2];
}

// A bin's box can hold a point of the solid: it meets the solid's world box and, for a plane
// (inside: local y < 0), its lowest local y is below 0.  The plane test keeps a margin of
// 1e-12 of the terms' size: a point on a wall lying in a mesh face has a local y of a few
// 1e-17 (the rotation's rounding), which decides it and which this sum's rounding can lose.
__device__ inline bool poly_bin_hit(const SolidK& K, const double* bb, const double* lo, const double* hi) {
    if (!poly_box_hit(bb, lo, hi)) return false;
    if (K.shape != SDFIBM_SHAPE_PLANE) return true;
    double y = 0.0, r = 0.0, m = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double n = K.rt[3 + d], c = 0.5 * (bb[d] + bb[3 + d]) - K.c[d], h = 0.5 * (bb[3 + d] - bb[d]);
        y += n * c;
        r += fabs(n) * h;
        m += fabs(n) * (fabs(bb[d]) + fabs(bb[3 + d]) + fabs(K.c[d]));
    }
    return y - r < 1e-12 * m + 1e-300;
}

// The world box of the mesh points on the inside of a plane (local y < 0): the
// box of the mesh box's corners on the inside of a relaxed plane, f(p) < delta, and
// of its edges' crossings of f = delta (that part of the mesh box is the convex hull
// of those points).  delta (1e-12 of the terms' size) bounds the rounding of the
// inside test's local y, which decides points lying in the plane (a wall on a mesh
// face): the relaxed half-space holds every point that test can call inside.
__device__ inline void poly_plane_box(const PolyDev& P, const SolidK& K, double* lo, double* hi) {
    const double n[3] = {K.rt[3], K.rt[4], K.rt[5]};
    double delta = 1e-300;
    for (int d = 0; d < 3; ++d)
        delta += 1e-12 * fabs(n[d]) * (fmax(fabs(P.mlo[d]), fabs(P.mhi[d])) + fabs(K.c[d]));
    auto g = [&](const double* p) {
        return n[0] * (p[0] - K.c[0]) + n[1] * (p[1] - K.c[1]) + n[2] * (p[2] - K.c[2]) - delta;
    };
    for (int d = 0; d < 3; ++d) {
        lo[d] = __builtin_huge_val();
        hi[d] = -__builtin_huge_val();
    }
    auto add = [&](const double* p) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], p[d]);
            hi[d] = fmax(hi[d], p[d]);
        }
    };
    for (int v = 0; v < 8; ++v) {
        double a[3];
        for (int d = 0; d < 3; ++d) a[d] = (v >> d) & 1 ? P.mhi[d] : P.mlo[d];
        const double ga = g(a);
        if (ga < 0) add(a);
        for (int d = 0; d < 3; ++d) {  // the edge to the corner across axis d (each edge once)
            if ((v >> d) & 1) continue;
            double b[3] = {a[0], a[1], a[2]};
            b[d] = P.mhi[d];
            const double gb = g(b);
            if ((ga < 0) != (gb < 0)) {
                double x[3] = {a[0], a[1], a[2]};
                x[d] = fmin(fmax(a[d] + ga / (ga - gb) * (b[d] - a[d]), a[d]), b[d]);
                add(x);
            }
        }
    }
    // the crossings' own rounding: widen by a hair of the mesh size
    for (int d = 0; d < 3; ++d) {
        const double e = 1e-9 * (P.mhi[d] - P.mlo[d]) + 1e-12 * fmax(fabs(P.mlo[d]), fabs(P.mhi[d])) + 1e-300;
        lo[d] -= e;
        hi[d] += e;
    }
}

// The bin index box of a solid's world box widened by `reach` bins (clamped to the
// grid), and its number of bins (0 for an empty box).
constexpr int kPolyChunk = 512;  // bins per visit chunk (a wave takes 64 at a time)
__device__ inline long long poly_solid_bins(const PolyDev& P, const double* lo, const double* hi, int reach, int* b0o,
                                            int* b1o) {
    int b0[3], b1[3];
    const bool any = poly_bin_box(P, lo, hi, b0, b1);
    for (int d = 0; d < 3; ++d) {
        b0[d] = max(0, b0[d] - reach);
        b1[d] = min(P.bn[d] - 1, b1[d] + reach);
        if (b0o) b0o[d] = b0[d];
        if (b1o) b1o[d] = b1[d];
    }
    return any ? (long long)(b1[0] - b0[0] + 1) * (b1[1] - b0[1] + 1) * (b1[2] - b0[2] + 1) : 0;
}

// Per solid (one wave; 4 per workgroup): constants, world box (a cell can hold an
// inside point only if its point box meets it), the nearest cell, and the number
// of cells in the bins whose cell boxes meet the solid's box.  A bin's cells can
// reach beyond the bin (the bin boxes hold their point boxes), so every bin is a
// candidate and its box decides; `reach` widens the bin index box by the cells'
// largest extent beyond their bin, in bins.
__global__ void __launch_bounds__(256) k_poly_setup(const sdfibm_solid* __restrict__ in, int ns, PolyDev P,
                                                    SolidK* __restrict__ Ks, double* __restrict__ box,
                                                    int* __restrict__ seed, long long* __restrict__ vcnt,
                                                    double* __restrict__ ft, int* __restrict__ first,
                                                    const CsgK* __restrict__ csg, int reach) {
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= ns) return;  // wave-uniform
    const SolidK K = make_solid_k(in[s], csg);
    double wc[3], wh[3], lo[3], hi[3];
    const bool bounded = world_bounds(K, wc, wh);
    for (int d = 0; d < 3; ++d) {
        lo[d] = bounded && wh[d] < 1e300 ? wc[d] - wh[d] : -__builtin_huge_val();
        hi[d] = bounded && wh[d] < 1e300 ? wc[d] + wh[d] : __builtin_huge_val();
    }
    if (K.shape == SDFIBM_SHAPE_PLANE) poly_plane_box(P, K, lo, hi);
    if (lane == 0) {
        Ks[s] = K;
        for (int d = 0; d < 3; ++d) {
            box[6 * s + d] = lo[d];
            box[6 * s + 3 + d] = hi[d];
        }
        for (int q = 0; q < 6; ++q) ft[6 * s + q] = 0.0;
        first[s] = 0x7fffffff;
    }
    // the bins of the solid's (widened) index box, in chunks of kPolyChunk
    if (lane == 0) vcnt[s] = (poly_solid_bins(P, lo, hi, reach, nullptr, nullptr) + kPolyChunk - 1) / kPolyChunk;

    // nearest cell centre (meshSearch::findNearestCell): rings of bins around the centre's
    // bin until no unvisited bin can hold a nearer centre; ties (squared distances within a
    // relative 1e-10 of the smallest) go to the lowest cell index
    int c0[3];
    for (int d = 0; d < 3; ++d) {
        double f = floor((K.c[d] - P.blo[d]) * P.binv[d]);
        f = f < 0 ? 0 : (f > P.bn[d] - 1 ? P.bn[d] - 1 : f);
        c0[d] = (int)f;
    }
    const int maxr = max(P.bn[0], max(P.bn[1], P.bn[2]));
    double best = __builtin_huge_val();
    int r = 0;
    // the block of bins within Chebyshev distance r of c0, clamped to the grid (so a ring costs
    // at most the grid's bins)
    auto clamp_box = [&](int rr, int* lo3, int* ex3) {
        for (int d = 0; d < 3; ++d) {
            lo3[d] = max(0, c0[d] - rr);
            ex3[d] = min(P.bn[d] - 1, c0[d] + rr) - lo3[d] + 1;
        }
    };
    for (;; ++r) {
        // the ring of bins at Chebyshev distance r, over the lanes
        int lo3[3], ex3[3];
        clamp_box(r, lo3, ex3);
        const long long nring = (long long)ex3[0] * ex3[1] * ex3[2];
        for (long long t = lane; t < nring; t += 64) {
            const int i = lo3[0] + (int)(t % ex3[0]), j = lo3[1] + (int)((t / ex3[0]) % ex3[1]);
            const int k = lo3[2] + (int)(t / ((long long)ex3[0] * ex3[1]));
            if (max(abs(i - c0[0]), max(abs(j - c0[1]), abs(k - c0[2]))) != r) continue;
            const long long b = i + (long long)P.bn[0] * (j + (long long)P.bn[1] * k);
            for (int q = P.b_off[b]; q < P.b_off[b + 1]; ++q) {
                const int c = P.b_cells[q];
                const double dx = P.cc[3 * c] - K.c[0], dy = P.cc[3 * c + 1] - K.c[1], dz = P.cc[3 * c + 2] - K.c[2];
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 < best) best = d2;
            }
        }
        for (int off = 32; off >= 1; off >>= 1) best = fmin(best, __shfl_xor(best, off, 64));
        // distance from the centre to the nearest face of the visited block that has bins beyond it
        double lb = __builtin_huge_val();
        bool more = false;
        for (int d = 0; d < 3; ++d) {
            if (c0[d] - r > 0) {
                more = true;
                lb = fmin(lb, K.c[d] - (P.blo[d] + (c0[d] - r) * P.bw[d]));
            }
            if (c0[d] + r + 1 < P.bn[d]) {
                more = true;
                lb = fmin(lb, (P.blo[d] + (c0[d] + r + 1) * P.bw[d]) - K.c[d]);
            }
        }
        const double lim = best * (1.0 + kNearTie) + 1e-300;
        if (!more || r >= maxr || (best < __builtin_huge_val() && lb > 0 && lb * lb > lim)) break;
    }
    // lowest index among the near ties over the visited block
    const double lim = best * (1.0 + kNearTie) + 1e-300;
    int bc = 0x7fffffff;
    int lo3[3], ex3[3];
    clamp_box(r, lo3, ex3);
    const long long nbox = (long long)ex3[0] * ex3[1] * ex3[2];
    for (long long t = lane; t < nbox; t += 64) {
        const int i = lo3[0] + (int)(t % ex3[0]), j = lo3[1] + (int)((t / ex3[0]) % ex3[1]);
        const int k = lo3[2] + (int)(t / ((long long)ex3[0] * ex3[1]));
        const long long b = i + (long long)P.bn[0] * (j + (long long)P.bn[1] * k);
        for (int q = P.b_off[b]; q < P.b_off[b + 1]; ++q) {
            const int c = P.b_cells[q];
            const double dx = P.cc[3 * c] - K.c[0], dy = P.cc[3 * c + 1] - K.c[1], dz = P.cc[3 * c + 2] - K.c[2];
            if (dx * dx + dy * dy + dz * dz <= lim && c < bc) bc = c;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) bc = min(bc, __shfl_xor(bc, off, 64));
    if (lane == 0) seed[s] = bc == 0x7fffffff ? -1 : bc;
}

// Chunk -> solid map: chunk coff[s] + k of solid s, k < coff[s + 1] - coff[s]
// (one wave per solid).
__global__ void __launch_bounds__(256) k_poly_chunk_map(int ns, const long long* __restrict__ coff,
                                                        int* __restrict__ csol) {
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= ns) return;
    for (long long c = coff[s] + lane; c < coff[s + 1]; c += 64) csol[c] = s;
}

// The visits of one chunk of kPolyChunk bins of a solid's index box (one wave
// per chunk): its bins in bin index order, 64 at a time, one per lane; the
// cells of the bins whose boxes can hold an inside point, each bin's cells
// ascending.  count: the number of cells only (into cnt[chunk]); else the
// cells and solids written from voff[chunk].
template <bool COUNT>
__global__ void __launch_bounds__(256) k_poly_visit(long long nchunk, PolyDev P, const SolidK* __restrict__ Ks,
                                                    const double* __restrict__ box, const int* __restrict__ csol,
                                                    const long long* __restrict__ coff, long long* __restrict__ cnt,
                                                    const long long* __restrict__ voff, int* __restrict__ vcell,
                                                    int* __restrict__ vsolid, int reach) {
    const long long ch = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (ch >= nchunk) return;
    const int s = csol[ch];
    const SolidK& K = Ks[s];
    const double* lo = box + 6 * s;
    const double* hi = lo + 3;
    int b0[3], b1[3];
    const long long nb = poly_solid_bins(P, lo, hi, reach, b0, b1);
    const int ex = b1[0] - b0[0] + 1, ey = b1[1] - b0[1] + 1;
    const long long t_lo = (ch - coff[s]) * kPolyChunk, t_hi = min(nb, t_lo + kPolyChunk);
    long long o = COUNT ? 0 : voff[ch];
    for (long long t0 = t_lo; t0 < t_hi; t0 += 64) {
        const long long t = t0 + lane;
        int q0 = 0, n = 0;
        if (t < t_hi) {
            const int bi = b0[0] + (int)(t % ex), bj = b0[1] + (int)((t / ex) % ey), bk = b0[2] + (int)(t / ((long long)ex * ey));
            const long long b = bi + (long long)P.bn[0] * (bj + (long long)P.bn[1] * bk);
            if (poly_bin_hit(K, P.b_box + 6 * b, lo, hi)) {
                q0 = P.b_off[b];
                n = P.b_off[b + 1] - q0;
            }
        }
        int incl = n;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (!COUNT) {
            const long long at = o + incl - n;
            for (int q = 0; q < n; ++q) {
                vcell[at + q] = P.b_cells[q0 + q];
                vsolid[at + q] = s;
            }
        }
        o += __shfl(incl, 63, 64);
    }
    if (COUNT && lane == 0) cnt[ch] = o;
}

// A solid's first visit: the visit offset of its first chunk (s = ns: the total).
__global__ void k_poly_solid_voff(int ns, const long long* __restrict__ coff, const long long* __restrict__ cvoff,
                                  long long* __restrict__ voff) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= ns) voff[s] = cvoff[coff[s]];
}

// Strict inside test of each visited cell's points (FP64, the reference's
// operation order): act = 1 when >= 1 point is inside, all = 1 when every one
// is; the lowest active cell of each solid.
__global__ void __launch_bounds__(256) k_poly_test(PolyDev P, const SolidK* __restrict__ Ks, long long nv,
                                                   const int* __restrict__ vcell, const int* __restrict__ vsolid,
                                                   unsigned char* __restrict__ act, unsigned char* __restrict__ all,
                                                   int* __restrict__ first) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (long long)gridDim.x * blockDim.x) {
        const int c = vcell[v], s = vsolid[v];
        const SolidK K = Ks[s];  // the fields the test reads, once per visit
        int nin = 0;
        const int p0 = P.c_poff[c], p1 = P.c_poff[c + 1];
        if (p1 - p0 == 8) {  // hexahedra: the eight labels, then the eight points, in flight together
            int id[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) id[q] = P.c_pts[p0 + q];
            d3 pt[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) pt[q] = poly_point(P, id[q]);
#pragma unroll
            for (int q = 0; q < 8; ++q) nin += inside_world(K, pt[q]) ? 1 : 0;
        } else {
            for (int t = p0; t < p1; ++t) nin += inside_world(K, poly_point(P, P.c_pts[t])) ? 1 : 0;
        }
        act[v] = nin > 0 ? 1 : 0;
        all[v] = nin == p1 - p0 ? 1 : 0;
        if (nin > 0 && c < __hip_atomic_load(&first[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&first[s], c);
    }
}

// Active visits -> pairs (at their scanned positions; a solid's pairs stay
// contiguous and in solid order), labels = own index, hash table entries.
__global__ void __launch_bounds__(256) k_poly_compact(long long nv, const int* __restrict__ vcell,
                                                      const int* __restrict__ vsolid,
                                                      const unsigned char* __restrict__ act,
                                                      const unsigned char* __restrict__ all,
                                                      const long long* __restrict__ vpos, int* __restrict__ pcell,
                                                      int* __restrict__ psolid, unsigned char* __restrict__ pall,
                                                      int* __restrict__ plabel, unsigned long long* __restrict__ hk,
                                                      int* __restrict__ hv, int bits) {
    const unsigned long long mask = (1ull << bits) - 1ull;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (long long)gridDim.x * blockDim.x) {
        if (!act[v]) continue;
        const int p = (int)vpos[v];
        const int c = vcell[v], s = vsolid[v];
        pcell[p] = c;
        psolid[p] = s;
        pall[p] = all[v];
        plabel[p] = p;
        const unsigned long long key = poly_key(s, c);
        for (unsigned long long h = poly_hash(key, bits);; h = (h + 1) & mask) {
            const unsigned long long old = atomicCAS(&hk[h], kPolyEmpty, key);
            if (old == kPolyEmpty) {
                hv[h] = p;
                break;
            }
        }
    }
}

__device__ inline int poly_find(const int* L, int x) {
    int p = L[x];
    while (p != x) {
        x = p;
        p = L[x];
    }
    return x;
}

// Union of active pairs of one solid across each internal face (each face once:
// from the lower cell).
__global__ void __launch_bounds__(256) k_poly_hook(PolyDev P, long long np, const int* __restrict__ pcell,
                                                   const int* __restrict__ psolid, int* __restrict__ L,
                                                   const unsigned long long* __restrict__ hk, const int* __restrict__ hv,
                                                   int bits) {
    auto unite = [&](int x, int y) {
        int a = poly_find(L, x), b = poly_find(L, y);
        while (a != b) {
            const int hi = a > b ? a : b, lo = a > b ? b : a;
            const int old = atomicMin(&L[hi], lo);
            if (old == hi) break;  // hi was a root and now points at lo
            a = poly_find(L, old);
            b = poly_find(L, lo);
        }
    };
    const unsigned long long mask = (1ull << bits) - 1ull;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (long long)gridDim.x * blockDim.x) {
        const int c = pcell[p], s = psolid[p];
        const int n0 = P.c_noff[c], nn = P.c_noff[c + 1] - n0;
        if (nn <= 8) {
            // up to eight neighbours: their labels, then the first probes of their hash slots,
            // each in flight together; a probe that neither hits nor finds an empty slot goes on
            int d[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) d[t] = t < nn ? P.c_nbs[n0 + t] : -1;
            unsigned long long key[8], h[8], k0[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                key[t] = poly_key(s, d[t] > c ? d[t] : 0);
                h[t] = poly_hash(key[t], bits);
                k0[t] = d[t] > c ? hk[h[t]] : kPolyEmpty;
            }
            int q[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) q[t] = k0[t] == key[t] && d[t] > c ? hv[h[t]] : -1;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                if (d[t] <= c || k0[t] == kPolyEmpty) continue;
                if (k0[t] != key[t]) {  // a collision: the rest of the probe sequence
                    q[t] = -1;
                    for (unsigned long long hh = (h[t] + 1) & mask;; hh = (hh + 1) & mask) {
                        const unsigned long long k = hk[hh];
                        if (k == key[t]) {
                            q[t] = hv[hh];
                            break;
                        }
                        if (k == kPolyEmpty) break;
                    }
                }
                if (q[t] >= 0) unite((int)p, q[t]);
            }
            continue;
        }
        for (int t = P.c_noff[c]; t < P.c_noff[c + 1]; ++t) {
            const int d = P.c_nbs[t];
            if (d < c) continue;
            const int q = poly_lookup(hk, hv, bits, s, d);
            if (q < 0) continue;
            int a = poly_find(L, (int)p), b = poly_find(L, q);
            while (a != b) {
                const int hi = a > b ? a : b, lo = a > b ? b : a;
                const int old = atomicMin(&L[hi], lo);
                if (old == hi) break;  // hi was a root and now points at lo
                a = poly_find(L, old);
                b = poly_find(L, lo);
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_poly_compress(long long np, int* __restrict__ L) {
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (long long)gridDim.x * blockDim.x)
        L[p] = poly_find(L, (int)p);
}

// Apex of a point list (geometrictools.cpp:25-45): A = the first point, B = the
// first later one whose distance has the other sign (or is zero), else the last.
// Coordinates relative to w0.
template <class R>
__device__ inline r3<R> poly_apex(const PolyDev& P, const SolidK& K, const int* lab, int n, d3 w0, R* phiA_out) {
    const d3 a = poly_point(P, lab[0]);
    const R fA = sdf_fast<R>(K, a);
    r3<R> B{R(0), R(0), R(0)};
    R fB = R(0);
    for (int t = 1; t < n; ++t) {
        const d3 b = poly_point(P, lab[t]);
        B = {R(b.x - w0.x), R(b.y - w0.y), R(b.z - w0.z)};
        fB = sdf_fast<R>(K, b);
        if (fA * fB <= 0) break;
    }
    if (phiA_out) *phiA_out = fA;
    return apex_point(r3<R>{R(a.x - w0.x), R(a.y - w0.y), R(a.z - w0.z)}, fA, B, fB);
}

// Solid fraction of a general polyhedral cell (calcCellVolume / V,
// geometrictools.cpp:47-72, with calcFaceAreaFraction and calcFaceArea,
// geometrictools.cpp:74-116): sum over the cell's faces of (1/3) eps_f
// |(apex - fc).Sf|.  two_d: the cell apex's world z is 0.
template <class R>
__device__ inline R poly_fraction(const PolyDev& P, const SolidK& K, int c) {
    const int* cp = P.c_pts + P.c_poff[c];
    const int ncp = P.c_poff[c + 1] - P.c_poff[c];
    const d3 w0 = poly_point(P, cp[0]);
    r3<R> ap = poly_apex<R>(P, K, cp, ncp, w0, nullptr);
    if (P.two_d) ap.z = R(0.0 - w0.z);
    R vol = R(0);
    for (int t = P.c_foff[c]; t < P.c_foff[c + 1]; ++t) {
        const int f = P.c_faces[t];
        const int* fp = P.f_pts + P.f_off[f];
        const int nf = P.f_off[f + 1] - P.f_off[f];
        int ss = 0;
        for (int u = 0; u < nf; ++u) ss += sdf_fast<R>(K, poly_point(P, fp[u])) > 0 ? 1 : -1;
        R eps;
        if (ss == nf) {
            eps = R(0);
        } else if (ss == -nf) {
            eps = R(1);
        } else {
            const r3<R> fa = poly_apex<R>(P, K, fp, nf, w0, nullptr);
            R area = R(0);
            for (int u = 0; u < nf; ++u) {
                const d3 o = poly_point(P, fp[u]), a = poly_point(P, fp[(u + 1) % nf]);
                const R po = sdf_fast<R>(K, o), pa = sdf_fast<R>(K, a);
                const r3<R> O{R(o.x - w0.x), R(o.y - w0.y), R(o.z - w0.z)};
                const r3<R> A{R(a.x - w0.x), R(a.y - w0.y), R(a.z - w0.z)};
                area += fabs(R(0.5) * rmag(rcross(sub(A, O), sub(fa, O)))) * line_fraction<R>(po, pa);
            }
            const r3<R> S{R(P.sf[3 * f]), R(P.sf[3 * f + 1]), R(P.sf[3 * f + 2])};
            eps = area / rmag(S);
        }
        if (eps != R(0)) {
            const r3<R> fcr{R(P.fc[3 * f] - w0.x), R(P.fc[3 * f + 1] - w0.y), R(P.fc[3 * f + 2] - w0.z)};
            const r3<R> S{R(P.sf[3 * f]), R(P.sf[3 * f + 1]), R(P.sf[3 * f + 2])};
            vol += (R(1) / R(3)) * eps * fabs(rdot(sub(ap, fcr), S));
        }
    }
    return vol / R(P.vol[c]);
}

// poly_fraction for a cell of eight points (hexahedra, whatever its faces), the
// same arithmetic on values computed once: the eight points' coordinates
// relative to the first one and their signed distances, kept in this thread's
// LDS slots (element q at [q * 64 + lane]); a face point's slot is found by
// binary search (cellPoints are ascending labels).  The labels and points are
// loaded together, not one dependent load after another.
template <class R>
__device__ inline R poly_fraction8(const PolyDev& P, const SolidK& K, int c, R* qx, R* qy, R* qz, R* qf, int* ql,
                                   int lane) {
    const int* cp = P.c_pts + P.c_poff[c];
    int lab[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) lab[q] = cp[q];
    d3 pt[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) pt[q] = poly_point(P, lab[q]);
    const d3 w0 = pt[0];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int o = q * 64 + lane;
        ql[o] = lab[q];
        qx[o] = R(pt[q].x - w0.x);
        qy[o] = R(pt[q].y - w0.y);
        qz[o] = R(pt[q].z - w0.z);
        qf[o] = sdf_fast<R>(K, pt[q]);
    }
    auto rel = [&](int q) { return r3<R>{qx[q * 64 + lane], qy[q * 64 + lane], qz[q * 64 + lane]}; };
    auto slot = [&](int label) {  // its position among the eight ascending labels
        int q = 0;
#pragma unroll
        for (int st = 4; st >= 1; st >>= 1)
            if (ql[(q + st) * 64 + lane] <= label) q += st;
        return q;
    };
    // the cell apex over cellPoints: A the first point, B the first later one across the surface
    const R fA = qf[lane];
    int b = 7;
#pragma unroll
    for (int q = 1; q < 8; ++q)
        if (b == 7 && fA * qf[q * 64 + lane] <= 0) b = q;
    r3<R> ap = apex_point(rel(0), fA, rel(b), qf[b * 64 + lane]);
    if (P.two_d) ap.z = R(0.0 - w0.z);
    R vol = R(0);
    for (int t = P.c_foff[c]; t < P.c_foff[c + 1]; ++t) {
        const int f = P.c_faces[t];
        const int* fp = P.f_pts + P.f_off[f];
        const int nf = P.f_off[f + 1] - P.f_off[f];
        int ss = 0;
        for (int u = 0; u < nf; ++u) ss += qf[slot(fp[u]) * 64 + lane] > 0 ? 1 : -1;
        R eps;
        if (ss == nf) {
            eps = R(0);
        } else if (ss == -nf) {
            eps = R(1);
        } else {
            // the face apex over the face's own point order
            const int s0 = slot(fp[0]);
            const R gA = qf[s0 * 64 + lane];
            int sb = s0;
            for (int u = 1; u < nf; ++u) {
                sb = slot(fp[u]);
                if (gA * qf[sb * 64 + lane] <= 0) break;
            }
            const r3<R> fa = apex_point(rel(s0), gA, rel(sb), qf[sb * 64 + lane]);
            R area = R(0);
            int so = s0;
            for (int u = 0; u < nf; ++u) {
                const int sa = slot(fp[(u + 1) % nf]);
                const r3<R> O = rel(so), A = rel(sa);
                area += fabs(R(0.5) * rmag(rcross(sub(A, O), sub(fa, O)))) *
                        line_fraction<R>(qf[so * 64 + lane], qf[sa * 64 + lane]);
                so = sa;
            }
            const r3<R> S{R(P.sf[3 * f]), R(P.sf[3 * f + 1]), R(P.sf[3 * f + 2])};
            eps = area / rmag(S);
        }
        if (eps != R(0)) {
            const r3<R> fcr{R(P.fc[3 * f] - w0.x), R(P.fc[3 * f + 1] - w0.y), R(P.fc[3 * f + 2] - w0.z)};
            const r3<R> S{R(P.sf[3 * f]), R(P.sf[3 * f + 1]), R(P.sf[3 * f + 2])};
            vol += (R(1) / R(3)) * eps * fabs(rdot(sub(ap, fcr), S));
        }
    }
    return vol / R(P.vol[c]);
}

// Per pair: member of its solid's seed component?  Then its code (+inf
// ALL_INSIDE; the fraction, negative when the centre is outside) and its force
// and torque contribution (FP64: f = alpha (uf - us), f V / dt, (cc - c) x f V / dt,
// solidcloud.cpp:384-422); else NaN and zeros.
template <class R>
__global__ void __launch_bounds__(64) k_poly_force(PolyDev P, const SolidK* __restrict__ Ks, long long np,
                                                    const int* __restrict__ pcell,
                                                    const int* __restrict__ psolid, const unsigned char* __restrict__ pall,
                                                    const int* __restrict__ L, const int* __restrict__ seed,
                                                    const int* __restrict__ first, const unsigned long long* __restrict__ hk,
                                                    const int* __restrict__ hv, int bits, const R* __restrict__ Ux,
                                                    const R* __restrict__ Uy, const R* __restrict__ Uz,
                                                    R* __restrict__ pcode, double* __restrict__ pf, double dt) {
    __shared__ R qx[8 * 64], qy[8 * 64], qz[8 * 64], qf[8 * 64];  // hexahedra: per thread its cell's points
    __shared__ int ql[8 * 64];
    const int lane = threadIdx.x;
    const double dtINV = 1.0 / dt;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (long long)gridDim.x * blockDim.x) {
        const int c = pcell[p], s = psolid[p];
        // the seed pair: the nearest cell when it is active, else the lowest active cell
        int sp = seed[s] >= 0 ? poly_lookup(hk, hv, bits, s, seed[s]) : -1;
        if (sp < 0) sp = poly_lookup(hk, hv, bits, s, first[s]);
        double* out = pf + 6 * p;
        if (sp < 0 || L[p] != L[sp]) {
            pcode[p] = code_none<R>();
            for (int t = 0; t < 6; ++t) out[t] = 0.0;
            continue;
        }
        const SolidK& K = Ks[s];
        double alpha = 1.0;
        if (pall[p]) {
            pcode[p] = R(__builtin_huge_val());
        } else {
            const R a = P.c_poff[c + 1] - P.c_poff[c] == 8 ? poly_fraction8<R>(P, K, c, qx, qy, qz, qf, ql, lane)
                                                            : poly_fraction<R>(P, K, c);
            const bool cin = inside_world(K, d3{P.cc[3 * c], P.cc[3 * c + 1], P.cc[3 * c + 2]});
            pcode[p] = cin ? a : -a;
            alpha = (double)a;
        }
        const d3 rv{P.cc[3 * c] - K.c[0], P.cc[3 * c + 1] - K.c[1], P.cc[3 * c + 2] - K.c[2]};
        const d3 us = d3{K.v[0], K.v[1], K.v[2]} + cross(d3{K.w[0], K.w[1], K.w[2]}, rv);
        const d3 uf{(double)Ux[c], (double)Uy[c], (double)Uz[c]};
        const d3 f = alpha * (uf - us);
        const d3 tq = cross(rv, f);
        const double V = P.vol[c];
        out[0] = (f.x * V) * dtINV;
        out[1] = (f.y * V) * dtINV;
        out[2] = (f.z * V) * dtINV;
        out[3] = (tq.x * V) * dtINV;
        out[4] = (tq.y * V) * dtINV;
        out[5] = (tq.z * V) * dtINV;
    }
}

// Pair chunks: a solid's pairs p0 <= p < p1 (p0 = the active scan at its first
// visit, p1 at the next solid's) in chunks of kPolyPairChunk; per solid its
// pair range and its number of chunks.
constexpr int kPolyPairChunk = 4096;
__global__ void k_poly_pair_chunks(int ns, const long long* __restrict__ voff, const long long* __restrict__ vpos,
                                   long long* __restrict__ prange, long long* __restrict__ nch) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const long long p0 = vpos[voff[s]], p1 = vpos[voff[s + 1]];
    prange[2 * s] = p0;
    prange[2 * s + 1] = p1;
    nch[s] = (p1 - p0 + kPolyPairChunk - 1) / kPolyPairChunk;
}

// One wave per pair chunk (chunks beyond the total exit): the chunk's
// contributions (lanes strided over its pairs, then a fixed shuffle tree) and
// its candidate counts, into part[8 * chunk] (six sums, ALL_INSIDE, cut).
template <class R>
__global__ void __launch_bounds__(256) k_poly_pair_part(long long maxch, const int* __restrict__ csol,
                                                        const long long* __restrict__ coff,
                                                        const long long* __restrict__ prange,
                                                        const R* __restrict__ pcode, const double* __restrict__ pf,
                                                        double* __restrict__ part, int ns) {
    const long long ch = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (ch >= maxch || ch >= coff[ns]) return;
    const int s = csol[ch];
    const long long p0 = prange[2 * s] + (ch - coff[s]) * kPolyPairChunk;
    const long long p1 = min(prange[2 * s + 1], p0 + kPolyPairChunk);
    double a[6] = {0, 0, 0, 0, 0, 0};
    int nall = 0, ncut = 0;
    for (long long p = p0 + lane; p < p1; p += 64) {
        const R v = pcode[p];
        if (v != v) continue;
        nall += isinf(v) ? 1 : 0;
        ncut += isinf(v) ? 0 : 1;
        for (int t = 0; t < 6; ++t) a[t] += pf[6 * p + t];
    }
    for (int off = 32; off >= 1; off >>= 1) {
        for (int t = 0; t < 6; ++t) a[t] += __shfl_xor(a[t], off, 64);
        nall += __shfl_xor(nall, off, 64);
        ncut += __shfl_xor(ncut, off, 64);
    }
    if (lane == 0) {
        for (int t = 0; t < 6; ++t) part[8 * ch + t] = a[t];
        part[8 * ch + 6] = (double)nall;
        part[8 * ch + 7] = (double)ncut;
    }
}

// Per solid (one wave): force and torque = the sum of its chunks' partial sums
// (lanes strided over the chunks in order, then a fixed shuffle tree) x rho_f;
// its candidate counts (ALL_INSIDE, cut).
__global__ void __launch_bounds__(256) k_poly_reduce(int ns, const long long* __restrict__ coff,
                                                     const double* __restrict__ part, double rho_f,
                                                     double* __restrict__ ft, int2* __restrict__ ncand) {
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= ns) return;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long long ch = coff[s] + lane; ch < coff[s + 1]; ch += 64)
        for (int t = 0; t < 8; ++t) a[t] += part[8 * ch + t];
    for (int off = 32; off >= 1; off >>= 1)
        for (int t = 0; t < 8; ++t) a[t] += __shfl_xor(a[t], off, 64);
    if (lane == 0) {
        for (int t = 0; t < 6; ++t) ft[6 * s + t] = a[t] * rho_f;
        ncand[s] = make_int2((int)a[6], (int)a[7]);
    }
}

// Member pairs per cell: counted (ccnt, zeroed before), then listed at the
// cell's scanned offset (any order: the field pass sorts them).
template <class R>
__global__ void __launch_bounds__(256) k_poly_cells(long long np, const int* __restrict__ pcell,
                                                    const R* __restrict__ pcode, int* __restrict__ ccnt,
                                                    const long long* __restrict__ coff, int* __restrict__ clist) {
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (long long)gridDim.x * blockDim.x) {
        const R v = pcode[p];
        if (v != v) continue;
        const int c = pcell[p];
        const int k = atomicAdd(&ccnt[c], 1);
        if (coff) clist[coff[c] + k] = (int)p;
    }
}

// Every cell once: As (clipped at 1), Fs, Ts, Ct over its member pairs in pair
// order (= solid index order; the last solid's class wins Ct): the forcing of
// solidcloud.cpp:384-422 per (solid, cell).
template <class R>
__global__ void __launch_bounds__(256) k_poly_fields(PolyDev P, const SolidK* __restrict__ Ks,
                                                     const long long* __restrict__ coff, int* __restrict__ clist,
                                                     const int* __restrict__ psolid, const R* __restrict__ pcode,
                                                     const R* __restrict__ Ux, const R* __restrict__ Uy,
                                                     const R* __restrict__ Uz, R* __restrict__ As, R* __restrict__ Fx,
                                                     R* __restrict__ Fy, R* __restrict__ Fz, R* __restrict__ Tsf,
                                                     int* __restrict__ Ct, double dt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.n_cells) return;
    double sA = 0.0, sT = 0.0, sF[3] = {0.0, 0.0, 0.0};
    int ct = 0;
    {
        int* l = clist + coff[c];
        const int n = (int)(coff[c + 1] - coff[c]);
        for (int a = 1; a < n; ++a) {  // insertion sort by pair index (short lists)
            const int x = l[a];
            int b = a - 1;
            while (b >= 0 && l[b] > x) {
                l[b + 1] = l[b];
                --b;
            }
            l[b + 1] = x;
        }
        const d3 cc{P.cc[3 * c], P.cc[3 * c + 1], P.cc[3 * c + 2]};
        const d3 uf{(double)Ux[c], (double)Uy[c], (double)Uz[c]};
        const double dtINV = 1.0 / dt;
        for (int a = 0; a < n; ++a) {
            const int p = l[a];
            const int s = psolid[p];
            const R v = pcode[p];
            const bool all_in = isinf(v);
            const double alpha = all_in ? 1.0 : (double)fabs(v);
            const SolidK& K = Ks[s];
            const d3 r{cc.x - K.c[0], cc.y - K.c[1], cc.z - K.c[2]};
            const d3 us = d3{K.v[0], K.v[1], K.v[2]} + cross(d3{K.w[0], K.w[1], K.w[2]}, r);
            const d3 f = alpha * (uf - us);
            sA += alpha;
            sT += alpha;
            sF[0] += f.x * dtINV;
            sF[1] += f.y * dtINV;
            sF[2] += f.z * dtINV;
            ct = all_in ? s + 4 : (signbit(v) ? 3 : 2);
        }
    }
    As[c] = R(sA < 1.0 ? sA : 1.0);
    Fx[c] = R(sF[0]);
    Fy[c] = R(sF[1]);
    Fz[c] = R(sF[2]);
    Tsf[c] = R(sT);
    Ct[c] = ct;
}

// U = us (the solid's rigid velocity at the cell centre) where Ct >= 4: cells
// wholly inside solid Ct - 4 (SolidCloud::fixInternal, solidcloud.cpp:288-301).
template <class R>
__global__ void __launch_bounds__(256) k_poly_fix(PolyDev P, const int* __restrict__ Ct,
                                                  const sdfibm_solid* __restrict__ sol, int ns, R* __restrict__ Ux,
                                                  R* __restrict__ Uy, R* __restrict__ Uz) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.n_cells) return;
    const int t = Ct[c];
    if (t < 4 || t - 4 >= ns) return;
    const sdfibm_solid& S = sol[t - 4];
    const d3 r{P.cc[3 * c] - S.center[0], P.cc[3 * c + 1] - S.center[1], P.cc[3 * c + 2] - S.center[2]};
    const d3 us = d3{S.vel[0], S.vel[1], S.vel[2]} + cross(d3{S.omega[0], S.omega[1], S.omega[2]}, r);
    Ux[c] = R(us.x);
    Uy[c] = R(us.y);
    Uz[c] = R(us.z);
}
