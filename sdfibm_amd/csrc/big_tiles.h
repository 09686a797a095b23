// big_tiles.h -- the tile pass for solids whose cell tile is too large for the
// LDS bit rows of k_classify (more than 63 cells along x or more than 4096
// vertex rows), e.g. the 40-80-cell CSG solids of a polydisperse cloud.
//
// Included by interact.hip, inside its kernel namespace, after the helpers it
// uses (divmod_small, block_exscan2, block_reduce, uf_find, code_none,
// cell_force).  The results are those of the
// reference's enumeration (src/cellenumerator.cpp:6-78) for these solids, as
// for every other solid: the cells with >= 1 strictly inside vertex, connected
// through faces to the seed (the nearest cell if it has an inside vertex, else
// the first such cell in index order), ALL_INSIDE when all 8 are inside.  The
// outputs go where k_classify puts them -- tile codes at the tile's offset, the
// candidate list (ALL_INSIDE cells ascending, then the cut cells ascending) in
// the list buffer at the same offset, the two counts -- so k_fields reads these
// tiles like any other.  The list holds tile-local linear indices here (the
// small tiles' packed (i, j, k) does not fit 80-cell tiles).
//
// All big solids of a step are processed together, over flat index spaces
// (their tile vertices, their tile cells, their candidate chunks) laid out by
// one scan kernel; no host round trip, so the step stays capturable.
//   k_big_layout    offsets of each big solid's vertex bits, cells, totals
//   k_big_vertices  strict inside test per tile vertex (FP64) -> bit words
//   k_big_cells     cells with >= 1 inside vertex -> union-find labels (in the
//                   candidate-list buffer), lowest such cell
//   k_big_hook      union over +x, +y, +z neighbours of active cells
//   k_big_compress  labels -> roots
//   k_big_member    codes: NaN outside the seed's component, +inf ALL_INSIDE,
//                   0 cut (placeholder); a second look at every edge checks the
//                   union converged (EF_LABEL otherwise)
//   k_big_list      candidate lists and counts, one workgroup per solid
//   k_big_chunks    chunks of candidates (the chunk size is a run-time choice), their offsets
//   k_big_force     fractions, codes and forcing per chunk; FP64 partial sums
//   k_big_ft        partials summed in chunk order (x rho_f) -> force/torque
#pragma once

constexpr int kBigMaxSolids = 4096;  // big tiles per step (SDFIBM_ERR_CAPACITY beyond)
constexpr int kBigChunk = 1024;      // default candidate entries per chunk of k_big_force (the autotuner's choice)

struct BigMeta {
    int solid;
    int first;       // lowest tile-local index of a cell with an inside vertex (INT_MAX: none)
    long long woff;  // first 64-bit word of its vertex bits
    long long vbase; // first flat vertex index (multiple of 64: one solid per bit word)
    long long cbase; // first flat cell index
    int chunk0;      // first chunk of its candidate entries
    int nchunk;
};

// Totals of the step's big tiles (BigTot fields as long long):
// [0] big solids, [1] flat vertices (padded), [2] flat cells, [3] bit words, [4] chunks,
// [5] bit words needed (for the host's buffer growth), [6] chunks needed
constexpr int kBigTotN = 8;

// b with base[b] <= t < base[b + 1] (base ascending, n entries, base[0] = 0)
template <class F>
__device__ inline int big_find(int n, long long t, F base) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base(mid) <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Layout of the step's big tiles: one workgroup of 1024 threads, up to 4 big
// solids per thread, exclusive scans of vertex words, cells.
__global__ void __launch_bounds__(1024) k_big_layout(int ns, const TileInfo* __restrict__ Ts,
                                                     const int* __restrict__ qmeta, const int* __restrict__ qlist,
                                                     BigMeta* __restrict__ meta, long long* __restrict__ tot,
                                                     long long word_cap, int* __restrict__ err) {
    __shared__ long long part[2][1024];
    const int tid = threadIdx.x;
    int nbig = qmeta[4];
    if (nbig > kBigMaxSolids) {
        if (tid == 0) atomicOr(err, EF_CODECAP);
        nbig = 0;
    }
    long long w[4], c[4];
    long long sw = 0, sc = 0;
    for (int u = 0; u < 4; ++u) {
        const int b = tid * 4 + u;
        w[u] = c[u] = 0;
        if (b < nbig) {
            const TileInfo T = Ts[qlist[4LL * ns + b]];
            const long long V = (long long)(T.n[0] + 1) * (T.n[1] + 1) * (T.n[2] + 1);
            w[u] = (V + 63) / 64;
            c[u] = (long long)T.n[0] * T.n[1] * T.n[2];
        }
        sw += w[u];
        sc += c[u];
    }
    part[0][tid] = sw;
    part[1][tid] = sc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan of the thread sums
        const long long a = tid >= off ? part[0][tid - off] : 0, b2 = tid >= off ? part[1][tid - off] : 0;
        __syncthreads();
        part[0][tid] += a;
        part[1][tid] += b2;
        __syncthreads();
    }
    long long ow = part[0][tid] - sw, oc = part[1][tid] - sc;
    for (int u = 0; u < 4; ++u) {
        const int b = tid * 4 + u;
        if (b < nbig) {
            BigMeta M;
            M.solid = qlist[4LL * ns + b];
            M.first = 0x7fffffff;
            M.woff = ow;
            M.vbase = ow * 64;
            M.cbase = oc;
            M.chunk0 = 0;
            M.nchunk = 0;
            meta[b] = M;
        }
        ow += w[u];
        oc += c[u];
    }
    if (tid == 1023) {
        const long long words = part[0][1023], cells = part[1][1023];
        const bool fits = words <= word_cap;
        if (!fits) atomicOr(err, EF_CODECAP);
        tot[0] = fits ? nbig : 0;
        tot[1] = fits ? words * 64 : 0;
        tot[2] = fits ? cells : 0;
        tot[3] = words;
        tot[5] = words;
    }
}

// Strict inside test at every big-tile vertex (FP64, the reference's operation
// order, as k_classify); a wave covers one 64-bit word of one solid.
template <int MK>
__global__ void __launch_bounds__(256) k_big_vertices(MeshDev m, const SolidK* __restrict__ Ks,
                                                      const TileInfo* __restrict__ Ts, const BigMeta* __restrict__ meta,
                                                      const long long* __restrict__ tot,
                                                      unsigned long long* __restrict__ vbits) {
    const int nbig = (int)tot[0];
    const long long nv = tot[1];
    const int lane = threadIdx.x & 63;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t - lane < nv;
         t += (long long)gridDim.x * blockDim.x) {
        const int b = big_find(nbig, t, [&](int x) { return meta[x].vbase; });  // wave-uniform
        const BigMeta M = meta[b];
        const TileInfo T = Ts[M.solid];
        const int VX = T.n[0] + 1, VY = T.n[1] + 1;
        const long long v = t - M.vbase;
        bool in = false;
        if (v < (long long)VX * VY * (T.n[2] + 1)) {
            int r, vi, vj, vk;
            divmod_small((int)v, VX, 1.0f / (float)VX, r, vi);
            divmod_small(r, VY, 1.0f / (float)VY, vk, vj);
            in = inside_world(Ks[M.solid], vertex_pos<MK>(m, T.lo[0] + vi, T.lo[1] + vj, T.lo[2] + vk));
        }
        const unsigned long long bal = __ballot(in);
        if (lane == 0) vbits[M.woff + v / 64] = bal;
    }
}

__device__ inline bool big_vbit(const unsigned long long* w, long long v) { return (w[v >> 6] >> (v & 63)) & 1ull; }

// Inside-vertex count of tile cell (i, j, k) from the vertex bits.
__device__ inline int big_nin(const unsigned long long* w, int i, int j, int k, int VX, int VY) {
    const long long v0 = i + (long long)VX * (j + (long long)VY * k);
    const long long dy = VX, dz = (long long)VX * VY;
    return (int)big_vbit(w, v0) + (int)big_vbit(w, v0 + 1) + (int)big_vbit(w, v0 + dy) + (int)big_vbit(w, v0 + dy + 1) +
           (int)big_vbit(w, v0 + dz) + (int)big_vbit(w, v0 + dz + 1) + (int)big_vbit(w, v0 + dz + dy) +
           (int)big_vbit(w, v0 + dz + dy + 1);
}

// Iterates the flat cells of the step's big tiles: f(meta, tile, q, i, j, k).
template <class F>
__device__ inline void big_cells_loop(const BigMeta* meta, const TileInfo* Ts, const long long* tot, long long code_cap,
                                      F f) {
    const int nbig = (int)tot[0];
    const long long nc = tot[2];
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nc; t += (long long)gridDim.x * blockDim.x) {
        const int b = big_find(nbig, t, [&](int x) { return meta[x].cbase; });
        const BigMeta& M = meta[b];
        const TileInfo T = Ts[M.solid];
        if (T.off + (long long)T.n[0] * T.n[1] * T.n[2] > code_cap) continue;  // flagged by the tile offsets
        const int q = (int)(t - M.cbase);
        int r, i, j, k;
        divmod_small(q, T.n[0], 1.0f / (float)T.n[0], r, i);
        divmod_small(r, T.n[1], 1.0f / (float)T.n[1], k, j);
        f(b, M, T, q, i, j, k);
    }
}

// Cells with >= 1 inside vertex: label = own index, else -1; the lowest one.
__global__ void __launch_bounds__(256) k_big_cells(const TileInfo* __restrict__ Ts, BigMeta* __restrict__ meta,
                                                   const long long* __restrict__ tot,
                                                   const unsigned long long* __restrict__ vbits, int* __restrict__ cand,
                                                   long long code_cap) {
    big_cells_loop(meta, Ts, tot, code_cap, [&](int b, const BigMeta& M, const TileInfo& T, int q, int i, int j, int k) {
        const int nin = big_nin(vbits + M.woff, i, j, k, T.n[0] + 1, T.n[1] + 1);
        cand[T.off + q] = nin > 0 ? q : -1;
        // the lowest active cell: the grid walks the cells in ascending order, so after the
        // first few waves of a solid the plain read already rules out almost every atomic
        if (nin > 0 && q < __hip_atomic_load(&meta[b].first, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&meta[b].first, q);
    });
}

// Union over the +x, +y, +z neighbours (hook the larger root under the smaller:
// the same concurrent union as the planes' labelling).
__global__ void __launch_bounds__(256) k_big_hook(const TileInfo* __restrict__ Ts, const BigMeta* __restrict__ meta,
                                                  const long long* __restrict__ tot, int* __restrict__ cand,
                                                  long long code_cap) {
    big_cells_loop(meta, Ts, tot, code_cap, [&](int, const BigMeta&, const TileInfo& T, int q, int i, int j, int k) {
        int* L = cand + T.off;
        if (L[q] < 0) return;
        const int nx = T.n[0], nxy = T.n[0] * T.n[1];
        int nb[3], nn = 0;
        if (i + 1 < T.n[0]) nb[nn++] = q + 1;
        if (j + 1 < T.n[1]) nb[nn++] = q + nx;
        if (k + 1 < T.n[2]) nb[nn++] = q + nxy;
        for (int u = 0; u < nn; ++u) {
            if (L[nb[u]] < 0) continue;
            int a = uf_find(L, q), b2 = uf_find(L, nb[u]);
            while (a != b2) {
                const int hi = a > b2 ? a : b2, lo = a > b2 ? b2 : a;
                const int old = atomicMin(&L[hi], lo);
                if (old == hi) break;  // hi was a root and now points at lo
                a = uf_find(L, old);
                b2 = uf_find(L, lo);
            }
        }
    });
}

__global__ void __launch_bounds__(256) k_big_compress(const TileInfo* __restrict__ Ts, const BigMeta* __restrict__ meta,
                                                      const long long* __restrict__ tot, int* __restrict__ cand,
                                                      long long code_cap) {
    big_cells_loop(meta, Ts, tot, code_cap, [&](int, const BigMeta&, const TileInfo& T, int q, int, int, int) {
        int* L = cand + T.off;
        if (L[q] >= 0) L[q] = uf_find(L, q);
    });
}

// Codes of the seed's component; every active cell's label must equal its active
// neighbours' (a union that did not converge raises EF_LABEL).
template <class R>
__global__ void __launch_bounds__(256) k_big_member(const MeshDev m, const TileInfo* __restrict__ Ts,
                                                    const BigMeta* __restrict__ meta, const long long* __restrict__ tot,
                                                    const unsigned long long* __restrict__ vbits,
                                                    const int* __restrict__ cand, R* __restrict__ codes,
                                                    long long code_cap, int* __restrict__ err) {
    big_cells_loop(meta, Ts, tot, code_cap, [&](int, const BigMeta& M, const TileInfo& T, int q, int i, int j, int k) {
        const int* L = cand + T.off;
        R* tc = codes + T.off;
        const int lq = L[q];
        if (lq < 0) {
            tc[q] = code_none<R>();
            return;
        }
        int seed = -1;
        if (T.seed >= 0) {
            const int si = T.seed % m.nx, sj = (T.seed / m.nx) % m.ny, sk = T.seed / (m.nx * m.ny);
            const int li = si - T.lo[0], lj = sj - T.lo[1], lk = sk - T.lo[2];
            if (li >= 0 && li < T.n[0] && lj >= 0 && lj < T.n[1] && lk >= 0 && lk < T.n[2]) {
                const int sq = li + T.n[0] * (lj + T.n[1] * lk);
                if (L[sq] >= 0) seed = sq;
            }
        }
        if (seed < 0) seed = M.first;  // an active cell exists: this one
        const bool member = L[seed] == lq;
        const int nx = T.n[0], nxy = T.n[0] * T.n[1];
        const bool bad = (i + 1 < T.n[0] && L[q + 1] >= 0 && L[q + 1] != lq) ||
                         (j + 1 < T.n[1] && L[q + nx] >= 0 && L[q + nx] != lq) ||
                         (k + 1 < T.n[2] && L[q + nxy] >= 0 && L[q + nxy] != lq);
        if (bad) atomicOr(err, EF_LABEL);
        if (!member) {
            tc[q] = code_none<R>();
        } else {
            const int nin = big_nin(vbits + M.woff, i, j, k, T.n[0] + 1, T.n[1] + 1);
            tc[q] = nin == 8 ? R(__builtin_huge_val()) : R(0);
        }
    });
}

// Candidate lists of the big tiles: one workgroup per solid; each thread a
// contiguous run of the tile's cells; ALL_INSIDE cells first, then cut cells,
// each ascending (tile-local linear indices).
template <class R>
__global__ void __launch_bounds__(1024) k_big_list(const TileInfo* __restrict__ Ts, BigMeta* __restrict__ meta,
                                                   const long long* __restrict__ tot, const R* __restrict__ codes,
                                                   int* __restrict__ cand, int2* __restrict__ ncand,
                                                   long long code_cap, int chunk_entries) {
    __shared__ int scan_lds[32];
    const int nbig = (int)tot[0];
    const int tid = threadIdx.x;
    for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
        const int s = meta[b].solid;
        const TileInfo T = Ts[s];
        const int C = T.n[0] * T.n[1] * T.n[2];
        if (T.off + (long long)C > code_cap) continue;
        const R* tc = codes + T.off;
        const int chunk = (C + 1023) / 1024;
        const int q0 = min(tid * chunk, C), q1 = min(q0 + chunk, C);
        int nin = 0, ncut = 0;
        for (int q = q0; q < q1; ++q) {
            const R v = tc[q];
            nin += (v == R(__builtin_huge_val())) ? 1 : 0;
            ncut += (v == R(0)) ? 1 : 0;
        }
        int pi, pc, tin, tcut;
        block_exscan2<1024>(nin, ncut, pi, pc, tin, tcut, scan_lds);
        pc += tin;
        int* tl = cand + T.off;
        for (int q = q0; q < q1; ++q) {
            const R v = tc[q];
            if (v == R(__builtin_huge_val()))
                tl[pi++] = q;
            else if (v == R(0))
                tl[pc++] = q;
        }
        if (tid == 0) {
            ncand[s] = make_int2(tin, tcut);
            meta[b].nchunk = (tin + tcut + chunk_entries - 1) / chunk_entries;
        }
        __syncthreads();
    }
}

// Chunk offsets (one workgroup, exclusive scan over the big solids).
__global__ void __launch_bounds__(1024) k_big_chunks(BigMeta* __restrict__ meta, long long* __restrict__ tot,
                                                     long long chunk_cap, int* __restrict__ err) {
    __shared__ int part[1024];
    const int nbig = (int)tot[0];
    const int tid = threadIdx.x;
    int c[4], sum = 0;
    for (int u = 0; u < 4; ++u) {
        const int b = tid * 4 + u;
        c[u] = b < nbig ? meta[b].nchunk : 0;
        sum += c[u];
    }
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int a = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += a;
        __syncthreads();
    }
    int o = part[tid] - sum;
    for (int u = 0; u < 4; ++u) {
        const int b = tid * 4 + u;
        if (b < nbig) meta[b].chunk0 = o;
        o += c[u];
    }
    if (tid == 1023) {
        const long long total = part[1023];
        tot[6] = total;
        if (total > chunk_cap) {
            atomicOr(err, EF_CODECAP);
            tot[4] = 0;
        } else {
            tot[4] = total;
        }
    }
}

// Fractions, codes and forcing of one chunk of a big solid's candidates per
// workgroup iteration; FP64 partial force/torque per chunk.
template <class R, int MK, int NT>
__global__ void __launch_bounds__(NT) k_big_force(MeshDev m, const SolidK* __restrict__ Ks,
                                                  const TileInfo* __restrict__ Ts, const BigMeta* __restrict__ meta,
                                                  const long long* __restrict__ tot, const int* __restrict__ cand,
                                                  const int2* __restrict__ ncand, const R* __restrict__ Ux,
                                                  const R* __restrict__ Uy, const R* __restrict__ Uz,
                                                  R* __restrict__ codes, long long code_cap, double dt,
                                                  double* __restrict__ part, int chunk_entries) {
    __shared__ double red[6 * 16];
    const int nbig = (int)tot[0];
    const long long nch = tot[4];
    const double dtINV = 1.0 / dt;
    const int tid = threadIdx.x;
    for (long long ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const int b = big_find(nbig, ch, [&](int x) { return (long long)meta[x].chunk0; });
        const int s = meta[b].solid;
        const TileInfo T = Ts[s];
        if (T.off + (long long)T.n[0] * T.n[1] * T.n[2] > code_cap) continue;  // flagged; uniform per workgroup
        const SolidK K = Ks[s];
        const int2 cnt = ncand[s];
        const int nin = cnt.x, ntot = cnt.x + cnt.y;
        const int e0 = (int)(ch - meta[b].chunk0) * chunk_entries, e1 = min(e0 + chunk_entries, ntot);
        const int* tl = cand + T.off;
        R* tcodes = codes + T.off;
        const float rn0 = 1.0f / (float)T.n[0], rn1 = 1.0f / (float)T.n[1];
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int base = e0; base < e1; base += NT) {
            const int e = base + tid;
            if (e >= e1) continue;
            const int q = tl[e];
            int r, i, j, k;
            divmod_small(q, T.n[0], rn0, r, i);
            divmod_small(r, T.n[1], rn1, k, j);
            cell_force<R, MK>(m, K, T.lo[0] + i, T.lo[1] + j, T.lo[2] + k, e >= nin, Ux, Uy, Uz, tcodes + q, dtINV, acc);
        }
        block_reduce<6>(acc, red);
        if (tid == 0)
            for (int t = 0; t < 6; ++t) part[6 * ch + t] = acc[t];
    }
}

// Force and torque of the big solids: chunk partials in chunk order, x rho_f.
__global__ void __launch_bounds__(256) k_big_ft(const BigMeta* __restrict__ meta, const long long* __restrict__ tot,
                                                const double* __restrict__ part, double rho_f,
                                                double* __restrict__ ft) {
    const int nbig = (int)tot[0];
    const bool ok = tot[4] > 0 || tot[6] == 0;  // chunk partials exist (or there are no candidates)
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nbig; b += gridDim.x * blockDim.x) {
        const BigMeta M = meta[b];
        double a[6] = {0, 0, 0, 0, 0, 0};
        if (ok)
            for (int c = 0; c < M.nchunk; ++c)
                for (int t = 0; t < 6; ++t) a[t] += part[6LL * (M.chunk0 + c) + t];
        for (int t = 0; t < 6; ++t) ft[6 * M.solid + t] = a[t] * rho_f;
    }
}
