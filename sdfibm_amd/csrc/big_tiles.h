// big_tiles.h -- the tile pass for solids whose cell tile is too large for the
// LDS bit rows of k_classify (more than 63 cells along x or more than 1024
// vertex rows), e.g. the 40-80-cell CSG solids of a polydisperse cloud: one
// workgroup per solid would make such a tile the step's long pole, so these
// tiles are spread over the whole chip.
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
// (their tile vertices; their tile cells in blocks of kBigCPB, one workgroup
// iteration each, so a workgroup finds its solid once; for the activity pass,
// their rows in units of up to 64 cells, a wave each; their candidate chunks)
// laid out by one scan kernel; no host round trip, so the step stays capturable.
//   k_big_layout    offsets of each big solid's vertex bits, cells, cell and unit blocks, totals
//   k_big_vertices  strict inside test per tile vertex (FP64) -> bit words
//   k_big_cells     cells with >= 1 inside vertex -> union-find labels (in the
//                   candidate-list buffer): the first cell of their run along x;
//                   lowest such cell
//   k_big_hook      union of the runs over +y, +z neighbours (once per overlap)
//   k_big_compress  labels -> roots
//   k_big_member    codes: NaN outside the seed's component, +inf ALL_INSIDE,
//                   0 cut (placeholder), and each cell block's two counts; a
//                   second look at every edge checks the union converged
//                   (EF_LABEL otherwise)
//   k_big_scan      per solid, the blocks' counts scanned in block order: list
//                   offsets, the solid's counts and candidate chunks
//   k_big_write     candidate lists (stable compaction of each block's cells)
//   k_big_chunks    chunks of candidates (the chunk size is a run-time choice), their offsets
//   k_big_force     fractions, codes and forcing per chunk; FP64 partial sums
//   k_big_ft        partials summed in a fixed order, a wave per solid (x rho_f)
//                   -> force/torque
#pragma once

constexpr int kBigPer = 16;                  // big solids per thread of the one-workgroup scans
constexpr int kBigMaxSolids = 1024 * kBigPer;  // big tiles per step (SDFIBM_ERR_CAPACITY beyond)
constexpr int kBigChunk = 1024;      // default candidate entries per chunk of k_big_force (the autotuner's choice)
constexpr int kBigCPB = 1024;        // tile cells per block of the flat-cell kernels (4 per thread of 256)
constexpr int kBigWPB = 32;          // vertex bit words per block of k_big_vertices (8 per wave of 4)
constexpr int kBigUPB = 16;          // units (up to 64 cells of one row) per block of k_big_cells

struct BigMeta {
    int solid;
    int first;       // lowest tile-local index of a cell with an inside vertex (INT_MAX: none)
    long long woff;  // first 64-bit word of its vertex bits
    long long vbase; // first flat vertex index (multiple of 64: one solid per bit word)
    long long cbase; // first flat cell index
    int chunk0;      // first chunk of its candidate entries
    int nchunk;
    int blk0;        // first cell block (kBigCPB tile cells each)
    int nblk;
    int vblk0;       // first vertex block (kBigWPB bit words each)
    int ublk0;       // first unit block (kBigUPB units each)
};

// Totals of the step's big tiles (BigTot fields as long long):
// [0] big solids, [1] flat vertices (padded), [2] flat cells, [3] bit words, [4] chunks,
// [5] bit words needed (for the host's buffer growth), [6] chunks needed, [7] cell blocks,
// [8] vertex blocks, [9] unit blocks
constexpr int kBigTotN = 10;

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

// Layout of the step's big tiles: one workgroup of 1024 threads, up to kBigPer
// big solids per thread, exclusive scans of vertex words, cells, cell blocks.
__global__ void __launch_bounds__(1024) k_big_layout(int ns, const TileInfo* __restrict__ Ts,
                                                     const int* __restrict__ qmeta, const int* __restrict__ qlist,
                                                     BigMeta* __restrict__ meta, long long* __restrict__ tot,
                                                     long long word_cap, int* __restrict__ err) {
    __shared__ long long part[5][1024];
    const int tid = threadIdx.x;
    int nbig = qmeta[4];
    if (nbig > kBigMaxSolids) {
        if (tid == 0) atomicOr(err, EF_CODECAP);
        nbig = 0;
    }
    auto totals = [&](long long words, long long cells, long long blocks, long long vblocks, long long ublocks) {
        // the block counts live in a buffer of words / 16 + kBigMaxSolids entries (a tile has at
        // least as many vertices as cells, so blocks <= words * 64 / kBigCPB + big solids)
        const bool fits = words + 1 <= word_cap && blocks <= word_cap / 16 + kBigMaxSolids;  // + the guard word
        if (!fits) atomicOr(err, EF_CODECAP);
        tot[0] = fits ? nbig : 0;
        tot[1] = fits ? words * 64 : 0;
        tot[2] = fits ? cells : 0;
        tot[3] = words;
        tot[5] = words;
        tot[7] = fits ? blocks : 0;
        tot[8] = fits ? vblocks : 0;
        tot[9] = fits ? ublocks : 0;
    };
    if (nbig <= 1024) {  // a solid per thread: wave scans by shuffles, the 16 wave totals through LDS
        const int lane = tid & 63, wv = tid >> 6;
        long long v[5] = {0, 0, 0, 0, 0}, un = 0;
        int solid = 0;
        if (tid < nbig) {
            solid = qlist[4LL * ns + tid];
            const TileInfo T = Ts[solid];
            const long long V = (long long)(T.n[0] + 1) * (T.n[1] + 1) * (T.n[2] + 1);
            v[0] = (V + 63) / 64;
            v[1] = (long long)T.n[0] * T.n[1] * T.n[2];
            un = (long long)T.n[1] * T.n[2] * ((T.n[0] + 63) >> 6);
            v[2] = (v[1] + kBigCPB - 1) / kBigCPB;
            v[3] = (v[0] + kBigWPB - 1) / kBigWPB;
            v[4] = (un + kBigUPB - 1) / kBigUPB;
        }
        long long x[5];
        for (int r = 0; r < 5; ++r) {
            x[r] = v[r];
            for (int d = 1; d < 64; d <<= 1) {
                const long long y = __shfl_up(x[r], d, 64);
                if (lane >= d) x[r] += y;
            }
            if (lane == 63) part[r][wv] = x[r];
        }
        __syncthreads();
        long long o[5], all[5];
        for (int r = 0; r < 5; ++r) {
            o[r] = x[r] - v[r];
            all[r] = 0;
            for (int w = 0; w < 16; ++w) {
                const long long t = part[r][w];
                if (w < wv) o[r] += t;
                all[r] += t;
            }
        }
        if (tid < nbig) {
            BigMeta M;
            M.solid = solid;
            M.first = 0x7fffffff;
            M.woff = o[0];
            M.vbase = o[0] * 64;
            M.cbase = o[1];
            M.chunk0 = 0;
            M.nchunk = 0;
            M.blk0 = (int)o[2];
            M.nblk = (int)v[2];
            M.vblk0 = (int)o[3];
            M.ublk0 = (int)o[4];
            meta[tid] = M;
        }
        if (tid == 0) totals(all[0], all[1], all[2], all[3], all[4]);
        return;
    }
    long long w[kBigPer], c[kBigPer], un[kBigPer];
    long long sw = 0, sc = 0, sk = 0, sv = 0, su = 0;
    for (int u = 0; u < kBigPer; ++u) {
        const int b = tid * kBigPer + u;
        w[u] = c[u] = un[u] = 0;
        if (b < nbig) {
            const TileInfo T = Ts[qlist[4LL * ns + b]];
            const long long V = (long long)(T.n[0] + 1) * (T.n[1] + 1) * (T.n[2] + 1);
            w[u] = (V + 63) / 64;
            c[u] = (long long)T.n[0] * T.n[1] * T.n[2];
            un[u] = (long long)T.n[1] * T.n[2] * ((T.n[0] + 63) >> 6);
        }
        sw += w[u];
        sc += c[u];
        sk += (c[u] + kBigCPB - 1) / kBigCPB;
        sv += (w[u] + kBigWPB - 1) / kBigWPB;
        su += (un[u] + kBigUPB - 1) / kBigUPB;
    }
    part[0][tid] = sw;
    part[1][tid] = sc;
    part[2][tid] = sk;
    part[3][tid] = sv;
    part[4][tid] = su;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan of the thread sums
        long long a[5];
        for (int r = 0; r < 5; ++r) a[r] = tid >= off ? part[r][tid - off] : 0;
        __syncthreads();
        for (int r = 0; r < 5; ++r) part[r][tid] += a[r];
        __syncthreads();
    }
    long long ow = part[0][tid] - sw, oc = part[1][tid] - sc, ok = part[2][tid] - sk, ov = part[3][tid] - sv,
              ou = part[4][tid] - su;
    for (int u = 0; u < kBigPer; ++u) {
        const int b = tid * kBigPer + u;
        if (b < nbig) {
            BigMeta M;
            M.solid = qlist[4LL * ns + b];
            M.first = 0x7fffffff;
            M.woff = ow;
            M.vbase = ow * 64;
            M.cbase = oc;
            M.chunk0 = 0;
            M.nchunk = 0;
            M.blk0 = (int)ok;
            M.nblk = (int)((c[u] + kBigCPB - 1) / kBigCPB);
            M.vblk0 = (int)ov;
            M.ublk0 = (int)ou;
            meta[b] = M;
        }
        ow += w[u];
        oc += c[u];
        ok += (c[u] + kBigCPB - 1) / kBigCPB;
        ov += (w[u] + kBigWPB - 1) / kBigWPB;
        ou += (un[u] + kBigUPB - 1) / kBigUPB;
    }
    if (tid == 1023) totals(part[0][1023], part[1][1023], part[2][1023], part[3][1023], part[4][1023]);
}

// A CSG solid's shape constants copied to LDS for one workgroup iteration (the
// whole workgroup; its previous readers are done first): read from there, they
// are not loaded again after each of the loop's stores, which the compiler
// cannot tell apart from the global table.
__device__ inline void big_csg_to_lds(SolidK& K, CsgK& sh, int tid, int nt) {
    if (K.shape != SDFIBM_SHAPE_CSG) return;  // uniform
    __syncthreads();
    const unsigned* src = reinterpret_cast<const unsigned*>(K.csg);
    unsigned* dst = reinterpret_cast<unsigned*>(&sh);
    for (int t = tid; t < (int)(sizeof(CsgK) / 4); t += nt) dst[t] = src[t];
    __syncthreads();
    K.csg = &sh;
}

// Strict inside test at every big-tile vertex (FP64, the reference's operation
// order, as k_classify): blocks of kBigWPB bit words of one solid per workgroup
// iteration (the solid found once), a wave per word (one ballot).
template <int MK>
__global__ void __launch_bounds__(256) k_big_vertices(MeshDev m, const SolidK* __restrict__ Ks,
                                                      const TileInfo* __restrict__ Ts, const BigMeta* __restrict__ meta,
                                                      const long long* __restrict__ tot,
                                                      unsigned long long* __restrict__ vbits) {
    __shared__ CsgK sh_csg;
    const int nbig = (int)tot[0];
    const long long nvb = tot[8];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (long long blk = blockIdx.x; blk < nvb; blk += gridDim.x) {
        const int b = big_find(nbig, blk, [&](int x) { return (long long)meta[x].vblk0; });  // uniform
        const BigMeta M = meta[b];
        const TileInfo T = Ts[M.solid];
        SolidK K = Ks[M.solid];
        big_csg_to_lds(K, sh_csg, threadIdx.x, 256);
        const int VX = T.n[0] + 1, VY = T.n[1] + 1;
        const long long V = (long long)VX * VY * (T.n[2] + 1), words = (V + 63) / 64;
        const float rvx = 1.0f / (float)VX, rvy = 1.0f / (float)VY;
        for (int r = 0; r < kBigWPB / 4; ++r) {
            const long long word = (blk - M.vblk0) * kBigWPB + 4 * r + wv;
            if (word >= words) break;  // wave-uniform
            const long long v = word * 64 + lane;
            bool in = false;
            if (v < V) {
                int q, vi, vj, vk;
                divmod_small((int)v, VX, rvx, q, vi);
                divmod_small(q, VY, rvy, vk, vj);
                in = inside_world(K, vertex_pos<MK>(m, T.lo[0] + vi, T.lo[1] + vj, T.lo[2] + vk));
            }
            const unsigned long long bal = __ballot(in);
            if (lane == 0) vbits[M.woff + word] = bal;
        }
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

// Iterates the cell blocks of the step's big tiles, one workgroup iteration per
// block (the block's solid found once, uniform): f(b, meta, tile, block, q0, q1)
// for the tile-local cells [q0, q1).
template <class F>
__device__ inline void big_blocks_loop(const BigMeta* meta, const TileInfo* Ts, const long long* tot, long long code_cap,
                                       F f) {
    const int nbig = (int)tot[0];
    const long long nblk = tot[7];
    for (long long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const int b = big_find(nbig, blk, [&](int x) { return (long long)meta[x].blk0; });
        const BigMeta M = meta[b];
        const TileInfo T = Ts[M.solid];
        const int C = T.n[0] * T.n[1] * T.n[2];
        if (T.off + (long long)C > code_cap) continue;  // flagged by the tile offsets
        const int q0 = (int)(blk - M.blk0) * kBigCPB;
        f(b, M, T, blk, q0, min(q0 + kBigCPB, C));
    }
}

// Iterates the flat cells of the step's big tiles: f(b, meta, tile, q, i, j, k).
template <class F>
__device__ inline void big_cells_loop(const BigMeta* meta, const TileInfo* Ts, const long long* tot, long long code_cap,
                                      F f) {
    big_blocks_loop(meta, Ts, tot, code_cap, [&](int b, const BigMeta& M, const TileInfo& T, long long, int q0, int q1) {
        const float r0 = 1.0f / (float)T.n[0], r1 = 1.0f / (float)T.n[1];
        for (int q = q0 + (int)threadIdx.x; q < q1; q += blockDim.x) {
            int r, i, j, k;
            divmod_small(q, T.n[0], r0, r, i);
            divmod_small(r, T.n[1], r1, k, j);
            f(b, M, T, q, i, j, k);
        }
    });
}

// 64 bits of a bit array starting at bit pos (the word after pos's word is read:
// the vertex-bit buffer keeps one guard word after the last solid's words).
__device__ inline unsigned long long big_bits64(const unsigned long long* w, long long pos) {
    const long long wi = pos >> 6;
    const int sh = (int)(pos & 63);
    const unsigned long long lo = w[wi] >> sh;
    return sh == 0 ? lo : (lo | (w[wi + 1] << (64 - sh)));
}

// First cell of the run of active cells along x that holds active cell i of row
// (j, k): one past the highest inactive cell below i (0 if none).  Cell p is
// inactive when none of its 8 vertices is inside: with R(p) the OR of the row's
// four vertex rows at vertex p, when R(p) = R(p + 1) = 0.  Windows of 63 cells.
__device__ inline int big_run_start(const unsigned long long* w, int i, int j, int k, int VX, int VY) {
    const long long r00 = (long long)VX * (j + (long long)VY * k), r10 = r00 + VX;
    const long long r01 = r00 + (long long)VX * VY, r11 = r01 + VX;
    int pos = i - 1;  // highest candidate cell
    while (pos >= 0) {
        const int lo = pos >= 62 ? pos - 62 : 0;
        const unsigned long long R = big_bits64(w, r00 + lo) | big_bits64(w, r10 + lo) | big_bits64(w, r01 + lo) |
                                     big_bits64(w, r11 + lo);
        // cells lo .. pos inactive where R(p) = R(p + 1) = 0
        const unsigned long long Z = ~R & ~(R >> 1) & ((2ull << (pos - lo)) - 1ull);
        if (Z) return lo + (63 - __builtin_clzll(Z)) + 1;
        if (lo == 0) return 0;
        pos = lo - 1;
    }
    return 0;
}

// A unit: up to 64 consecutive cells i0 .. i0 + len - 1 of tile row (j, k) --
// one wave, a lane per cell.  Its masks from the vertex bits (loads at
// wave-uniform addresses): act = cells with >= 1 inside vertex, all8 = cells
// with all 8 inside.  A cell's vertices are i and i + 1 of the four vertex rows
// (j, k), (j + 1, k), (j, k + 1), (j + 1, k + 1).
__device__ inline void big_unit_masks(const unsigned long long* w, int VX, int VY, int j, int k, int i0, int len,
                                      unsigned long long& act, unsigned long long& all8) {
    const long long r00 = (long long)VX * (j + (long long)VY * k) + i0, r10 = r00 + VX;
    const long long r01 = r00 + (long long)VX * VY, r11 = r01 + VX;
    const unsigned long long a = big_bits64(w, r00), b = big_bits64(w, r10), c = big_bits64(w, r01),
                             d = big_bits64(w, r11);
    const unsigned long long O = a | b | c | d, A = a & b & c & d;
    unsigned long long hiO = 0, hiA = 0;  // vertex i0 + 64: the upper vertex of cell 63
    if (len == 64) {
        const bool x0 = big_vbit(w, r00 + 64), x1 = big_vbit(w, r10 + 64), x2 = big_vbit(w, r01 + 64),
                   x3 = big_vbit(w, r11 + 64);
        hiO = (x0 || x1 || x2 || x3) ? 1ull : 0ull;
        hiA = (x0 && x1 && x2 && x3) ? 1ull : 0ull;
    }
    const unsigned long long valid = len == 64 ? ~0ull : ((1ull << len) - 1ull);
    act = (O | (O >> 1) | (hiO << 63)) & valid;
    all8 = (A & ((A >> 1) | (hiA << 63))) & valid;
}

// Unit u of a tile: its row (j, k), first cell i0, length and first tile index.
// Units run along each row (ceil(n0 / 64) per row), rows in index order, so
// consecutive units hold consecutive tile cells.
struct BigUnit {
    int j, k, i0, len, q0;
};
__device__ inline BigUnit big_unit(const TileInfo& T, int u) {
    const int nseg = (T.n[0] + 63) >> 6;
    const int r = u / nseg, seg = u - r * nseg;
    BigUnit U;
    U.j = r % T.n[1];
    U.k = r / T.n[1];
    U.i0 = seg << 6;
    U.len = min(64, T.n[0] - U.i0);
    U.q0 = r * T.n[0] + U.i0;
    return U;
}
__device__ inline int big_units(const TileInfo& T) { return T.n[1] * T.n[2] * ((T.n[0] + 63) >> 6); }

// Iterates the unit blocks of the step's big tiles, one workgroup iteration per
// block of kBigUPB units (the block's solid found once, uniform); wave w takes
// the block's units kBigUPB / 4 * w ...: f(b, meta, tile, block, unit index in
// the block, unit).  pre(b, meta, tile, block) and post(...) run once per block
// with every thread (barriers allowed).
template <class F, class Pre, class Post>
__device__ inline void big_units_loop(const BigMeta* meta, const TileInfo* Ts, const long long* tot, long long code_cap,
                                      F f, Pre pre, Post post) {
    const int nbig = (int)tot[0];
    const long long nblk = tot[9];
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (long long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const int b = big_find(nbig, blk, [&](int x) { return (long long)meta[x].ublk0; });
        const BigMeta M = meta[b];
        const TileInfo T = Ts[M.solid];
        if (T.off + (long long)T.n[0] * T.n[1] * T.n[2] > code_cap) continue;  // flagged by the tile offsets
        const int units = big_units(T);
        const int ub = (int)(blk - M.ublk0) * kBigUPB;
        pre(b, M, T, blk);
        for (int t = 0; t < kBigUPB / 4; ++t) {
            const int ui = wv * (kBigUPB / 4) + t;
            if (ub + ui >= units) break;  // wave-uniform
            f(b, M, T, blk, ui, big_unit(T, ub + ui));
        }
        post(b, M, T, blk);
    }
}
struct BigNone {
    __device__ void operator()(int, const BigMeta&, const TileInfo&, long long) const {}
};

// Cells with >= 1 inside vertex: label = the first cell of their run along x
// (so that the runs start out united), else -1; the lowest such cell (one atomic
// per unit).  The runs that start in the unit come from its mask; a run that
// enters from the unit before is traced back from the vertex bits once.
__global__ void __launch_bounds__(256) k_big_cells(const TileInfo* __restrict__ Ts, BigMeta* __restrict__ meta,
                                                   const long long* __restrict__ tot,
                                                   const unsigned long long* __restrict__ vbits, int* __restrict__ cand,
                                                   long long code_cap) {
    const int lane = threadIdx.x & 63;
    big_units_loop(
        meta, Ts, tot, code_cap,
        [&](int b, const BigMeta& M, const TileInfo& T, long long, int, const BigUnit U) {
            const unsigned long long* w = vbits + M.woff;
            const int VX = T.n[0] + 1, VY = T.n[1] + 1;
            unsigned long long act, all8;
            big_unit_masks(w, VX, VY, U.j, U.k, U.i0, U.len, act, all8);
            unsigned long long starts = act & ~(act << 1);
            int enter = 0;
            if (U.i0 > 0 && (act & 1ull)) {
                const int s0 = big_run_start(w, U.i0, U.j, U.k, VX, VY);
                if (s0 < U.i0) {
                    starts &= ~1ull;
                    enter = U.q0 - U.i0 + s0;
                }
            }
            const unsigned long long mine = starts & ((2ull << lane) - 1ull);
            if (lane < U.len) {
                const bool a = (act >> lane) & 1ull;
                cand[T.off + U.q0 + lane] = !a ? -1 : (mine ? U.q0 + (63 - __builtin_clzll(mine)) : enter);
            }
            if (act && lane == 0) {
                const int lowest = U.q0 + __builtin_ctzll(act);
                if (lowest < __hip_atomic_load(&meta[b].first, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    atomicMin(&meta[b].first, lowest);
            }
        },
        BigNone{}, BigNone{});
}

// Union of the runs over the +y and +z neighbours (hook the larger root under
// the smaller: the same concurrent union as the planes' labelling), once per
// overlap of two runs -- at its first cell, where the cells before the pair are
// not both active.
__global__ void __launch_bounds__(256) k_big_hook(const TileInfo* __restrict__ Ts, const BigMeta* __restrict__ meta,
                                                  const long long* __restrict__ tot, int* __restrict__ cand,
                                                  long long code_cap) {
    big_cells_loop(meta, Ts, tot, code_cap, [&](int, const BigMeta&, const TileInfo& T, int q, int i, int j, int k) {
        int* L = cand + T.off;
        if (L[q] < 0) return;
        const int nx = T.n[0], nxy = T.n[0] * T.n[1];
        const bool left = i > 0 && L[q - 1] >= 0;
        int nb[2], nn = 0;
        if (j + 1 < T.n[1] && L[q + nx] >= 0 && !(left && L[q - 1 + nx] >= 0)) nb[nn++] = q + nx;
        if (k + 1 < T.n[2] && L[q + nxy] >= 0 && !(left && L[q - 1 + nxy] >= 0)) nb[nn++] = q + nxy;
        for (int u = 0; u < nn; ++u) {
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
// neighbours' (a union that did not converge raises EF_LABEL).  bcnt[block] =
// the block's ALL_INSIDE and cut counts.
template <class R>
__global__ void __launch_bounds__(256) k_big_member(const MeshDev m, const TileInfo* __restrict__ Ts,
                                                    const BigMeta* __restrict__ meta, const long long* __restrict__ tot,
                                                    const unsigned long long* __restrict__ vbits,
                                                    const int* __restrict__ cand, R* __restrict__ codes,
                                                    int2* __restrict__ bcnt, long long code_cap, int* __restrict__ err) {
    __shared__ int scan_lds[32];
    big_blocks_loop(meta, Ts, tot, code_cap, [&](int, const BigMeta& M, const TileInfo& T, long long blk, int q0, int q1) {
        const int* L = cand + T.off;
        R* tc = codes + T.off;
        // the seed: the nearest cell when it has an inside vertex, else the lowest such cell
        int seed = -1;
        if (T.seed >= 0) {
            const int si = T.seed % m.nx, sj = (T.seed / m.nx) % m.ny, sk = T.seed / (m.nx * m.ny);
            const int li = si - T.lo[0], lj = sj - T.lo[1], lk = sk - T.lo[2];
            if (li >= 0 && li < T.n[0] && lj >= 0 && lj < T.n[1] && lk >= 0 && lk < T.n[2]) {
                const int sq = li + T.n[0] * (lj + T.n[1] * lk);
                if (L[sq] >= 0) seed = sq;
            }
        }
        if (seed < 0) seed = M.first;
        const int lseed = seed >= 0 && seed != 0x7fffffff ? L[seed] : -2;
        const int nx = T.n[0], nxy = T.n[0] * T.n[1];
        const float r0 = 1.0f / (float)T.n[0], r1 = 1.0f / (float)T.n[1];
        int nall = 0, ncut = 0;
        for (int q = q0 + (int)threadIdx.x; q < q1; q += blockDim.x) {
            const int lq = L[q];
            if (lq < 0) {
                tc[q] = code_none<R>();
                continue;
            }
            int r, i, j, k;
            divmod_small(q, T.n[0], r0, r, i);
            divmod_small(r, T.n[1], r1, k, j);
            const bool bad = (i + 1 < T.n[0] && L[q + 1] >= 0 && L[q + 1] != lq) ||
                             (j + 1 < T.n[1] && L[q + nx] >= 0 && L[q + nx] != lq) ||
                             (k + 1 < T.n[2] && L[q + nxy] >= 0 && L[q + nxy] != lq);
            if (bad) atomicOr(err, EF_LABEL);
            if (lq != lseed) {
                tc[q] = code_none<R>();
            } else {
                const bool all8 = big_nin(vbits + M.woff, i, j, k, T.n[0] + 1, T.n[1] + 1) == 8;
                tc[q] = all8 ? R(__builtin_huge_val()) : R(0);
                nall += all8 ? 1 : 0;
                ncut += all8 ? 0 : 1;
            }
        }
        int pa, pc, ta, tcut;
        block_exscan2<256>(nall, ncut, pa, pc, ta, tcut, scan_lds);
        if (threadIdx.x == 0) bcnt[blk] = make_int2(ta, tcut);
        __syncthreads();  // scan_lds is reused by the next block
    });
}

// Per big solid (one workgroup each): its blocks' counts scanned in block order
// -> each block's list offsets (ALL_INSIDE, cut), the solid's two counts and
// its number of candidate chunks.
__global__ void __launch_bounds__(1024) k_big_scan(const TileInfo* __restrict__ Ts, BigMeta* __restrict__ meta,
                                                   const long long* __restrict__ tot, const int2* __restrict__ bcnt,
                                                   int2* __restrict__ boff, int2* __restrict__ ncand,
                                                   long long code_cap, int chunk_entries) {
    __shared__ int scan_lds[32];
    const int nbig = (int)tot[0];
    for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
        const BigMeta M = meta[b];
        const TileInfo T = Ts[M.solid];
        if (T.off + (long long)T.n[0] * T.n[1] * T.n[2] > code_cap) continue;  // flagged; uniform
        int ca = 0, cc = 0;  // carries over rounds of 1024 blocks
        for (int base = 0; base < M.nblk; base += 1024) {
            const int t = base + (int)threadIdx.x;
            const int2 v = t < M.nblk ? bcnt[M.blk0 + t] : make_int2(0, 0);
            int pa, pc, ta, tc;
            block_exscan2<1024>(v.x, v.y, pa, pc, ta, tc, scan_lds);
            if (t < M.nblk) boff[M.blk0 + t] = make_int2(ca + pa, cc + pc);
            ca += ta;
            cc += tc;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            ncand[M.solid] = make_int2(ca, cc);
            meta[b].nchunk = (ca + cc + chunk_entries - 1) / chunk_entries;
        }
    }
}

// Candidate lists of the big tiles: ALL_INSIDE cells first, then cut cells,
// each ascending (tile-local linear indices); per block a stable compaction of
// its cells (4 consecutive per thread) at the offsets k_big_scan gave it.
template <class R>
__global__ void __launch_bounds__(256) k_big_write(const TileInfo* __restrict__ Ts, const BigMeta* __restrict__ meta,
                                                   const long long* __restrict__ tot, const R* __restrict__ codes,
                                                   const int2* __restrict__ boff, const int2* __restrict__ ncand,
                                                   int* __restrict__ cand, long long code_cap) {
    static_assert(kBigCPB == 4 * 256, "4 cells per thread");
    __shared__ int scan_lds[32];
    big_blocks_loop(meta, Ts, tot, code_cap, [&](int, const BigMeta& M, const TileInfo& T, long long blk, int q0, int q1) {
        const R* tc = codes + T.off;
        const int qa = q0 + 4 * (int)threadIdx.x;
        int cls[4];
        int na = 0, nc = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = qa + u;
            const R v = q < q1 ? tc[q] : code_none<R>();
            cls[u] = v == R(__builtin_huge_val()) ? 1 : (v == R(0) ? 2 : 0);
            na += cls[u] == 1 ? 1 : 0;
            nc += cls[u] == 2 ? 1 : 0;
        }
        int pa, pc, ta, tcut;
        block_exscan2<256>(na, nc, pa, pc, ta, tcut, scan_lds);
        const int2 o = boff[blk];
        int* tl = cand + T.off;
        int wa = o.x + pa, wc = ncand[M.solid].x + o.y + pc;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (cls[u] == 1) tl[wa++] = qa + u;
            if (cls[u] == 2) tl[wc++] = qa + u;
        }
        __syncthreads();
    });
}

// Chunk offsets (one workgroup, exclusive scan over the big solids).
__global__ void __launch_bounds__(1024) k_big_chunks(BigMeta* __restrict__ meta, long long* __restrict__ tot,
                                                     long long chunk_cap, int* __restrict__ err) {
    __shared__ int part[1024];
    const int nbig = (int)tot[0];
    const int tid = threadIdx.x;
    int c[kBigPer], sum = 0;
    for (int u = 0; u < kBigPer; ++u) {
        const int b = tid * kBigPer + u;
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
    for (int u = 0; u < kBigPer; ++u) {
        const int b = tid * kBigPer + u;
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
    __shared__ CsgK sh_csg;
    const int nbig = (int)tot[0];
    const long long nch = tot[4];
    const double dtINV = 1.0 / dt;
    const int tid = threadIdx.x;
    for (long long ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const int b = big_find(nbig, ch, [&](int x) { return (long long)meta[x].chunk0; });
        const int s = meta[b].solid;
        const TileInfo T = Ts[s];
        if (T.off + (long long)T.n[0] * T.n[1] * T.n[2] > code_cap) continue;  // flagged; uniform per workgroup
        SolidK K = Ks[s];
        big_csg_to_lds(K, sh_csg, tid, NT);
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

// Force and torque of the big solids: chunk partials in a fixed order (lanes
// strided over the chunks, then a shuffle tree), x rho_f.
__global__ void __launch_bounds__(256) k_big_ft(const BigMeta* __restrict__ meta, const long long* __restrict__ tot,
                                                const double* __restrict__ part, double rho_f,
                                                double* __restrict__ ft) {
    const int nbig = (int)tot[0];
    const bool ok = tot[4] > 0 || tot[6] == 0;  // chunk partials exist (or there are no candidates)
    const int lane = threadIdx.x & 63;
    const int nw = gridDim.x * (blockDim.x >> 6);
    // a wave per solid: lanes strided over its chunks in order, then a fixed shuffle tree
    for (int b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); b < nbig; b += nw) {
        const BigMeta M = meta[b];
        double a[6] = {0, 0, 0, 0, 0, 0};
        if (ok)
            for (int c = lane; c < M.nchunk; c += 64)
                for (int t = 0; t < 6; ++t) a[t] += part[6LL * (M.chunk0 + c) + t];
        for (int off = 32; off >= 1; off >>= 1)
            for (int t = 0; t < 6; ++t) a[t] += __shfl_xor(a[t], off, 64);
        if (lane == 0)
            for (int t = 0; t < 6; ++t) ft[6 * M.solid + t] = a[t] * rho_f;
    }
}
