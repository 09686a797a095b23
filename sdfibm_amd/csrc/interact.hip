// interact.hip -- MI355X (gfx950) kernels of the solid-fluid coupling step and
// the C ABI of include/sdfibm_hip.h.
//
// One interact (SolidCloud::interact, src/solidcloud.cpp:435-464) runs as:
//
//   k_setup        one thread per solid: shape constants, the solid's cell tile
//                  (index box holding every cell with a vertex strictly inside),
//                  the seed cell (nearest cell centre, solidcloud.cpp:363).
//   k_scan         tile sizes -> offsets into the per-tile code buffer.
//   k_classify     one workgroup per finite solid: inside-vertex bits of the
//                  tile in LDS (64-bit rows along x), cells with >= 1 inside
//                  vertex, the connected component grown from the seed (or the
//                  first such cell when the seed has none; cellenumerator.cpp:
//                  6-78) by bit-parallel flood fill, ALL_INSIDE / cut
//                  classification, compacted candidate lists.
//   k_cells        one workgroup per finite solid over its candidate list:
//                  volume fractions of the cut cells, the centre test, the
//                  forcing and the solid's force and torque (FP64 block
//                  reduction).  Together the two write one code per tile cell.
//   k_bin_*        per-brick lists of the solids whose tiles overlap the brick.
//   k_fields       one workgroup per brick of cells: every cell of the mesh is
//                  written exactly once -- As (clipped to 1), Fs, Ts, Ct summed
//                  over the brick's solids in index order, planes evaluated in
//                  place.  This fuses the zero-fill, the per-solid accumulation
//                  and checkAlpha (solidcloud.cpp:438-441, 403-422, 564-570).
//   k_plane_reduce force/torque of plane solids from per-brick partials.
//
// Tile codes (element type R): NaN = not a candidate; +inf = ALL_INSIDE
// (alpha 1); finite >= 0 with sign bit clear = CENTER_INSIDE with that alpha;
// sign bit set = CENTER_OUTSIDE with alpha = |code|.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <mutex>
#include <vector>

#include "../../include/sdfibm_hip.h"
#include "dynamics.h"
#include "mesh_dev.h"
#include "surrogate.h"

using namespace sdfibm_dev;

namespace {

constexpr int kBlock = 256;
constexpr int kMaxPerBrick = 128;  // solids per brick handled by k_fields
#ifndef SDFIBM_FILL_REP
#define SDFIBM_FILL_REP 1
#endif
constexpr int kFillRep = SDFIBM_FILL_REP;  // row updates per flood-fill pass
constexpr int kMaxPlanes = 32;  // plane and oversized solids (unbounded path)
constexpr int kContactBlockSolids = 1024;   // evolve with contacts in one workgroup up to this many solids
constexpr int kContactBlockCells = 1 << 16;  // ... and contact grid cells

// Staging area of the allocation-free velocity upload
// (sdfibm_set_velocity_cells_noalloc): kStageCells cells of FP64 AoS velocity,
// 24 MiB, held by every GPU the library runs on until the process ends.  Host
// code never passes this name to a kernel or copy: it uses the device address
// that hipGetSymbolAddress gives for the current device (sdfibm_ctx::stage).
constexpr long long kStageCells = SDFIBM_NOALLOC_MAX_CELLS;
static __device__ __align__(256) double scratch[3 * kStageCells];

enum ErrFlags { EF_TILE = 1, EF_CODECAP = 2, EF_BRICK = 8, EF_PLANE_MOVED = 16, EF_LABEL = 32 };

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ inline double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

template <int NV>
__device__ inline void wave_reduce(double* v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
        for (int t = 0; t < NV; ++t) v[t] += shfl_xor_d(v[t], off);
}

// Block-wide sum of NV doubles (blockDim.x multiple of 64); result valid in thread 0.
template <int NV>
__device__ inline void block_reduce(double* v, double* lds /* [NV * nwaves] */) {
    wave_reduce<NV>(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int t = 0; t < NV; ++t) lds[t * 16 + wid] = v[t];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int t = 0; t < NV; ++t) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += lds[t * 16 + w];
            v[t] = s;
        }
    }
    __syncthreads();
}

template <class R>
__device__ inline R code_none() {
    return R(__builtin_nan(""));
}

// q = a / b, r = a % b for 0 <= a < 2^22 and 1 <= b < 2^12, rb = 1.0f / b: the
// truncated float product is within one of the quotient (relative error below
// 2^-23 on a value below 2^22); one step corrects it.
__device__ inline void divmod_small(int a, int b, float rb, int& q, int& r) {
    q = (int)((float)a * rb);
    r = a - q * b;
    if (r < 0) {
        --q;
        r += b;
    } else if (r >= b) {
        ++q;
        r -= b;
    }
}

// Run fill inside a row: bits of `a` reachable from `s` (s subset of a) along x.
__device__ inline unsigned long long run_fill(unsigned long long a, unsigned long long s) {
    const unsigned long long up = (((a + s) ^ a) & a) | s;
    const unsigned long long ra = __builtin_bitreverse64(a), rs = __builtin_bitreverse64(s);
    const unsigned long long dn = __builtin_bitreverse64((((ra + rs) ^ ra) & ra) | rs);
    return up | dn;
}

// LDS words read and written while other threads of the workgroup may write
// them (the flood fill's relaxation): relaxed atomics keep every access a
// single 64-bit LDS operation that the compiler neither drops nor reorders.
__device__ inline unsigned long long lds_load_relaxed(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_store_relaxed(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int* p, long long n, int v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Work queues of the per-solid tile pass, one per tile class (0 small, 1 medium,
// 2 unused; k_setup moves small tiles of at most kRowsTiny vertex rows to 3, tiny;
// 4 big, the big-tile pass): qmeta[q] = number of solids queued, qlist[q * ns + t]
// = the t-th of them.
constexpr int kQueues = 5;
__host__ __device__ inline int tile_queue(int cls) {
    return cls == TC_SMALL ? 0 : (cls == TC_MEDIUM ? 1 : (cls == TC_BIG ? 4 : -1));
}

// The tiles' code and list offsets: one atomic per wave on the step's running
// total (a bump allocation; the layout does not depend on the order, every tile
// is addressed through its own offset), the wave's tiles in lane order.  On a
// rectilinear mesh of up to kSetupStage vertex coordinates the three coordinate
// arrays are staged in LDS first (dynamic LDS of setup_lds_bytes): each lane's
// tile bounds and nearest cell are searches whose probes depend on each other,
// served from LDS instead of memory.
constexpr int kSetupStage = 3 * 1025;

__host__ inline size_t setup_lds_bytes(const MeshDev& m, int mk) {
    const long long n = (long long)m.nx + m.ny + m.nz + 3;
    return mk == MK_RECT && n <= kSetupStage ? (size_t)n * sizeof(double) : 0;
}

template <int MK>
__global__ void __launch_bounds__(64) k_setup(const sdfibm_solid* __restrict__ in, int ns, MeshDev m, SolidK* __restrict__ K,
                        TileInfo* __restrict__ T, long long* __restrict__ total, long long code_cap,
                        double* __restrict__ ft, int* __restrict__ qmeta, int* __restrict__ qlist,
                        const unsigned char* __restrict__ unb, int* __restrict__ err, const CsgK* __restrict__ csg,
                        int big_on) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool live = s < ns;  // the whole wave takes part in the offsets' scan
    MeshDev mm = m;
    if constexpr (MK == MK_RECT) {
        extern __shared__ double sc[];
        const int n0 = m.nx + 1, n1 = m.ny + 1, n2 = m.nz + 1;
        if (n0 + n1 + n2 <= kSetupStage) {  // the same for the whole workgroup
            for (int v = threadIdx.x; v < n0; v += blockDim.x) sc[v] = m.xs[v];
            for (int v = threadIdx.x; v < n1; v += blockDim.x) sc[n0 + v] = m.ys[v];
            for (int v = threadIdx.x; v < n2; v += blockDim.x) sc[n0 + n1 + v] = m.zs[v];
            __syncthreads();
            mm.xs = sc;
            mm.ys = sc + n0;
            mm.zs = sc + n0 + n1;
        }
    }
    TileInfo t{};
    t.cls = TC_EMPTY;
    if (live) {
        const SolidK k = make_solid_k(in[s], csg);
        K[s] = k;
        t = make_tile<MK>(mm, k);
        // solids set_solids put on the unbounded path stay there (their tile box is the cull)
        if (unb[s] && t.cls != TC_PLANE) t.cls = t.cls == TC_EMPTY ? TC_EMPTY : TC_HUGE;
        // structured meshes: the seed search (a box of cell centres) is spread over a wave by k_seed
        t.seed = t.cls == TC_EMPTY ? -1 : (MK == MK_RECT ? nearest_cell<MK>(mm, k.c) : -2);
        // a tile beyond 2^22 vertices, or a big tile when set_solids found none possible (the
        // big-tile pass is not launched): interact classifies the solids again
        if (t.cls == TC_TOOBIG || (t.cls == TC_BIG && !big_on)) atomicOr(err, EF_TILE);
        for (int q = 0; q < 6; ++q) ft[6 * s + q] = 0.0;  // solids that touch no cell exert nothing
    }
    const long long tc = live && tile_is_finite(t.cls) ? (long long)t.n[0] * t.n[1] * t.n[2] : 0;
    long long x = tc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const long long y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    const long long wtot = __shfl(x, 63, 64);
    // the solid's queue (the lanes of a wave are consecutive solids, queued in index order)
    int q = live ? tile_queue(t.cls) : -1;
    if (q == 0 && (t.n[1] + 1) * (t.n[2] + 1) <= kRowsTiny) q = 3;  // tiny tiles: a wave each
    unsigned long long bq = 0;  // lane 1 + qq: the wave's members of queue qq
#pragma unroll
    for (int qq = 0; qq < kQueues; ++qq) {
        const unsigned long long b = __ballot(q == qq);
        if (lane == 1 + qq) bq = b;
    }
    // the wave's atomics in one round: lane 0 the tiles' running total, lane 1 + qq queue qq
    long long got = 0;
    if (lane == 0 && wtot > 0)
        got = (long long)atomicAdd((unsigned long long*)total, (unsigned long long)wtot);
    else if (lane >= 1 && lane <= kQueues && bq)
        got = atomicAdd(&qmeta[lane - 1], __popcll(bq));
    const long long base = __shfl(got, 0, 64);
    const int qbase = (int)__shfl(got, q >= 0 ? 1 + q : 0, 64);
    const unsigned long long qmask = __shfl(bq, q >= 0 ? 1 + q : 0, 64);
    if (lane == 0 && wtot > 0 && base + wtot > code_cap) atomicOr(err, EF_CODECAP);
    if (!live) return;
    t.off = base + x - tc;
    T[s] = t;
    if (q >= 0) qlist[(long long)q * ns + qbase + __popcll(qmask & ((1ull << lane) - 1ull))] = s;
}

// The seed cell of every non-empty solid on a structured mesh (nearest_cell's
// rule, one wave per solid): the smallest cell index among the cells whose
// centre lies within a relative 1e-10 of the smallest squared distance, over
// the box of cells the jitter allows around the nominal cell.
__global__ void __launch_bounds__(256) k_seed(MeshDev m, int ns, const SolidK* __restrict__ Ks,
                                              TileInfo* __restrict__ Ts) {
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= ns || Ts[s].seed != -2) return;  // wave-uniform
    const double* p = Ks[s].c;
    const int nn[3] = {m.nx, m.ny, m.nz};
    int c0[3];
    for (int d = 0; d < 3; ++d) {
        double f = floor((p[d] - m.o[d]) / m.h[d]);
        if (f < 0) f = 0;
        if (f > nn[d] - 1) f = nn[d] - 1;
        c0[d] = (int)f;
    }
    const int r = (int)ceil(2.0 * m.jit) + 1, w = 2 * r + 1, box = w * w * w;
    double best = __builtin_huge_val();
    for (int t = lane; t < box; t += 64) {
        const int i = c0[0] - r + t % w, j = c0[1] - r + (t / w) % w, k = c0[2] - r + t / (w * w);
        if (i < 0 || i >= m.nx || j < 0 || j >= m.ny || k < 0 || k >= m.nz) continue;
        const long long c = cell_id(m, i, j, k);
        const double dx = m.ccx[c] - p[0], dy = m.ccy[c] - p[1], dz = m.ccz[c] - p[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < best) best = d2;
    }
    for (int off = 32; off >= 1; off >>= 1) best = fmin(best, __shfl_xor(best, off, 64));
    const double lim = best * (1.0 + kNearTie) + 1e-300;
    long long bc = LLONG_MAX;
    for (int t = lane; t < box; t += 64) {
        const int i = c0[0] - r + t % w, j = c0[1] - r + (t / w) % w, k = c0[2] - r + t / (w * w);
        if (i < 0 || i >= m.nx || j < 0 || j >= m.ny || k < 0 || k >= m.nz) continue;
        const long long c = cell_id(m, i, j, k);
        const double dx = m.ccx[c] - p[0], dy = m.ccy[c] - p[1], dz = m.ccz[c] - p[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 <= lim && c < bc) bc = c;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const long long o = __shfl_xor(bc, off, 64);
        bc = o < bc ? o : bc;
    }
    if (lane == 0) Ts[s].seed = bc == LLONG_MAX ? -1 : (int)bc;
}

// Exclusive scan of n values in three passes: tile sums, a scan of the tile
// sums, tile-local scans.  out[n] = total.  kScanTile values per workgroup.
constexpr int kScanTile = 4096;

template <class T>
__global__ void __launch_bounds__(256) k_scan_sums(const T* __restrict__ in, long long n, long long* __restrict__ sums) {
    __shared__ long long part[256];
    const long long base = (long long)blockIdx.x * kScanTile;
    long long s = 0;
    for (int t = threadIdx.x; t < kScanTile; t += 256) {
        const long long i = base + t;
        if (i < n) s += (long long)in[i];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = part[0];
}

// In-place exclusive scan of nb tile sums (one workgroup of 1024); sums[nb] = total.
__global__ void __launch_bounds__(1024) k_scan_top(long long* __restrict__ sums, int nb) {
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const int chunk = (nb + 1023) / 1024;
    const int b = t * chunk, e = min(b + chunk, nb);
    long long s = 0;
    for (int i = b; i < e; ++i) s += sums[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const long long v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = t > 0 ? part[t - 1] : 0;
    for (int i = b; i < e; ++i) {
        const long long x = sums[i];
        sums[i] = run;
        run += x;
    }
    if (t == 1023) sums[nb] = part[1023];
}

template <class T>
__global__ void __launch_bounds__(256) k_scan_tiles(const T* __restrict__ in, long long n,
                                                    const long long* __restrict__ sums, int nb,
                                                    long long* __restrict__ out) {
    constexpr int PER = kScanTile / 256;
    __shared__ long long part[256];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * PER;
    long long v[PER];
    long long s = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const long long i = base + q;
        v[q] = i < n ? (long long)in[i] : 0;
        s += v[q];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const long long x = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    long long run = sums[blockIdx.x] + (threadIdx.x > 0 ? part[threadIdx.x - 1] : 0);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const long long i = base + q;
        if (i < n) out[i] = run;
        run += v[q];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = sums[nb];
}

template <class T>
void scan(const T* in, long long n, long long* out, long long* tmp, hipStream_t st) {
    const int nb = (int)((n + kScanTile - 1) / kScanTile);
    if (nb == 0) {
        k_scan_top<<<1, 1024, 0, st>>>(tmp, 0);
        k_scan_tiles<T><<<1, 256, 0, st>>>(in, 0, tmp, 0, out);
        return;
    }
    k_scan_sums<T><<<nb, 256, 0, st>>>(in, n, tmp);
    k_scan_top<<<1, 1024, 0, st>>>(tmp, nb);
    k_scan_tiles<T><<<nb, 256, 0, st>>>(in, n, tmp, nb, out);
}


// Per-step resets: error flags, queue counters, brick counts.
__global__ void k_flags_out(const int* __restrict__ err, int* __restrict__ host_word) {
    if (threadIdx.x == 0) *(volatile int*)host_word = *err;
}

__global__ void k_reset(int* __restrict__ err, int* __restrict__ qmeta, int* __restrict__ bcount, long long nbricks,
                        long long* __restrict__ total) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *err = 0;
    if (i < 8) qmeta[i] = 0;
    if (i == 0) *total = 0;  // the tiles' running total (k_setup)
    if (i < nbricks) bcount[i] = 0;
}

// ---------------------------------------------------------------------------
// per-solid tile pass, in two kernels
//
//   k_classify  (workgroup per solid, LDS rows, few registers): inside bit of
//               every tile vertex, cells with >= 1 inside vertex, the seed's
//               connected component (flood fill), codes of the non-candidate and
//               ALL_INSIDE cells, and the compacted candidate list of the tile:
//               ALL_INSIDE cells first, then the cut cells, each part ascending.
//   k_cells     (workgroup per solid, no LDS rows, the arithmetic): volume
//               fraction and centre test of every cut cell, the forcing of every
//               candidate and the solid's FP64 force/torque.
//
// Both take their solids from work queues (persistent workgroups).
// ---------------------------------------------------------------------------

// Block-wide exclusive scan of two per-thread counts (NT threads): wave prefix
// sums by shuffles, then the wave totals through LDS.
template <int NT>
__device__ inline void block_exscan2(int a, int b, int& pa, int& pb, int& ta, int& tb, int* lds /* [32] */) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int ia = a, ib = b;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int ya = __shfl_up(ia, off, 64), yb = __shfl_up(ib, off, 64);
        if (lane >= off) {
            ia += ya;
            ib += yb;
        }
    }
    if (lane == 63) {
        lds[wv] = ia;
        lds[16 + wv] = ib;
    }
    __syncthreads();
    int ba = 0, bb = 0, sa = 0, sb = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const int x = lds[w], y = lds[16 + w];
        if (w < wv) {
            ba += x;
            bb += y;
        }
        sa += x;
        sb += y;
    }
    pa = ba + ia - a;
    pb = bb + ib - b;
    ta = sa;
    tb = sb;
    __syncthreads();
}

// Solid of item t of the tile pass, largest tile class first.
__device__ inline int queue_item(const int* qmeta, const int* qlist, int ns, int t) {
    const int n2 = qmeta[2], n1 = qmeta[1], n0 = qmeta[0];
    if (t < n2) return qlist[2LL * ns + t];
    t -= n2;
    if (t < n1) return qlist[1LL * ns + t];
    t -= n1;
    if (t < n0) return qlist[t];
    return qlist[3LL * ns + (t - n0)];
}

// Candidate entry: the cell's tile-local (i, j, k) packed as i | j << 6 | k << 19
// (i < 64, j and k < 2^13: a tile has at most 1024 vertex rows); the ALL_INSIDE
// entries come first in the list, the count of each part is in ncand[s] =
// {inside, cut}.
constexpr int kStageYZ = 256;  // tile coordinates staged in LDS up to this many per axis

// Workgroup b takes the solids b, b + G, b + 2G, ... of queue Q (G = grid size;
// for the small class G is the number of solids, one each, and the hardware
// dispatcher balances the uneven tiles).  MAXR bounds the vertex rows
// (ny+1)(nz+1) of the tiles of the queue.  No shared work counter: that would
// cost a device-scope atomic (slow across the XCDs) per solid.
// Register budget: 64 VGPRs in FP32 (eight waves per SIMD; the kernel waits on dependent
// loads more than it computes), 80 in FP64 (no spills).
template <class R, int MK, int MAXR, int Q, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 8 : 6))) k_classify(MeshDev m, int ns, const SolidK* __restrict__ Ks,
                                                     const TileInfo* __restrict__ Ts, int* __restrict__ qmeta,
                                                     const int* __restrict__ qlist, R* __restrict__ codes,
                                                     int* __restrict__ cand, long long code_cap,
                                                     int2* __restrict__ ncand) {
    __shared__ unsigned long long vin[MAXR];
    __shared__ unsigned long long act[MAXR];
    __shared__ unsigned long long all8[MAXR];
    __shared__ unsigned long long reach[MAXR];
    constexpr int NSX = MK == MK_RECT ? kMaxTileX + 2 : 1;
    constexpr int NSYZ = MK == MK_RECT ? (MAXR / 2 + 1 < kStageYZ ? MAXR / 2 + 1 : kStageYZ) : 1;
    __shared__ double sx[NSX], sy[NSYZ], sz[NSYZ];
    __shared__ int scan_lds[32];
    __shared__ int sh_first, sh_changed, sh_flag[3];
    // connectivity summary of the active rows: nonzero rows, j-links, rows of
    // more than one run; per slice k: nonzero, linked to slice k - 1
    constexpr int kSW = (MAXR / 2 + 63) / 64;
    __shared__ int sh_cnt[3];
    __shared__ unsigned long long sh_knz[kSW], sh_kl[kSW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nq = qmeta[Q];
    for (int item = blockIdx.x; item < nq; item += gridDim.x) {
    __syncthreads();  // LDS of the previous solid is free
    const int s = __builtin_amdgcn_readfirstlane(qlist[(long long)Q * ns + item]);
    const TileInfo T = Ts[s];
    if (T.off + (long long)T.n[0] * T.n[1] * T.n[2] > code_cap) continue;

    const SolidK K = Ks[s];
    const int nx = T.n[0], ny = T.n[1], nz = T.n[2];
    const int VX = nx + 1, VY = ny + 1;
    const int VR = VY * (nz + 1), CR = ny * nz;
    const unsigned long long cmask = (nx >= 64) ? ~0ull : ((1ull << nx) - 1ull);
    const float rnx = 1.0f / (float)nx, rny = 1.0f / (float)ny, rVY = 1.0f / (float)VY;

    // tile coordinates of a rectilinear mesh in LDS
    const bool staged = MK == MK_RECT && VY <= NSYZ && nz + 1 <= NSYZ;
    if (staged) {
        for (int t = tid; t < VX; t += NT) sx[t] = m.xs[T.lo[0] + t];
        for (int t = tid; t < VY; t += NT) sy[t] = m.ys[T.lo[1] + t];
        for (int t = tid; t <= nz; t += NT) sz[t] = m.zs[T.lo[2] + t];
    }
    if (tid == 0) sh_first = 0x7fffffff;
    if (tid < 3) sh_cnt[tid] = 0;
    for (int t = tid; t < kSW; t += NT) sh_knz[t] = sh_kl[t] = 0ull;
    __syncthreads();

    // 1. inside test at every tile vertex (FP64, reference operation order).  A
    //    wave covers g = 64 / VX vertex rows per pass; one ballot gives their bits.
    //    Quadric solids on a staged rectilinear tile: a row per thread, the
    //    interval between the roots of the row's quadratic (row_inside_bits).
    if (staged && row_quadric(K.shape)) {
        for (int r = tid; r < VR; r += NT) {
            int vj, vk;
            divmod_small(r, VY, rVY, vk, vj);
            vin[r] = row_inside_bits(K, sx, VX, sy[vj], sz[vk]);
        }
    } else {
        const int g = 64 / VX;
        const int rr = lane / VX, vi = lane - rr * VX;
        const unsigned long long vmask = VX >= 64 ? ~0ull : ((1ull << VX) - 1ull);
        for (int base = wv * g; base < VR; base += (NT / 64) * g) {
            const int r = base + rr;
            const bool mine = rr < g && r < VR;
            bool in = false;
            if (mine) {
                int vj, vk;
                divmod_small(r, VY, rVY, vk, vj);
                const d3 p = staged ? d3{sx[vi], sy[vj], sz[vk]}
                                    : vertex_pos<MK>(m, T.lo[0] + vi, T.lo[1] + vj, T.lo[2] + vk);
                in = inside_world(K, p);
            }
            const unsigned long long bal = __ballot(in);
            if (mine && vi == 0) vin[r] = (bal >> (rr * VX)) & vmask;
        }
    }
    __syncthreads();

    // 2. per cell row: >= 1 inside vertex / all 8 inside; the first such cell
    //    (one LDS atomic per wave)
    {
        bool found = false;
        for (int r0 = wv * 64; r0 < CR; r0 += NT) {
            const int r = r0 + lane;
            unsigned long long av = 0ull;
            if (r < CR) {
                int j, k;
                divmod_small(r, ny, rny, k, j);
                const unsigned long long a = vin[k * VY + j], b = vin[k * VY + j + 1], c = vin[(k + 1) * VY + j],
                                         d = vin[(k + 1) * VY + j + 1];
                const unsigned long long o = a | b | c | d, n = a & b & c & d;
                av = (o | (o >> 1)) & cmask;
                act[r] = av;
                all8[r] = (n & (n >> 1)) & cmask;
                reach[r] = 0ull;
            }
            const unsigned long long b = __ballot(av != 0ull);
            if (b && !found) {
                found = true;
                if (lane == __builtin_ctzll(b)) atomicMin(&sh_first, r * 64 + __builtin_ctzll(av));
            }
        }
    }
    __syncthreads();

    // 2b. is the active set connected?  It is when every nonzero row is one run
    //     of cells, the nonzero rows of each slice k form one chain of rows that
    //     overlap their j-neighbour (rows - links = 1 per slice), the nonempty
    //     slices are consecutive, and each overlaps the one before it in some
    //     row.  Then the flood fill from any active seed reaches every active
    //     cell; otherwise the fill below runs.  (Convex solids pass this test.)
    {
        for (int r0 = wv * 64; r0 < CR; r0 += NT) {
            const int r = r0 + lane;
            bool nzr = false, jl = false, multi = false, kl = false;
            int k = 0;
            if (r < CR) {
                int j;
                divmod_small(r, ny, rny, k, j);
                const unsigned long long a = act[r];
                nzr = a != 0ull;
                if (nzr) {
                    const unsigned long long t = a >> __builtin_ctzll(a);
                    multi = (t & (t + 1ull)) != 0ull;
                    jl = j > 0 && (a & act[r - 1]) != 0ull;
                    kl = k > 0 && (a & act[r - ny]) != 0ull;
                }
            }
            const unsigned long long bn = __ballot(nzr), bj = __ballot(jl), bm = __ballot(multi);
            if (lane == 0 && bn) {
                atomicAdd(&sh_cnt[0], __popcll(bn));
                atomicAdd(&sh_cnt[1], __popcll(bj));
                atomicAdd(&sh_cnt[2], __popcll(bm));
            }
            if (nzr) atomicOr(&sh_knz[k >> 6], 1ull << (k & 63));
            if (kl) atomicOr(&sh_kl[k >> 6], 1ull << (k & 63));
        }
    }
    __syncthreads();

    // 3. seed: nearest cell if it has an inside vertex, else the first cell (index
    //    order) that has one.  Tile raster order is ascending cell index.
    if (tid == 0) {
        int seed_r = -1, seed_b = -1;
        if (T.seed >= 0) {
            const int si = T.seed % m.nx, sj = (T.seed / m.nx) % m.ny, sk = T.seed / (m.nx * m.ny);
            const int li = si - T.lo[0], lj = sj - T.lo[1], lk = sk - T.lo[2];
            if (li >= 0 && li < nx && lj >= 0 && lj < ny && lk >= 0 && lk < nz) {
                const int r = lk * ny + lj;
                if ((act[r] >> li) & 1ull) {
                    seed_r = r;
                    seed_b = li;
                }
            }
        }
        if (seed_r < 0 && sh_first != 0x7fffffff) {
            seed_r = sh_first / 64;
            seed_b = sh_first % 64;
        }
        bool conn = sh_cnt[2] == 0;
        if (conn) {
            int nslice = 0, kmin = 0x7fffffff, kmax = -1;
            unsigned long long carry = 0ull;
            for (int w = 0; w <= (nz - 1) >> 6; ++w) {
                const unsigned long long z = sh_knz[w];
                nslice += __popcll(z);
                if (z) {
                    kmin = min(kmin, w * 64 + __builtin_ctzll(z));
                    kmax = w * 64 + 63 - __builtin_clzll(z);
                }
                if (z & ((z << 1) | carry) & ~sh_kl[w]) conn = false;  // slices k - 1, k both nonempty, not linked
                carry = z >> 63;
            }
            conn = conn && nslice > 0 && kmax - kmin + 1 == nslice && sh_cnt[0] - sh_cnt[1] == nslice;
        }
        if (seed_r >= 0) reach[seed_r] = run_fill(act[seed_r], 1ull << seed_b);
        // 0: no seed, 1: fill from the seed, 2: every active cell is reached
        sh_changed = seed_r < 0 ? 0 : (conn ? 2 : 1);
        sh_flag[0] = 0;
    }
    __syncthreads();

    // 4. flood fill over face neighbours within cells that have an inside vertex,
    //    by chaotic relaxation: between two barriers every thread applies the row
    //    update kFillRep times, reading its neighbours' rows as the other threads
    //    leave them (a row only gains bits, so any mix of old and new values is a
    //    subset of the component).  A pass in which no thread changes a row has
    //    read only the rows as they stood at its start, so the rows are then the
    //    seed's component.  Three flags rotate: a pass's flag is cleared two
    //    passes ahead of its use.
    if (sh_changed == 2) {
        for (int r = tid; r < CR; r += NT) reach[r] = act[r];
        __syncthreads();
    } else if (sh_changed == 1) {
        for (int it = 0;; ++it) {
            if (tid == 0) sh_flag[(it + 1) % 3] = 0;
            int ch = 0;
#pragma unroll 1
            for (int rep = 0; rep < kFillRep; ++rep) {
                for (int r = tid; r < CR; r += NT) {
                    int j, k;
                    divmod_small(r, ny, rny, k, j);
                    unsigned long long nb = lds_load_relaxed(&reach[r]);
                    const unsigned long long own = nb;
                    if (j > 0) nb |= lds_load_relaxed(&reach[r - 1]);
                    if (j < ny - 1) nb |= lds_load_relaxed(&reach[r + 1]);
                    if (k > 0) nb |= lds_load_relaxed(&reach[r - ny]);
                    if (k < nz - 1) nb |= lds_load_relaxed(&reach[r + ny]);
                    const unsigned long long a = act[r];
                    const unsigned long long sd = nb & a;
                    if (sd & ~own) {
                        lds_store_relaxed(&reach[r], run_fill(a, sd));
                        ch = 1;
                    }
                }
            }
            if (ch) sh_flag[it % 3] = 1;
            __syncthreads();
            if (!sh_flag[it % 3]) break;
        }
    }

    // 5. candidate list: each thread takes a contiguous run of rows; ALL_INSIDE
    //    cells first, then the cut cells (reached, not all inside), each ascending
    {
        const int chunk = (CR + NT - 1) / NT;
        const int r0 = min(tid * chunk, CR), r1 = min(r0 + chunk, CR);
        int nin = 0, ncut = 0;
        for (int r = r0; r < r1; ++r) {
            const unsigned long long b = reach[r], a = all8[r];
            nin += __builtin_popcountll(b & a);
            ncut += __builtin_popcountll(b & ~a);
        }
        int qi, qc, tin, tcut;
        block_exscan2<NT>(nin, ncut, qi, qc, tin, tcut, scan_lds);
        qc += tin;
        int* tl = cand + T.off;
        for (int r = r0; r < r1; ++r) {
            const unsigned long long a = all8[r];
            unsigned long long b = reach[r];
            int j, k;
            divmod_small(r, ny, rny, k, j);
            const int jk = (j << 6) | (k << 19);
            while (b) {
                const int i = __builtin_ctzll(b);
                b &= b - 1ull;
                if ((a >> i) & 1ull)
                    tl[qi++] = jk | i;
                else
                    tl[qc++] = jk | i;
            }
        }
        if (tid == 0) ncand[s] = make_int2(tin, tcut);
    }

    // 6. codes of the non-candidate and ALL_INSIDE cells (k_cells writes the cut ones)
    R* tcodes = codes + T.off;
    const int NC = nx * CR;
    for (int c = tid; c < NC; c += NT) {
        int i, r;
        divmod_small(c, nx, rnx, r, i);
        const unsigned long long bit = 1ull << i;
        if (!(reach[r] & bit))
            tcodes[c] = code_none<R>();
        else if (all8[r] & bit)
            tcodes[c] = R(__builtin_huge_val());
    }
    }  // solids of the workgroup
}

// The forcing of one candidate cell (solidcloud.cpp:384-422): f = alpha (uf - us),
// force += f V / dt, torque += (cc - c) x f V / dt, summed in FP64 (each cell's
// terms in FP32 when the fields are FP32).
template <class R, int MK>
__device__ inline void add_force(const MeshDev& m, const SolidK& K, int gi, int gj, int gk, double alpha,
                                 const R* __restrict__ Ux, const R* __restrict__ Uy, const R* __restrict__ Uz,
                                 double dtINV, double* acc) {
    const long long c = cell_id(m, gi, gj, gk);
    const d3 cc = cell_centre<MK>(m, gi, gj, gk);
    const double V = cell_volume<MK>(m, gi, gj, gk);
    const d3 rv = {cc.x - K.c[0], cc.y - K.c[1], cc.z - K.c[2]};
    if constexpr (sizeof(R) == sizeof(float)) {
        // FP32 fields: the cell's force and torque in FP32 (the field's own precision), summed in FP64
        const float rx = (float)rv.x, ry = (float)rv.y, rz = (float)rv.z;
        const float wx = (float)K.w[0], wy = (float)K.w[1], wz = (float)K.w[2];
        const float usx = (float)K.v[0] + (wy * rz - wz * ry), usy = (float)K.v[1] + (wz * rx - wx * rz),
                    usz = (float)K.v[2] + (wx * ry - wy * rx);
        const float a = (float)alpha;
        const float fx = a * (Ux[c] - usx), fy = a * (Uy[c] - usy), fz = a * (Uz[c] - usz);
        const float sc = (float)(V * dtINV);
        acc[0] += (double)(fx * sc);
        acc[1] += (double)(fy * sc);
        acc[2] += (double)(fz * sc);
        acc[3] += (double)((ry * fz - rz * fy) * sc);
        acc[4] += (double)((rz * fx - rx * fz) * sc);
        acc[5] += (double)((rx * fy - ry * fx) * sc);
        return;
    }
    const d3 us = d3{K.v[0], K.v[1], K.v[2]} + cross(d3{K.w[0], K.w[1], K.w[2]}, rv);
    const d3 uf{(double)Ux[c], (double)Uy[c], (double)Uz[c]};
    const d3 f = alpha * (uf - us);
    const d3 tq = cross(rv, f);
    acc[0] += (f.x * V) * dtINV;
    acc[1] += (f.y * V) * dtINV;
    acc[2] += (f.z * V) * dtINV;
    acc[3] += (tq.x * V) * dtINV;
    acc[4] += (tq.y * V) * dtINV;
    acc[5] += (tq.z * V) * dtINV;
}

// One candidate cell: a cut cell's fraction and centre test (its code: the
// fraction, negative when the centre is outside), then the forcing (alpha = 1
// for ALL_INSIDE cells, solidcloud.cpp:408-410).
template <class R, int MK>
__device__ inline void cell_force(const MeshDev& m, const SolidK& K, int gi, int gj, int gk, bool cut,
                                  const R* __restrict__ Ux, const R* __restrict__ Uy, const R* __restrict__ Uz,
                                  R* code, double dtINV, double* acc) {
    double alpha = 1.0;
    if (cut) {
        const R a = cell_fraction<R, MK, false>(m, K, gi, gj, gk);  // finite solids only
        const bool cin = inside_world(K, cell_centre<MK>(m, gi, gj, gk));
        *code = cin ? a : -a;
        alpha = (double)a;
    }
    add_force<R, MK>(m, K, gi, gj, gk, alpha, Ux, Uy, Uz, dtINV, acc);
}

// Forcing, fractions and the force/torque reduction over the candidate lists:
// one workgroup per queued solid, largest tile class first.  SG: the learned
// surrogate gives the fraction of a cut cell (FP32, rectilinear mesh, 3-D,
// interior cells of quadric solids) unless its prediction lies in [sg_lo,
// sg_hi] while the centre gradient and the corner-difference gradient differ by
// more than the angle whose cosine is sg_cos; then, and for every other cell,
// the exact box fraction.
// The plain FP32 rectilinear pass fits 96 VGPRs without spills: five waves per SIMD.
template <class R, int MK, bool SG, int NT>
__global__ void __launch_bounds__(NT)
    __attribute__((amdgpu_waves_per_eu(!SG && MK == MK_RECT && sizeof(R) == 4 ? 5 : 1))) k_cells(MeshDev m, int ns, const SolidK* __restrict__ Ks,
                                                  const TileInfo* __restrict__ Ts, int* __restrict__ qmeta,
                                                  const int* __restrict__ qlist, const int* __restrict__ cand,
                                                  const int2* __restrict__ ncand, const R* __restrict__ Ux,
                                                  const R* __restrict__ Uy, const R* __restrict__ Uz,
                                                  R* __restrict__ codes, long long code_cap, double dt,
                                                  double rho_f, double* __restrict__ ft, const SgNet* __restrict__ sgnet,
                                                  float sg_lo, float sg_hi, float sg_cos, int* __restrict__ sg_count,
                                                  int part, int lm_cap) {
    __shared__ double red[6 * 16];
    constexpr int NSG = SG ? 1 : 0;
    __shared__ SgNet sn[NSG ? 1 : 1];
    __shared__ unsigned short feat[SG ? NT * kSgPad : 1];
    const int tid = threadIdx.x;
    // part 0: every item; 1: the large and medium tiles' items below lm_cap; 2: the rest
    const int n_lm = qmeta[2] + qmeta[1];
    int item = blockIdx.x;
    if (part == 1 && item >= min(n_lm, lm_cap)) return;
    if (part == 2) item += min(n_lm, lm_cap);
    if (item >= qmeta[0] + qmeta[1] + qmeta[2] + qmeta[3]) return;
    const double dtINV = 1.0 / dt;
    const int s = __builtin_amdgcn_readfirstlane(queue_item(qmeta, qlist, ns, item));
    const TileInfo T = Ts[s];
    if (T.off + (long long)T.n[0] * T.n[1] * T.n[2] > code_cap) return;
    const SolidK K = Ks[s];
    const int2 cnt = ncand[s];
    const int nin = cnt.x, ntot = cnt.x + cnt.y;
    const int nx = T.n[0], ny = T.n[1];
    const int* tl = cand + T.off;
    R* tcodes = codes + T.off;

    SgRegs sgr;
    bool sg_shape = false;
    if constexpr (SG) {
        for (int t = tid; t < (int)(sizeof(SgNet) / 4); t += NT)
            reinterpret_cast<unsigned*>(&sn[0])[t] = reinterpret_cast<const unsigned*>(sgnet)[t];
        __syncthreads();
        sg_load(sn[0], sgr, tid & 63);
        sg_shape = K.shape == SDFIBM_SHAPE_SPHERE || K.shape == SDFIBM_SHAPE_ELLIPSOID;
    }
    int n_net = 0;

    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int base = 0; base < ntot; base += NT) {  // the same trip count for every thread
        const int q = base + tid;
        const bool live = q < ntot;
        int i = 0, j = 0, k = 0;
        if (live) {
            const int e = tl[q];
            i = e & 63, j = (e >> 6) & 8191, k = e >> 19;
        }
        const bool cut = live && q >= nin;
        const int li = i + nx * (j + ny * k);
        const int gi = T.lo[0] + i, gj = T.lo[1] + j, gk = T.lo[2] + k;
        if constexpr (SG) {
            double alpha = 1.0;
            Phi8<float> h{0, 0, 0, 0, 0, 0, 0, 0};
            d3 lc{0, 0, 0};
            float g[3] = {0.f, 0.f, 0.f};
            bool ok = false;
            float ihx = 0.f, ihy = 0.f, ihz = 0.f;
            if (cut) {
                box_corner_phi(m, K, gi, gj, gk, h, lc);
                if (K.shape == SDFIBM_SHAPE_SPHERE) sphere_corner_phi(m, K, gi, gj, gk, h);  // as cell_fraction
                ihx = (float)(1.0 / (m.xs[gi + 1] - m.xs[gi]));
                ihy = (float)(1.0 / (m.ys[gj + 1] - m.ys[gj]));
                ihz = (float)(1.0 / (m.zs[gk + 1] - m.zs[gk]));
                ok = sg_shape && m.two_d == 0 && gi > 0 && gj > 0 && gk > 0 && quadric_grad(K, lc, g);
            }
            const float phi8[8] = {h.f0, h.f1, h.f2, h.f3, h.f4, h.f5, h.f6, h.f7};
            const float ih = cbrtf(ihx * ihy * ihz);
            unsigned short* fw = feat + (tid & ~63) * kSgPad;
            sg_features(phi8, ok ? ih : 0.f, g, feat + tid * kSgPad);
            float pred = 0.f;
            if (__any(ok)) pred = sg_infer(sn[0], sgr, fw, tid & 63);  // whole wave
            if (ok && pred >= sg_lo && pred <= sg_hi && corner_alignment(phi8, g, ihx, ihy, ihz) < sg_cos) ok = false;
            if (cut) {
                const double z0 = m.zs[gk], z1 = m.zs[gk + 1];
                // a prediction is clamped to [0, 1]: the code's sign carries the centre test
                const float a = ok ? fminf(fmaxf(pred, 0.f), 1.f)
                                   : box_fraction<float>(h, gi == 0, gj == 0, gk == 0, m.two_d != 0,
                                                         (float)((0.0 - z0) / (z1 - z0)));
                n_net += ok ? 1 : 0;
                const bool cin = inside_world(K, cell_centre<MK>(m, gi, gj, gk));
                tcodes[li] = R(cin ? a : -a);
                alpha = (double)a;
            }
            if (live) add_force<R, MK>(m, K, gi, gj, gk, alpha, Ux, Uy, Uz, dtINV, acc);
        } else {
            if (live) cell_force<R, MK>(m, K, gi, gj, gk, cut, Ux, Uy, Uz, tcodes + li, dtINV, acc);
        }
    }
    if constexpr (SG) {
        for (int off = 32; off >= 1; off >>= 1) n_net += __shfl_xor(n_net, off, 64);
        if ((tid & 63) == 0 && n_net > 0) atomicAdd(sg_count, n_net);
    }
    block_reduce<6>(acc, red);
    if (tid == 0)
        for (int t = 0; t < 6; ++t) ft[6 * s + t] = acc[t] * rho_f;
}

// Brick lists: one wave per solid, its lanes over the bricks its tile overlaps.
// Brick b's list is blist[b * kMaxPerBrick + t], t < count[b] (in no order).
template <int BX, int BY, int BZ>
__global__ void k_bin(int ns, const TileInfo* __restrict__ T, int nbx, int nby, int* __restrict__ count,
                      int* __restrict__ blist, int* __restrict__ err) {
    const int s = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= ns) return;
    const TileInfo t = T[s];
    if (!tile_is_finite(t.cls)) return;
    const int b0x = t.lo[0] / BX, b1x = (t.lo[0] + t.n[0] - 1) / BX;
    const int b0y = t.lo[1] / BY, b1y = (t.lo[1] + t.n[1] - 1) / BY;
    const int b0z = t.lo[2] / BZ, b1z = (t.lo[2] + t.n[2] - 1) / BZ;
    const int ex = b1x - b0x + 1, ey = b1y - b0y + 1, ez = b1z - b0z + 1;
    const int nb = ex * ey * ez;
    for (int q = lane; q < nb; q += 64) {
        const int bx = b0x + q % ex, by = b0y + (q / ex) % ey, bz = b0z + q / (ex * ey);
        const long long b = bx + (long long)nbx * (by + (long long)nby * bz);
        const int slot = atomicAdd(&count[b], 1);
        if (slot < kMaxPerBrick)
            blist[b * kMaxPerBrick + slot] = s;
        else
            atomicOr(err, EF_BRICK);
    }
}

// Brick culling for a plane (inside: local y < 0, local y = n . (p - c) with
// n = R(q) e_y): 0 = no vertex of the box can be inside, 1 = every vertex is
// inside, 2 = undecided.  The margin covers the rounding of the exact test.
__device__ inline int plane_box_state(const SolidK& K, const double* bb) {
    double Rm[3][3];
    rot_of(K.q, Rm);
    const double n0 = Rm[0][1], n1 = Rm[1][1], n2 = Rm[2][1];
    const double cx = 0.5 * (bb[0] + bb[1]) - K.c[0], cy = 0.5 * (bb[2] + bb[3]) - K.c[1],
                 cz = 0.5 * (bb[4] + bb[5]) - K.c[2];
    const double ext = fabs(n0) * 0.5 * (bb[1] - bb[0]) + fabs(n1) * 0.5 * (bb[3] - bb[2]) +
                       fabs(n2) * 0.5 * (bb[5] - bb[4]);
    const double mid = n0 * cx + n1 * cy + n2 * cz;
    const double tol = 1e-9 * (fabs(cx) + fabs(cy) + fabs(cz) + ext + fabs(K.c[0]) + fabs(K.c[1]) + fabs(K.c[2])) + 1e-300;
    if (mid - ext > tol) return 0;
    if (mid + ext < -tol) return 1;
    return 2;
}

// Cull state of a brick (cell ranges [i0, i1] x [j0, j1] x [k0, k1], vertex box
// bb) for a solid of the unbounded path: a plane by its half-space, a finite
// (oversized) solid by its tile box -- 0: no vertex inside, 2: may have some.
__device__ inline int unbounded_box_state(const SolidK& K, const TileInfo& T, const double* bb, int i0, int i1, int j0,
                                          int j1, int k0, int k1) {
    if (K.shape == SDFIBM_SHAPE_PLANE) return plane_box_state(K, bb);
    const bool hit = i0 < T.lo[0] + T.n[0] && i1 >= T.lo[0] && j0 < T.lo[1] + T.n[1] && j1 >= T.lo[1] &&
                     k0 < T.lo[2] + T.n[2] && k1 >= T.lo[2];
    return hit ? 2 : 0;
}

template <int MK, int BX, int BY, int BZ>
__device__ inline void brick_vertex_box(const MeshDev& m, const double* brick_box, long long b, int bx, int by, int bz,
                                        double* bb) {
    if constexpr (MK == MK_RECT) {
        const int i0 = bx * BX, j0 = by * BY, k0 = bz * BZ;
        const int i1 = min(i0 + BX, m.nx), j1 = min(j0 + BY, m.ny), k1 = min(k0 + BZ, m.nz);
        bb[0] = m.xs[i0];
        bb[1] = m.xs[i1];
        bb[2] = m.ys[j0];
        bb[3] = m.ys[j1];
        bb[4] = m.zs[k0];
        bb[5] = m.zs[k1];
    } else {
        for (int t = 0; t < 6; ++t) bb[t] = brick_box[6 * b + t];
    }
}

__device__ inline bool bit_get(const unsigned* bits, long long c) { return (bits[c >> 5] >> (c & 31)) & 1u; }

// Cells of the mesh with >= 1 inside vertex of a plane: bitmap, list, lowest index.
template <int MK, int BX, int BY, int BZ>
__global__ void __launch_bounds__(BX* BY* BZ)
    k_plane_active(MeshDev m, int nbx, int nby, const double* __restrict__ brick_box, const SolidK* __restrict__ Ks,
                   const TileInfo* __restrict__ Ts, int sp, unsigned* __restrict__ act, int* __restrict__ list,
                   int* __restrict__ nlist, long long cap, int* __restrict__ minidx) {
    const SolidK K = Ks[sp];
    const int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    const long long b = bx + (long long)nbx * (by + (long long)nby * bz);
    double bb[6];
    brick_vertex_box<MK, BX, BY, BZ>(m, brick_box, b, bx, by, bz, bb);
    const int st = unbounded_box_state(K, Ts[sp], bb, bx * BX, bx * BX + BX - 1, by * BY, by * BY + BY - 1, bz * BZ,
                                       bz * BZ + BZ - 1);
    if (st == 0) return;
    const int tid = threadIdx.x;
    const int i = bx * BX + tid % BX, j = by * BY + (tid / BX) % BY, k = bz * BZ + tid / (BX * BY);
    if (i >= m.nx || j >= m.ny || k >= m.nz) return;
    int nin = 8;
    if (st == 2) {
        d3 P[8];
        cell_vertices<MK>(m, i, j, k, P);
        nin = 0;
#pragma unroll
        for (int v = 0; v < 8; ++v) nin += inside_world(K, P[v]) ? 1 : 0;
    }
    if (nin == 0) return;
    const long long c = cell_id(m, i, j, k);
    atomicOr(&act[c >> 5], 1u << (c & 31));
    const int slot = atomicAdd(nlist, 1);
    if (slot < cap) list[slot] = (int)c;
    atomicMin(minidx, (int)c);
}

__device__ inline int uf_find(const int* L, int x) {
    int p = L[x];
    while (p != x) {
        x = p;
        p = L[x];
    }
    return x;
}

__global__ void k_ccl_init(const int* __restrict__ list, const int* __restrict__ nlist, int* __restrict__ L) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < *nlist) L[list[t]] = list[t];
}

// Union over face neighbours (+x, +y, +z) of active cells: hook the larger root
// under the smaller one.
__global__ void k_ccl_hook(MeshDev m, const int* __restrict__ list, const int* __restrict__ nlist,
                           const unsigned* __restrict__ act, int* L, int* changed) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *nlist) return;
    const int c = list[t];
    const int i = c % m.nx, j = (c / m.nx) % m.ny, k = c / (m.nx * m.ny);
    int nb[3];
    int nn = 0;
    if (i + 1 < m.nx) nb[nn++] = c + 1;
    if (j + 1 < m.ny) nb[nn++] = c + m.nx;
    if (k + 1 < m.nz) nb[nn++] = c + m.nx * m.ny;
    for (int q = 0; q < nn; ++q) {
        if (!bit_get(act, nb[q])) continue;
        int a = uf_find(L, c), b2 = uf_find(L, nb[q]);
        while (a != b2) {
            const int hi = a > b2 ? a : b2, lo = a > b2 ? b2 : a;
            const int old = atomicMin(&L[hi], lo);
            *changed = 1;
            if (old == hi) break;  // hi was a root and now points at lo
            a = uf_find(L, old);
            b2 = uf_find(L, lo);
        }
    }
}

__global__ void k_ccl_compress(const int* __restrict__ list, const int* __restrict__ nlist, int* L) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *nlist) return;
    const int c = list[t];
    L[c] = uf_find(L, c);
}

// Cells of the seed's component (seed: the nearest cell if active, else the
// lowest-index active cell) -> member bitmap of the plane.
__global__ void k_plane_member(const int* __restrict__ list, const int* __restrict__ nlist, const int* __restrict__ L,
                               const unsigned* __restrict__ act, const TileInfo* __restrict__ T, int solid,
                               const int* __restrict__ minidx, unsigned* __restrict__ member,
                               int* __restrict__ seed_used) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *nlist) return;
    const int seed = T[solid].seed;
    const int s = (seed >= 0 && bit_get(act, seed)) ? seed : *minidx;
    const int root = uf_find(L, s);
    const int c = list[t];
    if (uf_find(L, c) == root) atomicOr(&member[c >> 5], 1u << (c & 31));
    if (t == 0) *seed_used = s;
}

// Codes of a plane's member cells (ALL_INSIDE: +inf; cut: the fraction, negative
// when the centre is outside), computed once per plane labelling so that the
// field pass only reads them.  Cells that are not members are left as they are.
template <class R, int MK>
__global__ void __launch_bounds__(256) k_plane_alpha(MeshDev m, const int* __restrict__ list, const int* __restrict__ nlist,
                              const SolidK* __restrict__ Ks, int solid, const unsigned* __restrict__ member,
                              R* __restrict__ pcode) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *nlist) return;
    const int c = list[t];
    if (!bit_get(member, c)) return;
    const int i = c % m.nx, j = (c / m.nx) % m.ny, k = c / (m.nx * m.ny);
    const SolidK K = Ks[solid];
    d3 P[8];
    cell_vertices<MK>(m, i, j, k, P);
    int nin = 0;
#pragma unroll
    for (int v = 0; v < 8; ++v) nin += inside_world(K, P[v]) ? 1 : 0;
    R code = code_none<R>();
    if (nin == 8) {
        code = R(__builtin_huge_val());
    } else if (nin > 0) {
        const R a = cell_fraction<R, MK>(m, K, i, j, k);
        code = inside_world(K, cell_centre<MK>(m, i, j, k)) ? a : -a;
    }
    pcode[c] = code;
}

#include "big_tiles.h"
#include "poly_mesh.h"

// Field pass: one workgroup per brick of BX*BY*BZ cells, CPT cells per thread
// (the thread's cells are CPT layers apart along z: lane (ti, tj) of layer tk
// takes layers tk, tk + BZ/CPT, ...).  Each wave covers 64/BX whole x-rows of
// its layers; it skips, with scalar tests on the solid's tile record
// (scalar-cache loads), the listed solids whose tile misses its rows.
template <class R, int MK, int BX, int BY, int BZ, int CPT, bool HASP>
__global__ void __launch_bounds__(BX* BY* BZ / CPT)
    k_fields(MeshDev m, int nbx, int nby, const int* __restrict__ bcount, const int* __restrict__ blist,
             const SolidK* __restrict__ Ks, const TileInfo* __restrict__ Ts,
             const R* __restrict__ codes, long long code_cap, const int* __restrict__ planes, int nplanes,
             const unsigned* __restrict__ plane_member, long long member_words, const R* __restrict__ pcodes,
             const double* __restrict__ brick_box, const R* __restrict__ Ux, const R* __restrict__ Uy,
             const R* __restrict__ Uz, R* __restrict__ As, R* __restrict__ Fx, R* __restrict__ Fy,
             R* __restrict__ Fz, R* __restrict__ Tsf, int* __restrict__ Ct, double dt,
             double* __restrict__ plane_partial, long long nbricks, int* __restrict__ err) {
    constexpr int NT = BX * BY * BZ / CPT;
    constexpr int LZ = BZ / CPT;  // z distance between a thread's cells
    static_assert(BZ % CPT == 0 && 64 % BX == 0 && (BX * BY) % 64 == 0 && NT % 64 == 0,
                  "a wave covers whole x-rows of its layers");
    __shared__ int raw[kMaxPerBrick];
    __shared__ int lst[kMaxPerBrick];
    __shared__ double red[6 * 16];

    // grid (nbx, nby, nbz): no integer division for the brick coordinates
    const int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    const long long b = bx + (long long)nbx * (by + (long long)nby * bz);
    const int tid = threadIdx.x;
    const int ti = tid % BX, tj = (tid / BX) % BY, tk = tid / (BX * BY);
    const int i = bx * BX + ti, j = by * BY + tj;
    int kq[CPT];
    bool valid[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        kq[q] = bz * BZ + tk + q * LZ;
        valid[q] = i < m.nx && j < m.ny && kq[q] < m.nz;
    }

    // list entries loaded together with the count (the first 32, one cache line, speculatively)
    constexpr int kSpec = 32;
    const int* bl = blist + b * kMaxPerBrick;
    const int spec = tid < kSpec ? bl[tid] : 0;
    const int cnt = min(bcount[b], kMaxPerBrick);  // k_bin flagged an overflow
    // the cells' velocity, centre and volume (used if a listed solid or a plane takes them);
    // a cell's terms in the fields' precision A (FP32 fields: FP32 terms and sums of the few
    // solids sharing the cell; the planes' force and torque partials stay FP64)
    using A = typename std::conditional<sizeof(R) == sizeof(float), float, double>::type;
    A ufx[CPT], ufy[CPT], ufz[CPT];
    d3 cc[CPT];
    double V[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        ufx[q] = ufy[q] = ufz[q] = A(0);
        cc[q] = d3{0, 0, 0};
        V[q] = 0.0;
        if (valid[q]) {  // issued with the list loads, before the count is known
            const long long c = cell_id(m, i, j, kq[q]);
            ufx[q] = A(Ux[c]);
            ufy[q] = A(Uy[c]);
            ufz[q] = A(Uz[c]);
            cc[q] = cell_centre<MK>(m, i, j, kq[q]);
            V[q] = cell_volume<MK>(m, i, j, kq[q]);
        }
    }
    if (tid < min(cnt, kSpec)) raw[tid] = spec;
    for (int t = kSpec + tid; t < cnt; t += NT) raw[t] = bl[t];
    __syncthreads();
    for (int t = tid; t < cnt; t += NT) {  // rank sort (solid ids are distinct)
        const int v = raw[t];
        int rank = 0;
        for (int u = 0; u < cnt; ++u) rank += raw[u] < v;
        lst[rank] = v;
    }
    __syncthreads();
    // the wave's cell rows (uniform): x in [wi0, wi1], y in [wj0, wj1], layers wk + q LZ
    const int w0 = (tid >> 6) << 6;
    const int wk = bz * BZ + w0 / (BX * BY);
    const int wj0 = by * BY + (w0 % (BX * BY)) / BX, wj1 = wj0 + 64 / BX - 1;
    const int wi0 = bx * BX, wi1 = wi0 + BX - 1;

    // brick vertex box, for plane culling
    double bb[6];
    if constexpr (HASP) brick_vertex_box<MK, BX, BY, BZ>(m, brick_box, b, bx, by, bz, bb);

    const double dtINV = 1.0 / dt;
    const A dti = A(dtINV);
    A sA[CPT], sT[CPT], sF[CPT][3];
    int ct[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        sA[q] = sT[q] = sF[q][0] = sF[q][1] = sF[q][2] = A(0);
        ct[q] = 0;
    }
    // contribute: a solid's share in cell q; a plane also adds its force and torque to pa
    auto contribute = [&](int q, const double* cvw, double alpha, int ctype, double* pa) {
        const A rx = A(cc[q].x - cvw[0]), ry = A(cc[q].y - cvw[1]), rz = A(cc[q].z - cvw[2]);
        const A wx = A(cvw[6]), wy = A(cvw[7]), wz = A(cvw[8]);
        const A usx = A(cvw[3]) + (wy * rz - wz * ry), usy = A(cvw[4]) + (wz * rx - wx * rz),
                usz = A(cvw[5]) + (wx * ry - wy * rx);
        const A a = A(alpha);
        const A fx = a * (ufx[q] - usx), fy = a * (ufy[q] - usy), fz = a * (ufz[q] - usz);
        sA[q] += a;
        sT[q] += a;
        sF[q][0] += fx * dti;
        sF[q][1] += fy * dti;
        sF[q][2] += fz * dti;
        ct[q] = ctype;
        if (HASP && pa != nullptr) {
            const A tx = ry * fz - rz * fy, ty = rz * fx - rx * fz, tz = rx * fy - ry * fx;
            double* p = pa;
            p[0] += ((double)fx * V[q]) * dtINV;
            p[1] += ((double)fy * V[q]) * dtINV;
            p[2] += ((double)fz * V[q]) * dtINV;
            p[3] += ((double)tx * V[q]) * dtINV;
            p[4] += ((double)ty * V[q]) * dtINV;
            p[5] += ((double)tz * V[q]) * dtINV;
        }
    };

    if (!HASP) nplanes = 0;
    // The listed solids and the planes are taken in ascending solid index.  The listed solids
    // between two planes form a segment; a segment is walked in batches of kBatch solids whose
    // codes are loaded together (independent loads), then consumed in an unrolled pass.
    constexpr int kBatch = 8 / CPT;
    int ia = 0;
    for (int ip = 0;; ++ip) {
        const bool plane_next = HASP && ip < nplanes;
        const int sp = plane_next ? planes[ip] : 0x7fffffff;
        int iend = cnt;
        if (plane_next) {
            iend = ia;
            while (iend < cnt && __builtin_amdgcn_readfirstlane(lst[iend]) < sp) ++iend;
        }
        for (int cb0 = ia; cb0 < iend; cb0 += kBatch) {
            R cb[kBatch][CPT];
            int sid[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
#pragma unroll
                for (int q = 0; q < CPT; ++q) cb[u][q] = code_none<R>();
                sid[u] = 0;
                const int a = cb0 + u;
                if (a < iend) {
                    const int sq = __builtin_amdgcn_readfirstlane(lst[a]);
                    sid[u] = sq;
                    const TileInfo T = Ts[sq];  // uniform: scalar loads
                    const bool hit = T.lo[0] <= wi1 && T.lo[0] + T.n[0] > wi0 && T.lo[1] <= wj1 &&
                                     T.lo[1] + T.n[1] > wj0 &&
                                     T.off + (long long)T.n[0] * T.n[1] * T.n[2] <= code_cap;
                    const unsigned li = (unsigned)(i - T.lo[0]), lj = (unsigned)(j - T.lo[1]);
                    const bool in_xy = li < (unsigned)T.n[0] && lj < (unsigned)T.n[1];
#pragma unroll
                    for (int q = 0; q < CPT; ++q) {
                        const int lk = wk + q * LZ - T.lo[2];
                        if (hit && lk >= 0 && lk < T.n[2] && valid[q] && in_xy)
                            cb[u][q] = codes[T.off + (long long)(li + T.n[0] * (lj + T.n[1] * lk))];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                if (cb0 + u >= iend) break;
                bool any = false;
#pragma unroll
                for (int q = 0; q < CPT; ++q) any = any || cb[u][q] == cb[u][q];
                if (!__any(any)) continue;  // no candidate in the wave (NaN codes)
                const int sa = sid[u];
                const SolidK& Ka = Ks[sa];
                const double cvw[9] = {Ka.c[0], Ka.c[1], Ka.c[2], Ka.v[0], Ka.v[1], Ka.v[2], Ka.w[0], Ka.w[1], Ka.w[2]};
#pragma unroll
                for (int q = 0; q < CPT; ++q) {
                    const R cq = cb[u][q];
                    if (cq != cq) continue;
                    // one instruction stream for ALL_INSIDE (alpha 1) and cut cells
                    const bool all_in = isinf(cq);
                    contribute(q, cvw, all_in ? 1.0 : (double)fabs(cq), all_in ? sa + 4 : (signbit(cq) ? 3 : 2),
                               nullptr);
                }
            }
        }
        ia = iend;
        if (!plane_next) break;
        if constexpr (HASP) {
            const int slot = ip;
            const SolidK K = Ks[sp];
            // plane: only cells of the seed's component (member bitmap); cull the brick first
            // (uniform over the workgroup: the brick's vertex box)
            const int st = unbounded_box_state(K, Ts[sp], bb, bx * BX, bx * BX + BX - 1, by * BY, by * BY + BY - 1,
                                               bz * BZ, bz * BZ + BZ - 1);
            double* part = plane_partial + ((long long)slot * nbricks + b) * 6;
            if (st == 0) {
                if (tid == 0)
                    for (int t = 0; t < 6; ++t) part[t] = 0.0;
                continue;  // no vertex of the brick can be inside
            }
            double pa[6] = {0, 0, 0, 0, 0, 0};
            const double cvw[9] = {K.c[0], K.c[1], K.c[2], K.v[0], K.v[1], K.v[2], K.w[0], K.w[1], K.w[2]};
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                if (!valid[q]) continue;
                const int k = kq[q];
                const long long c = cell_id(m, i, j, k);
                if (!bit_get(plane_member + (long long)slot * member_words, c)) continue;
                const R pc = pcodes[(long long)slot * m.nx * m.ny * m.nz + c];  // k_plane_alpha
                if (pc != pc) continue;
                const bool all_in = isinf(pc);
                contribute(q, cvw, all_in ? 1.0 : (double)fabs(pc), all_in ? sp + 4 : (signbit(pc) ? 3 : 2), pa);
            }
            block_reduce<6>(pa, red);  // the brick's share of the plane's force and torque
            if (tid == 0)
                for (int t = 0; t < 6; ++t) part[t] = pa[t];
        }
    }

#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        if (!valid[q]) continue;
        const long long c = cell_id(m, i, j, kq[q]);
        // streaming stores: the fields are not read again in this step
        __builtin_nontemporal_store(R(sA[q] < A(1) ? sA[q] : A(1)), &As[c]);
        __builtin_nontemporal_store(R(sF[q][0]), &Fx[c]);
        __builtin_nontemporal_store(R(sF[q][1]), &Fy[c]);
        __builtin_nontemporal_store(R(sF[q][2]), &Fz[c]);
        __builtin_nontemporal_store(R(sT[q]), &Tsf[c]);
        __builtin_nontemporal_store(ct[q], &Ct[c]);
    }
}

__global__ void k_plane_reduce(const double* __restrict__ partial, long long nbricks, const int* __restrict__ planes,
                               double rho_f, double* __restrict__ ft) {
    __shared__ double red[6 * 16];
    const int p = blockIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    // thread t sums bricks t, t + blockDim, ... (neighbouring threads read neighbouring bricks)
    const double* base = partial + (long long)p * nbricks * 6;
    for (long long b = threadIdx.x; b < nbricks; b += blockDim.x)
        for (int t = 0; t < 6; ++t) acc[t] += base[b * 6 + t];
    block_reduce<6>(acc, red);
    if (threadIdx.x == 0)
        for (int t = 0; t < 6; ++t) ft[6 * planes[p] + t] = acc[t] * rho_f;
}

// Candidate codes of a plane over the whole mesh (parity / diagnostics only).
template <class R, int MK>
__global__ void __launch_bounds__(256) k_plane_codes(MeshDev m, const SolidK* __restrict__ Ks, int sp, const unsigned* __restrict__ member,
                              R* __restrict__ out) {
    const SolidK K = Ks[sp];
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long N = (long long)m.nx * m.ny * m.nz;
    if (c >= N) return;
    if (!bit_get(member, c)) {
        out[c] = code_none<R>();
        return;
    }
    int i, j, k;
    cell_ijk(m, c, i, j, k);
    d3 P[8];
    cell_vertices<MK>(m, i, j, k, P);
    int nin = 0;
    for (int v = 0; v < 8; ++v) nin += inside_world(K, P[v]) ? 1 : 0;
    if (nin == 0) {
        out[c] = code_none<R>();
    } else if (nin == 8) {
        out[c] = R(__builtin_huge_val());
    } else {
        const R a = cell_fraction<R, MK>(m, K, i, j, k);
        out[c] = inside_world(K, cell_centre<MK>(m, i, j, k)) ? a : -a;
    }
}

// U -= Fs*dt (src/main.cpp:70), FP64 arithmetic, stored in R.
template <class R>
__global__ void k_update_velocity(long long n, R* __restrict__ Ux, R* __restrict__ Uy, R* __restrict__ Uz,
                                  const R* __restrict__ Fx, const R* __restrict__ Fy, const R* __restrict__ Fz,
                                  double dt) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    Ux[c] = R((double)Ux[c] - (double)Fx[c] * dt);
    Uy[c] = R((double)Uy[c] - (double)Fy[c] * dt);
    Uz[c] = R((double)Uz[c] - (double)Fz[c] * dt);
}

// fixInternal (src/solidcloud.cpp:288-301): where Ct >= 4 the cell lies wholly in
// solid Ct-4 and U takes that solid's rigid velocity at the cell centre.
template <class R, int MK>
__global__ void __launch_bounds__(256) k_fix_internal(MeshDev m, const int* __restrict__ Ct, const sdfibm_solid* __restrict__ S, int ns,
                               R* __restrict__ Ux, R* __restrict__ Uy, R* __restrict__ Uz) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long N = (long long)m.nx * m.ny * m.nz;
    if (c >= N) return;
    const int t = Ct[c];
    if (t < 4 || t - 4 >= ns) return;
    const sdfibm_solid& s = S[t - 4];
    int i, j, k;
    cell_ijk(m, c, i, j, k);
    const d3 cc = cell_centre<MK>(m, i, j, k);
    const d3 r{cc.x - s.center[0], cc.y - s.center[1], cc.z - s.center[2]};
    const d3 u = d3{s.vel[0], s.vel[1], s.vel[2]} + cross(d3{s.omega[0], s.omega[1], s.omega[2]}, r);
    Ux[c] = R(u.x);
    Uy[c] = R(u.y);
    Uz[c] = R(u.z);
}

// U -= Fs*dt then fixInternal, in one pass over the cells, for the coupled loop
// whose order is interact, U -= Fs*dt, evolve, fixInternal: evolve reads no
// velocity field, so the two cell passes can both follow it.  A cell wholly in
// a solid takes that solid's rigid velocity (its U - Fs*dt would be
// overwritten); every other cell gets k_update_velocity's arithmetic.
template <class R, int MK>
__global__ void __launch_bounds__(256) k_update_fix(MeshDev m, const int* __restrict__ Ct,
                                                    const sdfibm_solid* __restrict__ S, int ns, R* __restrict__ Ux,
                                                    R* __restrict__ Uy, R* __restrict__ Uz, const R* __restrict__ Fx,
                                                    const R* __restrict__ Fy, const R* __restrict__ Fz, double dt,
                                                    const int* __restrict__ gate) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long N = (long long)m.nx * m.ny * m.nz;
    if (c >= N || (gate && *gate)) return;
    const int t = Ct[c];
    if (t < 4 || t - 4 >= ns) {
        Ux[c] = R((double)Ux[c] - (double)Fx[c] * dt);
        Uy[c] = R((double)Uy[c] - (double)Fy[c] * dt);
        Uz[c] = R((double)Uz[c] - (double)Fz[c] * dt);
        return;
    }
    const sdfibm_solid& s = S[t - 4];
    int i, j, k;
    cell_ijk(m, c, i, j, k);
    const d3 cc = cell_centre<MK>(m, i, j, k);
    const d3 r{cc.x - s.center[0], cc.y - s.center[1], cc.z - s.center[2]};
    const d3 u = d3{s.vel[0], s.vel[1], s.vel[2]} + cross(d3{s.omega[0], s.omega[1], s.omega[2]}, r);
    Ux[c] = R(u.x);
    Uy[c] = R(u.y);
    Uz[c] = R(u.z);
}

// SolidCloud::evolve (solidcloud.cpp:521-562), one thread per solid.
struct Vec3 {
    double v[3];
};
__global__ void __launch_bounds__(64) k_evolve(sdfibm_solid* __restrict__ S, SolidDyn* __restrict__ D, const double* __restrict__ ft,
                         int ns, int n_sub, double dt, Vec3 g, double rho_f, const int* __restrict__ gate) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns || (gate && *gate)) return;
    sdfibm_solid sol = S[s];
    SolidDyn d = D[s];
    evolve_solid(sol, d, ft + 6LL * s, n_sub, dt, g.v, rho_f);
    S[s] = sol;
    D[s] = d;
}

// Uniform grid of the contact search: cubic cells of twice the largest
// bounding radius over the mesh's bounding box (the grid the reference sizes
// UGrid(bbox, 2 m_radiusB) for, ugrid.cpp:5-22); centres outside the box fall
// into the border cells, which keeps every pair closer than one cell size in
// neighbouring cells.
struct ContactGrid {
    double lo[3];
    double inv_cell;
    int n[3];
};

__device__ inline int contact_cell(const ContactGrid& g, const double* c, int* ijk) {
    for (int a = 0; a < 3; ++a) {
        const double t = floor((c[a] - g.lo[a]) * g.inv_cell);
        ijk[a] = t >= 0.0 ? (t < (double)(g.n[a] - 1) ? (int)t : g.n[a] - 1) : 0;  // NaN -> 0
    }
    return ijk[0] + g.n[0] * (ijk[1] + g.n[1] * ijk[2]);
}

// Cell lists of the finite solids (bounding radius > 0): head per cell, next per solid.
__global__ void __launch_bounds__(64) k_contact_insert(const sdfibm_solid* __restrict__ S, const double* __restrict__ rb, int ns,
                                 ContactGrid g, int* __restrict__ head, int* __restrict__ next,
                                 const int* __restrict__ gate) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns || !(rb[s] > 0) || (gate && *gate)) return;
    int ijk[3];
    const int cell = contact_cell(g, S[s].center, ijk);
    next[s] = atomicExch(head + cell, s);
}

// One sub-step of SolidCloud::evolve with contacts, one thread per solid:
// loads, contact forces from the state at the start of the sub-step (Sin, cell
// lists head/next), Solid::move into Sout, then (unless it is the last
// sub-step) the moved solid goes into the next sub-step's lists (head_ins,
// next_ins).  The contacts of a solid are summed in ascending partner index
// (finite solids), then the planes in list order -- the sum does not depend on
// the order the cell lists were built in.  Three head arrays and two next
// arrays rotate over the sub-steps; this launch also resets the head array the
// sub-step after the next one fills (the previous sub-step read it).
// One sub-step of one solid with contacts (see k_contact_step); returns the
// moved solid in sol and the updated dynamics in d.
__device__ inline void contact_substep(const sdfibm_solid* __restrict__ Sin, int s, sdfibm_solid& sol, SolidDyn& d,
                                       const double* __restrict__ ft, const double* __restrict__ rb, ContactGrid g,
                                       const int* __restrict__ head, const int* __restrict__ next,
                                       const int* __restrict__ planes, int nplanes, ContactLaw law, double dts,
                                       Vec3 gv, double rho_f, int last) {
    const BodyConst b = body_const(sol, d);
    const double* f6 = ft + 6LL * s;
    const d3 ff{f6[0], f6[1], f6[2]}, fq{f6[3], f6[4], f6[5]};
    d3 force, torque;
    substep_loads(d, ff, fq, d3{gv.v[0], gv.v[1], gv.v[2]}, rho_f, b, force, torque);
    const double r = rb[s];
    if (r > 0) {
        d3 fc{0, 0, 0};
        int ijk[3];
        contact_cell(g, sol.center, ijk);
        // the heads of the 27 neighbour cells, loaded together (independent loads)
        int h27[27];
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const int ii = ijk[0] + t % 3 - 1, jj = ijk[1] + (t / 3) % 3 - 1, kk = ijk[2] + t / 9 - 1;
            const bool in = ii >= 0 && ii < g.n[0] && jj >= 0 && jj < g.n[1] && kk >= 0 && kk < g.n[2];
            h27[t] = in ? head[ii + g.n[0] * (jj + g.n[1] * kk)] : -1;
        }
        int prev = -1;
        for (;;) {  // next partner in contact: the smallest index above prev
            int best = INT_MAX;
#pragma unroll
            for (int t = 0; t < 27; ++t)
                for (int j = h27[t]; j >= 0; j = next[j])
                    if (j > prev && j < best && j != s && pair_overlap(sol, r, Sin[j], rb[j]) > 0) best = j;
            if (best == INT_MAX) break;
            fc = fc + pair_force(sol, r, Sin[best], rb[best], law);
            prev = best;
        }
        for (int p = 0; p < nplanes; ++p) fc = fc + plane_force(sol, r, Sin[planes[p]], law);
        force = force + fc;
    }
    move_solid(sol, d, b, force, torque, dts);
    if (last) store_old_force(d, ff, fq);
}

// One sub-step of SolidCloud::evolve with contacts, one thread per solid:
// loads, contact forces from the state at the start of the sub-step (Sin, cell
// lists head/next), Solid::move into Sout, then (unless it is the last
// sub-step) the moved solid goes into the next sub-step's lists (head_ins,
// next_ins).  The contacts of a solid are summed in ascending partner index
// (finite solids), then the planes in list order -- the sum does not depend on
// the order the cell lists were built in.  Three head arrays and two next
// arrays rotate over the sub-steps; this launch also resets the head array the
// sub-step after the next one fills (the previous sub-step read it).
__global__ void __launch_bounds__(64) k_contact_step(const sdfibm_solid* __restrict__ Sin, sdfibm_solid* __restrict__ Sout,
                               SolidDyn* __restrict__ D, const double* __restrict__ ft, const double* __restrict__ rb,
                               int ns, ContactGrid g, const int* __restrict__ head, const int* __restrict__ next,
                               int* __restrict__ head_ins, int* __restrict__ next_ins, int* __restrict__ head_reset,
                               int ncells, const int* __restrict__ planes, int nplanes, ContactLaw law, double dts,
                               Vec3 gv, double rho_f, int last, const int* __restrict__ gate) {
    if (gate && *gate) return;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (int t = tid; t < ncells; t += gridDim.x * blockDim.x) head_reset[t] = -1;
    if (tid >= ns) return;
    const int s = tid;
    sdfibm_solid sol = Sin[s];
    SolidDyn d = D[s];
    contact_substep(Sin, s, sol, d, ft, rb, g, head, next, planes, nplanes, law, dts, gv, rho_f, last);
    Sout[s] = sol;
    D[s] = d;
    if (!last && rb[s] > 0) {
        int ijk[3];
        const int cell = contact_cell(g, sol.center, ijk);
        next_ins[s] = atomicExch(head_ins + cell, s);
    }
}

// All sub-steps of an evolve with contacts in one workgroup (few solids, small
// grid): per sub-step build the cell lists, move every solid, clear the lists,
// with workgroup barriers between the phases instead of kernel boundaries.
// The same per-solid arithmetic and partner order as k_contact_step.
template <int NT>
__global__ void __launch_bounds__(NT) k_contact_evolve_block(
    sdfibm_solid* __restrict__ S0, sdfibm_solid* __restrict__ S1, SolidDyn* __restrict__ D,
    const double* __restrict__ ft, const double* __restrict__ rb, int ns, ContactGrid g, int* __restrict__ head,
    int* __restrict__ next, int ncells, const int* __restrict__ planes, int nplanes, ContactLaw law, double dts,
    Vec3 gv, double rho_f, int n_sub, const int* __restrict__ gate) {
    if (gate && *gate) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int t = tid; t < ncells; t += nt) head[t] = -1;
    __syncthreads();
    for (int it = 0; it < n_sub; ++it) {
        const sdfibm_solid* Sin = (it & 1) ? S1 : S0;
        sdfibm_solid* Sout = (it & 1) ? S0 : S1;
        for (int s = tid; s < ns; s += nt)
            if (rb[s] > 0) {
                int ijk[3];
                const int cell = contact_cell(g, Sin[s].center, ijk);
                next[s] = atomicExch(head + cell, s);
            }
        __syncthreads();
        for (int s = tid; s < ns; s += nt) {
            sdfibm_solid sol = Sin[s];
            SolidDyn d = D[s];
            contact_substep(Sin, s, sol, d, ft, rb, g, head, next, planes, nplanes, law, dts, gv, rho_f,
                            it == n_sub - 1);
            Sout[s] = sol;
            D[s] = d;
        }
        __syncthreads();
        for (int t = tid; t < ncells; t += nt) head[t] = -1;
        __syncthreads();
    }
}

// Plane (and oversized) solids whose pose differs from the one their cell sets
// were last labelled for: bit p of *dirty; the cached pose is updated.  A pose
// that is bitwise the same keeps the member bitmaps and codes (NaN in the cache
// forces a labelling).
__global__ void k_plane_pose(const sdfibm_solid* __restrict__ S, const int* __restrict__ planes, int np,
                             double* __restrict__ cache, int* __restrict__ dirty) {
    const int p = threadIdx.x;
    if (p == 0) dirty[0] = 0;
    __syncthreads();
    if (p >= np) return;
    const sdfibm_solid& s = S[planes[p]];
    bool same = true;
    for (int a = 0; a < 3; ++a) same = same && s.center[a] == cache[8 * p + a];
    for (int a = 0; a < 4; ++a) same = same && s.q[a] == cache[8 * p + 3 + a];
    if (!same) {
        for (int a = 0; a < 3; ++a) cache[8 * p + a] = s.center[a];
        for (int a = 0; a < 4; ++a) cache[8 * p + 3 + a] = s.q[a];
        atomicOr(dirty, 1 << p);
    }
}

// Graph replays: keep every step's error flags, and flag a plane solid whose
// pose moved away from the one its membership bitmaps were computed for (the
// quaternion normalisation of Solid::move may change a fixed plane's pose by
// rounding; that much is accepted).
__global__ void k_flag_or(const int* __restrict__ err, int* __restrict__ sticky) {
    if (threadIdx.x == 0 && err[0] != 0) sticky[0] |= err[0];
}

__global__ void k_plane_static(const sdfibm_solid* __restrict__ S, const int* __restrict__ planes, int np,
                               const double* __restrict__ pose, int* __restrict__ sticky) {
    const int p = threadIdx.x;
    if (p >= np) return;
    const sdfibm_solid& s = S[planes[p]];
    bool same = true;
    for (int a = 0; a < 3; ++a) same = same && fabs(s.center[a] - pose[7 * p + a]) <= 1e-12 * (1.0 + fabs(pose[7 * p + a]));
    for (int a = 0; a < 4; ++a) same = same && fabs(s.q[a] - pose[7 * p + 3 + a]) <= 1e-12;
    if (!same) atomicOr(sticky, (int)EF_PLANE_MOVED);
}

template <class R>
__global__ void k_aos_to_soa(const double* __restrict__ in, long long n, R* __restrict__ x, R* __restrict__ y,
                             R* __restrict__ z) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    x[c] = R(in[3 * c]);
    y[c] = R(in[3 * c + 1]);
    z[c] = R(in[3 * c + 2]);
}

template <class R>
__global__ void k_fields_to_f64(const R* __restrict__ As, const R* __restrict__ Fx, const R* __restrict__ Fy,
                                const R* __restrict__ Fz, const R* __restrict__ Tsf, const int* __restrict__ Ct,
                                long long n, double* __restrict__ oA, double* __restrict__ oF,
                                double* __restrict__ oT, double* __restrict__ oC) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    oA[c] = (double)As[c];
    oF[3 * c] = (double)Fx[c];
    oF[3 * c + 1] = (double)Fy[c];
    oF[3 * c + 2] = (double)Fz[c];
    oT[c] = (double)Tsf[c];
    oC[c] = (double)Ct[c];
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
constexpr int kBX3 = 32, kBY3 = 4, kBZ3 = 4, kCPT3 = 2;  // brick of a 3-D mesh, cells per thread
constexpr int kBX2 = 32, kBY2 = 8, kBZ2 = 1;  // brick of a one-cell-thick (2-D) mesh

struct sdfibm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipStream_t aux = nullptr;                   // second stream of the tile classification
    hipEvent_t fork = nullptr, join = nullptr;  // its dependencies on the main stream
    hipEvent_t joinc = nullptr;                 // ... the classification of its tiles done
    int* h_flags = nullptr;  // pinned, mapped host words written by k_flags_out: [0] error flags, [1] moved planes
    int* d_hflags = nullptr;  // its device address
    int kind = 0, prec = 0, two_d = 0;
    int nx = 0, ny = 0, nz = 0;
    long long N = 0;
    double rho_f = 1.0;
    int n_cu = 256;
    MeshDev m{};   // device pointers
    MeshDev hm{};  // host pointers (tile estimates)
    std::vector<double> h_xs, h_ys, h_zs;
    double box_lo[3] = {0, 0, 0}, box_hi[3] = {0, 0, 0};  // the mesh's bounding box (contact grid)
    // device buffers
    std::vector<void*> allocs;
    void* U[3] = {nullptr, nullptr, nullptr};
    void* As = nullptr;
    void* Fs[3] = {nullptr, nullptr, nullptr};
    void* Tsf = nullptr;
    int* Ct = nullptr;
    double* brick_box = nullptr;
    int bx = 0, by = 0, bz = 0;
    int nbx = 0, nby = 0, nbz = 0;
    long long nbricks = 0;
    int* bcount = nullptr;          // solids per brick
    long long* scan_tmp = nullptr;  // tile sums of the scans
    int* qmeta = nullptr;           // tile-class work queues: counts [0..2]
    int* qlist = nullptr;
    int* err = nullptr;
    // solids
    long long ns = 0;
    bool have_solids = false, have_step = false, have_ct = false;
    std::vector<sdfibm_solid> h_solids;
    sdfibm_solid* d_solids = nullptr;
    SolidK* d_K = nullptr;
    TileInfo* d_T = nullptr;
    long long* d_tile_total = nullptr;  // the step's tile cells (k_setup's running total)
    long long solid_cap = 0;
    void* d_codes = nullptr;
    int* d_cand = nullptr;  // candidate lists, same offsets as the codes
    long long code_cap = 0;
    int2* d_ncand = nullptr;  // per solid: ALL_INSIDE and cut candidates
    int* d_blist = nullptr;  // brick lists, kMaxPerBrick slots per brick
    double* d_ft = nullptr;
    SolidDyn* d_dyn = nullptr;  // rigid-body state (set_dynamics)
    SgNet* d_sg = nullptr;      // learned fraction surrogate (set_surrogate), null = exact fractions
    int* d_sg_count = nullptr;  // cut cells the surrogate served in the last interact
    float sg_lo = 0.02f, sg_hi = 0.98f, sg_cos = 0.8660254f;
    long long dyn_n = -1;       // solids the dynamics were set for
    long long dyn_cap = 0;
    // contacts (set_contact): law, bounding radii, sub-step solid buffer, cell lists
    ContactLaw law{0.0, 0.0};
    double* d_rb = nullptr;
    sdfibm_solid* d_solids_alt = nullptr;
    int* d_chead = nullptr;  // three head arrays of ncells each
    // d_cnext: two next arrays of contact_cap each
    int* d_cnext = nullptr;
    long long contact_cap = 0, chead_cap = 0;
    bool contact_ready = false;  // contact buffers sized for the current solids (contact_prepare)
    bool small_tiles = false;  // mean tile of the finite solids under 512 cells (set_solids)
    ContactGrid cgrid{};
    int ncells_c = 0;
    // the coupled step captured as a HIP graph (run_steps)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph_i = nullptr;  // interact alone, captured (interact_steps)
    hipGraphExec_t gexec_i = nullptr;
    double graph_i_dt = 0.0;
    double graph_dt = 0.0, graph_g[3] = {0, 0, 0};
    int* d_sticky = nullptr;
    double* d_plane_pose = nullptr;  // center + q of each plane at capture
    std::vector<int> h_planes;
    int* d_planes = nullptr;
    unsigned char* d_unb = nullptr;  // per solid: 1 = unbounded path (planes and oversized solids)
    double* d_plane_partial = nullptr;
    // plane components: active bitmap, active list, union-find labels, member bitmaps
    unsigned* d_pact = nullptr;
    int* d_plist = nullptr;
    int* d_plabel = nullptr;
    int* d_pscalars = nullptr;  // [0] list size, [1] lowest active index, [2] changed flag, [3] moved planes
    unsigned* d_pmember = nullptr;
    void* d_pcode = nullptr;  // per plane: codes of its member cells (N each)
    double* d_plane_cache = nullptr;  // per plane: the pose its cell sets were labelled for (8 doubles)
    int* d_pseed = nullptr;
    long long member_words = 0;
    double* d_ustage = nullptr;  // set_velocity's AoS staging buffer, N * 3 doubles (first call)
    double* stage = nullptr;     // this device's address of `scratch` (resolved by sdfibm_create)
    // CSG shapes (set_csg_shapes): the caller's records and their derived constants
    std::vector<sdfibm_csg_shape> h_csg_raw;
    std::vector<CsgK> h_csg;
    CsgK* d_csg = nullptr;
    // big-tile pass (tiles beyond the on-chip tile pass): launched when some solid can have one
    bool big_on = false;
    BigMeta* d_big_meta = nullptr;  // kBigMaxSolids
    long long* d_big_tot = nullptr;  // kBigTotN
    bool phase_timing = false;  // HIP events between the step's phases (sdfibm_set_phase_timing)
    int lm_grid = 0;  // workgroups of the cut-cell pass's 256-thread part (solids whose tile can be medium or big)
    // general polyhedral meshes (SDFIBM_MESH_POLY): the mesh, per-solid, per-visit, per-pair
    // and per-cell buffers of the step
    PolyDev pm{};
    int poly_reach = 0;  // bins a cell's points reach beyond its own
    double* g_box = nullptr;
    int* g_seed = nullptr;
    int* g_first = nullptr;
    long long* g_vcnt = nullptr;
    long long* g_voff = nullptr;
    long long* g_prange = nullptr;
    long long* g_ccoff = nullptr;  // per solid: its first visit chunk (ns + 1)
    long long* g_pnch = nullptr;   // per solid: its pair chunks
    long long* g_pcoff = nullptr;  // per solid: its first pair chunk (ns + 1)
    int* g_csol = nullptr;         // visit chunk -> solid
    long long* g_ccnt_v = nullptr; // visit chunk -> its cells
    long long* g_cvoff = nullptr;  // visit chunk -> its first visit
    int* g_psol = nullptr;         // pair chunk -> solid
    double* g_part = nullptr;      // pair chunk -> six partial sums and two counts
    long long g_csol_cap = 0, g_ccnt_cap = 0, g_cvoff_cap = 0, g_psol_cap = 0, g_part_cap = 0;
    int* g_vcell = nullptr;
    int* g_vsolid = nullptr;
    unsigned char* g_act = nullptr;
    unsigned char* g_all = nullptr;
    long long* g_vpos = nullptr;
    long long g_vcap = 0;
    int* g_pcell = nullptr;
    int* g_psolid = nullptr;
    int* g_label = nullptr;
    unsigned char* g_pall = nullptr;
    void* g_pcode = nullptr;
    double* g_pf = nullptr;
    int* g_clist = nullptr;
    long long g_pcap = 0;
    unsigned long long* g_hk = nullptr;
    int* g_hv = nullptr;
    int g_bits = 0;
    int* g_ccnt = nullptr;
    long long* g_coff = nullptr;
    long long* g_tmp = nullptr;
    long long g_tmpcap = 0;
    long long g_np = 0;
    unsigned long long* d_big_vbits = nullptr;
    int2* d_big_bcnt = nullptr;  // per cell block: ALL_INSIDE and cut counts
    int2* d_big_boff = nullptr;  // per cell block: its list offsets
    double* d_big_part = nullptr;
    long long big_word_cap = 0, big_chunk_cap = 0;
    // [0..4] phases of the main stream, [5], [6] the big-tile pass, [7] end of the cut-cell pass
    hipEvent_t ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double timing[7] = {0, 0, 0, 0, 0, 0, 0};
    // run-time variants (the autotuner's choices): candidates per chunk and workgroups per CU
    // of the big tiles' cut-cell pass; the small tiles' cut-cell pass one wave per solid
    // (1), 256 threads (0), or by the mean tile size (-1)
    int big_chunk = kBigChunk, big_grid = 4, cells_wave = -1;
    std::string tile_key;  // histogram of the solids' tile sizes at set_solids (the autotuner's cache key)
    std::string msg;
};

namespace {

thread_local std::string g_noctx_msg;

int fail(sdfibm_ctx* c, int code, const std::string& m) {
    if (c)
        c->msg = m;
    else
        g_noctx_msg = m;
    return code;
}

#define HIPCHK(ctx, call)                                                                                     \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return fail((ctx), SDFIBM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

int dalloc(sdfibm_ctx* c, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 8;
    // failure injection for the error-path tests: allocations above this many bytes fail
    if (const char* lim = std::getenv("SDFIBM_TEST_ALLOC_LIMIT"))
        if (*lim && bytes > (size_t)std::strtoull(lim, nullptr, 10))
            return fail(c, SDFIBM_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + ") refused by SDFIBM_TEST_ALLOC_LIMIT");
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(c, SDFIBM_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
    }
    c->allocs.push_back(*p);
    return SDFIBM_OK;
}

// Frees a context allocation and nulls the caller's pointer, so a later failing
// allocation never leaves a freed address behind in the context.
template <class T>
void dfree(sdfibm_ctx* c, T*& p) {
    if (!p) return;
    void* v = (void*)p;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), v);
    if (it != c->allocs.end()) c->allocs.erase(it);
    (void)hipFree(v);
    p = nullptr;
}

size_t rsize(const sdfibm_ctx* c) { return c->prec == SDFIBM_FP64 ? sizeof(double) : sizeof(float); }

unsigned grid_for(long long n, int block) { return (unsigned)((n + block - 1) / block); }

bool finite3(const double* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

int validate_solid(const sdfibm_solid& s, long long idx, std::string& why, long long n_csg) {
    if (s.shape < 0 || s.shape > SDFIBM_SHAPE_CSG) {
        why = "solid " + std::to_string(idx) + ": unknown shape id " + std::to_string(s.shape);
        return SDFIBM_ERR_SHAPE;
    }
    if (!finite3(s.prm, 6) || !finite3(s.center, 3) || !finite3(s.q, 4) || !finite3(s.vel, 3) ||
        !finite3(s.omega, 3)) {
        why = "solid " + std::to_string(idx) + ": non-finite parameter or state";
        return SDFIBM_ERR_ARG;
    }
    const double qn = s.q[0] * s.q[0] + s.q[1] * s.q[1] + s.q[2] * s.q[2] + s.q[3] * s.q[3];
    if (!(qn > 0.5 && qn < 2.0)) {
        why = "solid " + std::to_string(idx) + ": quaternion is not (close to) unit length";
        return SDFIBM_ERR_ARG;
    }
    int npos = 0;
    switch (s.shape) {
    case SDFIBM_SHAPE_SPHERE:
    case SDFIBM_SHAPE_CIRCLE: npos = 1; break;
    case SDFIBM_SHAPE_ELLIPSOID:
    case SDFIBM_SHAPE_BOX: npos = 3; break;
    case SDFIBM_SHAPE_ELLIPSE:
    case SDFIBM_SHAPE_RECTANGLE: npos = 2; break;
    case SDFIBM_SHAPE_CIRCLE_TAIL:
    case SDFIBM_SHAPE_CIRCLE_TWOTAIL: npos = 1; break;
    default: npos = 0;
    }
    for (int t = 0; t < npos; ++t)
        if (!(s.prm[t] > 0)) {
            why = "solid " + std::to_string(idx) + ": shape size parameter " + std::to_string(t) + " must be > 0";
            return SDFIBM_ERR_SHAPE;
        }
    if ((s.shape == SDFIBM_SHAPE_CIRCLE_TAIL || s.shape == SDFIBM_SHAPE_CIRCLE_TWOTAIL) &&
        !(s.prm[1] >= 0 && s.prm[2] > 0)) {
        why = "solid " + std::to_string(idx) + ": tail ratio must be >= 0 and thickness > 0";
        return SDFIBM_ERR_SHAPE;
    }
    if (s.shape == SDFIBM_SHAPE_CSG && !(s.prm[0] >= 0 && s.prm[0] < (double)n_csg && s.prm[0] == std::floor(s.prm[0]))) {
        why = "solid " + std::to_string(idx) + ": CSG shape index " + std::to_string(s.prm[0]) +
              " is not in the table of set_csg_shapes (" + std::to_string(n_csg) + " shapes)";
        return SDFIBM_ERR_SHAPE;
    }
    return SDFIBM_OK;
}

// A CSG shape record: primitive count, kinds, operations, axes, sizes, mass properties.
int validate_csg(const sdfibm_csg_shape& g, long long idx, std::string& why) {
    auto bad = [&](const std::string& w) {
        why = "CSG shape " + std::to_string(idx) + ": " + w;
        return SDFIBM_ERR_SHAPE;
    };
    if (g.n_prims < 1 || g.n_prims > SDFIBM_CSG_MAX_PRIMS) return bad("1 to 8 primitives");
    if (!(g.volume > 0) || !std::isfinite(g.volume) || !(g.radius_b >= 0) || !std::isfinite(g.radius_b))
        return bad("volume must be > 0 and the bounding radius >= 0");
    for (int a = 0; a < 3; ++a)
        if (!(g.moi[a] > 0) || !std::isfinite(g.moi[a])) return bad("moments of inertia must be > 0");
    for (int t = 0; t < g.n_prims; ++t) {
        const sdfibm_csg_prim& p = g.prim[t];
        if (p.kind < SDFIBM_CSG_BOX || p.kind > SDFIBM_CSG_SPHERE) return bad("unknown primitive kind");
        if (p.op < SDFIBM_CSG_UNION || p.op > SDFIBM_CSG_INTERSECTION) return bad("unknown operation");
        if (p.kind == SDFIBM_CSG_CYLINDER && (p.axis < 0 || p.axis > 2)) return bad("cylinder axis must be 0, 1 or 2");
        if (!finite3(p.center, 3) || !finite3(p.size, 3)) return bad("non-finite primitive");
        const int npos = p.kind == SDFIBM_CSG_SPHERE ? 1 : (p.kind == SDFIBM_CSG_CYLINDER ? 2 : 3);
        for (int a = 0; a < npos; ++a)
            if (!(p.size[a] > 0)) return bad("primitive sizes must be > 0");
    }
    return SDFIBM_OK;
}

// The big-tile pass (big_tiles.h) on stream st: every big solid of the step, in
// persistent grids that take their work from the device-side layout.
template <class R, int MK>
void launch_big(sdfibm_ctx* c, double dt, hipStream_t st) {
    const int ns = (int)c->ns;
    // flat kernels: 8 workgroups of 256 per CU (32 waves: their loads are latency-bound)
    const unsigned G = (unsigned)(8 * c->n_cu), GF = (unsigned)(c->big_grid * c->n_cu);
    R* codes = (R*)c->d_codes;
    k_big_layout<<<1, 1024, 0, st>>>(ns, c->d_T, c->qmeta, c->qlist, c->d_big_meta, c->d_big_tot, c->big_word_cap,
                                     c->err);
    k_big_vertices<MK><<<G, 256, 0, st>>>(c->m, c->d_K, c->d_T, c->d_big_meta, c->d_big_tot, c->d_big_vbits);
    k_big_cells<<<G, 256, 0, st>>>(c->d_T, c->d_big_meta, c->d_big_tot, c->d_big_vbits, c->d_cand, c->code_cap);
    k_big_hook<<<G, 256, 0, st>>>(c->d_T, c->d_big_meta, c->d_big_tot, c->d_cand, c->code_cap);
    k_big_compress<<<G, 256, 0, st>>>(c->d_T, c->d_big_meta, c->d_big_tot, c->d_cand, c->code_cap);
    k_big_member<R><<<G, 256, 0, st>>>(c->m, c->d_T, c->d_big_meta, c->d_big_tot, c->d_big_vbits, c->d_cand, codes,
                                       c->d_big_bcnt, c->code_cap, c->err);
    k_big_scan<<<(unsigned)c->n_cu, 1024, 0, st>>>(c->d_T, c->d_big_meta, c->d_big_tot, c->d_big_bcnt, c->d_big_boff,
                                                   c->d_ncand, c->code_cap, c->big_chunk);
    k_big_write<R><<<G, 256, 0, st>>>(c->d_T, c->d_big_meta, c->d_big_tot, codes, c->d_big_boff, c->d_ncand, c->d_cand,
                                      c->code_cap);
    k_big_chunks<<<1, 1024, 0, st>>>(c->d_big_meta, c->d_big_tot, c->big_chunk_cap, c->err);
    k_big_force<R, MK, kBlock><<<GF, kBlock, 0, st>>>(c->m, c->d_K, c->d_T, c->d_big_meta, c->d_big_tot, c->d_cand,
                                                     c->d_ncand, (const R*)c->U[0], (const R*)c->U[1],
                                                     (const R*)c->U[2], codes, c->code_cap, dt, c->d_big_part,
                                                     c->big_chunk);
    k_big_ft<<<64, 256, 0, st>>>(c->d_big_meta, c->d_big_tot, c->d_big_part, c->rho_f, c->d_ft);
}

// graph: the launches are being captured into a HIP graph -- no timing events
// and no plane labelling (it needs host round trips); the plane membership
// bitmaps of the last eager pass are used, the planes being fixed.
template <class R, int MK>
int launch_step(sdfibm_ctx* c, double dt, bool graph) {
    // timing events only on request: each record orders the kernels around it (about 25 us
    // per step on the cloud3d workload)
    const bool ev_on = !graph && c->phase_timing;
    const int ns = (int)c->ns;
    hipStream_t st = c->stream;
    if (ev_on) (void)hipEventRecord(c->ev[0], st);
    k_reset<<<grid_for(std::max(c->nbricks, 8LL), 256), 256, 0, st>>>(c->err, c->qmeta, c->bcount, c->nbricks,
                                                                     c->d_tile_total);
    if (ns > 0) {
        k_setup<MK><<<grid_for(ns, 64), 64, setup_lds_bytes(c->m, MK), st>>>(c->d_solids, ns, c->m, c->d_K, c->d_T, c->d_tile_total,
                                                       c->code_cap, c->d_ft, c->qmeta, c->qlist, c->d_unb, c->err,
                                                       c->d_csg, c->big_on ? 1 : 0);
        if (MK != MK_RECT) k_seed<<<grid_for(ns, 4), 256, 0, st>>>(c->m, ns, c->d_K, c->d_T);
    }
    if (ns > 0) {
        const unsigned gb = grid_for(64LL * ns, 256);
        if (c->nz == 1)
            k_bin<kBX2, kBY2, kBZ2><<<gb, 256, 0, st>>>(ns, c->d_T, c->nbx, c->nby, c->bcount, c->d_blist, c->err);
        else
            k_bin<kBX3, kBY3, kBZ3><<<gb, 256, 0, st>>>(ns, c->d_T, c->nbx, c->nby, c->bcount, c->d_blist, c->err);
    }
    if (ev_on) (void)hipEventRecord(c->ev[1], st);
    R* codes = (R*)c->d_codes;
    if (ns > 0) {
        // the class sizes live on the device: the small-tile grid covers every solid (one
        // each), the medium grid (big LDS, few solids) loops over its queue; the medium
        // tiles (long per-solid latency) and the big-tile pass run beside the small ones
        // on a second stream
        const unsigned g = (unsigned)ns, cus = (unsigned)c->n_cu;
        if (hipEventRecord(c->fork, st) != hipSuccess || hipStreamWaitEvent(c->aux, c->fork, 0) != hipSuccess)
            return -1;
        k_classify<R, MK, kRowsMedium, 1, 1024><<<std::min(g, 4 * cus), 1024, 0, c->aux>>>(
            c->m, ns, c->d_K, c->d_T, c->qmeta, c->qlist, codes, c->d_cand, c->code_cap, c->d_ncand);
        // the cut-cell pass needs these tiles, the field pass also the big ones
        if (hipEventRecord(c->joinc, c->aux) != hipSuccess) return -1;
        if (c->big_on) {
            if (ev_on) (void)hipEventRecord(c->ev[5], c->aux);
            launch_big<R, MK>(c, dt, c->aux);
            if (ev_on) (void)hipEventRecord(c->ev[6], c->aux);
        }
        if (hipEventRecord(c->join, c->aux) != hipSuccess) return -1;
        k_classify<R, MK, kRowsSmall, 0, kBlock><<<g, kBlock, 0, st>>>(c->m, ns, c->d_K, c->d_T, c->qmeta, c->qlist, codes,
                                                               c->d_cand, c->code_cap, c->d_ncand);
        k_classify<R, MK, kRowsTiny, 3, 64><<<g, 64, 0, st>>>(c->m, ns, c->d_K, c->d_T, c->qmeta, c->qlist, codes,
                                                              c->d_cand, c->code_cap, c->d_ncand);
        if (hipStreamWaitEvent(st, c->joinc, 0) != hipSuccess) return -1;
        if (ev_on) (void)hipEventRecord(c->ev[4], st);
        constexpr bool kSgOk = MK == MK_RECT && sizeof(R) == sizeof(float);
        const bool sg = kSgOk && c->d_sg != nullptr;
        // same workgroup sizes with and without the surrogate: identical reduction order when
        // every cell falls back
        if (sg) k_fill_i32<<<1, 64, 0, st>>>(c->d_sg_count, 1, 0);
        const bool one_wave = c->cells_wave == 1 || (c->cells_wave < 0 && c->small_tiles);
#define SDFIBM_CELLS(SG_, NT_, GRID_, PART_)                                                                    \
    k_cells<R, MK, SG_, NT_><<<GRID_, NT_, 0, st>>>(                                                           \
        c->m, ns, c->d_K, c->d_T, c->qmeta, c->qlist, c->d_cand, c->d_ncand, (const R*)c->U[0],               \
        (const R*)c->U[1], (const R*)c->U[2], codes, c->code_cap, dt, c->rho_f, c->d_ft, SG_ ? c->d_sg : nullptr, \
        c->sg_lo, c->sg_hi, c->sg_cos, SG_ ? c->d_sg_count : nullptr, PART_, c->lm_grid)
        // one wave per small or tiny tile (a few dozen cut cells each); the large and medium
        // tiles -- up to lm_grid of them, the number set_solids allows for (beyond it, one
        // wave each) -- keep 256 threads
        if (sg) {
            if (one_wave) {
                if (c->lm_grid > 0) SDFIBM_CELLS(kSgOk, kBlock, (unsigned)c->lm_grid, 1);
                SDFIBM_CELLS(kSgOk, 64, g, 2);
            } else {
                SDFIBM_CELLS(kSgOk, kBlock, g, 0);
            }
        } else {
            if (one_wave) {
                if (c->lm_grid > 0) SDFIBM_CELLS(false, kBlock, (unsigned)c->lm_grid, 1);
                SDFIBM_CELLS(false, 64, g, 2);
            } else {
                SDFIBM_CELLS(false, kBlock, g, 0);
            }
        }
#undef SDFIBM_CELLS
        if (ev_on) (void)hipEventRecord(c->ev[7], st);
    } else {
        if (ev_on) (void)hipEventRecord(c->ev[4], st);
        if (ev_on) (void)hipEventRecord(c->ev[7], st);
    }
    const int np = (int)c->h_planes.size();
    unsigned dirty = 0;
    if (!graph && np > 0) {
        // only the planes that moved since their last labelling are labelled again
        k_plane_pose<<<1, 64, 0, st>>>(c->d_solids, c->d_planes, np, c->d_plane_cache, c->d_pscalars + 3);
        k_flags_out<<<1, 64, 0, st>>>(c->d_pscalars + 3, c->d_hflags + 1);  // pinned mapped word
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
        dirty = (unsigned)*(volatile int*)(c->h_flags + 1);
    }
    for (int p = 0; p < (graph ? 0 : np); ++p) {
        if (!((dirty >> p) & 1u)) continue;
        const int sp = c->h_planes[p];
        unsigned* member = c->d_pmember + (long long)p * c->member_words;
        k_fill_i32<<<grid_for(c->member_words, 256), 256, 0, st>>>((int*)c->d_pact, c->member_words, 0);
        k_fill_i32<<<grid_for(c->member_words, 256), 256, 0, st>>>((int*)member, c->member_words, 0);
        k_fill_i32<<<1, 64, 0, st>>>(c->d_pscalars, 1, 0);
        k_fill_i32<<<1, 64, 0, st>>>(c->d_pscalars + 1, 1, 0x7fffffff);
        if (c->nz == 1)
            k_plane_active<MK, kBX2, kBY2, kBZ2><<<dim3(c->nbx, c->nby, c->nbz), kBX2 * kBY2 * kBZ2, 0, st>>>(
                c->m, c->nbx, c->nby, c->brick_box, c->d_K, c->d_T, sp, c->d_pact, c->d_plist, c->d_pscalars,
                c->N, c->d_pscalars + 1);
        else
            k_plane_active<MK, kBX3, kBY3, kBZ3><<<dim3(c->nbx, c->nby, c->nbz), kBX3 * kBY3 * kBZ3, 0, st>>>(
                c->m, c->nbx, c->nby, c->brick_box, c->d_K, c->d_T, sp, c->d_pact, c->d_plist, c->d_pscalars,
                c->N, c->d_pscalars + 1);
        int nact = 0;
        if (hipMemcpyAsync(&nact, c->d_pscalars, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return -1;
        if (nact == 0) continue;
        const unsigned g = grid_for(nact, 256);
        k_ccl_init<<<g, 256, 0, st>>>(c->d_plist, c->d_pscalars, c->d_plabel);
        int changed = 1, it = 0;
        for (; changed && it < 64; ++it) {
            k_fill_i32<<<1, 64, 0, st>>>(c->d_pscalars + 2, 1, 0);
            k_ccl_hook<<<g, 256, 0, st>>>(c->m, c->d_plist, c->d_pscalars, c->d_pact, c->d_plabel, c->d_pscalars + 2);
            k_ccl_compress<<<g, 256, 0, st>>>(c->d_plist, c->d_pscalars, c->d_plabel);
            if (hipMemcpyAsync(&changed, c->d_pscalars + 2, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return -1;
        }
        if (changed) return -2;
        k_plane_member<<<g, 256, 0, st>>>(c->d_plist, c->d_pscalars, c->d_plabel, c->d_pact, c->d_T, sp,
                                          c->d_pscalars + 1, member, c->d_pseed + p);
        k_plane_alpha<R, MK><<<g, 256, 0, st>>>(c->m, c->d_plist, c->d_pscalars, c->d_K, sp, member,
                                                (R*)c->d_pcode + (long long)p * c->N);
    }
    if (ns > 0 && hipStreamWaitEvent(st, c->join, 0) != hipSuccess) return -1;  // the big-tile pass
    if (ev_on) (void)hipEventRecord(c->ev[2], st);
#define SDFIBM_FIELDS(BX_, BY_, BZ_, CPT_, HP_)                                                                 \
    k_fields<R, MK, BX_, BY_, BZ_, CPT_, HP_><<<dim3(c->nbx, c->nby, c->nbz), BX_ * BY_ * BZ_ / CPT_, 0, st>>>(     \
        c->m, c->nbx, c->nby, c->bcount, c->d_blist, c->d_K, c->d_T, codes, c->code_cap, c->d_planes, np,          \
        c->d_pmember, c->member_words, (const R*)c->d_pcode, c->brick_box, (const R*)c->U[0], (const R*)c->U[1],   \
        (const R*)c->U[2],                                                                                         \
        (R*)c->As, (R*)c->Fs[0], (R*)c->Fs[1], (R*)c->Fs[2], (R*)c->Tsf, c->Ct, dt, c->d_plane_partial, c->nbricks, \
        c->err)
    if (c->nz == 1) {
        if (np > 0)
            SDFIBM_FIELDS(kBX2, kBY2, kBZ2, 1, true);
        else
            SDFIBM_FIELDS(kBX2, kBY2, kBZ2, 1, false);
    } else {
        if (np > 0)
            SDFIBM_FIELDS(kBX3, kBY3, kBZ3, kCPT3, true);
        else
            SDFIBM_FIELDS(kBX3, kBY3, kBZ3, kCPT3, false);
    }
#undef SDFIBM_FIELDS
    if (np > 0) k_plane_reduce<<<np, 1024, 0, st>>>(c->d_plane_partial, c->nbricks, c->d_planes, c->rho_f, c->d_ft);
    if (ev_on) (void)hipEventRecord(c->ev[3], st);
    return 0;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

int setup_mesh(sdfibm_ctx* c, const sdfibm_mesh_desc* d) {
    const long long nx = d->nx, ny = d->ny, nz = d->nz;
    const long long N = nx * ny * nz;
    const long long NP = (nx + 1) * (ny + 1) * (nz + 1);
    MeshDev m{};
    m.nx = (int)nx;
    m.ny = (int)ny;
    m.nz = (int)nz;
    m.two_d = c->two_d;
    int rc;
    if (d->kind == SDFIBM_MESH_RECTILINEAR) {
        if (!d->xs || !d->ys || !d->zs) return fail(c, SDFIBM_ERR_ARG, "rectilinear mesh needs xs, ys, zs");
        c->h_xs.assign(d->xs, d->xs + nx + 1);
        c->h_ys.assign(d->ys, d->ys + ny + 1);
        c->h_zs.assign(d->zs, d->zs + nz + 1);
        for (auto* v : {&c->h_xs, &c->h_ys, &c->h_zs}) {
            for (size_t t = 0; t < v->size(); ++t) {
                if (!std::isfinite((*v)[t])) return fail(c, SDFIBM_ERR_MESH, "non-finite mesh coordinate");
                if (t > 0 && !((*v)[t] > (*v)[t - 1]))
                    return fail(c, SDFIBM_ERR_MESH, "mesh coordinates must be strictly increasing");
            }
        }
        double* p[3];
        const std::vector<double>* src[3] = {&c->h_xs, &c->h_ys, &c->h_zs};
        for (int a = 0; a < 3; ++a) {
            if ((rc = dalloc(c, (void**)&p[a], src[a]->size() * 8))) return rc;
            HIPCHK(c, hipMemcpy(p[a], src[a]->data(), src[a]->size() * 8, hipMemcpyHostToDevice));
        }
        m.xs = p[0];
        m.ys = p[1];
        m.zs = p[2];
        for (int a = 0; a < 3; ++a) {
            c->box_lo[a] = src[a]->front();
            c->box_hi[a] = src[a]->back();
        }
        c->hm = m;
        c->hm.xs = c->h_xs.data();
        c->hm.ys = c->h_ys.data();
        c->hm.zs = c->h_zs.data();
    } else if (d->kind == SDFIBM_MESH_STRUCTURED) {
        if (!d->points || !d->cell_centres || !d->cell_volumes)
            return fail(c, SDFIBM_ERR_ARG, "structured mesh needs points, cell_centres, cell_volumes");
        for (int f = 0; f < 3; ++f)
            if (!d->face_centres[f] || !d->face_areas[f])
                return fail(c, SDFIBM_ERR_ARG, "structured mesh needs the three face families");
        // nominal uniform grid: the box spanned by the corner points
        const double* P = d->points;
        auto pidx = [&](long long i, long long j, long long k) { return i + (nx + 1) * (j + (ny + 1) * k); };
        const long long pe = pidx(nx, ny, nz);
        double o[3], h[3];
        const long long nn[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a) {
            o[a] = P[a];
            h[a] = (P[3 * pe + a] - P[a]) / (double)nn[a];
            if (!(h[a] > 0)) return fail(c, SDFIBM_ERR_MESH, "structured mesh: corner points do not span a box");
            c->box_lo[a] = o[a];
            c->box_hi[a] = P[3 * pe + a];
        }
        double jit = 0.0;
        for (int a = 0; a < 3; ++a)
            for (long long k = 0; k <= nz; ++k)
                for (long long j = 0; j <= ny; ++j)
                    for (long long i = 0; i <= nx; ++i) {
                        const double v = P[3 * pidx(i, j, k) + a];
                        if (!std::isfinite(v)) return fail(c, SDFIBM_ERR_MESH, "non-finite point");
                        const long long idx = a == 0 ? i : (a == 1 ? j : k);
                        jit = std::max(jit, std::fabs((v - o[a]) / h[a] - (double)idx));
                    }
        double* dp = nullptr;  // AoS, as given
        if ((rc = dalloc(c, (void**)&dp, NP * 24))) return rc;
        HIPCHK(c, hipMemcpy(dp, P, NP * 24, hipMemcpyHostToDevice));
        if (jit > 0.45)
            return fail(c, SDFIBM_ERR_MESH,
                        "structured mesh: points deviate more than 0.45 cells from a uniform grid (" +
                            std::to_string(jit) + ")");
        m.pxyz = dp;
        std::vector<double> cs(N);
        double* dc[3];
        for (int a = 0; a < 3; ++a) {
            for (long long q = 0; q < N; ++q) cs[q] = d->cell_centres[3 * q + a];
            if ((rc = dalloc(c, (void**)&dc[a], N * 8))) return rc;
            HIPCHK(c, hipMemcpy(dc[a], cs.data(), N * 8, hipMemcpyHostToDevice));
        }
        m.ccx = dc[0];
        m.ccy = dc[1];
        m.ccz = dc[2];
        double* dv;
        if ((rc = dalloc(c, (void**)&dv, N * 8))) return rc;
        HIPCHK(c, hipMemcpy(dv, d->cell_volumes, N * 8, hipMemcpyHostToDevice));
        m.vol = dv;
        const long long nf[3] = {(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)};
        for (int f = 0; f < 3; ++f) {
            double *a, *b;
            if ((rc = dalloc(c, (void**)&a, nf[f] * 24))) return rc;
            if ((rc = dalloc(c, (void**)&b, nf[f] * 24))) return rc;
            HIPCHK(c, hipMemcpy(a, d->face_centres[f], nf[f] * 24, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(b, d->face_areas[f], nf[f] * 24, hipMemcpyHostToDevice));
            m.fc[f] = a;
            m.sf[f] = b;
        }
        for (int a = 0; a < 3; ++a) {
            m.o[a] = o[a];
            m.h[a] = h[a];
        }
        m.jit = jit;
        c->hm = m;  // host copy: only the nominal grid is used on the host
        // brick vertex boxes (plane culling)
        const int BX = nz == 1 ? kBX2 : kBX3, BY = nz == 1 ? kBY2 : kBY3, BZ = nz == 1 ? kBZ2 : kBZ3;
        const long long nbx = (nx + BX - 1) / BX, nby = (ny + BY - 1) / BY, nbz = (nz + BZ - 1) / BZ;
        std::vector<double> box(nbx * nby * nbz * 6);
        for (long long bz = 0; bz < nbz; ++bz)
            for (long long by = 0; by < nby; ++by)
                for (long long bx = 0; bx < nbx; ++bx) {
                    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
                    for (long long k = bz * BZ; k <= std::min(bz * BZ + BZ, nz); ++k)
                        for (long long j = by * BY; j <= std::min(by * BY + BY, ny); ++j)
                            for (long long i = bx * BX; i <= std::min(bx * BX + BX, nx); ++i) {
                                const long long q = pidx(i, j, k);
                                for (int a = 0; a < 3; ++a) {
                                    lo[a] = std::min(lo[a], P[3 * q + a]);
                                    hi[a] = std::max(hi[a], P[3 * q + a]);
                                }
                            }
                    double* bb = &box[6 * (bx + nbx * (by + nby * bz))];
                    bb[0] = lo[0];
                    bb[1] = hi[0];
                    bb[2] = lo[1];
                    bb[3] = hi[1];
                    bb[4] = lo[2];
                    bb[5] = hi[2];
                }
        if ((rc = dalloc(c, (void**)&c->brick_box, box.size() * 8))) return rc;
        HIPCHK(c, hipMemcpy(c->brick_box, box.data(), box.size() * 8, hipMemcpyHostToDevice));
    } else {
        return fail(c, SDFIBM_ERR_MESH, "unknown mesh kind " + std::to_string(d->kind));
    }
    c->m = m;
    return SDFIBM_OK;
}

template <class R, int MK>
int interact_impl(sdfibm_ctx* c, double dt, bool graph) {
    const int r = launch_step<R, MK>(c, dt, graph);
    HIPCHK(c, hipGetLastError());
    if (r == -1) return fail(c, SDFIBM_ERR_HIP, "plane component pass: HIP copy/synchronize failed");
    if (r == -2) return fail(c, SDFIBM_ERR_CAPACITY, "plane component labelling did not converge");
    return SDFIBM_OK;
}

int dispatch_interact(sdfibm_ctx* c, double dt, bool graph = false) {
    if (c->prec == SDFIBM_FP64)
        return c->kind == SDFIBM_MESH_RECTILINEAR ? interact_impl<double, MK_RECT>(c, dt, graph)
                                                  : interact_impl<double, MK_STRUCT>(c, dt, graph);
    return c->kind == SDFIBM_MESH_RECTILINEAR ? interact_impl<float, MK_RECT>(c, dt, graph)
                                              : interact_impl<float, MK_STRUCT>(c, dt, graph);
}

bool contact_on(const sdfibm_ctx* c);
int contact_prepare(sdfibm_ctx* c);
int classify_solids(sdfibm_ctx* c, const sdfibm_solid* solids, int64_t n);
int resolve_stage(sdfibm_ctx* c);

// Drop the captured step (its launches hold the buffers and settings of capture time).
void graph_reset(sdfibm_ctx* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    if (c->gexec_i) (void)hipGraphExecDestroy(c->gexec_i);
    if (c->graph_i) (void)hipGraphDestroy(c->graph_i);
    c->gexec_i = nullptr;
    c->graph_i = nullptr;
}

// Launch-only parts of the steps around interact (no synchronisation; also
// used inside graph capture).
void launch_update_velocity(sdfibm_ctx* c, double dt) {
    const unsigned g = grid_for(c->N, 256);
    if (c->prec == SDFIBM_FP64)
        k_update_velocity<double><<<g, 256, 0, c->stream>>>(c->N, (double*)c->U[0], (double*)c->U[1], (double*)c->U[2],
                                                            (const double*)c->Fs[0], (const double*)c->Fs[1],
                                                            (const double*)c->Fs[2], dt);
    else
        k_update_velocity<float><<<g, 256, 0, c->stream>>>(c->N, (float*)c->U[0], (float*)c->U[1], (float*)c->U[2],
                                                           (const float*)c->Fs[0], (const float*)c->Fs[1],
                                                           (const float*)c->Fs[2], dt);
}

void launch_fix_internal(sdfibm_ctx* c) {
    const unsigned g = grid_for(c->N, 256);
    const int ns = (int)c->ns;
#define SDFIBM_FIX(R_, MK_)                                                                                  \
    k_fix_internal<R_, MK_><<<g, 256, 0, c->stream>>>(c->m, c->Ct, c->d_solids, ns, (R_*)c->U[0], (R_*)c->U[1], \
                                                     (R_*)c->U[2])
    if (c->prec == SDFIBM_FP64) {
        if (c->kind == SDFIBM_MESH_RECTILINEAR)
            SDFIBM_FIX(double, MK_RECT);
        else
            SDFIBM_FIX(double, MK_STRUCT);
    } else {
        if (c->kind == SDFIBM_MESH_RECTILINEAR)
            SDFIBM_FIX(float, MK_RECT);
        else
            SDFIBM_FIX(float, MK_STRUCT);
    }
#undef SDFIBM_FIX
}

// U -= Fs*dt and fixInternal fused, after evolve (the coupled loop only)
// gate: a device word that skips the pass when nonzero (the graph loop's sticky error flags), or null.
void launch_update_fix(sdfibm_ctx* c, double dt, const int* gate) {
    const unsigned g = grid_for(c->N, 256);
    const int ns = (int)c->ns;
#define SDFIBM_UFIX(R_, MK_)                                                                                     \
    k_update_fix<R_, MK_><<<g, 256, 0, c->stream>>>(c->m, c->Ct, c->d_solids, ns, (R_*)c->U[0], (R_*)c->U[1],   \
                                                   (R_*)c->U[2], (const R_*)c->Fs[0], (const R_*)c->Fs[1],     \
                                                   (const R_*)c->Fs[2], dt, gate)
    if (c->prec == SDFIBM_FP64) {
        if (c->kind == SDFIBM_MESH_RECTILINEAR)
            SDFIBM_UFIX(double, MK_RECT);
        else
            SDFIBM_UFIX(double, MK_STRUCT);
    } else {
        if (c->kind == SDFIBM_MESH_RECTILINEAR)
            SDFIBM_UFIX(float, MK_RECT);
        else
            SDFIBM_UFIX(float, MK_STRUCT);
    }
#undef SDFIBM_UFIX
}

// gate: as for launch_update_fix.
void launch_evolve(sdfibm_ctx* c, double dt, Vec3 g, const int* gate) {
    const int ns = (int)c->ns;
    if (ns == 0) return;
    const int n_sub = ns == 1 ? 1 : 20;  // solidcloud.cpp:524-526
    if (contact_on(c) && ns > 1 && ns <= kContactBlockSolids && c->ncells_c <= kContactBlockCells) {
        // few solids: every sub-step in one workgroup (the result ends in d_solids: 20 sub-steps)
        // (up to 256 solids: a 256-thread workgroup, whose register budget holds the sub-step)
        if (ns <= 256)
            k_contact_evolve_block<256><<<1, 256, 0, c->stream>>>(
                c->d_solids, c->d_solids_alt, c->d_dyn, c->d_ft, c->d_rb, ns, c->cgrid, c->d_chead, c->d_cnext,
                c->ncells_c, c->d_planes, (int)c->h_planes.size(), c->law, dt / n_sub, g, c->rho_f, n_sub, gate);
        else
            k_contact_evolve_block<1024><<<1, 1024, 0, c->stream>>>(
                c->d_solids, c->d_solids_alt, c->d_dyn, c->d_ft, c->d_rb, ns, c->cgrid, c->d_chead, c->d_cnext,
                c->ncells_c, c->d_planes, (int)c->h_planes.size(), c->law, dt / n_sub, g, c->rho_f, n_sub, gate);
    } else if (contact_on(c) && ns > 1) {
        // sub-steps alternate between the two solid buffers (20: the result ends in d_solids);
        // cell lists: heads H[it % 3], nexts N[it % 2]
        const long long nc = c->ncells_c;
        k_fill_i32<<<grid_for(3 * nc, 256), 256, 0, c->stream>>>(c->d_chead, 3 * nc, -1);
        sdfibm_solid* buf[2] = {c->d_solids, c->d_solids_alt};
        const double dts = dt / n_sub;
        const int nplanes = (int)c->h_planes.size();
        k_contact_insert<<<grid_for(ns, 64), 64, 0, c->stream>>>(c->d_solids, c->d_rb, ns, c->cgrid, c->d_chead,
                                                                 c->d_cnext, gate);
        for (int it = 0; it < n_sub; ++it) {
            int* H[3] = {c->d_chead + (it % 3) * nc, c->d_chead + ((it + 1) % 3) * nc, c->d_chead + ((it + 2) % 3) * nc};
            int* Nr = c->d_cnext + (long long)(it & 1) * c->contact_cap;
            int* Nw = c->d_cnext + (long long)((it + 1) & 1) * c->contact_cap;
            k_contact_step<<<grid_for(ns, 64), 64, 0, c->stream>>>(
                buf[it & 1], buf[(it + 1) & 1], c->d_dyn, c->d_ft, c->d_rb, ns, c->cgrid, H[0], Nr, H[1], Nw, H[2],
                (int)nc, c->d_planes, nplanes, c->law, dts, g, c->rho_f, it == n_sub - 1, gate);
        }
    } else {
        k_evolve<<<grid_for(ns, 64), 64, 0, c->stream>>>(c->d_solids, c->d_dyn, c->d_ft, ns, n_sub, dt, g, c->rho_f,
                                                         gate);
    }
}

// Staged AoS FP64 velocities of cells [first, first + n) into the resident SoA U.
void launch_aos_to_soa(sdfibm_ctx* c, const double* staged, long long first, long long n) {
    if (n <= 0) return;
    if (c->prec == SDFIBM_FP64)
        k_aos_to_soa<double><<<grid_for(n, 256), 256, 0, c->stream>>>(
            staged, n, (double*)c->U[0] + first, (double*)c->U[1] + first, (double*)c->U[2] + first);
    else
        k_aos_to_soa<float><<<grid_for(n, 256), 256, 0, c->stream>>>(
            staged, n, (float*)c->U[0] + first, (float*)c->U[1] + first, (float*)c->U[2] + first);
}

std::string err_text(int flags) {
    std::string s;
    if (flags & EF_TILE) s += "a solid's cell tile exceeds 63 cells along x or 1024 vertex rows; ";
    if (flags & EF_CODECAP) s += "tile code buffer too small (or more than 16384 big tiles); ";
    if (flags & EF_BRICK) s += "more than 128 solids overlap one brick; ";
    if (flags & EF_PLANE_MOVED) s += "a plane solid moved during a graph replay; ";
    if (flags & EF_LABEL) s += "the big-tile component labelling did not converge; ";
    return s;
}

// ---------------------------------------------------------------------------
// General polyhedral meshes (SDFIBM_MESH_POLY): set-up and the step
// ---------------------------------------------------------------------------
// Face centre and area vector of a polygon as OpenFOAM's primitiveMesh forms
// them: triangles fanned about the average of the face's points, the centre
// their centroids weighted by the triangle areas projected on the face normal.
void poly_face_geometry(const double* P, const int32_t* lab, int n, double* fc, double* sf) {
    auto pt = [&](int t, int d) { return P[3 * (long long)lab[t] + d]; };
    if (n == 3) {
        double e1[3], e2[3];
        for (int d = 0; d < 3; ++d) {
            fc[d] = (1.0 / 3.0) * ((pt(0, d) + pt(1, d)) + pt(2, d));
            e1[d] = pt(1, d) - pt(0, d);
            e2[d] = pt(2, d) - pt(0, d);
        }
        sf[0] = 0.5 * (e1[1] * e2[2] - e1[2] * e2[1]);
        sf[1] = 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]);
        sf[2] = 0.5 * (e1[0] * e2[1] - e1[1] * e2[0]);
        return;
    }
    double m[3] = {0, 0, 0};
    for (int t = 0; t < n; ++t)
        for (int d = 0; d < 3; ++d) m[d] += pt(t, d);
    for (int d = 0; d < 3; ++d) m[d] /= (double)n;
    // triangle (p_t, p_t+1, m): normal (p_t+1 - p_t) x (m - p_t)
    auto tri_normal = [&](int t, double* nrm) {
        const int u = (t + 1) % n;
        double a[3], b[3];
        for (int d = 0; d < 3; ++d) {
            a[d] = pt(u, d) - pt(t, d);
            b[d] = m[d] - pt(t, d);
        }
        nrm[0] = a[1] * b[2] - a[2] * b[1];
        nrm[1] = a[2] * b[0] - a[0] * b[2];
        nrm[2] = a[0] * b[1] - a[1] * b[0];
    };
    double sum_n[3] = {0, 0, 0};
    for (int t = 0; t < n; ++t) {
        double nrm[3];
        tri_normal(t, nrm);
        for (int d = 0; d < 3; ++d) sum_n[d] += nrm[d];
    }
    const double mag = std::sqrt(sum_n[0] * sum_n[0] + sum_n[1] * sum_n[1] + sum_n[2] * sum_n[2]);
    for (int d = 0; d < 3; ++d) sf[d] = 0.5 * sum_n[d];
    if (!(mag > 1e-300)) {
        for (int d = 0; d < 3; ++d) fc[d] = m[d];
        return;
    }
    double w_sum = 0.0, c_sum[3] = {0, 0, 0};
    for (int t = 0; t < n; ++t) {
        double nrm[3];
        tri_normal(t, nrm);
        const double w = (nrm[0] * sum_n[0] + nrm[1] * sum_n[1] + nrm[2] * sum_n[2]) / mag;
        const int u = (t + 1) % n;
        w_sum += w;
        for (int d = 0; d < 3; ++d) c_sum[d] += w * ((pt(t, d) + pt(u, d)) + m[d]);
    }
    for (int d = 0; d < 3; ++d) fc[d] = ((1.0 / 3.0) * c_sum[d]) / w_sum;
}

// The mesh's addressing and geometry on the device; the cell bins.
int poly_setup(sdfibm_ctx* c, const sdfibm_polymesh_desc* d) {
    const long long nP = d->n_points, nF = d->n_faces, nI = d->n_internal_faces, nC = d->n_cells;
    if (nP < 4 || nF < 4 || nC < 1 || nI < 0 || nI > nF)
        return fail(c, SDFIBM_ERR_MESH, "polyMesh sizes are inconsistent");
    if (nP > INT_MAX || nF > INT_MAX || nC > INT_MAX) return fail(c, SDFIBM_ERR_MESH, "more than 2^31-1 points, faces or cells");
    if (!d->points || !d->face_offsets || !d->face_points || !d->owner || (nI > 0 && !d->neighbour))
        return fail(c, SDFIBM_ERR_ARG, "polyMesh arrays missing");
    const int64_t* fo = d->face_offsets;
    if (fo[0] != 0) return fail(c, SDFIBM_ERR_MESH, "face_offsets[0] must be 0");
    for (long long f = 0; f < nF; ++f)
        if (fo[f + 1] - fo[f] < 3) return fail(c, SDFIBM_ERR_MESH, "face " + std::to_string(f) + " has fewer than 3 points");
    if (fo[nF] > INT_MAX) return fail(c, SDFIBM_ERR_MESH, "more than 2^31-1 face points");
    for (long long t = 0; t < fo[nF]; ++t)
        if (d->face_points[t] < 0 || d->face_points[t] >= nP) return fail(c, SDFIBM_ERR_MESH, "face point label out of range");
    for (long long f = 0; f < nF; ++f) {
        if (d->owner[f] < 0 || d->owner[f] >= nC) return fail(c, SDFIBM_ERR_MESH, "owner label out of range");
        if (f < nI && (d->neighbour[f] < 0 || d->neighbour[f] >= nC || d->neighbour[f] == d->owner[f]))
            return fail(c, SDFIBM_ERR_MESH, "neighbour label out of range");
    }
    // cells(): owned faces ascending, then neighbour faces ascending (primitiveMesh::calcCells)
    std::vector<int> cfo(nC + 1, 0), cf;
    for (long long f = 0; f < nF; ++f) ++cfo[d->owner[f] + 1];
    for (long long f = 0; f < nI; ++f) ++cfo[d->neighbour[f] + 1];
    for (long long k = 0; k < nC; ++k) cfo[k + 1] += cfo[k];
    cf.resize(cfo[nC]);
    {
        std::vector<int> at(cfo.begin(), cfo.end() - 1);
        for (long long f = 0; f < nF; ++f) cf[at[d->owner[f]]++] = (int)f;
        for (long long f = 0; f < nI; ++f) cf[at[d->neighbour[f]]++] = (int)f;
    }
    for (long long k = 0; k < nC; ++k)
        if (cfo[k + 1] - cfo[k] < 4) return fail(c, SDFIBM_ERR_MESH, "cell " + std::to_string(k) + " has fewer than 4 faces");
    // cellPoints (ascending labels) and cellCells
    std::vector<int> cpo(nC + 1, 0), cp, cno(nC + 1, 0), cn;
    {
        std::vector<int> tmp;
        for (long long k = 0; k < nC; ++k) {
            tmp.clear();
            for (int t = cfo[k]; t < cfo[k + 1]; ++t)
                for (long long u = fo[cf[t]]; u < fo[cf[t] + 1]; ++u) tmp.push_back(d->face_points[u]);
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            cp.insert(cp.end(), tmp.begin(), tmp.end());
            cpo[k + 1] = (int)cp.size();
        }
        for (long long f = 0; f < nI; ++f) {
            ++cno[d->owner[f] + 1];
            ++cno[d->neighbour[f] + 1];
        }
        for (long long k = 0; k < nC; ++k) cno[k + 1] += cno[k];
        cn.resize(cno[nC]);
        std::vector<int> at(cno.begin(), cno.end() - 1);
        for (long long f = 0; f < nI; ++f) {
            cn[at[d->owner[f]]++] = d->neighbour[f];
            cn[at[d->neighbour[f]]++] = d->owner[f];
        }
    }
    // geometry: the caller's, else primitiveMesh's (face fan; cell pyramids about the average
    // of its face centres)
    std::vector<double> fcv, sfv, ccv, volv;
    const double *fcp = d->face_centres, *sfp = d->face_areas, *ccp = d->cell_centres, *vp = d->cell_volumes;
    if (!fcp || !sfp) {
        fcv.resize(3 * nF);
        sfv.resize(3 * nF);
        for (long long f = 0; f < nF; ++f)
            poly_face_geometry(d->points, d->face_points + fo[f], (int)(fo[f + 1] - fo[f]), &fcv[3 * f], &sfv[3 * f]);
        fcp = fcv.data();
        sfp = sfv.data();
    }
    if (!ccp || !vp) {
        ccv.assign(3 * nC, 0.0);
        volv.assign(nC, 0.0);
        for (long long k = 0; k < nC; ++k) {
            double est[3] = {0, 0, 0};
            const int nfc = cfo[k + 1] - cfo[k];
            for (int t = cfo[k]; t < cfo[k + 1]; ++t)
                for (int e = 0; e < 3; ++e) est[e] += fcp[3 * (long long)cf[t] + e];
            for (int e = 0; e < 3; ++e) est[e] /= (double)nfc;
            double vsum = 0.0, csum[3] = {0, 0, 0};
            for (int t = cfo[k]; t < cfo[k + 1]; ++t) {
                const long long f = cf[t];
                const double sgn = d->owner[f] == k ? 1.0 : -1.0;  // outward from this cell
                double h = 0.0;
                for (int e = 0; e < 3; ++e) h += sfp[3 * f + e] * (sgn * (fcp[3 * f + e] - est[e]));
                for (int e = 0; e < 3; ++e) csum[e] += h * ((3.0 / 4.0) * fcp[3 * f + e] + (1.0 / 4.0) * est[e]);
                vsum += h;
            }
            for (int e = 0; e < 3; ++e) ccv[3 * k + e] = std::fabs(vsum) > 1e-300 ? csum[e] / vsum : est[e];
            volv[k] = vsum * (1.0 / 3.0);
        }
        ccp = ccv.data();
        vp = volv.data();
    }
    for (long long k = 0; k < nC; ++k)
        if (!(vp[k] > 0)) return fail(c, SDFIBM_ERR_MESH, "cell " + std::to_string(k) + " has no positive volume");
    // bins: a uniform grid over the cell centres' box, about 4 cells per bin; each bin's box
    // holds its cells' point boxes
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (long long k = 0; k < nC; ++k)
        for (int e = 0; e < 3; ++e) {
            lo[e] = std::min(lo[e], ccp[3 * k + e]);
            hi[e] = std::max(hi[e], ccp[3 * k + e]);
        }
    double ext[3], vol_box = 1.0;
    int dims = 0;
    for (int e = 0; e < 3; ++e) {
        ext[e] = hi[e] - lo[e];
        if (ext[e] > 0) {
            vol_box *= ext[e];
            ++dims;
        }
    }
    const double want = std::max(1.0, (double)nC / 4.0);
    const double side = dims > 0 ? std::pow(vol_box / want, 1.0 / dims) : 1.0;
    PolyDev& P = c->pm;
    long long nb = 1;
    for (int e = 0; e < 3; ++e) {
        P.bn[e] = ext[e] > 0 ? (int)std::max(1.0, std::min(4096.0, std::floor(ext[e] / side))) : 1;
        P.bw[e] = ext[e] > 0 ? ext[e] / P.bn[e] : 1.0;
        P.binv[e] = 1.0 / P.bw[e];
        P.blo[e] = lo[e];
        nb *= P.bn[e];
    }
    std::vector<int> bid(nC), bo(nb + 1, 0), bc(nC);
    for (long long k = 0; k < nC; ++k) {
        long long b = 0;
        for (int e = 2; e >= 0; --e) {
            int i = (int)std::floor((ccp[3 * k + e] - P.blo[e]) * P.binv[e]);
            i = std::max(0, std::min(P.bn[e] - 1, i));
            b = b * P.bn[e] + i;
        }
        bid[k] = (int)b;
        ++bo[b + 1];
    }
    for (long long b = 0; b < nb; ++b) bo[b + 1] += bo[b];
    std::vector<double> bbox(6 * nb);
    for (long long b = 0; b < nb; ++b)
        for (int e = 0; e < 3; ++e) {
            bbox[6 * b + e] = 1e300;
            bbox[6 * b + 3 + e] = -1e300;
        }
    {
        std::vector<int> at(bo.begin(), bo.end() - 1);
        for (long long k = 0; k < nC; ++k) bc[at[bid[k]]++] = (int)k;  // ascending within a bin
    }
    // a cell's point box, and how far it reaches beyond its bin (in bins)
    int reach = 0;
    for (long long k = 0; k < nC; ++k) {
        double cl[3] = {1e300, 1e300, 1e300}, ch[3] = {-1e300, -1e300, -1e300};
        for (int t = cpo[k]; t < cpo[k + 1]; ++t)
            for (int e = 0; e < 3; ++e) {
                const double x = d->points[3 * (long long)cp[t] + e];
                cl[e] = std::min(cl[e], x);
                ch[e] = std::max(ch[e], x);
            }
        const long long b = bid[k];
        long long r = b;
        int ijk[3];
        for (int e = 0; e < 3; ++e) {
            ijk[e] = (int)(r % P.bn[e]);
            r /= P.bn[e];
        }
        for (int e = 0; e < 3; ++e) {
            bbox[6 * b + e] = std::min(bbox[6 * b + e], cl[e]);
            bbox[6 * b + 3 + e] = std::max(bbox[6 * b + 3 + e], ch[e]);
            const double blo = P.blo[e] + ijk[e] * P.bw[e], bhi = blo + P.bw[e];
            const double over = std::max(blo - cl[e], ch[e] - bhi);
            if (over > 0) reach = std::max(reach, (int)std::ceil(over * P.binv[e]));
        }
    }
    c->poly_reach = std::min(reach, 4096);
    // the mesh box (contact grid)
    for (int e = 0; e < 3; ++e) {
        c->box_lo[e] = 1e300;
        c->box_hi[e] = -1e300;
    }
    for (long long q = 0; q < nP; ++q)
        for (int e = 0; e < 3; ++e) {
            c->box_lo[e] = std::min(c->box_lo[e], d->points[3 * q + e]);
            c->box_hi[e] = std::max(c->box_hi[e], d->points[3 * q + e]);
        }
    for (int e = 0; e < 3; ++e) {
        P.mlo[e] = c->box_lo[e];
        P.mhi[e] = c->box_hi[e];
    }
    // uploads
    int rc;
    auto up = [&](const void* src, size_t bytes, void** dst) -> int {
        int r;
        if ((r = dalloc(c, dst, bytes))) return r;
        if (bytes && hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
            return fail(c, SDFIBM_ERR_HIP, "mesh upload failed");
        return SDFIBM_OK;
    };
    void* pts = nullptr;  // AoS: a point's three coordinates share a cache line (scattered gathers)
    if ((rc = up(d->points, nP * 24, &pts))) return rc;
    P.pxyz = (const double*)pts;
    std::vector<int> fo32(nF + 1), fp32(d->face_points, d->face_points + fo[nF]);
    for (long long f = 0; f <= nF; ++f) fo32[f] = (int)fo[f];
    void* p;
    if ((rc = up(fo32.data(), (nF + 1) * 4, &p))) return rc;
    P.f_off = (const int*)p;
    if ((rc = up(fp32.data(), fp32.size() * 4, &p))) return rc;
    P.f_pts = (const int*)p;
    if ((rc = up(fcp, nF * 24, &p))) return rc;
    P.fc = (const double*)p;
    if ((rc = up(sfp, nF * 24, &p))) return rc;
    P.sf = (const double*)p;
    if ((rc = up(cfo.data(), (nC + 1) * 4, &p))) return rc;
    P.c_foff = (const int*)p;
    if ((rc = up(cf.data(), cf.size() * 4, &p))) return rc;
    P.c_faces = (const int*)p;
    if ((rc = up(cpo.data(), (nC + 1) * 4, &p))) return rc;
    P.c_poff = (const int*)p;
    if ((rc = up(cp.data(), cp.size() * 4, &p))) return rc;
    P.c_pts = (const int*)p;
    if ((rc = up(cno.data(), (nC + 1) * 4, &p))) return rc;
    P.c_noff = (const int*)p;
    if ((rc = up(cn.data(), std::max<size_t>(cn.size(), 1) * 4, &p))) return rc;
    P.c_nbs = (const int*)p;
    if ((rc = up(ccp, nC * 24, &p))) return rc;
    P.cc = (const double*)p;
    if ((rc = up(vp, nC * 8, &p))) return rc;
    P.vol = (const double*)p;
    if ((rc = up(bo.data(), (nb + 1) * 4, &p))) return rc;
    P.b_off = (const int*)p;
    if ((rc = up(bc.data(), nC * 4, &p))) return rc;
    P.b_cells = (const int*)p;
    if ((rc = up(bbox.data(), nb * 48, &p))) return rc;
    P.b_box = (const double*)p;
    P.n_cells = (int)nC;
    P.n_faces = (int)nF;
    P.n_points = (int)nP;
    P.two_d = c->two_d;
    // per-cell work buffers of the step
    if ((rc = dalloc(c, (void**)&c->g_ccnt, nC * 4))) return rc;
    if ((rc = dalloc(c, (void**)&c->g_coff, (nC + 1) * 8))) return rc;
    return SDFIBM_OK;
}

// A buffer grown to hold at least n elements of `bytes` each (with headroom).
template <class T>
int poly_grow(sdfibm_ctx* c, T*& buf, long long& cap, long long n, size_t bytes) {
    if (n <= cap && buf) return SDFIBM_OK;
    dfree(c, buf);
    cap = 0;
    const long long want = n + n / 4 + 1024;
    int rc;
    if ((rc = dalloc(c, (void**)&buf, want * bytes))) return rc;
    cap = want;
    return SDFIBM_OK;
}

// Per-solid buffers of the step for n solids.
int poly_set_solids(sdfibm_ctx* c, const sdfibm_solid* solids, int64_t n) {
    for (int64_t s = 0; s < n; ++s) {
        std::string why;
        const int v = validate_solid(solids[s], s, why, (long long)c->h_csg.size());
        if (v) return fail(c, v, why);
    }
    c->have_solids = false;
    c->have_step = false;
    c->have_ct = false;
    c->ns = 0;
    c->contact_ready = false;
    int rc;
    if (n > c->solid_cap) {
        for (void** b : {(void**)&c->d_solids, (void**)&c->d_K, (void**)&c->d_ft, (void**)&c->d_ncand,
                         (void**)&c->g_box, (void**)&c->g_seed, (void**)&c->g_first, (void**)&c->g_vcnt,
                         (void**)&c->g_voff, (void**)&c->g_prange, (void**)&c->g_ccoff, (void**)&c->g_pnch,
                         (void**)&c->g_pcoff}) {
            void* q = *b;
            dfree(c, q);
            *b = nullptr;
        }
        c->solid_cap = 0;
        const long long cap = std::max<long long>(n, 16);
        if ((rc = dalloc(c, (void**)&c->d_solids, cap * sizeof(sdfibm_solid)))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_K, cap * sizeof(SolidK)))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_ft, cap * 48))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_ncand, cap * sizeof(int2)))) return rc;
        if ((rc = dalloc(c, (void**)&c->g_box, cap * 48))) return rc;
        if ((rc = dalloc(c, (void**)&c->g_seed, cap * 4))) return rc;
        if ((rc = dalloc(c, (void**)&c->g_first, cap * 4))) return rc;
        if ((rc = dalloc(c, (void**)&c->g_vcnt, cap * 8))) return rc;
        if ((rc = dalloc(c, (void**)&c->g_voff, (cap + 1) * 8))) return rc;
        if ((rc = dalloc(c, (void**)&c->g_prange, cap * 16))) return rc;
        if ((rc = dalloc(c, (void**)&c->g_ccoff, (cap + 1) * 8))) return rc;
        if ((rc = dalloc(c, (void**)&c->g_pnch, cap * 8))) return rc;
        if ((rc = dalloc(c, (void**)&c->g_pcoff, (cap + 1) * 8))) return rc;
        c->solid_cap = cap;
    }
    if (n > 0) {
        HIPCHK(c, hipMemcpy(c->d_solids, solids, n * sizeof(sdfibm_solid), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemset(c->d_ft, 0, n * 48));
    }
    c->h_solids.assign(solids, solids + n);
    c->ns = n;
    c->have_solids = true;
    return SDFIBM_OK;
}

// One step on a polyhedral mesh: pairs, labels, codes, force/torque, fields.
template <class R>
int poly_step(sdfibm_ctx* c, double dt) {
    const int ns = (int)c->ns;
    const long long nC = c->pm.n_cells;
    hipStream_t st = c->stream;
    const unsigned G = (unsigned)(4 * c->n_cu);
    int rc;
    (void)hipEventRecord(c->ev[0], st);
    long long np = 0;
    if (ns > 0) {
        if ((rc = poly_grow(c, c->g_tmp, c->g_tmpcap, std::max<long long>(ns, nC) / kScanTile + 2, 8))) return rc;
        k_poly_setup<<<grid_for(ns, 4), 256, 0, st>>>(c->d_solids, ns, c->pm, c->d_K, c->g_box, c->g_seed, c->g_vcnt,
                                                      c->d_ft, c->g_first, c->d_csg, c->poly_reach);
        // the solids' bins in chunks (a wave each): counted, scanned, then listed
        scan<long long>(c->g_vcnt, ns, c->g_ccoff, c->g_tmp, st);
        long long nchunk = 0;
        HIPCHK(c, hipMemcpyAsync(&nchunk, c->g_ccoff + ns, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if ((rc = poly_grow(c, c->g_csol, c->g_csol_cap, nchunk + 1, 4))) return rc;
        if ((rc = poly_grow(c, c->g_ccnt_v, c->g_ccnt_cap, nchunk + 1, 8))) return rc;
        if ((rc = poly_grow(c, c->g_cvoff, c->g_cvoff_cap, nchunk + 1, 8))) return rc;
        if ((rc = poly_grow(c, c->g_tmp, c->g_tmpcap, nchunk / kScanTile + 2, 8))) return rc;
        k_poly_chunk_map<<<grid_for(ns, 4), 256, 0, st>>>(ns, c->g_ccoff, c->g_csol);
        if (nchunk > 0)
            k_poly_visit<true><<<grid_for(nchunk, 4), 256, 0, st>>>(nchunk, c->pm, c->d_K, c->g_box, c->g_csol,
                                                                    c->g_ccoff, c->g_ccnt_v, nullptr, nullptr, nullptr,
                                                                    c->poly_reach);
        scan<long long>(c->g_ccnt_v, nchunk, c->g_cvoff, c->g_tmp, st);
        k_poly_solid_voff<<<grid_for(ns + 1, 256), 256, 0, st>>>(ns, c->g_ccoff, c->g_cvoff, c->g_voff);
        long long nv = 0;
        HIPCHK(c, hipMemcpyAsync(&nv, c->g_cvoff + nchunk, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (nv > INT_MAX) return fail(c, SDFIBM_ERR_CAPACITY, "more than 2^31-1 (solid, cell) visits");
        if (nv + 1 > c->g_vcap || !c->g_vcell) {  // the visit arrays, one capacity
            for (void** b : {(void**)&c->g_vcell, (void**)&c->g_vsolid, (void**)&c->g_act, (void**)&c->g_all,
                             (void**)&c->g_vpos}) {
                void* q = *b;
                dfree(c, q);
                *b = nullptr;
            }
            c->g_vcap = 0;
            const long long want = nv + 1 + (nv + 1) / 4 + 1024;
            if ((rc = dalloc(c, (void**)&c->g_vcell, want * 4))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_vsolid, want * 4))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_act, want))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_all, want))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_vpos, want * 8))) return rc;
            c->g_vcap = want;
        }
        if ((rc = poly_grow(c, c->g_tmp, c->g_tmpcap, nv / kScanTile + 2, 8))) return rc;
        if (nchunk > 0)
            k_poly_visit<false><<<grid_for(nchunk, 4), 256, 0, st>>>(nchunk, c->pm, c->d_K, c->g_box, c->g_csol,
                                                                     c->g_ccoff, nullptr, c->g_cvoff, c->g_vcell,
                                                                     c->g_vsolid, c->poly_reach);
        if (nv > 0)
            k_poly_test<<<G, 256, 0, st>>>(c->pm, c->d_K, nv, c->g_vcell, c->g_vsolid, c->g_act, c->g_all, c->g_first);
        scan<unsigned char>(c->g_act, nv, c->g_vpos, c->g_tmp, st);
        HIPCHK(c, hipMemcpyAsync(&np, c->g_vpos + nv, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (np > INT_MAX / 2) return fail(c, SDFIBM_ERR_CAPACITY, "more than 2^30 (solid, cell) pairs");
        if (np > c->g_pcap || !c->g_pcell) {  // the pair arrays, one capacity
            for (void** b : {(void**)&c->g_pcell, (void**)&c->g_psolid, (void**)&c->g_label, (void**)&c->g_pall,
                             (void**)&c->g_pcode, (void**)&c->g_pf, (void**)&c->g_clist}) {
                void* q = *b;
                dfree(c, q);
                *b = nullptr;
            }
            c->g_pcap = 0;
            const long long want = np + np / 4 + 1024;
            if ((rc = dalloc(c, (void**)&c->g_pcell, want * 4))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_psolid, want * 4))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_label, want * 4))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_pall, want))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_pcode, want * 8))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_pf, want * 48))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_clist, want * 4))) return rc;
            c->g_pcap = want;
        }
        // hash table: a power of two of at least twice the pairs
        int bits = 10;
        while ((1LL << bits) < 2 * np) ++bits;
        if (bits > c->g_bits) {
            dfree(c, c->g_hk);
            dfree(c, c->g_hv);
            c->g_bits = 0;
            if ((rc = dalloc(c, (void**)&c->g_hk, (1LL << bits) * 8))) return rc;
            if ((rc = dalloc(c, (void**)&c->g_hv, (1LL << bits) * 4))) return rc;
            c->g_bits = bits;
        }
        bits = c->g_bits;
        k_fill_i32<<<grid_for(2LL << bits, 256), 256, 0, st>>>((int*)c->g_hk, 2LL << bits, -1);  // kPolyEmpty keys
        if (np > 0) {
            k_poly_compact<<<G, 256, 0, st>>>(nv, c->g_vcell, c->g_vsolid, c->g_act, c->g_all, c->g_vpos, c->g_pcell,
                                              c->g_psolid, c->g_pall, c->g_label, c->g_hk, c->g_hv, bits);
            k_poly_hook<<<G, 256, 0, st>>>(c->pm, np, c->g_pcell, c->g_psolid, c->g_label, c->g_hk, c->g_hv, bits);
            k_poly_compress<<<G, 256, 0, st>>>(np, c->g_label);
            k_poly_force<R><<<4 * G, 64, 0, st>>>(c->pm, c->d_K, np, c->g_pcell, c->g_psolid, c->g_pall, c->g_label,
                                               c->g_seed, c->g_first, c->g_hk, c->g_hv, bits, (const R*)c->U[0],
                                               (const R*)c->U[1], (const R*)c->U[2], (R*)c->g_pcode, c->g_pf, dt);
        }
        // per solid: its pair range in chunks (a wave each: partial sums), then the chunks' sums
        // in chunk order
        k_poly_pair_chunks<<<grid_for(ns, 256), 256, 0, st>>>(ns, c->g_voff, c->g_vpos, c->g_prange, c->g_pnch);
        if ((rc = poly_grow(c, c->g_tmp, c->g_tmpcap, ns / kScanTile + 2, 8))) return rc;
        scan<long long>(c->g_pnch, ns, c->g_pcoff, c->g_tmp, st);
        const long long maxch = ns + np / kPolyPairChunk + 1;  // >= sum of ceil(pairs / chunk)
        if ((rc = poly_grow(c, c->g_psol, c->g_psol_cap, maxch, 4))) return rc;
        if ((rc = poly_grow(c, c->g_part, c->g_part_cap, 8 * maxch, 8))) return rc;
        k_poly_chunk_map<<<grid_for(ns, 4), 256, 0, st>>>(ns, c->g_pcoff, c->g_psol);
        k_poly_pair_part<R><<<grid_for(maxch, 4), 256, 0, st>>>(maxch, c->g_psol, c->g_pcoff, c->g_prange,
                                                                (const R*)c->g_pcode, c->g_pf, c->g_part, ns);
        k_poly_reduce<<<grid_for(ns, 4), 256, 0, st>>>(ns, c->g_pcoff, c->g_part, c->rho_f, c->d_ft, c->d_ncand);
    }
    c->g_np = np;
    (void)hipEventRecord(c->ev[1], st);
    // the member pairs of each cell
    if ((rc = poly_grow(c, c->g_tmp, c->g_tmpcap, nC / kScanTile + 2, 8))) return rc;
    k_fill_i32<<<grid_for(nC, 256), 256, 0, st>>>(c->g_ccnt, nC, 0);
    if (np > 0) k_poly_cells<R><<<G, 256, 0, st>>>(np, c->g_pcell, (const R*)c->g_pcode, c->g_ccnt, nullptr, nullptr);
    scan<int>(c->g_ccnt, nC, c->g_coff, c->g_tmp, st);
    k_fill_i32<<<grid_for(nC, 256), 256, 0, st>>>(c->g_ccnt, nC, 0);
    if (np > 0)
        k_poly_cells<R><<<G, 256, 0, st>>>(np, c->g_pcell, (const R*)c->g_pcode, c->g_ccnt, c->g_coff, c->g_clist);
    (void)hipEventRecord(c->ev[2], st);
    k_poly_fields<R><<<grid_for(nC, 256), 256, 0, st>>>(c->pm, c->d_K, c->g_coff, c->g_clist, c->g_psolid,
                                                        (const R*)c->g_pcode, (const R*)c->U[0], (const R*)c->U[1],
                                                        (const R*)c->U[2], (R*)c->As, (R*)c->Fs[0], (R*)c->Fs[1],
                                                        (R*)c->Fs[2], (R*)c->Tsf, c->Ct, dt);
    (void)hipEventRecord(c->ev[3], st);
    HIPCHK(c, hipGetLastError());
    return SDFIBM_OK;
}

int poly_interact(sdfibm_ctx* c, double dt, double* ft) {
    // until this step succeeds, nothing reads the pair buffers (a failed growth frees them)
    c->have_step = false;
    c->have_ct = false;
    const int rc = c->prec == SDFIBM_FP64 ? poly_step<double>(c, dt) : poly_step<float>(c, dt);
    if (rc) return rc;
    if (ft && c->ns > 0) HIPCHK(c, hipMemcpyAsync(ft, c->d_ft, c->ns * 48, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms[4] = {0, 0, 0, 0};
    (void)hipEventElapsedTime(&ms[0], c->ev[0], c->ev[1]);
    (void)hipEventElapsedTime(&ms[1], c->ev[1], c->ev[2]);
    (void)hipEventElapsedTime(&ms[2], c->ev[2], c->ev[3]);
    (void)hipEventElapsedTime(&ms[3], c->ev[0], c->ev[3]);
    for (int t = 0; t < 7; ++t) c->timing[t] = 0.0;
    c->timing[0] = ms[0];
    c->timing[1] = ms[1];
    c->timing[2] = ms[2];
    c->timing[3] = ms[3];
    c->have_step = true;
    c->have_ct = true;
    return SDFIBM_OK;
}

// Candidate cells of one solid (the ALL_INSIDE, centre-inside and centre-outside
// sets, each ascending) from the last step's pairs.
int poly_candidates(sdfibm_ctx* c, int64_t solid, int64_t* counts, int64_t* cells, double* alpha, int64_t capacity) {
    long long pr[2] = {0, 0};
    int seed = -1;
    HIPCHK(c, hipMemcpy(pr, c->g_prange + 2 * solid, 16, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(&seed, c->g_seed + solid, 4, hipMemcpyDeviceToHost));
    const long long n = pr[1] - pr[0];
    const size_t rs = rsize(c);
    std::vector<int> pc(std::max<long long>(n, 1));
    std::vector<unsigned char> raw(std::max<long long>(n, 1) * rs);
    if (n > 0) {
        HIPCHK(c, hipMemcpy(pc.data(), c->g_pcell + pr[0], n * 4, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(raw.data(), (unsigned char*)c->g_pcode + pr[0] * rs, n * rs, hipMemcpyDeviceToHost));
    }
    std::vector<std::pair<long long, double>> lists[3];
    for (long long q = 0; q < n; ++q) {
        const double v = rs == 8 ? ((const double*)raw.data())[q] : (double)((const float*)raw.data())[q];
        if (std::isnan(v)) continue;
        const int t = std::isinf(v) ? 0 : (std::signbit(v) ? 2 : 1);
        lists[t].push_back({pc[q], std::isinf(v) ? 1.0 : std::fabs(v)});
    }
    long long first = -1;
    bool near_in = false;
    for (int t = 0; t < 3; ++t) {
        std::sort(lists[t].begin(), lists[t].end());
        counts[t] = (int64_t)lists[t].size();
        for (auto& e : lists[t]) {
            if (first < 0 || e.first < first) first = e.first;
            if (e.first == seed) near_in = true;
        }
    }
    counts[3] = counts[0] + counts[1] + counts[2] > 0 ? (near_in ? seed : first) : -1;
    long long k = 0;
    for (int t = 0; t < 3; ++t)
        for (auto& e : lists[t]) {
            if (k >= capacity) break;
            if (cells) cells[k] = e.first;
            if (alpha) alpha[k] = e.second;
            ++k;
        }
    return SDFIBM_OK;
}

// Context set-up shared by both mesh kinds: the configuration and the device
// (create_begin), then, once the mesh is on the device, the streams, events,
// flag words, resident fields and step metadata (create_end).
int create_begin(const sdfibm_config* cfg, sdfibm_ctx** out) {
    *out = nullptr;
    if (cfg->precision != SDFIBM_FP32 && cfg->precision != SDFIBM_FP64)
        return fail(nullptr, SDFIBM_ERR_ARG, "precision must be SDFIBM_FP32 or SDFIBM_FP64");
    if (!std::isfinite(cfg->rho_f)) return fail(nullptr, SDFIBM_ERR_ARG, "rho_f must be finite");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return fail(nullptr, SDFIBM_ERR_NODEVICE, "no HIP device available");
    }
    sdfibm_ctx* c = new sdfibm_ctx;
    (void)hipGetDevice(&c->device);
    c->prec = cfg->precision;
    c->two_d = cfg->two_d ? 1 : 0;
    c->rho_f = cfg->rho_f;
    *out = c;
    return SDFIBM_OK;
}

int create_end(sdfibm_ctx* c) {
    int rc;
    if (hipStreamCreate(&c->stream) != hipSuccess) return fail(c, SDFIBM_ERR_HIP, "hipStreamCreate failed");
    c->own_stream = true;
    if ((rc = resolve_stage(c))) return rc;
    if (hipStreamCreate(&c->aux) != hipSuccess) return fail(c, SDFIBM_ERR_HIP, "hipStreamCreate failed");
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(c, SDFIBM_ERR_HIP, "hipEventCreate failed");
    if (hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->joinc, hipEventDisableTiming) != hipSuccess)
        return fail(c, SDFIBM_ERR_HIP, "hipEventCreate failed");
    if (hipHostMalloc((void**)&c->h_flags, 64, hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void**)&c->d_hflags, c->h_flags, 0) != hipSuccess)
        return fail(c, SDFIBM_ERR_HIP, "pinned host buffer for the error flags failed");
    const size_t rs = rsize(c);
    for (int a = 0; a < 3; ++a) {
        if ((rc = dalloc(c, &c->U[a], c->N * rs))) return rc;
        if ((rc = dalloc(c, &c->Fs[a], c->N * rs))) return rc;
        if (hipMemset(c->U[a], 0, c->N * rs) != hipSuccess) return fail(c, SDFIBM_ERR_HIP, "memset U");
    }
    if ((rc = dalloc(c, &c->As, c->N * rs))) return rc;
    if ((rc = dalloc(c, &c->Tsf, c->N * rs))) return rc;
    if ((rc = dalloc(c, (void**)&c->Ct, c->N * 4))) return rc;
    if ((rc = dalloc(c, (void**)&c->err, 64))) return rc;
    if ((rc = dalloc(c, (void**)&c->qmeta, 64))) return rc;
    if ((rc = dalloc(c, (void**)&c->d_tile_total, 64))) return rc;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
        c->n_cu = prop.multiProcessorCount;
    return SDFIBM_OK;
}

}  // namespace

extern "C" {

int sdfibm_abi_version(void) { return SDFIBM_ABI_VERSION; }

const char* sdfibm_status_string(int status) {
    switch (status) {
    case SDFIBM_OK: return "ok";
    case SDFIBM_ERR_ARG: return "invalid argument";
    case SDFIBM_ERR_HIP: return "HIP runtime error";
    case SDFIBM_ERR_NOMEM: return "device allocation failed";
    case SDFIBM_ERR_SHAPE: return "invalid shape";
    case SDFIBM_ERR_TILE: return "solid tile too large";
    case SDFIBM_ERR_STATE: return "call out of order";
    case SDFIBM_ERR_MESH: return "unsupported mesh";
    case SDFIBM_ERR_CAPACITY: return "buffer capacity exceeded";
    case SDFIBM_ERR_NODEVICE: return "no HIP device";
    default: return "unknown status";
    }
}

const char* sdfibm_last_error(const sdfibm_ctx* c) { return c ? c->msg.c_str() : g_noctx_msg.c_str(); }

int sdfibm_create(const sdfibm_mesh_desc* d, const sdfibm_config* cfg, sdfibm_ctx** out) {
    if (!out) return fail(nullptr, SDFIBM_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!d || !cfg) return fail(nullptr, SDFIBM_ERR_ARG, "mesh or config is NULL");
    if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(nullptr, SDFIBM_ERR_ARG, "mesh dimensions must be >= 1");
    if ((long long)d->nx * d->ny * d->nz > (1LL << 31) - 1)
        return fail(nullptr, SDFIBM_ERR_MESH, "more than 2^31-1 cells");
    if (d->kind != SDFIBM_MESH_RECTILINEAR && d->kind != SDFIBM_MESH_STRUCTURED)
        return fail(nullptr, SDFIBM_ERR_ARG, "mesh kind must be RECTILINEAR or STRUCTURED (sdfibm_create_polymesh: POLY)");
    sdfibm_ctx* c = nullptr;
    int rc = create_begin(cfg, &c);
    if (rc) return rc;
    auto bail = [&](int code) {
        g_noctx_msg = c->msg;
        sdfibm_destroy(c);
        return code;
    };
    c->kind = d->kind;
    c->nx = d->nx;
    c->ny = d->ny;
    c->nz = d->nz;
    c->N = (long long)d->nx * d->ny * d->nz;
    if ((rc = setup_mesh(c, d))) return bail(rc);
    if ((rc = create_end(c))) return bail(rc);
    const bool flat = c->nz == 1;
    c->bx = flat ? kBX2 : kBX3;
    c->by = flat ? kBY2 : kBY3;
    c->bz = flat ? kBZ2 : kBZ3;
    c->nbx = (c->nx + c->bx - 1) / c->bx;
    c->nby = (c->ny + c->by - 1) / c->by;
    c->nbz = (c->nz + c->bz - 1) / c->bz;
    c->nbricks = (long long)c->nbx * c->nby * c->nbz;
    if (c->nby > 65535 || c->nbz > 65535)
        return bail(fail(c, SDFIBM_ERR_MESH, "mesh too large along y or z for the brick grid"));
    if ((rc = dalloc(c, (void**)&c->bcount, c->nbricks * 4))) return bail(rc);
    if ((rc = dalloc(c, (void**)&c->d_blist, c->nbricks * kMaxPerBrick * 4))) return bail(rc);
    *out = c;
    return SDFIBM_OK;
}

int sdfibm_create_polymesh(const sdfibm_polymesh_desc* d, const sdfibm_config* cfg, sdfibm_ctx** out) {
    if (!out) return fail(nullptr, SDFIBM_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!d || !cfg) return fail(nullptr, SDFIBM_ERR_ARG, "mesh or config is NULL");
    sdfibm_ctx* c = nullptr;
    int rc = create_begin(cfg, &c);
    if (rc) return rc;
    auto bail = [&](int code) {
        g_noctx_msg = c->msg;
        sdfibm_destroy(c);
        return code;
    };
    c->kind = SDFIBM_MESH_POLY;
    c->nx = (int)std::min<long long>(d->n_cells, INT_MAX);
    c->ny = c->nz = 1;
    c->N = d->n_cells;
    if ((rc = poly_setup(c, d))) return bail(rc);
    if ((rc = create_end(c))) return bail(rc);
    *out = c;
    return SDFIBM_OK;
}

int sdfibm_destroy(sdfibm_ctx* c) {
    if (!c) return SDFIBM_ERR_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->aux) (void)hipStreamSynchronize(c->aux);
    graph_reset(c);
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->joinc) (void)hipEventDestroy(c->joinc);
    if (c->h_flags) (void)hipHostFree(c->h_flags);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SDFIBM_OK;
}

int sdfibm_set_stream(sdfibm_ctx* c, void* s) {
    if (!c) return SDFIBM_ERR_ARG;
    graph_reset(c);
    if (c->own_stream && c->stream) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
    }
    c->own_stream = false;
    c->stream = (hipStream_t)s;
    if (!s) {
        if (hipStreamCreate(&c->stream) != hipSuccess) return fail(c, SDFIBM_ERR_HIP, "hipStreamCreate failed");
        c->own_stream = true;
    }
    return SDFIBM_OK;
}

int sdfibm_set_solids(sdfibm_ctx* c, const sdfibm_solid* solids, int64_t n) {
    if (!c) return SDFIBM_ERR_ARG;
    if (n < 0 || (n > 0 && !solids)) return fail(c, SDFIBM_ERR_ARG, "bad solid array");
    if (n > (1LL << 30)) return fail(c, SDFIBM_ERR_ARG, "too many solids");
    if (c->kind == SDFIBM_MESH_POLY) return poly_set_solids(c, solids, n);
    // Per-step updates of the same solids (poses and velocities only; the reference's solver
    // hands over its moved solids every step): no tile estimates, buffers, classes or captured
    // step to rebuild -- the interact grows its code buffer itself if moving tiles need it.
    bool same = c->have_solids && n == c->ns && n > 0;
    for (int64_t s = 0; same && s < n; ++s)
        same = solids[s].shape == c->h_solids[s].shape &&
               std::memcmp(solids[s].prm, c->h_solids[s].prm, sizeof(solids[s].prm)) == 0;
    if (same) {
        for (int64_t s = 0; s < n; ++s) {
            std::string why;
            const int v = validate_solid(solids[s], s, why, (long long)c->h_csg.size());
            if (v) return fail(c, v, why);
        }
        HIPCHK(c, hipMemsetAsync(c->d_ft, 0, n * 48, c->stream));  // evolve before any interact: no fluid force
        HIPCHK(c, hipMemcpyAsync(c->d_solids, solids, n * sizeof(sdfibm_solid), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->have_step = false;
        return SDFIBM_OK;
    }
    for (int64_t s = 0; s < n; ++s) {
        std::string why;
        const int v = validate_solid(solids[s], s, why, (long long)c->h_csg.size());
        if (v) return fail(c, v, why);
    }
    return classify_solids(c, solids, n);
}

namespace {

// The device address of `scratch` on the context's device, looked up once per
// device (each GPU holds its own copy of the array) and kept for every context.
int resolve_stage(sdfibm_ctx* c) {
    static std::mutex mu;
    static std::vector<double*> per_device;
    std::lock_guard<std::mutex> lock(mu);
    if (c->device < 0) return fail(c, SDFIBM_ERR_NODEVICE, "no current HIP device");
    if ((int)per_device.size() <= c->device) per_device.resize(c->device + 1, nullptr);
    if (!per_device[c->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(c, SDFIBM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        per_device[c->device] = (double*)p;
    }
    c->stage = per_device[c->device];
    return SDFIBM_OK;
}

// Finest cell spacing per axis: the rectilinear coordinates' smallest step; a
// structured block's nominal spacing less twice the largest point displacement.
void min_spacing(const sdfibm_ctx* c, double* h) {
    if (c->kind == SDFIBM_MESH_RECTILINEAR) {
        const std::vector<double>* v[3] = {&c->h_xs, &c->h_ys, &c->h_zs};
        for (int d = 0; d < 3; ++d) {
            h[d] = __builtin_huge_val();
            for (size_t t = 1; t < v[d]->size(); ++t) h[d] = std::min(h[d], (*v[d])[t] - (*v[d])[t - 1]);
        }
    } else {
        for (int d = 0; d < 3; ++d) h[d] = c->hm.h[d] * std::max(0.05, 1.0 - 2.0 * c->hm.jit);
    }
}

// Buffers of the big-tile pass: metadata once, the vertex bits and chunk partials
// grown to at least the given sizes (with headroom).
int grow_big(sdfibm_ctx* c, long long words, long long chunks) {
    int rc;
    if (!c->d_big_meta && (rc = dalloc(c, (void**)&c->d_big_meta, kBigMaxSolids * sizeof(BigMeta)))) return rc;
    if (!c->d_big_tot && (rc = dalloc(c, (void**)&c->d_big_tot, kBigTotN * 8))) return rc;
    if (words > c->big_word_cap) {
        graph_reset(c);
        dfree(c, c->d_big_vbits);
        dfree(c, c->d_big_bcnt);
        dfree(c, c->d_big_boff);
        c->big_word_cap = 0;
        const long long want = words + words / 4 + 1024;
        if ((rc = dalloc(c, (void**)&c->d_big_vbits, want * 8))) return rc;
        // cell blocks: at most want / 16 + kBigMaxSolids (k_big_layout checks)
        const long long blocks = want / 16 + kBigMaxSolids;
        if ((rc = dalloc(c, (void**)&c->d_big_bcnt, blocks * sizeof(int2)))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_big_boff, blocks * sizeof(int2)))) return rc;
        c->big_word_cap = want;
    }
    if (chunks > c->big_chunk_cap) {
        graph_reset(c);
        dfree(c, c->d_big_part);
        c->big_chunk_cap = 0;
        const long long want = chunks + chunks / 4 + 256;
        if ((rc = dalloc(c, (void**)&c->d_big_part, want * 6 * 8))) return rc;
        c->big_chunk_cap = want;
    }
    return SDFIBM_OK;
}

// The full set-up of a solid array: tile classes, which solids take the planes' path, buffer
// sizes, uploads.  Used by set_solids for new solids and by interact when a solid's tile
// outgrew the tile pass since set_solids (it then takes the planes' path, as it would in a
// fresh context).  The solids are valid (validate_solid).  Until it succeeds the context holds
// no solids: a failed allocation leaves no freed buffer reachable from a later call.
int classify_solids(sdfibm_ctx* c, const sdfibm_solid* solids, int64_t n) {
    graph_reset(c);
    std::vector<int> planes;
    long long cells = 0, tile_sum = 0, n_tiles = 0, big_words = 0, big_chunks = 0, n_lm = 0;
    bool big_on = false;
    int hist[34] = {0};  // tiles per power of two of their cell count; [32] planes and oversized, [33] big
    double hmin[3];
    min_spacing(c, hmin);
    const long long nn[3] = {c->nx, c->ny, c->nz};
    for (int64_t s = 0; s < n; ++s) {
        const SolidK K = make_solid_k(solids[s], c->h_csg.data());
        if (solids[s].shape == SDFIBM_SHAPE_PLANE) {
            planes.push_back((int)s);
            continue;
        }
        TileInfo t = c->kind == SDFIBM_MESH_RECTILINEAR ? make_tile<MK_RECT>(c->hm, K) : make_tile<MK_STRUCT>(c->hm, K);
        if (t.cls == TC_TOOBIG) {  // oversized: the unbounded path, like a plane
            planes.push_back((int)s);
            ++hist[32];
            continue;
        }
        if (t.cls != TC_EMPTY) {
            const long long nt = (long long)t.n[0] * t.n[1] * t.n[2];
            ++hist[t.cls == TC_BIG ? 33 : std::min(31, 63 - __builtin_clzll((unsigned long long)nt))];
        }
        // can the solid's tile, in any pose, outgrow the on-chip tile pass?  (its bounding
        // sphere in cells of the finest spacing: about the box of its local bounds, or its
        // largest semi-axis for spheres and ellipsoids; the z-extruded 2-D shapes reach that far
        // in the plane normal to their local z, whose current direction they span whole)
        {
            d3 lc, lh;
            long long e[3] = {nn[0], nn[1], nn[2]};
            if (local_bounds(K, lc, lh)) {
                const bool quad = K.shape == SDFIBM_SHAPE_SPHERE || K.shape == SDFIBM_SHAPE_ELLIPSOID;
                const bool extruded = !(lh.z < 1e300);
                double R;
                if (quad)
                    R = std::sqrt(dot(lc, lc)) + std::max(lh.x, std::max(lh.y, lh.z));
                else if (extruded)
                    R = std::sqrt(lc.x * lc.x + lc.y * lc.y) + std::sqrt(lh.x * lh.x + lh.y * lh.y);
                else
                    R = std::sqrt(dot(lc, lc)) + std::sqrt(dot(lh, lh));
                double Rm[3][3];
                rot_of(K.q, Rm);
                for (int d = 0; d < 3; ++d) {
                    const bool along = extruded && std::fabs(Rm[d][2]) > 1e-12;  // spans the mesh
                    if (std::isfinite(R) && !along)
                        e[d] = std::min<long long>(nn[d], (long long)std::ceil(2.0 * R / hmin[d]) + 2);
                }
            }
            if (e[0] > kMaxTileX || (e[1] + 1) * (e[2] + 1) > kRowsSmall)
                ++n_lm;  // a medium (or big) tile in some pose
            if (e[0] > kMaxTileX || (e[1] + 1) * (e[2] + 1) > kRowsMedium) {
                big_on = true;
                big_words += (e[0] + 2) * (e[1] + 2) * (e[2] + 2) / 64 + 2;
                big_chunks += e[0] * e[1] * e[2] / kBigChunk + 2;
            }
        }
        if (t.cls == TC_EMPTY) continue;
        if (t.cls != TC_BIG) {
            tile_sum += (long long)t.n[0] * t.n[1] * t.n[2];
            ++n_tiles;
        }
        // allow one extra cell layer per axis for motion between set_solids calls
        const long long ex = std::min(t.n[0] + 1, c->nx), ey = std::min(t.n[1] + 1, c->ny),
                        ez = std::min(t.n[2] + 1, c->nz);
        cells += ex * ey * ez;
    }
    if ((int)planes.size() > kMaxPlanes)
        return fail(c, SDFIBM_ERR_ARG,
                    "at most " + std::to_string(kMaxPlanes) +
                        " plane solids and solids whose cell tile has more than 2^22 vertices");
    c->have_solids = false;
    c->have_step = false;
    c->have_ct = false;
    c->ns = 0;
    c->contact_ready = false;
    int rc;
    if (n > c->solid_cap) {
        dfree(c, c->d_solids);
        dfree(c, c->d_K);
        dfree(c, c->d_T);
        dfree(c, c->d_ft);
        c->solid_cap = 0;
        const long long cap = std::max<long long>(n, 16);
        if ((rc = dalloc(c, (void**)&c->d_solids, cap * sizeof(sdfibm_solid)))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_K, cap * sizeof(SolidK)))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_T, cap * sizeof(TileInfo)))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_ft, cap * 6 * 8))) return rc;
        dfree(c, c->d_ncand);
        if ((rc = dalloc(c, (void**)&c->d_ncand, cap * sizeof(int2)))) return rc;
        dfree(c, c->qlist);
        if ((rc = dalloc(c, (void**)&c->qlist, cap * kQueues * 4))) return rc;
        dfree(c, c->d_unb);
        if ((rc = dalloc(c, (void**)&c->d_unb, cap))) return rc;
        dfree(c, c->scan_tmp);
        if ((rc = dalloc(c, (void**)&c->scan_tmp, ((cap + kScanTile - 1) / kScanTile + 2) * 8))) return rc;
        c->solid_cap = cap;
    }
    const long long want_codes = cells + cells / 20 + 1024;
    if (want_codes > c->code_cap) {
        dfree(c, c->d_codes);
        dfree(c, c->d_cand);
        c->code_cap = 0;
        if ((rc = dalloc(c, &c->d_codes, want_codes * rsize(c)))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_cand, want_codes * 4))) return rc;
        c->code_cap = want_codes;
    }
    dfree(c, c->d_planes);
    dfree(c, c->d_plane_partial);
    if (!planes.empty()) {
        if ((rc = dalloc(c, (void**)&c->d_planes, planes.size() * 4))) return rc;
        HIPCHK(c, hipMemcpy(c->d_planes, planes.data(), planes.size() * 4, hipMemcpyHostToDevice));
        if ((rc = dalloc(c, (void**)&c->d_plane_partial, planes.size() * c->nbricks * 6 * 8))) return rc;
    }
    if (!planes.empty()) {  // mesh-sized labelling buffers, kept once allocated
        c->member_words = (c->N + 31) / 32;
        if (!c->d_pact && (rc = dalloc(c, (void**)&c->d_pact, c->member_words * 4))) return rc;
        if (!c->d_plist && (rc = dalloc(c, (void**)&c->d_plist, c->N * 4))) return rc;
        if (!c->d_plabel && (rc = dalloc(c, (void**)&c->d_plabel, c->N * 4))) return rc;
        if (!c->d_pscalars && (rc = dalloc(c, (void**)&c->d_pscalars, 64))) return rc;
    }
    dfree(c, c->d_pmember);
    dfree(c, c->d_pseed);
    dfree(c, c->d_pcode);
    if (!planes.empty()) {
        if ((rc = dalloc(c, (void**)&c->d_pmember, planes.size() * c->member_words * 4))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_pseed, planes.size() * 4))) return rc;
        if ((rc = dalloc(c, &c->d_pcode, (long long)planes.size() * c->N * rsize(c)))) return rc;
        if (!c->d_plane_cache && (rc = dalloc(c, (void**)&c->d_plane_cache, 8 * kMaxPlanes * 8))) return rc;
        const std::vector<double> nan(8 * kMaxPlanes, std::nan(""));  // every plane labelled at the next interact
        HIPCHK(c, hipMemcpy(c->d_plane_cache, nan.data(), nan.size() * 8, hipMemcpyHostToDevice));
    }
    c->h_planes = planes;
    c->small_tiles = n_tiles > 0 && tile_sum < 512 * n_tiles;
    c->lm_grid = (int)std::min<long long>(n_lm, n);
    c->big_on = big_on;
    if (big_on && (rc = grow_big(c, big_words, big_chunks))) return rc;
    {
        std::string key = std::to_string(c->nx) + "x" + std::to_string(c->ny) + "x" + std::to_string(c->nz) +
                          (c->kind == SDFIBM_MESH_RECTILINEAR ? ":rect:" : ":struct:") +
                          (c->prec == SDFIBM_FP64 ? "fp64:" : "fp32:") + std::to_string(n) + ":";
        for (int b = 0; b < 34; ++b)
            if (hist[b]) key += std::to_string(b) + "=" + std::to_string(hist[b]) + ",";
        c->tile_key = key;
    }
    if (n > 0) {
        std::vector<unsigned char> unb(n, 0);
        for (int p : planes) unb[p] = 1;
        HIPCHK(c, hipMemcpy(c->d_unb, unb.data(), n, hipMemcpyHostToDevice));
    }
    c->h_solids.assign(solids, solids + n);
    if (n > 0) HIPCHK(c, hipMemsetAsync(c->d_ft, 0, n * 48, c->stream));  // evolve before any interact: no fluid force
    if (c->dyn_n != n) c->dyn_n = -1;
    if (n > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_solids, solids, n * sizeof(sdfibm_solid), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ns = n;
    c->have_solids = true;
    c->have_step = false;
    if (contact_on(c)) return contact_prepare(c);
    return SDFIBM_OK;
}

}  // namespace

int sdfibm_set_csg_shapes(sdfibm_ctx* c, const sdfibm_csg_shape* shapes, int64_t n) {
    if (!c) return SDFIBM_ERR_ARG;
    if (n < 0 || (n > 0 && !shapes)) return fail(c, SDFIBM_ERR_ARG, "bad CSG shape array");
    if (n > (1LL << 24)) return fail(c, SDFIBM_ERR_ARG, "too many CSG shapes");
    for (int64_t t = 0; t < n; ++t) {
        std::string why;
        if (const int v = validate_csg(shapes[t], t, why)) return fail(c, v, why);
    }
    // the solids refer to the table by index: they must be set again
    graph_reset(c);
    c->have_solids = false;
    c->have_step = false;
    c->have_ct = false;
    c->h_csg_raw.assign(shapes, shapes + n);
    c->h_csg.resize(n);
    for (int64_t t = 0; t < n; ++t) c->h_csg[t] = make_csg_k(shapes[t]);
    dfree(c, c->d_csg);
    int rc;
    if (n > 0) {
        if ((rc = dalloc(c, (void**)&c->d_csg, n * sizeof(CsgK)))) return rc;
        HIPCHK(c, hipMemcpy(c->d_csg, c->h_csg.data(), n * sizeof(CsgK), hipMemcpyHostToDevice));
    }
    return SDFIBM_OK;
}

int sdfibm_set_velocity(sdfibm_ctx* c, const double* U) {
    if (!c) return SDFIBM_ERR_ARG;
    if (!U) return fail(c, SDFIBM_ERR_ARG, "U is NULL");
    int rc;
    // the AoS staging buffer is allocated by the first call and kept (a solver hands U over
    // every step)
    if (!c->d_ustage && (rc = dalloc(c, (void**)&c->d_ustage, c->N * 24))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_ustage, U, c->N * 24, hipMemcpyHostToDevice, c->stream));
    launch_aos_to_soa(c, c->d_ustage, 0, c->N);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SDFIBM_OK;
}

int sdfibm_set_velocity_cells_noalloc(sdfibm_ctx* c, int64_t first, int64_t count, const double* U) {
    if (!c) return SDFIBM_ERR_ARG;
    if (first < 0 || count < 0 || first > c->N - count)
        return fail(c, SDFIBM_ERR_ARG, "cells [first, first + count) must lie in the mesh");
    if (count > kStageCells)
        return fail(c, SDFIBM_ERR_ARG,
                    "count " + std::to_string(count) + " exceeds the staging area of " + std::to_string(kStageCells) +
                        " cells: upload in pieces");
    if (count > 0 && !U) return fail(c, SDFIBM_ERR_ARG, "U is NULL");
    if (!c->stage) return fail(c, SDFIBM_ERR_STATE, "no staging area on this device");
    if (count == 0) return SDFIBM_OK;
    // the copy overwrites the `count` staged values the kernel reads: nothing to clear
    HIPCHK(c, hipMemcpyAsync(c->stage, U, count * 24, hipMemcpyHostToDevice, c->stream));
    launch_aos_to_soa(c, c->stage, first, count);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SDFIBM_OK;
}

int sdfibm_interact(sdfibm_ctx* c, double time, double dt, double* ft) {
    (void)time;
    if (!c) return SDFIBM_ERR_ARG;
    if (!c->have_solids) return fail(c, SDFIBM_ERR_STATE, "interact before set_solids");
    if (!(dt > 0) || !std::isfinite(dt)) return fail(c, SDFIBM_ERR_ARG, "dt must be finite and > 0");
    if (c->kind == SDFIBM_MESH_POLY) return poly_interact(c, dt, ft);
    int rc = dispatch_interact(c, dt);
    if (rc) return rc;
    // the flags come back through a pinned mapped word (no copy-engine round trip)
    k_flags_out<<<1, 64, 0, c->stream>>>(c->err, c->d_hflags);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int flags = *(volatile int*)c->h_flags;
    if ((flags & EF_TILE) && c->ns > 0) {
        // a solid moved or turned since set_solids until its tile no longer fits the tile pass:
        // classify the current solids again (that solid takes the planes' path, as in a fresh
        // context given these solids) and run the step again
        std::vector<sdfibm_solid> cur(c->ns);
        HIPCHK(c, hipMemcpy(cur.data(), c->d_solids, c->ns * sizeof(sdfibm_solid), hipMemcpyDeviceToHost));
        if ((rc = classify_solids(c, cur.data(), (int64_t)cur.size()))) return rc;
        if ((rc = dispatch_interact(c, dt))) return rc;
        k_flags_out<<<1, 64, 0, c->stream>>>(c->err, c->d_hflags);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
        flags = *(volatile int*)c->h_flags;
    }
    for (int attempt = 0; attempt < 3 && (flags & EF_CODECAP) && !(flags & EF_TILE) && c->ns > 0; ++attempt) {
        // the solids moved or turned since set_solids and their tiles outgrew the code buffer
        // (or the big-tile pass's buffers): grow them to the sizes this step needs (with
        // headroom) and run the step again
        long long need = 0;
        HIPCHK(c, hipMemcpy(&need, c->d_tile_total, 8, hipMemcpyDeviceToHost));
        if (need > c->code_cap) {
            const long long want = need + need / 4 + 1024;
            graph_reset(c);
            dfree(c, c->d_codes);
            dfree(c, c->d_cand);
            c->code_cap = 0;
            if ((rc = dalloc(c, &c->d_codes, want * rsize(c)))) return rc;
            if ((rc = dalloc(c, (void**)&c->d_cand, want * 4))) return rc;
            c->code_cap = want;
        }
        if (c->big_on) {
            long long bt[kBigTotN];
            HIPCHK(c, hipMemcpy(bt, c->d_big_tot, sizeof(bt), hipMemcpyDeviceToHost));
            if ((rc = grow_big(c, bt[5], bt[6]))) return rc;
        }
        if ((rc = dispatch_interact(c, dt))) return rc;
        k_flags_out<<<1, 64, 0, c->stream>>>(c->err, c->d_hflags);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
        flags = *(volatile int*)c->h_flags;
    }
    if (ft && c->ns > 0) HIPCHK(c, hipMemcpyAsync(ft, c->d_ft, c->ns * 48, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms[5] = {0, 0, 0, 0, 0};
    if (c->phase_timing) {
        (void)hipEventElapsedTime(&ms[0], c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&ms[1], c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&ms[2], c->ev[2], c->ev[3]);
        (void)hipEventElapsedTime(&ms[3], c->ev[1], c->ev[4]);
        (void)hipEventElapsedTime(&ms[4], c->ev[4], c->ev[7]);
    }
    c->timing[0] = ms[0];
    c->timing[1] = ms[1];
    c->timing[2] = ms[2];
    c->timing[3] = (double)ms[0] + ms[1] + ms[2];
    c->timing[4] = ms[3];
    c->timing[5] = ms[4];
    float msb = 0.f;
    if (c->big_on && c->phase_timing) (void)hipEventElapsedTime(&msb, c->ev[5], c->ev[6]);
    c->timing[6] = msb;
    if (flags) {
        c->have_step = false;
        c->have_ct = false;
        return fail(c, (flags & EF_TILE) ? SDFIBM_ERR_TILE : SDFIBM_ERR_CAPACITY, err_text(flags));
    }
    c->have_step = true;
    c->have_ct = true;
    return SDFIBM_OK;
}

int sdfibm_update_velocity(sdfibm_ctx* c, double dt) {
    if (!c) return SDFIBM_ERR_ARG;
    if (!std::isfinite(dt)) return fail(c, SDFIBM_ERR_ARG, "dt must be finite");
    launch_update_velocity(c, dt);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SDFIBM_OK;
}

int sdfibm_fix_internal(sdfibm_ctx* c) {
    if (!c) return SDFIBM_ERR_ARG;
    if (!c->have_ct) return fail(c, SDFIBM_ERR_STATE, "fix_internal needs the Ct field of a successful interact");
    if (c->kind == SDFIBM_MESH_POLY) {
        const unsigned g = grid_for(c->N, 256);
        if (c->prec == SDFIBM_FP64)
            k_poly_fix<double><<<g, 256, 0, c->stream>>>(c->pm, c->Ct, c->d_solids, (int)c->ns, (double*)c->U[0],
                                                         (double*)c->U[1], (double*)c->U[2]);
        else
            k_poly_fix<float><<<g, 256, 0, c->stream>>>(c->pm, c->Ct, c->d_solids, (int)c->ns, (float*)c->U[0],
                                                        (float*)c->U[1], (float*)c->U[2]);
    } else {
        launch_fix_internal(c);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SDFIBM_OK;
}

int sdfibm_set_dynamics(sdfibm_ctx* c, const sdfibm_dynamics* dyn, int64_t n) {
    if (!c) return SDFIBM_ERR_ARG;
    graph_reset(c);
    if (!c->have_solids) return fail(c, SDFIBM_ERR_STATE, "set_dynamics before set_solids");
    if (n != c->ns || (n > 0 && !dyn)) return fail(c, SDFIBM_ERR_ARG, "need one dynamics record per solid");
    std::vector<SolidDyn> h(n);
    for (int64_t s = 0; s < n; ++s) {
        if (!(dyn[s].rho > 0) || !std::isfinite(dyn[s].rho))
            return fail(c, SDFIBM_ERR_ARG, "solid " + std::to_string(s) + ": density must be finite and > 0");
        if (dyn[s].motion_mask < -1 || dyn[s].motion_mask > 63)
            return fail(c, SDFIBM_ERR_ARG, "solid " + std::to_string(s) + ": motion mask must be -1 or 6 bits");
        h[s].rho = dyn[s].rho;
        h[s].mask = dyn[s].motion_mask;
        h[s].first_fluid = 1;
        for (int a = 0; a < 3; ++a) h[s].ff_old[a] = h[s].ft_old[a] = 0.0;
        h[s].vol = 0.0;
        h[s].moi[0] = h[s].moi[1] = h[s].moi[2] = 1.0;
        if (c->h_solids[s].shape == SDFIBM_SHAPE_CSG) {  // validated index (set_solids)
            const sdfibm_csg_shape& g = c->h_csg_raw[(size_t)c->h_solids[s].prm[0]];
            h[s].vol = g.volume;
            for (int a = 0; a < 3; ++a) h[s].moi[a] = g.moi[a];
        }
    }
    int rc;
    c->dyn_n = -1;  // no dynamics until the records below are on the device
    if (c->d_dyn == nullptr || c->dyn_cap < n) {
        dfree(c, c->d_dyn);
        c->dyn_cap = 0;
        if ((rc = dalloc(c, (void**)&c->d_dyn, std::max<long long>(n, 1) * sizeof(SolidDyn)))) return rc;
        c->dyn_cap = n;
    }
    if (n > 0) HIPCHK(c, hipMemcpy(c->d_dyn, h.data(), n * sizeof(SolidDyn), hipMemcpyHostToDevice));
    c->dyn_n = n;
    return SDFIBM_OK;
}

namespace {

bool contact_on(const sdfibm_ctx* c) { return c->law.k > 0 || c->law.c > 0; }

// Bounding radii, the contact grid and its buffers for the current solids.
int contact_prepare(sdfibm_ctx* c) {
    c->contact_ready = false;  // until the buffers below exist: evolve and run_steps refuse
    const long long n = c->ns;
    std::vector<double> rb(std::max<long long>(n, 1), 0.0);
    double rmax = 0.0;
    for (long long s = 0; s < n; ++s) {
        rb[s] = c->h_solids[s].shape == SDFIBM_SHAPE_CSG ? c->h_csg_raw[(size_t)c->h_solids[s].prm[0]].radius_b
                                                         : bounding_radius(c->h_solids[s]);
        rmax = std::max(rmax, rb[s]);
    }
    double cell = 2.0 * rmax, ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = c->box_hi[a] - c->box_lo[a];
    if (!(cell > 0)) cell = std::max(ext[0], std::max(ext[1], ext[2]));
    ContactGrid g{};
    long long total = 0;
    for (;;) {
        total = 1;
        for (int a = 0; a < 3; ++a) {
            g.n[a] = (int)std::max(1.0, std::min(1.0e6, std::ceil(ext[a] / cell)));
            total *= g.n[a];
        }
        if (total <= (1LL << 22)) break;
        cell *= 1.25;  // coarser cells: more candidates per solid, never a missed pair
    }
    for (int a = 0; a < 3; ++a) g.lo[a] = c->box_lo[a];
    g.inv_cell = 1.0 / cell;
    int rc;
    if (n > c->contact_cap) {
        dfree(c, c->d_rb);
        dfree(c, c->d_solids_alt);
        dfree(c, c->d_cnext);
        c->contact_cap = 0;
        if ((rc = dalloc(c, (void**)&c->d_rb, n * 8))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_solids_alt, n * sizeof(sdfibm_solid)))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_cnext, 2 * n * 4))) return rc;
        c->contact_cap = n;
    }
    if (total > c->chead_cap) {
        dfree(c, c->d_chead);
        c->chead_cap = 0;
        if ((rc = dalloc(c, (void**)&c->d_chead, 3 * total * 4))) return rc;
        c->chead_cap = total;
    }
    if (n > 0) HIPCHK(c, hipMemcpy(c->d_rb, rb.data(), n * 8, hipMemcpyHostToDevice));
    c->cgrid = g;
    c->ncells_c = (int)total;
    c->contact_ready = true;
    return SDFIBM_OK;
}

// evolve and run_steps: every buffer their launches read exists for the current solids.
int check_dynamics_ready(sdfibm_ctx* c, const char* who) {
    if (!c->have_solids) return fail(c, SDFIBM_ERR_STATE, std::string(who) + " before set_solids");
    if (c->dyn_n != c->ns || c->dyn_n < 0)
        return fail(c, SDFIBM_ERR_STATE, std::string(who) + " needs set_dynamics for these solids");
    if (contact_on(c) && !c->contact_ready)
        return fail(c, SDFIBM_ERR_STATE,
                    std::string(who) + ": the contact buffers are missing (an allocation failed); call set_contact again");
    return SDFIBM_OK;
}

}  // namespace

int sdfibm_set_contact(sdfibm_ctx* c, double stiffness, double damping) {
    if (!c) return SDFIBM_ERR_ARG;
    graph_reset(c);
    if (!(stiffness >= 0) || !(damping >= 0) || !std::isfinite(stiffness) || !std::isfinite(damping))
        return fail(c, SDFIBM_ERR_ARG, "contact stiffness and damping must be finite and >= 0");
    c->law = ContactLaw{stiffness, damping};
    if (contact_on(c) && c->have_solids) return contact_prepare(c);
    return SDFIBM_OK;
}

int sdfibm_evolve(sdfibm_ctx* c, double time, double dt, const double* gravity) {
    (void)time;
    if (!c) return SDFIBM_ERR_ARG;
    if (int rc = check_dynamics_ready(c, "evolve")) return rc;
    if (!(dt > 0) || !std::isfinite(dt)) return fail(c, SDFIBM_ERR_ARG, "dt must be finite and > 0");
    Vec3 g{{0.0, 0.0, 0.0}};
    if (gravity)
        for (int a = 0; a < 3; ++a) g.v[a] = gravity[a];
    if (c->ns == 0) return SDFIBM_OK;
    launch_evolve(c, dt, g, nullptr);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_step = false;  // the tiles of the last interact no longer match the solids
    return SDFIBM_OK;
}

int sdfibm_run_steps(sdfibm_ctx* c, double time, double dt, int64_t n_steps, const double* gravity, double* ft) {
    if (!c) return SDFIBM_ERR_ARG;
    if (c->kind == SDFIBM_MESH_POLY)
        return fail(c, SDFIBM_ERR_STATE, "run_steps needs a structured mesh (a polyMesh step is not captured)");
    if (int rc = check_dynamics_ready(c, "run_steps")) return rc;
    if (!(dt > 0) || !std::isfinite(dt)) return fail(c, SDFIBM_ERR_ARG, "dt must be finite and > 0");
    if (n_steps < 0) return fail(c, SDFIBM_ERR_ARG, "n_steps must be >= 0");
    Vec3 g{{0.0, 0.0, 0.0}};
    if (gravity)
        for (int a = 0; a < 3; ++a) g.v[a] = gravity[a];
    if (!std::isfinite(g.v[0]) || !std::isfinite(g.v[1]) || !std::isfinite(g.v[2]))
        return fail(c, SDFIBM_ERR_ARG, "gravity must be finite");
    const int ns = (int)c->ns;
    int rc;
    // an eager pass over the current state: plane memberships and the code buffer size (it
    // also moves a solid whose tile outgrew the tile pass to the planes' path)
    if (n_steps > 0 && (rc = sdfibm_interact(c, time, dt, nullptr))) return rc;
    const int np = (int)c->h_planes.size();
    // the plane memberships are computed once, before the replays: the planes must stay put
    std::vector<double> pose(7 * std::max(np, 1));
    for (int p = 0; p < np; ++p) {
        sdfibm_solid ps;
        SolidDyn pd;
        HIPCHK(c, hipMemcpy(&ps, c->d_solids + c->h_planes[p], sizeof(ps), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(&pd, c->d_dyn + c->h_planes[p], sizeof(pd), hipMemcpyDeviceToHost));
        bool still = pd.mask == 0;
        for (int a = 0; a < 3; ++a) still = still && ps.vel[a] == 0.0 && ps.omega[a] == 0.0;
        if (!still)
            return fail(c, SDFIBM_ERR_STATE,
                        "run_steps needs fixed plane solids and fixed oversized solids (motion mask b000000, "
                        "zero velocity and spin): solid " +
                            std::to_string(c->h_planes[p]));
        for (int a = 0; a < 3; ++a) pose[7 * p + a] = ps.center[a];
        for (int a = 0; a < 4; ++a) pose[7 * p + 3 + a] = ps.q[a];
    }
    if (n_steps == 0) return SDFIBM_OK;
    long long need = 0;
    if (ns > 0) HIPCHK(c, hipMemcpy(&need, c->d_tile_total, 8, hipMemcpyDeviceToHost));
    const long long want = need + need / 2 + 4096;  // headroom: tiles grow as the solids move and turn
    if (want > c->code_cap) {
        graph_reset(c);
        dfree(c, c->d_codes);
        dfree(c, c->d_cand);
        c->code_cap = 0;
        if ((rc = dalloc(c, &c->d_codes, want * rsize(c)))) return rc;
        if ((rc = dalloc(c, (void**)&c->d_cand, want * 4))) return rc;
        c->code_cap = want;
    }
    if (c->big_on) {  // the same headroom for the big-tile pass
        long long bt[kBigTotN];
        HIPCHK(c, hipMemcpy(bt, c->d_big_tot, sizeof(bt), hipMemcpyDeviceToHost));
        if ((rc = grow_big(c, bt[5] + bt[5] / 2 + 64, bt[6] + bt[6] / 2 + 16))) return rc;
    }
    if (!c->d_sticky && (rc = dalloc(c, (void**)&c->d_sticky, 64))) return rc;
    if (!c->d_plane_pose && (rc = dalloc(c, (void**)&c->d_plane_pose, 7 * kMaxPlanes * 8))) return rc;
    if (np > 0) HIPCHK(c, hipMemcpy(c->d_plane_pose, pose.data(), 7 * np * 8, hipMemcpyHostToDevice));
    if (c->gexec && (c->graph_dt != dt || c->graph_g[0] != g.v[0] || c->graph_g[1] != g.v[1] || c->graph_g[2] != g.v[2]))
        graph_reset(c);
    if (!c->gexec) {
        // one coupled step (main.cpp:66-84 without the PISO solve), captured once
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int r = dispatch_interact(c, dt, true);
        if (r == SDFIBM_OK) {
            // a step whose interact failed, and every later one, leaves the solids and U alone:
            // the state stays the last valid step's, and the caller can grow buffers and go on
            k_flag_or<<<1, 64, 0, c->stream>>>(c->err, c->d_sticky);
            launch_evolve(c, dt, g, c->d_sticky);
            launch_update_fix(c, dt, c->d_sticky);  // U -= Fs dt, then fixInternal with the evolved solids
            if (np > 0)
                k_plane_static<<<1, 64, 0, c->stream>>>(c->d_solids, c->d_planes, np, c->d_plane_pose, c->d_sticky);
        }
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &gr);
        if (r != SDFIBM_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return r;
        }
        if (e != hipSuccess) return fail(c, SDFIBM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(gr);
            c->gexec = nullptr;
            return fail(c, SDFIBM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
        }
        c->graph = gr;
        c->graph_dt = dt;
        for (int a = 0; a < 3; ++a) c->graph_g[a] = g.v[a];
    }
    k_fill_i32<<<1, 64, 0, c->stream>>>(c->d_sticky, 1, 0);
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    for (int64_t t = 0; t < n_steps; ++t) HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    int flags = 0;
    HIPCHK(c, hipMemcpyAsync(&flags, c->d_sticky, 4, hipMemcpyDeviceToHost, c->stream));
    if (ft && ns > 0) HIPCHK(c, hipMemcpyAsync(ft, c->d_ft, ns * 48, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[3]);
    for (double& t : c->timing) t = 0.0;
    c->timing[3] = ms;
    c->have_step = false;  // the solids moved after the last interact
    c->have_ct = flags == 0;
    if (flags)
        return fail(c,
                    (flags & EF_TILE)          ? SDFIBM_ERR_TILE
                    : (flags & EF_PLANE_MOVED) ? SDFIBM_ERR_STATE
                                               : SDFIBM_ERR_CAPACITY,
                    err_text(flags) + "the solids and U are those of the last valid step");
    return SDFIBM_OK;
}

int sdfibm_interact_steps(sdfibm_ctx* c, double time, double dt, int64_t n_steps, double* ft) {
    if (!c) return SDFIBM_ERR_ARG;
    if (c->kind == SDFIBM_MESH_POLY)
        return fail(c, SDFIBM_ERR_STATE, "interact_steps needs a structured mesh (a polyMesh step is not captured)");
    if (!c->have_solids) return fail(c, SDFIBM_ERR_STATE, "interact_steps before set_solids");
    if (!(dt > 0) || !std::isfinite(dt)) return fail(c, SDFIBM_ERR_ARG, "dt must be finite and > 0");
    if (n_steps < 0) return fail(c, SDFIBM_ERR_ARG, "n_steps must be >= 0");
    if (n_steps == 0) return SDFIBM_OK;
    // an eager pass: plane memberships (the replays keep them: nothing moves between
    // interacts) and buffer sizes
    int rc = sdfibm_interact(c, time, dt, nullptr);
    if (rc) return rc;
    if (!c->d_sticky && (rc = dalloc(c, (void**)&c->d_sticky, 64))) return rc;
    if (c->gexec_i && c->graph_i_dt != dt) {
        (void)hipGraphExecDestroy(c->gexec_i);
        (void)hipGraphDestroy(c->graph_i);
        c->gexec_i = nullptr;
        c->graph_i = nullptr;
    }
    if (!c->gexec_i) {
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int r = dispatch_interact(c, dt, true);
        if (r == SDFIBM_OK) k_flag_or<<<1, 64, 0, c->stream>>>(c->err, c->d_sticky);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &gr);
        if (r != SDFIBM_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return r;
        }
        if (e != hipSuccess) return fail(c, SDFIBM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&c->gexec_i, gr, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(gr);
            c->gexec_i = nullptr;
            return fail(c, SDFIBM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
        }
        c->graph_i = gr;
        c->graph_i_dt = dt;
    }
    k_fill_i32<<<1, 64, 0, c->stream>>>(c->d_sticky, 1, 0);
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    for (int64_t t = 0; t < n_steps; ++t) HIPCHK(c, hipGraphLaunch(c->gexec_i, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    int flags = 0;
    HIPCHK(c, hipMemcpyAsync(&flags, c->d_sticky, 4, hipMemcpyDeviceToHost, c->stream));
    if (ft && c->ns > 0) HIPCHK(c, hipMemcpyAsync(ft, c->d_ft, c->ns * 48, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[3]);
    for (double& t : c->timing) t = 0.0;
    c->timing[3] = ms;
    c->have_step = flags == 0;
    c->have_ct = flags == 0;
    if (flags)
        return fail(c, (flags & EF_TILE) ? SDFIBM_ERR_TILE : SDFIBM_ERR_CAPACITY, err_text(flags));
    return SDFIBM_OK;
}

int sdfibm_set_surrogate(sdfibm_ctx* c, const float* params, double band_lo, double band_hi, double max_angle_deg) {
    if (!c) return SDFIBM_ERR_ARG;
    if (c->kind == SDFIBM_MESH_POLY && params)
        return fail(c, SDFIBM_ERR_STATE, "the surrogate needs a rectilinear mesh");
    graph_reset(c);
    if (!params) {
        dfree(c, c->d_sg);
        c->d_sg = nullptr;
        return SDFIBM_OK;
    }
    if (!(band_lo <= band_hi) || !std::isfinite(band_lo) || !std::isfinite(band_hi) || !(max_angle_deg <= 180))
        return fail(c, SDFIBM_ERR_ARG, "bad surrogate band or angle");
    for (int p = 0; p < kSgParams; ++p)
        if (!std::isfinite(params[p])) return fail(c, SDFIBM_ERR_ARG, "non-finite surrogate parameter");
    SgNet h;
    sg_pack(params, h);
    int rc;
    if (!c->d_sg && (rc = dalloc(c, (void**)&c->d_sg, sizeof(SgNet)))) return rc;
    if (!c->d_sg_count && (rc = dalloc(c, (void**)&c->d_sg_count, 64))) return rc;
    HIPCHK(c, hipMemcpy(c->d_sg, &h, sizeof(SgNet), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->d_sg_count, 0, 4));
    c->sg_lo = (float)band_lo;
    c->sg_hi = (float)band_hi;
    // a negative angle: every cell whose prediction lies in the band falls back
    c->sg_cos = max_angle_deg < 0 ? 2.0f : (float)std::cos(max_angle_deg * 3.14159265358979323846 / 180.0);
    return SDFIBM_OK;
}

int sdfibm_surrogate_count(sdfibm_ctx* c, int64_t* n) {
    if (!c || !n) return SDFIBM_ERR_ARG;
    int v = 0;
    if (c->d_sg_count) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(&v, c->d_sg_count, 4, hipMemcpyDeviceToHost));
    }
    *n = v;
    return SDFIBM_OK;
}

int sdfibm_get_solids(sdfibm_ctx* c, sdfibm_solid* out, int64_t n) {
    if (!c) return SDFIBM_ERR_ARG;
    if (n != c->ns || (n > 0 && !out)) return fail(c, SDFIBM_ERR_ARG, "need room for every solid");
    if (n > 0) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(out, c->d_solids, n * sizeof(sdfibm_solid), hipMemcpyDeviceToHost));
    }
    return SDFIBM_OK;
}

int sdfibm_get_velocity(sdfibm_ctx* c, double* U) {
    if (!c) return SDFIBM_ERR_ARG;
    if (!U) return fail(c, SDFIBM_ERR_ARG, "U is NULL");
    const size_t rs = rsize(c);
    std::vector<unsigned char> buf(c->N * rs);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; ++a) {
        HIPCHK(c, hipMemcpy(buf.data(), c->U[a], c->N * rs, hipMemcpyDeviceToHost));
        for (long long q = 0; q < c->N; ++q)
            U[3 * q + a] = rs == 8 ? ((const double*)buf.data())[q] : (double)((const float*)buf.data())[q];
    }
    return SDFIBM_OK;
}

int sdfibm_buffer_capacity(sdfibm_ctx* c, int64_t* out, int n) {
    if (!c || !out || n < 0) return SDFIBM_ERR_ARG;
    const long long v[2] = {c->code_cap, c->solid_cap};
    const int k = n < 2 ? n : 2;
    for (int t = 0; t < k; ++t) out[t] = v[t];
    return k;
}

int sdfibm_set_phase_timing(sdfibm_ctx* c, int32_t on) {
    if (!c) return SDFIBM_ERR_ARG;
    c->phase_timing = on != 0;
    return SDFIBM_OK;
}

int sdfibm_last_timing(sdfibm_ctx* c, double* ms, int n) {
    if (!c || !ms || n < 0) return SDFIBM_ERR_ARG;
    const int k = n < 7 ? n : 7;
    for (int t = 0; t < k; ++t) ms[t] = c->timing[t];
    return k;
}

int sdfibm_get_fields(sdfibm_ctx* c, double* As, double* Fs, double* Ts, double* Ct) {
    if (!c) return SDFIBM_ERR_ARG;
    const long long N = c->N;
    const size_t rs = rsize(c);
    std::vector<unsigned char> buf(N * rs);
    auto fetch = [&](void* src, double* dst, int stride, int comp) -> int {
        HIPCHK(c, hipMemcpy(buf.data(), src, N * rs, hipMemcpyDeviceToHost));
        if (rs == 8) {
            const double* b = (const double*)buf.data();
            for (long long q = 0; q < N; ++q) dst[stride * q + comp] = b[q];
        } else {
            const float* b = (const float*)buf.data();
            for (long long q = 0; q < N; ++q) dst[stride * q + comp] = (double)b[q];
        }
        return SDFIBM_OK;
    };
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int rc;
    if (As && (rc = fetch(c->As, As, 1, 0))) return rc;
    if (Ts && (rc = fetch(c->Tsf, Ts, 1, 0))) return rc;
    if (Fs)
        for (int a = 0; a < 3; ++a)
            if ((rc = fetch(c->Fs[a], Fs, 3, a))) return rc;
    if (Ct) {
        std::vector<int> ci(N);
        HIPCHK(c, hipMemcpy(ci.data(), c->Ct, N * 4, hipMemcpyDeviceToHost));
        for (long long q = 0; q < N; ++q) Ct[q] = (double)ci[q];
    }
    return SDFIBM_OK;
}

int sdfibm_field_pointers(sdfibm_ctx* c, void** U3, void** As, void** Fs3, void** Ts, int32_t** Ct) {
    if (!c) return SDFIBM_ERR_ARG;
    if (U3)
        for (int a = 0; a < 3; ++a) U3[a] = c->U[a];
    if (Fs3)
        for (int a = 0; a < 3; ++a) Fs3[a] = c->Fs[a];
    if (As) *As = c->As;
    if (Ts) *Ts = c->Tsf;
    if (Ct) *Ct = c->Ct;
    return SDFIBM_OK;
}

int sdfibm_candidates(sdfibm_ctx* c, int64_t solid, int64_t* counts, int64_t* cells, double* alpha,
                      int64_t capacity) {
    if (!c) return SDFIBM_ERR_ARG;
    if (!c->have_step) return fail(c, SDFIBM_ERR_STATE, "candidates before a successful interact");
    if (solid < 0 || solid >= c->ns || !counts) return fail(c, SDFIBM_ERR_ARG, "bad solid index or counts");
    if (c->kind == SDFIBM_MESH_POLY) return poly_candidates(c, solid, counts, cells, alpha, capacity);
    const size_t rs = rsize(c);
    TileInfo T;
    HIPCHK(c, hipMemcpy(&T, c->d_T + solid, sizeof(TileInfo), hipMemcpyDeviceToHost));
    std::vector<unsigned char> raw;
    long long nt = 0;
    int lo[3] = {0, 0, 0}, n[3] = {0, 0, 0};
    if (T.cls == TC_PLANE || T.cls == TC_HUGE) {
        nt = c->N;
        raw.resize(nt * rs);
        void* tmp = nullptr;
        int rc;
        if ((rc = dalloc(c, &tmp, nt * rs))) return rc;
        const int sp = (int)solid;
        const int p = (int)(std::find(c->h_planes.begin(), c->h_planes.end(), (int)solid) - c->h_planes.begin());
        const unsigned* mem = c->d_pmember + (long long)p * c->member_words;
        if (c->prec == SDFIBM_FP64) {
            if (c->kind == SDFIBM_MESH_RECTILINEAR)
                k_plane_codes<double, MK_RECT><<<grid_for(nt, 256), 256, 0, c->stream>>>(c->m, c->d_K, sp, mem, (double*)tmp);
            else
                k_plane_codes<double, MK_STRUCT><<<grid_for(nt, 256), 256, 0, c->stream>>>(c->m, c->d_K, sp, mem, (double*)tmp);
        } else {
            if (c->kind == SDFIBM_MESH_RECTILINEAR)
                k_plane_codes<float, MK_RECT><<<grid_for(nt, 256), 256, 0, c->stream>>>(c->m, c->d_K, sp, mem, (float*)tmp);
            else
                k_plane_codes<float, MK_STRUCT><<<grid_for(nt, 256), 256, 0, c->stream>>>(c->m, c->d_K, sp, mem, (float*)tmp);
        }
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(raw.data(), tmp, nt * rs, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dfree(c, tmp);
        if (e != hipSuccess) return fail(c, SDFIBM_ERR_HIP, std::string("plane codes: ") + hipGetErrorString(e));
        n[0] = c->nx;
        n[1] = c->ny;
        n[2] = c->nz;
    } else if (tile_is_finite(T.cls)) {
        nt = (long long)T.n[0] * T.n[1] * T.n[2];
        raw.resize(nt * rs);
        HIPCHK(c, hipMemcpy(raw.data(), (unsigned char*)c->d_codes + T.off * rs, nt * rs, hipMemcpyDeviceToHost));
        for (int a = 0; a < 3; ++a) {
            lo[a] = T.lo[a];
            n[a] = T.n[a];
        }
    }
    std::vector<long long> lists[3];
    std::vector<double> al[3];
    for (long long q = 0; q < nt; ++q) {
        const double v = rs == 8 ? ((const double*)raw.data())[q] : (double)((const float*)raw.data())[q];
        if (std::isnan(v)) continue;
        const long long li = q % n[0], lj = (q / n[0]) % n[1], lk = q / ((long long)n[0] * n[1]);
        const long long cell = (lo[0] + li) + (long long)c->nx * ((lo[1] + lj) + (long long)c->ny * (lo[2] + lk));
        int t;
        double a;
        if (std::isinf(v)) {
            t = 0;
            a = 1.0;
        } else if (std::signbit(v)) {
            t = 2;
            a = -v;
        } else {
            t = 1;
            a = v;
        }
        lists[t].push_back(cell);
        al[t].push_back(a);
    }
    counts[0] = (int64_t)lists[0].size();
    counts[1] = (int64_t)lists[1].size();
    counts[2] = (int64_t)lists[2].size();
    counts[3] = -1;
    if (counts[0] + counts[1] + counts[2] > 0) {
        // the seed actually used: the nearest cell if it is a candidate, else the first candidate
        long long near = T.seed;
        long long first = -1;
        bool near_in = false;
        for (int t = 0; t < 3; ++t)
            for (long long v : lists[t]) {
                if (first < 0 || v < first) first = v;
                if (v == near) near_in = true;
            }
        counts[3] = near_in ? near : first;
    }
    long long k = 0;
    for (int t = 0; t < 3; ++t)
        for (size_t q = 0; q < lists[t].size(); ++q) {
            if (k < capacity) {
                if (cells) cells[k] = lists[t][q];
                if (alpha) alpha[k] = al[t][q];
            }
            ++k;
        }
    if (k > capacity && (cells || alpha)) return fail(c, SDFIBM_ERR_CAPACITY, "candidate buffer too small");
    return SDFIBM_OK;
}

int sdfibm_pair_counts(sdfibm_ctx* c, int64_t* out) {
    if (!c) return SDFIBM_ERR_ARG;
    if (!out) return fail(c, SDFIBM_ERR_ARG, "out is NULL");
    if (!c->have_ct) return fail(c, SDFIBM_ERR_STATE, "pair_counts before a successful interact");
    out[0] = out[1] = 0;
    const long long ns = c->ns;
    if (ns == 0) return SDFIBM_OK;
    std::vector<int2> nc(ns);
    HIPCHK(c, hipMemcpy(nc.data(), c->d_ncand, ns * sizeof(int2), hipMemcpyDeviceToHost));
    std::vector<char> is_plane(ns, 0);
    for (int p : c->h_planes) is_plane[p] = 1;
    for (long long s = 0; s < ns; ++s)
        if (!is_plane[s]) {
            out[0] += nc[s].x;
            out[1] += nc[s].y;
        }
    const bool had = c->have_step;
    c->have_step = true;  // the plane cell sets of the last interact are still in place
    for (int p : c->h_planes) {
        int64_t cnt[4] = {0, 0, 0, 0};
        const int rc = sdfibm_candidates(c, p, cnt, nullptr, nullptr, 0);
        if (rc) {
            c->have_step = had;
            return rc;
        }
        out[0] += cnt[0];
        out[1] += cnt[1] + cnt[2];
    }
    c->have_step = had;
    return SDFIBM_OK;
}


// ---------------------------------------------------------------------------
// run-time variants and their autotuner (no counterpart in the reference)
// ---------------------------------------------------------------------------
int sdfibm_set_tuning(sdfibm_ctx* c, int32_t cells_wave, int32_t big_chunk, int32_t big_grid) {
    if (!c) return SDFIBM_ERR_ARG;
    if (cells_wave < -1 || cells_wave > 1 || big_chunk < 64 || big_chunk > 16384 || big_grid < 1 || big_grid > 32)
        return fail(c, SDFIBM_ERR_ARG, "tuning: cells_wave -1..1, big_chunk 64..16384, big_grid 1..32");
    graph_reset(c);
    c->cells_wave = cells_wave;
    c->big_chunk = big_chunk;
    c->big_grid = big_grid;
    return SDFIBM_OK;
}

}  // extern "C"

namespace {

struct Tuning {
    int cells_wave, big_chunk, big_grid;
    double ms;
};

// Device time of one phase of the step (an index of sdfibm_ctx::timing) under the
// context's current variant: the median of `reps` interacts after two untimed ones.
int time_step(sdfibm_ctx* c, double dt, int reps, int phase, double& ms) {
    std::vector<double> t;
    const bool was = c->phase_timing;
    c->phase_timing = true;
    for (int r = 0; r < reps + 2; ++r) {
        const int rc = sdfibm_interact(c, 0.0, dt, nullptr);
        if (rc) {
            c->phase_timing = was;
            return rc;
        }
        if (r >= 2) t.push_back(c->timing[phase]);
    }
    c->phase_timing = was;
    std::sort(t.begin(), t.end());
    ms = t[t.size() / 2];
    return SDFIBM_OK;
}

// One line of the cache file for a key: {"key": "...", "cells_wave": a, "big_chunk": b, "big_grid": g, "ms": m}
bool cache_lookup(const char* path, const std::string& key, Tuning& out) {
    FILE* f = std::fopen(path, "r");
    if (!f) return false;
    char line[4096];
    bool hit = false;
    const std::string want = "\"key\": \"" + key + "\"";
    while (std::fgets(line, sizeof(line), f)) {
        const std::string l(line);
        if (l.find(want) == std::string::npos) continue;
        auto num = [&](const char* name, double& v) {
            const std::string k = std::string("\"") + name + "\": ";
            const size_t p = l.find(k);
            if (p == std::string::npos) return false;
            v = std::strtod(l.c_str() + p + k.size(), nullptr);
            return true;
        };
        double a, b, g, m;
        if (num("cells_wave", a) && num("big_chunk", b) && num("big_grid", g) && num("ms", m)) {
            out = Tuning{(int)a, (int)b, (int)g, m};
            hit = true;  // the last matching line wins
        }
    }
    std::fclose(f);
    return hit;
}

std::string tuning_json(const Tuning& t, const char* source, int tried) {
    char buf[256];
    std::snprintf(buf, sizeof(buf),
                  "{\"cells_wave\": %d, \"big_chunk\": %d, \"big_grid\": %d, \"ms\": %.6f, \"source\": \"%s\", "
                  "\"variants_timed\": %d}",
                  t.cells_wave, t.big_chunk, t.big_grid, t.ms, source, tried);
    return buf;
}

}  // namespace

extern "C" {

int sdfibm_autotune(sdfibm_ctx* c, double dt, const char* cache_path, char* chosen, int64_t chosen_cap) {
    if (!c) return SDFIBM_ERR_ARG;
    if (c->kind == SDFIBM_MESH_POLY) return fail(c, SDFIBM_ERR_STATE, "autotune: a polyMesh step has no variants");
    if (!c->have_solids) return fail(c, SDFIBM_ERR_STATE, "autotune before set_solids");
    if (!(dt > 0) || !std::isfinite(dt)) return fail(c, SDFIBM_ERR_ARG, "dt must be finite and > 0");
    Tuning best{c->cells_wave, c->big_chunk, c->big_grid, 0.0};
    int tried = 0;
    const char* source = "timed";
    int rc;
    if (cache_path && cache_lookup(cache_path, c->tile_key, best)) {
        source = "cache";
    } else {
        // per size bin, one knob at a time, each timed by its own phase: the small tiles'
        // cut-cell pass by its phase (timing[5]), then (with big tiles) the big pass's chunk
        // size and grid by the big pass (timing[6], on the second stream); the choice's step
        // time is measured last
        auto run = [&](Tuning t, int phase, double& ms) {
            if ((rc = sdfibm_set_tuning(c, t.cells_wave, t.big_chunk, t.big_grid))) return rc;
            ++tried;
            return time_step(c, dt, 5, phase, ms);
        };
        double bin = 1e300;
        best = Tuning{0, kBigChunk, 4, 0.0};
        for (int w : {0, 1}) {
            Tuning t{w, kBigChunk, 4, 0.0};
            double ms;
            if ((rc = run(t, 5, ms))) return rc;
            if (ms < bin) bin = ms, best = t;
        }
        if (c->big_on) {
            bin = 1e300;
            for (int ch : {256, 512, 1024, 2048, 4096}) {
                Tuning t{best.cells_wave, ch, best.big_grid, 0.0};
                double ms;
                if ((rc = run(t, 6, ms))) return rc;
                if (ms < bin) bin = ms, best = t;
            }
            for (int g : {2, 8}) {
                Tuning t{best.cells_wave, best.big_chunk, g, 0.0};
                double ms;
                if ((rc = run(t, 6, ms))) return rc;
                if (ms < bin) bin = ms, best = t;
            }
        }
        if ((rc = sdfibm_set_tuning(c, best.cells_wave, best.big_chunk, best.big_grid))) return rc;
        if ((rc = time_step(c, dt, 5, 3, best.ms))) return rc;
        if (cache_path) {
            if (FILE* f = std::fopen(cache_path, "a")) {
                std::fprintf(f, "{\"key\": \"%s\", \"cells_wave\": %d, \"big_chunk\": %d, \"big_grid\": %d, \"ms\": %.6f}\n",
                             c->tile_key.c_str(), best.cells_wave, best.big_chunk, best.big_grid, best.ms);
                std::fclose(f);
            }
        }
    }
    if ((rc = sdfibm_set_tuning(c, best.cells_wave, best.big_chunk, best.big_grid))) return rc;
    if (chosen && chosen_cap > 0) {
        const std::string j = tuning_json(best, source, tried);
        std::snprintf(chosen, (size_t)chosen_cap, "%s", j.c_str());
    }
    return SDFIBM_OK;
}

}  // extern "C"
