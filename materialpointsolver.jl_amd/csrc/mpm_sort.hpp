// mpm_sort.hpp -- periodic on-device sort of the particle streams by stencil base.
//
// A particle's stencil base is the lowest node of its S-node window per axis
// (floor(u - 1/2) for uGIMP, the reference's cell corner for the linear
// basis, u = (x - x1) / h); particles with the same base form a "dual cell"
// and scatter to the same S^D nodes.  Every sort_interval steps the particles
// are reordered by the key
//
//     (tile of the base) << 11 | (rank inside the dual cell) << 6 | (base inside the tile)
//
// with tiles of 4x4x4 bases in 3D (8x8 in 2D) and the rank counted with one
// integer atomic per particle.  Inside a tile, the particles of one rank then
// have pairwise distinct bases, which is what the tiled scatter kernels need
// (mpm_tiled.hpp).  The keys are placed by a counting sort: the per-base
// counts give each tile's layout in (rank group, base) order (k_tile_layout),
// a scan of the tile sizes gives the tiles' offsets, and every particle
// computes its position from its tile, rank and base (k_sort_dest); no
// comparison or radix passes.  The permutation is one gather kernel over all
// particle streams into each stream's own alternate buffer, then a pointer swap,
// so every stream flips between the same two buffers and the particle
// pointers take one of two configurations (parity), which is what lets
// captured step graphs be reused across sorts.  The original particle index
// travels with the particles (pid) so downloads return the caller's order.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/mpm_hip.h"
#include "mpm_kernels.hpp"
#include "mpm_tiled.hpp"

namespace mpm {

struct SortKeyGeom {
    int nb[3];   // stencil bases per axis (nodes - S + 1)
    int tsh;     // log2 tile edge in bases
    int tx, ty;  // tiles per axis (x, y)
};

// Sort key of a frozen particle; valid keys stay below it (the tile count is
// checked against the 32-bit key at set-up).
constexpr unsigned kFrozenKey = 0xffffffffu;

template <int DIM, int S>
__global__ __launch_bounds__(256) void k_sort_keys(GridGeom g, SortKeyGeom kg,
                                                   const double* __restrict__ x0,
                                                   const double* __restrict__ x1,
                                                   const double* __restrict__ x2,
                                                   unsigned* __restrict__ count,
                                                   unsigned* __restrict__ keys,
                                                   int* __restrict__ rank_out, int np,
                                                   int* __restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double* xs[3] = {x0, x1, x2};
    int b[3] = {0, 0, 0};
    bool inside = true, finite = true;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        const double u = (xs[a][i] - g.org[a]) * g.hinv[a];
        const double f = stencil_base_f<S>(a, u);
        finite = finite && isfinite(u);
        if (!(f >= 0.0 && f <= double(kg.nb[a] - 1))) inside = false;  // also NaN
        else b[a] = int(f);
    }
    if (!inside) {
        // a frozen particle (its stencil left the grid, or x is not finite):
        // placed after every tile by k_sort_dest, in no tile segment, and
        // reported as the kernels that skip it report it
        keys[i] = kFrozenKey;
        rank_out[i] = 0;
        atomicOr(flags, finite ? int(FLAG_RANGE) : int(FLAG_NONFINITE));
        return;
    }
    const unsigned lin = unsigned(b[0]) + unsigned(kg.nb[0]) * (unsigned(b[1]) + unsigned(kg.nb[1]) * unsigned(b[2]));
    const unsigned rank = atomicAdd(&count[lin], 1u);
    const int tm = (1 << kg.tsh) - 1;
    const unsigned tile = unsigned(b[0] >> kg.tsh) +
                          unsigned(kg.tx) * (unsigned(b[1] >> kg.tsh) + unsigned(kg.ty) * unsigned(b[2] >> kg.tsh));
    const unsigned local = unsigned(tile_local<DIM>(kg.tsh, b[0] & tm, b[1] & tm, b[2] & tm));
    keys[i] = (tile << kTileShift) | (min(rank, (1u << kRankBits) - 1) << kLocalBits) | local;
    rank_out[i] = int(rank);
}

// Counting sort by key.  The keys' order inside a tile is (rank group, base):
// for a rank r below the last group the r-th particle of base l goes to
//   start + off[r] + popcount(mask[r] & bases below l),
// with mask[r] the bases holding more than r particles and off[r] the
// particles of the lower groups; the last group (ranks >= kLastRank) holds the
// rest in base order, each base's block at off[kLastRank] + ovf[l].  One
// wavefront per tile, a lane per base, from the per-base counts.
constexpr int kLastRank = (1 << kRankBits) - 1;

template <int DIM>
__global__ __launch_bounds__(64) void k_tile_layout(SortKeyGeom kg, const unsigned* __restrict__ count,
                                                    unsigned long long* __restrict__ masks,
                                                    int* __restrict__ offs, int* __restrict__ ovf,
                                                    int* __restrict__ totals, int* __restrict__ rounds,
                                                    int ntiles) {
    const int t = blockIdx.x, lane = threadIdx.x;
    if (t >= ntiles) return;
    const int tm = (1 << kg.tsh) - 1;
    const int cx = t % kg.tx, cy = (t / kg.tx) % kg.ty, cz = DIM == 3 ? t / (kg.tx * kg.ty) : 0;
    int lx, ly, lz;
    tile_local_xyz<DIM>(kg.tsh, lane, lx, ly, lz);
    const int b0 = (cx << kg.tsh) + lx, b1 = (cy << kg.tsh) + ly, b2 = DIM == 3 ? (cz << kg.tsh) + lz : 0;
    const bool valid = b0 < kg.nb[0] && b1 < kg.nb[1] && b2 < kg.nb[2];
    const int c = valid ? int(count[b0 + kg.nb[0] * (b1 + kg.nb[1] * b2)]) : 0;
    int acc = 0;
    for (int r = 0; r < kLastRank; ++r) {
        const unsigned long long m = __ballot(c > r);
        if (m == 0ull) break;
        if (lane == 0) {
            masks[(long long)t * kLastRank + r] = m;
            offs[(long long)t * (kLastRank + 1) + r] = acc;
        }
        acc += __popcll(m);
    }
    // the last group: exclusive prefix over bases of max(0, c - kLastRank)
    const int o = c > kLastRank ? c - kLastRank : 0;
    int incl = o;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d);
        if (lane >= d) incl += v;
    }
    ovf[(long long)t * 64 + lane] = incl - o;
    const int tail = __shfl(incl, 63);
    // rounds of the base-owned scatter: the largest particle count of a base
    int cmax = c;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) cmax = max(cmax, __shfl_xor(cmax, d));
    if (lane == 0) {
        offs[(long long)t * (kLastRank + 1) + kLastRank] = acc;
        totals[t] = acc + tail;
        rounds[t] = cmax;
    }
}

// Tile segments of the sorted order from the scanned tile sizes.
// Scan of the tile sizes (hand-written, three launches): for every tile t,
// starts[t] = the particles of the tiles below t (exclusive sum of totals) and
// slot_of[t] = the non-empty tiles below t (the tile's place in the compacted
// tile list and its tile-partials slot), -1 for an empty tile; *tile_count =
// the non-empty tiles.  Stage 1 sums each block of kScanBlock tiles, stage 2
// scans the block sums in one workgroup (chunk by chunk, with a carry), stage
// 3 scans each block again and adds its offset.  Integer sums: the result does
// not depend on the order.
constexpr int kScanBlock = 1024;

// Inclusive scan of (a, b) over a workgroup of kScanBlock threads; returns the
// workgroup's totals in ta, tb.  red holds 2 (kScanBlock / 64) ints.
__device__ __forceinline__ void block_scan2(int& a, int& b, int* red, int& ta, int& tb) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int NW = kScanBlock / 64;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int ua = __shfl_up(a, d), ub = __shfl_up(b, d);
        if (lane >= d) {
            a += ua;
            b += ub;
        }
    }
    if (lane == 63) {
        red[wave] = a;
        red[NW + wave] = b;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        int wa = threadIdx.x < NW ? red[threadIdx.x] : 0;
        int wb = threadIdx.x < NW ? red[NW + threadIdx.x] : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int ua = __shfl_up(wa, d), ub = __shfl_up(wb, d);
            if (threadIdx.x >= d) {
                wa += ua;
                wb += ub;
            }
        }
        if (threadIdx.x < NW) {
            red[threadIdx.x] = wa;  // inclusive over waves
            red[NW + threadIdx.x] = wb;
        }
    }
    __syncthreads();
    if (wave > 0) {
        a += red[wave - 1];
        b += red[NW + wave - 1];
    }
    ta = red[NW - 1];
    tb = red[2 * NW - 1];
    __syncthreads();  // red is reused by the caller's next scan
}

__global__ __launch_bounds__(kScanBlock) void k_tile_scan_sums(const int* __restrict__ totals, int ntiles,
                                                               int* __restrict__ bsum) {
    __shared__ int red[2 * (kScanBlock / 64)];
    const int t = blockIdx.x * kScanBlock + threadIdx.x;
    int a = t < ntiles ? totals[t] : 0;
    int b = a > 0 ? 1 : 0;
    int ta, tb;
    block_scan2(a, b, red, ta, tb);
    if (threadIdx.x == 0) {
        bsum[2 * blockIdx.x] = ta;
        bsum[2 * blockIdx.x + 1] = tb;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_tile_scan_blocks(int* __restrict__ bsum, int nblocks,
                                                                 int* __restrict__ tile_count) {
    __shared__ int red[2 * (kScanBlock / 64)];
    int ca = 0, cb = 0;  // carry: sums of the earlier chunks
    for (int c0 = 0; c0 < nblocks; c0 += kScanBlock) {
        const int k = c0 + int(threadIdx.x);
        const int xa = k < nblocks ? bsum[2 * k] : 0, xb = k < nblocks ? bsum[2 * k + 1] : 0;
        int a = xa, b = xb, ta, tb;
        block_scan2(a, b, red, ta, tb);
        if (k < nblocks) {
            bsum[2 * k] = ca + a - xa;  // exclusive
            bsum[2 * k + 1] = cb + b - xb;
        }
        ca += ta;
        cb += tb;
    }
    if (threadIdx.x == 0) *tile_count = cb;
}

__global__ __launch_bounds__(kScanBlock) void k_tile_scan_apply(const int* __restrict__ totals, int ntiles,
                                                                const int* __restrict__ bsum,
                                                                int* __restrict__ starts,
                                                                int* __restrict__ slot_of) {
    __shared__ int red[2 * (kScanBlock / 64)];
    const int t = blockIdx.x * kScanBlock + threadIdx.x;
    const int xa = t < ntiles ? totals[t] : 0, xb = xa > 0 ? 1 : 0;
    int a = xa, b = xb, ta, tb;
    block_scan2(a, b, red, ta, tb);
    if (t < ntiles) {
        starts[t] = bsum[2 * blockIdx.x] + a - xa;
        slot_of[t] = xb ? bsum[2 * blockIdx.x + 1] + b - xb : -1;
    }
}

// Tile segments of the sorted order, and the compacted list of non-empty
// tiles (increasing tile order; tile t at slot_of[t]).
__global__ __launch_bounds__(256) void k_tile_segments(const int* __restrict__ starts,
                                                       const int* __restrict__ totals,
                                                       const int* __restrict__ slot_of, int ntiles,
                                                       int* __restrict__ seg_begin,
                                                       int* __restrict__ seg_end,
                                                       int* __restrict__ tile_list) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    seg_begin[t] = starts[t];
    seg_end[t] = starts[t] + totals[t];
    if (slot_of[t] >= 0) tile_list[slot_of[t]] = t;
}

// Destination of every particle in the sorted order: order[pos] = i and the
// sorted key at pos.
__global__ __launch_bounds__(256) void k_sort_dest(const unsigned* __restrict__ keys,
                                                   const int* __restrict__ rank,
                                                   const int* __restrict__ starts,
                                                   const unsigned long long* __restrict__ masks,
                                                   const int* __restrict__ offs,
                                                   const int* __restrict__ ovf,
                                                   const int* __restrict__ totals, int ntiles,
                                                   unsigned* __restrict__ frozen,
                                                   int* __restrict__ order,
                                                   unsigned* __restrict__ keys_sorted, int np) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const unsigned key = keys[i];
    if (key == kFrozenKey) {
        // after the last tile, in arrival order (frozen particles take part in
        // no computation, so their order does not matter)
        const int pos = starts[ntiles - 1] + totals[ntiles - 1] + int(atomicAdd(frozen, 1u));
        order[pos] = i;
        keys_sorted[pos] = key;
        return;
    }
    const long long t = key >> kTileShift;
    const int l = int(key & ((1u << kLocalBits) - 1));
    const int r = rank[i];
    int pos = starts[t];
    if (r < kLastRank)
        pos += offs[t * (kLastRank + 1) + r] +
               __popcll(masks[t * kLastRank + r] & ((1ull << l) - 1ull));
    else
        pos += offs[t * (kLastRank + 1) + kLastRank] + ovf[t * 64 + l] + (r - kLastRank);
    order[pos] = i;
    keys_sorted[pos] = key;
}

// Every permuted particle stream in one launch: dst_k[i] = src_k[order[i]]
// (bit copies of 4- or 8-byte elements).  One index read per particle and
// all streams' loads in flight together, instead of one launch per stream.
struct PermuteStreams {
    static constexpr int kMax = 48;
    const void* src[kMax];
    void* dst[kMax];
    int wide[kMax];  // 1: 8-byte elements, 0: 4-byte
    int count;
};

__global__ __launch_bounds__(256) void k_permute_all(PermuteStreams a, const int* __restrict__ order,
                                                     int np) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int j = order[i];
#pragma unroll 4
    for (int k = 0; k < a.count; ++k) {
        if (a.wide[k])
            static_cast<unsigned long long*>(a.dst[k])[i] =
                static_cast<const unsigned long long*>(a.src[k])[j];
        else
            static_cast<unsigned*>(a.dst[k])[i] = static_cast<const unsigned*>(a.src[k])[j];
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(unsigned* __restrict__ p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template <typename R>
struct SortState {
    std::string error;
    bool enabled = false;
    int dim = 3, S = 3;
    long long np = 0;
    long long nbases = 0;
    SortKeyGeom kg{};
    GridGeom geom{};
    TileGeom tg{};
    // tile segments of the sorted order (valid after the first sort)
    int *seg_begin = nullptr, *seg_end = nullptr, *tile_list = nullptr, *tile_count = nullptr;
    int* scan_sums = nullptr;  // (particles, non-empty tiles) per scan block
    bool tiles_valid = false;
    unsigned *keys = nullptr, *keys_alt = nullptr, *count = nullptr;
    int *idx = nullptr, *idx_alt = nullptr;  // per-particle rank; the sorted order
    // counting-sort layout per tile (k_tile_layout)
    unsigned long long* masks = nullptr;
    int *offs = nullptr, *ovf = nullptr, *totals = nullptr, *starts = nullptr;
    int* rounds = nullptr;  // per tile: the largest particle count of one base
    int* slot_of = nullptr;  // per tile: its place in tile_list, -1 when empty
    // Alternate buffers, one per permuted stream in for_each_stream order;
    // they swap roles with the particle streams on every permutation, so the
    // caller allocates and frees them together with the particle streams.
    std::vector<void*> alt;
    int parity = 0;  // number of permutations, mod 2

    // Sorted keys of the current order (tile | rank | base), read by the tiled kernels.
    const unsigned* sorted_keys() const { return keys_alt; }

    int init(const mpm_config& cfg, const GridGeom& g, long long n) {
        if (cfg.sort_interval <= 0) return MPM_OK;
        enabled = true;
        dim = cfg.dim;
        S = cfg.basis == MPM_LINEAR ? 2 : 3;
        np = n;
        geom = g;
        kg.tsh = dim == 3 ? 2 : 3;
        const int te = 1 << kg.tsh;
        for (int a = 0; a < 3; ++a) kg.nb[a] = a < dim ? g.nn[a] - S + 1 : 1;
        kg.tx = (kg.nb[0] + te - 1) / te;
        kg.ty = (kg.nb[1] + te - 1) / te;
        const int tz = dim == 3 ? (kg.nb[2] + te - 1) / te : 1;
        nbases = (long long)kg.nb[0] * kg.nb[1] * kg.nb[2];
        unsigned long long maxkey = ((unsigned long long)kg.tx * kg.ty * tz) << kTileShift;
        if (maxkey >= (1ull << 32)) {
            error = "grid too large for 32-bit sort keys";
            return MPM_E_ARG;
        }
        tg.tsh = kg.tsh;
        tg.tx = kg.tx;
        tg.ty = kg.ty;
        tg.tz = tz;
        tg.ntiles = kg.tx * kg.ty * tz;
        hipError_t e = hipMalloc((void**)&keys, sizeof(unsigned) * np);
        // per-base counts, then the frozen-particle counter
        if (e == hipSuccess) e = hipMalloc((void**)&count, sizeof(unsigned) * (nbases + 1));
        if (e == hipSuccess) e = hipMalloc((void**)&seg_begin, sizeof(int) * tg.ntiles);
        if (e == hipSuccess) e = hipMalloc((void**)&seg_end, sizeof(int) * tg.ntiles);
        if (e == hipSuccess) e = hipMalloc((void**)&tile_list, sizeof(int) * tg.ntiles);
        if (e == hipSuccess) e = hipMalloc((void**)&tile_count, sizeof(int));
        if (e == hipSuccess)
            e = hipMalloc((void**)&scan_sums, sizeof(int) * 2 * ((tg.ntiles + kScanBlock - 1) / kScanBlock));
        if (e == hipSuccess) e = hipMalloc((void**)&keys_alt, sizeof(unsigned) * np);
        if (e == hipSuccess) e = hipMalloc((void**)&idx, sizeof(int) * np);
        if (e == hipSuccess) e = hipMalloc((void**)&idx_alt, sizeof(int) * np);
        const long long nt = tg.ntiles;
        if (e == hipSuccess) e = hipMalloc((void**)&masks, sizeof(unsigned long long) * nt * kLastRank);
        if (e == hipSuccess) e = hipMalloc((void**)&offs, sizeof(int) * nt * (kLastRank + 1));
        if (e == hipSuccess) e = hipMalloc((void**)&ovf, sizeof(int) * nt * 64);
        if (e == hipSuccess) e = hipMalloc((void**)&totals, sizeof(int) * nt);
        if (e == hipSuccess) e = hipMalloc((void**)&starts, sizeof(int) * nt);
        if (e == hipSuccess) e = hipMalloc((void**)&rounds, sizeof(int) * nt);
        if (e == hipSuccess) e = hipMalloc((void**)&slot_of, sizeof(int) * nt);
        if (e != hipSuccess) {
            error = std::string("hipMalloc: ") + hipGetErrorString(e);
            return MPM_E_NOMEM;
        }
        return MPM_OK;
    }

    void release() {
        void* ps[] = {keys, keys_alt, idx, idx_alt, scan_sums, seg_begin, seg_end, tile_list,
                      tile_count, count, masks, offs, ovf, totals, starts, rounds, slot_of};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        keys = keys_alt = count = nullptr;
        idx = idx_alt = nullptr;
        masks = nullptr;
        offs = ovf = totals = starts = rounds = slot_of = nullptr;
        scan_sums = nullptr;
        seg_begin = seg_end = tile_list = tile_count = nullptr;
    }


    // Every stream the sort permutes, in a fixed order: f(pointer, element size).
    template <typename F>
    static void for_each_stream(PartPtrs<R>& pp, int dim, F&& f) {
        const int ns = dim == 2 ? 4 : 6;
        for (int a = 0; a < dim; ++a) f(pp.x[a]);
        f(pp.J);
        for (int a = 0; a < dim; ++a) f(pp.v[a]);
        for (int k = 0; k < ns; ++k) f(pp.sig[k]);
        for (int k = 0; k < ns; ++k) f(pp.eps[k]);
        f(pp.ms);
        f(pp.vol);
        f(pp.vol0);
        f(pp.eq);
        f(pp.ek);
        f(pp.ev);
        f(pp.nid);
        f(pp.pid);
    }

    // Sort and permute every particle stream, then rebuild the tile segments.
    // `c` is the context (its launch wrapper times k_sort and k_permute).
    template <typename Ctx>
    int run(Ctx* c, PartPtrs<R>& pp, hipStream_t s) {
        if (!enabled) return MPM_OK;
        const unsigned nb = unsigned((np + 255) / 256);
        const int n = int(np);
        int rc = c->launch(6, [&] {
            k_fill_u32<<<unsigned((nbases + 1 + 255) / 256), 256, 0, s>>>(count, nbases + 1);
            const double* x2 = dim == 3 ? pp.x[2] : pp.x[1];
            if (dim == 3) {
                if (S == 3) k_sort_keys<3, 3><<<nb, 256, 0, s>>>(geom, kg, pp.x[0], pp.x[1], x2, count, keys, idx, n, c->d_flags);
                else k_sort_keys<3, 2><<<nb, 256, 0, s>>>(geom, kg, pp.x[0], pp.x[1], x2, count, keys, idx, n, c->d_flags);
                k_tile_layout<3><<<unsigned(tg.ntiles), 64, 0, s>>>(kg, count, masks, offs, ovf, totals, rounds, tg.ntiles);
            } else {
                if (S == 3) k_sort_keys<2, 3><<<nb, 256, 0, s>>>(geom, kg, pp.x[0], pp.x[1], x2, count, keys, idx, n, c->d_flags);
                else k_sort_keys<2, 2><<<nb, 256, 0, s>>>(geom, kg, pp.x[0], pp.x[1], x2, count, keys, idx, n, c->d_flags);
                k_tile_layout<2><<<unsigned(tg.ntiles), 64, 0, s>>>(kg, count, masks, offs, ovf, totals, rounds, tg.ntiles);
            }
            const int nsb = int((tg.ntiles + kScanBlock - 1) / kScanBlock);
            k_tile_scan_sums<<<unsigned(nsb), kScanBlock, 0, s>>>(totals, int(tg.ntiles), scan_sums);
            k_tile_scan_blocks<<<1, kScanBlock, 0, s>>>(scan_sums, nsb, tile_count);
            k_tile_scan_apply<<<unsigned(nsb), kScanBlock, 0, s>>>(totals, int(tg.ntiles), scan_sums, starts,
                                                                   slot_of);
            k_sort_dest<<<nb, 256, 0, s>>>(keys, idx, starts, masks, offs, ovf, totals, int(tg.ntiles), count + nbases,
                                          idx_alt, keys_alt, n);
        });
        if (rc != MPM_OK) return rc;
        const int* order = idx_alt;
        PermuteStreams ps{};
        for_each_stream(pp, dim, [&](auto*& ptr) {
            ps.src[ps.count] = ptr;
            ps.dst[ps.count] = alt[ps.count];
            ps.wide[ps.count] = sizeof(*ptr) == 8 ? 1 : 0;
            ++ps.count;
        });
        rc = c->launch(7, [&] { k_permute_all<<<nb, 256, 0, s>>>(ps, order, n); });
        if (rc != MPM_OK) return rc;
        // the permuted copies become the particle streams; the old ones the spares
        size_t k = 0;
        for_each_stream(pp, dim, [&](auto*& ptr) {
            using T = std::remove_reference_t<decltype(*ptr)>;
            void* old = (void*)ptr;
            ptr = static_cast<T*>(alt[k]);
            alt[k] = old;
            ++k;
        });
        parity ^= 1;
        // tile segments of the new order and the ordered list of non-empty tiles
        rc = c->launch(6, [&] {
            const unsigned tb = unsigned((tg.ntiles + 255) / 256);
            k_tile_segments<<<tb, 256, 0, s>>>(starts, totals, slot_of, int(tg.ntiles), seg_begin, seg_end,
                                               tile_list);
        });
        if (rc != MPM_OK) return rc;
        tiles_valid = true;
        return MPM_OK;
    }

    // Grid size of the tiled kernels (one wavefront per workgroup): a multiple
    // of 8 so that each XCD gets the same share.
    unsigned tiled_grid() const {
        long long g = ((long long)tg.ntiles + 7) / 8 * 8;
        if (g > 16384) g = 16384;
        if (g < 8) g = 8;
        return unsigned(g);
    }
};

}  // namespace mpm
