// mpm_tiled.hpp -- tile kernels for base-sorted particles (LDS windows).
//
// After every sort (mpm_sort.hpp) the particles of each tile of stencil
// bases (4x4x4 in 3D, 8x8 in 2D) are contiguous, ordered by (rank in the
// dual cell, base).  One wavefront owns one tile at a time and an LDS window
// of the grid nodes the tile's stencils reached at the sort (window edge =
// tile edge + S - 1 nodes; kWinSlack bases of slack per side would widen it
// for particles that drift):
//
//  * scatters (P2G #1, P2G #2) accumulate into the window and flush it with
//    one global atomic per touched node and field;
//  * gathers (the MUSL re-map, G2P #2) first stage the grid fields of the
//    window in LDS and then read every stencil from there.
//
// Scatter rounds.  A round is at most 64 particles of one rank of one tile,
// so at the last sort their bases were pairwise distinct; when a rank group
// leaves lanes free, the round continues with the next group, which adds into
// a second window (kScatterWindows), so half-empty groups share a wavefront.
// Each lane claims its current base in a small LDS table (write a unique tag,
// read it back) to confirm that the bases are still distinct; the winners'
// targets are then distinct at every stencil slot and they add into the
// window with plain LDS read-add-write (volatile, so the slots stay in program
// order).  Losers (particles that drifted onto a base another lane holds)
// claim again among themselves and add in a second pass.  LDS float atomics
// would serialise per lane on this hardware.  A particle that lost both claims
// or drifted out of the window scatters straight to the grid with global
// atomics (and a gather reads such a particle's stencil from the grid), so
// results never depend on the sort interval.
//
// Tile entries are dealt to the 8 XCDs in contiguous chunks (a workgroup's
// XCD is its blockIdx modulo 8), so neighbouring tiles, which share halo
// nodes, run behind the same L2.
#pragma once
#include <hip/hip_runtime.h>

#include "mpm_kernels.hpp"

namespace mpm {

constexpr int kRankBits = 5;   // ranks 0..31 (a fuller dual cell shares rank 31)
constexpr int kLocalBits = 6;  // 64 bases per tile
constexpr int kTileShift = kRankBits + kLocalBits;

// Local index of a stencil base inside its tile (bx, by, bz in [0, 2^tsh)):
// the low bits of the sort key, so the lanes of a round hold the group's bases
// in this order.  3D tiles of 4x4x4 bases: l = bz + 4 (bx & 1) + 8 (by & 1) +
// 16 (bx >> 1) + 32 (by >> 1), so every 16 consecutive lanes of a full round
// hold a 2x2x4 block of bases.  In a 6x6x6 window (node x + 6 y + 36 z) one
// stencil slot of such a block lands on x + 6 y + 4 z mod 16 = 16 distinct
// values (x, y in {0, 1}, z in 0..3, plus the slot's constant offset), i.e.
// on 16 distinct bank pairs of a 64-bit LDS access of 16 lanes; with x
// fastest, two of a 4x4 block's 16 slots share a bank pair.  Otherwise (2D,
// other tile sizes) x fastest.
template <int DIM>
__device__ __forceinline__ int tile_local(int tsh, int bx, int by, int bz) {
    if (DIM == 3 && tsh == 2)
        return bz + 4 * (bx & 1) + 8 * (by & 1) + 16 * (bx >> 1) + 32 * (by >> 1);
    return bx + ((by + (DIM == 3 ? bz << tsh : 0)) << tsh);
}

template <int DIM>
__device__ __forceinline__ void tile_local_xyz(int tsh, int l, int& bx, int& by, int& bz) {
    if (DIM == 3 && tsh == 2) {
        bz = l & 3;
        bx = ((l >> 2) & 1) | (((l >> 4) & 1) << 1);
        by = ((l >> 3) & 1) | (((l >> 5) & 1) << 1);
        return;
    }
    const int tm = (1 << tsh) - 1;
    bx = l & tm;
    by = (l >> tsh) & tm;
    bz = DIM == 3 ? l >> (2 * tsh) : 0;
}
// Bases of slack on each side of a window.  0 measured fastest on the 3D
// column: with one base of slack the 7-field scatter window grows from 6^3 to
// 8^3 nodes, fewer wavefronts fit per CU, and P2G #1 and the re-map gather
// ran 6-7 % slower, while the few particles that drift out of a slack-free
// window between sorts cost little on the fallback path.
constexpr int kWinSlack = 0;
// Scatter windows per wavefront: a round is one rank group in window 0 plus,
// in the lanes it leaves free, the next group in window 1 (the two are summed
// before the flush).  Measured on the 3D column: P2G #1 161 -> 150 us early in
// the collapse (52.8 particles per group) and 386 -> 343 us at step 5000 (about
// 30 per group, as the deposit spreads and compacts unevenly).
#ifndef MPM_SCATTER_WINDOWS
#define MPM_SCATTER_WINDOWS 2
#endif
constexpr int kScatterWindows = MPM_SCATTER_WINDOWS;
// Wavefronts per SIMD the rank-round scatter is compiled for (the layered form: 3).
#ifndef MPM_SCATTER_WAVES
#define MPM_SCATTER_WAVES 4
#endif

struct TileGeom {
    int tsh;         // log2 of the tile edge in stencil bases
    int tx, ty, tz;  // tiles per axis
    int ntiles;
};

// Window geometry of a tile: NB bases and WW nodes per axis, x fastest, z
// stride ZS.  (Row and plane strides padded so that a round's 4x4x4 block of
// bases hits distinct LDS banks measured slower on the 3D column: the larger
// window costs resident wavefronts, and bank conflicts are not what bounds the
// scatter.  ZS stays WW * WW.)
template <int DIM, int S, int TSH>
struct Win {
    static constexpr int NB = (1 << TSH) + 2 * kWinSlack;
    static constexpr int WW = NB + S - 1;
    static constexpr int ZS = WW * WW;
    static constexpr int WN = DIM == 3 ? ZS * WW : WW * WW;
    static constexpr int NBN = DIM == 3 ? NB * NB * NB : NB * NB;
};

__device__ __forceinline__ void xcd_chunk(int count, int& e0, int& e1, int& step) {
    const int G = gridDim.x;  // multiple of 8
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3, nloc = G >> 3;
    const int C = (count + 7) >> 3;
    e0 = xcd * C + local;
    e1 = min(count, (xcd + 1) * C);
    step = nloc;
}

// First window node of a tile (kWinSlack bases below the tile).
template <int DIM>
__device__ __forceinline__ void tile_window_origin(const TileGeom& tg, int tile, int* w0) {
    const int cx = tile % tg.tx;
    const int cy = (tile / tg.tx) % tg.ty;
    const int cz = DIM == 3 ? tile / (tg.tx * tg.ty) : 0;
    w0[0] = (cx << tg.tsh) - kWinSlack;
    w0[1] = (cy << tg.tsh) - kWinSlack;
    w0[2] = DIM == 3 ? (cz << tg.tsh) - kWinSlack : 0;
}

// LDS-typed pointers: volatile accesses through a generic pointer would be
// emitted as FLAT instructions, so the windows are addressed in address space 3.
template <typename T>
using lds_ptr = volatile __attribute__((address_space(3))) T*;

// Window offsets of a stencil base (node offset wb, base offset bidx); false
// if the base is outside the window.
template <int DIM, int NB, int WW, int ZS>
__device__ __forceinline__ bool window_slot(const int* base, const int* w0, int& wb, int& bidx) {
    const int rx = base[0] - w0[0], ry = base[1] - w0[1], rz = DIM == 3 ? base[2] - w0[2] : 0;
    const bool inside = rx >= 0 && rx < NB && ry >= 0 && ry < NB && rz >= 0 && rz < NB;
    wb = DIM == 3 ? rz * ZS + ry * WW + rx : ry * WW + rx;
    bidx = DIM == 3 ? (rz * NB + ry) * NB + rx : ry * NB + rx;
    return inside;
}

// Window slot k -> node offsets (kx, ky, kz) inside the window; false for a
// padding slot.
template <int DIM, int WW, int ZS>
__device__ __forceinline__ bool window_node(int k, int& kx, int& ky, int& kz) {
    kz = DIM == 3 ? k / ZS : 0;
    const int r = k - kz * ZS;
    ky = r / WW;
    kx = r - ky * WW;
    return ky < WW;
}

// Stage NF grid fields of a tile's window into LDS (zero outside the grid).
template <int DIM, int WW, int ZS, int WN, int NF, typename R>
__device__ __forceinline__ void stage_window(const GridGeom& g, const int* w0, R* const* fields,
                                             R* lds) {
    for (int k = threadIdx.x; k < WN; k += blockDim.x) {
        int kx, ky, kz;
        if (!window_node<DIM, WW, ZS>(k, kx, ky, kz)) continue;
        const int gx = w0[0] + kx, gy = w0[1] + ky, gz = DIM == 3 ? w0[2] + kz : 0;
        const bool in = gx >= 0 && gx < g.nn[0] && gy >= 0 && gy < g.nn[1] &&
                        (DIM == 2 || (gz >= 0 && gz < g.nn[2]));
        const long long n = in ? node_id(g, gx, gy, gz) : 0;
#pragma unroll
        for (int f = 0; f < NF; ++f) lds[f * WN + k] = in ? fields[f][n] : R(0);
    }
}

// Staged windows of the tile gathers (G2P #2, the re-map gather): node
// (x, y, z) of the window at x + SY y + SZ z.  With the dense strides (WW,
// WW^2) the 4x4x4 tile's bases in lane order (tile_local) put two lanes of
// every 32 on one LDS bank at each stencil slot; FP32 3D uGIMP pads the rows
// to 16 and the layers to 100 nodes, which puts the 32 lanes of each half and
// the 64 lanes of a full rank group on distinct banks at all 27 slots (586
// nodes per field instead of 216).
template <int DIM, int S, int TSH, typename R>
struct GWin {
    using W = Win<DIM, S, TSH>;
    static constexpr bool PAD = DIM == 3 && S == 3 && TSH == 2 && sizeof(R) == 4;
    static constexpr int SY = PAD ? 16 : W::WW;
    static constexpr int SZ = PAD ? 100 : W::ZS;
    static constexpr int N = PAD ? (W::WW - 1) * (1 + SY + SZ) + 1 : W::WN;  // nodes per field
};

// stage_window into a GWin layout (only the window's nodes are written).
template <int DIM, int WW, int ZS, int WN, int SY, int SZ, int GN, int NF, typename R>
__device__ __forceinline__ void stage_window_strided(const GridGeom& g, const int* w0,
                                                     R* const* fields, R* lds) {
    for (int k = threadIdx.x; k < WN; k += blockDim.x) {
        int kx, ky, kz;
        if (!window_node<DIM, WW, ZS>(k, kx, ky, kz)) continue;
        const int gx = w0[0] + kx, gy = w0[1] + ky, gz = DIM == 3 ? w0[2] + kz : 0;
        const bool in = gx >= 0 && gx < g.nn[0] && gy >= 0 && gy < g.nn[1] &&
                        (DIM == 2 || (gz >= 0 && gz < g.nn[2]));
        const long long n = in ? node_id(g, gx, gy, gz) : 0;
        const int p = kx + SY * ky + SZ * kz;
#pragma unroll
        for (int f = 0; f < NF; ++f) lds[f * GN + p] = in ? fields[f][n] : R(0);
    }
}

// FP32 3D P2G #1 windows interleave field pairs: (m, px), (py, pz), (fx, fy) as
// 64-bit pairs per node and fz alone, so a stencil slot reads and writes three
// pairs and one value (4 + 4 LDS instructions instead of 7 + 7).  Field f of
// window node k sits at f < 6 ? (f / 2) 2 WN + 2 k + f % 2 : 6 WN + k; the
// window size is unchanged.
// The momentum scatter of the re-map (3 fields) pairs (px, py) the same way,
// pz alone.
// 64-bit LDS accesses need 8-byte alignment: every pair must start at an even
// float, so the paired layout is used only when the field stride WN is even
// (uGIMP windows: 6^3 and 10^2 nodes; the linear basis's 5^3 and 9^2 keep the
// one-field-per-access layout).
template <int DIM, int NF, typename R, int WN>
constexpr bool kPairedP2G = sizeof(R) == 4 && WN % 2 == 0 &&
                            ((DIM == 3 && (NF == 7 || NF == 3)) || (DIM == 2 && (NF == 5 || NF == 2)));

template <int NF>
__device__ __forceinline__ int paired_offset(int f, int k, int wn) {
    constexpr int NP = NF / 2;  // pairs; the last field is alone
    return f < 2 * NP ? (f >> 1) * 2 * wn + 2 * k + (f & 1) : 2 * NP * wn + k;
}

// Flush a scatter window.  With a tile-partials slot tb (the tile's place in
// the tile list), the window is stored there, node-major (store_partial), and
// the grid solve adds the partials of the up to 2^D windows that hold a node:
// plain stores and a fixed summation order.  Without (tb == nullptr): global
// atomic adds of every non-zero node (only window nodes inside the grid can be
// non-zero).
template <int DIM, int WW, int ZS, int WN, int NF, typename R>
__device__ __forceinline__ void flush_window(const GridGeom& g, const int* w0, lds_ptr<R> win,
                                             R* const* fields, R* tb = nullptr) {
    if (tb) {
        for (int k = threadIdx.x; k < WN; k += blockDim.x) {
            R vals[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f)
                vals[f] = win[kPairedP2G<DIM, NF, R, WN> ? paired_offset<NF>(f, k, WN) : f * WN + k];
            store_partial<NF, R>(tb + (long long)k * part_stride<NF>(), vals);
        }
        return;
    }
    for (int k = threadIdx.x; k < WN; k += blockDim.x) {
        int kx, ky, kz;
        if (!window_node<DIM, WW, ZS>(k, kx, ky, kz)) continue;
        R vals[NF];
        bool any = false;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            vals[f] = win[kPairedP2G<DIM, NF, R, WN> ? paired_offset<NF>(f, k, WN) : f * WN + k];
            any = any || vals[f] != R(0);
        }
        if (!any) continue;
        const long long n = node_id(g, w0[0] + kx, w0[1] + ky, DIM == 3 ? w0[2] + kz : 0);
#pragma unroll
        for (int f = 0; f < NF; ++f) atomic_add(&fields[f][n], vals[f]);
    }
}


// P2G #1 contribution of one particle to one node: m, p[DIM], f[DIM]
// (utils_OS.jl:466-494).
template <int DIM, typename R>
__device__ __forceinline__ void p2g_values(R N, R dNx, R dNy, R dNz, R m, R vol, const R* pv,
                                           const R* sg, const Gravity<R>& G, R* c) {
    const R NiM = m * N;
    c[0] = NiM;
    c[1] = N * pv[0];
    c[2] = N * pv[1];
    if constexpr (DIM == 2) {
        c[3] = -vol * (dNx * sg[0] + dNy * sg[3]) + NiM * G.g[0];
        c[4] = -vol * (dNy * sg[1] + dNx * sg[3]) + NiM * G.g[1];
    } else {
        c[3] = N * pv[2];
        c[4] = -vol * (dNx * sg[0] + dNy * sg[3] + dNz * sg[5]) + NiM * G.g[0];
        c[5] = -vol * (dNy * sg[1] + dNx * sg[3] + dNz * sg[4]) + NiM * G.g[1];
        c[6] = -vol * (dNz * sg[2] + dNx * sg[5] + dNy * sg[4]) + NiM * G.g[2];
    }
}

// P2G #1 into a window, 3D, with the force factored per (y, z) node pair:
// with A = Ny Nz, B = dNy Nz, C = Ny dNz the gradient is (dNx A, Nx B, Nx C),
// so f_k = dNx P_k + Nx Q_k with P = -Ω σ_{k x} A and
// Q_k = -Ω (σ_{k y} B + σ_{k z} C) + m A g_k (the same terms as p2g_values,
// utils_OS.jl:480-494, grouped to take the products out of the x loop).
// Slots stay in program order with each slot's 7 reads before its 7 writes.
template <int S, typename R, typename Win>
__device__ __forceinline__ void p2g3_window(const Stencil<3, S, R>& st, R m, R vol, const R* pv,
                                            const R* sg, const Gravity<R>& G, Win L0, int ws,
                                            int zs, int wn) {
    const R mv = -vol;
    const R sxx = mv * sg[0], sxy = mv * sg[3], sxz = mv * sg[5];
    const R syy = mv * sg[1], syz = mv * sg[4], szz = mv * sg[2];
    static_for<S>([&](auto c) {
        static_for<S>([&](auto b) {
            const R A = st.N[1][b] * st.N[2][c];
            const R B = st.dN[1][b] * st.N[2][c];
            const R C = st.N[1][b] * st.dN[2][c];
            const R mA = m * A;
            const R Px = sxx * A, Py = sxy * A, Pz = sxz * A;
            const R Qx = sxy * B + sxz * C + mA * G.g[0];
            const R Qy = syy * B + syz * C + mA * G.g[1];
            const R Qz = syz * B + szz * C + mA * G.g[2];
            static_for<S>([&](auto a) {
                const R Nx = st.N[0][a], dNx = st.dN[0][a];
                const R N = Nx * A;
                if (N != R(0)) {
                    R cv[7];
                    cv[0] = m * N;
                    cv[1] = N * pv[0];
                    cv[2] = N * pv[1];
                    cv[3] = N * pv[2];
                    cv[4] = dNx * Px + Nx * Qx;
                    cv[5] = dNx * Py + Nx * Qy;
                    cv[6] = dNx * Pz + Nx * Qz;
                    auto L = L0 + (int(c) * zs + int(b) * ws + int(a));
                    R old[7];
#pragma unroll
                    for (int f = 0; f < 7; ++f) old[f] = L[f * wn];
#pragma unroll
                    for (int f = 0; f < 7; ++f) L[f * wn] = old[f] + cv[f];
                }
                __builtin_amdgcn_sched_barrier(0);
            });
        });
    });
}

// p2g3_window on the paired FP32 layout (kPairedP2G): W is the window's first
// float, node0 the stencil base's node index in the window.
template <int S>
__device__ __forceinline__ void p2g3_window_paired(const Stencil<3, S, float>& st, float m,
                                                   float vol, const float* pv, const float* sg,
                                                   const Gravity<float>& G, lds_ptr<float> W,
                                                   int node0, int ws, int zs, int wn) {
    using u64p = volatile __attribute__((address_space(3))) unsigned long long*;
    auto lo = [](unsigned long long v) { return __uint_as_float(unsigned(v)); };
    auto hi = [](unsigned long long v) { return __uint_as_float(unsigned(v >> 32)); };
    auto pack = [](float a, float b) {
        return (unsigned long long)__float_as_uint(a) |
               ((unsigned long long)__float_as_uint(b) << 32);
    };
    const float mv = -vol;
    const float sxx = mv * sg[0], sxy = mv * sg[3], sxz = mv * sg[5];
    const float syy = mv * sg[1], syz = mv * sg[4], szz = mv * sg[2];
    static_for<S>([&](auto c) {
        static_for<S>([&](auto b) {
            const float A = st.N[1][b] * st.N[2][c];
            const float B = st.dN[1][b] * st.N[2][c];
            const float C = st.N[1][b] * st.dN[2][c];
            const float mA = m * A;
            const float Px = sxx * A, Py = sxy * A, Pz = sxz * A;
            const float Qx = sxy * B + sxz * C + mA * G.g[0];
            const float Qy = syy * B + syz * C + mA * G.g[1];
            const float Qz = syz * B + szz * C + mA * G.g[2];
            static_for<S>([&](auto a) {
                const float Nx = st.N[0][a], dNx = st.dN[0][a];
                const float N = Nx * A;
                if (N != 0.0f) {
                    const int k = node0 + int(c) * zs + int(b) * ws + int(a);
                    u64p P0 = reinterpret_cast<u64p>(W + 2 * k);
                    u64p P1 = reinterpret_cast<u64p>(W + 2 * wn + 2 * k);
                    u64p P2 = reinterpret_cast<u64p>(W + 4 * wn + 2 * k);
                    lds_ptr<float> L6 = W + 6 * wn + k;
                    const unsigned long long o0 = *P0, o1 = *P1, o2 = *P2;
                    const float o6 = *L6;
                    *P0 = pack(lo(o0) + m * N, hi(o0) + N * pv[0]);
                    *P1 = pack(lo(o1) + N * pv[1], hi(o1) + N * pv[2]);
                    *P2 = pack(lo(o2) + (dNx * Px + Nx * Qx), hi(o2) + (dNx * Py + Nx * Qy));
                    *L6 = o6 + (dNx * Pz + Nx * Qz);
                }
                __builtin_amdgcn_sched_barrier(0);
            });
        });
    });
}

// Layered scatter windows (FP32 3D uGIMP, the paired layout).  Three windows
// per wavefront, one per stencil z offset c: layer window c holds what the
// slots (a, b, c) add, at node (x, y, z - c), so it spans the tile's NB base
// layers in z (WW x WW x NB nodes) and node0 (the base's offset in a full
// window) is also the base's offset in every layer window.  A round's lanes
// hold distinct bases, so slots of one layer can collide across lanes only
// with slots of the same layer: the three layers are independent chains, and
// each (a, b) step issues the reads of all three, then the three updates --
// one LDS round trip per step instead of one per slot (9 per particle instead
// of 27).  The layers are summed when the window is flushed.
template <int DIM, int S, int TSH>
struct WinLayers {
    using W = Win<DIM, S, TSH>;
    static constexpr int WNL = W::ZS * W::NB;  // nodes per layer window
};

// P2G #1 into the layer windows (7 paired fields; layer c at W + c 7 wnl).
template <int S>
__device__ __forceinline__ void p2g3_window_layered(const Stencil<3, S, float>& st, float m,
                                                    float vol, const float* pv, const float* sg,
                                                    const Gravity<float>& G, lds_ptr<float> W,
                                                    int node0, int ws, int wnl) {
    using u64p = volatile __attribute__((address_space(3))) unsigned long long*;
    auto lo = [](unsigned long long v) { return __uint_as_float(unsigned(v)); };
    auto hi = [](unsigned long long v) { return __uint_as_float(unsigned(v >> 32)); };
    auto pack = [](float a, float b) {
        return (unsigned long long)__float_as_uint(a) |
               ((unsigned long long)__float_as_uint(b) << 32);
    };
    const float mv = -vol;
    const float sxx = mv * sg[0], sxy = mv * sg[3], sxz = mv * sg[5];
    const float syy = mv * sg[1], syz = mv * sg[4], szz = mv * sg[2];
    static_for<S>([&](auto b) {
        // per c: A = Ny Nz and the force factors P, Q of p2g3_window
        float A[S], Px[S], Py[S], Pz[S], Qx[S], Qy[S], Qz[S];
        static_for<S>([&](auto c) {
            const float a_ = st.N[1][b] * st.N[2][c];
            const float B = st.dN[1][b] * st.N[2][c];
            const float C = st.N[1][b] * st.dN[2][c];
            const float mA = m * a_;
            A[c] = a_;
            Px[c] = sxx * a_;
            Py[c] = sxy * a_;
            Pz[c] = sxz * a_;
            Qx[c] = sxy * B + sxz * C + mA * G.g[0];
            Qy[c] = syy * B + syz * C + mA * G.g[1];
            Qz[c] = syz * B + szz * C + mA * G.g[2];
        });
        static_for<S>([&](auto a) {
            const float Nx = st.N[0][a], dNx = st.dN[0][a];
            if (Nx * st.N[1][b] != 0.0f) {
                const int k = node0 + int(b) * ws + int(a);
                unsigned long long o0[S], o1[S], o2[S];
                float o6[S];
                static_for<S>([&](auto c) {
                    lds_ptr<float> L = W + int(c) * 7 * wnl;
                    o0[c] = *reinterpret_cast<u64p>(L + 2 * k);
                    o1[c] = *reinterpret_cast<u64p>(L + 2 * wnl + 2 * k);
                    o2[c] = *reinterpret_cast<u64p>(L + 4 * wnl + 2 * k);
                    o6[c] = L[6 * wnl + k];
                });
                static_for<S>([&](auto c) {
                    lds_ptr<float> L = W + int(c) * 7 * wnl;
                    const float N = Nx * A[c];
                    *reinterpret_cast<u64p>(L + 2 * k) = pack(lo(o0[c]) + m * N, hi(o0[c]) + N * pv[0]);
                    *reinterpret_cast<u64p>(L + 2 * wnl + 2 * k) =
                        pack(lo(o1[c]) + N * pv[1], hi(o1[c]) + N * pv[2]);
                    *reinterpret_cast<u64p>(L + 4 * wnl + 2 * k) =
                        pack(lo(o2[c]) + (dNx * Px[c] + Nx * Qx[c]),
                             hi(o2[c]) + (dNx * Py[c] + Nx * Qy[c]));
                    L[6 * wnl + k] = o6[c] + (dNx * Pz[c] + Nx * Qz[c]);
                });
            }
            __builtin_amdgcn_sched_barrier(0);
        });
    });
}

// The momentum scatter of the MUSL re-map into the layer windows (3 fields:
// (px, py) paired, pz alone).
template <int S>
__device__ __forceinline__ void p2g2_window_layered(const Stencil<3, S, float>& st, const float* pv,
                                                    lds_ptr<float> W, int node0, int ws, int wnl) {
    using u64p = volatile __attribute__((address_space(3))) unsigned long long*;
    static_for<S>([&](auto b) {
        static_for<S>([&](auto a) {
            const float NxNy = st.N[0][a] * st.N[1][b];
            if (NxNy != 0.0f) {
                const int k = node0 + int(b) * ws + int(a);
                unsigned long long o0[S];
                float o2[S];
                static_for<S>([&](auto c) {
                    lds_ptr<float> L = W + int(c) * 3 * wnl;
                    o0[c] = *reinterpret_cast<u64p>(L + 2 * k);
                    o2[c] = L[2 * wnl + k];
                });
                static_for<S>([&](auto c) {
                    lds_ptr<float> L = W + int(c) * 3 * wnl;
                    const float N = NxNy * st.N[2][c];
                    const float x = __uint_as_float(unsigned(o0[c])) + N * pv[0];
                    const float y = __uint_as_float(unsigned(o0[c] >> 32)) + N * pv[1];
                    *reinterpret_cast<u64p>(L + 2 * k) =
                        (unsigned long long)__float_as_uint(x) | ((unsigned long long)__float_as_uint(y) << 32);
                    L[2 * wnl + k] = o2[c] + N * pv[2];
                });
            }
            __builtin_amdgcn_sched_barrier(0);
        });
    });
}

// Flush of the layer windows: window node (x, y, z) is the sum of layer c's
// node (x, y, z - c) over the layers that hold it (c ascending), then as
// flush_window (stored as the tile's partials, or added with global atomics).
template <int DIM, int S, int TSH, int NF, typename R>
__device__ __forceinline__ void flush_layers(const GridGeom& g, const int* w0, lds_ptr<R> win,
                                             R* const* fields, R* tb) {
    using W = Win<DIM, S, TSH>;
    constexpr int WW = W::WW, ZS = W::ZS, WN = W::WN, NB = W::NB;
    constexpr int WNL = WinLayers<DIM, S, TSH>::WNL;
    for (int k = threadIdx.x; k < WN; k += blockDim.x) {
        int kx, ky, kz;
        window_node<DIM, WW, ZS>(k, kx, ky, kz);
        R vals[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) vals[f] = R(0);
#pragma unroll
        for (int c = 0; c < S; ++c) {
            const int zl = kz - c;
            if (zl < 0 || zl >= NB) continue;
            const int kl = zl * ZS + ky * WW + kx;
#pragma unroll
            for (int f = 0; f < NF; ++f) vals[f] += win[c * NF * WNL + paired_offset<NF>(f, kl, WNL)];
        }
        if (tb) {
            store_partial<NF, R>(tb + (long long)k * part_stride<NF>(), vals);
            continue;
        }
        bool any = false;
#pragma unroll
        for (int f = 0; f < NF; ++f) any = any || vals[f] != R(0);
        if (!any) continue;
        const long long n = node_id(g, w0[0] + kx, w0[1] + ky, w0[2] + kz);
#pragma unroll
        for (int f = 0; f < NF; ++f) atomic_add(&fields[f][n], vals[f]);
    }
}

// ---------------------------------------------------------------------------
// Tile scatters.  MODE 0: P2G #1 (P2G_OS!, utils_OS.jl:424-496): mass,
// momentum and internal + external force.  MODE 1: the momentum scatter of the
// MUSL re-map (doublemapping2_OS!, utils_OS.jl:942-959) from the new particle
// momentum that k_g2p1_tiled stored in pp.pm, with the step-start basis.
// MODE 2 / 3: the whole MUSL re-map (doublemapping1(_a)_OS! +
// doublemapping2_OS!, utils_OS.jl:698-959) in the rounds: the tile's vsT, as
// (MODE 3: and vs, see pic_field) are staged in LDS, each particle gathers
// from there (from the grid outside the window), moves (v and the deferred
// position increment), and scatters its new momentum in the same round with
// the same step-start basis; the position is read once and the momentum never
// goes through memory.
// ---------------------------------------------------------------------------
template <int DIM, int S, int TSH, typename R, int MODE, bool LAY = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(LAY ? 3 : MPM_SCATTER_WAVES))) void k_scatter_tiled(
    GridGeom g, TileGeom tg, PartPtrs<R> pp, GridPtrs<R> gp, Gravity<R> G,
    const unsigned* __restrict__ keys, const int* __restrict__ tile_list,
    const int* __restrict__ seg_begin, const int* __restrict__ seg_end,
    const int* __restrict__ count_ptr, int* flags, R* __restrict__ tb, StepArgs<R> sa = {}) {
    using W = Win<DIM, S, TSH>;
    constexpr int WW = W::WW, WN = W::WN, NB = W::NB, ZS = W::ZS;
    constexpr int NF = MODE == 0 ? 1 + 2 * DIM : DIM;
    constexpr bool REMAP = MODE >= 2;           // gather + move + scatter
    constexpr bool VS = MODE == 3;              // the re-map gathers vs as well
    constexpr int NG = (VS ? 3 : 2) * DIM;      // gathered fields
    // LAY: three layer windows (WinLayers) and one rank group per round
    static_assert(!LAY || (DIM == 3 && S == 3 && (MODE == 0 || MODE == 1) &&
                           kPairedP2G<DIM, NF, R, WN>),
                  "layer windows: FP32 3D uGIMP scatters");
    constexpr int NWIN = LAY ? 1 : kScatterWindows;
    static_assert(NWIN == 1 || NWIN == 2, "one or two scatter windows");
    constexpr int WNL = WinLayers<DIM, S, TSH>::WNL;
    constexpr int WIN_FLOATS = LAY ? S * NF * WNL : NWIN * NF * WN;
    __shared__ __attribute__((aligned(16))) R win_[WIN_FLOATS];
    __shared__ unsigned claim_[NWIN * W::NBN];
    __shared__ R gwin[REMAP ? NG * WN : 1];
    R* const* vsf = pic_field(gp, sa);
    R* gfields[3 * DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        gfields[d] = gp.vsT[d];
        gfields[DIM + d] = gp.as[d];
        gfields[2 * DIM + d] = vsf[d];
    }
    lds_ptr<R> win = (lds_ptr<R>)win_;
    lds_ptr<unsigned> claim = (lds_ptr<unsigned>)claim_;
    R* fields[NF];
    if constexpr (MODE == 0) {
        fields[0] = gp.m;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            fields[1 + d] = gp.p[d];
            fields[1 + DIM + d] = gp.f[d];
        }
    } else {
#pragma unroll
        for (int d = 0; d < DIM; ++d) fields[d] = gp.p[d];
    }
    unsigned tag = threadIdx.x;  // unique per lane and round
    int flag = 0;
    int e0, e1, est;
    xcd_chunk(*count_ptr, e0, e1, est);
    for (int e = e0; e < e1; e += est) {
        const int tile = tile_list[e];
        const int beg = seg_begin[tile], end = seg_end[tile];
        int w0[3];
        tile_window_origin<DIM>(tg, tile, w0);
        for (int k = threadIdx.x; k < WIN_FLOATS; k += 64) win[k] = R(0);
        if constexpr (REMAP) {
            __syncthreads();  // the previous tile's gathers from gwin are done
            stage_window<DIM, WW, ZS, WN, NG, R>(g, w0, gfields, gwin);
            __syncthreads();
        }
        // Rounds are software-pipelined: the keys of the next round come from a
        // two-register sliding window over the sorted keys (a round has at most
        // 64 particles, so the next round lies in [cur, cur + 128)), and the next
        // round's particle data is loaded before this round's LDS scatter.
        const int lane = threadIdx.x;
        // A round: the keys up to the end of the first rank group (window 0)
        // and, with two windows, on to the end of the second group (window 1).
        // Returns the round's particle count; side is the lane's window.
        auto member = [&](unsigned ka, int c0, bool& act, int& side) -> int {
            const unsigned grp = unsigned(__shfl(int(ka), 0)) >> kLocalBits;
            const bool in = c0 + lane < end;
            const bool a0 = in && (ka >> kLocalBits) == grp;
            const int n0 = __popcll(__ballot(a0));
            side = 0;
            act = a0;
            if (NWIN == 1 || n0 >= 64) return n0;
            const unsigned grp2 = unsigned(__shfl(int(ka), n0)) >> kLocalBits;
            const bool a1 = in && lane >= n0 && (ka >> kLocalBits) == grp2;
            if (a1) side = 1;
            act = a0 || a1;
            return n0 + __popcll(__ballot(a1));
        };
        struct Pre {
            double x[3];
            R m, vol, v[3], sg[6];
        };
        auto load_pre = [&](int i, bool act, Pre& P) {
            if (!act) return;
            load_x<DIM>(pp.x, i, P.x);
            if constexpr (MODE == 0) {
                P.m = pp.ms[i];
                P.vol = pp.vol[i];
#pragma unroll
                for (int a = 0; a < DIM; ++a) P.v[a] = pp.v[a][i];
                P.sg[0] = pp.sig[0][i];
                P.sg[1] = pp.sig[1][i];
                P.sg[3] = pp.sig[3][i];
                if constexpr (DIM == 3) {
                    P.sg[2] = pp.sig[2][i];
                    P.sg[4] = pp.sig[4][i];
                    P.sg[5] = pp.sig[5][i];
                }
            } else if constexpr (REMAP) {
                P.m = pp.ms[i];
#pragma unroll
                for (int a = 0; a < DIM; ++a) P.v[a] = pp.v[a][i];
            } else {
#pragma unroll
                for (int a = 0; a < DIM; ++a) P.v[a] = pp.pm[a][i];
            }
        };
        int cur = beg;
        unsigned kA = cur + lane < end ? keys[cur + lane] : 0u;
        unsigned kB = cur + 64 + lane < end ? keys[cur + 64 + lane] : 0u;
        bool act = false;
        int side = 0;
        int cnt = cur < end ? member(kA, cur, act, side) : 0;
        Pre P{};
        load_pre(cur + lane, act, P);
        while (cur < end) {
            const int src = (cnt + lane) & 63;
            const unsigned a2 = unsigned(__shfl(int(kA), src)), b2 = unsigned(__shfl(int(kB), src));
            const unsigned kAn = (cnt + lane < 64) ? a2 : b2;
            const int curn = cur + cnt;
            const unsigned kBn = curn + 64 + lane < end ? keys[curn + 64 + lane] : 0u;
            bool actn = false;
            int siden = 0;
            const int cntn = curn < end ? member(kAn, curn, actn, siden) : 0;
            Pre Pn{};
            load_pre(curn + lane, actn, Pn);
            if (act) {
                const int i = cur + lane;
                Stencil<DIM, S, R> st;
                int fl2 = 0;
                const bool ok = make_stencil<DIM, S, R>(g, P.x, st, MODE == 0 ? flag : fl2);
                if (REMAP && !ok) {
#pragma unroll
                    for (int a = 0; a < DIM; ++a) pp.dxp[a][i] = R(0);  // frozen particle
                }
                if (ok) {
                    R pv[3] = {0, 0, 0};
                    const R m = P.m, vol = P.vol;
                    if constexpr (REMAP) {
                        // the re-map gather with the step-start basis, then the move
                        int gb, gidx;
                        const bool gin = window_slot<DIM, NB, WW, ZS>(st.base, w0, gb, gidx);
                        R A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, C[3] = {0, 0, 0};
                        if (gin) {
                            // branch-free: zero weights add exact zeros (staged values are finite)
                            for_stencil_local<DIM, S, R>(st, WW, ZS, [&](int off, R N, R, R, R) {
                                const R* L = gwin + gb + off;
                                R vT[3], vA[3], vS[3];
#pragma unroll
                                for (int d = 0; d < DIM; ++d) {
                                    vT[d] = L[d * WN];
                                    vA[d] = L[(DIM + d) * WN];
                                    vS[d] = VS ? L[(2 * DIM + d) * WN] : R(0);
                                }
                                g2p1_accumulate<DIM, R>(sa.va, N, vT, vA, vS, A, B, C);
                            });
                        } else {
                            for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R, R, R) {
                                if (N != R(0)) {
                                    R vT[3], vA[3], vS[3];
#pragma unroll
                                    for (int d = 0; d < DIM; ++d) {
                                        vT[d] = gp.vsT[d][n];
                                        vA[d] = gp.as[d][n];
                                        vS[d] = VS ? vsf[d][n] : R(0);
                                    }
                                    g2p1_accumulate<DIM, R>(sa.va, N, vT, vA, vS, A, B, C);
                                }
                            });
                        }
                        g2p1_commit<DIM, R>(pp, sa, i, A, B, C, pv, m, P.v);
                    } else {
#pragma unroll
                        for (int a = 0; a < DIM; ++a) pv[a] = MODE == 0 ? m * P.v[a] : P.v[a];
                    }
                    auto contrib = [&](R N, R dNx, R dNy, R dNz, R* c) {
                        if constexpr (MODE == 0) {
                            p2g_values<DIM, R>(N, dNx, dNy, dNz, m, vol, pv, P.sg, G, c);
                        } else {
#pragma unroll
                            for (int d = 0; d < DIM; ++d) c[d] = pv[d] * N;
                        }
                    };
                    int wb, bidx;
                    const bool inside = window_slot<DIM, NB, WW, ZS>(st.base, w0, wb, bidx);
                    bidx += side * W::NBN;
                    wb += side * NF * WN;
                    if (inside) claim[bidx] = tag;
                    const bool won = inside && claim[bidx] == tag;
                    tag += 64u;
                    // A particle that drifted onto a base another lane of the round
                    // holds claims again among the losers and adds in a second pass.
                    bool won2 = false;
                    if (inside && !won) {
                        claim[bidx] = tag;
                        won2 = claim[bidx] == tag;
                    }
                    tag += 64u;
                    auto lds_pass = [&] {
                        if constexpr (LAY && MODE == 0) {
                            p2g3_window_layered<S>(st, m, vol, pv, P.sg, G, win, wb, WW, WNL);
                        } else if constexpr (LAY) {
                            p2g2_window_layered<S>(st, pv, win, wb, WW, WNL);
                        } else if constexpr (MODE == 0 && DIM == 3) {
                            if constexpr (kPairedP2G<DIM, NF, R, WN>)
                                p2g3_window_paired<S>(st, m, vol, pv, P.sg, G, win + side * NF * WN,
                                                      wb - side * NF * WN, WW, ZS, WN);
                            else
                                p2g3_window<S, R>(st, m, vol, pv, P.sg, G, win + wb, WW, ZS, WN);
                        } else if constexpr (kPairedP2G<DIM, NF, R, WN>) {
                            // field pairs as 64-bit read-add-writes, an odd last field alone
                            using u64p = volatile __attribute__((address_space(3))) unsigned long long*;
                            constexpr int NP = NF / 2;
                            lds_ptr<R> Wb = win + side * NF * WN;
                            const int node0 = wb - side * NF * WN;
                            for_stencil_local<DIM, S, R>(st, WW, ZS, [&](int off, R N, R dNx, R dNy, R dNz) {
                                if (N != R(0)) {
                                    R c[NF];
                                    contrib(N, dNx, dNy, dNz, c);
                                    const int k = node0 + off;
                                    unsigned long long o[NP];
#pragma unroll
                                    for (int q = 0; q < NP; ++q)
                                        o[q] = *reinterpret_cast<u64p>(Wb + 2 * q * WN + 2 * k);
                                    R o1 = R(0);
                                    if constexpr (NF % 2) o1 = Wb[2 * NP * WN + k];
#pragma unroll
                                    for (int q = 0; q < NP; ++q) {
                                        const float a = __uint_as_float(unsigned(o[q])) + c[2 * q];
                                        const float b = __uint_as_float(unsigned(o[q] >> 32)) + c[2 * q + 1];
                                        *reinterpret_cast<u64p>(Wb + 2 * q * WN + 2 * k) =
                                            (unsigned long long)__float_as_uint(a) |
                                            ((unsigned long long)__float_as_uint(b) << 32);
                                    }
                                    if constexpr (NF % 2) Wb[2 * NP * WN + k] = o1 + c[NF - 1];
                                }
                                __builtin_amdgcn_sched_barrier(0);
                            });
                        } else {
                            for_stencil_local<DIM, S, R>(st, WW, ZS, [&](int off, R N, R dNx, R dNy, R dNz) {
                                if (N != R(0)) {
                                    R c[NF];
                                    contrib(N, dNx, dNy, dNz, c);
                                    lds_ptr<R> L = win + wb + off;
                                    R old[NF];
#pragma unroll
                                    for (int f = 0; f < NF; ++f) old[f] = L[f * WN];
#pragma unroll
                                    for (int f = 0; f < NF; ++f) L[f * WN] = old[f] + c[f];
                                }
                                // keep each slot's arithmetic next to its LDS update
                                __builtin_amdgcn_sched_barrier(0);
                            });
                        }
                    };
                    // the two passes in order: the ballot is wave-wide, so the passes
                    // cannot be merged into one (their lanes may share bases)
                    if (won) lds_pass();
                    if (__ballot(won2) != 0ull) {
                        if (won2) lds_pass();
                    }
                    if (!won && !won2) {
                        // drifted out of the window, or lost both claims
                        for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R dNx, R dNy, R dNz) {
                            if (N != R(0)) {
                                R c[NF];
                                contrib(N, dNx, dNy, dNz, c);
#pragma unroll
                                for (int f = 0; f < NF; ++f) atomic_add(&fields[f][n], c[f]);
                            }
                            __builtin_amdgcn_sched_barrier(0);
                        });
                    }
                    (void)i;
                }
            }
            P = Pn;
            act = actn;
            side = siden;
            cnt = cntn;
            cur = curn;
            kA = kAn;
            kB = kBn;
        }
        if constexpr (LAY) {
            flush_layers<DIM, S, TSH, NF, R>(g, w0, win, fields,
                                             tb ? tb + (long long)e * part_stride<NF>() * WN : nullptr);
            continue;
        }
        if constexpr (NWIN == 2) {
            for (int k = threadIdx.x; k < NF * WN; k += 64) win[k] = win[k] + win[NF * WN + k];
        }
        flush_window<DIM, WW, ZS, WN, NF, R>(g, w0, win, fields,
                                             tb ? tb + (long long)e * part_stride<NF>() * WN : nullptr);
    }
    raise_flag(flags, flag);
}

// ---------------------------------------------------------------------------
// G2P #2 with the Δu window staged in LDS (G2P_OS!, utils_OS.jl:1025-1153),
// then the strain / constitutive / MVL / position commit of g2p2_commit.
// Four wavefronts share the window of a tile.
// ---------------------------------------------------------------------------
// Wavefronts per SIMD the FP32 gathers are compiled for.  Left to itself the
// compiler gives G2P #2 (DP) 189 registers (2 wavefronts per SIMD, its commit
// loads all in flight) and the re-map gather 130 (3); held to 3 wavefronts
// (168 registers, two spilled in G2P #2) they measured 120.8 -> 105.1 us and
// 76.9 -> 66.7 us at step 40 of the 3D column, 174.8 -> 150.6 and 115.0 -> 97.2
// at step 5,040; at 4 G2P #2 spills 50 registers and takes twice as long.  The
// FP64 forms keep the compiler's choice (3 wavefronts would spill 50-140).
constexpr int kGatherWaves = 3;
template <typename R>
constexpr int gather_waves() { return sizeof(R) == 4 ? kGatherWaves : 1; }

template <int DIM, int S, int TSH, typename R, int CM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(gather_waves<R>()))) void k_g2p2_tiled(
    GridGeom g, TileGeom tg, PartPtrs<R> pp, GridPtrs<R> gp, const Material<R>* __restrict__ mats,
    StepArgs<R> sa, const int* __restrict__ tile_list, const int* __restrict__ seg_begin,
    const int* __restrict__ seg_end, const int* __restrict__ count_ptr, int* flags) {
    using W = Win<DIM, S, TSH>;
    using GW = GWin<DIM, S, TSH, R>;
    constexpr int WW = W::WW, WN = W::WN, NB = W::NB, ZS = W::ZS;
    constexpr int SY = GW::SY, SZ = GW::SZ, GN = GW::N;
    constexpr int NF = DIM * DIM;
    __shared__ R uwin[DIM * GN];
    R* ufields[3] = {gp.du[0], gp.du[1], DIM == 3 ? gp.du[2] : gp.du[1]};
    int flag = 0;
    int e0, e1, est;
    xcd_chunk(*count_ptr, e0, e1, est);
    for (int e = e0; e < e1; e += est) {
        const int tile = tile_list[e];
        const int beg = seg_begin[tile], end = seg_end[tile];
        int w0[3];
        tile_window_origin<DIM>(tg, tile, w0);
        __syncthreads();
        stage_window_strided<DIM, WW, ZS, WN, SY, SZ, GN, DIM, R>(g, w0, ufields, uwin);
        __syncthreads();
        for (int i = beg + int(threadIdx.x); i < end; i += blockDim.x) {
            double x[3];
            load_x<DIM>(pp.x, i, x);
            CommitIn<DIM, R> cin;
            if constexpr (kCommitLoadsEarly<R, CM>) load_commit_in<DIM, R, CM>(pp, sa, i, cin);
            Stencil<DIM, S, R> st;
            if (!make_stencil<DIM, S, R>(g, x, st, flag)) {
                if (sa.mvl) pp.p2c[i] = -1;  // frozen particle: no cell
                continue;
            }
            int wb, bidx;
            const bool inside = window_slot<DIM, NB, SY, SZ>(st.base, w0, wb, bidx);
            R F[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k) F[k] = R(0);
            auto acc = [&](const R* u, R dNx, R dNy, R dNz) {
                if constexpr (DIM == 2) {
                    F[0] += u[0] * dNx;
                    F[1] += u[0] * dNy;
                    F[2] += u[1] * dNx;
                    F[3] += u[1] * dNy;
                } else {
                    F[0] += u[0] * dNx; F[1] += u[0] * dNy; F[2] += u[0] * dNz;
                    F[3] += u[1] * dNx; F[4] += u[1] * dNy; F[5] += u[1] * dNz;
                    F[6] += u[2] * dNx; F[7] += u[2] * dNy; F[8] += u[2] * dNz;
                }
            };
            if (inside) {
                // branch-free: a zero weight comes with zero gradients for uGIMP;
                // the linear basis keeps the test (its gradient is not zero there)
                for_stencil_local<DIM, S, R>(st, SY, SZ, [&](int off, R N, R dNx, R dNy, R dNz) {
                    if (S == 3 || N != R(0)) {
                        R u[3];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) u[d] = uwin[d * GN + wb + off];
                        acc(u, dNx, dNy, dNz);
                    }
                });
            } else {
                for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R dNx, R dNy, R dNz) {
                    if (N != R(0)) {
                        R u[3];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) u[d] = gp.du[d][n];
                        acc(u, dNx, dNy, dNz);
                    }
                });
            }
            if constexpr (!kCommitLoadsEarly<R, CM>) load_commit_in<DIM, R, CM>(pp, sa, i, cin);
            g2p2_commit<DIM, R, CM>(g, pp, gp, mats, sa, i, x, F, cin);
        }
    }
    raise_flag(flags, flag);
}

// ---------------------------------------------------------------------------
// MUSL re-map, gather half (doublemapping1(_a)_OS!, utils_OS.jl:698-908): four
// wavefronts share one staged window of vsT, as, vs per tile and take the
// tile's particles in order (no rounds: gathers do not conflict); the new
// momentum m v goes to pp.pm for the momentum scatter k_scatter_tiled<..., 1>.
// ---------------------------------------------------------------------------
// VS: whether vs is gathered.  The :a form with PIC = 0 multiplies its sum by
// zero, so that window drops a third of its fields (the default FLIP = 1).
template <int DIM, int S, int TSH, typename R, bool VS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(gather_waves<R>()))) void k_g2p1_tiled(GridGeom g, TileGeom tg, PartPtrs<R> pp,
                                                   GridPtrs<R> gp, StepArgs<R> sa,
                                                   const int* __restrict__ tile_list,
                                                   const int* __restrict__ seg_begin,
                                                   const int* __restrict__ seg_end,
                                                   const int* __restrict__ count_ptr) {
    using W = Win<DIM, S, TSH>;
    using GW = GWin<DIM, S, TSH, R>;
    constexpr int WW = W::WW, WN = W::WN, NB = W::NB, ZS = W::ZS;
    constexpr int SY = GW::SY, SZ = GW::SZ, GN = GW::N;
    constexpr int NG = (VS ? 3 : 2) * DIM;  // gathered: vsT, as[, vs (or vsT again, see pic_field)]
    __shared__ R gwin[NG * GN];
    R* const* vsf = pic_field(gp, sa);
    R* gfields[3 * DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        gfields[d] = gp.vsT[d];
        gfields[DIM + d] = gp.as[d];
        gfields[2 * DIM + d] = vsf[d];
    }
    int e0, e1, est;
    xcd_chunk(*count_ptr, e0, e1, est);
    for (int e = e0; e < e1; e += est) {
        const int tile = tile_list[e];
        const int beg = seg_begin[tile], end = seg_end[tile];
        int w0[3];
        tile_window_origin<DIM>(tg, tile, w0);
        __syncthreads();  // the previous tile's reads of gwin are done
        stage_window_strided<DIM, WW, ZS, WN, SY, SZ, GN, NG, R>(g, w0, gfields, gwin);
        __syncthreads();
        for (int i = beg + int(threadIdx.x); i < end; i += blockDim.x) {
            double x[3];
            load_x<DIM>(pp.x, i, x);
            const R m0 = pp.ms[i];
            R v0[3];
#pragma unroll
            for (int d = 0; d < DIM; ++d) v0[d] = pp.v[d][i];
            Stencil<DIM, S, R> st;
            int flag = 0;  // reported by the P2G #1 kernel
            if (!make_stencil<DIM, S, R>(g, x, st, flag)) {
#pragma unroll
                for (int a = 0; a < DIM; ++a) {
                    pp.dxp[a][i] = R(0);
                    pp.pm[a][i] = R(0);
                }
                continue;
            }
            int wb, bidx;
            const bool inside = window_slot<DIM, NB, SY, SZ>(st.base, w0, wb, bidx);
            R A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, C[3] = {0, 0, 0};
            if (inside) {
                // branch-free: zero weights add exact zeros (staged values are finite)
                for_stencil_local<DIM, S, R>(st, SY, SZ, [&](int off, R N, R, R, R) {
                    const R* L = gwin + wb + off;
                    R vT[3], vA[3], vS[3];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        vT[d] = L[d * GN];
                        vA[d] = L[(DIM + d) * GN];
                        vS[d] = VS ? L[(2 * DIM + d) * GN] : R(0);
                    }
                    g2p1_accumulate<DIM, R>(sa.va, N, vT, vA, vS, A, B, C);
                });
            } else {
                for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R, R, R) {
                    if (N != R(0)) {
                        R vT[3], vA[3], vS[3];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) {
                            vT[d] = gp.vsT[d][n];
                            vA[d] = gp.as[d][n];
                            vS[d] = VS ? vsf[d][n] : R(0);
                        }
                        g2p1_accumulate<DIM, R>(sa.va, N, vT, vA, vS, A, B, C);
                    }
                });
            }
            R pv[3] = {0, 0, 0};
            g2p1_commit<DIM, R>(pp, sa, i, A, B, C, pv, m0, v0);
#pragma unroll
            for (int a = 0; a < DIM; ++a) pp.pm[a][i] = pv[a];
        }
    }
}

}  // namespace mpm
