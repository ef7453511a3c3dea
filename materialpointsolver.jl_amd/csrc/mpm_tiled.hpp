// mpm_tiled.hpp -- LDS-binned scatter kernels for cell-sorted particles.
//
// After every cell sort the particles of each tile (4x4x4 cells in 3D, 8x8 in
// 2D) are contiguous.  A wavefront takes one tile at a time, accumulates the
// scatter of its particles into an LDS window covering every node their
// stencils can reach (tile edge + 3 nodes per axis), and then adds each touched
// window node to the grid with one global atomic per field, instead of one per
// particle-node-field.  Particles that drifted out of their tile's window
// since the last sort scatter straight to the grid, so the result does not
// depend on the sort interval.
//
// Inside a tile the particles are ordered by (quadrant of the cell, cell), so
// 64 consecutive particles (one coalesced wavefront round) usually have 64
// distinct stencil bases.
//
// Tile entries are dealt to the 8 XCDs in contiguous chunks (a workgroup's
// XCD is its blockIdx modulo 8), so neighbouring tiles, which share halo nodes
// and gather the same grid values, run behind the same L2.
#pragma once
#include <hip/hip_runtime.h>

#include "mpm_kernels.hpp"

namespace mpm {

struct TileGeom {
    int tsh;         // log2 of the tile edge in cells
    int tx, ty, tz;  // tiles per axis
    int ntiles;
};

__device__ __forceinline__ void xcd_chunk(int count, int& e0, int& e1, int& step) {
    const int G = gridDim.x;  // multiple of 8
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3, nloc = G >> 3;
    const int C = (count + 7) >> 3;
    e0 = xcd * C + local;
    e1 = min(count, (xcd + 1) * C);
    step = nloc;
}

template <int DIM>
__device__ __forceinline__ void tile_window_origin(const TileGeom& tg, int tile, int* w0) {
    const int cx = tile % tg.tx;
    const int cy = (tile / tg.tx) % tg.ty;
    const int cz = DIM == 3 ? tile / (tg.tx * tg.ty) : 0;
    w0[0] = (cx << tg.tsh) - 1;
    w0[1] = (cy << tg.tsh) - 1;
    w0[2] = (cz << tg.tsh) - 1;
}

// Distinct-base rounds.  A wavefront holds one tile's LDS window.  In each
// round its 64 lanes take 64 consecutive particles of the tile; a lane
// "claims" its stencil base in a small LDS table and keeps the claim if no
// other lane overwrote it.  The winners have pairwise distinct bases, so at
// every stencil slot their target window nodes are distinct, and they add
// into the window with plain LDS read-add-write (volatile, so the slots stay
// in program order).  A lane that lost the claim (another particle of the
// round shares its base, which the (quadrant, cell) order makes rare) scatters
// straight to the grid.  This replaces LDS float atomics, which serialise per
// lane on this hardware.
template <int DIM, int S, int WW>
struct WindowShape {
    static constexpr int WN = DIM == 3 ? WW * WW * WW : WW * WW;  // window nodes
    static constexpr int NB = WW - S + 1;                         // bases per axis
    static constexpr int NBN = DIM == 3 ? NB * NB * NB : NB * NB;
};

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// LDS-typed pointers: volatile accesses through a generic pointer would be
// emitted as FLAT instructions, so the window is addressed in address space 3.
template <typename T>
using lds_ptr = volatile __attribute__((address_space(3))) T*;

// P2G #1 with LDS tile bins: m, p[DIM], f[DIM] (NF = 1 + 2*DIM fields).
template <int DIM, int S, int TSH, typename R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_p2g_tiled(GridGeom g, TileGeom tg, PartPtrs<R> pp,
                                                  GridPtrs<R> gp, R G,
                                                  const int* __restrict__ tile_list,
                                                  const int* __restrict__ seg_begin,
                                                  const int* __restrict__ seg_end,
                                                  const int* __restrict__ count_ptr, int* flags) {
    constexpr int WW = (1 << TSH) + 3;
    using WS = WindowShape<DIM, S, WW>;
    constexpr int WN = WS::WN, NB = WS::NB;
    constexpr int NF = 1 + 2 * DIM;
    __shared__ R win_[NF * WN];
    __shared__ unsigned claim_[WS::NBN];
    lds_ptr<R> win = (lds_ptr<R>)win_;
    lds_ptr<unsigned> claim = (lds_ptr<unsigned>)claim_;
    const int lane = threadIdx.x;
    unsigned tag = unsigned(lane);
    int flag = 0;
    int e0, e1, est;
    xcd_chunk(*count_ptr, e0, e1, est);
    for (int e = e0; e < e1; e += est) {
        const int tile = tile_list[e];
        const int beg = seg_begin[tile], P = seg_end[tile] - beg;
        int w0[3];
        tile_window_origin<DIM>(tg, tile, w0);
        for (int k = lane; k < NF * WN; k += 64) win[k] = R(0);
        for (int r0 = 0; r0 < P; r0 += 64) {
            const int q = r0 + lane;
            if (q >= P) continue;
            const int i = beg + q;
            double x[3];
            load_x<DIM>(pp.x, i, x);
            Stencil<DIM, S, R> st;
            if (!make_stencil<DIM, S, R>(g, x, st, flag)) continue;
            const R m = pp.ms[i], vol = pp.vol[i];
            R pv[3] = {0, 0, 0}, sg[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int a = 0; a < DIM; ++a) pv[a] = m * pp.v[a][i];
            sg[0] = pp.sig[0][i];
            sg[1] = pp.sig[1][i];
            sg[3] = pp.sig[3][i];
            if constexpr (DIM == 3) {
                sg[2] = pp.sig[2][i];
                sg[4] = pp.sig[4][i];
                sg[5] = pp.sig[5][i];
            }
            const int rx = st.base[0] - w0[0], ry = st.base[1] - w0[1];
            const int rz = DIM == 3 ? st.base[2] - w0[2] : 0;
            const bool inside = rx >= 0 && rx < NB && ry >= 0 && ry < NB && rz >= 0 && rz < NB;
            const int bidx = DIM == 3 ? (rz * NB + ry) * NB + rx : ry * NB + rx;
            // distinct-base claim (see above); tags are unique per lane and round
            if (inside) claim[bidx] = tag;
            const bool won = inside && claim[bidx] == tag;
            tag += 64u;
            auto contrib = [&](R N, R dNx, R dNy, R dNz, R* c) {
                const R NiM = m * N;
                c[0] = NiM;
                c[1] = N * pv[0];
                c[2] = N * pv[1];
                if constexpr (DIM == 2) {
                    c[3] = -vol * (dNx * sg[0] + dNy * sg[3]);
                    c[4] = -vol * (dNy * sg[1] + dNx * sg[3]) + NiM * G;
                } else {
                    c[3] = N * pv[2];
                    c[4] = -vol * (dNx * sg[0] + dNy * sg[3] + dNz * sg[5]);
                    c[5] = -vol * (dNy * sg[1] + dNx * sg[3] + dNz * sg[4]);
                    c[6] = -vol * (dNz * sg[2] + dNx * sg[5] + dNy * sg[4]) + NiM * G;
                }
            };
            if (won) {
                const int wb = DIM == 3 ? (rz * WW + ry) * WW + rx : ry * WW + rx;
                for_stencil_local<DIM, S, R>(st, WW, [&](int off, R N, R dNx, R dNy, R dNz) {
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
            } else {
                // drifted out of the tile's window since the last sort, or lost
                // the claim to a particle with the same stencil base
                for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R dNx, R dNy, R dNz) {
                    if (N != R(0)) {
                        R c[NF];
                        contrib(N, dNx, dNy, dNz, c);
                        atomic_add(&gp.m[n], c[0]);
#pragma unroll
                        for (int d = 0; d < DIM; ++d) atomic_add(&gp.p[d][n], c[1 + d]);
#pragma unroll
                        for (int d = 0; d < DIM; ++d) atomic_add(&gp.f[d][n], c[1 + DIM + d]);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                });
            }
        }
        // flush the window: one global atomic per touched node and field
        for (int k = lane; k < WN; k += 64) {
            R vals[NF];
            bool any = false;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                vals[f] = win[f * WN + k];
                any = any || vals[f] != R(0);
            }
            if (!any) continue;
            const int kx = k % WW, ky = (k / WW) % WW, kz = DIM == 3 ? k / (WW * WW) : 0;
            const long long n = node_id(g, w0[0] + kx, w0[1] + ky, DIM == 3 ? w0[2] + kz : 0);
            atomic_add(&gp.m[n], vals[0]);
#pragma unroll
            for (int d = 0; d < DIM; ++d) atomic_add(&gp.p[d][n], vals[1 + d]);
#pragma unroll
            for (int d = 0; d < DIM; ++d) atomic_add(&gp.f[d][n], vals[1 + DIM + d]);
        }
    }
    raise_flag(flags, flag);
}

// MUSL re-map with LDS tile bins for the momentum scatter (P2G #2).
template <int DIM, int S, int TSH, typename R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_g2p1_p2g2_tiled(GridGeom g, TileGeom tg, PartPtrs<R> pp,
                                                        GridPtrs<R> gp, StepArgs<R> sa,
                                                        const int* __restrict__ tile_list,
                                                        const int* __restrict__ seg_begin,
                                                        const int* __restrict__ seg_end,
                                                        const int* __restrict__ count_ptr) {
    constexpr int WW = (1 << TSH) + 3;
    using WS = WindowShape<DIM, S, WW>;
    constexpr int WN = WS::WN, NB = WS::NB;
    __shared__ R win_[DIM * WN];
    __shared__ unsigned claim_[WS::NBN];
    lds_ptr<R> win = (lds_ptr<R>)win_;
    lds_ptr<unsigned> claim = (lds_ptr<unsigned>)claim_;
    const int lane = threadIdx.x;
    unsigned tag = unsigned(lane);
    int e0, e1, est;
    xcd_chunk(*count_ptr, e0, e1, est);
    for (int e = e0; e < e1; e += est) {
        const int tile = tile_list[e];
        const int beg = seg_begin[tile], P = seg_end[tile] - beg;
        int w0[3];
        tile_window_origin<DIM>(tg, tile, w0);
        for (int k = lane; k < DIM * WN; k += 64) win[k] = R(0);
        for (int r0 = 0; r0 < P; r0 += 64) {
            const int q = r0 + lane;
            if (q >= P) continue;
            const int i = beg + q;
            double x[3];
            load_x<DIM>(pp.x, i, x);
            Stencil<DIM, S, R> st;
            int flag = 0;
            if (!make_stencil<DIM, S, R>(g, x, st, flag)) {
#pragma unroll
                for (int a = 0; a < DIM; ++a) pp.dxp[a][i] = R(0);
                continue;
            }
            R A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, C[3] = {0, 0, 0};
            for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R, R, R) {
                if (N != R(0)) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        R vT = gp.vsT[d][n];
                        if (sa.va == 0) {
                            A[d] += N * vT;
                            B[d] += N * gp.as[d][n];
                            C[d] += N * gp.vs[d][n];
                        } else {
                            A[d] += N * vT;
                            B[d] += N * (vT - gp.vs[d][n]);
                            C[d] += N * vT;
                        }
                    }
                }
            });
            const R m = pp.ms[i];
            R pv[3] = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                R v = pp.v[d][i];
                R vn = (sa.va == 0) ? sa.flip * v + sa.pic * C[d] + B[d] * sa.dt
                                    : sa.flip * (v + B[d]) + sa.pic * C[d];
                pp.v[d][i] = vn;
                pp.dxp[d][i] = A[d] * sa.dt;
                pv[d] = m * vn;
            }
            const int rx = st.base[0] - w0[0], ry = st.base[1] - w0[1];
            const int rz = DIM == 3 ? st.base[2] - w0[2] : 0;
            const bool inside = rx >= 0 && rx < NB && ry >= 0 && ry < NB && rz >= 0 && rz < NB;
            const int bidx = DIM == 3 ? (rz * NB + ry) * NB + rx : ry * NB + rx;
            if (inside) claim[bidx] = tag;
            const bool won = inside && claim[bidx] == tag;
            tag += 64u;
            if (won) {
                const int wb = DIM == 3 ? (rz * WW + ry) * WW + rx : ry * WW + rx;
                for_stencil_local<DIM, S, R>(st, WW, [&](int off, R N, R, R, R) {
                    if (N != R(0)) {
                        lds_ptr<R> L = win + wb + off;
                        R old[3];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) old[d] = L[d * WN];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) L[d * WN] = old[d] + pv[d] * N;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                });
            } else {
                for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R, R, R) {
                    if (N != R(0)) {
#pragma unroll
                        for (int d = 0; d < DIM; ++d) atomic_add(&gp.p[d][n], pv[d] * N);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                });
            }
        }
        for (int k = lane; k < WN; k += 64) {
            R vals[3];
            bool any = false;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                vals[d] = win[d * WN + k];
                any = any || vals[d] != R(0);
            }
            if (!any) continue;
            const int kx = k % WW, ky = (k / WW) % WW, kz = DIM == 3 ? k / (WW * WW) : 0;
            const long long n = node_id(g, w0[0] + kx, w0[1] + ky, DIM == 3 ? w0[2] + kz : 0);
#pragma unroll
            for (int d = 0; d < DIM; ++d) atomic_add(&gp.p[d][n], vals[d]);
        }
    }
}

// ---------------------------------------------------------------------------
// Tile segments of the sorted particle order
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tile_segments(const unsigned* __restrict__ keys, int np,
                                                       int shift, int* __restrict__ seg_begin,
                                                       int* __restrict__ seg_end,
                                                       unsigned char* __restrict__ nonempty) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const unsigned t = keys[i] >> shift;
    if (i == 0 || (keys[i - 1] >> shift) != t) {
        seg_begin[t] = i;
        nonempty[t] = 1;
    }
    if (i == np - 1 || (keys[i + 1] >> shift) != t) seg_end[t] = i + 1;
}

__global__ __launch_bounds__(256) void k_fill_u8(unsigned char* __restrict__ p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Ordered compaction of the non-empty tiles (one workgroup of 1024 threads).
__global__ __launch_bounds__(1024) void k_compact_tiles(const unsigned char* __restrict__ nonempty,
                                                        int ntiles, int* __restrict__ list,
                                                        int* __restrict__ count) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (ntiles + 1023) / 1024;
    const int b = t * per, e = min(ntiles, b + per);
    int c = 0;
    for (int k = b; k < e; ++k) c += nonempty[k];
    part[t] = c;
    __syncthreads();
    // inclusive Hillis-Steele scan
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int pos = part[t] - c;
    for (int k = b; k < e; ++k)
        if (nonempty[k]) list[pos++] = k;
    if (t == 1023) *count = part[1023];
}

}  // namespace mpm
