// mpm_owned.hpp -- base-owned tile scatters: each lane owns one stencil base.
//
// After a sort (mpm_sort.hpp) the particles of a tile of 4x4x4 stencil bases
// (8x8 in 2D) are laid out in rank groups: group r holds the r-th particle of
// every base that has more than r particles, in base order, and the groups of
// ranks >= 31 are merged into one block per base.  One wavefront takes one
// tile, and lane L owns base L of the tile: in round r it loads the r-th
// particle of its base (a coalesced load, the group is contiguous) and adds
// that particle's contributions to the S^D nodes of its stencil in REGISTERS.
// Only after the last round does the lane add its sums into an LDS window of
// the tile's nodes, one stencil slot at a time: at a fixed slot the 64 bases
// of a tile reach 64 distinct nodes, so the wave's read-add-write of one slot
// never collides, and the window holds each node's sum over the whole tile.
// The window is then flushed with one global atomic per touched node and
// field, as in the rank-round scatter (mpm_tiled.hpp).
//
// Against the rank-round scatter this moves the per-particle work from LDS
// read-add-write (27 dependent slots of 7 fields per particle in 3D) to
// register FMAs; LDS sees one read-add-write pass per base and tile.
//
// A particle that left its base since the sort cannot use its lane's
// registers (their node offsets are those of the lane's base).  It goes to a
// drift list in LDS; batches of 64 drifted particles claim their bases in a
// small LDS table (write a unique tag, read it back), the winners add into
// the window with the same lockstep read-add-write, and the rest claim again;
// a drifted particle whose stencil left the window scatters to the grid with
// global atomics.  Results never depend on the sort interval.
//
// The accumulators must fit the register file: 3D P2G #1 sums mass and
// momentum in one pass over the tile's particles and the force in a second
// (FP32), and FP64 3D also takes the z layers of the stencil one at a time
// (OwnedPasses).
//
// The file also holds the pair rounds (k_scatter_pairs, FP32 3D uGIMP P2G #1
// by default): between the two, a lane holds two particles of one base and
// adds both per stencil slot with one window read-add-write.
#pragma once
#include <hip/hip_runtime.h>

#include "mpm_kernels.hpp"
#include "mpm_tiled.hpp"

namespace mpm {

// Window read-add-write of one node: c[f] added to field f of window node k
// (paired FP32 layout or one field per access, as kPairedP2G selects).
template <int NF, typename R, int WN, bool PAIRED>
__device__ __forceinline__ void window_add_node(lds_ptr<R> W, int k, const R* c) {
    if constexpr (PAIRED) {
        using u64p = volatile __attribute__((address_space(3))) unsigned long long*;
        constexpr int NP = NF / 2;
        unsigned long long o[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) o[q] = *reinterpret_cast<u64p>(W + 2 * q * WN + 2 * k);
        R o1 = R(0);
        if constexpr (NF % 2) o1 = W[2 * NP * WN + k];
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const float a = __uint_as_float(unsigned(o[q])) + c[2 * q];
            const float b = __uint_as_float(unsigned(o[q] >> 32)) + c[2 * q + 1];
            *reinterpret_cast<u64p>(W + 2 * q * WN + 2 * k) =
                (unsigned long long)__float_as_uint(a) | ((unsigned long long)__float_as_uint(b) << 32);
        }
        if constexpr (NF % 2) W[2 * NP * WN + k] = o1 + c[NF - 1];
    } else {
        R o[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) o[f] = W[f * WN + k];
#pragma unroll
        for (int f = 0; f < NF; ++f) W[f * WN + k] = o[f] + c[f];
    }
}

// Hides the values of a register array from the optimiser (an empty asm
// that may change them), so computations on them stay where they are written.
template <typename T, int N>
__device__ __forceinline__ void opaque_regs(T (&a)[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) asm volatile("" : "+v"(a[k]));
}
template <typename T, int N, int M>
__device__ __forceinline__ void opaque_regs(T (&a)[N][M]) {
#pragma unroll
    for (int k = 0; k < N; ++k) opaque_regs(a[k]);
}

// Particle data a scatter reads.
template <typename R>
struct ScatterIn {
    double x[3];
    R m, vol, v[3], sg[6];
};

// MOM / FRC (MODE 0): the pass sums mass and momentum / force, so it reads
// the velocity / the volume and stress (the mass is read by both).
template <int DIM, typename R, int MODE, bool MOM = true, bool FRC = true>
__device__ __forceinline__ void load_scatter_in(const PartPtrs<R>& pp, int i, ScatterIn<R>& P) {
    load_x<DIM>(pp.x, i, P.x);
    if constexpr (MODE == 0) {
        P.m = pp.ms[i];
        if constexpr (MOM) {
#pragma unroll
            for (int a = 0; a < DIM; ++a) P.v[a] = pp.v[a][i];
        }
        if constexpr (FRC) {
            P.vol = pp.vol[i];
            P.sg[0] = pp.sig[0][i];
            P.sg[1] = pp.sig[1][i];
            P.sg[3] = pp.sig[3][i];
            if constexpr (DIM == 3) {
                P.sg[2] = pp.sig[2][i];
                P.sg[4] = pp.sig[4][i];
                P.sg[5] = pp.sig[5][i];
            }
        }
    } else {
#pragma unroll
        for (int a = 0; a < DIM; ++a) P.v[a] = pp.pm[a][i];
    }
}

// Passes over a tile's particles.  SPLIT 0: one pass, every field and z
// layer.  SPLIT 1 (3D P2G #1): mass and momentum, then force.  SPLIT 2: the
// field groups of SPLIT 1, each one z layer of the stencil at a time.  A
// pass p sums fields [F0, F1) on z layers [Z0, Z0 + NZP) in registers.
template <int DIM, int S, int MODE, int SPLIT>
struct OwnedPasses {
    static constexpr int NF = MODE == 0 ? 1 + 2 * DIM : DIM;
    static constexpr int NZ = DIM == 3 ? S : 1;
    static constexpr int NG = (MODE == 0 && DIM == 3 && SPLIT >= 1) ? 2 : 1;  // field groups
    static constexpr int NL = (DIM == 3 && SPLIT == 2) ? NZ : 1;              // layer groups
    static constexpr int NPASS = NG * NL;
    static constexpr int NZP = NZ / NL;
    static constexpr int F0(int p) { return NG == 1 ? 0 : (p / NL == 0 ? 0 : 1 + DIM); }
    static constexpr int F1(int p) { return NG == 1 ? NF : (p / NL == 0 ? 1 + DIM : NF); }
    static constexpr int Z0(int p) { return (p % NL) * NZP; }
    static constexpr int MAXF = NG == 1 ? NF : (1 + DIM > DIM ? 1 + DIM : DIM);
};

// Register accumulators of one pass: NFP fields x NZP z layers x S rows x S
// columns (2D: NZP = 1, the row is y).
template <int NFP, int NZP, int S, typename R>
struct OwnedAcc {
    R v[NFP][NZP][S][S];
};

// Adds one particle's contributions to fields [F0, F1) on the z layers
// [Z0, Z0 + NZP) of the lane's accumulators.  MODE 0 is P2G #1 (P2G_OS!,
// utils_OS.jl:424-496): mass, momentum and internal + external force, with
// the force factored per (y, z) node pair as in p2g3_window: with A = Ny Nz,
// B = dNy Nz, C = Ny dNz the gradient is (dNx A, Nx B, Nx C), so
// f_k = dNx P_k + Nx Q_k, P = -Ω σ_{kx} A, Q_k = -Ω (σ_{ky} B + σ_{kz} C) + m A g_k.
// MODE 1 is the re-map momentum scatter (doublemapping2_OS!,
// utils_OS.jl:917-959): N pm.  The stencil's weights are zero for a lane
// that sits the round out, so the sums are branch-free.
template <int DIM, int S, typename R, int MODE, int F0, int F1, int Z0, int NZP>
__device__ __forceinline__ void owned_accumulate(const Stencil<DIM, S, R>& st, const ScatterIn<R>& P,
                                                 const Gravity<R>& G, OwnedAcc<F1 - F0, NZP, S, R>& acc) {
    auto& s = acc.v;
    if constexpr (MODE == 0) {
        constexpr bool MOM = F0 == 0;          // mass and momentum in this pass
        constexpr bool FRC = F1 == 1 + 2 * DIM;  // force in this pass
        const R m = P.m;
        const R mv = -P.vol;
        R pv[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) pv[d] = m * P.v[d];
        // -Ω σ
        const R sxx = mv * P.sg[0], syy = mv * P.sg[1], sxy = mv * P.sg[3];
        const R szz = DIM == 3 ? mv * P.sg[2] : R(0);
        const R syz = DIM == 3 ? mv * P.sg[4] : R(0);
        const R sxz = DIM == 3 ? mv * P.sg[5] : R(0);
        R mNx[S], pN[S][3];
        static_for<S>([&](auto a) {
            mNx[a] = m * st.N[0][a];
#pragma unroll
            for (int d = 0; d < DIM; ++d) pN[a][d] = pv[d] * st.N[0][a];
        });
        static_for<NZP>([&](auto cc) {
            constexpr int c = Z0 + decltype(cc)::value;
            static_for<S>([&](auto b) {
                R A, B, Cz;
                if constexpr (DIM == 3) {
                    A = st.N[1][b] * st.N[2][c];
                    B = st.dN[1][b] * st.N[2][c];
                    Cz = st.N[1][b] * st.dN[2][c];
                } else {
                    A = st.N[1][b];
                    B = st.dN[1][b];
                    Cz = R(0);
                }
                R Pf[3] = {0, 0, 0}, Qf[3] = {0, 0, 0};
                if constexpr (FRC) {
                    const R mA = m * A;
                    if constexpr (DIM == 3) {
                        Pf[0] = sxx * A; Pf[1] = sxy * A; Pf[2] = sxz * A;
                        Qf[0] = sxy * B + sxz * Cz + mA * G.g[0];
                        Qf[1] = syy * B + syz * Cz + mA * G.g[1];
                        Qf[2] = syz * B + szz * Cz + mA * G.g[2];
                    } else {
                        Pf[0] = sxx * A; Pf[1] = sxy * A;
                        Qf[0] = sxy * B + mA * G.g[0];
                        Qf[1] = syy * B + mA * G.g[1];
                    }
                }
                static_for<S>([&](auto a) {
                    if constexpr (MOM) {
                        s[0][cc][b][a] += mNx[a] * A;
#pragma unroll
                        for (int d = 0; d < DIM; ++d) s[1 + d][cc][b][a] += pN[a][d] * A;
                    }
                    if constexpr (FRC) {
                        constexpr int o = 1 + DIM - F0;
#pragma unroll
                        for (int d = 0; d < DIM; ++d)
                            s[o + d][cc][b][a] += st.dN[0][a] * Pf[d] + st.N[0][a] * Qf[d];
                    }
                });
            });
        });
    } else {
        R pN[S][3];
        static_for<S>([&](auto a) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) pN[a][d] = P.v[d] * st.N[0][a];
        });
        static_for<NZP>([&](auto cc) {
            constexpr int c = Z0 + decltype(cc)::value;
            static_for<S>([&](auto b) {
                const R A = DIM == 3 ? st.N[1][b] * st.N[2][c] : st.N[1][b];
                static_for<S>([&](auto a) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) s[d][cc][b][a] += pN[a][d] * A;
                });
            });
        });
    }
}

// Window read-add-write of fields [F0, F1) of node k (c[f - F0]); a pair of
// the paired layout with one field in the range is accessed as one float.
template <int NF, int F0, int F1, typename R, int WN, bool PAIRED>
__device__ __forceinline__ void window_add_fields(lds_ptr<R> W, int k, const R* c) {
    if constexpr (PAIRED && F0 == 0 && F1 == NF) {
        window_add_node<NF, R, WN, true>(W, k, c);
    } else {
        constexpr int NFP = F1 - F0;
        R o[NFP];
        static_for<NFP>([&](auto q) {
            constexpr int f = F0 + decltype(q)::value;
            o[q] = W[PAIRED ? paired_offset<NF>(f, k, WN) : f * WN + k];
        });
        static_for<NFP>([&](auto q) {
            constexpr int f = F0 + decltype(q)::value;
            W[PAIRED ? paired_offset<NF>(f, k, WN) : f * WN + k] = o[q] + c[q];
        });
    }
}

// Contribution of one particle to one node (drifted particles).
template <int DIM, typename R, int MODE>
__device__ __forceinline__ void owned_node_values(R N, R dNx, R dNy, R dNz, const ScatterIn<R>& P,
                                                  const Gravity<R>& G, R* c) {
    if constexpr (MODE == 0) {
        R pv[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) pv[d] = P.m * P.v[d];
        p2g_values<DIM, R>(N, dNx, dNy, dNz, P.m, P.vol, pv, P.sg, G, c);
    } else {
#pragma unroll
        for (int d = 0; d < DIM; ++d) c[d] = P.v[d] * N;
    }
}

// Stencil loop of the drift paths: a runtime loop over the S^(D-1) (y, z)
// rows with x unrolled (the rare drifted particles need few registers more
// than they need speed); body(row offset, column a, N, dNx, dNy, dNz) in the
// factor order of for_stencil.
template <int DIM, int S, typename R, typename F>
__device__ __forceinline__ void for_stencil_rows(const Stencil<DIM, S, R>& st, F&& body) {
    constexpr int NROW = DIM == 3 ? S * S : S;
#pragma unroll 1
    for (int q = 0; q < NROW; ++q) {
        const int b = q % S, c = DIM == 3 ? q / S : 0;
        R Ny = st.N[1][0], dNy = st.dN[1][0], Nz = R(1), dNz = R(0);
        static_for<S>([&](auto k) {
            if (b == int(k)) {
                Ny = st.N[1][k];
                dNy = st.dN[1][k];
            }
            if (DIM == 3 && c == int(k)) {
                Nz = st.N[2][k];
                dNz = st.dN[2][k];
            }
        });
        static_for<S>([&](auto a) {
            const R Nx = st.N[0][a];
            R N, dNx, dNyv, dNzv;
            if constexpr (DIM == 3) {
                N = Nx * Ny * Nz;
                dNx = st.dN[0][a] * Ny * Nz;
                dNyv = dNy * Nx * Nz;
                dNzv = dNz * Nx * Ny;
            } else {
                N = Nx * Ny;
                dNx = st.dN[0][a] * Ny;
                dNyv = dNy * Nx;
                dNzv = R(0);
            }
            body(b, c, int(a), N, dNx, dNyv, dNzv);
            __builtin_amdgcn_sched_barrier(0);
        });
    }
}

// ---------------------------------------------------------------------------
// Pieces shared by the base-owned kernels.
// ---------------------------------------------------------------------------
constexpr int kOwnedLast = (1 << kRankBits) - 1;  // rank of the merged last group
constexpr int kDriftCap = 256;                    // drift list entries (LDS)

// The rounds of one tile as seen by the lane that owns base `lane`.
struct TileRounds {
    const unsigned long long* tmask;
    const int* toff;
    unsigned long long lt;  // lanes below this one
    int lane, beg, end, nr, xbeg, xcnt;

    __device__ __forceinline__ void init(int tile, int lane_, const unsigned long long* masks,
                                         const int* offs, const int* ovf, const int* rounds,
                                         const int* seg_begin, const int* seg_end) {
        lane = lane_;
        lt = (1ull << lane) - 1ull;
        beg = seg_begin[tile];
        end = seg_end[tile];
        nr = rounds[tile];
        tmask = masks + (long long)tile * kOwnedLast;
        toff = offs + (long long)tile * (kOwnedLast + 1);
        // the merged last group: this base's particles beyond rank 30
        xbeg = 0;
        xcnt = 0;
        if (nr > kOwnedLast) {
            const int o = ovf[(long long)tile * 64 + lane];
            const int o1 = lane < 63 ? ovf[(long long)tile * 64 + lane + 1] : end - beg - toff[kOwnedLast];
            xbeg = beg + toff[kOwnedLast] + o;
            xcnt = o1 - o;
        }
    }
    // particle index of round r (0 <= r < nr) for this lane; false if it sits out
    __device__ __forceinline__ bool index(int r, int& i) const {
        if (r < kOwnedLast) {
            const unsigned long long mk = tmask[r];
            i = beg + toff[r] + __popcll(mk & lt);
            return (mk >> lane) & 1ull;
        }
        i = xbeg + (r - kOwnedLast);
        return r - kOwnedLast < xcnt;
    }
};

// Fields a scatter adds to: MODE 0 m, p, f; MODE 1 p.
template <int DIM, typename R, int MODE, int NF>
__device__ __forceinline__ void scatter_fields(const GridPtrs<R>& gp, R* (&fields)[NF]) {
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
}

// The lane's sums of fields [F0, F1) on z layers [Z0, Z0 + NZP) into the
// window, one stencil slot at a time (the 64 bases reach 64 distinct nodes at
// each slot).  The empty asm keeps the compiler from holding the slot
// addresses across the rounds.
template <int DIM, int S, int NF, int F0, int F1, int Z0, int NZP, typename R, int WW, int ZS,
          int WN, bool PAIRED>
__device__ __forceinline__ void owned_add_sums(lds_ptr<R> win, int wbase0,
                                               const OwnedAcc<F1 - F0, NZP, S, R>& acc, bool any) {
    int wbase = wbase0;
    asm volatile("" : "+v"(wbase));
    if (!any) return;
    static_for<NZP>([&](auto cc) {
        static_for<S>([&](auto b) {
            static_for<S>([&](auto a) {
                constexpr int C = Z0 + decltype(cc)::value;
                const int k = wbase + (DIM == 3 ? C * ZS : 0) + int(b) * WW + int(a);
                R c[F1 - F0];
#pragma unroll
                for (int f = 0; f < F1 - F0; ++f) c[f] = acc.v[f][cc][b][a];
                window_add_fields<NF, F0, F1, R, WN, PAIRED>(win, k, c);
                __builtin_amdgcn_sched_barrier(0);
            });
        });
    });
}

template <int NFP, int NZP, int S, typename R>
__device__ __forceinline__ void owned_zero(OwnedAcc<NFP, NZP, S, R>& acc) {
#pragma unroll
    for (int f = 0; f < NFP; ++f)
#pragma unroll
        for (int cc = 0; cc < NZP; ++cc)
#pragma unroll
            for (int b = 0; b < S; ++b)
#pragma unroll
                for (int a = 0; a < S; ++a) acc.v[f][cc][b][a] = R(0);
}

// Drifted particles drift[k0, k0 + n), n <= 64, one per lane: claim rounds
// into the window (write a unique tag to the base's claim slot, read it back;
// the winners add with the lockstep read-add-write, the rest claim again),
// global atomics for a stencil outside the window.  Every lane calls this
// (wave-uniform n).  MODE 1 reads the momentum from pp.pm.
template <int DIM, int S, int NB, int WW, int ZS, int WN, typename R, int MODE, int NF, bool PAIRED>
__device__ __forceinline__ void owned_drift_batch(const GridGeom& g, const int* w0,
                                                  const PartPtrs<R>& pp, const Gravity<R>& G,
                                                  R* const* fields, lds_ptr<R> win,
                                                  lds_ptr<unsigned> claim, lds_ptr<int> drift,
                                                  int k0, int n, unsigned& tag) {
    const int lane = threadIdx.x;
    ScatterIn<R> P;
    Stencil<DIM, S, R> st;
    bool pend = false;
    int node0 = 0, bidx = 0;
    if (lane < n) {
        const int i = drift[k0 + lane];
        load_scatter_in<DIM, R, MODE>(pp, i, P);
        int fl2 = 0;
        if (make_stencil<DIM, S, R>(g, P.x, st, fl2)) {
            const bool inside = window_slot<DIM, NB, WW, ZS>(st.base, w0, node0, bidx);
            if (inside) {
                pend = true;
            } else {
                for_stencil_rows<DIM, S, R>(st, [&](int b, int c, int a, R N, R dNx, R dNy, R dNz) {
                    if (N != R(0)) {
                        R cv[NF];
                        owned_node_values<DIM, R, MODE>(N, dNx, dNy, dNz, P, G, cv);
                        const long long nn = node_id(g, st.base[0] + a, st.base[1] + b,
                                                     DIM == 3 ? st.base[2] + c : 0);
#pragma unroll
                        for (int f = 0; f < NF; ++f) atomic_add(&fields[f][nn], cv[f]);
                    }
                });
            }
        }
    }
    while (__ballot(pend) != 0ull) {
        bool won = false;
        if (pend) {
            claim[bidx] = tag;
            won = claim[bidx] == tag;
        }
        tag += 64u;
        if (won) {
            for_stencil_rows<DIM, S, R>(st, [&](int b, int c, int a, R N, R dNx, R dNy, R dNz) {
                if (N != R(0)) {
                    R cv[NF];
                    owned_node_values<DIM, R, MODE>(N, dNx, dNy, dNz, P, G, cv);
                    window_add_node<NF, R, WN, PAIRED>(win, node0 + c * ZS + b * WW + a, cv);
                }
            });
            pend = false;
        }
    }
}

// Stencil weights of a lane that is not home zeroed (its sums add nothing),
// and the derivative of a zero axis weight zeroed: the reference skips nodes
// with Ni == 0 (utils_OS.jl:431,470), and a zero axis weight zeroes every node
// of its row (the linear basis has a non-zero derivative on a node).
template <int DIM, int S, typename R>
__device__ __forceinline__ void owned_mask_stencil(Stencil<DIM, S, R>& st, bool home) {
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const R n = home ? st.N[d][k] : R(0);
            st.dN[d][k] = n == R(0) ? R(0) : st.dN[d][k];
            st.N[d][k] = n;
        }
}

// ---------------------------------------------------------------------------
// The scatter kernel.  MODE 0: P2G #1; MODE 1: the MUSL re-map momentum
// scatter (from pp.pm).  SPLIT: passes over the tile (OwnedPasses).
// ---------------------------------------------------------------------------
template <int DIM, int S, int TSH, typename R, int MODE, int SPLIT>
__global__ __launch_bounds__(64) void k_scatter_owned(
    GridGeom g, TileGeom tg, PartPtrs<R> pp, GridPtrs<R> gp, Gravity<R> G,
    const unsigned long long* __restrict__ masks, const int* __restrict__ offs,
    const int* __restrict__ ovf, const int* __restrict__ rounds,
    const int* __restrict__ tile_list, const int* __restrict__ seg_begin,
    const int* __restrict__ seg_end, const int* __restrict__ count_ptr, int* flags,
    R* __restrict__ tb) {
    using W = Win<DIM, S, TSH>;
    using PS = OwnedPasses<DIM, S, MODE, SPLIT>;
    static_assert(kWinSlack == 0, "lane-owned bases assume a slack-free window");
    static_assert(W::NBN == 64, "one base per lane");
    constexpr int WW = W::WW, WN = W::WN, NB = W::NB, ZS = W::ZS;
    constexpr int NF = PS::NF;
    constexpr bool PAIRED = kPairedP2G<DIM, NF, R, WN>;
    __shared__ __attribute__((aligned(16))) R win_[NF * WN];
    __shared__ unsigned claim_[64];
    __shared__ int drift_[kDriftCap];
    lds_ptr<R> win = (lds_ptr<R>)win_;
    lds_ptr<unsigned> claim = (lds_ptr<unsigned>)claim_;
    lds_ptr<int> drift = (lds_ptr<int>)drift_;
    R* fields[NF];
    scatter_fields<DIM, R, MODE, NF>(gp, fields);
    const int lane = threadIdx.x;
    const unsigned long long lt = (1ull << lane) - 1ull;
    // the lane's base inside the tile (tile_local order), and its node in the window
    int lx, ly, lz;
    tile_local_xyz<DIM>(TSH, lane, lx, ly, lz);
    const int wbase0 = DIM == 3 ? lz * ZS + ly * WW + lx : ly * WW + lx;
    unsigned tag = unsigned(lane);
    int flag = 0;
    int e0, e1, est;
    xcd_chunk(*count_ptr, e0, e1, est);
    for (int e = e0; e < e1; e += est) {
        const int tile = tile_list[e];
        TileRounds tr;
        tr.init(tile, lane, masks, offs, ovf, rounds, seg_begin, seg_end);
        int w0[3];
        tile_window_origin<DIM>(tg, tile, w0);
        const int myb[3] = {w0[0] + lx, w0[1] + ly, w0[2] + lz};
        for (int k = lane; k < NF * WN; k += 64) win[k] = R(0);
        int ndrift = 0;
        auto drift_all = [&] {
            for (int k0 = 0; k0 < ndrift; k0 += 64)
                owned_drift_batch<DIM, S, NB, WW, ZS, WN, R, MODE, NF, PAIRED>(
                    g, w0, pp, G, fields, win, claim, drift, k0, min(64, ndrift - k0), tag);
            ndrift = 0;
        };
        static_for<PS::NPASS>([&](auto pass) {
            constexpr int p = decltype(pass)::value;
            constexpr int F0 = PS::F0(p), F1 = PS::F1(p), Z0 = PS::Z0(p), NZP = PS::NZP;
            constexpr bool FIRST = p == 0;
            OwnedAcc<F1 - F0, NZP, S, R> acc;
            owned_zero(acc);
            constexpr bool MOM = MODE != 0 || F0 == 0, FRC = MODE != 0 || F1 == NF;
            bool any = false;  // the lane's base holds a particle
            // Every lane loads (a lane that sits out reads the tile's first
            // particle) and sums with zero weights unless its particle is home.
            // Software pipeline: round r + 1's loads are issued before round r's sums.
            int in_ = 0;
            bool actn = tr.nr > 0 && tr.index(0, in_);
            ScatterIn<R> Pn;
            load_scatter_in<DIM, R, MODE, MOM, FRC>(pp, actn ? in_ : tr.beg, Pn);
            for (int r = 0; r < tr.nr; ++r) {
                const int i = in_;
                const bool act = actn;
                const ScatterIn<R> P = Pn;
                if (r + 1 < tr.nr) {
                    actn = tr.index(r + 1, in_);
                    load_scatter_in<DIM, R, MODE, MOM, FRC>(pp, actn ? in_ : tr.beg, Pn);
                }
                Stencil<DIM, S, R> st;
                int fl2 = 0;
                const bool ok = make_stencil<DIM, S, R>(g, P.x, st, fl2);
                if (FIRST && MODE == 0 && act) flag |= fl2;
                const bool home = act && ok && st.base[0] == myb[0] && st.base[1] == myb[1] &&
                                  (DIM == 2 || st.base[2] == myb[2]);
                any = any || home;
                owned_mask_stencil(st, home);
                owned_accumulate<DIM, S, R, MODE, F0, F1, Z0, NZP>(st, P, G, acc);
                if constexpr (FIRST) {
                    const bool drifted = act && ok && !home;
                    const unsigned long long dm = __ballot(drifted);
                    if (dm != 0ull) {
                        if (drifted) drift[ndrift + __popcll(dm & lt)] = i;
                        ndrift += __popcll(dm);
                        if (ndrift > kDriftCap - 64) {
                            // The list is nearly full.  The sums so far go to the
                            // window first, so the batch does not hold registers
                            // beside them.
                            owned_add_sums<DIM, S, NF, F0, F1, Z0, NZP, R, WW, ZS, WN, PAIRED>(
                                win, wbase0, acc, any);
                            owned_zero(acc);
                            any = false;
                            drift_all();
                        }
                    }
                }
            }
            owned_add_sums<DIM, S, NF, F0, F1, Z0, NZP, R, WW, ZS, WN, PAIRED>(win, wbase0, acc, any);
        });
        drift_all();
        flush_window<DIM, WW, ZS, WN, NF, R>(g, w0, win, fields,
                                             tb ? tb + (long long)e * part_stride<NF>() * WN : nullptr);
    }
    raise_flag(flags, flag);
}

// FP32 uGIMP per-axis constants of make_stencil staged in LDS (kStencilK per
// axis: h, 2h, 1/h, lp^2, c4, c1, lp/2, h + lp/2, h - lp/2, as ugimp3_f32 forms
// them).  Read back with volatile loads at each use, they are not hoisted out
// of the particle loop into registers the sums need.
constexpr int kStencilK = 9;
__device__ __forceinline__ void stage_stencil_consts(const GridGeom& g, lds_ptr<float> kc) {
    const int t = threadIdx.x;
    if (t < 3) {
        const float h = float(g.h[t]), half = float(g.half[t]), lp = float(g.lp[t]);
        lds_ptr<float> k = kc + t * kStencilK;
        k[0] = h;
        k[1] = 2.0f * h;
        k[2] = float(g.hinv[t]);
        k[3] = lp * lp;
        k[4] = float(g.c4[t]);
        k[5] = float(g.c1[t]);
        k[6] = half;
        k[7] = h + half;
        k[8] = h - half;
    }
}

// make_stencil's FP32 3D uGIMP path with the per-axis constants from LDS.
__device__ __forceinline__ bool make_stencil3_lds(const GridGeom& g, const double* x,
                                                  Stencil<3, 3, float>& st, int& flag,
                                                  lds_ptr<float> kc) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double u = (x[a] - g.org[a]) * g.hinv[a];
        if (!isfinite(u)) {
            flag |= FLAG_NONFINITE;
            return false;
        }
        const double fb = stencil_base_f<3>(a, u);
        int b = 0;
        if (fb < 0.0 || fb > double(g.nn[a] - 3)) ok = false;
        else b = int(fb);
        st.base[a] = b;
        const float d0 = float(x[a] - (g.org[a] + fb * g.h[a]));
        lds_ptr<float> k = kc + a * kStencilK;
        const float h = k[0], h2 = k[1], hinv = k[2], lp2 = k[3], c4 = k[4], c1 = k[5],
                    half = k[6], hph = k[7], hmh = k[8];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float d = j == 0 ? d0 : (j == 1 ? d0 - h : d0 - h2);
            const float ad = fabsf(d);
            const float t = hph - ad;
            const float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
            const float Nc = 1.0f - (4.0f * d * d + lp2) * c4;
            const float Nl = 1.0f - ad * hinv;
            const float Nq = t * t * 0.5f * c1;
            const float dNc = -8.0f * d * c4;
            const float dNl = -sg * hinv;
            const float dNq = -sg * t * c1;
            const bool in_c = ad < half, in_l = ad < hmh, in_q = ad < hph;
            st.N[a][j] = in_c ? Nc : (in_l ? Nl : (in_q ? Nq : 0.0f));
            st.dN[a][j] = in_c ? dNc : (in_l ? dNl : (in_q ? dNq : 0.0f));
        }
    }
    if (!ok) flag |= FLAG_RANGE;
    return ok;
}

// ---------------------------------------------------------------------------
// Pair rounds (FP32 3D uGIMP).  A lane takes two particles of one stencil
// base, ranks 2k and 2k + 1 of the last sort, and per stencil slot adds both
// contributions in registers, then does one read-add-write of the window node:
// half the window traffic and adds of the rank rounds (one read-add-write per
// particle and slot), with two particles' stencils in registers instead of the
// base-owned 27 x 7 sums.  A round holds the bases of pair group k (the bases
// with more than 2k particles, distinct, so their targets are distinct at
// every slot) and, when both fit in 64 lanes, those of group k + 1 in a second
// window (the two are summed before the flush).  Lane j of group k holds
// particle start + off[2k] + j, of the j-th base of mask[2k] (found through a
// 64-entry LDS table the base's lane writes), and its partner
// start + off[2k + 1] + popcount(mask[2k + 1] & bases below) (mpm_sort.hpp's
// layout).  Ranks from 30 on (a dual cell with more than 30 particles) are
// taken with one lane per base.  A particle off its sort-time base goes to a
// drift list and claims its base as in k_scatter_owned.  One wavefront per tile.
// ---------------------------------------------------------------------------
#ifndef MPM_PAIRS_WAVES
#define MPM_PAIRS_WAVES 3
#endif
constexpr int kPairsWaves = MPM_PAIRS_WAVES;  // P2G #1's wavefronts per SIMD
template <int S, int TSH, int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MODE == 0 ? kPairsWaves : 4))) void k_scatter_pairs(
    GridGeom g, TileGeom tg, PartPtrs<float> pp, GridPtrs<float> gp, Gravity<float> G,
    const unsigned long long* __restrict__ masks, const int* __restrict__ offs,
    const int* __restrict__ ovf, const int* __restrict__ rounds,
    const int* __restrict__ tile_list, const int* __restrict__ seg_begin,
    const int* __restrict__ seg_end, const int* __restrict__ count_ptr, int* flags,
    float* __restrict__ tb) {
    using R = float;
    using W = Win<3, S, TSH>;
    static_assert(S == 3, "uGIMP");
    static_assert(kWinSlack == 0, "lane-owned bases assume a slack-free window");
    static_assert(W::NBN == 64 && kLocalBits == 6, "64 bases per tile");
    constexpr int WW = W::WW, WN = W::WN, NB = W::NB, ZS = W::ZS;
    constexpr int NF = MODE == 0 ? 7 : 3;
    static_assert(kPairedP2G<3, NF, R, WN>, "paired window");
    constexpr int kPairGroups = 15;  // pair groups over ranks 0..29 (masks hold ranks < kOwnedLast)
    static_assert(2 * kPairGroups < kOwnedLast, "a pair group's ranks have masks");
    constexpr int kCap = 192;        // drift list entries (a round appends at most 128)
    __shared__ __attribute__((aligned(16))) R win_[2 * NF * WN];
    __shared__ unsigned claim_[64];
    __shared__ int drift_[kCap];
    __shared__ float kc_[3 * kStencilK];
    __shared__ int item_base_[64];
    lds_ptr<R> win = (lds_ptr<R>)win_;
    lds_ptr<unsigned> claim = (lds_ptr<unsigned>)claim_;
    lds_ptr<int> drift = (lds_ptr<int>)drift_;
    lds_ptr<float> kc = (lds_ptr<float>)kc_;
    lds_ptr<int> item_base = (lds_ptr<int>)item_base_;
    stage_stencil_consts(g, kc);
    __syncthreads();
    R* fields[NF];
    scatter_fields<3, R, MODE, NF>(gp, fields);
    const int lane = threadIdx.x;
    const unsigned long long lt = (1ull << lane) - 1ull;
    unsigned tag = unsigned(lane);
    int flag = 0;
    int e0, e1, est;
    xcd_chunk(*count_ptr, e0, e1, est);
    for (int e = e0; e < e1; e += est) {
        const int tile = tile_list[e];
        TileRounds tr;
        tr.init(tile, lane, masks, offs, ovf, rounds, seg_begin, seg_end);
        int w0[3];
        tile_window_origin<3>(tg, tile, w0);
        // the second window only for a tile with a round of two groups
        const int np = min((tr.nr + 1) >> 1, kPairGroups);
        bool two_any = false;
        for (int k = 0; k + 1 < np; ++k)
            two_any = two_any || __popcll(tr.tmask[2 * k]) + __popcll(tr.tmask[2 * k + 2]) <= 64;
        using f4 = float __attribute__((ext_vector_type(4)));
        using lds_f4 = volatile __attribute__((address_space(3))) f4*;
        lds_f4 win4 = reinterpret_cast<lds_f4>(win);
        constexpr int WN4 = NF * WN / 4;
        static_assert(NF * WN % 4 == 0, "whole float4 windows");
        for (int k = lane; k < (two_any ? 2 : 1) * WN4; k += 64) win4[k] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        int ndrift = 0;
        // drifted particles drift[k0, k0 + n): claim rounds into window 0,
        // global atomics for a stencil outside the window
        auto drift_batch = [&](int k0, int n) __attribute__((always_inline)) {
            ScatterIn<R> P;
            Stencil<3, S, R> st;
            bool pend = false;
            int node0 = 0, bidx = 0;
            if (lane < n) {
                const int i = drift[k0 + lane];
                load_scatter_in<3, R, MODE>(pp, i, P);
                int fl2 = 0;
                if (make_stencil3_lds(g, P.x, st, fl2, kc)) {
                    if (window_slot<3, NB, WW, ZS>(st.base, w0, node0, bidx)) {
                        pend = true;
                    } else {
                        for_stencil_rows<3, S, R>(st, [&](int b, int c, int a, R N, R dNx, R dNy, R dNz) {
                            if (N != R(0)) {
                                R cv[NF];
                                owned_node_values<3, R, MODE>(N, dNx, dNy, dNz, P, G, cv);
                                const long long nn = node_id(g, st.base[0] + a, st.base[1] + b, st.base[2] + c);
#pragma unroll
                                for (int f = 0; f < NF; ++f) atomic_add(&fields[f][nn], cv[f]);
                            }
                        });
                    }
                }
            }
            while (__ballot(pend) != 0ull) {
                bool won = false;
                if (pend) {
                    claim[bidx] = tag;
                    won = claim[bidx] == tag;
                }
                tag += 64u;
                if (won) {
                    for_stencil_rows<3, S, R>(st, [&](int b, int c, int a, R N, R dNx, R dNy, R dNz) {
                        if (N != R(0)) {
                            R cv[NF];
                            owned_node_values<3, R, MODE>(N, dNx, dNy, dNz, P, G, cv);
                            window_add_node<NF, R, WN, true>(win, node0 + c * ZS + b * WW + a, cv);
                        }
                    });
                    pend = false;
                }
            }
        };
        auto drift_all = [&]() __attribute__((always_inline)) {
            for (int k0 = 0; k0 < ndrift; k0 += 64) drift_batch(k0, min(64, ndrift - k0));
            ndrift = 0;
        };
        // One lane's pair: particles i0 (a0) and i1 (a1) of tile-local base l,
        // added into window side.  An inactive particle is replaced by the
        // tile's first one with zero weights.
        auto pair = [&](int i0, bool a0, int i1, bool a1, int l, int side) __attribute__((always_inline)) {
            int lx, ly, lz;
            tile_local_xyz<3>(TSH, l, lx, ly, lz);
            const int wb = lz * ZS + ly * WW + lx;
            lds_ptr<R> ws = win + side * NF * WN;
            const int bx = w0[0] + lx, by = w0[1] + ly, bz = w0[2] + lz;
            ScatterIn<R> P0, P1;
            load_scatter_in<3, R, MODE>(pp, a0 ? i0 : tr.beg, P0);
            load_scatter_in<3, R, MODE>(pp, a1 ? i1 : tr.beg, P1);
            Stencil<3, S, R> s0, s1;
            int f0 = 0, f1 = 0;
            const bool ok0 = make_stencil3_lds(g, P0.x, s0, f0, kc);
            const bool ok1 = make_stencil3_lds(g, P1.x, s1, f1, kc);
            if (MODE == 0) {
                if (a0) flag |= f0;
                if (a1) flag |= f1;
            }
            const bool h0 = a0 && ok0 && s0.base[0] == bx && s0.base[1] == by && s0.base[2] == bz;
            const bool h1 = a1 && ok1 && s1.base[0] == bx && s1.base[1] == by && s1.base[2] == bz;
            owned_mask_stencil(s0, h0);
            owned_mask_stencil(s1, h1);
            if (h0 || h1) {
                if constexpr (MODE == 0) {
                    // per particle f_k = dNx P_k + Nx Q_k per (y, z) row, as in p2g3_window
                    const R m0 = P0.m, m1 = P1.m, v0 = -P0.vol, v1 = -P1.vol;
                    const R p00 = m0 * P0.v[0], p01 = m0 * P0.v[1], p02 = m0 * P0.v[2];
                    const R p10 = m1 * P1.v[0], p11 = m1 * P1.v[1], p12 = m1 * P1.v[2];
                    static_for<S>([&](auto c) {
                        static_for<S>([&](auto b) {
                            const R A0 = s0.N[1][b] * s0.N[2][c], A1 = s1.N[1][b] * s1.N[2][c];
                            const R B0 = s0.dN[1][b] * s0.N[2][c], B1 = s1.dN[1][b] * s1.N[2][c];
                            const R C0 = s0.N[1][b] * s0.dN[2][c], C1 = s1.N[1][b] * s1.dN[2][c];
                            const R mA0 = m0 * A0, mA1 = m1 * A1;
                            const R Px0 = v0 * P0.sg[0] * A0, Py0 = v0 * P0.sg[3] * A0,
                                    Pz0 = v0 * P0.sg[5] * A0;
                            const R Px1 = v1 * P1.sg[0] * A1, Py1 = v1 * P1.sg[3] * A1,
                                    Pz1 = v1 * P1.sg[5] * A1;
                            const R Qx0 = v0 * P0.sg[3] * B0 + v0 * P0.sg[5] * C0 + mA0 * G.g[0];
                            const R Qy0 = v0 * P0.sg[1] * B0 + v0 * P0.sg[4] * C0 + mA0 * G.g[1];
                            const R Qz0 = v0 * P0.sg[4] * B0 + v0 * P0.sg[2] * C0 + mA0 * G.g[2];
                            const R Qx1 = v1 * P1.sg[3] * B1 + v1 * P1.sg[5] * C1 + mA1 * G.g[0];
                            const R Qy1 = v1 * P1.sg[1] * B1 + v1 * P1.sg[4] * C1 + mA1 * G.g[1];
                            const R Qz1 = v1 * P1.sg[4] * B1 + v1 * P1.sg[2] * C1 + mA1 * G.g[2];
                            static_for<S>([&](auto a) {
                                const R Nx0 = s0.N[0][a], dNx0 = s0.dN[0][a];
                                const R Nx1 = s1.N[0][a], dNx1 = s1.dN[0][a];
                                const R N0 = Nx0 * A0, N1 = Nx1 * A1;
                                if (N0 != R(0) || N1 != R(0)) {
                                    R cv[7];
                                    cv[0] = m0 * N0 + m1 * N1;
                                    cv[1] = p00 * N0 + p10 * N1;
                                    cv[2] = p01 * N0 + p11 * N1;
                                    cv[3] = p02 * N0 + p12 * N1;
                                    cv[4] = fmaf(Nx1, Qx1, fmaf(dNx1, Px1, fmaf(Nx0, Qx0, dNx0 * Px0)));
                                    cv[5] = fmaf(Nx1, Qy1, fmaf(dNx1, Py1, fmaf(Nx0, Qy0, dNx0 * Py0)));
                                    cv[6] = fmaf(Nx1, Qz1, fmaf(dNx1, Pz1, fmaf(Nx0, Qz0, dNx0 * Pz0)));
                                    window_add_node<7, R, WN, true>(
                                        ws, wb + int(c) * ZS + int(b) * WW + int(a), cv);
                                }
                                __builtin_amdgcn_sched_barrier(0);
                            });
                        });
                    });
                } else {
                    static_for<S>([&](auto c) {
                        static_for<S>([&](auto b) {
                            const R A0 = s0.N[1][b] * s0.N[2][c], A1 = s1.N[1][b] * s1.N[2][c];
                            static_for<S>([&](auto a) {
                                const R N0 = s0.N[0][a] * A0, N1 = s1.N[0][a] * A1;
                                if (N0 != R(0) || N1 != R(0)) {
                                    R cv[3];
#pragma unroll
                                    for (int d = 0; d < 3; ++d) cv[d] = P0.v[d] * N0 + P1.v[d] * N1;
                                    window_add_node<3, R, WN, true>(
                                        ws, wb + int(c) * ZS + int(b) * WW + int(a), cv);
                                }
                                __builtin_amdgcn_sched_barrier(0);
                            });
                        });
                    });
                }
            }
            const bool d0 = a0 && ok0 && !h0, d1 = a1 && ok1 && !h1;
            const unsigned long long dm0 = __ballot(d0), dm1 = __ballot(d1);
            if ((dm0 | dm1) != 0ull) {
                if (d0) drift[ndrift + __popcll(dm0 & lt)] = i0;
                ndrift += __popcll(dm0);
                if (d1) drift[ndrift + __popcll(dm1 & lt)] = i1;
                ndrift += __popcll(dm1);
                if (ndrift > kCap - 128) drift_all();
            }
        };
        // Packed rounds over the pair groups of ranks 0..29, then ranks from 30
        // on with one lane per base; one call site of pair().  The masks and
        // offsets are read at wave-uniform ranks and selected per lane.
        int k = 0, rt = 2 * kPairGroups;
        while (k < np || rt < tr.nr) {
            int i0 = 0, i1 = 0, l = lane, side = 0;
            bool a0 = false, a1 = false;
            if (k < np) {
                const int rA = 2 * k;
                const bool hasB = k + 1 < np;
                const unsigned long long mA = tr.tmask[rA];
                const unsigned long long mB = hasB ? tr.tmask[rA + 2] : 0ull;
                const unsigned long long pA = rA + 1 < tr.nr ? tr.tmask[rA + 1] : 0ull;
                const unsigned long long pB = hasB && rA + 3 < tr.nr ? tr.tmask[rA + 3] : 0ull;
                const int oA = tr.toff[rA], oA1 = tr.toff[rA + 1];
                const int oB = hasB ? tr.toff[rA + 2] : 0, oB1 = hasB ? tr.toff[rA + 3] : 0;
                const int nA = __popcll(mA), nB = __popcll(mB);
                const bool two = hasB && nA + nB <= 64;
                // item -> base: the base's lane writes its base at its item index
                if ((mA >> lane) & 1ull) item_base[__popcll(mA & lt)] = lane;
                if (two && ((mB >> lane) & 1ull)) item_base[nA + __popcll(mB & lt)] = lane;
                a0 = lane < nA || (two && lane < nA + nB);
                side = lane < nA ? 0 : 1;
                if (a0) {
                    l = item_base[lane];
                    const unsigned long long p = side ? pB : pA;
                    i0 = tr.beg + (side ? oB + lane - nA : oA + lane);
                    a1 = (p >> l) & 1ull;
                    i1 = tr.beg + (side ? oB1 : oA1) + __popcll(p & ((1ull << l) - 1ull));
                } else {
                    side = 0;
                }
                k += two ? 2 : 1;
            } else {
                a0 = tr.index(rt, i0);
                a1 = rt + 1 < tr.nr && tr.index(rt + 1, i1);
                rt += 2;
            }
            pair(i0, a0, i1, a1, l, side);
        }
        drift_all();
        if (two_any) {
            for (int k = lane; k < WN4; k += 64) {
                const f4 a = win4[k], b = win4[WN4 + k];
                win4[k] = a + b;
            }
        }
        flush_window<3, WW, ZS, WN, NF, R>(g, w0, win, fields,
                                           tb ? tb + (long long)e * part_stride<NF>() * WN : nullptr);
    }
    raise_flag(flags, flag);
}

// ---------------------------------------------------------------------------
// The MUSL re-map in one pass (doublemapping1(_a)_OS! + doublemapping2_OS!,
// utils_OS.jl:698-959): the tile's vsT, as (and vs when VS) are staged in LDS,
// each lane gathers its round's particle from there (a particle outside the
// window from the grid), moves it (v, the deferred position increment) and
// adds the new momentum N m v of a particle on its own base to its register
// sums, all with the step-start basis.  A drifted particle stores its
// momentum in pp.pm for the drift batch.  Replaces k_g2p1_tiled +
// k_scatter_owned<..., 1>: the particle position is read once and the
// momentum never goes through memory.
// ---------------------------------------------------------------------------
template <int DIM, int S, int TSH, typename R, bool VS>
__global__ __launch_bounds__(64) void k_remap_owned(
    GridGeom g, TileGeom tg, PartPtrs<R> pp, GridPtrs<R> gp, StepArgs<R> sa,
    const unsigned long long* __restrict__ masks, const int* __restrict__ offs,
    const int* __restrict__ ovf, const int* __restrict__ rounds,
    const int* __restrict__ tile_list, const int* __restrict__ seg_begin,
    const int* __restrict__ seg_end, const int* __restrict__ count_ptr, R* __restrict__ tb) {
    using W = Win<DIM, S, TSH>;
    static_assert(kWinSlack == 0, "lane-owned bases assume a slack-free window");
    static_assert(W::NBN == 64, "one base per lane");
    constexpr int WW = W::WW, WN = W::WN, NB = W::NB, ZS = W::ZS;
    constexpr int NF = DIM;                    // momentum sums
    constexpr int NG = (VS ? 3 : 2) * DIM;     // gathered: vsT, as[, vs]
    constexpr bool PAIRED = kPairedP2G<DIM, NF, R, WN>;
    __shared__ __attribute__((aligned(16))) R win_[NF * WN];
    __shared__ R gwin[NG * WN];
    __shared__ unsigned claim_[64];
    __shared__ int drift_[kDriftCap];
    lds_ptr<R> win = (lds_ptr<R>)win_;
    lds_ptr<unsigned> claim = (lds_ptr<unsigned>)claim_;
    lds_ptr<int> drift = (lds_ptr<int>)drift_;
    R* fields[NF];
    scatter_fields<DIM, R, 1, NF>(gp, fields);
    R* const* vsf = pic_field(gp, sa);
    R* gfields[3 * DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        gfields[d] = gp.vsT[d];
        gfields[DIM + d] = gp.as[d];
        gfields[2 * DIM + d] = vsf[d];
    }
    const Gravity<R> G0{};  // unused by the momentum scatter
    const int lane = threadIdx.x;
    const unsigned long long lt = (1ull << lane) - 1ull;
    int lx, ly, lz;
    tile_local_xyz<DIM>(TSH, lane, lx, ly, lz);
    const int wbase0 = DIM == 3 ? lz * ZS + ly * WW + lx : ly * WW + lx;
    unsigned tag = unsigned(lane);
    int e0, e1, est;
    xcd_chunk(*count_ptr, e0, e1, est);
    for (int e = e0; e < e1; e += est) {
        const int tile = tile_list[e];
        TileRounds tr;
        tr.init(tile, lane, masks, offs, ovf, rounds, seg_begin, seg_end);
        int w0[3];
        tile_window_origin<DIM>(tg, tile, w0);
        const int myb[3] = {w0[0] + lx, w0[1] + ly, w0[2] + lz};
        __syncthreads();  // the previous tile's reads of gwin are done
        for (int k = lane; k < NF * WN; k += 64) win[k] = R(0);
        stage_window<DIM, WW, ZS, WN, NG, R>(g, w0, gfields, gwin);
        __syncthreads();
        int ndrift = 0;
        auto drift_all = [&] {
            for (int k0 = 0; k0 < ndrift; k0 += 64)
                owned_drift_batch<DIM, S, NB, WW, ZS, WN, R, 1, NF, PAIRED>(
                    g, w0, pp, G0, fields, win, claim, drift, k0, min(64, ndrift - k0), tag);
            ndrift = 0;
        };
        OwnedAcc<NF, DIM == 3 ? S : 1, S, R> acc;
        owned_zero(acc);
        bool any = false;
        for (int r = 0; r < tr.nr; ++r) {
            int i = 0;
            const bool act = tr.index(r, i);
            const int il = act ? i : tr.beg;
            ScatterIn<R> P;  // P.x: position; P.v: the new momentum
            load_x<DIM>(pp.x, il, P.x);
            const R m0 = pp.ms[il];
            R v0[3];
#pragma unroll
            for (int d = 0; d < DIM; ++d) v0[d] = pp.v[d][il];
            Stencil<DIM, S, R> st;
            int fl2 = 0;  // reported by P2G #1
            const bool ok = make_stencil<DIM, S, R>(g, P.x, st, fl2);
            int node0 = 0, bidx = 0;
            const bool inside = ok && window_slot<DIM, NB, WW, ZS>(st.base, w0, node0, bidx);
            R pv[3] = {0, 0, 0};
            if (act && ok) {
                R A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, C[3] = {0, 0, 0};
                if (inside) {
                    // branch-free: zero weights add exact zeros (staged values are finite)
                    for_stencil_local<DIM, S, R>(st, WW, ZS, [&](int off, R N, R, R, R) {
                        const R* L = gwin + node0 + off;
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
                g2p1_commit<DIM, R>(pp, sa, i, A, B, C, pv, m0, v0);
            } else if (act) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) pp.dxp[d][i] = R(0);  // frozen particle
            }
#pragma unroll
            for (int d = 0; d < DIM; ++d) P.v[d] = pv[d];
            const bool home = act && ok && st.base[0] == myb[0] && st.base[1] == myb[1] &&
                              (DIM == 2 || st.base[2] == myb[2]);
            any = any || home;
            const bool drifted = act && ok && !home;
            if (drifted) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) pp.pm[d][i] = pv[d];
            }
            owned_mask_stencil(st, home);
            owned_accumulate<DIM, S, R, 1, 0, NF, 0, DIM == 3 ? S : 1>(st, P, G0, acc);
            const unsigned long long dm = __ballot(drifted);
            if (dm != 0ull) {
                if (drifted) drift[ndrift + __popcll(dm & lt)] = i;
                ndrift += __popcll(dm);
                if (ndrift > kDriftCap - 64) {
                    owned_add_sums<DIM, S, NF, 0, NF, 0, DIM == 3 ? S : 1, R, WW, ZS, WN, PAIRED>(
                        win, wbase0, acc, any);
                    owned_zero(acc);
                    any = false;
                    drift_all();
                }
            }
        }
        owned_add_sums<DIM, S, NF, 0, NF, 0, DIM == 3 ? S : 1, R, WW, ZS, WN, PAIRED>(win, wbase0,
                                                                                  acc, any);
        drift_all();
        flush_window<DIM, WW, ZS, WN, NF, R>(g, w0, win, fields,
                                             tb ? tb + (long long)e * part_stride<NF>() * WN : nullptr);
    }
}

}  // namespace mpm
