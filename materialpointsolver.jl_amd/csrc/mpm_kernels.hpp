// mpm_kernels.hpp -- gfx950 kernels of the explicit single-phase step.
//
// A MUSL step is five launches (six with MVL):
//   k_p2g        P2G #1: mass, momentum, internal + external force
//                (P2G_OS!, src/solvers/utils_OS.jl:424-496, with the basis of
//                resetmpstatus_OS!, utils_OS.jl:178-415, evaluated on the fly)
//   k_grid1      grid solve #1 (solvegrid_a_OS! / solvegrid_OS!,
//                utils_OS.jl:507-687); also clears f for the next step
//   k_g2p1_p2g2  MUSL re-map: gather vsT, as, vs, move the particle, scatter
//                the new momentum with the SAME step-start basis
//                (doublemapping1(_a)_OS! + doublemapping2_OS!, utils_OS.jl:698-959)
//   k_grid2      grid solve #2 (doublemapping3_OS!, utils_OS.jl:968-1016); also
//                clears m and p for the next step
//   k_g2p2       G2P with the step-start basis, strain / J update, constitutive
//                update and position commit (G2P_OS!, utils_OS.jl:1025-1153;
//                liE!, dpP!, mcP!)
//   k_mvl2       MVL second half (vollock2_OS!, utils_OS.jl:1202-1256); the
//                first half (vollock1_OS!) is fused into k_g2p2
//
// USL (OS_USL.jl:102-144,325-366) drops the re-map: k_grid1 also writes the
// nodal displacement and clears m, and k_g2p1_p2g2 / the tiled gather skip
// the scatter.  USF (OS_USF.jl:11-97) is USL with k_g2p2 moved to the start
// of the next step (k_mvl1 then does MVL's cell sums after the re-map).
//
// The basis is never stored: each particle pass recomputes the S per-axis
// weights (S = 3 for uGIMP, 2 for linear) from the particle position, so a
// pass streams only the particle state it needs (SoA, coalesced).  Particles
// whose stencil would leave the grid, or whose position is not finite, are
// frozen and reported through a device flag word instead of being allowed to
// read or write outside the grid arrays.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <stdint.h>

#include <utility>

namespace mpm {

// Compile-time loop: f(std::integral_constant<int, 0>) ... f(ic<N-1>).  The
// stencil loops use it so their register arrays are always indexed with
// constants (a loop the compiler declines to unroll would move them to scratch).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

constexpr double kEpsF64 = 2.220446049250313e-16;  // eps(Float64): massless-node cut

enum : int { FLAG_RANGE = 1, FLAG_NONFINITE = 2 };

// ---------------------------------------------------------------------------
// Parameters passed by value
// ---------------------------------------------------------------------------
struct GridGeom {
    int nn[3];        // nodes per axis (nn[2] = 1 in 2D)
    int nc[3];        // cells per axis
    long long ni;     // nn[0]*nn[1]*nn[2]
    double org[3];    // x1, y1, z1
    double h[3];      // dx, dy, dz
    double hinv[3];   // 1/h
    double lp[3];     // particle domain size
    // uGIMP constants per axis (mpbasisfunc.jl:56-120)
    double c4[3];     // 1/(4 h lp)
    double c1[3];     // 1/(h lp)
    double half[3];   // 0.5 lp
};

template <typename R>
struct PartPtrs {
    double* x[3];     // positions (FP64 in every precision)
    R* v[3];          // velocity
    R* sig[6];        // Cauchy stress, Voigt [xx,yy,zz,xy,yz,xz] (2D: [xx,yy,zz,xy])
    R* ms;            // mass (constant: Ω·ρs at upload)
    R* vol;           // Ω = J Ω0
    R* vol0;          // Ω0
    double* J;        // det F (FP64 in every precision)
    R* eq;            // ϵq
    R* ek;            // ϵk
    R* ev;            // ϵv
    R* eps[6];        // ϵijs (engineering shear)
    R* dxp[3];        // MUSL scratch: position increment of the current step
    int* nid;         // 0-based material id
    int* pid;         // original particle index (for downloads after sorting)
    int* p2c;         // step-start cell id (MVL only)
    R* de[6];         // strain increment of the step (deferred 3D Mohr-Coulomb return only)
    R* pm[3];         // MUSL scratch: new momentum between the re-map's gather and scatter
    R* pr[6];         // learned return: trial principal stresses (3) and rotation (3)
};

template <typename R>
struct GridPtrs {
    R* m;
    R* p[3];
    R* f[3];
    R* vs[3];
    R* vsT[3];
    R* as[3];
    R* du[3];
    unsigned char* bcf;   // bit d set: axis d velocity fixed
    R* bcv[3];            // fixed values
    R* cell_sm;           // MVL: Σ Ω σm per cell
    R* cell_vol;          // MVL: Σ Ω per cell
};

// Material constants, precomputed on the host in FP64 with the reference's
// expressions (druckerprager.jl:178-183, linearelastic.jl:48-49) and stored
// in the working precision.
template <typename R>
struct Material {
    R Ks, Gs, Dt, Dd;
    R kphi, qphi, qpsi, sigt, ab, tb;   // Drucker-Prager
    R eq_shear;                          // sqrt(0.333333 + 0.222222 qψ²)
    R sinphi, cosphi, tanphi, c, Hp, cr; // Mohr-Coulomb
    R sinpsi;                            // Mohr-Coulomb (3D)
    R phi, psi;                          // angles (rad), inputs of the learned return map
    R phir;                              // residual friction (3D Mohr-Coulomb softening)
    R eqr;                               // residual plastic strain; 0: phi, psi do not soften
};

// Body force per unit mass on each axis.  The reference applies gravity on
// the last axis only (utils_OS.jl:445-448,490-494); the other components are
// an extension for slope-aligned frames and are zero by default.
template <typename R>
struct Gravity {
    R g[3];
};

template <typename R>
struct StepArgs {
    R dt, G, flip, pic, zeta;
    int va;            // 0 = :a, 1 = :v
    int pic_new;       // USL/USF :a: the PIC term gathers vsT (the updated velocity), not vs
    int remap;         // 1: MUSL (momentum re-map); 0: USL/USF (grid solve #1 writes du)
    int commit_dx;     // k_g2p2 adds the step's position increment (0: first USF step)
    int plastic;       // Ti >= Te
    int constitutive;  // 0 LE, 1 DP, 2 MC
    int mvl;
    // 3D Mohr-Coulomb only: defer_return leaves the return mapping to the
    // learned-return kernel (trial stress in sig, strain increment in de);
    // log_* (when log_x is set) records return-mapping samples of every
    // log_stride-th particle that yields: 16 inputs and 7 outputs each.
    int defer_return;
    int log_stride;
    float* log_x;
    float* log_y;
    unsigned long long* log_count;
    long long log_cap;
};

// Tile partials of a scatter (the tile kernels' windows stored per tile, see
// flush_window): node-major, tb[slot][k][f] with the NF fields of window node
// k padded to part_stride<NF>() values (one or two vector accesses per node),
// the tile's slot from slot_of (-1: no particles), windows of ww nodes per
// axis starting at tile * 2^tsh.  tb == nullptr: the scatter added atomically
// into the grid fields.
template <int NF>
constexpr int part_stride() {
    return NF <= 1 ? 1 : NF <= 2 ? 2 : NF <= 4 ? 4 : 8;
}

// The NF values of one partial node (p: the node's first value, aligned to
// part_stride<NF>() values), in vector accesses; the padding is written as 0.
template <int NF, typename R>
__device__ __forceinline__ void load_partial(const R* p, R* v) {
    constexpr int NS = part_stride<NF>();
    using V2 = std::conditional_t<sizeof(R) == 4, float2, double2>;
    if constexpr (sizeof(R) == 4 && NS >= 4) {
#pragma unroll
        for (int q = 0; q < NS / 4; ++q) {
            const float4 t = reinterpret_cast<const float4*>(p)[q];
            const float e[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (4 * q + j < NF) v[4 * q + j] = R(e[j]);
        }
    } else if constexpr (NS >= 2) {
#pragma unroll
        for (int q = 0; q < NS / 2; ++q) {
            const V2 t = reinterpret_cast<const V2*>(p)[q];
            if (2 * q < NF) v[2 * q] = R(t.x);
            if (2 * q + 1 < NF) v[2 * q + 1] = R(t.y);
        }
    } else {
        v[0] = p[0];
    }
}

template <int NF, typename R>
__device__ __forceinline__ void store_partial(R* p, const R* v) {
    constexpr int NS = part_stride<NF>();
    using V2 = std::conditional_t<sizeof(R) == 4, float2, double2>;
    auto at = [&](int j) { return j < NF ? v[j] : R(0); };
    if constexpr (sizeof(R) == 4 && NS >= 4) {
#pragma unroll
        for (int q = 0; q < NS / 4; ++q)
            reinterpret_cast<float4*>(p)[q] =
                make_float4(float(at(4 * q)), float(at(4 * q + 1)), float(at(4 * q + 2)), float(at(4 * q + 3)));
    } else if constexpr (NS >= 2) {
#pragma unroll
        for (int q = 0; q < NS / 2; ++q) {
            V2 t;
            t.x = at(2 * q);
            t.y = at(2 * q + 1);
            reinterpret_cast<V2*>(p)[q] = t;
        }
    } else {
        p[0] = v[0];
    }
}
template <typename R>
struct TilePartials {
    const R* tb;
    const int* slot_of;
    int tsh, ww, wn;
    int tn[3];  // tiles per axis
};

// Adds to acc[0, NF) the partials of node n from every tile window that holds
// it, in a fixed order (tiles with lower indices first).  A window spans
// ww <= 2^(tsh+1) nodes per axis, so at most two tiles per axis hold a node
// (2^DIM candidates): every candidate's slot is loaded first and then every
// held window's values, so a node waits for two memory round trips rather
// than two per candidate.
template <int DIM, int NF, typename R>
__device__ __forceinline__ void add_partials(const GridGeom& g, const TilePartials<R>& tp,
                                             long long n, R* acc) {
    // node coordinates with 32-bit divisions (grids stay below 2^31 nodes);
    // 64-bit divisions here cost more than the rest of the grid solve
    const unsigned nx = unsigned(g.nn[0]), ny = unsigned(g.nn[1]);
    const unsigned un = unsigned(n);
    const unsigned row = un / nx;
    int i[3];
    i[0] = int(un - row * nx);
    if constexpr (DIM == 3) {
        const unsigned lay = row / ny;
        i[1] = int(row - lay * ny);
        i[2] = int(lay);
    } else {
        i[1] = int(row);
        i[2] = 0;
    }
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a < DIM) {
            // windows [t 2^tsh, t 2^tsh + ww - 1] that contain i
            hi[a] = min(i[a] >> tp.tsh, tp.tn[a] - 1);
            const int l = i[a] - tp.ww + 1;
            lo[a] = l <= 0 ? 0 : ((l + (1 << tp.tsh) - 1) >> tp.tsh);
        } else {
            lo[a] = hi[a] = 0;
        }
    }
    constexpr int NC = 1 << DIM;
    int koff[NC];
    int slot[NC];
    // candidate c = (dz, dy, dx) bits: tile hi - d; c = NC - 1 is the lowest tile
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = DIM == 3 ? (c >> 2) & 1 : 0;
        const int tx = hi[0] - dx, ty = hi[1] - dy, tz = hi[2] - dz;
        const bool ok = tx >= lo[0] && ty >= lo[1] && tz >= lo[2];
        slot[c] = ok ? tp.slot_of[tx + tp.tn[0] * (ty + tp.tn[1] * tz)] : -1;
        koff[c] = (i[0] - (tx << tp.tsh)) +
                  tp.ww * ((i[1] - (ty << tp.tsh)) + (DIM == 3 ? tp.ww * (i[2] - (tz << tp.tsh)) : 0));
    }
    constexpr int NS = part_stride<NF>();
    R v[NC][NF];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if (slot[c] >= 0) {
            load_partial<NF, R>(tp.tb + ((long long)slot[c] * tp.wn + koff[c]) * NS, v[c]);
        } else {
#pragma unroll
            for (int f = 0; f < NF; ++f) v[c][f] = R(0);
        }
    }
    // lower tile indices first: tz, then ty, then tx ascending
#pragma unroll
    for (int c = NC - 1; c >= 0; --c)
        if (slot[c] >= 0) {
#pragma unroll
            for (int f = 0; f < NF; ++f) acc[f] += v[c][f];
        }
}

// ---------------------------------------------------------------------------
// Basis functions
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R sgn(R x) {
    return x > R(0) ? R(1) : (x < R(0) ? R(-1) : x);
}

// uGIMP weight and derivative (uGIMPbasis, mpbasisfunc.jl:40-53).  The FP64
// form keeps the reference's expression so parity runs agree to rounding;
// the FP32 form uses the reciprocal constants of uGIMPbasisx
// (mpbasisfunc.jl:56-76).
template <typename R>
__device__ __forceinline__ void ugimp(R d, R h, R lp, R c4, R c1, R half, R& N, R& dN) {
    R a = d < R(0) ? -d : d;
    N = R(0);
    dN = R(0);
    if constexpr (sizeof(R) == 8) {
        if (a < R(0.5) * lp) {
            N = R(1) - ((R(4) * d * d + lp * lp) / (R(4) * h * lp));
            dN = -((R(8) * d) / (R(4) * h * lp));
        } else if (R(0.5) * lp <= a && a < (h - R(0.5) * lp)) {
            N = R(1) - (a / h);
            dN = sgn(d) * (R(-1) / h);
        } else if ((h - R(0.5) * lp) <= a && a < (h + R(0.5) * lp)) {
            N = ((h + R(0.5) * lp - a) * (h + R(0.5) * lp - a)) / (R(2) * h * lp);
            dN = -sgn(d) * ((h + R(0.5) * lp - a) / (h * lp));
        }
    } else {
        if (a < half) {
            N = R(1) - (R(4) * d * d + lp * lp) * c4;
            dN = R(-8) * d * c4;
        } else if (a < h - half) {
            N = R(1) - a / h;
            dN = -sgn(d) / h;
        } else if (a < h + half) {
            R t = h + half - a;
            N = t * t * R(0.5) * c1;
            dN = -sgn(d) * t * c1;
        }
    }
}

// FP32 uGIMP weights of the three nodes of one axis, branch-free, from the
// offset d0 = x - x_base (in [h/2, 3h/2)): node k sits at d0 - k h.  The same
// piecewise function as ugimp (uGIMPbasisx, mpbasisfunc.jl:56-76) with every
// division by h replaced by a product with 1/h.
__device__ __forceinline__ void ugimp3_f32(float d0, float h, float hinv, float lp, float c4,
                                           float c1, float half, float* N, float* dN) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float d = d0 - float(k) * h;
        const float a = fabsf(d);
        const float t = (h + half) - a;
        const float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
        const float Nc = 1.0f - (4.0f * d * d + lp * lp) * c4;
        const float Nl = 1.0f - a * hinv;
        const float Nq = t * t * 0.5f * c1;
        const float dNc = -8.0f * d * c4;
        const float dNl = -sg * hinv;
        const float dNq = -sg * t * c1;
        const bool in_c = a < half, in_l = a < h - half, in_q = a < h + half;
        N[k] = in_c ? Nc : (in_l ? Nl : (in_q ? Nq : 0.0f));
        dN[k] = in_c ? dNc : (in_l ? dNl : (in_q ? dNq : 0.0f));
    }
}

// Per-axis stencil: S nodes starting at base[a].
template <int DIM, int S, typename R>
struct Stencil {
    int base[3];
    R N[3][S];
    R dN[3][S];
};

// Builds the stencil of a particle at x.  uGIMP (S = 3): the nodes with
// |x - x_n| < h + lp/2 lie in {floor(u - 1/2) + k, k = 0..2}, u = (x - x1)/h,
// for any lp <= h; the weight function itself returns zero outside the
// support, so the window may hold one zero-weight node.  Linear (S = 2): the
// two corners of the reference's cell (fld on x and z, cld on y,
// utils_OS.jl:124-127).  Returns false when the window leaves the grid or x is
// not finite.
// Lowest node of the S-node window on axis a, as a float (u = (x - x1) / h):
// floor(u - 1/2) for uGIMP; for the linear basis the reference's cell corner
// (fld on x and z, cld on y).
template <int S>
__device__ __forceinline__ double stencil_base_f(int a, double u) {
    if constexpr (S == 3) return floor(u - 0.5);
    else return (a == 1) ? ceil(u) - 1.0 : floor(u);
}

template <int DIM, int S, typename R>
__device__ __forceinline__ bool make_stencil(const GridGeom& g, const double* x,
                                             Stencil<DIM, S, R>& st, int& flag) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        double u = (x[a] - g.org[a]) * g.hinv[a];
        if (!isfinite(u)) {
            flag |= FLAG_NONFINITE;
            return false;
        }
        int b;
        const double fb = stencil_base_f<S>(a, u);
        if (fb < 0.0 || fb > double(g.nn[a] - S)) {
            ok = false;
            b = 0;
        } else {
            b = int(fb);
        }
        st.base[a] = b;
        if constexpr (S == 3 && sizeof(R) == 4) {
            // FP32: the offset from the base node once in FP64, the other two
            // nodes' offsets in FP32 (d0 - h is exact for d0 in [h/2, 3h/2)),
            // and the weights branch-free with the reciprocal constants
            const float d0 = float(x[a] - (g.org[a] + fb * g.h[a]));
            const float h = float(g.h[a]), hinv = float(g.hinv[a]);
            ugimp3_f32(d0, h, hinv, float(g.lp[a]), float(g.c4[a]), float(g.c1[a]),
                       float(g.half[a]), st.N[a], st.dN[a]);
            continue;
        }
#pragma unroll
        for (int k = 0; k < S; ++k) {
            double xn = g.org[a] + double(b + k) * g.h[a];
            R d = R(x[a] - xn);
            if constexpr (S == 3) {
                ugimp<R>(d, R(g.h[a]), R(g.lp[a]), R(g.c4[a]), R(g.c1[a]), R(g.half[a]),
                         st.N[a][k], st.dN[a][k]);
            } else {
                R hd = R(1) / R(g.h[a]);
                R ad = d < R(0) ? -d : d;
                st.N[a][k] = R(1) - ad * hd;
                st.dN[a][k] = -sgn(d) * hd;
            }
        }
    }
    if (!ok) flag |= FLAG_RANGE;
    return ok;
}

__device__ __forceinline__ long long node_id(const GridGeom& g, int ix, int iy, int iz) {
    return (long long)ix + (long long)g.nn[0] * ((long long)iy + (long long)g.nn[1] * iz);
}

template <typename R>
__device__ __forceinline__ void atomic_add(R* a, R v) {
    atomicAdd(a, v);  // global_atomic_add_f32 / _f64 (built with -munsafe-fp-atomics)
}

__device__ __forceinline__ void raise_flag(int* flags, int f) {
    if (f) atomicOr(flags, f);
}

// Iterate the DIM-dimensional stencil; body(n, N, dNx, dNy, dNz).  The weight
// products follow the reference's factor order (utils_OS.jl:351-354).
template <int DIM, int S, typename R, typename F>
__device__ __forceinline__ void for_stencil(const GridGeom& g, const Stencil<DIM, S, R>& st,
                                            F&& body) {
    if constexpr (DIM == 2) {
        static_for<S>([&](auto a) {
            static_for<S>([&](auto b) {
                R Nx = st.N[0][a], Ny = st.N[1][b];
                R N = Nx * Ny;
                R dNx = st.dN[0][a] * Ny;
                R dNy = st.dN[1][b] * Nx;
                long long n = node_id(g, st.base[0] + a, st.base[1] + b, 0);
                body(n, N, dNx, dNy, R(0));
            });
        });
    } else {
        static_for<S>([&](auto c) {
            static_for<S>([&](auto b) {
                static_for<S>([&](auto a) {
                    R Nx = st.N[0][a], Ny = st.N[1][b], Nz = st.N[2][c];
                    R N = Nx * Ny * Nz;
                    R dNx = st.dN[0][a] * Ny * Nz;
                    R dNy = st.dN[1][b] * Nx * Nz;
                    R dNz = st.dN[2][c] * Nx * Ny;
                    long long n = node_id(g, st.base[0] + a, st.base[1] + b, st.base[2] + c);
                    body(n, N, dNx, dNy, dNz);
                });
            });
        });
    }
}

// Same iteration as for_stencil, with the node given as an offset inside a
// local window (row stride WW, layer stride ZS) whose corner is the stencil base.
template <int DIM, int S, typename R, typename F>
__device__ __forceinline__ void for_stencil_local(const Stencil<DIM, S, R>& st, int WW, int ZS,
                                                  F&& body) {
    if constexpr (DIM == 2) {
        static_for<S>([&](auto a) {
            static_for<S>([&](auto b) {
                R Nx = st.N[0][a], Ny = st.N[1][b];
                body(b * WW + a, Nx * Ny, st.dN[0][a] * Ny, st.dN[1][b] * Nx, R(0));
            });
        });
    } else {
        static_for<S>([&](auto c) {
            static_for<S>([&](auto b) {
                static_for<S>([&](auto a) {
                    R Nx = st.N[0][a], Ny = st.N[1][b], Nz = st.N[2][c];
                    body(c * ZS + b * WW + a, Nx * Ny * Nz, st.dN[0][a] * Ny * Nz,
                         st.dN[1][b] * Nx * Nz, st.dN[2][c] * Nx * Ny);
                });
            });
        });
    }
}

template <int DIM>
__device__ __forceinline__ void load_x(const double* const* xs, int i, double* x) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) x[a] = xs[a][i];
}

// ---------------------------------------------------------------------------
// P2G #1 (P2G_OS!, utils_OS.jl:424-496)
// ---------------------------------------------------------------------------
template <int DIM, int S, typename R>
__global__ __launch_bounds__(256) void k_p2g(GridGeom g, PartPtrs<R> pp, GridPtrs<R> gp,
                                             Gravity<R> G, int np, int* flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double x[3];
    load_x<DIM>(pp.x, i, x);
    Stencil<DIM, S, R> st;
    int flag = 0;
    if (!make_stencil<DIM, S, R>(g, x, st, flag)) {
        raise_flag(flags, flag);
        return;
    }
    const R m = pp.ms[i], vol = pp.vol[i];
    R pv[3];
#pragma unroll
    for (int a = 0; a < DIM; ++a) pv[a] = m * pp.v[a][i];
    R s0 = pp.sig[0][i], s1 = pp.sig[1][i], s2 = R(0), s3 = pp.sig[3][i], s4 = R(0), s5 = R(0);
    if constexpr (DIM == 3) {
        s2 = pp.sig[2][i];
        s4 = pp.sig[4][i];
        s5 = pp.sig[5][i];
    }
    for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R dNx, R dNy, R dNz) {
        if (N != R(0)) {
            R NiM = m * N;
            atomic_add(&gp.m[n], NiM);
            atomic_add(&gp.p[0][n], N * pv[0]);
            atomic_add(&gp.p[1][n], N * pv[1]);
            if constexpr (DIM == 2) {
                atomic_add(&gp.f[0][n], -vol * (dNx * s0 + dNy * s3) + NiM * G.g[0]);
                atomic_add(&gp.f[1][n], -vol * (dNy * s1 + dNx * s3) + NiM * G.g[1]);
            } else {
                atomic_add(&gp.p[2][n], N * pv[2]);
                atomic_add(&gp.f[0][n], -vol * (dNx * s0 + dNy * s3 + dNz * s5) + NiM * G.g[0]);
                atomic_add(&gp.f[1][n], -vol * (dNy * s1 + dNx * s3 + dNz * s4) + NiM * G.g[1]);
                atomic_add(&gp.f[2][n], -vol * (dNz * s2 + dNx * s5 + dNy * s4) + NiM * G.g[2]);
            }
        }
    });
}

// ---------------------------------------------------------------------------
// Grid solve #1 (solvegrid_a_OS! utils_OS.jl:547-640; solvegrid_OS! 507-535,650-687)
// Leaves p = 0 (reference) and f = 0 (next step's P2G accumulates into it);
// clears the MVL cell sums (resetgridstatus_OS!, utils_OS.jl:82-99).
// ---------------------------------------------------------------------------
template <int DIM, typename R>
__global__ __launch_bounds__(256) void k_grid1(GridGeom g, GridPtrs<R> gp, StepArgs<R> sa,
                                               int ncells, TilePartials<R> tp) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < ncells && sa.mvl) {
        gp.cell_sm[n] = R(0);
        gp.cell_vol[n] = R(0);
    }
    if (n >= g.ni) return;
    // the grid fields hold what was added atomically (every P2G without tile
    // partials; with them, the particles that left their tile's window)
    R m = gp.m[n];
    R acc[1 + 2 * DIM];
#pragma unroll
    for (int f = 0; f < 1 + 2 * DIM; ++f) acc[f] = R(0);
    if (tp.tb) {
        add_partials<DIM, 1 + 2 * DIM, R>(g, tp, n, acc);
        m += acc[0];
    }
    // A node without mass received nothing from P2G (every particle with a
    // non-zero weight on it adds m_p N > 0), so p and f are still zero, and no
    // gather reads it (gathers skip zero weights): nothing to do.
    if (m == R(0)) return;
    if (tp.tb && sa.remap) gp.m[n] = m;  // for grid solve #2
    const R msd = m < R(kEpsF64) ? R(0) : R(1) / m;
    const unsigned bc = gp.bcf[n];
    R P[3], Fc[3], vs[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) P[d] = gp.p[d][n] + acc[1 + d];
    R ss = R(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        R fd = gp.f[d][n] + acc[1 + DIM + d];
        Fc[d] = fd;
        ss += fd * fd;
    }
    const R damp = -sa.zeta * sqrt(ss);
    R bv[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) bv[d] = (bc & (1u << d)) ? gp.bcv[d][n] : R(0);
    R vT[3];
    if (sa.va == 0) {
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            if (bc & (1u << d)) P[d] = bv[d];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            vs[d] = P[d] * msd;
            R Fs = Fc[d] + damp * sgn(vs[d]);
            vT[d] = vs[d] + Fs * sa.dt * msd;
            R ac = Fs * msd;
            if (bc & (1u << d)) {
                vT[d] = bv[d];
                ac = bv[d];
            }
            gp.vs[d][n] = vs[d];
            gp.vsT[d][n] = vT[d];
            gp.as[d][n] = ac;
        }
    } else {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            vs[d] = P[d] * msd;
            R Fs = Fc[d] + damp * sgn(vs[d]);
            vT[d] = (P[d] + Fs * sa.dt) * msd;
            if (bc & (1u << d)) vT[d] = bv[d];
            gp.vs[d][n] = vs[d];
            gp.vsT[d][n] = vT[d];
        }
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        gp.p[d][n] = R(0);
        gp.f[d][n] = R(0);
    }
    if (!sa.remap) {
        // USL/USF: Δus = v_new Δt (solvegrid(_a)_USL_OS!, OS_USL.jl:63-100,225-269);
        // nothing else reads this node's mass in the step
#pragma unroll
        for (int d = 0; d < DIM; ++d) gp.du[d][n] = vT[d] * sa.dt;
        gp.m[n] = R(0);
    }
}

// ---------------------------------------------------------------------------
// MUSL re-map (doublemapping1(_a)_OS! + doublemapping2_OS!, utils_OS.jl:698-959)
// ---------------------------------------------------------------------------
// One node's share of the re-map sums: A = Σ N vsT, and for :a B = Σ N as,
// C = Σ N vs; for :v B = Σ N (vsT - vs), C = Σ N vsT.  (USL/USF :a gather
// vsT in place of vs, see pic_field: doublemapping1_a_USL_OS! reads the
// already updated nodal velocity, OS_USL.jl:271-313.)
template <int DIM, typename R>
__device__ __forceinline__ void g2p1_accumulate(int va, R N, const R* vT, const R* vA, const R* vS,
                                                R* A, R* B, R* C) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        if (va == 0) {
            A[d] += N * vT[d];
            B[d] += N * vA[d];
            C[d] += N * vS[d];
        } else {
            A[d] += N * vT[d];
            B[d] += N * (vT[d] - vS[d]);
            C[d] += N * vT[d];
        }
    }
}

// New particle velocity and position increment; returns the new momentum.
template <int DIM, typename R>
__device__ __forceinline__ void g2p1_commit(PartPtrs<R>& pp, const StepArgs<R>& sa, int i,
                                            const R* A, const R* B, const R* C, R* pv, R m,
                                            const R* v) {
    // m and v are loaded by the caller before the gather: a load after a store
    // to another stream would wait for it
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        R vn = (sa.va == 0) ? sa.flip * v[d] + sa.pic * C[d] + B[d] * sa.dt
                            : sa.flip * (v[d] + B[d]) + sa.pic * C[d];
        pp.v[d][i] = vn;
        pp.dxp[d][i] = A[d] * sa.dt;
        pv[d] = m * vn;
    }
}

// The grid field whose gather feeds the re-map's C sum (see g2p1_accumulate).
template <typename R>
__device__ __forceinline__ R* const* pic_field(const GridPtrs<R>& gp, const StepArgs<R>& sa) {
    return sa.pic_new ? gp.vsT : gp.vs;
}

template <int DIM, int S, typename R>
__global__ __launch_bounds__(256) void k_g2p1_p2g2(GridGeom g, PartPtrs<R> pp, GridPtrs<R> gp,
                                                   StepArgs<R> sa, int np, int* flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double x[3];
    load_x<DIM>(pp.x, i, x);
    const R m0 = pp.ms[i];
    R v0[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) v0[d] = pp.v[d][i];
    Stencil<DIM, S, R> st;
    int flag = 0;
    if (!make_stencil<DIM, S, R>(g, x, st, flag)) {
#pragma unroll
        for (int a = 0; a < DIM; ++a) pp.dxp[a][i] = R(0);
        return;  // already reported by k_p2g
    }
    R A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, C[3] = {0, 0, 0};
    R* const* vsf = pic_field(gp, sa);
    const bool need_vs = sa.va != 0 || sa.pic != R(0);  // :a with PIC = 0 multiplies it by 0
    for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R, R, R) {
        if (N != R(0)) {
            R vT[3], vA[3], vS[3];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                vT[d] = gp.vsT[d][n];
                vA[d] = gp.as[d][n];
                vS[d] = need_vs ? vsf[d][n] : R(0);
            }
            g2p1_accumulate<DIM, R>(sa.va, N, vT, vA, vS, A, B, C);
        }
    });
    R pv[3];
    g2p1_commit<DIM, R>(pp, sa, i, A, B, C, pv, m0, v0);
    if (!sa.remap) return;  // USL/USF: no momentum re-map
    for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R, R, R) {
        if (N != R(0)) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) atomic_add(&gp.p[d][n], pv[d] * N);
        }
    });
}

// ---------------------------------------------------------------------------
// Grid solve #2 (doublemapping3_OS!, utils_OS.jl:968-1016); clears m and p.
// ---------------------------------------------------------------------------
template <int DIM, typename R>
__global__ __launch_bounds__(256) void k_grid2(GridGeom g, GridPtrs<R> gp, R dt, TilePartials<R> tp) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.ni) return;
    const R m = gp.m[n];
    if (m == R(0)) return;  // no mass: p is zero and no G2P weight reads du here
    const R msd = m < R(kEpsF64) ? R(0) : R(1) / m;
    const unsigned bc = gp.bcf[n];
    R acc[3] = {0, 0, 0};
    if (tp.tb) add_partials<DIM, DIM, R>(g, tp, n, acc);
    // loads before stores (a load after a store to another field waits for it)
    R P[3], bv[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        P[d] = gp.p[d][n] + acc[d];
        bv[d] = (bc & (1u << d)) ? gp.bcv[d][n] : R(0);
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        R v = P[d] * msd;
        if (bc & (1u << d)) v = bv[d];
        gp.du[d][n] = v * dt;
        gp.p[d][n] = R(0);
    }
    gp.m[n] = R(0);
}

// ---------------------------------------------------------------------------
// Constitutive updates on registers
// ---------------------------------------------------------------------------
template <int NS, typename R>
__device__ __forceinline__ R mean3(const R* s) {
    return (s[0] + s[1] + s[2]) * R(0.333333);
}

// liE! (linearelastic.jl:22-69 2D, 78-140 3D).  F holds ΔF row-major
// (2D: [dF1 dF2; dF3 dF4], 3D: 9 entries), e the strain increment.
template <int DIM, typename R>
__device__ __forceinline__ void liE(R* s, const R* F, const R* e, const Material<R>& M) {
    if constexpr (DIM == 2) {
        R wxy = R(0.5) * (F[1] - F[2]);
        R s1 = s[0], s2 = s[1], s4 = s[3];
        s[0] += wxy * s4 * R(2);
        s[1] += -wxy * s4 * R(2);
        s[3] += wxy * (s2 - s1);
        s[0] += M.Dt * e[0] + M.Dd * e[1] + M.Dd * e[2];
        s[1] += M.Dd * e[0] + M.Dt * e[1] + M.Dd * e[2];
        s[2] += M.Dd * e[0] + M.Dd * e[1] + M.Dt * e[2];
        s[3] += M.Gs * e[3];
    } else {
        R wxy = R(0.5) * (F[3] - F[1]);
        R wyz = R(0.5) * (F[7] - F[5]);
        R wxz = R(0.5) * (F[6] - F[2]);
        R s1 = s[0], s2 = s[1], s3 = s[2], s4 = s[3], s5 = s[4], s6 = s[5];
        s[0] += R(2) * (s4 * wxy + s6 * wxz);
        s[1] += -R(2) * (s4 * wxy - s5 * wyz);
        s[2] += -R(2) * (s6 * wxz + s5 * wyz);
        s[3] += wxy * (s2 - s1) + wxz * s5 + wyz * s6;
        s[4] += wyz * (s3 - s2) - wxz * s4 - wxy * s6;
        s[5] += wxz * (s3 - s1) + wxy * s5 - wyz * s4;
        s[0] += M.Dt * e[0] + M.Dd * e[1] + M.Dd * e[2];
        s[1] += M.Dd * e[0] + M.Dt * e[1] + M.Dd * e[2];
        s[2] += M.Dd * e[0] + M.Dd * e[1] + M.Dt * e[2];
        s[3] += M.Gs * e[3];
        s[4] += M.Gs * e[4];
        s[5] += M.Gs * e[5];
    }
}

// dpP! (druckerprager.jl:22-80 2D, 89-155 3D).
template <int DIM, typename R>
__device__ __forceinline__ void dpP(R* s, R& eq, R& ek, const Material<R>& M) {
    constexpr int NS = DIM == 2 ? 4 : 6;
    R sm = mean3<NS>(s);
    R sij[NS];
    sij[0] = s[0] - sm;
    sij[1] = s[1] - sm;
    sij[2] = s[2] - sm;
#pragma unroll
    for (int k = 3; k < NS; ++k) sij[k] = s[k];
    R q = R(0.5) * (sij[0] * sij[0] + sij[1] * sij[1] + sij[2] * sij[2]);
#pragma unroll
    for (int k = 3; k < NS; ++k) q += sij[k] * sij[k];
    const R tau = sqrt(q);
    const R st = M.sigt;
    const R fs = tau + M.qphi * sm - M.kphi;
    const R ft = sm - st;
    const R BF = (tau - M.tb) - (M.ab * (sm - st));
    if (((sm < st) && (fs > R(0))) || ((sm >= st) && (BF > R(0)))) {
        R dls = fs / (M.Gs + M.Ks * M.qphi * M.qpsi);
        R t1 = sm - M.Ks * M.qpsi * dls;
        R t2 = (M.kphi - M.qphi * t1) / tau;
        s[0] = sij[0] * t2 + t1;
        s[1] = sij[1] * t2 + t1;
        s[2] = sij[2] * t2 + t1;
#pragma unroll
        for (int k = 3; k < NS; ++k) s[k] = sij[k] * t2;
        eq += dls * M.eq_shear;
        ek += dls * M.qpsi;
    }
    if ((sm >= st) && (BF <= R(0))) {
        R dlt = ft / M.Ks;
        s[0] = sij[0] + st;
        s[1] = sij[1] + st;
        s[2] = sij[2] + st;
        eq += dlt * R(0.333333) * R(1.414214);
        ek += dlt;
    }
}

// mcP! 2D plane strain (mohrcoulomb.jl:22-86).  3D: mcP3 below.
template <typename R>
__device__ __forceinline__ void mcP2(R* s, R& eq, R& ek, const Material<R>& M) {
    R c = M.c + M.Hp * eq;
    c = (c > M.cr || c != c) ? c : M.cr;  // max(c + Hp ϵq, cr)
    const R ds = s[0] - s[1];
    const R tau = sqrt(R(0.25) * (ds * ds) + s[3] * s[3]);
    const R sg = R(0.5) * (s[0] + s[1]);
    const R f = tau + sg * M.sinphi - c * M.cosphi;
    R sn1 = s[0], sn2 = s[1], sn3 = s[2], sn4 = s[3];
    R bb = c * M.cosphi - sg * M.sinphi;
    const R beta = (bb < R(0) ? -bb : bb) / tau;
    const R dsigA = R(0.5) * beta * ds;
    const R dsigB = c / M.tanphi;
    if ((sg <= dsigB) && (f > R(0))) {
        sn1 = sg + dsigA;
        sn2 = sg - dsigA;
        sn4 = beta * s[3];
    }
    if ((sg > dsigB) && (f > R(0))) {
        sn1 = dsigB;
        sn2 = dsigB;
        sn4 = R(0);
    }
    const R d1 = sn1 - s[0], d2 = sn2 - s[1], d3 = sn3 - s[2], d4 = sn4 - s[3];
    s[0] = sn1;
    s[1] = sn2;
    s[2] = sn3;
    s[3] = sn4;
    const R Dt = M.Dt, Dd = M.Dd;
    const R base = R(1) / ((Dd - Dt) * (R(2) * Dd + Dt));
    const R exx = -(Dd * d1 + Dt * d1 - Dd * d2 - Dd * d3) * base;
    const R eyy = -(-Dd * d1 + Dd * d2 + Dt * d2 - Dd * d3) * base;
    const R ezz = -(-Dd * d1 - Dd * d2 + Dd * d3 + Dt * d3) * base;
    const R exy = d4 / M.Gs;
    ek = exx + eyy + ezz;
    eq += sqrt(R(0.666667) * (exx * exx + eyy * eyy + ezz * ezz + exy * exy * R(2)));
}

// Symmetric 3x3 eigen-decomposition of [xx,yy,zz,xy,yz,xz] by cyclic Jacobi
// rotations (Golub & Van Loan, Matrix Computations, sec. 8.5).  lam: descending
// eigenvalues; V[r][k]: component r of eigenvector k.
template <typename R>
__device__ __forceinline__ void sym3_eig(const R* s, R* lam, R (*V)[3]) {
    R A[3][3] = {{s[0], s[3], s[5]}, {s[3], s[1], s[4]}, {s[5], s[4], s[2]}};
    R Q[3][3] = {{R(1), R(0), R(0)}, {R(0), R(1), R(0)}, {R(0), R(0), R(1)}};
    const R tol = sizeof(R) == 4 ? R(1e-13) : R(1e-30);
    for (int sweep = 0; sweep < 10; ++sweep) {
        const R off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const R dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= tol * dia || off == R(0)) break;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
#pragma unroll
            for (int q = p + 1; q < 3; ++q) {
                const R apq = A[p][q];
                if (apq == R(0)) continue;
                const R th = (A[q][q] - A[p][p]) / (R(2) * apq);
                const R t = (th >= R(0) ? R(1) : R(-1)) / (fabs(th) + sqrt(th * th + R(1)));
                const R c = R(1) / sqrt(t * t + R(1)), sn = t * c;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const R akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - sn * akq;
                    A[k][q] = sn * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const R apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - sn * aqk;
                    A[q][k] = sn * apk + c * aqk;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const R qkp = Q[k][p], qkq = Q[k][q];
                    Q[k][p] = c * qkp - sn * qkq;
                    Q[k][q] = sn * qkp + c * qkq;
                }
            }
        }
    }
    int o0 = 0, o1 = 1, o2 = 2;
    const R d[3] = {A[0][0], A[1][1], A[2][2]};
    if (d[o0] < d[o1]) { int t = o0; o0 = o1; o1 = t; }
    if (d[o1] < d[o2]) { int t = o1; o1 = o2; o2 = t; }
    if (d[o0] < d[o1]) { int t = o0; o0 = o1; o1 = t; }
    lam[0] = d[o0];
    lam[1] = d[o1];
    lam[2] = d[o2];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        V[r][0] = Q[r][o0];
        V[r][1] = Q[r][o1];
        V[r][2] = Q[r][o2];
    }
}

// Principal-space Mohr-Coulomb return of p (descending), perfect plasticity
// with dilation psi (de Souza Neto, Peric & Owen, Computational Methods for
// Plasticity, 2008, sec. 8.2): main plane, then the edge whose ordering
// constraint the main-plane return crosses first, then the apex c cot(phi).
// Planes f = (s1 - s3) + (s1 + s3) sin(phi) - 2 c cos(phi); flow direction
// (1 + sin(psi), 0, -(1 - sin(psi))).  Returns 0 when elastic.
template <typename R>
__device__ __forceinline__ int mc3_principal(R* p, R Dt, R Dd, R sphi, R cphi, R spsi, R c) {
    const R k2 = R(2) * c * cphi;
    const R f = (p[0] - p[2]) + (p[0] + p[2]) * sphi - k2;
    if (!(f > R(0))) return 0;
    auto dmul = [&](const R* n, R* out) {
        out[0] = Dt * n[0] + Dd * n[1] + Dd * n[2];
        out[1] = Dd * n[0] + Dt * n[1] + Dd * n[2];
        out[2] = Dd * n[0] + Dd * n[1] + Dt * n[2];
    };
    auto dot = [](const R* a, const R* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    const R Ga[3] = {R(1) + sphi, R(0), -(R(1) - sphi)};
    const R Na[3] = {R(1) + spsi, R(0), -(R(1) - spsi)};
    R DNa[3], DNb[3], q[3];
    dmul(Na, DNa);
    const R a11 = dot(Ga, DNa);
    const R g = f / a11;
#pragma unroll
    for (int k = 0; k < 3; ++k) q[k] = p[k] - g * DNa[k];
    if (q[0] >= q[1] && q[1] >= q[2]) {
        p[0] = q[0];
        p[1] = q[1];
        p[2] = q[2];
        return 1;
    }
    const bool edge12 = (R(1) - spsi) * p[0] - R(2) * p[1] + (R(1) + spsi) * p[2] < R(0);
    R Gb[3], Nb[3], fb;
    if (edge12) {  // s1 = s2: second plane (s2 - s3)
        Gb[0] = R(0); Gb[1] = R(1) + sphi; Gb[2] = -(R(1) - sphi);
        Nb[0] = R(0); Nb[1] = R(1) + spsi; Nb[2] = -(R(1) - spsi);
        fb = (p[1] - p[2]) + (p[1] + p[2]) * sphi - k2;
    } else {       // s2 = s3: second plane (s1 - s2)
        Gb[0] = R(1) + sphi; Gb[1] = -(R(1) - sphi); Gb[2] = R(0);
        Nb[0] = R(1) + spsi; Nb[1] = -(R(1) - spsi); Nb[2] = R(0);
        fb = (p[0] - p[1]) + (p[0] + p[1]) * sphi - k2;
    }
    dmul(Nb, DNb);
    const R a12 = dot(Ga, DNb), a21 = dot(Gb, DNa), a22 = dot(Gb, DNb);
    const R det = a11 * a22 - a12 * a21;
    if (det > R(0)) {
        const R ga = (a22 * f - a12 * fb) / det;
        const R gb = (a11 * fb - a21 * f) / det;
#pragma unroll
        for (int k = 0; k < 3; ++k) q[k] = p[k] - ga * DNa[k] - gb * DNb[k];
        if ((ga >= R(0) && gb >= R(0) && q[0] >= q[2]) || !(sphi > R(0))) {
            p[0] = q[0];
            p[1] = q[1];
            p[2] = q[2];
            return edge12 ? 2 : 3;
        }
    }
    const R apex = c * cphi / sphi;
    p[0] = p[1] = p[2] = apex;
    return 4;
}

// 3D Mohr-Coulomb, defined by this library: the reference's 3D mcP! body is
// commented out (mohrcoulomb.jl:152-455).  Conventions of its 2D mcP!
// (mohrcoulomb.jl:22-86): cohesion softening max(c + Hp ϵq, cr), plastic
// strain from the stress correction through the inverse elastic matrix, ϵk
// assigned, ϵq accumulated.  s = [xx,yy,zz,xy,yz,xz].
template <typename R>
__device__ __forceinline__ R mc_cohesion(const Material<R>& M, R eq) {
    const R c = M.c + M.Hp * eq;
    return (c > M.cr || c != c) ? c : M.cr;  // max(c + Hp ϵq, cr)
}

// Strength in effect at plastic strain eq (3D): the linear softening rule of
// mohr_coulomb.tcc:8-25 with one peak/residual strain pair for cohesion,
// friction and dilation, peak at eq = 0.  Cohesion follows max(c + Hp ϵq, cr);
// when Hp < 0, c > cr and ϕr > 0 (eqr = (cr - c) / Hp > 0, set at upload)
// friction goes linearly from ϕ to ϕr and dilation from ψ to 0 over [0, eqr].
template <typename R>
struct MCStrength {
    R c, phi, psi, sinphi, cosphi, sinpsi;
};

template <typename R>
__device__ __forceinline__ MCStrength<R> mc_strength(const Material<R>& M, R eq) {
    MCStrength<R> p;
    p.c = mc_cohesion<R>(M, eq);
    if (M.eqr > R(0)) {
        const R t = eq >= M.eqr ? R(1) : eq / M.eqr;
        p.phi = M.phi + (M.phir - M.phi) * t;
        p.psi = M.psi - M.psi * t;
        p.sinphi = sin(p.phi);
        p.cosphi = cos(p.phi);
        p.sinpsi = sin(p.psi);
    } else {
        p.phi = M.phi;
        p.psi = M.psi;
        p.sinphi = M.sinphi;
        p.cosphi = M.cosphi;
        p.sinpsi = M.sinpsi;
    }
    return p;
}

// Rotation (eigenvectors in the columns of V, any signs) <-> unit quaternion
// (x, y, z; w >= 0 is implied).  A reflection is made a rotation by turning
// the last column, which leaves V diag(p) V^T unchanged.
template <typename R>
__device__ __forceinline__ void quat_from_eigvecs(R (*V)[3], R* q3) {
    const R det = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1]) -
                  V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0]) +
                  V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
    if (det < R(0))
#pragma unroll
        for (int r = 0; r < 3; ++r) V[r][2] = -V[r][2];
    const R tr = V[0][0] + V[1][1] + V[2][2];
    R w, x, y, z;
    if (tr > R(0)) {
        const R s = sqrt(tr + R(1)) * R(2);
        w = R(0.25) * s;
        x = (V[2][1] - V[1][2]) / s;
        y = (V[0][2] - V[2][0]) / s;
        z = (V[1][0] - V[0][1]) / s;
    } else if (V[0][0] > V[1][1] && V[0][0] > V[2][2]) {
        const R s = sqrt(R(1) + V[0][0] - V[1][1] - V[2][2]) * R(2);
        w = (V[2][1] - V[1][2]) / s;
        x = R(0.25) * s;
        y = (V[0][1] + V[1][0]) / s;
        z = (V[0][2] + V[2][0]) / s;
    } else if (V[1][1] > V[2][2]) {
        const R s = sqrt(R(1) + V[1][1] - V[0][0] - V[2][2]) * R(2);
        w = (V[0][2] - V[2][0]) / s;
        x = (V[0][1] + V[1][0]) / s;
        y = R(0.25) * s;
        z = (V[1][2] + V[2][1]) / s;
    } else {
        const R s = sqrt(R(1) + V[2][2] - V[0][0] - V[1][1]) * R(2);
        w = (V[1][0] - V[0][1]) / s;
        x = (V[0][2] + V[2][0]) / s;
        y = (V[1][2] + V[2][1]) / s;
        z = R(0.25) * s;
    }
    const R sg = w < R(0) ? R(-1) : R(1);
    q3[0] = sg * x;
    q3[1] = sg * y;
    q3[2] = sg * z;
}

template <typename R>
__device__ __forceinline__ void eigvecs_from_quat(const R* q3, R (*V)[3]) {
    const R x = q3[0], y = q3[1], z = q3[2];
    const R ww = R(1) - (x * x + y * y + z * z);
    const R w = ww > R(0) ? sqrt(ww) : R(0);
    V[0][0] = R(1) - R(2) * (y * y + z * z);
    V[0][1] = R(2) * (x * y - z * w);
    V[0][2] = R(2) * (x * z + y * w);
    V[1][0] = R(2) * (x * y + z * w);
    V[1][1] = R(1) - R(2) * (x * x + z * z);
    V[1][2] = R(2) * (y * z - x * w);
    V[2][0] = R(2) * (x * z - y * w);
    V[2][1] = R(2) * (y * z + x * w);
    V[2][2] = R(1) - R(2) * (x * x + y * y);
}

// Inputs of the learned 3D Mohr-Coulomb return (16): the trial stress as its
// principal values p (descending) over the scale S = max|p| + c cos(phi) and
// its principal frame as a quaternion (3), the strain increment (Voigt,
// engineering shears) times 2G / S, the accumulated plastic strain, and the
// strength in effect (phi, c / S, psi).  Scale-free where the return is.
template <typename R>
__device__ __forceinline__ void mc_features(const R* p, const R* q3, const R* de, R eq,
                                            const MCStrength<R>& P, R Gs, float* x) {
    const R ap0 = p[0] < R(0) ? -p[0] : p[0], ap2 = p[2] < R(0) ? -p[2] : p[2];
    const R S = (ap0 > ap2 ? ap0 : ap2) + P.c * P.cosphi + R(1e-30);
    const R iS = R(1) / S;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = float(p[k] * iS);
        x[3 + k] = float(q3[k]);
    }
    const R ge = R(2) * Gs * iS;
#pragma unroll
    for (int k = 0; k < 6; ++k) x[6 + k] = float(de[k] * ge);
    x[12] = float(eq);
    x[13] = float(P.phi);
    x[14] = float(P.c * iS);
    x[15] = float(P.psi);
}

// The principal-space return on one given branch (1 main plane, 2 edge
// s1 = s2, 3 edge s2 = s3, 4 apex) of a yielding trial stress p, accepted
// only when that branch's own conditions hold: the main-plane return keeps
// the principal order; an edge return has non-negative multipliers of both
// planes and keeps s1 >= s3 (mc3_principal's edge test); the apex is taken
// only where mc3_principal itself ends at the apex.  Returns true with p the
// returned stresses, false with p unchanged.  The closed forms are
// mc3_principal's, so a main-plane or apex branch that is accepted gives its
// stresses exactly; an accepted edge is the one whose conditions hold.
template <typename R>
__device__ __forceinline__ bool mc3_principal_branch(R* p, int kind, R Dt, R Dd, R sphi, R cphi,
                                                     R spsi, R c) {
    const R k2 = R(2) * c * cphi;
    auto dmul = [&](const R* n, R* out) {
        out[0] = Dt * n[0] + Dd * n[1] + Dd * n[2];
        out[1] = Dd * n[0] + Dt * n[1] + Dd * n[2];
        out[2] = Dd * n[0] + Dd * n[1] + Dt * n[2];
    };
    auto dot = [](const R* a, const R* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    const R f = (p[0] - p[2]) + (p[0] + p[2]) * sphi - k2;
    const R Ga[3] = {R(1) + sphi, R(0), -(R(1) - sphi)};
    const R Na[3] = {R(1) + spsi, R(0), -(R(1) - spsi)};
    R DNa[3], q[3];
    dmul(Na, DNa);
    const R a11 = dot(Ga, DNa);
    if (kind == 1) {
        const R g = f / a11;
#pragma unroll
        for (int k = 0; k < 3; ++k) q[k] = p[k] - g * DNa[k];
        if (!(q[0] >= q[1] && q[1] >= q[2])) return false;
    } else if (kind == 2 || kind == 3) {
        R Gb[3], Nb[3], DNb[3], fb;
        if (kind == 2) {  // s1 = s2: second plane (s2 - s3)
            Gb[0] = R(0); Gb[1] = R(1) + sphi; Gb[2] = -(R(1) - sphi);
            Nb[0] = R(0); Nb[1] = R(1) + spsi; Nb[2] = -(R(1) - spsi);
            fb = (p[1] - p[2]) + (p[1] + p[2]) * sphi - k2;
        } else {          // s2 = s3: second plane (s1 - s2)
            Gb[0] = R(1) + sphi; Gb[1] = -(R(1) - sphi); Gb[2] = R(0);
            Nb[0] = R(1) + spsi; Nb[1] = -(R(1) - spsi); Nb[2] = R(0);
            fb = (p[0] - p[1]) + (p[0] + p[1]) * sphi - k2;
        }
        dmul(Nb, DNb);
        const R a12 = dot(Ga, DNb), a21 = dot(Gb, DNa), a22 = dot(Gb, DNb);
        const R det = a11 * a22 - a12 * a21;
        if (!(det > R(0))) return false;
        const R ga = (a22 * f - a12 * fb) / det;
        const R gb = (a11 * fb - a21 * f) / det;
#pragma unroll
        for (int k = 0; k < 3; ++k) q[k] = p[k] - ga * DNa[k] - gb * DNb[k];
        if (!((ga >= R(0) && gb >= R(0) && q[0] >= q[2]) || !(sphi > R(0)))) return false;
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) q[k] = p[k];
        if (mc3_principal<R>(q, Dt, Dd, sphi, cphi, spsi, c) != 4) return false;
    }
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
    return true;
}

// Plastic strain bookkeeping of a stress correction d = s_new - s_trial
// (inverse elastic matrix; ϵk assigned, ϵq accumulated).
template <typename R>
__device__ __forceinline__ void mc_strain(const R* d, R& eq, R& ek, const Material<R>& M) {
    const R Dt = M.Dt, Dd = M.Dd;
    const R base = R(1) / ((Dd - Dt) * (R(2) * Dd + Dt));
    const R exx = -(Dd * d[0] + Dt * d[0] - Dd * d[1] - Dd * d[2]) * base;
    const R eyy = -(-Dd * d[0] + Dd * d[1] + Dt * d[1] - Dd * d[2]) * base;
    const R ezz = -(-Dd * d[0] - Dd * d[1] + Dd * d[2] + Dt * d[2]) * base;
    const R exy = d[3] / M.Gs, eyz = d[4] / M.Gs, exz = d[5] / M.Gs;
    ek = exx + eyy + ezz;
    eq += sqrt(R(0.666667) * (exx * exx + eyy * eyy + ezz * ezz +
                              R(2) * (exy * exy + eyz * eyz + exz * exz)));
}

// Returns the kind of return (0 elastic, 1 main plane, 2/3 edge, 4 apex).
template <typename R>
__device__ __forceinline__ int mcP3(R* s, R& eq, R& ek, const Material<R>& M) {
    const MCStrength<R> P = mc_strength<R>(M, eq);
    R lam[3], V[3][3];
    sym3_eig<R>(s, lam, V);
    const int kind = mc3_principal<R>(lam, M.Dt, M.Dd, P.sinphi, P.cosphi, P.sinpsi, P.c);
    if (kind == 0) {
        ek = R(0);
        return 0;
    }
    const int ri[6] = {0, 1, 2, 0, 1, 0}, ci[6] = {0, 1, 2, 1, 2, 2};
    R d[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const R sn = lam[0] * V[ri[k]][0] * V[ci[k]][0] + lam[1] * V[ri[k]][1] * V[ci[k]][1] +
                     lam[2] * V[ri[k]][2] * V[ci[k]][2];
        d[k] = sn - s[k];
        s[k] = sn;
    }
    mc_strain<R>(d, eq, ek, M);
    return kind;
}

// Appends one return-mapping sample: the learned return's 16 inputs
// (mc_features of the trial stress s0, strain increment e, plastic strain eq0
// and the strength in effect) and its 7 targets: the branch the exact return
// took as a one-hot over (main plane, edge s1 = s2, edge s2 = s3, apex) and
// three zeros.
template <typename R>
__device__ __forceinline__ void log_return(const StepArgs<R>& sa, const R* s0, const R* e, R eq0,
                                           int kind, const Material<R>& M) {
    const unsigned long long slot = atomicAdd(sa.log_count, 1ull);
    if ((long long)slot >= sa.log_cap) return;
    float* x = sa.log_x + slot * 16;
    float* y = sa.log_y + slot * 7;
    R lam[3], V[3][3], q3[3];
    sym3_eig<R>(s0, lam, V);
    quat_from_eigvecs<R>(V, q3);
    const MCStrength<R> P = mc_strength<R>(M, eq0);
    mc_features<R>(lam, q3, e, eq0, P, M.Gs, x);
#pragma unroll
    for (int k = 0; k < 7; ++k) y[k] = (k == kind - 1) ? 1.0f : 0.0f;
}

// Cell id of the reference's p2c (fld on x and z, cld on y), 0-based, in this
// library's x-fastest cell order.
template <int DIM>
__device__ __forceinline__ int cell_of(const GridGeom& g, const double* x) {
    int cx = int(floor((x[0] - g.org[0]) * g.hinv[0]));
    int cy = int(ceil((x[1] - g.org[1]) * g.hinv[1])) - 1;
    int cz = DIM == 3 ? int(floor((x[2] - g.org[2]) * g.hinv[2])) : 0;
    return cx + g.nc[0] * (cy + g.nc[1] * cz);
}

// Everything G2P #2 does after gathering ΔF (row-major, DIM x DIM): strain
// rate and increment, J, Ω, the constitutive update, MVL's cell sums and the
// position commit (G2P_OS!, utils_OS.jl:1025-1153; liE!; dpP!; mcP!;
// vollock1_OS!).  x is the step-start position.  CM is the constitutive
// model (0 linear elastic, 1 Drucker-Prager, 2 Mohr-Coulomb), a template
// parameter so that each kernel holds the registers of one model only.
// The particle state G2P #2's commit reads.  Every load of the commit is
// issued before any of its stores: the particle streams are separate arrays,
// but the compiler cannot know they do not alias, so a load placed after a
// store waits for it.  Where registers allow, callers load it before the
// gather, so the loads are in flight while the stencil is summed.
template <int DIM, typename R>
struct CommitIn {
    static constexpr int NS = DIM == 2 ? 4 : 6;
    R eps0[NS], s[NS];
    double J0;
    R vol0, eq, ek, dxp[3];
    int nid;
};

// Whether a kernel loads the commit state before its gather: FP32 Drucker-
// Prager / linear-elastic keep 3 (tiled) wavefronts per SIMD with the state
// live across the gather (166 registers); the 3D Mohr-Coulomb and FP64 forms
// would lose a wavefront per SIMD, so they load it after the gather.
template <typename R, int CM>
constexpr bool kCommitLoadsEarly = sizeof(R) == 4 && CM != 2;

template <int DIM, typename R, int CM>
__device__ __forceinline__ void load_commit_in(const PartPtrs<R>& pp, const StepArgs<R>& sa, int i,
                                               CommitIn<DIM, R>& c) {
    constexpr int NS = DIM == 2 ? 4 : 6;
#pragma unroll
    for (int k = 0; k < NS; ++k) c.eps0[k] = (DIM == 2 && k == 2) ? R(0) : pp.eps[k][i];
#pragma unroll
    for (int k = 0; k < NS; ++k) c.s[k] = pp.sig[k][i];
    c.J0 = pp.J[i];
    c.vol0 = pp.vol0[i];
    c.nid = pp.nid[i];
    c.eq = R(0);
    c.ek = R(0);
    if (CM != 0 && sa.plastic) {
        c.eq = pp.eq[i];
        c.ek = pp.ek[i];
    }
    c.dxp[0] = c.dxp[1] = c.dxp[2] = R(0);
    if (sa.commit_dx) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) c.dxp[d] = pp.dxp[d][i];
    }
}

template <int DIM, typename R, int CM>
__device__ __forceinline__ void g2p2_commit(const GridGeom& g, PartPtrs<R>& pp, GridPtrs<R>& gp,
                                            const Material<R>* __restrict__ mats,
                                            const StepArgs<R>& sa, int i, const double* x, R* F,
                                            const CommitIn<DIM, R>& cin) {
    constexpr int NS = DIM == 2 ? 4 : 6;
    R eps0[NS], s[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        eps0[k] = cin.eps0[k];
        s[k] = cin.s[k];
    }
    const double J0 = cin.J0;
    const R vol0 = cin.vol0;
    R eq = cin.eq, ek = cin.ek;
    const R* dxp = cin.dxp;
    const Material<R> M = mats[cin.nid];

    const R rdt = R(1) / sa.dt;
    R e[NS];
    double detI;
    if constexpr (DIM == 2) {
        R exx = F[0] * rdt, eyy = F[3] * rdt, exy = R(0.5) * (F[1] + F[2]) * rdt;
        pp.ev[i] = sqrt(exx * exx + eyy * eyy + R(2) * exy * exy);
        e[0] = F[0];
        e[1] = F[3];
        e[2] = R(0);
        e[3] = F[1] + F[2];
        pp.eps[0][i] = eps0[0] + e[0];
        pp.eps[1][i] = eps0[1] + e[1];
        pp.eps[3][i] = eps0[3] + e[3];
        double a = 1.0 + double(F[0]), b = double(F[1]), c = double(F[2]), d = 1.0 + double(F[3]);
        detI = a * d - b * c;
    } else {
        R exx = F[0] * rdt, eyy = F[4] * rdt, ezz = F[8] * rdt;
        R exy = R(0.5) * (F[1] + F[3]) * rdt;
        R eyz = R(0.5) * (F[5] + F[7]) * rdt;
        R exz = R(0.5) * (F[2] + F[6]) * rdt;
        pp.ev[i] = sqrt(exx * exx + eyy * eyy + ezz * ezz + R(2) * (exy * exy + eyz * eyz + exz * exz));
        e[0] = F[0];
        e[1] = F[4];
        e[2] = F[8];
        e[3] = F[1] + F[3];
        e[4] = F[5] + F[7];
        e[5] = F[2] + F[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) pp.eps[k][i] = eps0[k] + e[k];
        double A[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) A[k] = double(F[k]);
        A[0] += 1.0;
        A[4] += 1.0;
        A[8] += 1.0;
        detI = A[0] * A[4] * A[8] + A[1] * A[5] * A[6] + A[2] * A[3] * A[7] -
               A[6] * A[4] * A[2] - A[7] * A[5] * A[0] - A[8] * A[3] * A[1];
    }
    const double J = J0 * detI;
    pp.J[i] = J;
    const R vol = R(J * double(vol0));
    pp.vol[i] = vol;

    liE<DIM, R>(s, F, e, M);
    if (CM != 0 && sa.plastic) {
        if constexpr (CM == 1) {
            dpP<DIM, R>(s, eq, ek, M);
        } else if constexpr (DIM == 2) {
            mcP2<R>(s, eq, ek, M);
        } else {
            if (sa.defer_return) {
                // trial stress (stored below), strain increment and the trial
                // stress's principal values and frame: the learned-return
                // kernel finishes without an eigen-solve of its own
#pragma unroll
                for (int k = 0; k < 6; ++k) pp.de[k][i] = e[k];
                R lam[3], V[3][3], q3[3];
                sym3_eig<R>(s, lam, V);
                quat_from_eigvecs<R>(V, q3);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    pp.pr[k][i] = lam[k];
                    pp.pr[3 + k][i] = q3[k];
                }
            } else if (sa.log_x && i % sa.log_stride == 0) {
                R s0[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) s0[k] = s[k];
                const R eq0 = eq;
                const int kind = mcP3<R>(s, eq, ek, M);
                if (kind != 0) log_return<R>(sa, s0, e, eq0, kind, M);
            } else {
                mcP3<R>(s, eq, ek, M);
            }
        }
        if (!(DIM == 3 && CM == 2) || !sa.defer_return) {
            pp.eq[i] = eq;
            pp.ek[i] = ek;
        }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) pp.sig[k][i] = s[k];
    if (sa.mvl) {
        int c = cell_of<DIM>(g, x);
        pp.p2c[i] = c;
        R sm = mean3<NS>(s);
        atomic_add(&gp.cell_sm[c], vol * sm);
        atomic_add(&gp.cell_vol[c], vol);
    }
    if (sa.commit_dx) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) pp.x[d][i] = x[d] + double(dxp[d]);
    }
}

// ---------------------------------------------------------------------------
// G2P #2 + constitutive + commit (G2P_OS!, utils_OS.jl:1025-1153; liE!; dpP!;
// mcP!; vollock1_OS!)
// ---------------------------------------------------------------------------
template <int DIM, int S, typename R, int CM>
__global__ __launch_bounds__(256) void k_g2p2(GridGeom g, PartPtrs<R> pp, GridPtrs<R> gp,
                                              const Material<R>* __restrict__ mats,
                                              StepArgs<R> sa, int np, int* flags) {
    constexpr int NS = DIM == 2 ? 4 : 6;
    constexpr int NF = DIM * DIM;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double x[3];
    load_x<DIM>(pp.x, i, x);
    CommitIn<DIM, R> cin;
    if constexpr (kCommitLoadsEarly<R, CM>) load_commit_in<DIM, R, CM>(pp, sa, i, cin);
    Stencil<DIM, S, R> st;
    int flag = 0;
    if (!make_stencil<DIM, S, R>(g, x, st, flag)) {
        if (sa.mvl) pp.p2c[i] = -1;  // frozen particle: no cell
        return;
    }
    R F[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) F[k] = R(0);
    for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R dNx, R dNy, R dNz) {
        if (N != R(0)) {
            if constexpr (DIM == 2) {
                R u0 = gp.du[0][n], u1 = gp.du[1][n];
                F[0] += u0 * dNx;
                F[1] += u0 * dNy;
                F[2] += u1 * dNx;
                F[3] += u1 * dNy;
            } else {
                R u0 = gp.du[0][n], u1 = gp.du[1][n], u2 = gp.du[2][n];
                F[0] += u0 * dNx; F[1] += u0 * dNy; F[2] += u0 * dNz;
                F[3] += u1 * dNx; F[4] += u1 * dNy; F[5] += u1 * dNz;
                F[6] += u2 * dNx; F[7] += u2 * dNy; F[8] += u2 * dNz;
            }
        }
    });
    if constexpr (!kCommitLoadsEarly<R, CM>) load_commit_in<DIM, R, CM>(pp, sa, i, cin);
    g2p2_commit<DIM, R, CM>(g, pp, gp, mats, sa, i, x, F, cin);
    raise_flag(flags, flag);
}

// vollock1_OS! (utils_OS.jl:1162-1193) on its own, for USF, whose MVL runs
// after the position update (OS_USF.jl:49-52,92-95): cell sums of Ω σm and Ω
// by the step-start cell (x is not committed until the next step).
template <int DIM, int S, typename R>
__global__ __launch_bounds__(256) void k_mvl1(GridGeom g, PartPtrs<R> pp, GridPtrs<R> gp, int np) {
    constexpr int NS = DIM == 2 ? 4 : 6;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double x[3];
    load_x<DIM>(pp.x, i, x);
    Stencil<DIM, S, R> st;
    int flag = 0;
    if (!make_stencil<DIM, S, R>(g, x, st, flag)) {
        pp.p2c[i] = -1;  // frozen particle: no cell
        return;
    }
    R s[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = pp.sig[k][i];
    const int c = cell_of<DIM>(g, x);
    pp.p2c[i] = c;
    const R vol = pp.vol[i];
    atomic_add(&gp.cell_sm[c], vol * mean3<NS>(s));
    atomic_add(&gp.cell_vol[c], vol);
}

// vollock2_OS! (utils_OS.jl:1202-1256)
template <int DIM, typename R>
__global__ __launch_bounds__(256) void k_mvl2(PartPtrs<R> pp, GridPtrs<R> gp, int np) {
    constexpr int NS = DIM == 2 ? 4 : 6;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    R s[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = pp.sig[k][i];
    R sm0 = mean3<NS>(s);
    int c = pp.p2c[i];
    if (c < 0) return;
    R smc = gp.cell_sm[c] / gp.cell_vol[c];
    pp.sig[0][i] = (s[0] - sm0) + smc;
    pp.sig[1][i] = (s[1] - sm0) + smc;
    pp.sig[2][i] = (s[2] - sm0) + smc;
}

// Marks every node with a non-zero weight from some particle (diagnostic).
template <int DIM, int S, typename R>
__global__ __launch_bounds__(256) void k_mark_active(GridGeom g, PartPtrs<R> pp,
                                                     unsigned char* __restrict__ mark, int np) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double x[3];
    load_x<DIM>(pp.x, i, x);
    Stencil<DIM, S, R> st;
    int flag = 0;
    if (!make_stencil<DIM, S, R>(g, x, st, flag)) return;
    for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R, R, R) {
        if (N != R(0)) mark[n] = 1;
    });
}

// k_mark_active on a bitmap: bit n % 32 of word n / 32 set for every node n
// with a non-zero weight.
template <int DIM, int S, typename R>
__global__ __launch_bounds__(256) void k_mark_active_bits(GridGeom g, PartPtrs<R> pp,
                                                          unsigned* __restrict__ words, int np) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double x[3];
    load_x<DIM>(pp.x, i, x);
    Stencil<DIM, S, R> st;
    int flag = 0;
    if (!make_stencil<DIM, S, R>(g, x, st, flag)) return;
    for_stencil<DIM, S, R>(g, st, [&](long long n, R N, R, R, R) {
        if (N != R(0)) atomicOr(&words[n >> 5], 1u << (n & 31));
    });
}

// Set bits of words[0, nw), added to *out.
__global__ __launch_bounds__(256) void k_count_bits(const unsigned* __restrict__ words, long long nw,
                                                    unsigned long long* out) {
    __shared__ unsigned long long part[256];
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    unsigned long long c = 0;
    for (; i < nw; i += stride) c += __popc(words[i]);
    part[threadIdx.x] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && part[0]) atomicAdd(out, part[0]);
}

__global__ __launch_bounds__(256) void k_count_bytes(const unsigned char* __restrict__ mark,
                                                     long long n, unsigned long long* out) {
    __shared__ unsigned long long part[256];
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    unsigned long long c = 0;
    for (; i < n; i += stride) c += mark[i];
    part[threadIdx.x] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, part[0]);
}

}  // namespace mpm
