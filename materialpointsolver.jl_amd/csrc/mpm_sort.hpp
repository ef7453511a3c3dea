// mpm_sort.hpp -- periodic on-device cell sort of the particle streams.
//
// Every sort_interval steps the particles are reordered by a tile-blocked cell
// key (tiles of 4x4x4 cells in 3D, 8x8 in 2D; x fastest inside a tile and
// between tiles), so that consecutive lanes of a wavefront touch the same or
// neighbouring grid nodes in every scatter and gather.  The key sort is the
// rocPRIM radix sort (through hipCUB) over (key, index) pairs; the permutation
// is one gather kernel per particle stream into a scratch buffer of the same
// element size, followed by a pointer swap.  The original particle index
// travels with the particles (pid) so downloads return the caller's order.
#pragma once
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>
#include <utility>

#include "../../include/mpm_hip.h"
#include "mpm_kernels.hpp"
#include "mpm_tiled.hpp"

namespace mpm {

struct SortKeyGeom {
    int ncx, ncy, ncz;   // cells per axis
    int tx, ty, tz;      // tiles per axis
    int tsh;             // log2 tile edge (2 in 3D, 3 in 2D)
};

template <int DIM>
__global__ __launch_bounds__(256) void k_sort_keys(GridGeom g, SortKeyGeom kg,
                                                   const double* __restrict__ x0,
                                                   const double* __restrict__ x1,
                                                   const double* __restrict__ x2,
                                                   unsigned* __restrict__ keys,
                                                   int* __restrict__ idx, int np) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double* xs[3] = {x0, x1, x2};
    const int ncv[3] = {kg.ncx, kg.ncy, kg.ncz};
    int cc[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        double u = (xs[a][i] - g.org[a]) * g.hinv[a];
        double f = floor(u);
        if (!(f >= 0.0)) f = 0.0;  // also maps NaN to 0
        if (f > double(ncv[a] - 1)) f = double(ncv[a] - 1);
        cc[a] = int(f);
    }
    const int te = 1 << kg.tsh, tm = te - 1;
    unsigned tile = unsigned(cc[0] >> kg.tsh) +
                    unsigned(kg.tx) * (unsigned(cc[1] >> kg.tsh) + unsigned(kg.ty) * unsigned(cc[2] >> kg.tsh));
    unsigned local = unsigned(cc[0] & tm) + unsigned(te) * (unsigned(cc[1] & tm) + unsigned(te) * unsigned(cc[2] & tm));
    // quadrant of the particle inside its cell (upper half per axis): particles
    // of one quadrant share the stencil offset, so ordering tiles by
    // (quadrant, cell) gives 64 consecutive particles 64 distinct stencil bases
    unsigned quad = 0;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        double u = (xs[a][i] - g.org[a]) * g.hinv[a];
        if (u - floor(u) >= 0.5) quad |= 1u << a;
    }
    keys[i] = (tile << (DIM + 6)) | (quad << 6) | local;
    idx[i] = i;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather(const T* __restrict__ src, const int* __restrict__ idx,
                                                T* __restrict__ dst, int np) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) dst[i] = src[idx[i]];
}

template <typename R>
struct SortState {
    std::string error;
    bool enabled = false;
    int dim = 3;
    long long np = 0;
    SortKeyGeom kg{};
    GridGeom geom{};
    int key_bits = 32;
    TileGeom tg{};
    // tile segments of the sorted order (valid after the first sort)
    int *seg_begin = nullptr, *seg_end = nullptr, *tile_list = nullptr, *tile_count = nullptr;
    unsigned char* nonempty = nullptr;
    bool tiles_valid = false;
    unsigned *keys = nullptr, *keys_alt = nullptr;
    int *idx = nullptr, *idx_alt = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    // Scratch streams swap roles with particle streams on every permutation, so
    // the caller allocates and frees them together with the particle streams.
    void* scratchR = nullptr;   // np * sizeof(R)
    void* scratch8 = nullptr;   // np * 8 (only when R is 4 bytes)
    void* scratch4 = nullptr;   // np * 4 (int streams, only when R is 8 bytes)

    int init(const mpm_config& cfg, const GridGeom& g, long long n) {
        if (cfg.sort_interval <= 0) return MPM_OK;
        enabled = true;
        dim = cfg.dim;
        np = n;
        geom = g;
        kg.ncx = g.nc[0];
        kg.ncy = g.nc[1];
        kg.ncz = dim == 3 ? g.nc[2] : 1;
        kg.tsh = dim == 3 ? 2 : 3;
        const int te = 1 << kg.tsh;
        kg.tx = (kg.ncx + te - 1) / te;
        kg.ty = (kg.ncy + te - 1) / te;
        kg.tz = dim == 3 ? (kg.ncz + te - 1) / te : 1;
        unsigned long long maxkey = (unsigned long long)kg.tx * kg.ty * kg.tz << (dim + 6);
        if (maxkey >= (1ull << 32)) {
            error = "grid too large for 32-bit sort keys";
            return MPM_E_ARG;
        }
        key_bits = 1;
        while ((1ull << key_bits) < maxkey) ++key_bits;
        tg.tsh = kg.tsh;
        tg.tx = kg.tx;
        tg.ty = kg.ty;
        tg.tz = kg.tz;
        tg.ntiles = kg.tx * kg.ty * kg.tz;
        hipError_t e = hipMalloc((void**)&keys, sizeof(unsigned) * np);
        if (e == hipSuccess) e = hipMalloc((void**)&seg_begin, sizeof(int) * tg.ntiles);
        if (e == hipSuccess) e = hipMalloc((void**)&seg_end, sizeof(int) * tg.ntiles);
        if (e == hipSuccess) e = hipMalloc((void**)&tile_list, sizeof(int) * tg.ntiles);
        if (e == hipSuccess) e = hipMalloc((void**)&tile_count, sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void**)&nonempty, tg.ntiles);
        if (e == hipSuccess) e = hipMalloc((void**)&keys_alt, sizeof(unsigned) * np);
        if (e == hipSuccess) e = hipMalloc((void**)&idx, sizeof(int) * np);
        if (e == hipSuccess) e = hipMalloc((void**)&idx_alt, sizeof(int) * np);
        if (e != hipSuccess) {
            error = std::string("hipMalloc: ") + hipGetErrorString(e);
            return MPM_E_NOMEM;
        }
        e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys_alt, idx, idx_alt,
                                               int(np), 0, key_bits);
        if (e == hipSuccess) e = hipMalloc(&temp, temp_bytes ? temp_bytes : 1);
        if (e != hipSuccess) {
            error = std::string("radix sort workspace: ") + hipGetErrorString(e);
            return MPM_E_NOMEM;
        }
        return MPM_OK;
    }

    void release() {
        void* ps[] = {keys, keys_alt, idx, idx_alt, temp, seg_begin, seg_end, tile_list,
                      tile_count, nonempty};
        for (void* p : ps)
            if (p) hipFree(p);
        keys = keys_alt = nullptr;
        idx = idx_alt = nullptr;
        temp = nullptr;
    }

    template <typename T>
    hipError_t permute(T*& arr, void*& scratch, const int* order, hipStream_t s) {
        const unsigned nb = unsigned((np + 255) / 256);
        k_gather<T><<<nb, 256, 0, s>>>(arr, order, (T*)scratch, int(np));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        T* old = arr;
        arr = (T*)scratch;
        scratch = (void*)old;
        return hipSuccess;
    }

    // Sort and permute every particle stream.  `c` is the context (for the
    // launch wrapper that times k_sort and k_permute).
    template <typename Ctx>
    int run(Ctx* c, PartPtrs<R>& pp, hipStream_t s) {
        if (!enabled) return MPM_OK;
        const unsigned nb = unsigned((np + 255) / 256);
        const int n = int(np);
        int rc = c->launch(6, [&] {
            if (dim == 3)
                k_sort_keys<3><<<nb, 256, 0, s>>>(geom, kg, pp.x[0], pp.x[1], pp.x[2], keys, idx, n);
            else
                k_sort_keys<2><<<nb, 256, 0, s>>>(geom, kg, pp.x[0], pp.x[1], pp.x[1], keys, idx, n);
            hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_alt, idx, idx_alt, n, 0,
                                               key_bits, s);
        });
        if (rc != MPM_OK) return rc;
        const int* order = idx_alt;
        const int ns = dim == 2 ? 4 : 6;
        hipError_t e = hipSuccess;
        rc = c->launch(7, [&] {
            void*& s8 = sizeof(R) == 8 ? scratchR : scratch8;
            void*& sR = scratchR;
            void*& s4 = sizeof(R) == 4 ? scratchR : scratch4;
            for (int a = 0; a < dim && e == hipSuccess; ++a) e = permute(pp.x[a], s8, order, s);
            if (e == hipSuccess) e = permute(pp.J, s8, order, s);
            for (int a = 0; a < dim && e == hipSuccess; ++a) e = permute(pp.v[a], sR, order, s);
            for (int k = 0; k < ns && e == hipSuccess; ++k) e = permute(pp.sig[k], sR, order, s);
            for (int k = 0; k < ns && e == hipSuccess; ++k) e = permute(pp.eps[k], sR, order, s);
            if (e == hipSuccess) e = permute(pp.ms, sR, order, s);
            if (e == hipSuccess) e = permute(pp.vol, sR, order, s);
            if (e == hipSuccess) e = permute(pp.vol0, sR, order, s);
            if (e == hipSuccess) e = permute(pp.eq, sR, order, s);
            if (e == hipSuccess) e = permute(pp.ek, sR, order, s);
            if (e == hipSuccess) e = permute(pp.ev, sR, order, s);
            if (e == hipSuccess) e = permute(pp.nid, s4, order, s);
            if (e == hipSuccess) e = permute(pp.pid, s4, order, s);
        });
        if (rc != MPM_OK) return rc;
        if (e != hipSuccess) return c->hip_fail(e, "k_permute");
        // tile segments of the new order and the ordered list of non-empty tiles
        rc = c->launch(6, [&] {
            const unsigned tb = unsigned((tg.ntiles + 255) / 256);
            k_fill_u8<<<tb, 256, 0, s>>>(nonempty, tg.ntiles);
            k_tile_segments<<<nb, 256, 0, s>>>(keys_alt, n, dim + 6, seg_begin, seg_end, nonempty);
            k_compact_tiles<<<1, 1024, 0, s>>>(nonempty, tg.ntiles, tile_list, tile_count);
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
