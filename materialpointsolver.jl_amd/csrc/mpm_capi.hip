// mpm_capi.hip -- C ABI of libmpm_hip.so (declared in include/mpm_hip.h).
//
// Host side of the gfx950 MUSL / USL / USF step: context and device buffers, uploads from
// the reference's Julia layouts, the per-step launch sequence, downloads in
// the caller's particle order, HIP-event kernel timing and the triad probe.
#include <hip/hip_runtime.h>
#include <type_traits>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpm_hip.h"
#include "mpm_kernels.hpp"
#include "mpm_owned.hpp"
#include "mpm_sort.hpp"
#include "mpm_surrogate.hpp"

using namespace mpm;

namespace {

std::string g_create_error = "no error";

const char* const kKernelNames[] = {"k_p2g",  "k_grid1", "k_g2p1_p2g2", "k_grid2",
                                    "k_g2p2", "k_mvl2",  "k_sort",      "k_permute",
                                    "k_learned_return", "k_p2g2", "k_mvl1"};
constexpr int kNumKernels = int(sizeof(kKernelNames) / sizeof(kKernelNames[0]));
enum { KID_P2G = 0, KID_GRID1, KID_G2P1, KID_GRID2, KID_G2P2, KID_MVL2, KID_SORT, KID_PERMUTE,
       KID_LEARNED, KID_P2G2, KID_MVL1 };
// parts of a step: MUSL/USL run FULL; USF runs HEAD (the stress update of the
// previous step's displacement), then the sort when due, then BODY
enum { PH_FULL = 0, PH_HEAD = 1, PH_BODY = 2 };

constexpr int kBlock = 256;

inline unsigned nblocks(long long n) { return unsigned((n + kBlock - 1) / kBlock); }

// Julia's Base.min for IEEE floats (NaN propagates).
inline double jl_min(double x, double y) {
    if (std::isnan(x)) return x;
    if (std::isnan(y)) return y;
    if (y < x || (std::signbit(y) && !std::signbit(x))) return y;
    return x;
}

struct ProfRec {
    int kid;
    hipEvent_t e0, e1;
};

}  // namespace

// Scratch of mpm_count_active_nodes_noalloc: a bitmap of the grid nodes (one
// bit per node, at most kNoAllocMaxNodes nodes) followed by a 64-bit counter.
// Each GPU holds its own copy for the life of the process; host code reaches
// it only through the per-device address from hipGetSymbolAddress
// (g_scratch_dev), never through the array's name.
constexpr long long kNoAllocMaxNodes = 1ll << 25;            // 33,554,432 nodes
constexpr long long kNoAllocWords = kNoAllocMaxNodes / 32;   // bitmap words (4 MiB)
static __device__ __align__(256) unsigned scratch[kNoAllocWords + 2];
constexpr int kMaxDevices = 64;
static void* g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Per-device address of `scratch`, looked up once per device with that device
// current (the library's init call, mpm_create, does it).
static hipError_t scratch_address(int dev, void** out) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

struct ImplBase;

struct mpm_ctx {
    mpm_config cfg{};
    int device = 0;               // the device current at mpm_create
    void* scratch_ptr = nullptr;  // that device's address of `scratch`
    hipStream_t stream = nullptr;
    std::string err = "no error";
    ImplBase* impl = nullptr;
    int* d_flags = nullptr;
    bool have_particles = false, have_materials = false, have_bc = false;
    long long steps = 0;
    bool resort = true;  // particle order changed outside the sort (upload)
    bool prof_on = false;
    uint64_t prof_mask = 0;  // kernel ids to time; 0 = all
    bool use_tiles = true;  // LDS-binned scatters when the particles are tile-sorted
    int scatter = -1;                  // tile scatter form, -1: automatic (MPM_OPT_SCATTER)
    int remap_scatter = -1;            // re-map scatter form, -1: automatic (MPM_OPT_REMAP_SCATTER)
    bool fused_remap = false;         // one-pass MUSL re-map (MPM_OPT_FUSED_REMAP)
    bool tile_partials = true;        // tile scatters store windows (MPM_OPT_TILE_PARTIALS)
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> pool;
    // captured step graphs of mpm_run_graph, keyed by what their kernel
    // arguments depend on
    struct Graph {
        long long len;
        int parity, tiled;
        mpm_args args;
        double dt;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    void drop_graphs() {
        for (auto& g : graphs) hipGraphExecDestroy(g.exec);
        graphs.clear();
    }

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return MPM_E_HIP;
    }
    hipEvent_t event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    // Launch wrapper: optional event pair around the launch on the context stream.
    template <typename Fn>
    int launch(int kid, Fn&& fn) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool rec = prof_on && (prof_mask == 0 || ((prof_mask >> kid) & 1ull));
        if (rec) {
            e0 = event();
            if (e0) hipEventRecord(e0, stream);
        }
        fn();
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, kKernelNames[kid]);
        if (rec && e0) {
            e1 = event();
            if (e1) {
                hipEventRecord(e1, stream);
                recs.push_back({kid, e0, e1});
            }
        }
        return MPM_OK;
    }
};

struct ImplBase {
    virtual ~ImplBase() {}
    virtual int upload_particles(mpm_ctx*, const double*, const double*, const double*,
                                 const double*, const double*, const double*, const double*,
                                 const double*, const double*, const double*, const double*,
                                 const int64_t*) = 0;
    virtual int upload_materials(mpm_ctx*, const double*, const double*, const double*,
                                 const double*, const double*, const double*, const double*,
                                 const double*, const double*) = 0;
    virtual int upload_bc(mpm_ctx*, const int64_t*, const double*, const int64_t*, const double*,
                          const int64_t*, const double*) = 0;
    virtual int step(mpm_ctx*, const mpm_args& a, double dt, double Ti) = 0;
    virtual int run_graph(mpm_ctx*, const mpm_args& a, double dt, double Ti, int64_t nsteps,
                          double* Ti_out) = 0;
    virtual int download(mpm_ctx*, double*, double*, double*, double*, double*, double*,
                         double*, double*, double*, double*) = 0;
    virtual int count_active(mpm_ctx*, int64_t*) = 0;
    virtual int count_active_noalloc(mpm_ctx*, int64_t*) = 0;
    virtual int set_surrogate(mpm_ctx*, mpm_mlp*, double tol) = 0;
    virtual int surrogate_stats(mpm_ctx*, int64_t*, int64_t*) = 0;
    virtual int return_log(mpm_ctx*, int64_t cap, int64_t stride) = 0;
    virtual int return_log_read(mpm_ctx*, float*, float*, int64_t, int64_t*) = 0;
};

namespace {

template <typename T>
hipError_t dalloc(T** p, size_t n) {
    return hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
}

// dst[pid[i]] = double(src[i])
template <typename T>
__global__ __launch_bounds__(256) void k_unpermute(const T* __restrict__ src,
                                                   const int* __restrict__ pid,
                                                   double* __restrict__ dst, int np) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) dst[pid[i]] = double(src[i]);
}

// dst[pid[i]] = x[i] + dx[i]: positions with a pending USF increment
template <typename R>
__global__ __launch_bounds__(256) void k_unpermute_add(const double* __restrict__ x,
                                                       const R* __restrict__ dx,
                                                       const int* __restrict__ pid,
                                                       double* __restrict__ dst, int np) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) dst[pid[i]] = x[i] + double(dx[i]);
}

// σm = (σxx + σyy + σzz) * 0.333333 of the stored stress (liE!/dpP!/mcP! all
// end with this assignment, linearelastic.jl:61-62), scattered to caller order.
template <typename R>
__global__ __launch_bounds__(256) void k_sigm(const R* s0, const R* s1, const R* s2,
                                              const int* __restrict__ pid,
                                              double* __restrict__ dst, int np) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) dst[pid[i]] = double((s0[i] + s1[i] + s2[i]) * R(0.333333));
}

// Streaming probe a = b + c, one float4 per thread and no loop: on MI355X this
// form measured 6.1 TB/s at 1 GiB per array, against 4.4-4.9 TB/s for a
// grid-stride loop over 2,048-32,768 workgroups.
__global__ __launch_bounds__(256) void k_triad(const float4* __restrict__ b,
                                               const float4* __restrict__ c,
                                               float4* __restrict__ a, long long n4) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) {
        const float4 x = b[i], y = c[i];
        a[i] = make_float4(x.x + y.x, x.y + y.y, x.z + y.z, x.w + y.w);
    }
}

template <typename R>
struct Impl : ImplBase {
    int dim, ns, nf;
    long long np, ni, nc;
    GridGeom geom{};
    PartPtrs<R> pp{};
    GridPtrs<R> gp{};
    Material<R>* d_mats = nullptr;
    double* d_stage = nullptr;  // np doubles, download staging
    R* tile_part = nullptr;     // tile partials: ntiles x window nodes x 8 (padded fields)
    std::vector<void*> owned;
    SortState<R> sort;
    // learned 3D Mohr-Coulomb return (FP32 only)
    mpm_mlp* sur = nullptr;
    float sur_tol = 1e-3f;
    unsigned long long* d_sur_stats = nullptr;  // [plastic, fallbacks]
    bool have_de = false;
    // USF: the last step's position increment (dxp) and nodal displacement are
    // waiting for the next step's stress update
    bool usf_pending = false;
    // return-mapping sample log
    float *log_x = nullptr, *log_y = nullptr;
    unsigned long long* d_log_count = nullptr;
    long long log_cap = 0;
    int log_stride = 1;

    ~Impl() override {
        for (void* p : owned) hipFree(p);
        if (log_x) hipFree(log_x);
        if (log_y) hipFree(log_y);
        sort.release();
    }

    template <typename T>
    int alloc(mpm_ctx* c, T** p, size_t n, const char* what) {
        hipError_t e = dalloc(p, n);
        if (e != hipSuccess) return c->fail(MPM_E_NOMEM, std::string("hipMalloc ") + what + ": " +
                                                             hipGetErrorString(e));
        owned.push_back((void*)*p);
        return MPM_OK;
    }

    int init(mpm_ctx* c) {
        const mpm_config& cf = c->cfg;
        dim = cf.dim;
        ns = dim == 2 ? 4 : 6;
        nf = dim * dim;
        np = cf.np;
        geom.nn[0] = int(cf.nnx);
        geom.nn[1] = int(cf.nny);
        geom.nn[2] = dim == 3 ? int(cf.nnz) : 1;
        geom.nc[0] = geom.nn[0] - 1;
        geom.nc[1] = geom.nn[1] - 1;
        geom.nc[2] = dim == 3 ? geom.nn[2] - 1 : 1;
        ni = (long long)geom.nn[0] * geom.nn[1] * geom.nn[2];
        nc = (long long)geom.nc[0] * geom.nc[1] * geom.nc[2];
        geom.ni = ni;
        const double org[3] = {cf.x1, cf.y1, cf.z1}, h[3] = {cf.dx, cf.dy, cf.dz},
                     lp[3] = {cf.lpx, cf.lpy, cf.lpz};
        for (int a = 0; a < 3; ++a) {
            geom.org[a] = org[a];
            geom.h[a] = h[a];
            geom.hinv[a] = 1.0 / h[a];
            geom.lp[a] = lp[a];
            geom.c4[a] = 1.0 / (4.0 * h[a] * lp[a]);
            geom.c1[a] = 1.0 / (h[a] * lp[a]);
            geom.half[a] = 0.5 * lp[a];
        }
        int rc;
#define A(ptr, n, what) \
    if ((rc = alloc(c, &(ptr), (n), what)) != MPM_OK) return rc;
        for (int a = 0; a < dim; ++a) {
            A(pp.x[a], np, "x");
            A(pp.v[a], np, "v");
            A(pp.dxp[a], np, "dx");
        }
        for (int k = 0; k < ns; ++k) {
            A(pp.sig[k], np, "sigma");
            A(pp.eps[k], np, "eps");
        }
        A(pp.ms, np, "ms");
        A(pp.vol, np, "vol");
        A(pp.vol0, np, "vol0");
        A(pp.J, np, "J");
        A(pp.eq, np, "eq");
        A(pp.ek, np, "ek");
        A(pp.ev, np, "ev");
        A(pp.nid, np, "nid");
        A(pp.pid, np, "pid");
        A(pp.p2c, np, "p2c");
        if (cf.sort_interval > 0)
            for (int a = 0; a < dim; ++a) A(pp.pm[a], np, "momentum scratch");
        A(gp.m, ni, "grid m");
        for (int d = 0; d < dim; ++d) {
            A(gp.p[d], ni, "grid p");
            A(gp.f[d], ni, "grid f");
            A(gp.vs[d], ni, "grid vs");
            A(gp.vsT[d], ni, "grid vsT");
            A(gp.as[d], ni, "grid as");
            A(gp.du[d], ni, "grid du");
            A(gp.bcv[d], ni, "grid bc value");
        }
        A(gp.bcf, ni, "grid bc flags");
        A(gp.cell_sm, nc, "cell sm");
        A(gp.cell_vol, nc, "cell vol");
        A(d_stage, np, "staging");
        A(d_mats, cf.nmat, "materials");
#undef A
        // zero every grid field once; afterwards each step leaves m, p, f at zero
        hipError_t e = hipSuccess;
        e = hipMemset(gp.m, 0, sizeof(R) * ni);
        for (int d = 0; d < dim && e == hipSuccess; ++d) {
            e = hipMemset(gp.p[d], 0, sizeof(R) * ni);
            if (e == hipSuccess) e = hipMemset(gp.f[d], 0, sizeof(R) * ni);
            if (e == hipSuccess) e = hipMemset(gp.bcv[d], 0, sizeof(R) * ni);
            if (e == hipSuccess) e = hipMemset(gp.du[d], 0, sizeof(R) * ni);
            // gathers over LDS windows read nodes of zero weight too: keep every
            // field finite from the start
            if (e == hipSuccess) e = hipMemset(gp.vs[d], 0, sizeof(R) * ni);
            if (e == hipSuccess) e = hipMemset(gp.vsT[d], 0, sizeof(R) * ni);
            if (e == hipSuccess) e = hipMemset(gp.as[d], 0, sizeof(R) * ni);
        }
        if (e == hipSuccess) e = hipMemset(gp.bcf, 0, ni);
        if (e == hipSuccess) e = hipMemset(gp.cell_sm, 0, sizeof(R) * nc);
        if (e == hipSuccess) e = hipMemset(gp.cell_vol, 0, sizeof(R) * nc);
        if (e != hipSuccess) return c->hip_fail(e, "hipMemset");
        if ((rc = sort.init(c->cfg, geom, np)) != MPM_OK)
            return c->fail(rc, "sort buffers: " + sort.error);
        if (sort.enabled) {
            hipError_t ea = hipSuccess;
            SortState<R>::for_each_stream(pp, dim, [&](auto*& ptr) {
                using T = std::remove_pointer_t<std::remove_reference_t<decltype(ptr)>>;
                T* q = nullptr;
                if (ea == hipSuccess) ea = dalloc(&q, np);
                if (ea == hipSuccess) {
                    owned.push_back((void*)q);
                    sort.alt.push_back((void*)q);
                }
            });
            if (ea != hipSuccess) return c->fail(MPM_E_NOMEM, std::string("sort buffers: ") + hipGetErrorString(ea));
            // tile partials of the widest scatter (P2G #1: 1 + 2 dim fields,
            // padded to part_stride values per node)
            const int S = c->cfg.basis == MPM_LINEAR ? 2 : 3;
            const long long ww = (1ll << sort.tg.tsh) + S - 1;
            const long long wn = dim == 3 ? ww * ww * ww : ww * ww;
            const long long ns = dim == 3 ? part_stride<7>() : part_stride<5>();
            const long long cnt = (long long)sort.tg.ntiles * ns * wn;
            if ((rc = alloc(c, &tile_part, cnt, "tile partials")) != MPM_OK) return rc;
        }
        return MPM_OK;
    }

    int put(mpm_ctx* c, R* dst, const double* src, const char* what) {
        std::vector<R> tmp(np);
        for (long long i = 0; i < np; ++i) tmp[i] = R(src[i]);
        hipError_t e = hipMemcpy(dst, tmp.data(), sizeof(R) * np, hipMemcpyHostToDevice);
        return e == hipSuccess ? MPM_OK : c->hip_fail(e, what);
    }

    int upload_particles(mpm_ctx* c, const double* xi, const double* vs, const double* sigma,
                         const double* eps_s, const double* F, const double* omega,
                         const double* omega0, const double* rho, const double* rho0,
                         const double* eps_q, const double* eps_k, const int64_t* nid) override {
        if (!xi || !vs || !sigma || !eps_s || !F || !omega || !omega0 || !rho || !rho0 ||
            !eps_q || !eps_k || !nid)
            return c->fail(MPM_E_ARG, "mpm_upload_particles: NULL array");
        int rc;
        hipError_t e;
        // every check runs before the first copy, so a rejected upload leaves
        // the device state as it was
        std::vector<double> J(np), ms(np);
        std::vector<int> nid32(np), pid(np);
        for (long long i = 0; i < np; ++i) {
            const double* f = F + i;
            auto Fk = [&](int k) { return f[(long long)k * np]; };
            if (dim == 2) {
                J[i] = Fk(0) * Fk(3) - Fk(1) * Fk(2);
            } else {
                J[i] = Fk(0) * Fk(4) * Fk(8) + Fk(1) * Fk(5) * Fk(6) + Fk(2) * Fk(3) * Fk(7) -
                       Fk(6) * Fk(4) * Fk(2) - Fk(7) * Fk(5) * Fk(0) - Fk(8) * Fk(3) * Fk(1);
            }
            ms[i] = omega[i] * rho[i];
            if (nid[i] < 1 || nid[i] > c->cfg.nmat)
                return c->fail(MPM_E_ARG, "mpm_upload_particles: nid[" + std::to_string(i) +
                                              "] = " + std::to_string(nid[i]) +
                                              " outside 1.." + std::to_string(c->cfg.nmat));
            nid32[i] = int(nid[i] - 1);
            pid[i] = int(i);
            if (!(std::isfinite(xi[i]) && std::isfinite(xi[np + i]) &&
                  (dim == 2 || std::isfinite(xi[2 * np + i]))))
                return c->fail(MPM_E_RANGE, "mpm_upload_particles: particle " + std::to_string(i + 1) +
                                                " has a non-finite position");
        }
        for (int a = 0; a < dim; ++a) {
            e = hipMemcpy(pp.x[a], xi + a * np, sizeof(double) * np, hipMemcpyHostToDevice);
            if (e != hipSuccess) return c->hip_fail(e, "upload xi");
            if ((rc = put(c, pp.v[a], vs + a * np, "upload vs")) != MPM_OK) return rc;
        }
        for (int k = 0; k < ns; ++k) {
            if ((rc = put(c, pp.sig[k], sigma + k * np, "upload sigma")) != MPM_OK) return rc;
            if ((rc = put(c, pp.eps[k], eps_s + k * np, "upload eps_s")) != MPM_OK) return rc;
        }
        e = hipMemcpy(pp.J, J.data(), sizeof(double) * np, hipMemcpyHostToDevice);
        if (e != hipSuccess) return c->hip_fail(e, "upload J");
        if ((rc = put(c, pp.ms, ms.data(), "upload ms")) != MPM_OK) return rc;
        if ((rc = put(c, pp.vol, omega, "upload omega")) != MPM_OK) return rc;
        if ((rc = put(c, pp.vol0, omega0, "upload omega0")) != MPM_OK) return rc;
        if ((rc = put(c, pp.eq, eps_q, "upload eps_q")) != MPM_OK) return rc;
        if ((rc = put(c, pp.ek, eps_k, "upload eps_k")) != MPM_OK) return rc;
        e = hipMemcpy(pp.nid, nid32.data(), sizeof(int) * np, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(pp.pid, pid.data(), sizeof(int) * np, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(pp.ev, 0, sizeof(R) * np);
        if (e == hipSuccess) e = hipMemset(pp.p2c, 0xff, sizeof(int) * np);  // -1: no cell
        for (int a = 0; a < dim && e == hipSuccess; ++a) e = hipMemset(pp.dxp[a], 0, sizeof(R) * np);
        if (e != hipSuccess) return c->hip_fail(e, "upload nid/pid");
        sort.tiles_valid = false;  // caller order: sort again before the next step
        c->resort = true;
        usf_pending = false;
        return MPM_OK;
    }

    int upload_materials(mpm_ctx* c, const double* Ks, const double* Gs, const double* phi,
                         const double* psi, const double* coh, const double* sigt,
                         const double* Hp, const double* cr, const double* phir) override {
        if (!Ks || !Gs || !phi || !psi || !coh || !sigt || !Hp || !cr)
            return c->fail(MPM_E_ARG, "mpm_upload_materials: NULL array");
        const long long nm = c->cfg.nmat;
        std::vector<Material<R>> M(nm);
        for (long long k = 0; k < nm; ++k) {
            // druckerprager.jl:178-183, same expressions in FP64
            double kphi = (6.0 * coh[k] * std::cos(phi[k])) / (1.732051 * (3.0 + std::sin(phi[k])));
            double qphi = (6.0 * std::sin(phi[k])) / (1.732051 * (3.0 + std::sin(phi[k])));
            double qpsi = (6.0 * std::sin(psi[k])) / (1.732051 * (3.0 + std::sin(psi[k])));
            double st = jl_min(sigt[k], kphi / qphi);
            double ab = std::sqrt(1.0 + qphi * qphi) - qphi;
            double tb = kphi - qphi * st;
            Material<R>& m = M[k];
            m.Ks = R(Ks[k]);
            m.Gs = R(Gs[k]);
            m.Dt = R(Ks[k] + 1.333333 * Gs[k]);
            m.Dd = R(Ks[k] - 0.666667 * Gs[k]);
            m.kphi = R(kphi);
            m.qphi = R(qphi);
            m.qpsi = R(qpsi);
            m.sigt = R(st);
            m.ab = R(ab);
            m.tb = R(tb);
            m.eq_shear = R(std::sqrt(0.333333 + 0.222222 * qpsi * qpsi));
            m.sinphi = R(std::sin(phi[k]));
            m.cosphi = R(std::cos(phi[k]));
            m.tanphi = R(std::tan(phi[k]));
            m.sinpsi = R(std::sin(psi[k]));
            m.phi = R(phi[k]);
            m.psi = R(psi[k]);
            m.c = R(coh[k]);
            m.Hp = R(Hp[k]);
            m.cr = R(cr[k]);
            // 3D Mohr-Coulomb softening of friction and dilation (mc_strength):
            // over the plastic strain in which cohesion softens from c to cr
            const double pr = phir ? phir[k] : 0.0;
            m.phir = R(pr);
            m.eqr = (Hp[k] < 0.0 && coh[k] > cr[k] && pr > 0.0) ? R((cr[k] - coh[k]) / Hp[k]) : R(0);
        }
        hipError_t e = hipMemcpy(d_mats, M.data(), sizeof(Material<R>) * nm, hipMemcpyHostToDevice);
        return e == hipSuccess ? MPM_OK : c->hip_fail(e, "upload materials");
    }

    // Reference node id (0-based) -> this library's x-fastest id.
    long long node_map(long long n) const {
        const long long nnx = geom.nn[0], nny = geom.nn[1];
        if (dim == 2) {
            long long r = n % nny, col = n / nny;  // r: row from the top
            long long iy = nny - 1 - r;
            return col + nnx * iy;
        }
        long long iy = n % nny, rest = n / nny;
        long long ix = rest % nnx, iz = rest / nnx;
        return ix + nnx * (iy + nny * iz);
    }

    int upload_bc(mpm_ctx* c, const int64_t* vxi, const double* vxv, const int64_t* vyi,
                  const double* vyv, const int64_t* vzi, const double* vzv) override {
        const int64_t* idx[3] = {vxi, vyi, vzi};
        const double* val[3] = {vxv, vyv, vzv};
        for (int d = 0; d < dim; ++d)
            if (!idx[d] || !val[d]) return c->fail(MPM_E_ARG, "mpm_upload_bc: NULL array");
        std::vector<unsigned char> flags(ni, 0);
        std::vector<R> v[3];
        for (int d = 0; d < dim; ++d) v[d].assign(ni, R(0));
        for (long long n = 0; n < ni; ++n) {
            long long q = node_map(n);
            for (int d = 0; d < dim; ++d) {
                if (idx[d][n] != 0) flags[q] |= (unsigned char)(1u << d);
                v[d][q] = R(val[d][n]);
            }
        }
        hipError_t e = hipMemcpy(gp.bcf, flags.data(), ni, hipMemcpyHostToDevice);
        for (int d = 0; d < dim && e == hipSuccess; ++d)
            e = hipMemcpy(gp.bcv[d], v[d].data(), sizeof(R) * ni, hipMemcpyHostToDevice);
        return e == hipSuccess ? MPM_OK : c->hip_fail(e, "upload bc");
    }

    template <int DIM, int S>
    int step_impl(mpm_ctx* c, const mpm_args& ar, double dt, double Ti, int phase) {
        const mpm_config& cf = c->cfg;
        // OS_MUSL.jl:70 gravity ramp and :88 plasticity gate, evaluated in FP64
        double G = Ti < ar.Te ? ar.gravity / ar.Te * Ti : ar.gravity;
        Gravity<R> Gv{};
        const double gx = Ti < ar.Te ? ar.gravity_x / ar.Te * Ti : ar.gravity_x;
        const double gy = Ti < ar.Te ? ar.gravity_y / ar.Te * Ti : ar.gravity_y;
        Gv.g[0] = R(gx);
        Gv.g[1] = R(DIM == 3 ? gy : G);
        Gv.g[2] = R(DIM == 3 ? G : 0.0);
        StepArgs<R> sa{};
        sa.dt = R(dt);
        sa.G = R(G);
        sa.flip = R(ar.flip);
        sa.pic = R(ar.pic);
        sa.zeta = R(ar.zeta_s);
        sa.va = ar.va;
        sa.pic_new = (ar.va == MPM_VA_A && ar.scheme != MPM_MUSL) ? 1 : 0;
        sa.remap = ar.scheme == MPM_MUSL ? 1 : 0;
        // a USF context's first stress update has no position increment to commit
        sa.commit_dx = (phase != PH_HEAD || usf_pending) ? 1 : 0;
        sa.plastic = Ti >= ar.Te ? 1 : 0;
        sa.constitutive = ar.constitutive;
        sa.mvl = ar.mvl;
        const bool mc3 = DIM == 3 && ar.constitutive == MPM_MOHRCOULOMB && sa.plastic;
        sa.defer_return = (mc3 && sur) ? 1 : 0;
        if (sa.defer_return && ar.mvl)
            return c->fail(MPM_E_UNSUPPORTED, "mpm_step: the learned return does not combine with MVL");
        if (mc3 && !sur && log_cap > 0) {
            sa.log_x = log_x;
            sa.log_y = log_y;
            sa.log_count = d_log_count;
            sa.log_cap = log_cap;
            sa.log_stride = log_stride;
        }
        const int n = int(np);
        const unsigned pb = nblocks(np), gb = nblocks(ni > nc ? ni : nc);
        hipStream_t s = c->stream;
        int* fl = c->d_flags;
        int rc;
        (void)cf;
        const bool tiled = sort.enabled && sort.tiles_valid && c->use_tiles;
        // layer windows for the rank rounds: FP32 3D uGIMP (elsewhere the
        // two-window rounds run)
        constexpr bool kLayerable = DIM == 3 && S == 3 && sizeof(R) == 4;
        // scatter forms of P2G #1 and of the re-map's momentum scatter; by
        // default (measured fastest) FP32 3D uGIMP runs P2G #1 in pair rounds
        // and the re-map scatter base-owned, everything else in rank rounds;
        // the re-map's form follows P2G #1's unless set (MPM_OPT_REMAP_SCATTER)
        const int form1 = c->scatter >= 0 ? c->scatter
                                          : (kLayerable ? int(MPM_SCATTER_PAIRS) : int(MPM_SCATTER_ROUNDS));
        const int form2 = c->remap_scatter >= 0 ? c->remap_scatter
                                                : (kLayerable ? int(MPM_SCATTER_OWNED) : form1);
        // MPM_SCATTER_PAIRS is FP32 3D uGIMP only (elsewhere it runs as MPM_SCATTER_OWNED)
        auto is_owned = [](int f) {
            return f == MPM_SCATTER_OWNED || f == MPM_SCATTER_OWNED_LAYERS || f == MPM_SCATTER_PAIRS;
        };
        const bool owned = is_owned(form1), owned2 = is_owned(form2);
        const bool layered = kLayerable && form1 == MPM_SCATTER_ROUNDS_LAYERED;
        const bool layered2 = kLayerable && form2 == MPM_SCATTER_ROUNDS_LAYERED;
        constexpr int TSH = DIM == 3 ? 2 : 3;
        // tile partials: the tile scatters store their windows, the grid solves
        // add them up (no flush atomics; MPM_OPT_TILE_PARTIALS)
        R* const tbp = (tiled && c->tile_partials) ? tile_part : nullptr;
        TilePartials<R> tpart{};
        if (tbp) {
            constexpr int WWt = Win<DIM, S, TSH>::WW;
            // add_partials reads at most two tile windows per axis
            static_assert(WWt <= (2 << TSH), "a tile window spans at most two tiles per axis");
            tpart.tb = tbp;
            tpart.slot_of = sort.slot_of;
            tpart.tsh = TSH;
            tpart.ww = WWt;
            tpart.wn = Win<DIM, S, TSH>::WN;
            tpart.tn[0] = sort.tg.tx;
            tpart.tn[1] = sort.tg.ty;
            tpart.tn[2] = sort.tg.tz;
        }
        // Base-owned scatter passes (OwnedPasses in mpm_owned.hpp): FP64 3D
        // always takes one z layer per pass (its FP64 sums would not fit the
        // register file); FP32 3D P2G #1 sums mass and momentum, then force
        // (MPM_SCATTER_OWNED) or also one z layer per pass
        // (MPM_SCATTER_OWNED_LAYERS); 2D sums everything in one pass.
        auto owned_scatter = [&](auto mode) {
            constexpr int MODE = decltype(mode)::value;
            const bool layers =
                DIM == 3 && (sizeof(R) == 8 || (MODE == 0 ? form1 : form2) == MPM_SCATTER_OWNED_LAYERS);
            constexpr int SP_FIELDS = MODE == 0 ? 1 : 0;
            constexpr bool F64 = sizeof(R) == 8;
            if constexpr (kLayerable) {
                if ((MODE == 0 ? form1 : form2) == MPM_SCATTER_PAIRS) {
                    k_scatter_pairs<S, TSH, MODE><<<sort.tiled_grid(), 64, 0, s>>>(
                        geom, sort.tg, pp, gp, Gv, sort.masks, sort.offs, sort.ovf, sort.rounds,
                        sort.tile_list, sort.seg_begin, sort.seg_end, sort.tile_count, fl, tbp);
                    return;
                }
            }
            if (DIM == 3 && (F64 || layers))
                k_scatter_owned<DIM, S, TSH, R, MODE, DIM == 3 ? 2 : 0><<<sort.tiled_grid(), 64, 0, s>>>(
                    geom, sort.tg, pp, gp, Gv, sort.masks, sort.offs, sort.ovf, sort.rounds,
                    sort.tile_list, sort.seg_begin, sort.seg_end, sort.tile_count, fl, tbp);
            else if constexpr (!F64 || DIM == 2)
                k_scatter_owned<DIM, S, TSH, R, MODE, DIM == 3 ? SP_FIELDS : 0><<<sort.tiled_grid(), 64, 0, s>>>(
                    geom, sort.tg, pp, gp, Gv, sort.masks, sort.offs, sort.ovf, sort.rounds,
                    sort.tile_list, sort.seg_begin, sort.seg_end, sort.tile_count, fl, tbp);
        };
        if (phase != PH_HEAD) {
            if ((rc = c->launch(KID_P2G, [&] {
                     if (tiled && owned)
                         owned_scatter(std::integral_constant<int, 0>{});
                     else if (tiled && layered) {
                         if constexpr (kLayerable)
                             k_scatter_tiled<DIM, S, TSH, R, 0, true><<<sort.tiled_grid(), 64, 0, s>>>(
                                 geom, sort.tg, pp, gp, Gv, sort.sorted_keys(), sort.tile_list,
                                 sort.seg_begin, sort.seg_end, sort.tile_count, fl, tbp);
                     } else if (tiled)
                         k_scatter_tiled<DIM, S, TSH, R, 0><<<sort.tiled_grid(), 64, 0, s>>>(
                             geom, sort.tg, pp, gp, Gv, sort.sorted_keys(), sort.tile_list,
                             sort.seg_begin, sort.seg_end, sort.tile_count, fl, tbp);
                     else
                         k_p2g<DIM, S, R><<<pb, kBlock, 0, s>>>(geom, pp, gp, Gv, n, fl);
                 })) != MPM_OK)
                return rc;
            if ((rc = c->launch(KID_GRID1, [&] {
                     k_grid1<DIM, R><<<gb, kBlock, 0, s>>>(geom, gp, sa, int(nc), tpart);
                 })) != MPM_OK)
                return rc;
            // re-map gather (tiled: shared window, 256 threads); MUSL then scatters the
            // new momentum (tiled: its own kernel, or one pass with the gather in the
            // scatter's rounds or in the base-owned sums; otherwise fused, global
            // atomics).  The base-owned one-pass form's FP64 3D uGIMP sums do not
            // fit the register file: there FP64 3D uGIMP keeps the two kernels.
            const bool remap1 = tiled && sa.remap && c->fused_remap &&
                                (!owned2 || !(DIM == 3 && S == 3 && sizeof(R) == 8));
            if ((rc = c->launch(KID_G2P1, [&] {
                     if (remap1 && !owned2) {
                         // the rank-round scatter with the gather and move in its rounds
                         if (sa.va != 0 || sa.pic != R(0))
                             k_scatter_tiled<DIM, S, TSH, R, 3><<<sort.tiled_grid(), 64, 0, s>>>(
                                 geom, sort.tg, pp, gp, Gv, sort.sorted_keys(), sort.tile_list,
                                 sort.seg_begin, sort.seg_end, sort.tile_count, fl, tbp, sa);
                         else
                             k_scatter_tiled<DIM, S, TSH, R, 2><<<sort.tiled_grid(), 64, 0, s>>>(
                                 geom, sort.tg, pp, gp, Gv, sort.sorted_keys(), sort.tile_list,
                                 sort.seg_begin, sort.seg_end, sort.tile_count, fl, tbp, sa);
                     } else if (remap1) {
                         if constexpr (!(DIM == 3 && S == 3 && sizeof(R) == 8)) {
                             if (sa.va != 0 || sa.pic != R(0))
                                 k_remap_owned<DIM, S, TSH, R, true><<<sort.tiled_grid(), 64, 0, s>>>(
                                     geom, sort.tg, pp, gp, sa, sort.masks, sort.offs, sort.ovf,
                                     sort.rounds, sort.tile_list, sort.seg_begin, sort.seg_end,
                                     sort.tile_count, tbp);
                             else
                                 k_remap_owned<DIM, S, TSH, R, false><<<sort.tiled_grid(), 64, 0, s>>>(
                                     geom, sort.tg, pp, gp, sa, sort.masks, sort.offs, sort.ovf,
                                     sort.rounds, sort.tile_list, sort.seg_begin, sort.seg_end,
                                     sort.tile_count, tbp);
                         }
                     } else if (tiled)
                         if (sa.va != 0 || sa.pic != R(0))
                             k_g2p1_tiled<DIM, S, TSH, R, true><<<sort.tiled_grid(), 256, 0, s>>>(
                                 geom, sort.tg, pp, gp, sa, sort.tile_list, sort.seg_begin,
                                 sort.seg_end, sort.tile_count);
                         else
                             k_g2p1_tiled<DIM, S, TSH, R, false><<<sort.tiled_grid(), 256, 0, s>>>(
                                 geom, sort.tg, pp, gp, sa, sort.tile_list, sort.seg_begin,
                                 sort.seg_end, sort.tile_count);
                     else
                         k_g2p1_p2g2<DIM, S, R><<<pb, kBlock, 0, s>>>(geom, pp, gp, sa, n, fl);
                 })) != MPM_OK)
                return rc;
            if (tiled && sa.remap && !remap1) {
                if ((rc = c->launch(KID_P2G2, [&] {
                         if (owned2) {
                             owned_scatter(std::integral_constant<int, 1>{});
                         } else if (layered2) {
                             if constexpr (kLayerable)
                                 k_scatter_tiled<DIM, S, TSH, R, 1, true><<<sort.tiled_grid(), 64, 0, s>>>(
                                     geom, sort.tg, pp, gp, Gv, sort.sorted_keys(), sort.tile_list,
                                     sort.seg_begin, sort.seg_end, sort.tile_count, fl, tbp);
                         } else {
                             k_scatter_tiled<DIM, S, TSH, R, 1><<<sort.tiled_grid(), 64, 0, s>>>(
                                 geom, sort.tg, pp, gp, Gv, sort.sorted_keys(), sort.tile_list,
                                 sort.seg_begin, sort.seg_end, sort.tile_count, fl, tbp);
                         }
                     })) != MPM_OK)
                    return rc;
            }
            if (sa.remap) {
                if ((rc = c->launch(KID_GRID2, [&] {
                         k_grid2<DIM, R><<<nblocks(ni), kBlock, 0, s>>>(geom, gp, R(dt), tpart);
                     })) != MPM_OK)
                    return rc;
            }
            if (phase == PH_BODY) {
                // USF: MVL after the position update, on the stress of this step's head
                if (ar.mvl) {
                    if ((rc = c->launch(KID_MVL1, [&] {
                             k_mvl1<DIM, S, R><<<pb, kBlock, 0, s>>>(geom, pp, gp, n);
                         })) != MPM_OK)
                        return rc;
                    if ((rc = c->launch(KID_MVL2, [&] {
                             k_mvl2<DIM, R><<<pb, kBlock, 0, s>>>(pp, gp, n);
                         })) != MPM_OK)
                        return rc;
                }
                return MPM_OK;
            }
        }  // phase != PH_HEAD
        // the USF head's k_g2p2 leaves MVL to the body
        StepArgs<R> sg = sa;
        if (phase == PH_HEAD) sg.mvl = 0;
        auto g2p2 = [&](auto cm) {
            constexpr int CM = decltype(cm)::value;
            if (tiled)
                k_g2p2_tiled<DIM, S, TSH, R, CM><<<sort.tiled_grid(), 256, 0, s>>>(
                    geom, sort.tg, pp, gp, d_mats, sg, sort.tile_list, sort.seg_begin,
                    sort.seg_end, sort.tile_count, fl);
            else
                k_g2p2<DIM, S, R, CM><<<pb, kBlock, 0, s>>>(geom, pp, gp, d_mats, sg, n, fl);
        };
        if ((rc = c->launch(KID_G2P2, [&] {
                 if (ar.constitutive == MPM_DRUCKERPRAGER)
                     g2p2(std::integral_constant<int, 1>{});
                 else if (ar.constitutive == MPM_MOHRCOULOMB)
                     g2p2(std::integral_constant<int, 2>{});
                 else
                     g2p2(std::integral_constant<int, 0>{});
             })) != MPM_OK)
            return rc;
        if constexpr (std::is_same<R, float>::value) {
            if (sa.defer_return) {
                if ((rc = c->launch(KID_LEARNED, [&] {
                         mpm::mlp::k_mlp_forward<SurrogateIn, SurrogateOut>
                             <<<sur->forward_grid(np), 256, 0, s>>>(
                                 sur->sw, sur->P, np, SurrogateIn{pp, d_mats},
                                 SurrogateOut{pp, d_mats, sur_tol, d_sur_stats});
                     })) != MPM_OK)
                    return rc;
            }
        }
        if (ar.mvl && phase == PH_FULL) {
            if ((rc = c->launch(KID_MVL2, [&] {
                     k_mvl2<DIM, R><<<pb, kBlock, 0, s>>>(pp, gp, n);
                 })) != MPM_OK)
                return rc;
        }
        return MPM_OK;
    }

    // the kernels of one step, or of one part of a USF step (no sort)
    int kernels(mpm_ctx* c, const mpm_args& a, double dt, double Ti, int phase) {
        const bool lin = c->cfg.basis == MPM_LINEAR;
        if (dim == 2)
            return lin ? step_impl<2, 2>(c, a, dt, Ti, phase) : step_impl<2, 3>(c, a, dt, Ti, phase);
        return lin ? step_impl<3, 2>(c, a, dt, Ti, phase) : step_impl<3, 3>(c, a, dt, Ti, phase);
    }

    // the periodic cell sort, when due before step number c->steps
    int sort_if_due(mpm_ctx* c) {
        const int K = c->cfg.sort_interval;
        if (K > 0 && (c->resort || (c->steps % K) == 0)) {
            int rc = sort.run(c, pp, c->stream);
            if (rc != MPM_OK) return rc;
            c->resort = false;
        }
        return MPM_OK;
    }

    // One step.  USF: the head (stress update of the pending displacement and
    // position commit) runs before the sort, so no increment is pending while
    // the particles are permuted.
    int step(mpm_ctx* c, const mpm_args& a, double dt, double Ti) override {
        const bool usf = a.scheme == MPM_USF;
        int rc = usf ? kernels(c, a, dt, Ti, PH_HEAD) : MPM_OK;
        if (rc == MPM_OK) rc = sort_if_due(c);
        if (rc == MPM_OK) rc = kernels(c, a, dt, Ti, usf ? PH_BODY : PH_FULL);
        if (rc == MPM_OK) {
            c->steps += 1;
            usf_pending = usf;
        }
        return rc;
    }

    // Runs of steps between two sorts whose kernel arguments do not change
    // (Ti >= Te: full gravity, plasticity on) are captured once into a HIP
    // graph and replayed; sorts, ramp steps and runs of one step are eager.
    // The graph's pointers are those of the sort parity it was captured at.
    // USF: each run starts with an eager head (before the sort); the graph
    // holds body, then head + body for the following steps.  A context's
    // first USF step is eager.
    int run_graph(mpm_ctx* c, const mpm_args& a, double dt, double Ti, int64_t nsteps,
                  double* Ti_out) override {
        const int K = c->cfg.sort_interval;
        const bool usf = a.scheme == MPM_USF;
        const int body = usf ? PH_BODY : PH_FULL;
        int64_t k = 0;
        int rc = MPM_OK;
        while (k < nsteps && rc == MPM_OK) {
            const bool first_usf = usf && !usf_pending;
            if (usf && (rc = kernels(c, a, dt, Ti, PH_HEAD)) != MPM_OK) break;
            if ((rc = sort_if_due(c)) != MPM_OK) break;
            long long L = nsteps - k;
            if (K > 0) L = std::min<long long>(L, K - (c->steps % K));
            if (Ti < a.Te || L < 2 || c->prof_on || first_usf) {
                if ((rc = kernels(c, a, dt, Ti, body)) != MPM_OK) break;
                c->steps += 1;
                usf_pending = usf;
                Ti += dt;
                k += 1;
                continue;
            }
            const int tiled = (sort.enabled && sort.tiles_valid && c->use_tiles) ? 1 : 0;
            hipGraphExec_t exec = nullptr;
            for (auto& g : c->graphs)
                if (g.len == L && g.parity == sort.parity && g.tiled == tiled && g.dt == dt &&
                    std::memcmp(&g.args, &a, sizeof a) == 0)
                    exec = g.exec;
            if (!exec) {
                if (c->graphs.size() >= 8) c->drop_graphs();
                hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) return c->hip_fail(e, "hipStreamBeginCapture");
                double Tc = Ti;
                for (long long j = 0; j < L && rc == MPM_OK; ++j) {
                    if (usf && j > 0) rc = kernels(c, a, dt, Tc, PH_HEAD);
                    if (rc == MPM_OK) rc = kernels(c, a, dt, Tc, body);
                    Tc += dt;
                }
                hipGraph_t graph = nullptr;
                e = hipStreamEndCapture(c->stream, &graph);
                if (rc != MPM_OK) {
                    if (graph) hipGraphDestroy(graph);
                    break;
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                if (graph) hipGraphDestroy(graph);
                if (e != hipSuccess) return c->hip_fail(e, "step graph capture");
                c->graphs.push_back({L, sort.parity, tiled, a, dt, exec});
            }
            hipError_t e = hipGraphLaunch(exec, c->stream);
            if (e != hipSuccess) return c->hip_fail(e, "hipGraphLaunch");
            for (long long j = 0; j < L; ++j) Ti += dt;
            c->steps += L;
            k += L;
            usf_pending = usf;
        }
        if (Ti_out) *Ti_out = Ti;
        return rc;
    }

    template <int DIM, int S>
    int count_impl(mpm_ctx* c, int64_t* out) {
        unsigned char* mark = nullptr;
        unsigned long long* cnt = nullptr;
        hipError_t e = hipMallocAsync((void**)&mark, ni, c->stream);
        if (e == hipSuccess) e = hipMallocAsync((void**)&cnt, sizeof(unsigned long long), c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(mark, 0, ni, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, sizeof(unsigned long long), c->stream);
        if (e == hipSuccess) {
            k_mark_active<DIM, S, R><<<nblocks(np), kBlock, 0, c->stream>>>(geom, pp, mark, int(np));
            k_count_bytes<<<1024, 256, 0, c->stream>>>(mark, ni, cnt);
            e = hipGetLastError();
        }
        unsigned long long h = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream);
        if (mark) hipFreeAsync(mark, c->stream);
        if (cnt) hipFreeAsync(cnt, c->stream);
        hipError_t e2 = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = e2;
        if (e != hipSuccess) return c->hip_fail(e, "mpm_count_active_nodes");
        *out = (int64_t)h;
        return MPM_OK;
    }

    int set_surrogate(mpm_ctx* c, mpm_mlp* m, double tol) override {
        if (!std::is_same<R, float>::value && m)
            return c->fail(MPM_E_UNSUPPORTED, "mpm_set_surrogate: the learned return runs in FP32 mode only");
        if (m && dim != 3)
            return c->fail(MPM_E_UNSUPPORTED, "mpm_set_surrogate: the learned return is 3D Mohr-Coulomb only");
        if (m && !(tol > 0.0 && tol < 1.0)) return c->fail(MPM_E_ARG, "mpm_set_surrogate: tol outside (0, 1)");
        int rc;
        if (m && !have_de) {
            for (int k = 0; k < 6; ++k)
                if ((rc = alloc(c, &pp.de[k], np, "strain increment")) != MPM_OK) return rc;
            for (int k = 0; k < 6; ++k)
                if ((rc = alloc(c, &pp.pr[k], np, "trial principal stresses and frame")) != MPM_OK) return rc;
            have_de = true;
        }
        if (!d_sur_stats && (rc = alloc(c, &d_sur_stats, 2, "surrogate stats")) != MPM_OK) return rc;
        hipError_t e = hipMemsetAsync(d_sur_stats, 0, 2 * sizeof(unsigned long long), c->stream);
        if (e == hipSuccess && m) e = hipStreamSynchronize(m->stream);  // weights ready
        if (e != hipSuccess) return c->hip_fail(e, "mpm_set_surrogate");
        sur = m;
        sur_tol = float(tol);
        c->drop_graphs();
        return MPM_OK;
    }

    int surrogate_stats(mpm_ctx* c, int64_t* plastic, int64_t* fallbacks) override {
        unsigned long long h[2] = {0, 0};
        if (d_sur_stats) {
            hipError_t e = hipMemcpyAsync(h, d_sur_stats, sizeof h, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) return c->hip_fail(e, "mpm_surrogate_stats");
        }
        if (plastic) *plastic = int64_t(h[0]);
        if (fallbacks) *fallbacks = int64_t(h[1]);
        return MPM_OK;
    }

    int return_log(mpm_ctx* c, int64_t cap, int64_t stride) override {
        if (cap < 0 || stride < 1) return c->fail(MPM_E_ARG, "mpm_return_log: cap >= 0, stride >= 1");
        hipError_t e = hipStreamSynchronize(c->stream);
        if (log_x) hipFree(log_x);
        if (log_y) hipFree(log_y);
        log_x = log_y = nullptr;
        log_cap = 0;
        if (e == hipSuccess && cap > 0) {
            e = hipMalloc((void**)&log_x, sizeof(float) * 16 * cap);
            if (e == hipSuccess) e = hipMalloc((void**)&log_y, sizeof(float) * 7 * cap);
        }
        int rc;
        if (!d_log_count && (rc = alloc(c, &d_log_count, 1, "log counter")) != MPM_OK) return rc;
        if (e == hipSuccess) e = hipMemset(d_log_count, 0, sizeof(unsigned long long));
        if (e != hipSuccess) return c->hip_fail(e, "mpm_return_log");
        log_cap = cap;
        log_stride = int(stride);
        c->drop_graphs();
        return MPM_OK;
    }

    int return_log_read(mpm_ctx* c, float* x, float* y, int64_t max_rows, int64_t* rows) override {
        unsigned long long cnt = 0;
        hipError_t e = hipSuccess;
        if (d_log_count) {
            e = hipMemcpyAsync(&cnt, d_log_count, sizeof cnt, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        }
        long long n = (long long)cnt < log_cap ? (long long)cnt : log_cap;
        if (n > max_rows) n = max_rows;
        if (e == hipSuccess && n > 0 && x) e = hipMemcpy(x, log_x, sizeof(float) * 16 * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess && n > 0 && y) e = hipMemcpy(y, log_y, sizeof(float) * 7 * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return c->hip_fail(e, "mpm_return_log_read");
        if (rows) *rows = n;
        return MPM_OK;
    }

    int count_active(mpm_ctx* c, int64_t* out) override {
        const bool lin = c->cfg.basis == MPM_LINEAR;
        if (dim == 2) return lin ? count_impl<2, 2>(c, out) : count_impl<2, 3>(c, out);
        return lin ? count_impl<3, 2>(c, out) : count_impl<3, 3>(c, out);
    }

    // The same count with the node bitmap and the counter in the file-scope
    // scratch array: clear (a fill kernel, not hipMemsetAsync, so the calls
    // stay capturable), mark, count, read back.  No allocation.
    template <int DIM, int S>
    int count_noalloc_impl(mpm_ctx* c, int64_t* out) {
        if (ni > kNoAllocMaxNodes)
            return c->fail(MPM_E_RANGE, "mpm_count_active_nodes_noalloc: " + std::to_string(ni) +
                                            " grid nodes exceed its maximum of " +
                                            std::to_string(kNoAllocMaxNodes));
        if (!c->scratch_ptr) return c->fail(MPM_E_STATE, "mpm_count_active_nodes_noalloc: no scratch address");
        unsigned* words = static_cast<unsigned*>(c->scratch_ptr);
        unsigned long long* cnt = reinterpret_cast<unsigned long long*>(words + kNoAllocWords);
        const long long nw = (ni + 31) / 32;
        k_fill_u32<<<unsigned((nw + 255) / 256), 256, 0, c->stream>>>(words, nw);
        k_fill_u32<<<1, 256, 0, c->stream>>>(words + kNoAllocWords, 2);
        k_mark_active_bits<DIM, S, R><<<nblocks(np), kBlock, 0, c->stream>>>(geom, pp, words, int(np));
        k_count_bits<<<1024, 256, 0, c->stream>>>(words, nw, cnt);
        hipError_t e = hipGetLastError();
        unsigned long long h = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream);
        hipError_t e2 = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = e2;
        if (e != hipSuccess) return c->hip_fail(e, "mpm_count_active_nodes_noalloc");
        *out = (int64_t)h;
        return MPM_OK;
    }

    int count_active_noalloc(mpm_ctx* c, int64_t* out) override {
        const bool lin = c->cfg.basis == MPM_LINEAR;
        if (dim == 2) return lin ? count_noalloc_impl<2, 2>(c, out) : count_noalloc_impl<2, 3>(c, out);
        return lin ? count_noalloc_impl<3, 2>(c, out) : count_noalloc_impl<3, 3>(c, out);
    }

    template <typename T>
    int fetch(mpm_ctx* c, const T* src, double* host, const char* what) {
        const int n = int(np);
        k_unpermute<T><<<nblocks(np), kBlock, 0, c->stream>>>(src, pp.pid, d_stage, n);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(host, d_stage, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        return e == hipSuccess ? MPM_OK : c->hip_fail(e, what);
    }

    int download(mpm_ctx* c, double* xi, double* vs, double* sigma, double* sigma_m,
                 double* omega, double* eps_q, double* eps_k, double* eps_v, double* eps_s,
                 double* ms) override {
        int rc;
        for (int a = 0; a < dim; ++a) {
            if (xi && usf_pending) {
                // a USF position increment is committed by the next step's head
                k_unpermute_add<R><<<nblocks(np), kBlock, 0, c->stream>>>(pp.x[a], pp.dxp[a],
                                                                          pp.pid, d_stage, int(np));
                hipError_t e = hipGetLastError();
                if (e == hipSuccess)
                    e = hipMemcpyAsync(xi + a * np, d_stage, sizeof(double) * np,
                                       hipMemcpyDeviceToHost, c->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
                if (e != hipSuccess) return c->hip_fail(e, "download xi");
            } else if (xi && (rc = fetch(c, pp.x[a], xi + a * np, "download xi")) != MPM_OK) {
                return rc;
            }
            if (vs && (rc = fetch(c, pp.v[a], vs + a * np, "download vs")) != MPM_OK) return rc;
        }
        for (int k = 0; k < ns; ++k) {
            if (sigma && (rc = fetch(c, pp.sig[k], sigma + k * np, "download sigma")) != MPM_OK)
                return rc;
            if (eps_s && (rc = fetch(c, pp.eps[k], eps_s + k * np, "download eps_s")) != MPM_OK)
                return rc;
        }
        if (sigma_m) {
            k_sigm<R><<<nblocks(np), kBlock, 0, c->stream>>>(pp.sig[0], pp.sig[1], pp.sig[2],
                                                             pp.pid, d_stage, int(np));
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(sigma_m, d_stage, sizeof(double) * np, hipMemcpyDeviceToHost,
                                   c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) return c->hip_fail(e, "download sigma_m");
        }
        if (omega && (rc = fetch(c, pp.vol, omega, "download omega")) != MPM_OK) return rc;
        if (eps_q && (rc = fetch(c, pp.eq, eps_q, "download eps_q")) != MPM_OK) return rc;
        if (eps_k && (rc = fetch(c, pp.ek, eps_k, "download eps_k")) != MPM_OK) return rc;
        if (eps_v && (rc = fetch(c, pp.ev, eps_v, "download eps_v")) != MPM_OK) return rc;
        if (ms && (rc = fetch(c, pp.ms, ms, "download ms")) != MPM_OK) return rc;
        return MPM_OK;
    }
};

int validate(const mpm_config* c, std::string& why) {
    if (!c) { why = "NULL config"; return MPM_E_ARG; }
    if (c->dim != 2 && c->dim != 3) { why = "dim must be 2 or 3"; return MPM_E_ARG; }
    if (c->precision != MPM_FP64 && c->precision != MPM_FP32) { why = "precision must be MPM_FP64 or MPM_FP32"; return MPM_E_ARG; }
    if (c->basis != MPM_UGIMP && c->basis != MPM_LINEAR) { why = "basis must be uGIMP or linear (gslinear has no kernels in the reference either)"; return MPM_E_UNSUPPORTED; }
    if (c->np <= 0 || c->np > 0x7fffffffLL) { why = "np must be in 1..2^31-1"; return MPM_E_ARG; }
    if (c->nmat <= 0) { why = "nmat must be positive"; return MPM_E_ARG; }
    if (c->nnx < 3 || c->nny < 3 || (c->dim == 3 && c->nnz < 3)) { why = "grid needs at least 3 nodes per axis"; return MPM_E_ARG; }
    double ni = double(c->nnx) * double(c->nny) * double(c->dim == 3 ? c->nnz : 1);
    if (ni > 2.0e9) { why = "grid larger than 2e9 nodes"; return MPM_E_ARG; }
    if (!(c->dx > 0 && c->dy > 0 && (c->dim == 2 || c->dz > 0))) { why = "grid spacing must be positive"; return MPM_E_ARG; }
    if (c->basis == MPM_UGIMP) {
        if (!(c->lpx > 0 && c->lpy > 0 && (c->dim == 2 || c->lpz > 0))) { why = "particle spacing must be positive"; return MPM_E_ARG; }
        if (c->lpx > c->dx || c->lpy > c->dy || (c->dim == 3 && c->lpz > c->dz)) { why = "uGIMP needs particle spacing <= grid spacing"; return MPM_E_ARG; }
    }
    if (c->sort_interval < 0) { why = "sort_interval must be >= 0"; return MPM_E_ARG; }
    return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_abi_version(void) { return MPM_ABI_VERSION; }

int mpm_create(const mpm_config* cfg, mpm_ctx** out) {
    if (!out) {
        g_create_error = "mpm_create: NULL out";
        return MPM_E_ARG;
    }
    *out = nullptr;
    std::string why;
    int rc = validate(cfg, why);
    if (rc != MPM_OK) {
        g_create_error = "mpm_create: " + why;
        return rc;
    }
    mpm_ctx* c = new mpm_ctx();
    c->cfg = *cfg;
    {
        hipError_t e0 = hipGetDevice(&c->device);
        if (e0 == hipSuccess) e0 = scratch_address(c->device, &c->scratch_ptr);
        if (e0 != hipSuccess) {
            g_create_error = std::string("scratch address: ") + hipGetErrorString(e0);
            delete c;
            return MPM_E_HIP;
        }
    }
    if (c->cfg.dim == 2) {
        c->cfg.nnz = 1;
        c->cfg.z1 = 0.0;
        c->cfg.dz = 1.0;
        c->cfg.lpz = 1.0;
    }
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_create_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        delete c;
        return MPM_E_HIP;
    }
    e = hipMalloc((void**)&c->d_flags, sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->d_flags, 0, sizeof(int));
    if (e != hipSuccess) {
        g_create_error = std::string("hipMalloc flags: ") + hipGetErrorString(e);
        hipStreamDestroy(c->stream);
        delete c;
        return MPM_E_HIP;
    }
    if (c->cfg.precision == MPM_FP64) {
        auto* im = new Impl<double>();
        c->impl = im;
        rc = im->init(c);
    } else {
        auto* im = new Impl<float>();
        c->impl = im;
        rc = im->init(c);
    }
    if (rc != MPM_OK) {
        g_create_error = "mpm_create: " + c->err;
        mpm_destroy(c);
        return rc;
    }
    *out = c;
    return MPM_OK;
}

int mpm_upload_particles(mpm_ctx* c, const double* xi, const double* vs, const double* sigma,
                         const double* eps_s, const double* F, const double* omega,
                         const double* omega0, const double* rho, const double* rho0,
                         const double* eps_q, const double* eps_k, const int64_t* nid) {
    if (!c) return MPM_E_ARG;
    int rc = c->impl->upload_particles(c, xi, vs, sigma, eps_s, F, omega, omega0, rho, rho0,
                                       eps_q, eps_k, nid);
    if (rc == MPM_OK) c->have_particles = true;
    return rc;
}

int mpm_upload_materials(mpm_ctx* c, const double* Ks, const double* Gs, const double* phi,
                         const double* psi, const double* coh, const double* sigma_t,
                         const double* Hp, const double* cr, const double* phir) {
    if (!c) return MPM_E_ARG;
    int rc = c->impl->upload_materials(c, Ks, Gs, phi, psi, coh, sigma_t, Hp, cr, phir);
    if (rc == MPM_OK) c->have_materials = true;
    return rc;
}

int mpm_upload_bc(mpm_ctx* c, const int64_t* vx_idx, const double* vx_val, const int64_t* vy_idx,
                  const double* vy_val, const int64_t* vz_idx, const double* vz_val) {
    if (!c) return MPM_E_ARG;
    int rc = c->impl->upload_bc(c, vx_idx, vx_val, vy_idx, vy_val, vz_idx, vz_val);
    if (rc == MPM_OK) c->have_bc = true;
    return rc;
}

static int mpm_step_check(mpm_ctx* c, const mpm_args* a, double dt) {
    if (!a) return c->fail(MPM_E_ARG, "mpm_step: NULL args");
    if (!c->have_particles || !c->have_materials || !c->have_bc)
        return c->fail(MPM_E_STATE, "mpm_step: upload particles, materials and bc first");
    if (!(dt > 0.0) || !std::isfinite(dt)) return c->fail(MPM_E_ARG, "mpm_step: dt must be positive and finite");
    if (a->scheme != MPM_MUSL && a->scheme != MPM_USL && a->scheme != MPM_USF)
        return c->fail(MPM_E_UNSUPPORTED, "mpm_step: scheme must be MUSL, USL or USF");
    if (a->va != MPM_VA_A && a->va != MPM_VA_V) return c->fail(MPM_E_ARG, "mpm_step: va must be :a or :v");
    if (a->constitutive < 0 || a->constitutive > 2)
        return c->fail(MPM_E_UNSUPPORTED, "mpm_step: constitutive must be linearelastic, druckerprager or mohrcoulomb");
    return MPM_OK;
}

int mpm_step(mpm_ctx* c, const mpm_args* a, double dt, double Ti) {
    if (!c) return MPM_E_ARG;
    int rc = mpm_step_check(c, a, dt);
    if (rc != MPM_OK) return rc;
    return c->impl->step(c, *a, dt, Ti);
}

int mpm_run(mpm_ctx* c, const mpm_args* a, double dt, double Ti, int64_t nsteps, double* Ti_out) {
    if (!c) return MPM_E_ARG;
    if (nsteps < 0) return c->fail(MPM_E_ARG, "mpm_run: nsteps < 0");
    for (int64_t k = 0; k < nsteps; ++k) {
        int rc = mpm_step(c, a, dt, Ti);
        if (rc != MPM_OK) return rc;
        Ti += dt;
    }
    if (Ti_out) *Ti_out = Ti;
    return MPM_OK;
}

int mpm_run_graph(mpm_ctx* c, const mpm_args* a, double dt, double Ti, int64_t nsteps,
                  double* Ti_out) {
    if (!c) return MPM_E_ARG;
    if (nsteps < 0) return c->fail(MPM_E_ARG, "mpm_run_graph: nsteps < 0");
    if (nsteps == 0) {
        if (Ti_out) *Ti_out = Ti;
        return MPM_OK;
    }
    // the same argument checks as mpm_step, once
    int rc = mpm_step_check(c, a, dt);
    if (rc != MPM_OK) return rc;
    return c->impl->run_graph(c, *a, dt, Ti, nsteps, Ti_out);
}

int mpm_sync(mpm_ctx* c) {
    if (!c) return MPM_E_ARG;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "mpm_sync");
    int flags = 0;
    e = hipMemcpy(&flags, c->d_flags, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "mpm_sync flags");
    if (flags) {
        hipMemset(c->d_flags, 0, sizeof(int));
        std::string what;
        if (flags & FLAG_RANGE) what += "a particle's basis stencil left the grid (pad the grid or enlarge it)";
        if (flags & FLAG_NONFINITE) what += std::string(what.empty() ? "" : "; ") + "a particle position became non-finite";
        return c->fail(MPM_E_RANGE, "mpm_sync: " + what + "; such particles are frozen");
    }
    return MPM_OK;
}

int mpm_download_particles(mpm_ctx* c, double* xi, double* vs, double* sigma, double* sigma_m,
                           double* omega, double* eps_q, double* eps_k, double* eps_v,
                           double* eps_s, double* ms) {
    if (!c) return MPM_E_ARG;
    if (!c->have_particles) return c->fail(MPM_E_STATE, "mpm_download_particles: nothing uploaded");
    return c->impl->download(c, xi, vs, sigma, sigma_m, omega, eps_q, eps_k, eps_v, eps_s, ms);
}

void mpm_destroy(mpm_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    c->drop_graphs();
    delete c->impl;
    for (auto& r : c->recs) {
        hipEventDestroy(r.e0);
        hipEventDestroy(r.e1);
    }
    for (auto e : c->pool) hipEventDestroy(e);
    if (c->d_flags) hipFree(c->d_flags);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char* mpm_last_error(const mpm_ctx* c) {
    return c ? c->err.c_str() : g_create_error.c_str();
}

int mpm_profile_enable(mpm_ctx* c, int on) {
    if (!c) return MPM_E_ARG;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "mpm_profile_enable");
    for (auto& r : c->recs) {
        c->pool.push_back(r.e0);
        c->pool.push_back(r.e1);
    }
    c->recs.clear();
    c->prof_on = on != 0;
    return MPM_OK;
}

int mpm_profile_read(mpm_ctx* c, double* total_ms, int64_t* launches, int n) {
    if (!c || n < 0) return MPM_E_ARG;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "mpm_profile_read");
    for (int k = 0; k < n; ++k) {
        if (total_ms) total_ms[k] = 0.0;
        if (launches) launches[k] = 0;
    }
    for (auto& r : c->recs) {
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, r.e0, r.e1);
        if (e != hipSuccess) return c->hip_fail(e, "hipEventElapsedTime");
        if (r.kid < n) {
            if (total_ms) total_ms[r.kid] += ms;
            if (launches) launches[r.kid] += 1;
        }
    }
    return MPM_OK;
}

const char* mpm_profile_kernel_name(int k) {
    return (k >= 0 && k < kNumKernels) ? kKernelNames[k] : "";
}

int mpm_profile_kernel_count(void) { return kNumKernels; }

int mpm_set_option(mpm_ctx* c, int option, int64_t value) {
    if (!c) return MPM_E_ARG;
    if (option == MPM_OPT_SORT_INTERVAL) {
        if (c->cfg.sort_interval <= 0)
            return c->fail(MPM_E_STATE, "mpm_set_option: the context was created without sorting");
        if (value < 1 || value > (1 << 30)) return c->fail(MPM_E_ARG, "mpm_set_option: sort interval >= 1");
        c->cfg.sort_interval = int32_t(value);
    } else if (option == MPM_OPT_TILES) {
        c->use_tiles = value != 0;
    } else if (option == MPM_OPT_FUSED_REMAP) {
        c->fused_remap = value != 0;
    } else if (option == MPM_OPT_TILE_PARTIALS) {
        c->tile_partials = value != 0;
    } else if (option == MPM_OPT_SCATTER) {
        if (value != -1 && value != MPM_SCATTER_ROUNDS && value != MPM_SCATTER_OWNED &&
            value != MPM_SCATTER_OWNED_LAYERS && value != MPM_SCATTER_ROUNDS_LAYERED &&
            value != MPM_SCATTER_PAIRS)
            return c->fail(MPM_E_ARG, "mpm_set_option: scatter form is -1 (automatic), "
                                      "MPM_SCATTER_ROUNDS, MPM_SCATTER_OWNED, MPM_SCATTER_OWNED_LAYERS, "
                                      "MPM_SCATTER_ROUNDS_LAYERED or MPM_SCATTER_PAIRS");
        c->scatter = int(value);
    } else if (option == MPM_OPT_REMAP_SCATTER) {
        if (value != -1 && value != MPM_SCATTER_ROUNDS && value != MPM_SCATTER_OWNED &&
            value != MPM_SCATTER_OWNED_LAYERS && value != MPM_SCATTER_ROUNDS_LAYERED &&
            value != MPM_SCATTER_PAIRS)
            return c->fail(MPM_E_ARG, "mpm_set_option: re-map scatter form is -1 (automatic) or "
                                      "an MPM_SCATTER_* form");
        c->remap_scatter = int(value);
    } else if (option == MPM_OPT_PROFILE_MASK) {
        if (value < 0 || (kNumKernels < 63 && value >= (int64_t(1) << kNumKernels)))
            return c->fail(MPM_E_ARG, "mpm_set_option: profile mask names no such kernel");
        c->prof_mask = uint64_t(value);
        return MPM_OK;
    } else {
        return c->fail(MPM_E_ARG, "mpm_set_option: unknown option");
    }
    c->drop_graphs();
    return MPM_OK;
}

int mpm_set_surrogate(mpm_ctx* c, mpm_mlp* m, double tol) {
    if (!c) return MPM_E_ARG;
    return c->impl->set_surrogate(c, m, tol);
}

int mpm_surrogate_stats(mpm_ctx* c, int64_t* plastic, int64_t* fallbacks) {
    if (!c) return MPM_E_ARG;
    return c->impl->surrogate_stats(c, plastic, fallbacks);
}

int mpm_return_log(mpm_ctx* c, int64_t capacity, int64_t stride) {
    if (!c) return MPM_E_ARG;
    return c->impl->return_log(c, capacity, stride);
}

int mpm_return_log_read(mpm_ctx* c, float* x, float* y, int64_t max_rows, int64_t* rows) {
    if (!c || max_rows < 0) return MPM_E_ARG;
    return c->impl->return_log_read(c, x, y, max_rows, rows);
}

int mpm_count_active_nodes(mpm_ctx* c, int64_t* count) {
    if (!c || !count) return MPM_E_ARG;
    if (!c->have_particles) return c->fail(MPM_E_STATE, "mpm_count_active_nodes: nothing uploaded");
    return c->impl->count_active(c, count);
}

int mpm_count_active_nodes_noalloc(mpm_ctx* c, int64_t* count) {
    if (!c || !count) return MPM_E_ARG;
    if (!c->have_particles) return c->fail(MPM_E_STATE, "mpm_count_active_nodes_noalloc: nothing uploaded");
    return c->impl->count_active_noalloc(c, count);
}

int mpm_tpeak(int64_t nbytes, int reps, double* gbps) {
    if (nbytes < 1024 || reps < 1 || !gbps) return MPM_E_ARG;
    long long n4 = nbytes / 16;
    float4 *a = nullptr, *b = nullptr, *cc = nullptr;
    hipError_t e = hipMalloc((void**)&a, n4 * 16);
    if (e == hipSuccess) e = hipMalloc((void**)&b, n4 * 16);
    if (e == hipSuccess) e = hipMalloc((void**)&cc, n4 * 16);
    if (e == hipSuccess) e = hipMemset(b, 0, n4 * 16);
    if (e == hipSuccess) e = hipMemset(cc, 0, n4 * 16);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    double best = 0.0;
    const long long nb = (n4 + 255) / 256;
    if (nb > 0x7fffffffll) e = hipErrorInvalidValue;
    const unsigned blocks = unsigned(nb);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        hipEventRecord(e0, 0);
        k_triad<<<blocks, 256, 0, 0>>>(b, cc, a, n4);
        hipEventRecord(e1, 0);
        e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        double g = 3.0 * double(n4) * 16.0 / (double(ms) * 1e-3) / 1e9;
        if (g > best) best = g;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (a) hipFree(a);
    if (b) hipFree(b);
    if (cc) hipFree(cc);
    if (e != hipSuccess) return MPM_E_HIP;
    *gbps = best;
    return MPM_OK;
}

}  // extern "C"
