// mpm_surrogate.hpp -- host side of the learned return map (mpm_mlp.hpp):
// parameter storage, training on the device (minibatch Adam over MFMA GEMMs),
// gradient and forward entry points for the checks, timing.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/mpm_hip.h"
#include "mpm_kernels.hpp"
#include "mpm_mlp.hpp"

namespace mpm {

// Network inputs of particle p after a deferred 3D Mohr-Coulomb step: the
// mc_features of its trial stress (principal values and frame, stored by
// G2P #2), strain increment, plastic strain and strength in effect.  The
// kernel reads them in two halves of 8.
struct SurrogateIn {
    PartPtrs<float> pp;
    const Material<float>* mats;
    __device__ void load(long long p, int half, float* f) const {
        float pr[3], q3[3], de[6], x[16];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            pr[k] = pp.pr[k][p];
            q3[k] = pp.pr[3 + k][p];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) de[k] = pp.de[k][p];
        const Material<float>& M = mats[pp.nid[p]];
        const float eq = pp.eq[p];
        mc_features<float>(pr, q3, de, eq, mc_strength<float>(M, eq), M.Gs, x);
#pragma unroll
        for (int k = 0; k < 8; ++k) f[k] = x[8 * half + k];
    }
};

// The learned return of particle p.  The network scores the four branches
// of the principal-space return (main plane, edge s1 = s2, edge s2 = s3,
// apex); a yielding trial stress takes the highest-scoring branch when that
// branch's own conditions hold (mc3_principal_branch: the returned stress is
// that branch's closed form, on the Mohr-Coulomb surface) and its yield
// residual |f| <= tol (2 c cos(phi) + |s1 + s3| sin(phi)); otherwise the
// exact return (mc3_principal) decides.  The stress is rebuilt in the frame
// G2P #2 stored (no eigen-solve here), and the plastic strains follow from
// the stress correction as in mcP3.  Elastic trial stresses stay (ϵk = 0).
// stats[0] counts yielding particles, stats[1] fallbacks to the exact return.
struct SurrogateOut {
    PartPtrs<float> pp;
    const Material<float>* mats;
    float tol;
    unsigned long long* stats;
    __device__ void store(long long p, const float* y) const {
        float pr[3], q3[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            pr[k] = pp.pr[k][p];
            q3[k] = pp.pr[3 + k][p];
        }
        const Material<float> M = mats[pp.nid[p]];
        float eq = pp.eq[p], ek = 0.0f;
        const MCStrength<float> P = mc_strength<float>(M, eq);
        const float k2 = 2.0f * P.c * P.cosphi;
        const float f = (pr[0] - pr[2]) + (pr[0] + pr[2]) * P.sinphi - k2;
        const bool plastic = f > 0.0f;
        bool fallback = false;
        if (plastic) {
            int kind = 1;
            float best = y[0];
#pragma unroll
            for (int k = 1; k < 4; ++k)
                if (y[k] > best) {
                    best = y[k];
                    kind = k + 1;
                }
            float pn[3] = {pr[0], pr[1], pr[2]};
            bool ok = isfinite(best) &&
                      mc3_principal_branch<float>(pn, kind, M.Dt, M.Dd, P.sinphi, P.cosphi, P.sinpsi, P.c);
            if (ok) {
                const float F = (pn[0] - pn[2]) + (pn[0] + pn[2]) * P.sinphi - k2;
                float scale = k2 + fabsf(pn[0] + pn[2]) * P.sinphi;
                scale = fmaxf(scale, 1e-6f * (fabsf(pn[0]) + fabsf(pn[2])) + 1e-30f);
                ok = fabsf(F) <= tol * scale;
            }
            if (!ok) {
                fallback = true;
                pn[0] = pr[0];
                pn[1] = pr[1];
                pn[2] = pr[2];
                mc3_principal<float>(pn, M.Dt, M.Dd, P.sinphi, P.cosphi, P.sinpsi, P.c);
            }
            float V[3][3];
            eigvecs_from_quat<float>(q3, V);
            const int ri[6] = {0, 1, 2, 0, 1, 0}, ci[6] = {0, 1, 2, 1, 2, 2};
            float d[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const float sn = pn[0] * V[ri[k]][0] * V[ci[k]][0] + pn[1] * V[ri[k]][1] * V[ci[k]][1] +
                                 pn[2] * V[ri[k]][2] * V[ci[k]][2];
                const float s0 = pp.sig[k][p];
                d[k] = sn - s0;
                pp.sig[k][p] = sn;
            }
            mc_strain<float>(d, eq, ek, M);
            pp.eq[p] = eq;
        }
        pp.ek[p] = ek;
        const unsigned long long np_ = __popcll(__ballot(plastic));
        const unsigned long long nf_ = __popcll(__ballot(fallback));
        if ((threadIdx.x & 63) == 0) {
            if (np_) atomicAdd(&stats[0], np_);
            if (nf_) atomicAdd(&stats[1], nf_);
        }
    }
};

}  // namespace mpm

struct mpm_mlp {
    std::string error;
    hipStream_t stream = nullptr;
    int cap = 0;                 // batch capacity (rows)
    float *P = nullptr, *G = nullptr, *m = nullptr, *v = nullptr;
    mpm::mlp::bf16x8* sw = nullptr;
    __bf16* wb = nullptr;        // BF16 row-major copy of the parameters
    __bf16 *xb = nullptr, *h1 = nullptr, *h2 = nullptr, *h3 = nullptr;
    __bf16 *d1 = nullptr, *d2 = nullptr, *d3 = nullptr, *d4 = nullptr;
    float *z1 = nullptr, *z2 = nullptr, *z3 = nullptr, *o = nullptr, *yb = nullptr;
    float *dsum = nullptr, *loss = nullptr;
    float* csum = nullptr;       // column-sum partials (kColsumBlocks x NH)
    float* splitk_ws = nullptr;  // split-K slices of a weight gradient (kSplitK x NH x NH)
    static constexpr int kSplitK = 32;
    float *X = nullptr, *Y = nullptr;  // dataset rows on the device
    long long ndata = 0, data_cap = 0;
    long long adam_t = 0;
    std::vector<void*> owned;

    int fail(int code, const std::string& msg) {
        error = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        error = std::string(what) + ": " + hipGetErrorString(e);
        return MPM_E_HIP;
    }
    template <typename T>
    hipError_t alloc(T** p, size_t count) {
        hipError_t e = hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) owned.push_back((void*)*p);
        return e;
    }
    void release() {
        for (void* p : owned) (void)hipFree(p);
        owned.clear();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    int init(int capacity) {
        using namespace mpm::mlp;
        cap = capacity;
        hipError_t e = hipStreamCreate(&stream);
        if (e == hipSuccess) e = alloc(&P, NTOTAL);
        if (e == hipSuccess) e = alloc(&G, NPARAM);
        if (e == hipSuccess) e = alloc(&m, NPARAM);
        if (e == hipSuccess) e = alloc(&v, NPARAM);
        if (e == hipSuccess) e = alloc(&sw, SWN);
        if (e == hipSuccess) e = alloc(&wb, NPARAM);
        if (e == hipSuccess) e = alloc(&xb, (size_t)cap * NIN);
        __bf16** hb[] = {&h1, &h2, &h3, &d1, &d2, &d3};
        for (auto p : hb)
            if (e == hipSuccess) e = alloc(p, (size_t)cap * NH);
        if (e == hipSuccess) e = alloc(&d4, (size_t)cap * NOUT);
        float** zf[] = {&z1, &z2, &z3};
        for (auto p : zf)
            if (e == hipSuccess) e = alloc(p, (size_t)cap * NH);
        if (e == hipSuccess) e = alloc(&o, (size_t)cap * NOUT);
        if (e == hipSuccess) e = alloc(&yb, (size_t)cap * NOUT);
        if (e == hipSuccess) e = alloc(&dsum, (size_t)cap * NOUT);
        if (e == hipSuccess) e = alloc(&loss, 1);
        if (e == hipSuccess) e = alloc(&csum, (size_t)mpm::mlp::kColsumBlocks * NH);
        if (e == hipSuccess) e = alloc(&splitk_ws, (size_t)kSplitK * NH * NH);
        if (e == hipSuccess) e = hipMemset(P, 0, sizeof(float) * NTOTAL);
        if (e == hipSuccess) e = hipMemset(m, 0, sizeof(float) * NPARAM);
        if (e == hipSuccess) e = hipMemset(v, 0, sizeof(float) * NPARAM);
        if (e != hipSuccess) return hip_fail(e, "mpm_mlp_create");
        return MPM_OK;
    }

    // swizzled and BF16 copies of the current parameters
    hipError_t prepare() {
        using namespace mpm::mlp;
        k_mlp_swizzle<<<(SWN + 255) / 256, 256, 0, stream>>>(P, sw);
        k_weights_bf16<<<(NPARAM + 255) / 256, 256, 0, stream>>>(P, wb);
        return hipGetLastError();
    }

    int upload_data(const float* x, const float* y, long long n) {
        using namespace mpm::mlp;
        if (n > data_cap) {
            if (X) (void)hipFree(X);
            if (Y) (void)hipFree(Y);
            X = Y = nullptr;
            hipError_t e = hipMalloc((void**)&X, sizeof(float) * n * NIN);
            if (e == hipSuccess) e = hipMalloc((void**)&Y, sizeof(float) * n * NOUT);
            if (e != hipSuccess) return hip_fail(e, "dataset allocation");
            data_cap = n;
        }
        hipError_t e = hipMemcpyAsync(X, x, sizeof(float) * n * NIN, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess && y)
            e = hipMemcpyAsync(Y, y, sizeof(float) * n * NOUT, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return hip_fail(e, "dataset upload");
        ndata = n;
        return MPM_OK;
    }

    template <bool TA, bool TB, int EPI>
    void gemm(int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb,
              mpm::mlp::GemmEpi ep) {
        using namespace mpm::mlp;
        dim3 grid((N + 31) / 32, (M + 31) / 32);
        // the weight gradients (few output tiles, K = the batch): split K over
        // kSplitK slices so that the whole chip works on them
        if constexpr (EPI == EPI_STORE) {
            if (K >= 1024 && (long long)M * N <= (long long)NH * NH) {
                const int kc = ((K + kSplitK - 1) / kSplitK + 15) / 16 * 16;
                const int slices = (K + kc - 1) / kc;
                k_gemm_splitk<TA, TB><<<dim3(grid.x, grid.y, slices), 64, 0, stream>>>(
                    M, N, K, kc, A, lda, B, ldb, splitk_ws);
                k_splitk_reduce<<<(M * N + 255) / 256, 256, 0, stream>>>(splitk_ws, slices, M, N, ep.c,
                                                                          ep.ldc);
                return;
            }
        }
        k_gemm<TA, TB, EPI><<<grid, 64, 0, stream>>>(M, N, K, A, lda, B, ldb, ep);
    }

    // forward + backward on the B rows in xb / yb; gradient into G, loss added to *lossp
    hipError_t forward_backward(int B, float* lossp) {
        using namespace mpm::mlp;
        const __bf16 *W1 = wb + OW1, *W2 = wb + OW2, *W3 = wb + OW3, *W4 = wb + OW4;
        GemmEpi ep{};
        ep = GemmEpi{P + OB1, z1, h1, nullptr, 0};
        gemm<false, true, EPI_HIDDEN>(B, NH, NIN, xb, NIN, W1, NIN, ep);
        ep = GemmEpi{P + OB2, z2, h2, nullptr, 0};
        gemm<false, true, EPI_HIDDEN>(B, NH, NH, h1, NH, W2, NH, ep);
        ep = GemmEpi{P + OB3, z3, h3, nullptr, 0};
        gemm<false, true, EPI_HIDDEN>(B, NH, NH, h2, NH, W3, NH, ep);
        ep = GemmEpi{P + OB4, nullptr, nullptr, o, 0};
        gemm<false, true, EPI_OUTPUT>(B, NOUT, NH, h3, NH, W4, NH, ep);
        const int no = B * NOUT;
        k_mse<<<(no + 255) / 256, 256, 0, stream>>>(o, yb, no, d4, dsum, lossp);
        // layer 4
        ep = GemmEpi{nullptr, nullptr, nullptr, G + OW4, NH};
        gemm<true, false, EPI_STORE>(NOUT, NH, B, d4, NOUT, h3, NH, ep);
        colsum<float>(dsum, B, NOUT, G + OB4);
        ep = GemmEpi{nullptr, z3, d3, nullptr, 0};
        gemm<false, false, EPI_DELTA>(B, NH, NOUT, d4, NOUT, W4, NH, ep);
        // layer 3
        ep = GemmEpi{nullptr, nullptr, nullptr, G + OW3, NH};
        gemm<true, false, EPI_STORE>(NH, NH, B, d3, NH, h2, NH, ep);
        colsum<__bf16>(d3, B, NH, G + OB3);
        ep = GemmEpi{nullptr, z2, d2, nullptr, 0};
        gemm<false, false, EPI_DELTA>(B, NH, NH, d3, NH, W3, NH, ep);
        // layer 2
        ep = GemmEpi{nullptr, nullptr, nullptr, G + OW2, NH};
        gemm<true, false, EPI_STORE>(NH, NH, B, d2, NH, h1, NH, ep);
        colsum<__bf16>(d2, B, NH, G + OB2);
        ep = GemmEpi{nullptr, z1, d1, nullptr, 0};
        gemm<false, false, EPI_DELTA>(B, NH, NH, d2, NH, W2, NH, ep);
        // layer 1
        ep = GemmEpi{nullptr, nullptr, nullptr, G + OW1, NIN};
        gemm<true, false, EPI_STORE>(NH, NIN, B, d1, NH, xb, NIN, ep);
        colsum<__bf16>(d1, B, NH, G + OB1);
        return hipGetLastError();
    }

    template <typename T>
    void colsum(const T* m, int rows, int cols, float* g) {
        using namespace mpm::mlp;
        const int nb = rows < 256 ? 1 : kColsumBlocks;
        k_colsum_part<T><<<nb, 256, 0, stream>>>(m, rows, cols, csum);
        k_colsum_final<<<1, 256, 0, stream>>>(csum, nb, cols, g);
    }

    void batch(int B, unsigned long long seed, long long step, bool sequential) {
        mpm::mlp::k_batch<<<(B + 255) / 256, 256, 0, stream>>>(X, Y, ndata, B, seed, step, P, xb,
                                                                 yb, sequential ? 1 : 0);
    }

    unsigned forward_grid(long long n) const {
        long long g = (n + 255) / 256;
        if (g > 512) g = 512;  // two resident workgroups per CU
        return unsigned(g < 1 ? 1 : g);
    }
};

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int mpm_mlp_param_count(void) { return mpm::mlp::NTOTAL; }

int mpm_mlp_create(int64_t batch_capacity, mpm_mlp** out) {
    if (!out) return MPM_E_ARG;
    *out = nullptr;
    if (batch_capacity < 32 || batch_capacity > (1 << 22)) return MPM_E_ARG;
    mpm_mlp* m = new mpm_mlp();
    int rc = m->init(int(batch_capacity));
    if (rc != MPM_OK) {
        m->release();
        delete m;
        return rc;
    }
    *out = m;
    return MPM_OK;
}

void mpm_mlp_destroy(mpm_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    if (m->X) (void)hipFree(m->X);
    if (m->Y) (void)hipFree(m->Y);
    m->release();
    delete m;
}

const char* mpm_mlp_last_error(const mpm_mlp* m) {
    return m ? m->error.c_str() : "mpm_mlp: NULL handle";
}

int mpm_mlp_set_params(mpm_mlp* m, const float* params, int64_t n) {
    if (!m || !params) return MPM_E_ARG;
    if (n != mpm::mlp::NTOTAL) return m->fail(MPM_E_ARG, "mpm_mlp_set_params: wrong parameter count");
    hipError_t e = hipMemcpyAsync(m->P, params, sizeof(float) * n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->m, 0, sizeof(float) * mpm::mlp::NPARAM, m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->v, 0, sizeof(float) * mpm::mlp::NPARAM, m->stream);
    if (e == hipSuccess) e = m->prepare();
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) return m->hip_fail(e, "mpm_mlp_set_params");
    m->adam_t = 0;
    return MPM_OK;
}

int mpm_mlp_get_params(mpm_mlp* m, float* params, int64_t n) {
    if (!m || !params) return MPM_E_ARG;
    if (n != mpm::mlp::NTOTAL) return m->fail(MPM_E_ARG, "mpm_mlp_get_params: wrong parameter count");
    hipError_t e = hipMemcpyAsync(params, m->P, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    return e == hipSuccess ? MPM_OK : m->hip_fail(e, "mpm_mlp_get_params");
}

int mpm_mlp_forward(mpm_mlp* m, const float* x, int64_t n, float* y) {
    using namespace mpm::mlp;
    if (!m || !x || !y || n < 0) return MPM_E_ARG;
    if (n == 0) return MPM_OK;
    int rc = m->upload_data(x, nullptr, n);
    if (rc != MPM_OK) return rc;
    float* dy = nullptr;
    hipError_t e = hipMalloc((void**)&dy, sizeof(float) * n * NOUT);
    if (e != hipSuccess) return m->hip_fail(e, "mpm_mlp_forward");
    k_mlp_forward<RowsIn, RowsOut><<<m->forward_grid(n), 256, 0, m->stream>>>(
        m->sw, m->P, n, RowsIn{m->X}, RowsOut{dy});
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(y, dy, sizeof(float) * n * NOUT, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dy);
    return e == hipSuccess ? MPM_OK : m->hip_fail(e, "mpm_mlp_forward");
}

int mpm_mlp_grad(mpm_mlp* m, const float* x, const float* y, int64_t n, float* grad, float* loss) {
    using namespace mpm::mlp;
    if (!m || !x || !y || !grad || !loss) return MPM_E_ARG;
    if (n < 1 || n > m->cap) return m->fail(MPM_E_ARG, "mpm_mlp_grad: n outside 1..batch capacity");
    int rc = m->upload_data(x, y, n);
    if (rc != MPM_OK) return rc;
    hipError_t e = hipMemsetAsync(m->loss, 0, sizeof(float), m->stream);
    m->batch(int(n), 0, 0, true);
    if (e == hipSuccess) e = m->forward_backward(int(n), m->loss);
    if (e == hipSuccess)
        e = hipMemcpyAsync(grad, m->G, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    return e == hipSuccess ? MPM_OK : m->hip_fail(e, "mpm_mlp_grad");
}

int mpm_mlp_train(mpm_mlp* m, const float* x, const float* y, int64_t n, int64_t steps,
                  int64_t batch, double lr, uint64_t seed, float* loss_history) {
    using namespace mpm::mlp;
    if (!m || !x || !y || n < 1 || steps < 0) return MPM_E_ARG;
    if (batch < 32 || batch > m->cap) return m->fail(MPM_E_ARG, "mpm_mlp_train: batch outside 32..capacity");
    int rc = m->upload_data(x, y, n);
    if (rc != MPM_OK) return rc;
    float* hist = nullptr;
    hipError_t e = hipMalloc((void**)&hist, sizeof(float) * (steps ? steps : 1));
    if (e == hipSuccess) e = hipMemsetAsync(hist, 0, sizeof(float) * (steps ? steps : 1), m->stream);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int64_t s = 0; s < steps && e == hipSuccess; ++s) {
        m->batch(int(batch), seed, s, false);
        e = m->forward_backward(int(batch), hist + s);
        m->adam_t += 1;
        const float c1 = float(1.0 / (1.0 - std::pow(0.9, double(m->adam_t))));
        const float c2 = float(1.0 / (1.0 - std::pow(0.999, double(m->adam_t))));
        k_adam<<<(NPARAM + 255) / 256, 256, 0, m->stream>>>(m->P, m->G, m->m, m->v, float(lr), b1,
                                                             b2, eps, c1, c2);
        if (e == hipSuccess) e = m->prepare();
    }
    if (e == hipSuccess && loss_history && steps)
        e = hipMemcpyAsync(loss_history, hist, sizeof(float) * steps, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(hist);
    return e == hipSuccess ? MPM_OK : m->hip_fail(e, "mpm_mlp_train");
}

int mpm_mlp_bench(mpm_mlp* m, int64_t n, int reps, double* ms_per_launch) {
    using namespace mpm::mlp;
    if (!m || n < 1 || reps < 1 || !ms_per_launch) return MPM_E_ARG;
    std::vector<float> x((size_t)n * NIN);
    for (size_t i = 0; i < x.size(); ++i) x[i] = float((i * 2654435761u) % 1000) * 1e-3f - 0.5f;
    int rc = m->upload_data(x.data(), nullptr, n);
    if (rc != MPM_OK) return rc;
    float* dy = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void**)&dy, sizeof(float) * n * NOUT);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const unsigned grid = m->forward_grid(n);
    if (e == hipSuccess) {
        k_mlp_forward<RowsIn, RowsOut><<<grid, 256, 0, m->stream>>>(m->sw, m->P, n, RowsIn{m->X},
                                                                    RowsOut{dy});
        e = hipEventRecord(e0, m->stream);
    }
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        k_mlp_forward<RowsIn, RowsOut><<<grid, 256, 0, m->stream>>>(m->sw, m->P, n, RowsIn{m->X},
                                                                    RowsOut{dy});
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, m->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (dy) (void)hipFree(dy);
    if (e != hipSuccess) return m->hip_fail(e, "mpm_mlp_bench");
    *ms_per_launch = double(ms) / reps;
    return MPM_OK;
}

}  // extern "C"
