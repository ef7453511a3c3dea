// mpm_mlp.hpp -- the learned return map: a 16-128-128-128-7 SiLU MLP on the
// gfx950 matrix cores (BF16 operands, FP32 accumulation).
//
// Inputs (16): trial stress (6, after the elastic predictor), strain increment
// (6), accumulated plastic strain, friction angle, cohesion in effect, dilation
// angle; each standardised with a stored mean and inverse deviation.
// Outputs (7), in units of a stored per-output scale: the stress correction
// (σ_returned − σ_trial, 6) and the plastic strain increment.
//
// Matrix-core layout.  v_mfma_f32_32x32x16_bf16 multiplies A (32x16) by B
// (16x32) into D (32x32).  Lane l holds A[l%32][8(l/32)+v] and B[8(l/32)+v][l%32]
// for v = 0..7, and D[8(r/4) + 4(l/32) + r%4][l%32] in accumulator r = 0..15.
// A layer is H_out (units x particles) = W (units x inputs) · H_in, so the
// particles are the 32 columns of a tile and the weights are the A operand.
// Accumulators r = 8h..8h+7 of output tile t are then exactly the B operand of
// K-block 2t+h of the next layer once that layer's weight columns are permuted
// by perm_k below, so activations never leave the registers.  The permuted
// BF16 weights ("swizzled", 16 contiguous bytes per lane and MFMA) are built
// from the FP32 master parameters by k_mlp_swizzle.
#pragma once
#include <hip/hip_runtime.h>

namespace mpm {
namespace mlp {

constexpr int NIN = 16, NH = 128, NOUT = 7;
// flat FP32 parameter vector: W1[128][16] b1 W2[128][128] b2 W3[128][128] b3 W4[7][128] b4,
// then the normalisation block x_mean[16] x_istd[16] y_scale[7]
constexpr int OW1 = 0, OB1 = OW1 + NH * NIN, OW2 = OB1 + NH, OB2 = OW2 + NH * NH;
constexpr int OW3 = OB2 + NH, OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NOUT * NH;
constexpr int NPARAM = OB4 + NOUT;
constexpr int OXM = NPARAM, OXS = OXM + NIN, OYS = OXS + NIN, NTOTAL = OYS + NOUT;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// swizzled weights, in bf16x8 units (one per lane and MFMA): layer L has
// MT(L) x KB(L) blocks of 64 lanes
constexpr int SW1 = 4 * 1 * 64, SW2 = 4 * 8 * 64, SW3 = 4 * 8 * 64, SW4 = 1 * 8 * 64;
constexpr int SO1 = 0, SO2 = SO1 + SW1, SO3 = SO2 + SW2, SO4 = SO3 + SW3, SWN = SO4 + SW4;
constexpr int SW_LDS = SO4;  // layers 1-3 are staged in LDS (69,632 bytes)

// logistic function with the hardware reciprocal (1 ulp; the results are
// rounded to BF16 anyway)
__device__ __forceinline__ float sigmoid(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// Hidden unit held in element v of the B operand of K-block kb by lane half hi
// (= l / 32) when that operand is the previous layer's accumulator registers.
__host__ __device__ __forceinline__ int perm_k(int kb, int hi, int v) {
    return 32 * (kb >> 1) + 16 * (kb & 1) + 8 * (v >> 2) + 4 * hi + (v & 3);
}

__device__ __forceinline__ f32x16 mfma(const bf16x8& a, const bf16x8& b, const f32x16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Builds the swizzled BF16 weights from the FP32 master parameters.
__global__ __launch_bounds__(256) void k_mlp_swizzle(const float* __restrict__ P,
                                                     bf16x8* __restrict__ sw) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= SWN) return;
    int layer, rem;
    if (e < SO2) { layer = 1; rem = e - SO1; }
    else if (e < SO3) { layer = 2; rem = e - SO2; }
    else if (e < SO4) { layer = 3; rem = e - SO3; }
    else { layer = 4; rem = e - SO4; }
    const int KB = layer == 1 ? 1 : 8;
    const int l = rem & 63, blk = rem >> 6;
    const int mt = blk / KB, kb = blk % KB;
    const int row = 32 * mt + (l & 31), hi = l >> 5;
    const int nin = layer == 1 ? NIN : NH, nout = layer == 4 ? NOUT : NH;
    const float* W = P + (layer == 1 ? OW1 : layer == 2 ? OW2 : layer == 3 ? OW3 : OW4);
    bf16x8 out;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const int col = layer == 1 ? 8 * hi + v : perm_k(kb, hi, v);
        out[v] = (__bf16)(row < nout && col < nin ? W[row * nin + col] : 0.0f);
    }
    sw[e] = out;
}

// The network on 64 particles per wavefront, 4 wavefronts per workgroup.
// Pro::load(p, half, x8) fills x8 with the raw features 8*half..8*half+7 of
// particle p (p < n); Epi::store(p, y7) receives particle p's 7 de-normalised
// outputs (p < n).  Layers 1-3 are read from LDS, layer 4 from global memory.
template <class Pro, class Epi>
__global__ __launch_bounds__(256, 2) void k_mlp_forward(const bf16x8* __restrict__ sw,
                                                       const float* __restrict__ P,
                                                       long long n, Pro pro, Epi epi) {
    __shared__ bf16x8 wl[SW_LDS];
    __shared__ float bl[3 * NH + NOUT + NTOTAL - NPARAM];
    for (int i = threadIdx.x; i < SW_LDS; i += 256) wl[i] = sw[i];
    for (int i = threadIdx.x; i < NH; i += 256) {
        bl[i] = P[OB1 + i];
        bl[NH + i] = P[OB2 + i];
        bl[2 * NH + i] = P[OB3 + i];
    }
    for (int i = threadIdx.x; i < NOUT; i += 256) bl[3 * NH + i] = P[OB4 + i];
    for (int i = threadIdx.x; i < NTOTAL - NPARAM; i += 256) bl[3 * NH + NOUT + i] = P[NPARAM + i];
    __syncthreads();
    const float* xm = bl + 3 * NH + NOUT;
    const float* xs = xm + NIN;
    const float* ys = xs + NIN;
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, hi = l >> 5;
    const long long stride = (long long)gridDim.x * 4 * 64;
    for (long long base = ((long long)blockIdx.x * 4 + wave) * 64; base < n; base += stride) {
        // B operand of layer 1 for the two 32-particle column tiles
        bf16x8 x0[2];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const long long p = base + nt * 32 + (l & 31);
            float f[8];
            if (p < n) {
                pro.load(p, hi, f);
            } else {
#pragma unroll
                for (int v = 0; v < 8; ++v) f[v] = xm[8 * hi + v];
            }
#pragma unroll
            for (int v = 0; v < 8; ++v) x0[nt][v] = (__bf16)((f[v] - xm[8 * hi + v]) * xs[8 * hi + v]);
        }
        f32x16 acc[2][4];
        bf16x8 h[2][8];
        // accumulators start from the bias of their rows (4 contiguous rows per group)
        auto bias_tile = [&](const float* bias, int t) {
            f32x16 c;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float4 b4 = *(const float4*)(bias + 32 * t + 8 * g + 4 * hi);
                c[4 * g + 0] = b4.x;
                c[4 * g + 1] = b4.y;
                c[4 * g + 2] = b4.z;
                c[4 * g + 3] = b4.w;
            }
            return c;
        };
        // SiLU of output tile t into the B operands of K-blocks 2t and 2t+1
        auto activate = [&]() {
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        bf16x8 v8;
#pragma unroll
                        for (int v = 0; v < 8; ++v) v8[v] = (__bf16)silu(acc[nt][t][8 * hh + v]);
                        h[nt][2 * t + hh] = v8;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        };
        // layer 1
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const bf16x8 a = wl[SO1 + mt * 64 + l];
            const f32x16 b = bias_tile(bl, mt);
            acc[0][mt] = mfma(a, x0[0], b);
            acc[1][mt] = mfma(a, x0[1], b);
        }
        activate();
        // layers 2 and 3
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
            const int so = layer == 0 ? SO2 : SO3;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                acc[0][mt] = bias_tile(bl + (layer + 1) * NH, mt);
                acc[1][mt] = acc[0][mt];
            }
#pragma unroll
            for (int kb = 0; kb < 8; ++kb) {
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    const bf16x8 a = wl[so + (mt * 8 + kb) * 64 + l];
                    acc[0][mt] = mfma(a, h[0][kb], acc[0][mt]);
                    acc[1][mt] = mfma(a, h[1][kb], acc[1][mt]);
                }
            }
            activate();
        }
        // layer 4 (7 of 32 rows used)
        const f32x16 zero = {};
        f32x16 o0 = zero, o1 = zero;
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) {
            const bf16x8 a = sw[SO4 + kb * 64 + l];
            o0 = mfma(a, h[0][kb], o0);
            o1 = mfma(a, h[1][kb], o1);
        }
        // rows 0-3 sit in lanes 0-31 (r = 0..3), rows 4-6 in lanes 32-63 (r = 0..2);
        // lane l < 32 takes particle l of tile 0, lane l >= 32 particle l of tile 1
        float y[NOUT];
        float mine[4], part[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float a0 = o0[r], a1 = o1[r];
            const float s0 = __shfl_xor(a0, 32), s1 = __shfl_xor(a1, 32);
            mine[r] = hi == 0 ? a0 : a1;
            part[r] = hi == 0 ? s0 : s1;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) y[r] = hi == 0 ? mine[r] : part[r];
#pragma unroll
        for (int r = 0; r < 3; ++r) y[4 + r] = hi == 0 ? part[r] : mine[r];
#pragma unroll
        for (int k = 0; k < NOUT; ++k) y[k] = (y[k] + bl[3 * NH + k]) * ys[k];
        const long long p = base + l;
        if (p < n) epi.store(p, y);
    }
}

// ---------------------------------------------------------------------------
// Training: row-major activations (rows = samples), BF16 operands, FP32 sums.
// ---------------------------------------------------------------------------
enum { EPI_STORE = 0, EPI_HIDDEN = 1, EPI_OUTPUT = 2, EPI_DELTA = 3 };

struct GemmEpi {
    const float* bias;  // EPI_HIDDEN / EPI_OUTPUT: per column
    float* z;           // EPI_HIDDEN: pre-activation out (M x N); EPI_DELTA: pre-activation in
    __bf16* act;        // EPI_HIDDEN: silu(z) out; EPI_DELTA: delta out (BF16)
    float* c;           // EPI_STORE / EPI_OUTPUT: FP32 out
    int ldc;
};

// C (M x N) = op(A) (M x K) · op(B) (K x N); op(A)[i][k] = TA ? A[k lda + i] : A[i lda + k],
// op(B)[k][j] = TB ? B[j ldb + k] : B[k ldb + j].  One wavefront per 32x32 tile.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(64) void k_gemm(int M, int N, int K, const __bf16* __restrict__ A,
                                             int lda, const __bf16* __restrict__ B, int ldb,
                                             GemmEpi e) {
    const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32, l = threadIdx.x, hi = l >> 5;
    const int ia = i0 + (l & 31), jb = j0 + (l & 31);
    f32x16 acc = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const int k = k0 + 8 * hi + v;
            a[v] = (ia < M && k < K) ? (TA ? A[(long long)k * lda + ia] : A[(long long)ia * lda + k])
                                     : (__bf16)0.0f;
            b[v] = (jb < N && k < K) ? (TB ? B[(long long)jb * ldb + k] : B[(long long)k * ldb + jb])
                                     : (__bf16)0.0f;
        }
        acc = mfma(a, b, acc);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = i0 + 8 * (r >> 2) + 4 * hi + (r & 3), j = jb;
        if (i >= M || j >= N) continue;
        const float v = acc[r];
        const long long o = (long long)i * N + j;
        if constexpr (EPI == EPI_STORE) {
            e.c[(long long)i * e.ldc + j] = v;
        } else if constexpr (EPI == EPI_HIDDEN) {
            const float z = v + e.bias[j];
            e.z[o] = z;
            e.act[o] = (__bf16)silu(z);
        } else if constexpr (EPI == EPI_OUTPUT) {
            e.c[o] = v + e.bias[j];
        } else {
            e.act[o] = (__bf16)(v * dsilu(e.z[o]));
        }
    }
}

// Split-K form of k_gemm for the weight gradients (M, N <= 128, K = the batch):
// slice z of gridDim.z sums k in [z kc, (z + 1) kc) into ws[z][i][j]; then
// k_splitk_reduce adds the slices in slice order into C (deterministic).
template <bool TA, bool TB>
__global__ __launch_bounds__(64) void k_gemm_splitk(int M, int N, int K, int kc,
                                                    const __bf16* __restrict__ A, int lda,
                                                    const __bf16* __restrict__ B, int ldb,
                                                    float* __restrict__ ws) {
    const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32, l = threadIdx.x, hi = l >> 5;
    const int ia = i0 + (l & 31), jb = j0 + (l & 31);
    const int kb = blockIdx.z * kc, ke = min(K, kb + kc);
    f32x16 acc = {};
    for (int k0 = kb; k0 < ke; k0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const int k = k0 + 8 * hi + v;
            a[v] = (ia < M && k < ke) ? (TA ? A[(long long)k * lda + ia] : A[(long long)ia * lda + k])
                                      : (__bf16)0.0f;
            b[v] = (jb < N && k < ke) ? (TB ? B[(long long)jb * ldb + k] : B[(long long)k * ldb + jb])
                                      : (__bf16)0.0f;
        }
        acc = mfma(a, b, acc);
    }
    float* w = ws + (long long)blockIdx.z * M * N;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = i0 + 8 * (r >> 2) + 4 * hi + (r & 3), j = jb;
        if (i < M && j < N) w[(long long)i * N + j] = acc[r];
    }
}

__global__ __launch_bounds__(256) void k_splitk_reduce(const float* __restrict__ ws, int slices, int M,
                                                       int N, float* __restrict__ c, int ldc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * N) return;
    const int i = t / N, j = t - i * N;
    float s = 0.0f;
    for (int z = 0; z < slices; ++z) s += ws[(long long)z * M * N + t];
    c[(long long)i * ldc + j] = s;
}

// Output delta of the mean squared error over normalised targets:
// d = 2 (o - y) / (B * 7); loss partial sums per block.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ o, const float* __restrict__ y,
                                             int n, __bf16* __restrict__ d, float* __restrict__ dsum,
                                             float* __restrict__ loss) {
    __shared__ float part[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.0f;
    if (i < n) {
        const float r = o[i] - y[i];
        const float dv = 2.0f * r / float(n);
        d[i] = (__bf16)dv;
        dsum[i] = dv;
        s = r * r;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0] / float(n));
}

// Column sums of a rows x cols FP32 matrix (bias gradients), added to g.
// Column sums of a row-major rows x cols matrix (bias gradients), in two
// deterministic stages.  Stage 1: block b sums rows [b RB, (b + 1) RB) of every
// column, its threads laid out as row groups x columns (consecutive threads on
// consecutive columns, so each row group reads whole rows), eight independent
// loads in flight per thread, the groups combined in LDS in a fixed order;
// part[b][j] holds the block's sum.  Stage 2: one thread per column adds the
// blocks' sums in block order.  cols <= 256.
constexpr int kColsumBlocks = 64;

template <typename T>
__global__ __launch_bounds__(256) void k_colsum_part(const T* __restrict__ m, int rows, int cols,
                                                     float* __restrict__ part) {
    __shared__ float red[256];
    const int groups = 256 / cols;
    const int t = threadIdx.x, grp = t / cols, j = t - grp * cols;
    const int RB = (rows + gridDim.x - 1) / gridDim.x;
    const int r0 = blockIdx.x * RB, r1 = min(rows, r0 + RB);
    float s = 0.0f;
    if (grp < groups) {
        int r = r0 + grp;
        for (; r + 7 * groups < r1; r += 8 * groups) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = float(m[(long long)(r + u * groups) * cols + j]);
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; r < r1; r += groups) s += float(m[(long long)r * cols + j]);
    }
    red[t] = s;
    __syncthreads();
    if (t < cols) {
        float acc = 0.0f;
        for (int g2 = 0; g2 < groups; ++g2) acc += red[g2 * cols + t];
        part[blockIdx.x * cols + t] = acc;
    }
}

__global__ __launch_bounds__(256) void k_colsum_final(const float* __restrict__ part, int nblocks,
                                                      int cols, float* __restrict__ g) {
    const int j = threadIdx.x;
    if (j >= cols) return;
    float s = 0.0f;
    for (int b = 0; b < nblocks; ++b) s += part[b * cols + j];
    g[j] = s;
}

// BF16 row-major copies of W1..W4 for the training GEMMs.
__global__ __launch_bounds__(256) void k_weights_bf16(const float* __restrict__ P,
                                                      __bf16* __restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    w[i] = (__bf16)P[i];
}

// Gathers a minibatch: normalised BF16 inputs and normalised FP32 targets.
__global__ __launch_bounds__(256) void k_batch(const float* __restrict__ X, const float* __restrict__ Y,
                                               long long n, int B, unsigned long long seed,
                                               long long step, const float* __restrict__ P,
                                               __bf16* __restrict__ xb, float* __restrict__ yb,
                                               int sequential) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    long long src;
    if (sequential) {
        src = b < n ? b : n - 1;
    } else {
        // splitmix64 of (seed, step, b)
        unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(step * B + b + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        src = (long long)(z % (unsigned long long)n);
    }
#pragma unroll
    for (int k = 0; k < NIN; ++k)
        xb[b * NIN + k] = (__bf16)((X[src * NIN + k] - P[OXM + k]) * P[OXS + k]);
#pragma unroll
    for (int k = 0; k < NOUT; ++k) yb[b * NOUT + k] = Y[src * NOUT + k] / P[OYS + k];
}

// Adam (Kingma & Ba, 2015) on the NPARAM trainable parameters.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// Plain I/O for the forward kernel: raw features in, de-normalised outputs out.
struct RowsIn {
    const float* x;
    __device__ void load(long long p, int half, float* f) const {
#pragma unroll
        for (int v = 0; v < 8; ++v) f[v] = x[p * NIN + 8 * half + v];
    }
};
struct RowsOut {
    float* y;
    __device__ void store(long long p, const float* o) const {
#pragma unroll
        for (int k = 0; k < NOUT; ++k) y[p * NOUT + k] = o[k];
    }
};

}  // namespace mlp
}  // namespace mpm
