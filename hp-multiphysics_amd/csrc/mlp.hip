// Preconditioner MLP: training and inference kernels for gfx950.
//
// Network 11 -> 64 -> 64 -> 64 -> 9, GELU (tanh form) on the hidden layers.  The
// FP32 master weights are rounded to bfloat16 (nearest even) for every matrix
// product and activations are rounded to bfloat16 on input to a product;
// accumulation is FP32.  Inference runs on the matrix cores (bf16 MFMA, 16 samples
// per tile).  Training: per 64-sample block the backward pass keeps pre-activations and
// deltas in LDS and forms the weight gradients as sample reductions; per-block
// partial gradients are summed in a fixed order by the Adam kernel, so training
// is bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hpgpu.h"
#include "mlp.h"

namespace hpg {

constexpr int MI = MLP_IN, MH = 64, MO = MLP_OUT;
static_assert(MO <= 16, "the inference head is one 16-row MFMA tile");
constexpr int O_W1 = 0, O_B1 = O_W1 + MI * MH, O_W2 = O_B1 + MH, O_B2 = O_W2 + MH * MH, O_W3 = O_B2 + MH,
              O_B3 = O_W3 + MH * MH, O_W4 = O_B3 + MH, O_B4 = O_W4 + MH * MO, NPARAM = O_B4 + MO;
constexpr float C_GELU = 0.7978845608028654f;  // sqrt(2/pi)

__host__ __device__ inline float bf16r(float x) {
    unsigned u;
    std::memcpy(&u, &x, 4);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}
__device__ inline float gelu(float z) {
    const float t = tanhf(C_GELU * (z + 0.044715f * z * z * z));
    return 0.5f * z * (1.f + t);
}
__device__ inline float gelu_d(float z) {
    const float u = C_GELU * (z + 0.044715f * z * z * z);
    const float t = tanhf(u);
    const float du = C_GELU * (1.f + 3.f * 0.044715f * z * z);
    return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * du;
}

// one hidden layer: z[o] = b[o] + sum_i bf16(h[i]) * wq[i][o]
template <int NIN>
__device__ inline void layer(const float* __restrict__ wq, const float* __restrict__ h, float* z) {
    float hb[NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) hb[i] = bf16r(h[i]);
#pragma unroll
    for (int o = 0; o < MH; ++o) {
        float s = wq[NIN * MH + o];  // bias follows the weight block
#pragma unroll
        for (int i = 0; i < NIN; ++i) s = fmaf(hb[i], wq[i * MH + o], s);
        z[o] = s;
    }
}

__device__ inline void head(const float* __restrict__ wq, const float* h, float* y) {
    float hb[MH];
#pragma unroll
    for (int i = 0; i < MH; ++i) hb[i] = bf16r(h[i]);
#pragma unroll
    for (int o = 0; o < MO; ++o) {
        float s = wq[O_B4 + o];
#pragma unroll
        for (int i = 0; i < MH; ++i) s = fmaf(hb[i], wq[O_W4 + i * MO + o], s);
        y[o] = s;
    }
}

// ---- inference on the matrix cores -------------------------------------------------
// v_mfma_f32_16x16x16_bf16 with the hidden units as rows and 16 samples as columns.
// Operand layout: lane l holds A[l%16][4(l/16)+j] and B[4(l/16)+j][l%16] (j = 0..3);
// accumulator register i of lane l holds D[4(l/16)+i][l%16] (checked on the device
// by hpg_selftest_mfma_bf16).  The accumulators of output tile mt of one layer are
// therefore the B operand of K-step kt = mt of the next layer, lane for lane: the
// activations never leave registers.  Each wave holds every weight as A fragments
// (80 VGPRs) and sweeps 16-sample tiles; products of bfloat16 operands are exact in
// FP32 and accumulate in FP32, so only the summation order differs from the
// sample-per-lane forward pass of the training kernel below.
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ inline short bf16_bits(float x) {
    unsigned u;
    std::memcpy(&u, &x, 4);
    return (short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ inline f32x4_t mfma_bf16(bf16x4_t a, bf16x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// GELU of the accumulators, rounded to bfloat16: the B operands of the next layer
__device__ inline void gelu_fragments(const f32x4_t z[4], bf16x4_t h[4]) {
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int j = 0; j < 4; ++j) h[kt][j] = bf16_bits(gelu(z[kt][j]));
}

__global__ __launch_bounds__(256) void k_mlp_forward_mfma(int n, const float* __restrict__ X,
                                                          const float* __restrict__ wq, float* __restrict__ Y) {
    const int lane = threadIdx.x & 63, q = lane >> 4, c = lane & 15;
    bf16x4_t A1[4], A2[4][4], A3[4][4], A4[4];
    f32x4_t b1[4], b2[4], b3[4], b4;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * q + j;
            A1[mt][j] = k < MI ? bf16_bits(wq[O_W1 + k * MH + 16 * mt + c]) : (short)0;
        }
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = 16 * kt + 4 * q + j;
                A2[mt][kt][j] = bf16_bits(wq[O_W2 + k * MH + 16 * mt + c]);
                A3[mt][kt][j] = bf16_bits(wq[O_W3 + k * MH + 16 * mt + c]);
            }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            b1[mt][i] = wq[O_B1 + 16 * mt + 4 * q + i];
            b2[mt][i] = wq[O_B2 + 16 * mt + 4 * q + i];
            b3[mt][i] = wq[O_B3 + 16 * mt + 4 * q + i];
        }
    }
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 16 * kt + 4 * q + j;
            A4[kt][j] = c < MO ? bf16_bits(wq[O_W4 + k * MO + c]) : (short)0;
        }
#pragma unroll
    for (int i = 0; i < 4; ++i) b4[i] = 4 * q + i < MO ? wq[O_B4 + 4 * q + i] : 0.f;

    const long ntile = ((long)n + 15) / 16;
    for (long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6); tile < ntile; tile += (long)gridDim.x * 4) {
        const long s = tile * 16 + c;
        bf16x4_t h[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * q + j;
            h[0][j] = (s < n && k < MI) ? bf16_bits(X[s * MI + k]) : (short)0;
        }
        f32x4_t z[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) z[mt] = mfma_bf16(A1[mt], h[0], b1[mt]);
        gelu_fragments(z, h);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4_t acc = b2[mt];
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc = mfma_bf16(A2[mt][kt], h[kt], acc);
            z[mt] = acc;
        }
        gelu_fragments(z, h);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4_t acc = b3[mt];
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc = mfma_bf16(A3[mt][kt], h[kt], acc);
            z[mt] = acc;
        }
        gelu_fragments(z, h);
        f32x4_t y = b4;
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) y = mfma_bf16(A4[kt], h[kt], y);
        // register i of lane group q holds output 4q + i of sample column c
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (4 * q + i < MO && s < n) Y[s * MO + 4 * q + i] = y[i];
    }
}

// one v_mfma_f32_16x16x16_bf16 on A, B (16 x 16 row-major, rounded to bfloat16)
__global__ void k_mfma_probe_bf16(const float* A, const float* B, float* out) {
    const int l = threadIdx.x, q = l >> 4, c = l & 15;
    bf16x4_t a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        a[j] = bf16_bits(A[c * 16 + 4 * q + j]);
        b[j] = bf16_bits(B[(4 * q + j) * 16 + c]);
    }
    f32x4_t d = {0.f, 0.f, 0.f, 0.f};
    d = mfma_bf16(a, b, d);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[l * 4 + i] = d[i];
}

// backward of one 64-sample block: partial gradient (NPARAM) and partial loss sum
__global__ __launch_bounds__(64) void k_mlp_backward(int n, int nscale, const float* __restrict__ X,
                                                     const float* __restrict__ T, const float* __restrict__ wq,
                                                     float* __restrict__ partial, float* __restrict__ lossp) {
    // [layer][unit][sample], rows padded to 65 words so that lanes reading one
    // sample column of different units hit different banks
    __shared__ float sZ[3][MH][65];   // pre-activations; after the deltas, bf16(gelu(z))
    __shared__ float sD[3][MH][65];   // deltas of the hidden layers
    __shared__ float sX[MI][65];
    __shared__ float sY[MO][65];      // output deltas
    const int s = threadIdx.x;
    const int gs = blockIdx.x * 64 + s;
    const bool ok = gs < n;
    float x[MI], a[MH], b[MH];
#pragma unroll
    for (int i = 0; i < MI; ++i) x[i] = ok ? X[(size_t)gs * MI + i] : 0.f;
#pragma unroll
    for (int i = 0; i < MI; ++i) sX[i][s] = x[i];
    layer<MI>(wq + O_W1, x, a);
#pragma unroll
    for (int j = 0; j < MH; ++j) { sZ[0][j][s] = a[j]; a[j] = gelu(a[j]); }
    layer<MH>(wq + O_W2, a, b);
#pragma unroll
    for (int j = 0; j < MH; ++j) { sZ[1][j][s] = b[j]; b[j] = gelu(b[j]); }
    layer<MH>(wq + O_W3, b, a);
#pragma unroll
    for (int j = 0; j < MH; ++j) { sZ[2][j][s] = a[j]; a[j] = gelu(a[j]); }
    float y[MO];
    head(wq, a, y);
    float dy[MO], l = 0.f;
#pragma unroll
    for (int o = 0; o < MO; ++o) {
        const float d = ok ? y[o] - T[(size_t)gs * MO + o] : 0.f;
        l += d * d;
        dy[o] = 2.f * d / (float)nscale;
        sY[o][s] = dy[o];
    }
    // hidden deltas: dz3 = W4 dy * g'(z3); dz2 = W3 dz3 * g'(z2); dz1 = W2 dz2 * g'(z1)
#pragma unroll
    for (int j = 0; j < MH; ++j) {
        float t = 0.f;
#pragma unroll
        for (int o = 0; o < MO; ++o) t = fmaf(wq[O_W4 + j * MO + o], dy[o], t);
        a[j] = t * gelu_d(sZ[2][j][s]);
        sD[2][j][s] = a[j];
    }
#pragma unroll
    for (int i = 0; i < MH; ++i) {
        float t = 0.f;
#pragma unroll
        for (int j = 0; j < MH; ++j) t = fmaf(wq[O_W3 + i * MH + j], a[j], t);
        b[i] = t * gelu_d(sZ[1][i][s]);
        sD[1][i][s] = b[i];
    }
#pragma unroll
    for (int i = 0; i < MH; ++i) {
        float t = 0.f;
#pragma unroll
        for (int j = 0; j < MH; ++j) t = fmaf(wq[O_W2 + i * MH + j], b[j], t);
        sD[0][i][s] = t * gelu_d(sZ[0][i][s]);
    }
    // the weight gradients need the (rounded) activations, not the pre-activations;
    // each lane converts its own sample column
#pragma unroll
    for (int L = 0; L < 3; ++L)
#pragma unroll
        for (int j = 0; j < MH; ++j) sZ[L][j][s] = bf16r(gelu(sZ[L][j][s]));
    // block loss
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if (s == 0) lossp[blockIdx.x] = l;
    __syncthreads();
    // weight gradients: sample reductions, one parameter per lane per pass
    float* P = partial + (size_t)blockIdx.x * NPARAM;
    for (int p = s; p < NPARAM; p += 64) {
        float g = 0.f;
        if (p < O_B1) {                       // W1[i][j]
            const int i = p / MH, j = p % MH;
            for (int k = 0; k < 64; ++k) g = fmaf(bf16r(sX[i][k]), sD[0][j][k], g);
        } else if (p < O_W2) {                // b1
            const int j = p - O_B1;
            for (int k = 0; k < 64; ++k) g += sD[0][j][k];
        } else if (p < O_B2) {                // W2[i][j]: h1 = bf16(gelu(z1))
            const int q = p - O_W2, i = q / MH, j = q % MH;
            for (int k = 0; k < 64; ++k) g = fmaf(sZ[0][i][k], sD[1][j][k], g);
        } else if (p < O_W3) {
            const int j = p - O_B2;
            for (int k = 0; k < 64; ++k) g += sD[1][j][k];
        } else if (p < O_B3) {                // W3[i][j]: h2 = bf16(gelu(z2))
            const int q = p - O_W3, i = q / MH, j = q % MH;
            for (int k = 0; k < 64; ++k) g = fmaf(sZ[1][i][k], sD[2][j][k], g);
        } else if (p < O_W4) {
            const int j = p - O_B3;
            for (int k = 0; k < 64; ++k) g += sD[2][j][k];
        } else if (p < O_B4) {                // W4[i][o]: h3 = bf16(gelu(z3))
            const int q = p - O_W4, i = q / MO, o = q % MO;
            for (int k = 0; k < 64; ++k) g = fmaf(sZ[2][i][k], sY[o][k], g);
        } else {
            const int o = p - O_B4;
            for (int k = 0; k < 64; ++k) g += sY[o][k];
        }
        P[p] = g;
    }
}

// gradient = ordered sum of the block partials; optional Adam step on the masters
__global__ void k_mlp_reduce_adam(int nblocks, const float* __restrict__ partial, float* __restrict__ grad,
                                  float* master, float* wq, float* m1, float* m2, float lr, float bc1, float bc2,
                                  int do_step) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float g = 0.f;
    for (int b = 0; b < nblocks; ++b) g += partial[(size_t)b * NPARAM + p];
    grad[p] = g;
    if (!do_step) return;
    const float a = 0.9f * m1[p] + 0.1f * g;
    const float v = 0.999f * m2[p] + 0.001f * g * g;
    m1[p] = a;
    m2[p] = v;
    const float w = master[p] - lr * (a / bc1) / (sqrtf(v / bc2) + 1e-8f);
    master[p] = w;
    const bool is_bias = (p >= O_B1 && p < O_W2) || (p >= O_B2 && p < O_W3) || (p >= O_B3 && p < O_W4) || p >= O_B4;
    wq[p] = is_bias ? w : bf16r(w);
}

// mean-square loss from the block losses: the same double sum, in block order, and
// the same final division as the host side of hpg_mlp_loss_grad
__global__ void k_mlp_loss_sum(int nblocks, const float* __restrict__ lossp, int nscale, float* __restrict__ loss) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += lossp[b];
    loss[0] = (float)(s / (double)nscale);
}

// Scratch of the allocation-free loss/gradient entry point: the per-block partial
// gradients (NPARAM floats per 64-sample block) followed by the block losses.  One
// copy per GPU, held until the process ends; host code uses the address that
// hipGetSymbolAddress returns for the current device, never the symbol itself.
// Every call overwrites what it reads (each block writes all of its partials and
// its loss), and calls that overlap on one GPU race on it.
constexpr int NOALLOC_FLOATS = 4 << 20;  // 16 MiB
constexpr int NOALLOC_MAX_BLOCKS = NOALLOC_FLOATS / (NPARAM + 1);
static __device__ __align__(256) float scratch[NOALLOC_FLOATS];
static_assert(NOALLOC_MAX_BLOCKS * 64 == HPG_MLP_NOALLOC_MAX_SAMPLES, "header limit out of step with the scratch size");

}  // namespace hpg

using namespace hpg;

struct hpg_mlp {
    hipStream_t stream = nullptr;
    float *master = nullptr, *wq = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
    float *partial = nullptr, *lossp = nullptr, *X = nullptr, *T = nullptr, *Y = nullptr;
    int cap_blocks = 0, cap_n = 0, step = 0;
    int dev = 0;  // device the MLP was created on
    std::string err;
    ~hpg_mlp() {
        for (float* p : {master, wq, m1, m2, grad, partial, lossp, X, T, Y})
            if (p) hipFree(p);
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {
void ck(hipError_t e, const char* w) {
    if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}
void quantise(hpg_mlp* m, const std::vector<float>& w) {
    std::vector<float> q(w);
    for (int p = 0; p < NPARAM; ++p) {
        const bool is_bias = (p >= O_B1 && p < O_W2) || (p >= O_B2 && p < O_W3) || (p >= O_B3 && p < O_W4) || p >= O_B4;
        if (!is_bias) q[p] = bf16r(w[p]);
    }
    ck(hipMemcpyAsync(m->master, w.data(), NPARAM * 4, hipMemcpyHostToDevice, m->stream), "H2D");
    ck(hipMemcpyAsync(m->wq, q.data(), NPARAM * 4, hipMemcpyHostToDevice, m->stream), "H2D");
    ck(hipStreamSynchronize(m->stream), "sync");
}
void ensure_n(hpg_mlp* m, int n) {
    if (n <= m->cap_n) return;
    for (float** p : {&m->X, &m->T, &m->Y})
        if (*p) { hipFree(*p); *p = nullptr; }
    ck(hipMalloc(&m->X, (size_t)n * MI * 4), "hipMalloc");
    ck(hipMalloc(&m->T, (size_t)n * MO * 4), "hipMalloc");
    ck(hipMalloc(&m->Y, (size_t)n * MO * 4), "hipMalloc");
    m->cap_n = n;
}
// device address of `scratch` per GPU, resolved once per device with that device
// current (each GPU holds its own copy of the array at its own address)
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float* g_scratch[MAX_DEVICES] = {};

void resolve_scratch(int dev) {
    if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return;
    void* p = nullptr;
    ck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned null");
    g_scratch[dev] = static_cast<float*>(p);
}

// enqueue the backward pass of n samples (block partials and block losses into
// partial / lossp) and the ordered gradient sum into grad, no weight update
void enqueue_loss_grad(hpg_mlp* m, const float* X, const float* T, int n, float* partial, float* lossp,
                       float* grad) {
    const int nb = (n + 63) / 64;
    hipLaunchKernelGGL(k_mlp_backward, dim3(nb), dim3(64), 0, m->stream, n, n * MO, X, T, m->wq, partial, lossp);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, nb, partial, grad,
                       m->master, m->wq, m->m1, m->m2, 0.f, 1.f, 1.f, 0);
}

void ensure_blocks(hpg_mlp* m, int nb) {
    if (nb <= m->cap_blocks) return;
    if (m->partial) hipFree(m->partial);
    if (m->lossp) hipFree(m->lossp);
    ck(hipMalloc(&m->partial, (size_t)nb * NPARAM * 4), "hipMalloc");
    ck(hipMalloc(&m->lossp, (size_t)nb * 4), "hipMalloc");
    m->cap_blocks = nb;
}
}  // namespace

namespace hpg {
void mlp_forward_device(const hpg_mlp* m, const float* X, int n, float* Y, hipStream_t stream) {
    if (n <= 0) return;
    // four 16-sample tiles per block and pass; at most 2048 blocks, each wave loops
    const long tiles = ((long)n + 15) / 16;
    const long blocks = std::min<long>((tiles + 3) / 4, 2048);
    hipLaunchKernelGGL(k_mlp_forward_mfma, dim3((unsigned)blocks), dim3(256), 0, stream, n, X, m->wq, Y);
}
}  // namespace hpg

extern "C" {

int hpg_mlp_create(unsigned seed, hpg_mlp** out) {
    if (!out) return -2;
    std::unique_ptr<hpg_mlp> m(new hpg_mlp());
    try {
        ck(hipGetDevice(&m->dev), "hipGetDevice");
        resolve_scratch(m->dev);
        ck(hipStreamCreate(&m->stream), "hipStreamCreate");
        for (float** p : {&m->master, &m->wq, &m->m1, &m->m2, &m->grad}) ck(hipMalloc(p, NPARAM * 4), "hipMalloc");
        ck(hipMemsetAsync(m->m1, 0, NPARAM * 4, m->stream), "memset");
        ck(hipMemsetAsync(m->m2, 0, NPARAM * 4, m->stream), "memset");
        // He-style uniform init from a fixed LCG, biases zero
        std::vector<float> w(NPARAM, 0.f);
        unsigned long long st = 0x9E3779B97F4A7C15ull ^ seed;
        auto rnd = [&]() {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            return ((st >> 40) & 0xFFFFFF) / 16777216.0f * 2.f - 1.f;
        };
        const int fan[4] = {MI, MH, MH, MH}, offs[4] = {O_W1, O_W2, O_W3, O_W4}, nout[4] = {MH, MH, MH, MO};
        for (int L = 0; L < 4; ++L) {
            const float a = std::sqrt(6.f / fan[L]) * (L == 3 ? 0.1f : 1.f);
            for (int k = 0; k < fan[L] * nout[L]; ++k) w[offs[L] + k] = a * rnd();
        }
        quantise(m.get(), w);
    } catch (const std::exception&) {
        return -1;
    }
    *out = m.release();
    return 0;
}

int hpg_mlp_destroy(hpg_mlp* m) {
    delete m;
    return 0;
}

int hpg_mlp_num_params(void) { return NPARAM; }

int hpg_mlp_get_weights(hpg_mlp* m, float* w) {
    if (!m || !w) return -2;
    try {
        ck(hipMemcpyAsync(w, m->master, NPARAM * 4, hipMemcpyDeviceToHost, m->stream), "D2H");
        ck(hipStreamSynchronize(m->stream), "sync");
    } catch (const std::exception& e) { m->err = e.what(); return -1; }
    return 0;
}

int hpg_mlp_set_weights(hpg_mlp* m, const float* w) {
    if (!m || !w) return -2;
    try {
        quantise(m, std::vector<float>(w, w + NPARAM));
        m->step = 0;
        ck(hipMemsetAsync(m->m1, 0, NPARAM * 4, m->stream), "memset");
        ck(hipMemsetAsync(m->m2, 0, NPARAM * 4, m->stream), "memset");
    } catch (const std::exception& e) { m->err = e.what(); return -1; }
    return 0;
}

// MFMA bf16 16x16x16 lane-layout probe (test support; A, B 16x16 row-major, out 64x4)
int hpg_selftest_mfma_bf16(const float* A, const float* B, float* out) {
    if (!A || !B || !out) return -2;
    float *dA = nullptr, *dB = nullptr, *dO = nullptr;
    if (hipMalloc(&dA, 256 * 4) != hipSuccess || hipMalloc(&dB, 256 * 4) != hipSuccess ||
        hipMalloc(&dO, 256 * 4) != hipSuccess) {
        hipFree(dA); hipFree(dB);
        return -1;
    }
    hipMemcpy(dA, A, 256 * 4, hipMemcpyHostToDevice);
    hipMemcpy(dB, B, 256 * 4, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_mfma_probe_bf16, dim3(1), dim3(64), 0, 0, dA, dB, dO);
    const hipError_t e = hipMemcpy(out, dO, 256 * 4, hipMemcpyDeviceToHost);
    hipFree(dA); hipFree(dB); hipFree(dO);
    return e == hipSuccess ? 0 : -1;
}

int hpg_mlp_forward(hpg_mlp* m, const float* X, int n, float* Y) {
    if (!m || !X || !Y || n < 0) return -2;
    try {
        ensure_n(m, n);
        ck(hipMemcpyAsync(m->X, X, (size_t)n * MI * 4, hipMemcpyHostToDevice, m->stream), "H2D");
        mlp_forward_device(m, m->X, n, m->Y, m->stream);
        ck(hipMemcpyAsync(Y, m->Y, (size_t)n * MO * 4, hipMemcpyDeviceToHost, m->stream), "D2H");
        ck(hipStreamSynchronize(m->stream), "sync");
    } catch (const std::exception& e) { m->err = e.what(); return -1; }
    return 0;
}

// loss (mean square over samples and outputs) and its gradient for one batch, no update
int hpg_mlp_loss_grad(hpg_mlp* m, const float* X, const float* T, int n, float* loss, float* grad) {
    if (!m || !X || !T || n <= 0) return -2;
    try {
        ensure_n(m, n);
        const int nb = (n + 63) / 64;
        ensure_blocks(m, nb);
        ck(hipMemcpyAsync(m->X, X, (size_t)n * MI * 4, hipMemcpyHostToDevice, m->stream), "H2D");
        ck(hipMemcpyAsync(m->T, T, (size_t)n * MO * 4, hipMemcpyHostToDevice, m->stream), "H2D");
        enqueue_loss_grad(m, m->X, m->T, n, m->partial, m->lossp, m->grad);
        std::vector<float> lp(nb);
        ck(hipMemcpyAsync(lp.data(), m->lossp, nb * 4, hipMemcpyDeviceToHost, m->stream), "D2H");
        if (grad) ck(hipMemcpyAsync(grad, m->grad, NPARAM * 4, hipMemcpyDeviceToHost, m->stream), "D2H");
        ck(hipStreamSynchronize(m->stream), "sync");
        double s = 0.0;
        for (float v : lp) s += v;
        if (loss) *loss = (float)(s / ((double)n * MO));
    } catch (const std::exception& e) { m->err = e.what(); return -1; }
    return 0;
}

// Allocation-free form of hpg_mlp_loss_grad on device buffers: X (n x 11), T (n x 3),
// loss (1 float) and grad (NPARAM floats) are device pointers of the caller; the
// block partials live in the library's static scratch array.  Same kernels and
// summation order as hpg_mlp_loss_grad, so loss and gradient are bitwise equal.
int hpg_mlp_loss_grad_noalloc(hpg_mlp* m, const float* X, const float* T, int n, float* loss, float* grad) {
    if (!m || !X || !T || !loss || !grad || n <= 0) return -2;
    const int nb = (n + 63) / 64;
    if (nb > NOALLOC_MAX_BLOCKS) {
        m->err = "hpg_mlp_loss_grad_noalloc: n exceeds the scratch capacity of " +
                 std::to_string(HPG_MLP_NOALLOC_MAX_SAMPLES) + " samples";
        return -2;
    }
    try {
        float* sp = (m->dev >= 0 && m->dev < MAX_DEVICES) ? g_scratch[m->dev] : nullptr;
        if (!sp) throw std::runtime_error("scratch address not resolved for this device");
        float* lossp = sp + (size_t)nb * NPARAM;
        enqueue_loss_grad(m, X, T, n, sp, lossp, grad);
        hipLaunchKernelGGL(k_mlp_loss_sum, dim3(1), dim3(64), 0, m->stream, nb, lossp, n * MO, loss);
        ck(hipGetLastError(), "loss_grad_noalloc launch");
        ck(hipStreamSynchronize(m->stream), "sync");
    } catch (const std::exception& e) { m->err = e.what(); return -1; }
    return 0;
}

// Adam training: `epochs` passes over (X, T) in mini-batches of `batch` samples
// (taken in order), learning rate lr; returns the mean loss of the last epoch
int hpg_mlp_train(hpg_mlp* m, const float* X, const float* T, int n, int epochs, int batch, float lr, float* last) {
    if (!m || !X || !T || n <= 0 || batch <= 0 || epochs <= 0) return -2;
    try {
        ensure_n(m, n);
        batch = ((batch + 63) / 64) * 64;
        ensure_blocks(m, batch / 64);
        ck(hipMemcpyAsync(m->X, X, (size_t)n * MI * 4, hipMemcpyHostToDevice, m->stream), "H2D");
        ck(hipMemcpyAsync(m->T, T, (size_t)n * MO * 4, hipMemcpyHostToDevice, m->stream), "H2D");
        double lsum = 0.0;
        std::vector<float> lp(batch / 64);
        for (int ep = 0; ep < epochs; ++ep) {
            lsum = 0.0;
            for (int b0 = 0; b0 < n; b0 += batch) {
                const int nb_s = std::min(batch, n - b0), nb = (nb_s + 63) / 64;
                ++m->step;
                const float bc1 = 1.f - std::pow(0.9f, (float)m->step), bc2 = 1.f - std::pow(0.999f, (float)m->step);
                hipLaunchKernelGGL(k_mlp_backward, dim3(nb), dim3(64), 0, m->stream, nb_s, nb_s * MO,
                                   m->X + (size_t)b0 * MI, m->T + (size_t)b0 * MO, m->wq, m->partial, m->lossp);
                hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, nb,
                                   m->partial, m->grad, m->master, m->wq, m->m1, m->m2, lr, bc1, bc2, 1);
                if (ep == epochs - 1) {
                    ck(hipMemcpyAsync(lp.data(), m->lossp, nb * 4, hipMemcpyDeviceToHost, m->stream), "D2H");
                    ck(hipStreamSynchronize(m->stream), "sync");
                    for (int k = 0; k < nb; ++k) lsum += lp[k];
                }
            }
            ck(hipGetLastError(), "train launch");
        }
        ck(hipStreamSynchronize(m->stream), "sync");
        if (last) *last = (float)(lsum / ((double)n * MO));
    } catch (const std::exception& e) { m->err = e.what(); return -1; }
    return 0;
}

}  // extern "C"
