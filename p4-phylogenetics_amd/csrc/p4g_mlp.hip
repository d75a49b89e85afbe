// p4g_mlp.hip -- the learned proposal's network: a 3-layer MLP
// (F -> H -> H -> O, ReLU) with BF16 weights and BF16 MFMA operands,
// FP32 accumulation, FP32 master weights and an FP32 Adam update, all on
// the GPU.  It maps per-branch (and per-composition) features to the mean
// and log standard deviation of a Gaussian proposal; it is trained by
// minimising the Gaussian negative log likelihood of samples from
// plain-proposal chains.  The proposal that uses it is Metropolis-Hastings
// corrected (p4gpu/learned.py), so the network only affects mixing.
//
// GEMMs run on v_mfma_f32_16x16x16bf16_1k, one wavefront per 16x16 output
// tile.  Lane l holds A[l%16][4*(l/16)+i], B[4*(l/16)+i][l%16] (i < 4) and
// D rows 4*(l/16)+r of column l%16.
#include "../../include/p4gpu.h"
#include "p4g_internal.h"

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace p4g {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4m __attribute__((ext_vector_type(4)));

__host__ __device__ inline unsigned short f2bf(float f)
{
    unsigned u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h)
{
    return __uint_as_float(((unsigned)h) << 16);
}

// Element (m, k) of A and (k, n) of B; each operand is FP32 (rounded to
// BF16 for the MFMA) or BF16, row-major with a leading dimension, and may
// be transposed.
struct GemmOp {
    const void *p;
    int ld;
    int trans;   // 0: [rows][ld], 1: stored transposed
    int isBf16;
};

__device__ __forceinline__ unsigned short opLoad(const GemmOp &o, int r, int c, int R, int C)
{
    if (r >= R || c >= C) return 0;
    const size_t idx = o.trans ? (size_t)c * o.ld + r : (size_t)r * o.ld + c;
    if (o.isBf16) return ((const unsigned short *)o.p)[idx];
    return f2bf(((const float *)o.p)[idx]);
}

// C[M][N] = A[M][K] . B[K][N] (+ bias[n]) (ReLU) -- FP32 result.
__global__ __launch_bounds__(64) void k_gemm(GemmOp A, GemmOp B, float *C, int ldc, int M, int N, int K,
                                             const float *bias, float *reluOut)
{
    const int lane = threadIdx.x;
    const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
    const int row = lane & 15, kq = (lane >> 4) * 4;
    f4m acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        s4v a, b;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            a[i] = (short)opLoad(A, m0 + row, k0 + kq + i, M, K);
            b[i] = (short)opLoad(B, k0 + kq + i, n0 + row, K, N);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int n = n0 + row;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int m = m0 + kq + r;
        if (m < M && n < N) {
            const float v = acc[r] + (bias ? bias[n] : 0.f);
            C[(size_t)m * ldc + n] = v;
            if (reluOut) reluOut[(size_t)m * ldc + n] = fmaxf(v, 0.f);
        }
    }
}

// Gaussian negative log likelihood of targets y under (mu = out[:, :D],
// s = log sigma = out[:, D:2D]), masked per coordinate; writes dL/dout and
// the per-row loss.  Loss per row: sum_d mask (s + 0.5 (y - mu)^2 e^{-2s}).
__global__ void k_nll(const float *out, const float *y, const float *mask, int n, int D, float invCount,
                      float *dOut, float *rowLoss)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float L = 0.f;
    for (int d = 0; d < D; d++) {
        const float mk = mask[(size_t)i * D + d];
        const float mu = out[(size_t)i * 2 * D + d];
        const float s = out[(size_t)i * 2 * D + D + d];
        const float r = y[(size_t)i * D + d] - mu;
        const float w = expf(-2.f * s);
        L += mk * (s + 0.5f * r * r * w);
        dOut[(size_t)i * 2 * D + d] = -mk * r * w * invCount;
        dOut[(size_t)i * 2 * D + D + d] = mk * (1.f - r * r * w) * invCount;
    }
    rowLoss[i] = L;
}

// dZ = dH * (Z > 0)
__global__ void k_reluBack(float *dH, const float *Z, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(Z[i] > 0.f)) dH[i] = 0.f;
}

// column sums of X[n][N] (bias gradients)
__global__ void k_colsum(const float *X, int n, int N, float *out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float s = 0.f;
    for (int i = 0; i < n; i++) s += X[(size_t)i * N + c];
    out[c] = s;
}

__global__ void k_sum(const float *x, int n, float *out)
{
    __shared__ float sh[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Adam on the FP32 masters, then refresh the BF16 copies.
__global__ void k_adam(float *w, unsigned short *wb, const float *g, float *m, float *v, size_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

__global__ void k_tobf16(const float *w, unsigned short *wb, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf(w[i]);
}

}  // namespace p4g

using namespace p4g;

// Parameter block layout (FP32 masters, one contiguous array):
//   W1 [H][F], b1 [H], W2 [H][H], b2 [H], W3 [O][H], b3 [O]
struct p4g_mlp {
    int F = 0, H = 0, O = 0;
    size_t nW1, nW2, nW3, nParams;
    float *w = nullptr;           // masters
    unsigned short *wb = nullptr; // BF16 copy (biases unused)
    float *g = nullptr, *m = nullptr, *v = nullptr;
    long step = 0;
    // activations for a batch of up to cap rows
    int cap = 0;
    float *X = nullptr, *Z1 = nullptr, *Z2 = nullptr, *OUT = nullptr, *Y = nullptr, *MASK = nullptr;
    float *dOUT = nullptr, *dH2 = nullptr, *dH1 = nullptr, *rowLoss = nullptr, *loss = nullptr;
    size_t off(int which) const
    {
        switch (which) {
        case 0: return 0;                                  // W1
        case 1: return nW1;                                // b1
        case 2: return nW1 + H;                            // W2
        case 3: return nW1 + H + nW2;                      // b2
        case 4: return nW1 + H + nW2 + H;                  // W3
        default: return nW1 + H + nW2 + H + nW3;           // b3
        }
    }
};

namespace {

bool ensureCap(p4g_mlp *m, int n)
{
    if (n <= m->cap) return true;
    float **bufs[] = {&m->X, &m->Z1, &m->Z2, &m->OUT, &m->Y, &m->MASK, &m->dOUT, &m->dH2, &m->dH1, &m->rowLoss};
    for (auto b : bufs)
        if (*b) {
            (void)hipFree(*b);
            *b = nullptr;
        }
    const int D = m->O / 2;
    const size_t sizes[] = {(size_t)n * m->F, (size_t)n * m->H, (size_t)n * m->H, (size_t)n * m->O,
                            (size_t)n * D, (size_t)n * D, (size_t)n * m->O, (size_t)n * m->H,
                            (size_t)n * m->H, (size_t)n};
    for (int i = 0; i < 10; i++)
        if (!hipOk(hipMalloc((void **)bufs[i], sizes[i] * sizeof(float)), "hipMalloc mlp buffers")) return false;
    m->cap = n;
    return true;
}

hipError_t gemm(GemmOp A, GemmOp B, float *C, int ldc, int M, int N, int K, const float *bias,
                float *reluOut = nullptr)
{
    dim3 grid((M + 15) / 16, (N + 15) / 16);
    hipLaunchKernelGGL(k_gemm, grid, dim3(64), 0, stream(), A, B, C, ldc, M, N, K, bias, reluOut);
    return hipGetLastError();
}

// forward of n rows already in m->X
bool forward(p4g_mlp *m, int n)
{
    const unsigned short *wb = m->wb;
    const float *w = m->w;
    GemmOp X{m->X, m->F, 0, 0};
    GemmOp W1t{wb + m->off(0), m->F, 1, 1};  // B(k=f, n=h) = W1[h][f]
    // Z1 is kept for the backward pass, H1 = relu(Z1) goes to dH1's storage
    if (!hipOk(gemm(X, W1t, m->Z1, m->H, n, m->H, m->F, w + m->off(1), m->dH1), "mlp layer 1")) return false;
    GemmOp H1{m->dH1, m->H, 0, 0};
    GemmOp W2t{wb + m->off(2), m->H, 1, 1};
    if (!hipOk(gemm(H1, W2t, m->Z2, m->H, n, m->H, m->H, w + m->off(3), m->dH2), "mlp layer 2")) return false;
    GemmOp H2{m->dH2, m->H, 0, 0};
    GemmOp W3t{wb + m->off(4), m->H, 1, 1};
    return hipOk(gemm(H2, W3t, m->OUT, m->O, n, m->O, m->H, w + m->off(5)), "mlp layer 3");
}

// forward + loss + gradients into m->g for n rows in X, Y, MASK
bool backward(p4g_mlp *m, int n)
{
    if (!forward(m, n)) return false;
    const int D = m->O / 2;
    hipStream_t s = stream();
    hipLaunchKernelGGL(k_nll, dim3((n + 255) / 256), dim3(256), 0, s, m->OUT, m->Y, m->MASK, n, D,
                       1.0f / (float)n, m->dOUT, m->rowLoss);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, m->rowLoss, n, m->loss);
    float *g = m->g;
    // H1 (relu(Z1)) is in dH1, H2 (relu(Z2)) in dH2 after forward(); layer-3
    // gradients first, then overwrite dH2/dH1 with the backpropagated values
    GemmOp dOutT{m->dOUT, m->O, 1, 0};  // A(m=o, k=row) = dOUT[row][o]
    GemmOp H2{m->dH2, m->H, 0, 0};      // B(k=row, n=h)
    if (!hipOk(gemm(dOutT, H2, g + m->off(4), m->H, m->O, m->H, n, nullptr), "mlp dW3")) return false;
    hipLaunchKernelGGL(k_colsum, dim3((m->O + 63) / 64), dim3(64), 0, s, m->dOUT, n, m->O, g + m->off(5));
    GemmOp H1{m->dH1, m->H, 0, 0};
    // dH2 = dOUT . W3   (W3 [O][H]: B(k=o, n=h))
    GemmOp dOut{m->dOUT, m->O, 0, 0};
    GemmOp W3{m->wb + m->off(4), m->H, 0, 1};
    // dH2 must not overwrite H2 before dW3 is done: same stream, ordered
    if (!hipOk(gemm(dOut, W3, m->dH2, m->H, n, m->H, m->O, nullptr), "mlp dH2")) return false;
    const size_t cnt = (size_t)n * m->H;
    hipLaunchKernelGGL(k_reluBack, dim3((cnt + 255) / 256), dim3(256), 0, s, m->dH2, m->Z2, cnt);
    GemmOp dZ2T{m->dH2, m->H, 1, 0};
    if (!hipOk(gemm(dZ2T, H1, g + m->off(2), m->H, m->H, m->H, n, nullptr), "mlp dW2")) return false;
    hipLaunchKernelGGL(k_colsum, dim3((m->H + 63) / 64), dim3(64), 0, s, m->dH2, n, m->H, g + m->off(3));
    GemmOp dZ2{m->dH2, m->H, 0, 0};
    GemmOp W2{m->wb + m->off(2), m->H, 0, 1};
    if (!hipOk(gemm(dZ2, W2, m->dH1, m->H, n, m->H, m->H, nullptr), "mlp dH1")) return false;
    hipLaunchKernelGGL(k_reluBack, dim3((cnt + 255) / 256), dim3(256), 0, s, m->dH1, m->Z1, cnt);
    GemmOp dZ1T{m->dH1, m->H, 1, 0};
    GemmOp X{m->X, m->F, 0, 0};
    if (!hipOk(gemm(dZ1T, X, g + m->off(0), m->F, m->H, m->F, n, nullptr), "mlp dW1")) return false;
    hipLaunchKernelGGL(k_colsum, dim3((m->H + 63) / 64), dim3(64), 0, s, m->dH1, n, m->H, g + m->off(1));
    return hipOk(hipGetLastError(), "mlp backward");
}

}  // namespace

extern "C" {

p4g_mlp *p4g_mlpNew(int nIn, int nHidden, int nOut, unsigned long long seed)
{
    clearError();
    if (!deviceReady()) return nullptr;
    if (nIn < 1 || nHidden < 16 || nOut < 2 || (nOut & 1)) {
        setError("mlpNew: bad sizes (nOut holds means and log sds: even)");
        return nullptr;
    }
    auto *m = new p4g_mlp();
    m->F = nIn;
    m->H = nHidden;
    m->O = nOut;
    m->nW1 = (size_t)nHidden * nIn;
    m->nW2 = (size_t)nHidden * nHidden;
    m->nW3 = (size_t)nOut * nHidden;
    m->nParams = m->nW1 + nHidden + m->nW2 + nHidden + m->nW3 + nOut;
    std::vector<float> h(m->nParams, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](size_t off, size_t n, int fanIn, float gain) {
        std::normal_distribution<float> nd(0.f, gain * std::sqrt(2.0f / (float)fanIn));
        for (size_t i = 0; i < n; i++) h[off + i] = nd(rng);
    };
    fill(m->off(0), m->nW1, nIn, 1.f);
    fill(m->off(2), m->nW2, nHidden, 1.f);
    fill(m->off(4), m->nW3, nHidden, 0.1f);
    const size_t n = m->nParams;
    if (!hipOk(hipMalloc((void **)&m->w, n * 4), "hipMalloc mlp") ||
        !hipOk(hipMalloc((void **)&m->wb, n * 2), "hipMalloc mlp") ||
        !hipOk(hipMalloc((void **)&m->g, n * 4), "hipMalloc mlp") ||
        !hipOk(hipMalloc((void **)&m->m, n * 4), "hipMalloc mlp") ||
        !hipOk(hipMalloc((void **)&m->v, n * 4), "hipMalloc mlp") ||
        !hipOk(hipMalloc((void **)&m->loss, 8), "hipMalloc mlp")) {
        delete m;
        return nullptr;
    }
    hipStream_t s = stream();
    (void)hipMemcpyAsync(m->w, h.data(), n * 4, hipMemcpyHostToDevice, s);
    (void)hipMemsetAsync(m->m, 0, n * 4, s);
    (void)hipMemsetAsync(m->v, 0, n * 4, s);
    hipLaunchKernelGGL(k_tobf16, dim3((n + 255) / 256), dim3(256), 0, s, m->w, m->wb, n);
    if (!hipOk(hipStreamSynchronize(s), "mlp init")) {
        delete m;
        return nullptr;
    }
    return m;
}

void p4g_mlpFree(p4g_mlp *m)
{
    if (!m) return;
    (void)hipStreamSynchronize(stream());
    float *fs[] = {m->w, m->g, m->m, m->v, m->loss, m->X, m->Z1, m->Z2, m->OUT, m->Y, m->MASK, m->dOUT, m->dH2,
                   m->dH1, m->rowLoss};
    for (auto p : fs)
        if (p) (void)hipFree(p);
    if (m->wb) (void)hipFree(m->wb);
    delete m;
}

int p4g_mlpNumParams(const p4g_mlp *m) { return m ? (int)m->nParams : -1; }

int p4g_mlpGetParams(p4g_mlp *m, float *out)
{
    clearError();
    if (!m || !out) return 1;
    return hipOk(hipMemcpy(out, m->w, m->nParams * 4, hipMemcpyDeviceToHost), "mlp get params") ? 0 : 1;
}

int p4g_mlpSetParams(p4g_mlp *m, const float *in)
{
    clearError();
    if (!m || !in) return 1;
    hipStream_t s = stream();
    if (!hipOk(hipMemcpyAsync(m->w, in, m->nParams * 4, hipMemcpyHostToDevice, s), "mlp set params")) return 1;
    hipLaunchKernelGGL(k_tobf16, dim3((m->nParams + 255) / 256), dim3(256), 0, s, m->w, m->wb, m->nParams);
    return hipOk(hipStreamSynchronize(s), "mlp set params") ? 0 : 1;
}

int p4g_mlpForward(p4g_mlp *m, const float *x, int n, float *out)
{
    clearError();
    if (!m || !x || !out || n < 1) return 1;
    if (!ensureCap(m, n)) return 1;
    hipStream_t s = stream();
    if (!hipOk(hipMemcpyAsync(m->X, x, sizeof(float) * (size_t)n * m->F, hipMemcpyHostToDevice, s), "mlp input") ||
        !forward(m, n) ||
        !hipOk(hipMemcpyAsync(out, m->OUT, sizeof(float) * (size_t)n * m->O, hipMemcpyDeviceToHost, s), "mlp out"))
        return 1;
    return hipOk(hipStreamSynchronize(s), "mlp forward") ? 0 : 1;
}

static int uploadBatch(p4g_mlp *m, const float *x, const float *y, const float *mask, int n)
{
    const int D = m->O / 2;
    if (!ensureCap(m, n)) return 1;
    hipStream_t s = stream();
    if (!hipOk(hipMemcpyAsync(m->X, x, sizeof(float) * (size_t)n * m->F, hipMemcpyHostToDevice, s), "mlp x") ||
        !hipOk(hipMemcpyAsync(m->Y, y, sizeof(float) * (size_t)n * D, hipMemcpyHostToDevice, s), "mlp y") ||
        !hipOk(hipMemcpyAsync(m->MASK, mask, sizeof(float) * (size_t)n * D, hipMemcpyHostToDevice, s), "mlp mask"))
        return 1;
    return 0;
}

int p4g_mlpGradients(p4g_mlp *m, const float *x, const float *y, const float *mask, int n, float *grads,
                     float *loss)
{
    clearError();
    if (!m || !x || !y || !mask || n < 1) return 1;
    if (uploadBatch(m, x, y, mask, n) || !backward(m, n)) return 1;
    hipStream_t s = stream();
    if (grads && !hipOk(hipMemcpyAsync(grads, m->g, m->nParams * 4, hipMemcpyDeviceToHost, s), "mlp grads"))
        return 1;
    if (loss && !hipOk(hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, s), "mlp loss")) return 1;
    return hipOk(hipStreamSynchronize(s), "mlp gradients") ? 0 : 1;
}

int p4g_mlpTrainStep(p4g_mlp *m, const float *x, const float *y, const float *mask, int n, float lr,
                     float *loss)
{
    clearError();
    if (!m || !x || !y || !mask || n < 1) return 1;
    if (uploadBatch(m, x, y, mask, n) || !backward(m, n)) return 1;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    hipStream_t s = stream();
    hipLaunchKernelGGL(k_adam, dim3((m->nParams + 255) / 256), dim3(256), 0, s, m->w, m->wb, m->g, m->m, m->v,
                       m->nParams, lr, b1, b2, eps, c1, c2);
    if (loss && !hipOk(hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, s), "mlp loss")) return 1;
    return hipOk(hipStreamSynchronize(s), "mlp train step") ? 0 : 1;
}

}  // extern "C"
