// closure_mfma.hip -- the learned closure as configured for the rheology of
// dense chute flow: a multilayer perceptron 4 -> 128 -> 128 -> 128 -> 2
// (GELU hidden layers, linear output) with BF16 weights and FP32
// accumulation, trained by full-batch Adam on FP32 master weights.
//
// The two 128 x 128 layers are GEMMs over the samples and run on the matrix
// cores: v_mfma_f32_16x16x16_bf16 tiles, one 16 x 16 output tile per
// wavefront, for the forward pass (Z = H W^T), the data gradient
// (dH = dZ W) and the weight gradient (dW = dZ^T H, summed over the samples
// in fixed chunks, then over the chunks in order, so the gradient is
// deterministic).  The 4-wide input layer and the 2-wide output layer are too
// narrow for a tile and run on the vector units with the same BF16 weights.
// Activations fed to a matrix instruction are rounded to BF16 (round to
// nearest even); pre-activations, GELU and every sum stay FP32.
//
// Parameter layout (FP32, one flat array): W1[128][4], b1[128],
// W2[128][128], b2[128], W3[128][128], b3[128], W4[2][128], b4[2].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mercurydpm_closure.h"

namespace {

constexpr int kH = DEM_CLOSURE_HIDDEN;  // 128
constexpr int kIn = 4, kOut = 2;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kOut * kH, kP = oB4 + kOut;
static_assert(kP == DEM_CLOSURE_PARAMS, "parameter count");
constexpr int kBlock = 256;
constexpr int kChunks = 16;  // sample chunks of the weight-gradient sums

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite values)
__host__ __device__ inline uint16_t to_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// GELU, 0.5 z (1 + erf(z / sqrt 2)), and its derivative
__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__global__ void k_cast_bf16(const float* __restrict__ p, uint16_t* __restrict__ q, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = to_bf16(p[i]);
}

// input layer: Z1 = x W1^T + b1, H1 = bf16(gelu(Z1)); rows >= n are zero
__global__ void k_layer_in(const float* __restrict__ x, const uint16_t* __restrict__ wb, const float* __restrict__ p,
                           int n, int npad, float* __restrict__ z, uint16_t* __restrict__ h)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npad * kH) return;
    const int s = t / kH, o = t % kH;
    float v = 0.0f;
    if (s < n) {
        v = p[oB1 + o];
        for (int k = 0; k < kIn; ++k) v += x[kIn * s + k] * from_bf16(wb[oW1 + o * kIn + k]);
    }
    z[t] = v;
    h[t] = s < n ? to_bf16(gelu(v)) : (uint16_t)0;
}

// hidden layer on the matrix cores: Z = H_in W^T + b, H = bf16(gelu(Z)).
// One wavefront per 16 x 16 tile (16 samples x 16 outputs), K = 128 in eight
// 16-deep steps.  Operand layout of v_mfma_f32_16x16x16_bf16: lane l holds
// A[l % 16][4 (l / 16) + j] and B[4 (l / 16) + j][l % 16] (j = 0..3) and
// receives D[4 (l / 16) + i][l % 16] (i = 0..3).
__global__ void __launch_bounds__(kBlock) k_layer_mfma(const uint16_t* __restrict__ hin,
                                                       const uint16_t* __restrict__ w, const float* __restrict__ b,
                                                       int n, int ntiles, float* __restrict__ z,
                                                       uint16_t* __restrict__ h)
{
    const int tile = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (tile >= ntiles) return;
    const int lane = threadIdx.x & 63;
    const int st = tile / (kH / 16), ot = tile % (kH / 16);
    const int r = lane & 15, q = lane >> 4;
    const short4_t* arow = (const short4_t*)(hin + (size_t)(st * 16 + r) * kH + 4 * q);
    const short4_t* brow = (const short4_t*)(w + (size_t)(ot * 16 + r) * kH + 4 * q);
    float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < kH / 16; ++kk) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arow[4 * kk], brow[4 * kk], acc, 0, 0, 0);
    const int col = ot * 16 + r;
    const float bias = b[col];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int s = st * 16 + 4 * q + i;
        const float v = s < n ? acc[i] + bias : 0.0f;
        z[(size_t)s * kH + col] = v;
        h[(size_t)s * kH + col] = s < n ? to_bf16(gelu(v)) : (uint16_t)0;
    }
}

// output layer: y = H3 W4^T + b4
__global__ void k_layer_out(const uint16_t* __restrict__ h, const uint16_t* __restrict__ wb,
                            const float* __restrict__ p, int n, float* __restrict__ y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kOut) return;
    const int s = t / kOut, o = t % kOut;
    float v = p[oB4 + o];
    for (int k = 0; k < kH; ++k) v += from_bf16(h[(size_t)s * kH + k]) * from_bf16(wb[oW4 + o * kH + k]);
    y[t] = v;
}

// dY = 2 (y - t) / n (zero on padded rows) and the per-sample squared error
__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ tgt, int n, int npad,
                       float* __restrict__ dy, float* __restrict__ se)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npad) return;
    float e2 = 0.0f;
    for (int o = 0; o < kOut; ++o) {
        const float e = s < n ? y[kOut * s + o] - tgt[kOut * s + o] : 0.0f;
        dy[kOut * s + o] = 2.0f * e / (float)n;
        e2 += e * e;
    }
    se[s] = e2;
}

// back through the output layer: dZ3 = (dY W4) * gelu'(Z3), FP32 and BF16
__global__ void k_back_out(const float* __restrict__ dy, const uint16_t* __restrict__ wb,
                           const float* __restrict__ z3, int npad, float* __restrict__ dz, uint16_t* __restrict__ dzb)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npad * kH) return;
    const int s = t / kH, k = t % kH;
    float g = 0.0f;
    for (int o = 0; o < kOut; ++o) g += dy[kOut * s + o] * from_bf16(wb[oW4 + o * kH + k]);
    g *= gelu_grad(z3[t]);
    dz[t] = g;
    dzb[t] = to_bf16(g);
}

// data gradient on the matrix cores: dH = dZ W, then dZ_prev = dH *
// gelu'(Z_prev) (FP32 and BF16).  A = dZ (samples x outputs), B = W
// (outputs x inputs).
__global__ void __launch_bounds__(kBlock) k_back_data_mfma(const uint16_t* __restrict__ dzb,
                                                           const uint16_t* __restrict__ w,
                                                           const float* __restrict__ zprev, int ntiles,
                                                           float* __restrict__ dzp, uint16_t* __restrict__ dzpb)
{
    const int tile = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (tile >= ntiles) return;
    const int lane = threadIdx.x & 63;
    const int st = tile / (kH / 16), kt = tile % (kH / 16);
    const int r = lane & 15, q = lane >> 4;
    const short4_t* arow = (const short4_t*)(dzb + (size_t)(st * 16 + r) * kH + 4 * q);
    float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int oo = 0; oo < kH / 16; ++oo) {
        short4_t bv;
#pragma unroll
        for (int j = 0; j < 4; ++j) bv[j] = (short)w[(size_t)(oo * 16 + 4 * q + j) * kH + kt * 16 + r];
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arow[4 * oo], bv, acc, 0, 0, 0);
    }
    const int col = kt * 16 + r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const size_t at = (size_t)(st * 16 + 4 * q + i) * kH + col;
        const float g = acc[i] * gelu_grad(zprev[at]);
        dzp[at] = g;
        dzpb[at] = to_bf16(g);
    }
}

// weight gradient on the matrix cores, per sample chunk: dW_c = dZ^T H over
// the chunk's samples.  A[o][s] = dZ[s][o], B[s][k] = H[s][k]; one
// wavefront per (16 x 16 tile of dW, chunk).
__global__ void __launch_bounds__(kBlock) k_back_weight_mfma(const uint16_t* __restrict__ dzb,
                                                             const uint16_t* __restrict__ h, int nst,
                                                             float* __restrict__ part)
{
    const int item = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    constexpr int tiles = (kH / 16) * (kH / 16);
    if (item >= tiles * kChunks) return;
    const int c = item / tiles, tile = item % tiles;
    const int ot = tile / (kH / 16), kt = tile % (kH / 16);
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, q = lane >> 4;
    const int per = (nst + kChunks - 1) / kChunks;
    const int s0 = c * per, s1 = min(nst, s0 + per);
    float4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int st = s0; st < s1; ++st) {
        short4_t av, bv;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t row = (size_t)(st * 16 + 4 * q + j) * kH;
            av[j] = (short)dzb[row + ot * 16 + r];
            bv[j] = (short)h[row + kt * 16 + r];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    float* out = part + (size_t)c * kH * kH;
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(size_t)(ot * 16 + 4 * q + i) * kH + kt * 16 + r] = acc[i];
}

// the chunk partials of a weight gradient summed in chunk order
__global__ void k_sum_chunks(const float* __restrict__ part, int len, float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    float s = 0.0f;
    for (int c = 0; c < kChunks; ++c) s += part[(size_t)c * len + i];
    out[i] = s;
}

// column sums over the samples (bias gradients) and the small weight
// gradients of the input and output layers, one thread per entry, samples
// in order: out[o] = sum_s a[s][o] (bias, kind 0); dW1[o][k] = sum_s dZ1[s][o]
// x[s][k] (kind 1); dW4[o][k] = sum_s dY[s][o] H3[s][k] (kind 2)
__global__ void k_small_grads(const float* __restrict__ dz1, const float* __restrict__ dz2,
                              const float* __restrict__ dz3, const float* __restrict__ dy,
                              const float* __restrict__ x, const uint16_t* __restrict__ h3, int n,
                              float* __restrict__ g)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < kH * kIn) {  // dW1
        const int o = t / kIn, k = t % kIn;
        float s = 0.0f;
        for (int i = 0; i < n; ++i) s += dz1[(size_t)i * kH + o] * x[kIn * i + k];
        g[oW1 + t] = s;
    } else if (t < kH * kIn + 3 * kH) {  // b1, b2, b3
        const int u = t - kH * kIn, l = u / kH, o = u % kH;
        const float* a = l == 0 ? dz1 : (l == 1 ? dz2 : dz3);
        float s = 0.0f;
        for (int i = 0; i < n; ++i) s += a[(size_t)i * kH + o];
        g[(l == 0 ? oB1 : (l == 1 ? oB2 : oB3)) + o] = s;
    } else if (t < kH * kIn + 3 * kH + kOut * kH) {  // dW4
        const int u = t - kH * kIn - 3 * kH, o = u / kH, k = u % kH;
        float s = 0.0f;
        for (int i = 0; i < n; ++i) s += dy[kOut * i + o] * from_bf16(h3[(size_t)i * kH + k]);
        g[oW4 + u] = s;
    } else if (t < kH * kIn + 3 * kH + kOut * kH + kOut) {  // b4
        const int o = t - kH * kIn - 3 * kH - kOut * kH;
        float s = 0.0f;
        for (int i = 0; i < n; ++i) s += dy[kOut * i + o];
        g[oB4 + o] = s;
    }
}

__global__ void k_sum_loss(const float* __restrict__ se, int n, float* __restrict__ loss)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float s = 0.0f;
    for (int i = 0; i < n; ++i) s += se[i];
    loss[0] = s / (float)n;
}

__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, int n, float lr, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    m[i] = 0.9f * m[i] + 0.1f * gi;
    v[i] = 0.999f * v[i] + 0.001f * gi * gi;
    p[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + 1e-8f);
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

struct dem_closure {
    int device = 0;
    hipStream_t stream = nullptr;
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    uint16_t* wb = nullptr;
    int64_t step = 0;
    // per-sample buffers, grown on demand
    int64_t cap = 0;
    float *x = nullptr, *t = nullptr, *y = nullptr, *dy = nullptr, *se = nullptr, *loss = nullptr;
    float *z1 = nullptr, *z2 = nullptr, *z3 = nullptr, *d1 = nullptr, *d2 = nullptr, *d3 = nullptr, *part = nullptr;
    uint16_t *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *e1 = nullptr, *e2 = nullptr, *e3 = nullptr;
};

namespace {

#define CCHK(expr)                                  \
    do {                                            \
        if ((expr) != hipSuccess) return DEM_ERR_DEVICE; \
    } while (0)

void free_samples(dem_closure* c)
{
    for (void* a : {(void*)c->x, (void*)c->t, (void*)c->y, (void*)c->dy, (void*)c->se, (void*)c->loss, (void*)c->z1,
                    (void*)c->z2, (void*)c->z3, (void*)c->d1, (void*)c->d2, (void*)c->d3, (void*)c->part,
                    (void*)c->h1, (void*)c->h2, (void*)c->h3, (void*)c->e1, (void*)c->e2, (void*)c->e3})
        if (a) (void)hipFree(a);
    c->x = c->t = c->y = c->dy = c->se = c->loss = c->z1 = c->z2 = c->z3 = c->d1 = c->d2 = c->d3 = c->part = nullptr;
    c->h1 = c->h2 = c->h3 = c->e1 = c->e2 = c->e3 = nullptr;
    c->cap = 0;
}

int ensure(dem_closure* c, int64_t n)
{
    const int64_t npad = (n + 15) / 16 * 16;
    if (npad <= c->cap) return DEM_OK;
    CCHK(hipStreamSynchronize(c->stream));
    free_samples(c);
    const size_t a = (size_t)npad * kH;
    CCHK(hipMalloc(&c->x, sizeof(float) * kIn * npad));
    CCHK(hipMalloc(&c->t, sizeof(float) * kOut * npad));
    CCHK(hipMalloc(&c->y, sizeof(float) * kOut * npad));
    CCHK(hipMalloc(&c->dy, sizeof(float) * kOut * npad));
    CCHK(hipMalloc(&c->se, sizeof(float) * npad));
    CCHK(hipMalloc(&c->loss, sizeof(float)));
    for (float** b : {&c->z1, &c->z2, &c->z3, &c->d1, &c->d2, &c->d3}) CCHK(hipMalloc(b, sizeof(float) * a));
    for (uint16_t** b : {&c->h1, &c->h2, &c->h3, &c->e1, &c->e2, &c->e3}) CCHK(hipMalloc(b, sizeof(uint16_t) * a));
    CCHK(hipMalloc(&c->part, sizeof(float) * kChunks * kH * kH));
    c->cap = npad;
    return DEM_OK;
}

// the forward pass of the samples in c->x (n of them, npad rows)
void forward(dem_closure* c, int n)
{
    const int npad = (n + 15) / 16 * 16;
    const int ntiles = (npad / 16) * (kH / 16);
    hipLaunchKernelGGL(k_cast_bf16, dim3(blocks(kP)), dim3(kBlock), 0, c->stream, c->p, c->wb, kP);
    hipLaunchKernelGGL(k_layer_in, dim3(blocks((int64_t)npad * kH)), dim3(kBlock), 0, c->stream, c->x, c->wb, c->p, n,
                       npad, c->z1, c->h1);
    hipLaunchKernelGGL(k_layer_mfma, dim3((ntiles + 3) / 4), dim3(kBlock), 0, c->stream, c->h1, c->wb + oW2,
                       c->p + oB2, n, ntiles, c->z2, c->h2);
    hipLaunchKernelGGL(k_layer_mfma, dim3((ntiles + 3) / 4), dim3(kBlock), 0, c->stream, c->h2, c->wb + oW3,
                       c->p + oB3, n, ntiles, c->z3, c->h3);
    hipLaunchKernelGGL(k_layer_out, dim3(blocks((int64_t)n * kOut)), dim3(kBlock), 0, c->stream, c->h3, c->wb, c->p,
                       n, c->y);
}

// loss and gradient (into c->g) of the samples in c->x / c->t
void backward(dem_closure* c, int n)
{
    const int npad = (n + 15) / 16 * 16;
    const int ntiles = (npad / 16) * (kH / 16);
    const int nst = npad / 16;
    hipLaunchKernelGGL(k_loss, dim3(blocks(npad)), dim3(kBlock), 0, c->stream, c->y, c->t, n, npad, c->dy, c->se);
    hipLaunchKernelGGL(k_back_out, dim3(blocks((int64_t)npad * kH)), dim3(kBlock), 0, c->stream, c->dy, c->wb, c->z3,
                       npad, c->d3, c->e3);
    const unsigned wblocks = (unsigned)(((kH / 16) * (kH / 16) * kChunks + 3) / 4);
    // layer 3: dW3 = dZ3^T H2, dZ2 = (dZ3 W3) * gelu'(Z2)
    hipLaunchKernelGGL(k_back_weight_mfma, dim3(wblocks), dim3(kBlock), 0, c->stream, c->e3, c->h2, nst, c->part);
    hipLaunchKernelGGL(k_sum_chunks, dim3(blocks(kH * kH)), dim3(kBlock), 0, c->stream, c->part, kH * kH, c->g + oW3);
    hipLaunchKernelGGL(k_back_data_mfma, dim3((ntiles + 3) / 4), dim3(kBlock), 0, c->stream, c->e3, c->wb + oW3,
                       c->z2, ntiles, c->d2, c->e2);
    // layer 2: dW2 = dZ2^T H1, dZ1 = (dZ2 W2) * gelu'(Z1)
    hipLaunchKernelGGL(k_back_weight_mfma, dim3(wblocks), dim3(kBlock), 0, c->stream, c->e2, c->h1, nst, c->part);
    hipLaunchKernelGGL(k_sum_chunks, dim3(blocks(kH * kH)), dim3(kBlock), 0, c->stream, c->part, kH * kH, c->g + oW2);
    hipLaunchKernelGGL(k_back_data_mfma, dim3((ntiles + 3) / 4), dim3(kBlock), 0, c->stream, c->e2, c->wb + oW2,
                       c->z1, ntiles, c->d1, c->e1);
    hipLaunchKernelGGL(k_small_grads, dim3(blocks(kH * kIn + 3 * kH + kOut * kH + kOut)), dim3(kBlock), 0, c->stream,
                       c->d1, c->d2, c->d3, c->dy, c->x, c->h3, n, c->g);
    hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(1), 0, c->stream, c->se, n, c->loss);
}

int upload(dem_closure* c, int64_t n, const float* x, const float* t)
{
    if (ensure(c, n)) return DEM_ERR_DEVICE;
    CCHK(hipMemcpyAsync(c->x, x, sizeof(float) * kIn * n, hipMemcpyHostToDevice, c->stream));
    if (t) CCHK(hipMemcpyAsync(c->t, t, sizeof(float) * kOut * n, hipMemcpyHostToDevice, c->stream));
    return DEM_OK;
}

bool finite(const float* a, int64_t n)
{
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

}  // namespace

extern "C" {

int dem_closure_create(dem_closure** out, int32_t device)
{
    if (!out) return DEM_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return DEM_ERR_ARG;
    dem_closure* c = new dem_closure();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->p, sizeof(float) * kP) != hipSuccess || hipMalloc(&c->m, sizeof(float) * kP) != hipSuccess ||
        hipMalloc(&c->v, sizeof(float) * kP) != hipSuccess || hipMalloc(&c->g, sizeof(float) * kP) != hipSuccess ||
        hipMalloc(&c->wb, sizeof(uint16_t) * kP) != hipSuccess || hipMemset(c->p, 0, sizeof(float) * kP) != hipSuccess ||
        hipMemset(c->m, 0, sizeof(float) * kP) != hipSuccess || hipMemset(c->v, 0, sizeof(float) * kP) != hipSuccess) {
        dem_closure_destroy(c);
        return DEM_ERR_DEVICE;
    }
    *out = c;
    return DEM_OK;
}

int dem_closure_destroy(dem_closure* c)
{
    if (!c) return DEM_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_samples(c);
    for (void* a : {(void*)c->p, (void*)c->m, (void*)c->v, (void*)c->g, (void*)c->wb})
        if (a) (void)hipFree(a);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DEM_OK;
}

int dem_closure_set_params(dem_closure* c, const float* params)
{
    if (!c || !params || !finite(params, kP)) return DEM_ERR_ARG;
    CCHK(hipSetDevice(c->device));
    CCHK(hipMemcpyAsync(c->p, params, sizeof(float) * kP, hipMemcpyHostToDevice, c->stream));
    CCHK(hipMemsetAsync(c->m, 0, sizeof(float) * kP, c->stream));
    CCHK(hipMemsetAsync(c->v, 0, sizeof(float) * kP, c->stream));
    CCHK(hipStreamSynchronize(c->stream));
    c->step = 0;
    return DEM_OK;
}

int dem_closure_get_params(dem_closure* c, float* params)
{
    if (!c || !params) return DEM_ERR_ARG;
    CCHK(hipSetDevice(c->device));
    CCHK(hipMemcpyAsync(params, c->p, sizeof(float) * kP, hipMemcpyDeviceToHost, c->stream));
    CCHK(hipStreamSynchronize(c->stream));
    return DEM_OK;
}

int dem_closure_forward(dem_closure* c, int64_t n, const float* x, float* y)
{
    if (!c || n < 0 || n > (1 << 24) || (n > 0 && (!x || !y))) return DEM_ERR_ARG;
    if (n == 0) return DEM_OK;
    if (!finite(x, kIn * n)) return DEM_ERR_ARG;
    CCHK(hipSetDevice(c->device));
    int r = upload(c, n, x, nullptr);
    if (r) return r;
    forward(c, (int)n);
    CCHK(hipGetLastError());
    CCHK(hipMemcpyAsync(y, c->y, sizeof(float) * kOut * n, hipMemcpyDeviceToHost, c->stream));
    CCHK(hipStreamSynchronize(c->stream));
    return DEM_OK;
}

int dem_closure_gradient(dem_closure* c, int64_t n, const float* x, const float* target, float* grad, float* loss)
{
    if (!c || n <= 0 || n > (1 << 24) || !x || !target) return DEM_ERR_ARG;
    if (!finite(x, kIn * n) || !finite(target, kOut * n)) return DEM_ERR_ARG;
    CCHK(hipSetDevice(c->device));
    int r = upload(c, n, x, target);
    if (r) return r;
    forward(c, (int)n);
    backward(c, (int)n);
    CCHK(hipGetLastError());
    if (grad) CCHK(hipMemcpyAsync(grad, c->g, sizeof(float) * kP, hipMemcpyDeviceToHost, c->stream));
    if (loss) CCHK(hipMemcpyAsync(loss, c->loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    CCHK(hipStreamSynchronize(c->stream));
    return DEM_OK;
}

int dem_closure_train(dem_closure* c, int64_t n, const float* x, const float* target, int32_t epochs, float lr,
                      float* loss)
{
    if (!c || n <= 0 || n > (1 << 24) || !x || !target || epochs < 0 || !(lr > 0) || !std::isfinite(lr))
        return DEM_ERR_ARG;
    if (!finite(x, kIn * n) || !finite(target, kOut * n)) return DEM_ERR_ARG;
    CCHK(hipSetDevice(c->device));
    int r = upload(c, n, x, target);
    if (r) return r;
    for (int e = 0; e < epochs; ++e) {
        forward(c, (int)n);
        backward(c, (int)n);
        ++c->step;
        const float c1 = 1.0f - std::pow(0.9f, (float)c->step), c2 = 1.0f - std::pow(0.999f, (float)c->step);
        hipLaunchKernelGGL(k_adam, dim3(blocks(kP)), dim3(kBlock), 0, c->stream, c->p, c->m, c->v, c->g, kP, lr, c1,
                           c2);
    }
    CCHK(hipGetLastError());
    if (loss) CCHK(hipMemcpyAsync(loss, c->loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    CCHK(hipStreamSynchronize(c->stream));
    return DEM_OK;
}

}  // extern "C"
