// closure_mlp.hip -- the learned closure: a small multilayer perceptron
// 4 -> H -> H -> 2 (tanh hidden layers, linear output) that maps the
// coarse-grained state of a granular flow (inertial number, packing fraction,
// pressure, granular temperature; standardised by the caller) to the
// effective friction and packing fraction.  Forward pass, mean-squared-error
// gradients and Adam updates are hand-written kernels; the network is small
// enough (H = 32: 1,282 parameters) that one thread carries one sample
// through both layers in registers and FP64 costs nothing that matters.
//
// Parameter layout (one flat array, FP64): W1[H][4], b1[H], W2[H][H], b2[H],
// W3[2][H], b3[2] -- row-major, output index first.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/mercurydpm_closure.h"

namespace {

constexpr int kH = DEM_MLP_HIDDEN;
constexpr int kIn = 4, kOut = 2;
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH,
              kB3 = kW3 + kOut * kH, kP = kB3 + kOut;
static_assert(kP == DEM_MLP_PARAMS, "parameter count");
constexpr int kBlock = 256;

struct Acts {
    double h1[kH], h2[kH], y[kOut];
};

__device__ inline void forward(const double* __restrict__ p, const double* x, Acts& a)
{
    for (int j = 0; j < kH; ++j) {
        double s = p[kB1 + j];
        for (int i = 0; i < kIn; ++i) s += p[kW1 + j * kIn + i] * x[i];
        a.h1[j] = tanh(s);
    }
    for (int j = 0; j < kH; ++j) {
        double s = p[kB2 + j];
        for (int i = 0; i < kH; ++i) s += p[kW2 + j * kH + i] * a.h1[i];
        a.h2[j] = tanh(s);
    }
    for (int k = 0; k < kOut; ++k) {
        double s = p[kB3 + k];
        for (int i = 0; i < kH; ++i) s += p[kW3 + k * kH + i] * a.h2[i];
        a.y[k] = s;
    }
}

__global__ void __launch_bounds__(kBlock) k_mlp_forward(const double* __restrict__ p, int n,
                                                         const double* __restrict__ x, double* __restrict__ y)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double xi[kIn];
    for (int i = 0; i < kIn; ++i) xi[i] = x[kIn * s + i];
    Acts a;
    forward(p, xi, a);
    for (int k = 0; k < kOut; ++k) y[kOut * s + k] = a.y[k];
}

// Gradient of L = (1/n) sum_s sum_k (y_k - t_k)^2 with respect to every
// parameter.  Each thread back-propagates its sample; the block sums the
// per-sample gradients parameter by parameter through LDS (a fixed tree) and
// writes its partial sums, kP gradient entries and then the loss, to row
// blockIdx.x of `part`; k_mlp_sum_blocks adds the rows in block order, so the
// gradient is the same bits on every run.
constexpr int kRow = kP + 1;

__global__ void __launch_bounds__(kBlock) k_mlp_gradient(const double* __restrict__ p, int n,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ t, double* __restrict__ part)
{
    double* row = part + (size_t)blockIdx.x * kRow;
    __shared__ double red[kBlock];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = s < n;
    double xi[kIn] = {0, 0, 0, 0};
    double d3[kOut] = {0, 0};
    Acts a = {};
    double l = 0.0;
    if (live) {
        for (int i = 0; i < kIn; ++i) xi[i] = x[kIn * s + i];
        forward(p, xi, a);
        for (int k = 0; k < kOut; ++k) {
            const double e = a.y[k] - t[kOut * s + k];
            l += e * e;
            d3[k] = 2.0 * e / n;
        }
    }
    // deltas of the hidden layers (zero for threads past the end)
    double d2[kH], d1[kH];
    for (int j = 0; j < kH; ++j) {
        double g = 0.0;
        for (int k = 0; k < kOut; ++k) g += p[kW3 + k * kH + j] * d3[k];
        d2[j] = g * (1.0 - a.h2[j] * a.h2[j]);
    }
    for (int j = 0; j < kH; ++j) {
        double g = 0.0;
        for (int i = 0; i < kH; ++i) g += p[kW2 + i * kH + j] * d2[i];
        d1[j] = g * (1.0 - a.h1[j] * a.h1[j]);
    }
    if (!live) {
        for (int j = 0; j < kH; ++j) d1[j] = d2[j] = 0.0;
    }
    auto block_add = [&](double v, int index) {
        red[threadIdx.x] = v;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) row[index] = red[0];
        __syncthreads();
    };
    for (int j = 0; j < kH; ++j) {
        for (int i = 0; i < kIn; ++i) block_add(d1[j] * xi[i], kW1 + j * kIn + i);
        block_add(d1[j], kB1 + j);
    }
    for (int j = 0; j < kH; ++j) {
        for (int i = 0; i < kH; ++i) block_add(d2[j] * a.h1[i], kW2 + j * kH + i);
        block_add(d2[j], kB2 + j);
    }
    for (int k = 0; k < kOut; ++k) {
        for (int i = 0; i < kH; ++i) block_add(d3[k] * a.h2[i], kW3 + k * kH + i);
        block_add(d3[k], kB3 + k);
    }
    red[threadIdx.x] = l / n;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) row[kP] = red[0];
}

// grad[i] (i < kP) and *loss (i = kP): the sum of column i of the nb rows of
// `part`, in row order.
__global__ void k_mlp_sum_blocks(const double* __restrict__ part, int nb, double* __restrict__ grad,
                                 double* __restrict__ loss)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > kP) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[(size_t)b * kRow + i];
    if (i < kP)
        grad[i] = s;
    else
        *loss = s;
}

// Adam (Kingma & Ba 2015) with bias correction; `step` counts from 1.
__global__ void k_mlp_adam(double* __restrict__ p, const double* __restrict__ g, double* __restrict__ m,
                           double* __restrict__ v, double lr, double b1, double b2, double eps, int step)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kP) return;
    m[i] = b1 * m[i] + (1.0 - b1) * g[i];
    v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
    const double mh = m[i] / (1.0 - pow(b1, (double)step));
    const double vh = v[i] / (1.0 - pow(b2, (double)step));
    p[i] -= lr * mh / (sqrt(vh) + eps);
}

__global__ void k_zero(double* a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

}  // namespace

struct dem_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    double* p = nullptr;  // parameters, gradient, Adam moments, loss: one allocation
    double *g = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr;
    double *x = nullptr, *t = nullptr, *y = nullptr;  // sample buffers
    double* part = nullptr;                            // per-block partial gradients, kRow per block
    int64_t cap = 0;
    int adam_step = 0;
};

namespace {

int ensure(dem_mlp* c, int64_t n)
{
    if (n <= c->cap) return DEM_OK;
    if (c->x) (void)hipFree(c->x);
    c->x = c->t = c->y = nullptr;
    c->cap = 0;
    const size_t nb = (size_t)((n + kBlock - 1) / kBlock);
    if (hipMalloc(&c->x, sizeof(double) * ((size_t)n * (kIn + 2 * kOut) + nb * kRow)) != hipSuccess)
        return DEM_ERR_DEVICE;
    c->t = c->x + (size_t)n * kIn;
    c->y = c->t + (size_t)n * kOut;
    c->part = c->y + (size_t)n * kOut;
    c->cap = n;
    return DEM_OK;
}

unsigned blocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// the loss and gradient of the samples on the device, into c->loss and c->g
void launch_gradient(dem_mlp* c, int64_t n)
{
    hipLaunchKernelGGL(k_mlp_gradient, dim3(blocks(n)), dim3(kBlock), 0, c->stream, (const double*)c->p, (int)n,
                       (const double*)c->x, (const double*)c->t, c->part);
    hipLaunchKernelGGL(k_mlp_sum_blocks, dim3(blocks(kRow)), dim3(kBlock), 0, c->stream, (const double*)c->part,
                       (int)blocks(n), c->g, c->loss);
}

int gradient(dem_mlp* c, int64_t n, const double* x, const double* t, double* loss_out)
{
    if (ensure(c, n)) return DEM_ERR_DEVICE;
    if (hipMemcpyAsync(c->x, x, sizeof(double) * n * kIn, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->t, t, sizeof(double) * n * kOut, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return DEM_ERR_DEVICE;
    launch_gradient(c, n);
    if (hipGetLastError() != hipSuccess) return DEM_ERR_DEVICE;
    if (loss_out && (hipMemcpyAsync(loss_out, c->loss, sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
                         hipSuccess ||
                     hipStreamSynchronize(c->stream) != hipSuccess))
        return DEM_ERR_DEVICE;
    return DEM_OK;
}

}  // namespace

extern "C" {

int dem_mlp_create(dem_mlp** out, int32_t device)
{
    if (!out) return DEM_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return DEM_ERR_ARG;
    auto* c = new dem_mlp;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipMalloc(&c->p, sizeof(double) * (4 * kP + 1)) != hipSuccess) {
        dem_mlp_destroy(c);
        return DEM_ERR_DEVICE;
    }
    c->g = c->p + kP;
    c->m = c->g + kP;
    c->v = c->m + kP;
    c->loss = c->v + kP;
    hipLaunchKernelGGL(k_zero, dim3(blocks(4 * kP + 1)), dim3(kBlock), 0, c->stream, c->p, 4 * kP + 1);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        dem_mlp_destroy(c);
        return DEM_ERR_DEVICE;
    }
    *out = c;
    return DEM_OK;
}

int dem_mlp_destroy(dem_mlp* c)
{
    if (!c) return DEM_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->x) (void)hipFree(c->x);
    if (c->p) (void)hipFree(c->p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DEM_OK;
}

int dem_mlp_set_params(dem_mlp* c, const double* params)
{
    if (!c || !params) return DEM_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess ||
        hipMemcpy(c->p, params, sizeof(double) * kP, hipMemcpyHostToDevice) != hipSuccess)
        return DEM_ERR_DEVICE;
    hipLaunchKernelGGL(k_zero, dim3(blocks(2 * kP)), dim3(kBlock), 0, c->stream, c->m, 2 * kP);
    c->adam_step = 0;
    return hipStreamSynchronize(c->stream) == hipSuccess ? DEM_OK : DEM_ERR_DEVICE;
}

int dem_mlp_get_params(dem_mlp* c, double* params)
{
    if (!c || !params) return DEM_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess ||
        hipMemcpy(params, c->p, sizeof(double) * kP, hipMemcpyDeviceToHost) != hipSuccess)
        return DEM_ERR_DEVICE;
    return DEM_OK;
}

int dem_mlp_forward(dem_mlp* c, int64_t n, const double* x, double* y)
{
    if (!c || n < 0 || (n > 0 && (!x || !y)) || n > INT32_MAX) return DEM_ERR_ARG;
    if (n == 0) return DEM_OK;
    if (hipSetDevice(c->device) != hipSuccess || ensure(c, n)) return DEM_ERR_DEVICE;
    if (hipMemcpyAsync(c->x, x, sizeof(double) * n * kIn, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return DEM_ERR_DEVICE;
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks(n)), dim3(kBlock), 0, c->stream, (const double*)c->p, (int)n,
                       (const double*)c->x, c->y);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(y, c->y, sizeof(double) * n * kOut, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return DEM_ERR_DEVICE;
    return DEM_OK;
}

int dem_mlp_gradient(dem_mlp* c, int64_t n, const double* x, const double* target, double* grad, double* loss)
{
    if (!c || n <= 0 || !x || !target || n > INT32_MAX) return DEM_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return DEM_ERR_DEVICE;
    int r = gradient(c, n, x, target, loss);
    if (r) return r;
    if (grad && (hipMemcpy(grad, c->g, sizeof(double) * kP, hipMemcpyDeviceToHost) != hipSuccess)) return DEM_ERR_DEVICE;
    return DEM_OK;
}

int dem_mlp_train(dem_mlp* c, int64_t n, const double* x, const double* target, int32_t epochs, double lr,
                  double* loss)
{
    if (!c || n <= 0 || !x || !target || epochs < 0 || !(lr > 0) || n > INT32_MAX) return DEM_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return DEM_ERR_DEVICE;
    // the samples stay on the device; every epoch is one full-batch step
    if (ensure(c, n)) return DEM_ERR_DEVICE;
    if (hipMemcpyAsync(c->x, x, sizeof(double) * n * kIn, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->t, target, sizeof(double) * n * kOut, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return DEM_ERR_DEVICE;
    for (int e = 0; e < epochs; ++e) {
        launch_gradient(c, n);
        hipLaunchKernelGGL(k_mlp_adam, dim3(blocks(kP)), dim3(kBlock), 0, c->stream, c->p, (const double*)c->g,
                           c->m, c->v, lr, 0.9, 0.999, 1e-8, ++c->adam_step);
    }
    if (hipGetLastError() != hipSuccess) return DEM_ERR_DEVICE;
    if (loss && hipMemcpyAsync(loss, c->loss, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return DEM_ERR_DEVICE;
    return hipStreamSynchronize(c->stream) == hipSuccess ? DEM_OK : DEM_ERR_DEVICE;
}

}  // extern "C"
