// Kernels and host side of the learned sweep cost model (costmodel.h).
//
// Inference: one thread per row; the parameters are read at wave-uniform
// addresses (scalar loads), the hidden activations stay in registers.
// Training: one workgroup per 128 rows computes the forward pass, the loss
// gradient and the backward pass of its rows, stages activations and deltas
// in LDS, and reduces them to per-workgroup parameter gradients; a second
// kernel sums the workgroup partials in a fixed order (deterministic), and a
// third applies Adam.
#include "costmodel.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"

namespace q1t {

#define CM_HIP_OK(expr)                                                                            \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) throw Error(ERR_HIP, std::string("HIP error in cost model: ") + hipGetErrorString(e_)); \
    } while (0)

__global__ void __launch_bounds__(256) k_cm_forward(const float* __restrict__ P, const float* __restrict__ norm,
                                                    const float* __restrict__ x, int nrows, float* __restrict__ y)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    float xin[kCmFeat];
#pragma unroll
    for (int f = 0; f < kCmFeat; ++f) xin[f] = (x[size_t(r) * kCmFeat + f] - norm[f]) / norm[kCmFeat + f];
    float h1[kCmHid];
#pragma unroll
    for (int j = 0; j < kCmHid; ++j) {
        float a = P[kCmB1 + j];
#pragma unroll
        for (int f = 0; f < kCmFeat; ++f) a = fmaf(P[kCmW1 + j * kCmFeat + f], xin[f], a);
        h1[j] = tanhf(a);
    }
    float out = P[kCmB3];
    for (int j = 0; j < kCmHid; ++j) {
        float a = P[kCmB2 + j];
#pragma unroll
        for (int i = 0; i < kCmHid; ++i) a = fmaf(P[kCmW2 + j * kCmHid + i], h1[i], a);
        out = fmaf(P[kCmW3 + j], tanhf(a), out);
    }
    y[r] = out;
}

// Per-workgroup gradient of L = inv_n * sum_r (f(x_r) - t_r)^2 over rows
// [blockIdx.x * 128, +128) of the batch; partial[blockIdx.x][kCmParams],
// loss_partial[blockIdx.x].
__global__ void __launch_bounds__(kCmRowsPerBlock) k_cm_grad(const float* __restrict__ P,
                                                             const float* __restrict__ norm,
                                                             const float* __restrict__ x,
                                                             const float* __restrict__ t, int nrows, float inv_n,
                                                             float* __restrict__ partial,
                                                             double* __restrict__ loss_partial)
{
    constexpr int NR = kCmRowsPerBlock;
    __shared__ float xs[NR][kCmFeat];
    __shared__ float h1s[NR][kCmHid + 1];
    __shared__ float d2s[NR][kCmHid + 1];
    __shared__ float ws[NR][kCmHid + 1];   // g*h2 first, then d1
    __shared__ float gs[NR];
    __shared__ double ls[NR];
    const int tid = threadIdx.x;
    const int r = blockIdx.x * NR + tid;
    const bool live = r < nrows;

    float xin[kCmFeat];
#pragma unroll
    for (int f = 0; f < kCmFeat; ++f)
        xin[f] = live ? (x[size_t(r) * kCmFeat + f] - norm[f]) / norm[kCmFeat + f] : 0.0f;
    float h1[kCmHid], h2[kCmHid];
#pragma unroll
    for (int j = 0; j < kCmHid; ++j) {
        float a = P[kCmB1 + j];
#pragma unroll
        for (int f = 0; f < kCmFeat; ++f) a = fmaf(P[kCmW1 + j * kCmFeat + f], xin[f], a);
        h1[j] = tanhf(a);
    }
    float y = P[kCmB3];
#pragma unroll
    for (int j = 0; j < kCmHid; ++j) {
        float a = P[kCmB2 + j];
#pragma unroll
        for (int i = 0; i < kCmHid; ++i) a = fmaf(P[kCmW2 + j * kCmHid + i], h1[i], a);
        h2[j] = tanhf(a);
        y = fmaf(P[kCmW3 + j], h2[j], y);
    }
    const float e = live ? y - t[r] : 0.0f;
    const float g = 2.0f * e * inv_n;   // dL/dy
    ls[tid] = double(e) * double(e) * double(inv_n);
    gs[tid] = g;
#pragma unroll
    for (int f = 0; f < kCmFeat; ++f) xs[tid][f] = xin[f];
#pragma unroll
    for (int j = 0; j < kCmHid; ++j) {
        h1s[tid][j] = h1[j];
        ws[tid][j] = g * h2[j];
        d2s[tid][j] = g * P[kCmW3 + j] * (1.0f - h2[j] * h2[j]);
    }
    __syncthreads();

    float* out = partial + size_t(blockIdx.x) * kCmParams;
    // w3, b3, b2: column sums over the rows
    if (tid < kCmHid) {
        float s3 = 0.0f, s2 = 0.0f;
        for (int q = 0; q < NR; ++q) {
            s3 += ws[q][tid];
            s2 += d2s[q][tid];
        }
        out[kCmW3 + tid] = s3;
        out[kCmB2 + tid] = s2;
    }
    if (tid == 0) {
        float sb = 0.0f;
        double sl = 0.0;
        for (int q = 0; q < NR; ++q) {
            sb += gs[q];
            sl += ls[q];
        }
        out[kCmB3] = sb;
        loss_partial[blockIdx.x] = sl;
    }
    // W2[j][i] = sum_r d2[r][j] * h1[r][i]: thread -> (j = tid / 2, 32 i's)
    {
        const int j = tid >> 1, i0 = (tid & 1) * 32;
        float acc[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) acc[k] = 0.0f;
        for (int q = 0; q < NR; ++q) {
            const float dj = d2s[q][j];
#pragma unroll
            for (int k = 0; k < 32; ++k) acc[k] = fmaf(dj, h1s[q][i0 + k], acc[k]);
        }
#pragma unroll
        for (int k = 0; k < 32; ++k) out[kCmW2 + j * kCmHid + i0 + k] = acc[k];
    }
    __syncthreads();   // ws (g*h2) fully read: reuse it for d1

    // d1[i] = (1 - h1[i]^2) * sum_j W2[j][i] d2[j] for this thread's row
    {
        float d2r[kCmHid];
#pragma unroll
        for (int j = 0; j < kCmHid; ++j) d2r[j] = d2s[tid][j];
        for (int i = 0; i < kCmHid; ++i) {
            float a = 0.0f;
#pragma unroll
            for (int j = 0; j < kCmHid; ++j) a = fmaf(P[kCmW2 + j * kCmHid + i], d2r[j], a);
            const float hi = h1s[tid][i];
            ws[tid][i] = a * (1.0f - hi * hi);
        }
    }
    __syncthreads();
    // W1[i][f] = sum_r d1[r][i] * x[r][f]: thread -> (i = tid / 2, half of
    // the features); b1
    {
        static_assert(kCmFeat % 2 == 0, "two threads per hidden unit split the features");
        constexpr int FH = kCmFeat / 2;
        const int i = tid >> 1, f0 = (tid & 1) * FH;
        float acc[FH];
#pragma unroll
        for (int k = 0; k < FH; ++k) acc[k] = 0.0f;
        float sb = 0.0f;
        for (int q = 0; q < NR; ++q) {
            const float di = ws[q][i];
            sb += di;
#pragma unroll
            for (int k = 0; k < FH; ++k) acc[k] = fmaf(di, xs[q][f0 + k], acc[k]);
        }
#pragma unroll
        for (int k = 0; k < FH; ++k) out[kCmW1 + i * kCmFeat + f0 + k] = acc[k];
        if ((tid & 1) == 0) out[kCmB1 + i] = sb;
    }
}

__global__ void __launch_bounds__(256) k_cm_reduce(const float* __restrict__ partial, int nblocks,
                                                   float* __restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kCmParams) return;
    float s = 0.0f;
    for (int b = 0; b < nblocks; ++b) s += partial[size_t(b) * kCmParams + p];
    grad[p] = s;
}

__global__ void __launch_bounds__(256) k_cm_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                                                 const float* __restrict__ grad, float lr, float b1, float b2,
                                                 float eps, float bc1, float bc2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kCmParams) return;
    const float gr = grad[p];
    const float mp = b1 * m[p] + (1.0f - b1) * gr;
    const float vp = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mp;
    v[p] = vp;
    P[p] -= lr * (mp / bc1) / (sqrtf(vp / bc2) + eps);
}

// ------------------------------------------------------------------ host

CostModel::CostModel()
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(ERR_HIP, "No HIP device for the cost model");
    CM_HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    CM_HIP_OK(hipMalloc(&d_params_, kCmParams * sizeof(float)));
    CM_HIP_OK(hipMalloc(&d_m_, kCmParams * sizeof(float)));
    CM_HIP_OK(hipMalloc(&d_v_, kCmParams * sizeof(float)));
    CM_HIP_OK(hipMalloc(&d_grad_, kCmParams * sizeof(float)));
    CM_HIP_OK(hipMalloc(&d_norm_, 2 * kCmFeat * sizeof(float)));
    std::vector<float> mu(kCmFeat, 0.0f), sg(kCmFeat, 1.0f);
    set_norm(mu.data(), sg.data());
    init(1);
}

CostModel::~CostModel()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : {(void*)d_params_, (void*)d_m_, (void*)d_v_, (void*)d_grad_, (void*)d_norm_, (void*)d_partial_,
                    (void*)d_x_, (void*)d_t_, (void*)d_y_, (void*)d_loss_})
        if (p) (void)hipFree(p);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void CostModel::init(uint64_t seed)
{
    ++version_;
    std::vector<float> p(kCmParams, 0.0f);
    uint64_t s = seed;
    auto uni = [&]() {   // SplitMix64 -> U[0,1)
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        return double(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int count, int fan_in, int fan_out) {
        const double a = std::sqrt(6.0 / double(fan_in + fan_out));
        for (int k = 0; k < count; ++k) p[off + k] = float((2.0 * uni() - 1.0) * a);
    };
    fill(kCmW1, kCmHid * kCmFeat, kCmFeat, kCmHid);
    fill(kCmW2, kCmHid * kCmHid, kCmHid, kCmHid);
    fill(kCmW3, kCmHid, kCmHid, 1);
    set_params(p.data());
    reset_optimizer();
    trained_ = false;
}

void CostModel::reset_optimizer()
{
    CM_HIP_OK(hipMemsetAsync(d_m_, 0, kCmParams * sizeof(float), stream_));
    CM_HIP_OK(hipMemsetAsync(d_v_, 0, kCmParams * sizeof(float), stream_));
    CM_HIP_OK(hipStreamSynchronize(stream_));
    step_ = 0;
}

void CostModel::set_params(const float* p)
{
    ++version_;
    CM_HIP_OK(hipMemcpyAsync(d_params_, p, kCmParams * sizeof(float), hipMemcpyHostToDevice, stream_));
    CM_HIP_OK(hipStreamSynchronize(stream_));
}

void CostModel::get_params(float* p) const
{
    CM_HIP_OK(hipMemcpyAsync(p, d_params_, kCmParams * sizeof(float), hipMemcpyDeviceToHost, stream_));
    CM_HIP_OK(hipStreamSynchronize(stream_));
}

void CostModel::set_norm(const float* mu, const float* sigma)
{
    ++version_;
    float h[2 * kCmFeat];
    for (int f = 0; f < kCmFeat; ++f) {
        if (!(sigma[f] > 0.0f)) throw Error(ERR_INVALID_ARGUMENT, "Feature scale must be positive");
        h[f] = mu[f];
        h[kCmFeat + f] = sigma[f];
    }
    CM_HIP_OK(hipMemcpyAsync(d_norm_, h, sizeof(h), hipMemcpyHostToDevice, stream_));
    CM_HIP_OK(hipStreamSynchronize(stream_));
}

void CostModel::get_norm(float* mu, float* sigma) const
{
    float h[2 * kCmFeat];
    CM_HIP_OK(hipMemcpyAsync(h, d_norm_, sizeof(h), hipMemcpyDeviceToHost, stream_));
    CM_HIP_OK(hipStreamSynchronize(stream_));
    std::memcpy(mu, h, kCmFeat * sizeof(float));
    std::memcpy(sigma, h + kCmFeat, kCmFeat * sizeof(float));
}

void CostModel::ensure_rows(size_t nrows)
{
    if (nrows <= rows_cap_) return;
    CM_HIP_OK(hipStreamSynchronize(stream_));
    for (void** p : {(void**)&d_x_, (void**)&d_t_, (void**)&d_y_})
        if (*p) {
            CM_HIP_OK(hipFree(*p));
            *p = nullptr;
        }
    const size_t cap = std::max<size_t>(nrows, 1024);
    CM_HIP_OK(hipMalloc(&d_x_, cap * kCmFeat * sizeof(float)));
    CM_HIP_OK(hipMalloc(&d_t_, cap * sizeof(float)));
    CM_HIP_OK(hipMalloc(&d_y_, cap * sizeof(float)));
    rows_cap_ = cap;
}

void CostModel::ensure_partials(size_t nblocks)
{
    if (nblocks <= partial_blocks_) return;
    CM_HIP_OK(hipStreamSynchronize(stream_));
    if (d_partial_) CM_HIP_OK(hipFree(d_partial_));
    if (d_loss_) CM_HIP_OK(hipFree(d_loss_));
    d_partial_ = nullptr;
    d_loss_ = nullptr;
    const size_t cap = std::max<size_t>(nblocks, 64);
    CM_HIP_OK(hipMalloc(&d_partial_, cap * kCmParams * sizeof(float)));
    CM_HIP_OK(hipMalloc(&d_loss_, cap * sizeof(double)));
    partial_blocks_ = cap;
}

void CostModel::predict(const float* x, size_t nrows, float* y)
{
    if (!nrows) return;
    if (nrows > (size_t(1) << 30)) throw Error(ERR_INVALID_ARGUMENT, "Too many rows");
    ensure_rows(nrows);
    CM_HIP_OK(hipMemcpyAsync(d_x_, x, nrows * kCmFeat * sizeof(float), hipMemcpyHostToDevice, stream_));
    const unsigned nb = unsigned((nrows + 255) / 256);
    hipLaunchKernelGGL(k_cm_forward, dim3(nb), dim3(256), 0, stream_, d_params_, d_norm_, d_x_, int(nrows), d_y_);
    CM_HIP_OK(hipGetLastError());
    CM_HIP_OK(hipMemcpyAsync(y, d_y_, nrows * sizeof(float), hipMemcpyDeviceToHost, stream_));
    CM_HIP_OK(hipStreamSynchronize(stream_));
}

// gradient of rows [0, nrows) of d_x_/d_t_ into d_grad_ (no host sync)
static void grad_launch(hipStream_t st, const float* P, const float* norm, const float* x, const float* t,
                        size_t nrows, float* partial, double* lossp, float* grad)
{
    const unsigned nb = unsigned((nrows + kCmRowsPerBlock - 1) / kCmRowsPerBlock);
    hipLaunchKernelGGL(k_cm_grad, dim3(nb), dim3(kCmRowsPerBlock), 0, st, P, norm, x, t, int(nrows),
                       float(1.0 / double(nrows)), partial, lossp);
    CM_HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(k_cm_reduce, dim3((kCmParams + 255) / 256), dim3(256), 0, st, partial, int(nb), grad);
    CM_HIP_OK(hipGetLastError());
}

double CostModel::grad_device(size_t nrows)
{
    const size_t nb = (nrows + kCmRowsPerBlock - 1) / kCmRowsPerBlock;
    ensure_partials(nb);
    grad_launch(stream_, d_params_, d_norm_, d_x_, d_t_, nrows, d_partial_, d_loss_, d_grad_);
    std::vector<double> lp(nb);
    CM_HIP_OK(hipMemcpyAsync(lp.data(), d_loss_, nb * sizeof(double), hipMemcpyDeviceToHost, stream_));
    CM_HIP_OK(hipStreamSynchronize(stream_));
    double s = 0.0;
    for (double v : lp) s += v;
    return s;
}

double CostModel::loss_and_grad(const float* x, const float* t, size_t nrows, float* grad)
{
    if (!nrows) throw Error(ERR_INVALID_ARGUMENT, "Empty batch");
    ensure_rows(nrows);
    CM_HIP_OK(hipMemcpyAsync(d_x_, x, nrows * kCmFeat * sizeof(float), hipMemcpyHostToDevice, stream_));
    CM_HIP_OK(hipMemcpyAsync(d_t_, t, nrows * sizeof(float), hipMemcpyHostToDevice, stream_));
    const double l = grad_device(nrows);
    CM_HIP_OK(hipMemcpyAsync(grad, d_grad_, kCmParams * sizeof(float), hipMemcpyDeviceToHost, stream_));
    CM_HIP_OK(hipStreamSynchronize(stream_));
    return l;
}

static void adam_launch(hipStream_t st, float* P, float* m, float* v, const float* g, const AdamConfig& c,
                        uint64_t step)
{
    const float bc1 = float(1.0 - std::pow(double(c.beta1), double(step)));
    const float bc2 = float(1.0 - std::pow(double(c.beta2), double(step)));
    hipLaunchKernelGGL(k_cm_adam, dim3((kCmParams + 255) / 256), dim3(256), 0, st, P, m, v, g, c.lr, c.beta1,
                       c.beta2, c.eps, bc1, bc2);
    CM_HIP_OK(hipGetLastError());
}

double CostModel::adam_step(const float* x, const float* t, size_t nrows, const AdamConfig& cfg)
{
    ++version_;
    if (!nrows) throw Error(ERR_INVALID_ARGUMENT, "Empty batch");
    ensure_rows(nrows);
    CM_HIP_OK(hipMemcpyAsync(d_x_, x, nrows * kCmFeat * sizeof(float), hipMemcpyHostToDevice, stream_));
    CM_HIP_OK(hipMemcpyAsync(d_t_, t, nrows * sizeof(float), hipMemcpyHostToDevice, stream_));
    const double l = grad_device(nrows);
    adam_launch(stream_, d_params_, d_m_, d_v_, d_grad_, cfg, ++step_);
    CM_HIP_OK(hipStreamSynchronize(stream_));
    return l;
}

double CostModel::train(const float* x, const float* t, size_t nrows, int epochs, size_t batch,
                        const AdamConfig& cfg, uint64_t seed)
{
    if (!nrows || !batch) throw Error(ERR_INVALID_ARGUMENT, "Empty training set or batch");
    ++version_;
    batch = std::min(batch, nrows);
    ensure_rows(nrows);
    ensure_partials((batch + kCmRowsPerBlock - 1) / kCmRowsPerBlock);
    std::vector<float> xs(nrows * kCmFeat), ts(nrows);
    std::vector<size_t> perm(nrows);
    for (size_t i = 0; i < nrows; ++i) perm[i] = i;
    Xoshiro256 rng(seed);
    for (int e = 0; e < epochs; ++e) {
        for (size_t i = nrows - 1; i > 0; --i) std::swap(perm[i], perm[size_t(rng.next() % (i + 1))]);
        CM_HIP_OK(hipStreamSynchronize(stream_));   // the previous epoch's kernels read xs/ts
        for (size_t i = 0; i < nrows; ++i) {
            std::memcpy(&xs[i * kCmFeat], x + perm[i] * kCmFeat, kCmFeat * sizeof(float));
            ts[i] = t[perm[i]];
        }
        CM_HIP_OK(hipMemcpyAsync(d_x_, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
        CM_HIP_OK(hipMemcpyAsync(d_t_, ts.data(), ts.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
        for (size_t off = 0; off + batch <= nrows; off += batch) {
            grad_launch(stream_, d_params_, d_norm_, d_x_ + off * kCmFeat, d_t_ + off, batch, d_partial_, d_loss_,
                        d_grad_);
            adam_launch(stream_, d_params_, d_m_, d_v_, d_grad_, cfg, ++step_);
        }
    }
    CM_HIP_OK(hipStreamSynchronize(stream_));
    // full-set loss with the final parameters
    CM_HIP_OK(hipMemcpyAsync(d_x_, x, nrows * kCmFeat * sizeof(float), hipMemcpyHostToDevice, stream_));
    CM_HIP_OK(hipMemcpyAsync(d_t_, t, nrows * sizeof(float), hipMemcpyHostToDevice, stream_));
    ensure_partials((nrows + kCmRowsPerBlock - 1) / kCmRowsPerBlock);
    const double l = grad_device(nrows);
    trained_ = true;
    return l;
}

}  // namespace q1t
