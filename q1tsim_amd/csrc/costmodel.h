// Learned sweep cost model: a small MLP (20 features -> 64 -> 64 -> 1, tanh,
// FP32) that predicts the log of a sweep's run time per 1024 tiles from
// features of its program.  Inference and training (forward, backward,
// gradient reduction, Adam) run in the HIP kernels of costmodel.hip.  The
// planner uses it to choose between candidate programs of the same gates
// (per-gate register phases, or dense blocks of 3 to 6 qubits); it never
// changes amplitudes, only which equivalent program runs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace q1t {

constexpr int kCmFeat = 20;
constexpr int kCmHid = 64;
// parameter layout (one float array): W1[kCmHid][kCmFeat], b1[kCmHid],
// W2[kCmHid][kCmHid], b2[kCmHid], w3[kCmHid], b3
constexpr int kCmW1 = 0;
constexpr int kCmB1 = kCmW1 + kCmHid * kCmFeat;
constexpr int kCmW2 = kCmB1 + kCmHid;
constexpr int kCmB2 = kCmW2 + kCmHid * kCmHid;
constexpr int kCmW3 = kCmB2 + kCmHid;
constexpr int kCmB3 = kCmW3 + kCmHid;
constexpr int kCmParams = kCmB3 + 1;
constexpr int kCmRowsPerBlock = 128;   // rows per workgroup of the gradient kernel

struct AdamConfig {
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
};

// Device-side model state and its kernels.
class CostModel {
public:
    CostModel();
    ~CostModel();
    CostModel(const CostModel&) = delete;
    CostModel& operator=(const CostModel&) = delete;

    // parameters (host <-> device); `init` draws them from a seeded
    // Glorot-uniform distribution (SplitMix64-driven, reproducible)
    void init(uint64_t seed);
    void set_params(const float* p);
    void get_params(float* p) const;
    // feature standardisation x' = (x - mu) / sigma, applied by the kernels
    void set_norm(const float* mu, const float* sigma);
    void get_norm(float* mu, float* sigma) const;
    bool trained() const { return trained_; }
    void mark_trained(bool t)
    {
        trained_ = t;
        ++version_;
    }
    // changes with every change of parameters, scaling or training state
    // (plans chosen with the model are kept only while it is unchanged)
    uint64_t version() const { return version_; }

    // y[i] = f(x[i]) for nrows rows of kCmFeat features (host arrays)
    void predict(const float* x, size_t nrows, float* y);
    // mean squared error of f(x) against t and its gradient (host arrays;
    // gradient layout as the parameters)
    double loss_and_grad(const float* x, const float* t, size_t nrows, float* grad);
    // one Adam step on the batch (x, t); returns the batch loss before the step
    double adam_step(const float* x, const float* t, size_t nrows, const AdamConfig& cfg);
    // minibatch Adam over the data set for `epochs` passes (rows shuffled per
    // epoch by a seeded permutation); returns the final full-set loss
    double train(const float* x, const float* t, size_t nrows, int epochs, size_t batch, const AdamConfig& cfg,
                 uint64_t seed);
    void reset_optimizer();

private:
    hipStream_t stream_ = nullptr;
    float* d_params_ = nullptr;
    float* d_m_ = nullptr;
    float* d_v_ = nullptr;
    float* d_norm_ = nullptr;      // mu[kCmFeat], sigma[kCmFeat]
    float* d_grad_ = nullptr;
    float* d_partial_ = nullptr;   // per-workgroup gradient partials
    size_t partial_blocks_ = 0;
    float* d_x_ = nullptr;
    float* d_t_ = nullptr;
    float* d_y_ = nullptr;
    double* d_loss_ = nullptr;
    size_t rows_cap_ = 0;
    uint64_t step_ = 0;
    uint64_t version_ = 0;
    bool trained_ = false;

    void ensure_rows(size_t nrows);
    void ensure_partials(size_t nblocks);
    double grad_device(size_t nrows);   // x, t already in d_x_, d_t_; gradient in d_grad_
};

}  // namespace q1t
