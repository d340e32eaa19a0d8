// Column work of many states in one launch per kind (ItemQueue).
#include "items.h"

#include <algorithm>
#include <cstring>
#include <mutex>

#include "gvs.h"
#include "kernels.h"

namespace q1t {

#define IQ_OK(expr)                                                                                \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw Error(e_ == hipErrorOutOfMemory ? ERR_OUT_OF_MEMORY : ERR_HIP,                    \
                        std::string("HIP error in " #expr ": ") + hipGetErrorString(e_));          \
    } while (0)

// work units of one column item per workgroup step
static constexpr uint64_t kChunk = uint64_t(kBlock) * 4;
// inputs of one tree-level workgroup step (kBlock outputs)
static constexpr uint64_t kLevelChunk = uint64_t(kBlock) * kTreeFanout;
static constexpr unsigned kMaxGrid = 8192;

// The item whose chunk range holds chunk c: the last with chunk0 <= c.
template <typename I>
__device__ __forceinline__ uint32_t find_item(const I* items, uint32_t n, uint64_t c)
{
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (items[mid].chunk0 <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint64_t insert_zero(uint64_t j, int p)
{
    const uint64_t bit = uint64_t(1) << p;
    return ((j >> p) << (p + 1)) | (j & (bit - 1));
}

// ------------------------------------------------------------ marginals

// Outcome-0 weight of bit p over work units j0, j0 + step, ...: |a0|^2 (Z),
// |a0 + a1|^2 / 2 (X) or |a0 - i a1|^2 / 2 (Y) per pair, in FP64 with the
// fused multiply-adds spelled out.
template <typename V>
__device__ __forceinline__ double marginal_acc(const V* a, uint64_t nwork, int p, int basis, uint64_t j0,
                                               uint64_t step)
{
    const uint64_t bit = uint64_t(1) << p;
    double acc = 0.0;
    for (uint64_t j = j0; j < nwork; j += step) {
        const uint64_t i0 = insert_zero(j, p);
        if (basis == 0) {
            acc += weight(a[i0]);
        } else {
            const V x0 = a[i0], x1 = a[i0 | bit];
            double re, im;
            if (basis == 1) {
                re = double(x0.x) + double(x1.x);
                im = double(x0.y) + double(x1.y);
            } else {   // a0 - i a1
                re = double(x0.x) + double(x1.y);
                im = double(x0.y) - double(x1.x);
            }
            acc = __fma_rn(0.5, __fma_rn(re, re, __dmul_rn(im, im)), acc);
        }
    }
    return acc;
}

__device__ __forceinline__ double block_total(double v)
{
    __shared__ double red[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kBlock / 64; ++k) s += red[k];
    __syncthreads();   // red is reused by the next call
    return s;
}

// kNormBlocks partial sums per item: partial[item * kNormBlocks + b].
template <typename V>
__global__ void __launch_bounds__(kBlock) k_marginal_items(const MargItem* __restrict__ items, uint32_t n,
                                                           double* __restrict__ partial)
{
    const uint64_t step = uint64_t(kNormBlocks) * kBlock;
    for (uint32_t it = blockIdx.y; it < n; it += gridDim.y) {
        const MargItem m = items[it];
        const double acc = marginal_acc<V>(static_cast<const V*>(m.amp), m.nwork, m.p, m.basis,
                                           uint64_t(blockIdx.x) * kBlock + threadIdx.x, step);
        const double s = block_total(acc);
        if (threadIdx.x == 0) partial[uint64_t(it) * kNormBlocks + blockIdx.x] = s;
    }
}

// w0 of each item: its partial sums added one after another from 0.0.
__global__ void __launch_bounds__(kBlock) k_sum_items(const double* __restrict__ partial, uint32_t n,
                                                     double* __restrict__ w0)
{
    const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    double s = 0.0;
    for (unsigned b = 0; b < kNormBlocks; ++b) s += partial[uint64_t(it) * kNormBlocks + b];
    w0[it] = s;
}

// The same partial sums for `ncols` columns at stride col_stride (no item
// table: the allocation-free marginal).
template <typename V>
__global__ void __launch_bounds__(kBlock) k_marginal_cols(const V* __restrict__ amp, uint64_t col_stride,
                                                          uint32_t ncols, uint64_t nwork, int p, int basis,
                                                          double* __restrict__ partial)
{
    const uint64_t step = uint64_t(kNormBlocks) * kBlock;
    for (uint32_t c = blockIdx.y; c < ncols; c += gridDim.y) {
        const double acc = marginal_acc<V>(amp + uint64_t(c) * col_stride, nwork, p, basis,
                                           uint64_t(blockIdx.x) * kBlock + threadIdx.x, step);
        const double s = block_total(acc);
        if (threadIdx.x == 0) partial[uint64_t(c) * kNormBlocks + blockIdx.x] = s;
    }
}

void launch_marginal_cols(bool c64, const void* amp, uint64_t col_stride, uint32_t ncols, uint64_t nwork, int p,
                          int basis, double* partial, hipStream_t s)
{
    const dim3 g(kNormBlocks, std::min<uint32_t>(ncols, 65535));
    if (c64)
        hipLaunchKernelGGL(k_marginal_cols<float2>, g, dim3(kBlock), 0, s, static_cast<const float2*>(amp),
                           col_stride, ncols, nwork, p, basis, partial);
    else
        hipLaunchKernelGGL(k_marginal_cols<double2>, g, dim3(kBlock), 0, s, static_cast<const double2*>(amp),
                           col_stride, ncols, nwork, p, basis, partial);
    IQ_OK(hipGetLastError());
}

// ------------------------------------------------------------- collapses

// Collapse of one pair (a0, a1) of bit p.  Z: dst0 keeps a0 s0, dst1 a1 s1.
// X/Y (the basis change back included): X, outcome 0: a0' = a1' =
// (a0 + a1) s0/2; outcome 1: a0' = -a1' = (a0 - a1) s1/2.  Y, outcome 0:
// u = a0 - i a1, a0' = u s0/2, a1' = i u s0/2; outcome 1: v = a0 + i a1,
// a0' = v s1/2, a1' = -i v s1/2.
template <typename V>
__device__ __forceinline__ void collapse_pair(const CollItem& m, uint64_t j)
{
    using Rl = typename RealOf<V>::type;
    const V* src = static_cast<const V*>(m.src);
    V* d0 = static_cast<V*>(m.dst0);
    V* d1 = static_cast<V*>(m.dst1);
    const uint64_t bit = uint64_t(1) << m.p;
    const uint64_t i0 = insert_zero(j, m.p), i1 = i0 | bit;
    const V x0 = src[i0], x1 = src[i1];
    const Rl s0 = Rl(m.s0), s1 = Rl(m.s1);
    if (m.basis == 0) {
        const V zero = mkc<V>(0, 0);
        if (d0) {
            d0[i0] = mkc<V>(x0.x * s0, x0.y * s0);
            d0[i1] = zero;
        }
        if (d1) {
            d1[i0] = zero;
            d1[i1] = mkc<V>(x1.x * s1, x1.y * s1);
        }
        return;
    }
    const Rl h0 = Rl(0.5) * s0, h1 = Rl(0.5) * s1;
    if (m.basis == 1) {
        const V u = mkc<V>((x0.x + x1.x) * h0, (x0.y + x1.y) * h0);
        const V v = mkc<V>((x0.x - x1.x) * h1, (x0.y - x1.y) * h1);
        if (d0) {
            d0[i0] = u;
            d0[i1] = u;
        }
        if (d1) {
            d1[i0] = v;
            d1[i1] = mkc<V>(-v.x, -v.y);
        }
    } else {
        const V u = mkc<V>((x0.x + x1.y) * h0, (x0.y - x1.x) * h0);   // (a0 - i a1) s0/2
        const V v = mkc<V>((x0.x - x1.y) * h1, (x0.y + x1.x) * h1);   // (a0 + i a1) s1/2
        if (d0) {
            d0[i0] = u;
            d0[i1] = mkc<V>(-u.y, u.x);   // i u
        }
        if (d1) {
            d1[i0] = v;
            d1[i1] = mkc<V>(v.y, -v.x);   // -i v
        }
    }
}

template <typename V>
__global__ void __launch_bounds__(kBlock) k_collapse_items(const CollItem* __restrict__ items, uint32_t n,
                                                           uint64_t nchunks)
{
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const CollItem m = items[find_item(items, n, c)];
        const uint64_t j0 = (c - m.chunk0) * kChunk;
        const uint64_t j1 = j0 + kChunk < m.nwork ? j0 + kChunk : m.nwork;
        for (uint64_t j = j0 + threadIdx.x; j < j1; j += kBlock) collapse_pair<V>(m, j);
    }
}

// ------------------------------------------------------------ column gates

template <typename V> __device__ __forceinline__ V entry(const ColOp& o, int k)
{
    using Rl = typename RealOf<V>::type;
    return mkc<V>(Rl(o.m[2 * k]), Rl(o.m[2 * k + 1]));
}

template <typename V>
__device__ __forceinline__ void col_op_apply(V* a, const ColOp& o, const PosList& pl, uint64_t j)
{
    const uint64_t i = deposit(j, pl) | o.cmask;
    switch (o.kind) {
    case 0: {   // dense, one target
        const uint64_t i1 = i | (uint64_t(1) << o.t[0]);
        const V x = a[i], y = a[i1];
        a[i] = cadd(cmul(entry<V>(o, 0), x), cmul(entry<V>(o, 1), y));
        a[i1] = cadd(cmul(entry<V>(o, 2), x), cmul(entry<V>(o, 3), y));
        break;
    }
    case 1: {   // dense, two targets (t[0] the local MSB)
        const uint64_t b0 = uint64_t(1) << o.t[0], b1 = uint64_t(1) << o.t[1];
        const uint64_t ix[4] = {i, i | b1, i | b0, i | b0 | b1};
        V v[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) v[l] = a[ix[l]];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            V acc = cmul(entry<V>(o, r * 4), v[0]);
#pragma unroll
            for (int c = 1; c < 4; ++c) acc = cmac(acc, entry<V>(o, r * 4 + c), v[c]);
            a[ix[r]] = acc;
        }
        break;
    }
    case 2: {   // diagonal on up to 3 targets
        int l = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (q < o.nt) l = (l << 1) | int((i >> o.t[q]) & 1);
        a[i] = cmul(entry<V>(o, l), a[i]);
        break;
    }
    case 3:   // phase where every bit of cmask is 1
        a[i] = cmul(entry<V>(o, 0), a[i]);
        break;
    default: {   // swap of the two targets
        const uint64_t ba = uint64_t(1) << o.t[0], bb = uint64_t(1) << o.t[1];
        const V x = a[i | bb], y = a[i | ba];
        a[i | bb] = y;
        a[i | ba] = x;
        break;
    }
    }
}

template <typename V>
__global__ void __launch_bounds__(kBlock) k_op_items(const OpItem* __restrict__ items, uint32_t n, uint64_t nchunks,
                                                     const ColOp* __restrict__ ops)
{
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const OpItem it = items[find_item(items, n, c)];
        const ColOp& o = ops[it.op];
        PosList pl;
        pl.cnt = o.npos;
#pragma unroll
        for (int k = 0; k < 16; ++k) pl.s[k] = o.pos[k];
        V* a = static_cast<V*>(it.amp);
        const uint64_t j0 = (c - it.chunk0) * kChunk;
        const uint64_t j1 = j0 + kChunk < it.nwork ? j0 + kChunk : it.nwork;
        for (uint64_t j = j0 + threadIdx.x; j < j1; j += kBlock) col_op_apply<V>(a, o, pl, j);
    }
}

// --------------------------------------------------------- basis columns

// Column of |k> (rot 0), H^n|k> (rot 1) or S^n H^n|k> (rot 2): amplitude j =
// (-1)^popcount(j&k) i^popcount(j) (the i factor for rot 2 only) times mag.
template <typename V>
__global__ void __launch_bounds__(kBlock) k_basis_items(const BasisItem* __restrict__ items, uint32_t n,
                                                        uint64_t nchunks)
{
    using Rl = typename RealOf<V>::type;
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const BasisItem b = items[find_item(items, n, c)];
        V* a = static_cast<V*>(b.amp);
        const Rl mag = Rl(b.mag);
        const uint64_t j0 = (c - b.chunk0) * kChunk;
        const uint64_t j1 = j0 + kChunk < b.dim ? j0 + kChunk : b.dim;
        for (uint64_t j = j0 + threadIdx.x; j < j1; j += kBlock) {
            V v;
            if (b.rot == 0) {
                v = mkc<V>(j == b.k ? Rl(1) : Rl(0), Rl(0));
            } else {
                const Rl s = (__popcll(j & b.k) & 1) ? -mag : mag;
                const int q = b.rot == 2 ? (__popcll(j) & 3) : 0;   // i^q
                v = q == 0 ? mkc<V>(s, 0) : q == 1 ? mkc<V>(0, s) : q == 2 ? mkc<V>(-s, 0) : mkc<V>(0, -s);
            }
            a[j] = v;
        }
    }
}

// ------------------------------------------------------------- sampling

template <typename T> __device__ __forceinline__ double leaf(const T* in, uint64_t k);
template <> __device__ __forceinline__ double leaf<double2>(const double2* in, uint64_t k) { return weight(in[k]); }
template <> __device__ __forceinline__ double leaf<float2>(const float2* in, uint64_t k) { return weight(in[k]); }
template <> __device__ __forceinline__ double leaf<double>(const double* in, uint64_t k) { return in[k]; }

// One tree level of many columns: out[g] = left fold of in[16g .. 16g+15].
// A workgroup step folds 4096 inputs; loads are coalesced into LDS in a
// transposed, padded layout so each thread reads its 16 values bank-conflict
// free.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_level_items(const LevelItem* __restrict__ items, uint32_t n,
                                                        uint64_t nchunks)
{
    __shared__ double w[kTreeFanout * (kBlock + 1)];
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const LevelItem it = items[find_item(items, n, c)];
        const T* in = static_cast<const T*>(it.in);
        const uint64_t size = it.size;
        const uint64_t nout = (size + kTreeFanout - 1) / kTreeFanout;
        const uint64_t base = (c - it.chunk0) * kLevelChunk;
#pragma unroll
        for (int r = 0; r < kTreeFanout; ++r) {
            const int e = r * kBlock + threadIdx.x;
            const uint64_t k = base + e;
            const double v = k < size ? leaf<T>(in, k) : 0.0;
            w[(e & (kTreeFanout - 1)) * (kBlock + 1) + (e >> 4)] = v;
        }
        __syncthreads();
        const uint64_t g = base / kTreeFanout + threadIdx.x;
        if (g < nout) {
            double acc = w[threadIdx.x];
            for (int jj = 1; jj < kTreeFanout; ++jj)
                if (g * kTreeFanout + jj < size) acc += w[jj * (kBlock + 1) + threadIdx.x];
            it.out[g] = acc;
        }
        __syncthreads();
    }
}

template <typename V>
__device__ __forceinline__ double tree_at(const V* a, const TreeDesc& td, int l, uint64_t k, uint64_t coff)
{
    return l == 0 ? weight(a[k]) : td.lev[l][coff + k];
}

// Inverse-CDF sampling of one run per thread by the canonical tree walk:
// t = u * T; within each group keep offset += S_j while offset + S_j <= t,
// descend into the first child where that fails, or the last positive one.
template <typename V>
__global__ void __launch_bounds__(kBlock) k_sample_items(const SampleJob* __restrict__ jobs, uint32_t njobs,
                                                         const uint64_t* __restrict__ run0,
                                                         const uint64_t* __restrict__ first,
                                                         const double* __restrict__ u01, uint64_t nruns,
                                                         uint64_t* __restrict__ out)
{
    const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (run0[mid] <= r) lo = mid;
        else hi = mid - 1;
    }
    const SampleJob& jb = jobs[lo];
    const uint64_t rl = r - jb.run0;
    const uint64_t* fr = first + jb.first_off;
    int clo = 0, chi = jb.ncols - 1;
    while (clo < chi) {
        const int mid = (clo + chi + 1) >> 1;
        if (fr[mid] <= rl) clo = mid;
        else chi = mid - 1;
    }
    const V* a = static_cast<const V*>(jb.a0) + uint64_t(clo) * jb.col_stride;
    const uint64_t coff = uint64_t(clo) * jb.per_col;
    const TreeDesc& td = jb.td;
    const int top = td.nlev - 1;
    double total = tree_at(a, td, top, 0, coff);
    for (uint64_t k = 1; k < td.size[top]; ++k) total += tree_at(a, td, top, k, coff);
    const double t = u01[r] * total;
    double offset = 0.0;
    uint64_t gstart = 0, glen = td.size[top], idx = 0;
    for (int l = top; l >= 0; --l) {
        double v[kTreeFanout];
        int last = 0;
#pragma unroll
        for (int j = 0; j < kTreeFanout; ++j) {
            v[j] = (uint64_t(j) < glen) ? tree_at(a, td, l, gstart + j, coff) : 0.0;
            if (v[j] > 0.0) last = j;
        }
        int chosen = last;
        for (int j = 0; j < last; ++j) {
            const double nx = offset + v[j];
            if (nx <= t) {
                offset = nx;
            } else {
                chosen = j;
                break;
            }
        }
        idx = gstart + uint64_t(chosen);
        if (l == 0) break;
        gstart = idx * kTreeFanout;
        const uint64_t rem = td.size[l - 1] - gstart;
        glen = rem < uint64_t(kTreeFanout) ? rem : uint64_t(kTreeFanout);
    }
    out[r] = idx;
}

// Total weight of each column (its top tree level summed in order), one
// thread per column of all jobs: the norm check of measure_all.
template <typename V>
__global__ void __launch_bounds__(kBlock) k_top_items(const SampleJob* __restrict__ jobs, uint32_t njobs,
                                                      uint64_t ncols, double* __restrict__ tops)
{
    const uint64_t cg = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (cg >= ncols) return;
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].col0 <= cg) lo = mid;
        else hi = mid - 1;
    }
    const SampleJob& jb = jobs[lo];
    const uint64_t c = cg - jb.col0;
    const V* a = static_cast<const V*>(jb.a0) + c * jb.col_stride;
    const int top = jb.td.nlev - 1;
    double total = 0.0;
    for (uint64_t k = 0; k < jb.td.size[top]; ++k) total += tree_at(a, jb.td, top, k, c * jb.per_col);
    tops[cg] = total;
}

// ------------------------------------------------------------- the queue

ItemQueue::ItemQueue(bool c64) : c64_(c64) {}

ItemQueue::~ItemQueue()
{
    if (last_stream_) (void)hipStreamSynchronize(last_stream_);
    if (dbuf_) (void)hipFree(dbuf_);
    if (hin_) (void)hipHostFree(hin_);
    if (hout_) (void)hipHostFree(hout_);
    if (in_done_) (void)hipEventDestroy(in_done_);
}

size_t ItemQueue::add_marginal(const MargItem& m)
{
    marg_.push_back(m);
    return marg_.size() - 1;
}

void ItemQueue::add_collapse(const CollItem& c)
{
    if (!c.nwork) return;
    CollItem x = c;
    x.chunk0 = coll_chunks_;
    coll_chunks_ += (c.nwork + kChunk - 1) / kChunk;
    coll_.push_back(x);
}

uint32_t ItemQueue::add_op(const ColOp& op)
{
    ops_.push_back(op);
    return uint32_t(ops_.size() - 1);
}

void ItemQueue::add_op_item(void* amp, uint64_t nwork, uint32_t op)
{
    if (!nwork) return;
    OpItem it{amp, nwork, op_chunks_, op, 0};
    op_chunks_ += (nwork + kChunk - 1) / kChunk;
    opit_.push_back(it);
}

void ItemQueue::add_basis(const BasisItem& b)
{
    if (!b.dim) return;
    BasisItem x = b;
    x.chunk0 = basis_chunks_;
    basis_chunks_ += (b.dim + kChunk - 1) / kChunk;
    basis_.push_back(x);
}

size_t ItemQueue::add_sample(const SampleJob& job, const std::vector<LevelItem>& levels, const uint64_t* first,
                             const double* u, size_t nruns)
{
    if (job.td.nlev < 1 || job.td.nlev > 16 || levels.size() != size_t(job.td.nlev - 1) * size_t(job.ncols) ||
        job.ncols < 1)
        throw Error(ERR_INVALID_ARGUMENT, "sampling job: inconsistent tree");
    for (size_t i = 0; i < levels.size(); ++i) {
        const int l = int(i / size_t(job.ncols)) + 1;
        LevelItem it = levels[i];
        it.chunk0 = level_chunks_[l];
        level_chunks_[l] += (it.size + kLevelChunk - 1) / kLevelChunk;
        levels_[l].push_back(it);
    }
    SampleJob j = job;
    j.first_off = first_.size();
    j.run0 = u_.size();
    j.col0 = cols_;
    cols_ += uint64_t(job.ncols);
    first_.insert(first_.end(), first, first + job.ncols + 1);
    u_.insert(u_.end(), u, u + nruns);
    jobs_.push_back(j);
    return jobs_.size() - 1;
}

void ItemQueue::release_after(void* p, size_t bytes, hipStream_t s)
{
    if (p) rel_.push_back(Release{p, bytes, s});
}

bool ItemQueue::empty() const
{
    bool lv = false;
    for (const auto& v : levels_) lv = lv || !v.empty();
    return marg_.empty() && coll_.empty() && opit_.empty() && basis_.empty() && jobs_.empty() && !lv &&
           rel_.empty();
}

static size_t up256(size_t x) { return (x + 255) / 256 * 256; }

void ItemQueue::launch(hipStream_t s)
{
    // layout: inputs, then outputs (partials, indices, tops)
    size_t in = 0;
    auto take = [&](size_t bytes) {
        const size_t o = in;
        in = up256(in + bytes);
        return o;
    };
    const size_t o_marg = take(marg_.size() * sizeof(MargItem));
    const size_t o_coll = take(coll_.size() * sizeof(CollItem));
    const size_t o_ops = take(ops_.size() * sizeof(ColOp));
    const size_t o_opit = take(opit_.size() * sizeof(OpItem));
    const size_t o_basis = take(basis_.size() * sizeof(BasisItem));
    size_t o_lev[16] = {};
    for (int l = 1; l < 16; ++l) o_lev[l] = take(levels_[l].size() * sizeof(LevelItem));
    const size_t o_jobs = take(jobs_.size() * sizeof(SampleJob));
    run0_.clear();
    for (const SampleJob& j : jobs_) run0_.push_back(j.run0);
    run0_.push_back(u_.size());
    const size_t o_run0 = take(run0_.size() * sizeof(uint64_t));
    const size_t o_first = take(first_.size() * sizeof(uint64_t));
    const size_t o_u = take(u_.size() * sizeof(double));
    const size_t o_part = take(marg_.size() * size_t(kNormBlocks) * sizeof(double));   // device only
    const size_t in_bytes = o_part, dev_bytes = in;
    size_t out = 0;
    const size_t p_w0 = out;
    out = up256(out + marg_.size() * sizeof(double));
    const size_t p_idx = out;
    out = up256(out + u_.size() * sizeof(uint64_t));
    const size_t p_top = out;
    out = up256(out + cols_ * sizeof(double));
    const size_t out_bytes = out;

    // buffers: wait for the previous upload before the host inputs are
    // rewritten; grow the device buffer after the stream has drained
    if (in_done_) IQ_OK(hipEventSynchronize(in_done_));
    else IQ_OK(hipEventCreateWithFlags(&in_done_, hipEventDisableTiming));
    if (dev_bytes + out_bytes > dcap_) {
        if (last_stream_) IQ_OK(hipStreamSynchronize(last_stream_));
        IQ_OK(hipStreamSynchronize(s));
        if (dbuf_) IQ_OK(hipFree(dbuf_));
        dbuf_ = nullptr;
        dcap_ = 0;
        const size_t cap = std::max<size_t>(dev_bytes + out_bytes, size_t(1) << 20) * 3 / 2;
        IQ_OK(hipMalloc(reinterpret_cast<void**>(&dbuf_), cap));
        dcap_ = cap;
    }
    if (in_bytes > hin_cap_) {
        if (hin_) IQ_OK(hipHostFree(hin_));
        hin_ = nullptr;
        hin_cap_ = 0;
        const size_t cap = std::max<size_t>(in_bytes, size_t(1) << 20) * 3 / 2;
        IQ_OK(hipHostMalloc(reinterpret_cast<void**>(&hin_), cap, hipHostMallocDefault));
        hin_cap_ = cap;
    }
    if (out_bytes > hout_cap_) {
        if (last_stream_) IQ_OK(hipStreamSynchronize(last_stream_));
        if (hout_) IQ_OK(hipHostFree(hout_));
        hout_ = nullptr;
        hout_cap_ = 0;
        const size_t cap = std::max<size_t>(out_bytes, size_t(1) << 20) * 3 / 2;
        IQ_OK(hipHostMalloc(reinterpret_cast<void**>(&hout_), cap, hipHostMallocDefault));
        hout_cap_ = cap;
    }
    last_stream_ = s;
    auto put = [&](size_t off, const void* src, size_t bytes) {
        if (bytes) std::memcpy(hin_ + off, src, bytes);
    };
    put(o_marg, marg_.data(), marg_.size() * sizeof(MargItem));
    put(o_coll, coll_.data(), coll_.size() * sizeof(CollItem));
    put(o_ops, ops_.data(), ops_.size() * sizeof(ColOp));
    put(o_opit, opit_.data(), opit_.size() * sizeof(OpItem));
    put(o_basis, basis_.data(), basis_.size() * sizeof(BasisItem));
    for (int l = 1; l < 16; ++l) put(o_lev[l], levels_[l].data(), levels_[l].size() * sizeof(LevelItem));
    put(o_jobs, jobs_.data(), jobs_.size() * sizeof(SampleJob));
    put(o_run0, run0_.data(), run0_.size() * sizeof(uint64_t));
    put(o_first, first_.data(), first_.size() * sizeof(uint64_t));
    put(o_u, u_.data(), u_.size() * sizeof(double));
    if (in_bytes) {
        IQ_OK(hipMemcpyAsync(dbuf_, hin_, in_bytes, hipMemcpyHostToDevice, s));
        launches_ += 1;
    }
    IQ_OK(hipEventRecord(in_done_, s));
    char* d = dbuf_;
    char* dout = dbuf_ + dev_bytes;
    auto grid = [](uint64_t chunks) { return dim3(unsigned(std::min<uint64_t>(chunks, kMaxGrid))); };
    if (!basis_.empty()) {
        if (c64_) hipLaunchKernelGGL(k_basis_items<float2>, grid(basis_chunks_), dim3(kBlock), 0, s,
                                     reinterpret_cast<const BasisItem*>(d + o_basis), uint32_t(basis_.size()), basis_chunks_);
        else hipLaunchKernelGGL(k_basis_items<double2>, grid(basis_chunks_), dim3(kBlock), 0, s,
                                reinterpret_cast<const BasisItem*>(d + o_basis), uint32_t(basis_.size()), basis_chunks_);
        IQ_OK(hipGetLastError());
        launches_ += 1;
    }
    if (!coll_.empty()) {
        if (c64_) hipLaunchKernelGGL(k_collapse_items<float2>, grid(coll_chunks_), dim3(kBlock), 0, s,
                                     reinterpret_cast<const CollItem*>(d + o_coll), uint32_t(coll_.size()), coll_chunks_);
        else hipLaunchKernelGGL(k_collapse_items<double2>, grid(coll_chunks_), dim3(kBlock), 0, s,
                                reinterpret_cast<const CollItem*>(d + o_coll), uint32_t(coll_.size()), coll_chunks_);
        IQ_OK(hipGetLastError());
        launches_ += 1;
    }
    if (!opit_.empty()) {
        const ColOp* dops = reinterpret_cast<const ColOp*>(d + o_ops);
        if (c64_) hipLaunchKernelGGL(k_op_items<float2>, grid(op_chunks_), dim3(kBlock), 0, s,
                                     reinterpret_cast<const OpItem*>(d + o_opit), uint32_t(opit_.size()), op_chunks_, dops);
        else hipLaunchKernelGGL(k_op_items<double2>, grid(op_chunks_), dim3(kBlock), 0, s,
                                reinterpret_cast<const OpItem*>(d + o_opit), uint32_t(opit_.size()), op_chunks_, dops);
        IQ_OK(hipGetLastError());
        launches_ += 1;
    }
    if (!marg_.empty()) {
        const dim3 g(kNormBlocks, unsigned(std::min<size_t>(marg_.size(), 65535)));
        double* part = reinterpret_cast<double*>(d + o_part);
        if (c64_) hipLaunchKernelGGL(k_marginal_items<float2>, g, dim3(kBlock), 0, s,
                                     reinterpret_cast<const MargItem*>(d + o_marg), uint32_t(marg_.size()), part);
        else hipLaunchKernelGGL(k_marginal_items<double2>, g, dim3(kBlock), 0, s,
                                reinterpret_cast<const MargItem*>(d + o_marg), uint32_t(marg_.size()), part);
        IQ_OK(hipGetLastError());
        hipLaunchKernelGGL(k_sum_items, dim3(unsigned((marg_.size() + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                           part, uint32_t(marg_.size()), reinterpret_cast<double*>(dout + p_w0));
        IQ_OK(hipGetLastError());
        launches_ += 2;
    }
    for (int l = 1; l < 16; ++l) {
        if (levels_[l].empty()) continue;
        const LevelItem* li = reinterpret_cast<const LevelItem*>(d + o_lev[l]);
        const uint32_t nl = uint32_t(levels_[l].size());
        if (l > 1) hipLaunchKernelGGL(k_level_items<double>, grid(level_chunks_[l]), dim3(kBlock), 0, s, li, nl,
                                      level_chunks_[l]);
        else if (c64_) hipLaunchKernelGGL(k_level_items<float2>, grid(level_chunks_[l]), dim3(kBlock), 0, s, li, nl,
                                          level_chunks_[l]);
        else hipLaunchKernelGGL(k_level_items<double2>, grid(level_chunks_[l]), dim3(kBlock), 0, s, li, nl,
                                level_chunks_[l]);
        IQ_OK(hipGetLastError());
        launches_ += 1;
    }
    if (!jobs_.empty()) {
        const SampleJob* dj = reinterpret_cast<const SampleJob*>(d + o_jobs);
        const uint32_t nj = uint32_t(jobs_.size());
        const uint64_t nruns = u_.size();
        uint64_t* idx = reinterpret_cast<uint64_t*>(dout + p_idx);
        double* tops = reinterpret_cast<double*>(dout + p_top);
        if (nruns) {
            const dim3 g(unsigned((nruns + kBlock - 1) / kBlock));
            if (c64_) hipLaunchKernelGGL(k_sample_items<float2>, g, dim3(kBlock), 0, s, dj, nj,
                                         reinterpret_cast<const uint64_t*>(d + o_run0),
                                         reinterpret_cast<const uint64_t*>(d + o_first),
                                         reinterpret_cast<const double*>(d + o_u), nruns, idx);
            else hipLaunchKernelGGL(k_sample_items<double2>, g, dim3(kBlock), 0, s, dj, nj,
                                    reinterpret_cast<const uint64_t*>(d + o_run0),
                                    reinterpret_cast<const uint64_t*>(d + o_first),
                                    reinterpret_cast<const double*>(d + o_u), nruns, idx);
            IQ_OK(hipGetLastError());
            launches_ += 1;
        }
        const dim3 gt(unsigned((cols_ + kBlock - 1) / kBlock));
        if (c64_) hipLaunchKernelGGL(k_top_items<float2>, gt, dim3(kBlock), 0, s, dj, nj, cols_, tops);
        else hipLaunchKernelGGL(k_top_items<double2>, gt, dim3(kBlock), 0, s, dj, nj, cols_, tops);
        IQ_OK(hipGetLastError());
        launches_ += 1;
    }
    if (out_bytes) {
        IQ_OK(hipMemcpyAsync(hout_, dout, out_bytes, hipMemcpyDeviceToHost, s));
        launches_ += 1;
    }
    out_w0_ = p_w0;
    out_idx_ = p_idx;
    out_top_ = p_top;
    res_run0_.clear();
    res_col0_.clear();
    for (const SampleJob& j : jobs_) {
        res_run0_.push_back(j.run0);
        res_col0_.push_back(j.col0);
    }
    // the kernels that read the released blocks are enqueued: stream order
    // keeps later users of those blocks behind them
    for (const Release& r : rel_) DeviceCache::get().release(r.p, r.bytes, r.s);
    rel_.clear();
    marg_.clear();
    coll_.clear();
    ops_.clear();
    opit_.clear();
    basis_.clear();
    for (auto& v : levels_) v.clear();
    jobs_.clear();
    first_.clear();
    u_.clear();
    cols_ = 0;
    coll_chunks_ = op_chunks_ = basis_chunks_ = 0;
    for (uint64_t& c : level_chunks_) c = 0;
}

double ItemQueue::w0(size_t item) const { return reinterpret_cast<const double*>(hout_ + out_w0_)[item]; }

const uint64_t* ItemQueue::indices(size_t job) const
{
    return reinterpret_cast<const uint64_t*>(hout_ + out_idx_) + res_run0_[job];
}

const double* ItemQueue::tops(size_t job) const
{
    return reinterpret_cast<const double*>(hout_ + out_top_) + res_col0_[job];
}

void ItemQueue::detach()
{
    if (last_stream_) (void)hipStreamSynchronize(last_stream_);
    last_stream_ = nullptr;
    // the upload event belongs to the last owner's stream, which may be
    // destroyed and its handle reused: the next owner records a fresh one
    if (in_done_) (void)hipEventDestroy(in_done_);
    in_done_ = nullptr;
    for (const Release& r : rel_) DeviceCache::get().release(r.p, r.bytes, r.s);
    rel_.clear();
    marg_.clear();
    coll_.clear();
    ops_.clear();
    opit_.clear();
    basis_.clear();
    for (auto& v : levels_) v.clear();
    jobs_.clear();
    first_.clear();
    u_.clear();
    cols_ = 0;
    coll_chunks_ = op_chunks_ = basis_chunks_ = 0;
    for (uint64_t& c : level_chunks_) c = 0;
}

namespace {
struct QueuePool {
    std::mutex mu;
    std::vector<std::unique_ptr<ItemQueue>> free[2];
};
QueuePool& queue_pool()
{
    static QueuePool* p = new QueuePool();   // never destroyed (the HIP runtime may be gone at exit)
    return *p;
}
}  // namespace

std::unique_ptr<ItemQueue> ItemQueuePool::take(bool c64)
{
    QueuePool& p = queue_pool();
    {
        std::lock_guard<std::mutex> lk(p.mu);
        auto& v = p.free[c64 ? 1 : 0];
        if (!v.empty()) {
            std::unique_ptr<ItemQueue> q = std::move(v.back());
            v.pop_back();
            return q;
        }
    }
    return std::make_unique<ItemQueue>(c64);
}

void ItemQueuePool::give(std::unique_ptr<ItemQueue> q)
{
    if (!q) return;
    q->detach();
    QueuePool& p = queue_pool();
    std::lock_guard<std::mutex> lk(p.mu);
    auto& v = p.free[q->c64() ? 1 : 0];
    if (v.size() < 64) v.push_back(std::move(q));
}

bool col_op_from(const GateOp& g, ColOp& o)
{
    std::memset(&o, 0, sizeof(o));
    std::vector<int> pos;
    for (int b = 0; b < 64; ++b)
        if ((g.ctrl_mask >> b) & 1) pos.push_back(b);
    o.cmask = g.ctrl_mask;
    const int nt = int(g.targets.size());
    auto put = [&](size_t k, const cdouble& c) {
        o.m[2 * k] = c.real();
        o.m[2 * k + 1] = c.imag();
    };
    switch (g.kind) {
    case OpKind::Dense:
        if (nt < 1 || nt > 2) return false;
        o.kind = nt == 1 ? 0 : 1;
        for (int q = 0; q < nt; ++q) {
            o.t[q] = g.targets[q];
            pos.push_back(g.targets[q]);
        }
        for (size_t k = 0; k < g.m.size(); ++k) put(k, g.m[k]);
        break;
    case OpKind::Diag:
        if (nt < 1 || nt > 3) return false;
        o.kind = 2;
        for (int q = 0; q < nt; ++q) o.t[q] = g.targets[q];
        for (size_t k = 0; k < g.m.size(); ++k) put(k, g.m[k]);
        break;
    case OpKind::PhaseOnes:
        o.kind = 3;
        put(0, g.m[0]);
        break;
    case OpKind::Swap2:
        o.kind = 4;
        o.t[0] = g.targets[0];
        o.t[1] = g.targets[1];
        pos.push_back(g.targets[0]);
        pos.push_back(g.targets[1]);
        break;
    default:
        return false;
    }
    if (pos.size() > 16) return false;
    std::sort(pos.begin(), pos.end());
    o.nt = nt;
    o.npos = int(pos.size());
    for (size_t k = 0; k < pos.size(); ++k) o.pos[k] = pos[k];
    return true;
}

}  // namespace q1t
