// GpuVectorState: state storage, per-gate sweep kernels and measurement
// kernels for MI355X (gfx950).
//
// Layout in HBM: one contiguous 2^n-amplitude vector per branch column
// (column-major), interleaved complex (re, im) in FP64 (c128) or FP32 (c64).
// The reference stores the same matrix row-major (cmatrix.rs:20-24); a column
// per contiguous vector makes every gate sweep a unit-stride pass over one
// column and lets columns be added, dropped and copied without a transpose.
#include "gvs.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "kernels.h"
#include "sweep.h"

namespace q1t {

#define HIP_OK(expr)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw Error(e_ == hipErrorOutOfMemory ? ERR_OUT_OF_MEMORY : ERR_HIP,                    \
                        std::string("HIP error in " #expr ": ") + hipGetErrorString(e_));          \
    } while (0)

// ------------------------------------------------------------------ kernels

// One-qubit dense update on every (i0, i0|tbit) pair whose control bits are 1.
template <typename V>
__global__ void __launch_bounds__(kBlock) k_dense1(V* __restrict__ amp, uint64_t col_stride, uint64_t nwork,
                                                   PosList pos, uint64_t cmask, uint64_t tbit, V m0, V m1, V m2, V m3)
{
    V* a = amp + uint64_t(blockIdx.y) * col_stride;
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < nwork; j += step) {
        const uint64_t i0 = deposit(j, pos) | cmask, i1 = i0 | tbit;
        const V x = a[i0], y = a[i1];
        a[i0] = cadd(cmul(m0, x), cmul(m1, y));
        a[i1] = cadd(cmul(m2, x), cmul(m3, y));
    }
}

template <typename V> struct Mat4 {
    V m[16];
};

// Two-qubit dense update; t0 is the local MSB.
template <typename V>
__global__ void __launch_bounds__(kBlock) k_dense2(V* __restrict__ amp, uint64_t col_stride, uint64_t nwork,
                                                   PosList pos, uint64_t cmask, uint64_t b0, uint64_t b1, Mat4<V> m)
{
    V* a = amp + uint64_t(blockIdx.y) * col_stride;
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < nwork; j += step) {
        const uint64_t i = deposit(j, pos) | cmask;
        const uint64_t ix[4] = {i, i | b1, i | b0, i | b0 | b1};
        V v[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) v[l] = a[ix[l]];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            V acc = cmul(m.m[r * 4], v[0]);
#pragma unroll
            for (int c = 1; c < 4; ++c) acc = cmac(acc, m.m[r * 4 + c], v[c]);
            a[ix[r]] = acc;
        }
    }
}

// K-qubit dense update (K = 3..5), matrix staged in LDS.
template <typename V, int K>
__global__ void __launch_bounds__(kBlock) k_denseK(V* __restrict__ amp, uint64_t col_stride, uint64_t nwork,
                                                   PosList pos, uint64_t cmask, Targets tg, const V* __restrict__ mat)
{
    constexpr int D = 1 << K;
    __shared__ V sm[D * D];
    for (int t = threadIdx.x; t < D * D; t += blockDim.x) sm[t] = mat[t];
    __syncthreads();
    uint64_t off[D];
#pragma unroll
    for (int l = 0; l < D; ++l) {
        uint64_t o = 0;
#pragma unroll
        for (int q = 0; q < K; ++q)
            if ((l >> (K - 1 - q)) & 1) o |= uint64_t(1) << tg.t[q];
        off[l] = o;
    }
    V* a = amp + uint64_t(blockIdx.y) * col_stride;
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < nwork; j += step) {
        const uint64_t i = deposit(j, pos) | cmask;
        V v[D];
#pragma unroll
        for (int l = 0; l < D; ++l) v[l] = a[i | off[l]];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            V acc = cmul(sm[r * D], v[0]);
#pragma unroll
            for (int c = 1; c < D; ++c) acc = cmac(acc, sm[r * D + c], v[c]);
            a[i | off[r]] = acc;
        }
    }
}

// Dense update for 6..12 target qubits: one workgroup per amplitude group,
// the group staged in LDS, matrix rows read from global memory.
template <typename V>
__global__ void __launch_bounds__(kBlock) k_dense_big(V* __restrict__ amp, uint64_t col_stride, uint64_t ngroups,
                                                      PosList pos, uint64_t cmask, Targets tg,
                                                      const V* __restrict__ mat)
{
    extern __shared__ unsigned char smem_raw[];
    V* sg = reinterpret_cast<V*>(smem_raw);
    const int K = tg.cnt, D = 1 << K;
    V* a = amp + uint64_t(blockIdx.y) * col_stride;
    for (uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const uint64_t base = deposit(g, pos) | cmask;
        for (int l = threadIdx.x; l < D; l += blockDim.x) {
            uint64_t o = 0;
            for (int q = 0; q < K; ++q)
                if ((l >> (K - 1 - q)) & 1) o |= uint64_t(1) << tg.t[q];
            sg[l] = a[base | o];
        }
        __syncthreads();
        for (int r = threadIdx.x; r < D; r += blockDim.x) {
            V acc = mkc<V>(0, 0);
            for (int c = 0; c < D; ++c) acc = cmac(acc, mat[uint64_t(r) * D + c], sg[c]);
            uint64_t o = 0;
            for (int q = 0; q < K; ++q)
                if ((r >> (K - 1 - q)) & 1) o |= uint64_t(1) << tg.t[q];
            a[base | o] = acc;
        }
        __syncthreads();
    }
}

template <typename V> struct Diag8 {
    V d[8];
};

// Diagonal gate on up to 3 targets: a[i] *= d[local(i)].
template <typename V>
__global__ void __launch_bounds__(kBlock) k_diag(V* __restrict__ amp, uint64_t col_stride, uint64_t nwork,
                                                 PosList pos, uint64_t cmask, Targets tg, Diag8<V> dv)
{
    __shared__ V sd[8];
    if (threadIdx.x < 8) sd[threadIdx.x] = dv.d[threadIdx.x];
    __syncthreads();
    V* a = amp + uint64_t(blockIdx.y) * col_stride;
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < nwork; j += step) {
        const uint64_t i = deposit(j, pos) | cmask;
        int l = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (q < tg.cnt) l = (l << 1) | int((i >> tg.t[q]) & 1);
        a[i] = cmul(sd[l], a[i]);
    }
}

// Phase on the subspace where every bit of `mask` is 1 (Z, S, T, U1, CZ, CU1 ...).
template <typename V>
__global__ void __launch_bounds__(kBlock) k_phase_ones(V* __restrict__ amp, uint64_t col_stride, uint64_t nwork,
                                                       PosList pos, uint64_t mask, V ph)
{
    V* a = amp + uint64_t(blockIdx.y) * col_stride;
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < nwork; j += step) {
        const uint64_t i = deposit(j, pos) | mask;
        a[i] = cmul(ph, a[i]);
    }
}

// Swap: exchange |..0a..1b..> and |..1a..0b..> (controls set).
template <typename V>
__global__ void __launch_bounds__(kBlock) k_swap2(V* __restrict__ amp, uint64_t col_stride, uint64_t nwork,
                                                  PosList pos, uint64_t cmask, uint64_t ba, uint64_t bb)
{
    V* a = amp + uint64_t(blockIdx.y) * col_stride;
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < nwork; j += step) {
        const uint64_t i = deposit(j, pos) | cmask;
        const V x = a[i | bb], y = a[i | ba];
        a[i | bb] = y;
        a[i | ba] = x;
    }
}

__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double red[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < int(blockDim.x) / 64; ++k) s += red[k];
    return s;
}

// Partial sums of |a|^2 over the subspace given by pos (FP64).
template <typename V>
__global__ void __launch_bounds__(kBlock) k_norm_partial(const V* __restrict__ amp, uint64_t col_stride,
                                                         uint64_t nwork, PosList pos, double* __restrict__ partial)
{
    const V* a = amp + uint64_t(blockIdx.y) * col_stride;
    double acc = 0.0;
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < nwork; j += step) {
        const V v = a[deposit(j, pos)];
        acc += double(v.x) * double(v.x) + double(v.y) * double(v.y);
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[uint64_t(blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

// Measurement in the X or Y basis without the basis-change passes: the
// reference measures H.psi (X) or H.Sdg.psi (Y) on the qubit and rotates back
// (circuit.rs:669-748).  On a pair (a0, a1) of the qubit, the rotated
// amplitude of outcome 0 is b0 = (a0 + a1)/sqrt2 (X) or (a0 - i a1)/sqrt2 (Y),
// of outcome 1 b1 = (a0 - a1)/sqrt2 or (a0 + i a1)/sqrt2.
enum MeasBasis : int { BASIS_Z = 0, BASIS_X = 1, BASIS_Y = 2 };

// Product state: amplitude = b_{n-1}[x_{n-1}] * (... * (b_0[x_0] * 1)), the
// association order of successive kron_mat calls (vectorstate.rs:68-74).
template <typename V>
__global__ void __launch_bounds__(kBlock) k_product(V* __restrict__ amp, uint64_t N, int n, const double2* __restrict__ b)
{
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += step) {
        double re = 1.0, im = 0.0;
        for (int q = 0; q < n; ++q) {
            const double2 c = b[2 * q + ((i >> (n - 1 - q)) & 1)];
            const double nr = c.x * re - c.y * im, ni = c.x * im + c.y * re;
            re = nr;
            im = ni;
        }
        amp[i] = mkc<V>(typename RealOf<V>::type(re), typename RealOf<V>::type(im));
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// re, im ~ N(0,1): SplitMix64 counter hash of (seed, i) and Box-Muller.
template <typename V>
__global__ void __launch_bounds__(kBlock) k_random_normal(V* __restrict__ amp, uint64_t N, uint64_t seed)
{
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += step) {
        const uint64_t ha = mix64(seed + 0x9e3779b97f4a7c15ULL * (2 * i + 1));
        const uint64_t hb = mix64(seed + 0x9e3779b97f4a7c15ULL * (2 * i + 2));
        const double u1 = double((ha >> 11) + 1) * (1.0 / 9007199254740992.0);
        const double u2 = double(hb >> 11) * (1.0 / 9007199254740992.0);
        const double r = sqrt(-2.0 * log(u1));
        const double th = 2.0 * M_PI * u2;
        amp[i] = mkc<V>(typename RealOf<V>::type(r * cos(th)), typename RealOf<V>::type(r * sin(th)));
    }
}

template <typename V>
__global__ void __launch_bounds__(kBlock) k_scale(V* __restrict__ amp, uint64_t N, double f)
{
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += step) {
        const V v = amp[i];
        amp[i] = mkc<V>(typename RealOf<V>::type(v.x * f), typename RealOf<V>::type(v.y * f));
    }
}

// ------------------------------------------------------------- host helpers

static unsigned grid_for(uint64_t nwork)
{
    uint64_t g = (nwork + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    if (g == 0) g = 1;
    return unsigned(g);
}

static PosList make_pos(std::vector<int> v)
{
    std::sort(v.begin(), v.end());
    PosList p{};
    p.cnt = int(v.size());
    for (size_t k = 0; k < v.size(); ++k) p.s[k] = v[k];
    return p;
}

static std::vector<int> mask_bits(uint64_t m)
{
    std::vector<int> r;
    for (int b = 0; b < 64; ++b)
        if ((m >> b) & 1) r.push_back(b);
    return r;
}

template <typename V> static V to_v(const cdouble& c)
{
    return mkc<V>(typename RealOf<V>::type(c.real()), typename RealOf<V>::type(c.imag()));
}

// ------------------------------------------------------------ the state

// Never destroyed: at process exit the HIP runtime may already be gone, and
// the driver reclaims the cached blocks.
DeviceCache& DeviceCache::get()
{
    static DeviceCache* c = new DeviceCache();
    return *c;
}

size_t DeviceCache::size_class(size_t bytes)
{
    if (bytes <= 4096) return 4096;
    if (bytes <= (size_t(1) << 30)) {
        size_t c = 4096;
        while (c < bytes) c <<= 1;
        return c;
    }
    const size_t g = size_t(256) << 20;
    return (bytes + g - 1) / g * g;
}

static void drop_entry_event(hipEvent_t ev, bool wait)
{
    if (!ev) return;
    if (wait) HIP_OK(hipEventSynchronize(ev));
    HIP_OK(hipEventDestroy(ev));
}

void DeviceCache::trim()
{
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : free_)
        for (Entry& e : kv.second) {
            drop_entry_event(e.ev, true);
            HIP_OK(hipFree(e.p));
        }
    free_.clear();
    cached_ = 0;
}

size_t DeviceCache::default_cap()
{
    size_t f = 0, t = 0;
    HIP_OK(hipMemGetInfo(&f, &t));
    return t / 8 * 7;
}

void DeviceCache::evict_to(size_t limit)
{
    while (cached_ > limit) {
        // the oldest block over all classes
        auto oldest = free_.end();
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (!it->second.empty() && (oldest == free_.end() || it->second.front().seq < oldest->second.front().seq))
                oldest = it;
        if (oldest == free_.end()) break;
        Entry e = oldest->second.front();
        oldest->second.pop_front();
        if (oldest->second.empty()) free_.erase(oldest);
        drop_entry_event(e.ev, true);
        HIP_OK(hipFree(e.p));
        cached_ -= e.bytes;
    }
}

size_t DeviceCache::evict_for(size_t need, size_t cls)
{
    std::lock_guard<std::mutex> lk(mu_);
    size_t freed = 0;
    // the oldest blocks of other size classes first, then of `cls` itself
    for (int pass = 0; pass < 2 && freed < need; ++pass) {
        while (freed < need) {
            auto oldest = free_.end();
            for (auto it = free_.begin(); it != free_.end(); ++it) {
                if (it->second.empty() || (pass == 0 && it->first == cls)) continue;
                if (oldest == free_.end() || it->second.front().seq < oldest->second.front().seq) oldest = it;
            }
            if (oldest == free_.end()) break;
            Entry e = oldest->second.front();
            oldest->second.pop_front();
            if (oldest->second.empty()) free_.erase(oldest);
            drop_entry_event(e.ev, true);
            HIP_OK(hipFree(e.p));
            cached_ -= e.bytes;
            freed += e.bytes;
        }
    }
    return freed;
}

void DeviceCache::set_cap(size_t bytes)
{
    std::lock_guard<std::mutex> lk(mu_);
    cap_ = bytes ? bytes : default_cap();
    evict_to(cap_);
}

size_t DeviceCache::cached_bytes()
{
    std::lock_guard<std::mutex> lk(mu_);
    return cached_;
}

// Called with `s` synchronised and before it is destroyed: its blocks are
// idle, and their events must not outlive the stream they were recorded on.
void DeviceCache::forget_stream(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : free_)
        for (Entry& e : kv.second) {
            if (e.s != s) continue;
            if (e.ev) (void)hipEventDestroy(e.ev);
            e.ev = nullptr;
            e.s = nullptr;
        }
}

void* DeviceCache::alloc(size_t bytes, hipStream_t s)
{
    const size_t cls = size_class(bytes);
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.find(cls);
        if (it != free_.end()) {
            std::deque<Entry>& v = it->second;
            for (size_t i = 0; i < v.size(); ++i) {
                Entry& e = v[i];
                if (e.ev && e.s != s && hipEventQuery(e.ev) != hipSuccess) {
                    (void)hipGetLastError();   // a pending event is not an error: keep it out of the next check
                    continue;
                }
                void* p = e.p;
                if (e.ev) HIP_OK(hipEventDestroy(e.ev));
                cached_ -= e.bytes;
                v.erase(v.begin() + std::ptrdiff_t(i));   // keeps release order (oldest first)
                if (v.empty()) free_.erase(it);
                return p;
            }
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    while (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        // free just enough cached blocks and retry: emptying the whole cache
        // would make every later request of this pass miss and allocate anew
        if (evict_for(cls, cls) == 0) break;
        e = hipMalloc(&p, cls);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(e == hipErrorOutOfMemory ? ERR_OUT_OF_MEMORY : ERR_HIP,
                    "Device allocation of " + std::to_string(cls) + " bytes failed: " + hipGetErrorString(e));
    }
    return p;
}

void DeviceCache::release(void* p, size_t bytes, hipStream_t s)
{
    if (!p) return;
    const size_t cls = size_class(bytes);
    std::lock_guard<std::mutex> lk(mu_);
    if (!cap_) cap_ = default_cap();
    if (cls > cap_) {
        HIP_OK(hipStreamSynchronize(s));
        HIP_OK(hipFree(p));
        return;
    }
    // over the cap: free the oldest cached blocks, so the cache follows the
    // sizes in use now (a batch after a large state, c128 after c64)
    evict_to(cap_ - cls);
    hipEvent_t ev;
    HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIP_OK(hipEventRecord(ev, s));
    free_[cls].push_back(Entry{p, cls, s, ev, seq_++});
    cached_ += cls;
}

SharedStream::SharedStream() { HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }

SharedStream::~SharedStream()
{
    if (s) {
        (void)hipStreamSynchronize(s);
        DeviceCache::get().forget_stream(s);
        (void)hipStreamDestroy(s);
    }
}

GpuVectorState::GpuVectorState(size_t nr_bits, size_t nr_shots, Precision prec, std::shared_ptr<SharedStream> shared,
                               bool compact_zero)
    : nr_bits_(nr_bits), nr_shots_(nr_shots), prec_(prec)
{
    if (nr_bits > 40) throw Error(ERR_INVALID_ARGUMENT, "Too many qubits for a dense state");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw Error(ERR_HIP, "No HIP device available for the GPU statevector");
    HIP_OK(hipGetDevice(&device_));
    if (shared) {
        shared_ = std::move(shared);
        stream_ = shared_->s;
    } else {
        HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    }
    geom_ = tile_geometry(prec);
    if (compact_zero) set_zero_state_compact();
    else set_zero_state();
}

void GpuVectorState::set_zero_state_compact()
{
    queue_.clear();
    basis_ = true;
    basis_rot_ = 0;
    basis_idx_.assign(1, 0);
    counts_.assign(1, nr_shots_);
}

void GpuVectorState::prepare_batched_flush(ItemQueue& q)
{
    materialize_enqueue(q);
    stats_ = RunStats{};
}

void GpuVectorState::finish_batched_flush(const RunStats& s)
{
    stats_ = s;
    queue_.clear();
}

void GpuVectorState::set_geometry(const TileGeom& g)
{
    if (!tile_shape_supported(prec_, g.M, g.R) || g.L < 1 || g.L > g.M)
        throw Error(ERR_INVALID_ARGUMENT, "Unsupported tile geometry M=" + std::to_string(g.M) + " R=" +
                                              std::to_string(g.R) + " L=" + std::to_string(g.L));
    flush();
    geom_ = g;
    geom_set_ = true;
}

void GpuVectorState::set_jit(int mode)
{
    if (mode < 0 || mode > 3) throw Error(ERR_INVALID_ARGUMENT, "jit mode must be 0..3");
    jit_mode_ = mode;
    if (!geom_set_) geom_ = mode ? tile_geometry_specialised(prec_) : tile_geometry(prec_);
}

GpuVectorState::~GpuVectorState()
{
    // an own stream is destroyed below, so its work must end first; a shared
    // stream outlives the state, and the released blocks are event-guarded
    if (stream_ && !shared_) (void)hipStreamSynchronize(stream_);
    try {
        ItemQueuePool::give(std::move(iq_));   // synchronises the queue's last stream
    } catch (...) {
    }
    try {
        DeviceCache& dc = DeviceCache::get();
        for (FlushGraph& fg : fgraphs_) {
            if (fg.exec) (void)hipGraphExecDestroy(fg.exec);
            dc.release(fg.table, fg.cap, stream_);
        }
        fgraphs_.clear();
        dc.release(amp_, amp_cols_ * dim() * elem_bytes(), stream_);
        dc.release(d_partial_, partial_cap_ * sizeof(double), stream_);
        dc.release(d_tree_, tree_cap_ * sizeof(double), stream_);
        dc.release(d_small_, small_cap_, stream_);
        for (int s = 0; s < SLOT_COUNT; ++s) dc.release(d_slot_[s], slot_cap_[s], stream_);
    } catch (...) {
    }
    for (auto& ev : ev_pool_) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    if (stream_ && !shared_) {
        DeviceCache::get().forget_stream(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

void GpuVectorState::check_qbit(size_t q) const
{
    if (q >= nr_bits_) throw err_invalid_qbit(q);
}

void* GpuVectorState::col_ptr(size_t c) const
{
    return static_cast<char*>(amp_) + c * dim() * elem_bytes();
}

void* GpuVectorState::dev_alloc(size_t bytes) { return DeviceCache::get().alloc(bytes, stream_); }

void GpuVectorState::dev_release(void* p, size_t bytes) { DeviceCache::get().release(p, bytes, stream_); }

void GpuVectorState::alloc_cols(size_t ncols)
{
    if (ncols <= amp_cols_ && amp_) return;
    if (amp_) {
        dev_release(amp_, amp_cols_ * dim() * elem_bytes());
        amp_ = nullptr;
        amp_cols_ = 0;
    }
    amp_ = dev_alloc(ncols * dim() * elem_bytes());
    amp_cols_ = ncols;
}

void* GpuVectorState::small_buf(size_t bytes)
{
    if (bytes > small_cap_) {
        dev_release(d_small_, small_cap_);
        d_small_ = nullptr;
        small_cap_ = 0;
        size_t cap = std::max<size_t>(bytes, 1 << 20);
        d_small_ = dev_alloc(cap);
        small_cap_ = cap;
    }
    return d_small_;
}

double* GpuVectorState::partial_buf(size_t n)
{
    if (n > partial_cap_) {
        dev_release(d_partial_, partial_cap_ * sizeof(double));
        d_partial_ = nullptr;
        partial_cap_ = 0;
        const size_t cap = std::max<size_t>(n, 4096);
        d_partial_ = static_cast<double*>(dev_alloc(cap * sizeof(double)));
        partial_cap_ = cap;
    }
    return d_partial_;
}

void GpuVectorState::synchronize() { HIP_OK(hipStreamSynchronize(stream_)); }

void GpuVectorState::set_zero_state()
{
    queue_.clear();
    basis_ = false;
    counts_.assign(1, nr_shots_);
    alloc_cols(1);
    HIP_OK(hipMemsetAsync(amp_, 0, dim() * elem_bytes(), stream_));
    // static sources: the copies may run after this returns (same stream
    // order for everything that follows, no wait needed)
    static const double one64[2] = {1.0, 0.0};
    static const float one32[2] = {1.0f, 0.0f};
    if (prec_ == PREC_C128) HIP_OK(hipMemcpyAsync(amp_, one64, sizeof(one64), hipMemcpyHostToDevice, stream_));
    else HIP_OK(hipMemcpyAsync(amp_, one32, sizeof(one32), hipMemcpyHostToDevice, stream_));
}

void GpuVectorState::set_product_state(const cdouble* coefs)
{
    queue_.clear();
    basis_ = false;
    counts_.assign(1, nr_shots_);
    alloc_cols(1);
    // normalise each qubit as vectorstate.rs:71-72 does
    std::vector<double> b(4 * nr_bits_);
    for (size_t q = 0; q < nr_bits_; ++q) {
        cdouble c0 = coefs[2 * q], c1 = coefs[2 * q + 1];
        double norm = std::sqrt((c0.real() * c0.real() + c0.imag() * c0.imag()) +
                                (c1.real() * c1.real() + c1.imag() * c1.imag()));
        b[4 * q + 0] = c0.real() / norm;
        b[4 * q + 1] = c0.imag() / norm;
        b[4 * q + 2] = c1.real() / norm;
        b[4 * q + 3] = c1.imag() / norm;
    }
    void* d = small_buf(b.size() * sizeof(double));
    HIP_OK(hipMemcpyAsync(d, b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    if (prec_ == PREC_C128)
        hipLaunchKernelGGL(k_product<double2>, dim3(grid_for(dim())), dim3(kBlock), 0, stream_,
                           static_cast<double2*>(amp_), dim(), int(nr_bits_), static_cast<const double2*>(d));
    else
        hipLaunchKernelGGL(k_product<float2>, dim3(grid_for(dim())), dim3(kBlock), 0, stream_,
                           static_cast<float2*>(amp_), dim(), int(nr_bits_), static_cast<const double2*>(d));
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(stream_));
}

void GpuVectorState::set_amplitudes(const cdouble* amps)
{
    queue_.clear();
    basis_ = false;
    counts_.assign(1, nr_shots_);
    alloc_cols(1);
    if (prec_ == PREC_C128) {
        HIP_OK(hipMemcpyAsync(amp_, amps, dim() * sizeof(cdouble), hipMemcpyHostToDevice, stream_));
    } else {
        std::vector<float> f(2 * dim());
        for (uint64_t i = 0; i < dim(); ++i) {
            f[2 * i] = float(amps[i].real());
            f[2 * i + 1] = float(amps[i].imag());
        }
        HIP_OK(hipMemcpyAsync(amp_, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
    }
    // stream-ordered (earlier kernels may still use the buffer); the host
    // data must outlive the copy
    HIP_OK(hipStreamSynchronize(stream_));
}

void GpuVectorState::set_random_normal(uint64_t seed)
{
    queue_.clear();
    basis_ = false;
    counts_.assign(1, nr_shots_);
    alloc_cols(1);
    const unsigned g = grid_for(dim());
    PosList none{};
    double* part = partial_buf(512);
    if (prec_ == PREC_C128) {
        hipLaunchKernelGGL(k_random_normal<double2>, dim3(g), dim3(kBlock), 0, stream_, static_cast<double2*>(amp_), dim(), seed);
        hipLaunchKernelGGL(k_norm_partial<double2>, dim3(512), dim3(kBlock), 0, stream_, static_cast<const double2*>(amp_),
                           dim(), dim(), none, part);
    } else {
        hipLaunchKernelGGL(k_random_normal<float2>, dim3(g), dim3(kBlock), 0, stream_, static_cast<float2*>(amp_), dim(), seed);
        hipLaunchKernelGGL(k_norm_partial<float2>, dim3(512), dim3(kBlock), 0, stream_, static_cast<const float2*>(amp_),
                           dim(), dim(), none, part);
    }
    HIP_OK(hipGetLastError());
    std::vector<double> h(512);
    HIP_OK(hipMemcpyAsync(h.data(), part, 512 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    double ss = 0.0;
    for (double v : h) ss += v;
    const double f = 1.0 / std::sqrt(ss);
    if (prec_ == PREC_C128)
        hipLaunchKernelGGL(k_scale<double2>, dim3(g), dim3(kBlock), 0, stream_, static_cast<double2*>(amp_), dim(), f);
    else
        hipLaunchKernelGGL(k_scale<float2>, dim3(g), dim3(kBlock), 0, stream_, static_cast<float2*>(amp_), dim(), f);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(stream_));
}

ItemQueue& GpuVectorState::iq()
{
    if (!iq_) iq_ = ItemQueuePool::take(prec_ == PREC_C64);
    return *iq_;
}

void GpuVectorState::materialize()
{
    if (!basis_) return;
    ItemQueue& q = iq();
    materialize_enqueue(q);
    q.launch(stream_);
}

void GpuVectorState::materialize_enqueue(ItemQueue& q)
{
    if (!basis_) return;
    const size_t nc = counts_.size();
    const size_t bytes_needed = nc * dim() * elem_bytes();
    const size_t have = amp_ ? amp_cols_ * dim() * elem_bytes() : 0;
    if (bytes_needed > have && bytes_needed >= (size_t(1) << 28)) {
        // a large expansion: a clear error before the allocation is tried
        // (blocks idle in the device cache are freed by its allocation
        // retry, so they count as free)
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b));
        if (bytes_needed > free_b + have + DeviceCache::get().cached_bytes())
            throw Error(ERR_OUT_OF_MEMORY, "Not enough device memory to expand " + std::to_string(nc) +
                                               " measured basis columns into dense columns");
    }
    alloc_cols(nc);
    // every amplitude written by the kernel (no memset node if a caller
    // captures the launch)
    const double mag = (nr_bits_ & 1) ? std::ldexp(0.70710678118654752440, -int(nr_bits_ / 2))
                                      : std::ldexp(1.0, -int(nr_bits_ / 2));
    for (size_t c = 0; c < nc; ++c) q.add_basis(BasisItem{col_ptr(c), dim(), basis_idx_[c], 0, mag, basis_rot_, 0});
    basis_ = false;
    basis_rot_ = 0;
    basis_idx_.clear();
}

// --------------------------------------------------------------- gates

void GpuVectorState::apply_gate(const cdouble* mat, size_t nr_affected, const size_t* qbits, size_t nbits,
                                const std::string& desc)
{
    if (nr_affected != nbits) throw err_invalid_nr_bits(nbits, nr_affected, desc);
    for (size_t k = 0; k < nbits; ++k) check_qbit(qbits[k]);
    if (nbits > 12) throw Error(ERR_NOT_SUPPORTED, "Gates on more than 12 qubits are not supported");
    PendingGate pg;
    pg.op = analyze_gate(mat, nbits, qbits, nr_bits_);
    const size_t d = size_t(1) << nbits;
    pg.matrix.assign(mat, mat + d * d);
    for (size_t k = 0; k < nbits; ++k) pg.qbits.push_back(int(nr_bits_ - 1 - qbits[k]));
    queue_.push_back(std::move(pg));
}

void GpuVectorState::apply_unary_gate_all(const cdouble* mat2, const std::string& desc)
{
    for (size_t q = 0; q < nr_bits_; ++q) apply_gate(mat2, 1, &q, 1, desc);
}

void GpuVectorState::flush()
{
    if (queue_.empty()) return;
    materialize();
    auto t0 = std::chrono::steady_clock::now();
    stats_ = RunStats{};
    if (fusion_k_ > 0 || tiling_)
        run_planned(*this, queue_, 0, counts_.size(), fusion_k_, tiling_, stats_);
    else
        apply_queue_plain(0, counts_.size());
    HIP_OK(hipStreamSynchronize(stream_));
    stats_.gate_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    queue_.clear();
    prof_collect();
}

void GpuVectorState::apply_queue_plain(size_t cb, size_t ce)
{
    for (const PendingGate& pg : queue_) apply_op(pg.op, cb, ce);
}

template <typename V>
static void launch_op(const GateOp& op, V* amp, uint64_t col_stride, size_t ncols, int n, hipStream_t st,
                      GpuVectorState& gs, RunStats& stats)
{
    if (op.kind == OpKind::Identity) {
        stats.gates += 1;
        return;
    }
    const std::vector<int> ctrl = mask_bits(op.ctrl_mask);
    const double S = double(uint64_t(1) << n) * sizeof(V);
    const double frac = 1.0 / double(uint64_t(1) << ctrl.size());   // touched fraction
    for (size_t c0 = 0; c0 < ncols; c0 += 65535) {
        const unsigned ny = unsigned(std::min<size_t>(65535, ncols - c0));
        V* a = amp + c0 * col_stride;
        const char* kname = "";
        double bytes = 2.0 * S * frac * ny;
        int ps = -1;
        switch (op.kind) {
        case OpKind::PhaseOnes: {
            const uint64_t nwork = uint64_t(1) << (n - int(ctrl.size()));
            kname = "k_phase_ones";
            ps = gs.prof_begin();
            hipLaunchKernelGGL(k_phase_ones<V>, dim3(grid_for(nwork), ny), dim3(kBlock), 0, st, a, col_stride, nwork,
                               make_pos(ctrl), op.ctrl_mask, to_v<V>(op.m[0]));
            break;
        }
        case OpKind::Diag: {
            const uint64_t nwork = uint64_t(1) << (n - int(ctrl.size()));
            Targets tg{};
            tg.cnt = int(op.targets.size());
            for (int q = 0; q < tg.cnt; ++q) tg.t[q] = op.targets[q];
            Diag8<V> dv{};
            for (size_t l = 0; l < op.m.size(); ++l) dv.d[l] = to_v<V>(op.m[l]);
            kname = "k_diag";
            ps = gs.prof_begin();
            hipLaunchKernelGGL(k_diag<V>, dim3(grid_for(nwork), ny), dim3(kBlock), 0, st, a, col_stride, nwork,
                               make_pos(ctrl), op.ctrl_mask, tg, dv);
            break;
        }
        case OpKind::Swap2: {
            std::vector<int> pos = ctrl;
            pos.push_back(op.targets[0]);
            pos.push_back(op.targets[1]);
            const uint64_t nwork = uint64_t(1) << (n - int(pos.size()));
            kname = "k_swap2";
            bytes *= 0.5;
            ps = gs.prof_begin();
            hipLaunchKernelGGL(k_swap2<V>, dim3(grid_for(nwork), ny), dim3(kBlock), 0, st, a, col_stride, nwork,
                               make_pos(pos), op.ctrl_mask, uint64_t(1) << op.targets[0], uint64_t(1) << op.targets[1]);
            break;
        }
        case OpKind::Dense: {
            const int kt = int(op.targets.size());
            std::vector<int> pos = ctrl;
            for (int t : op.targets) pos.push_back(t);
            const uint64_t nwork = uint64_t(1) << (n - int(pos.size()));
            const PosList pl = make_pos(pos);
            if (kt == 1) {
                kname = "k_dense1";
                ps = gs.prof_begin();
                hipLaunchKernelGGL(k_dense1<V>, dim3(grid_for(nwork), ny), dim3(kBlock), 0, st, a, col_stride, nwork,
                                   pl, op.ctrl_mask, uint64_t(1) << op.targets[0], to_v<V>(op.m[0]),
                                   to_v<V>(op.m[1]), to_v<V>(op.m[2]), to_v<V>(op.m[3]));
            } else if (kt == 2) {
                Mat4<V> m4{};
                for (int k = 0; k < 16; ++k) m4.m[k] = to_v<V>(op.m[k]);
                kname = "k_dense2";
                ps = gs.prof_begin();
                hipLaunchKernelGGL(k_dense2<V>, dim3(grid_for(nwork), ny), dim3(kBlock), 0, st, a, col_stride, nwork,
                                   pl, op.ctrl_mask, uint64_t(1) << op.targets[0], uint64_t(1) << op.targets[1], m4);
            } else {
                Targets tg{};
                tg.cnt = kt;
                for (int q = 0; q < kt; ++q) tg.t[q] = op.targets[q];
                const size_t d = size_t(1) << kt;
                std::vector<V> hm(d * d);
                for (size_t k = 0; k < d * d; ++k) hm[k] = to_v<V>(op.m[k]);
                V* dm = static_cast<V*>(gs.scratch_matrix(hm.size() * sizeof(V)));
                if (hipMemcpyAsync(dm, hm.data(), hm.size() * sizeof(V), hipMemcpyHostToDevice, st) != hipSuccess ||
                    hipStreamSynchronize(st) != hipSuccess)
                    throw Error(ERR_HIP, "matrix upload failed");
                kname = kt <= 5 ? "k_denseK" : "k_dense_big";
                ps = gs.prof_begin();
                if (kt == 3)
                    hipLaunchKernelGGL((k_denseK<V, 3>), dim3(grid_for(nwork), ny), dim3(kBlock), 0, st, a, col_stride,
                                       nwork, pl, op.ctrl_mask, tg, dm);
                else if (kt == 4)
                    hipLaunchKernelGGL((k_denseK<V, 4>), dim3(grid_for(nwork), ny), dim3(kBlock), 0, st, a, col_stride,
                                       nwork, pl, op.ctrl_mask, tg, dm);
                else if (kt == 5)
                    hipLaunchKernelGGL((k_denseK<V, 5>), dim3(grid_for(nwork), ny), dim3(kBlock), 0, st, a, col_stride,
                                       nwork, pl, op.ctrl_mask, tg, dm);
                else {
                    // a group of 2^kt amplitudes in dynamic LDS: 64 KiB for 12 complex128 targets
                    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_dense_big<V>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int(d * sizeof(V))) != hipSuccess)
                        throw Error(ERR_HIP, "LDS attribute of the wide dense kernel");
                    hipLaunchKernelGGL(k_dense_big<V>, dim3(grid_for(nwork * kBlock), ny), dim3(kBlock), d * sizeof(V),
                                       st, a, col_stride, nwork, pl, op.ctrl_mask, tg, dm);
                }
            }
            break;
        }
        default:
            break;
        }
        if (hipGetLastError() != hipSuccess) throw Error(ERR_HIP, "gate kernel launch failed");
        gs.prof_end(ps, kname, bytes);
        stats.bytes += bytes;
        stats.launches += 1;
    }
    stats.gates += 1;
    stats.sweeps += 1;
}

void* GpuVectorState::scratch(int slot, size_t bytes)
{
    if (bytes > slot_cap_[slot]) {
        dev_release(d_slot_[slot], slot_cap_[slot]);
        d_slot_[slot] = nullptr;
        slot_cap_[slot] = 0;
        size_t cap = size_t(1) << 16;
        while (cap < bytes) cap <<= 1;
        d_slot_[slot] = dev_alloc(cap);
        slot_cap_[slot] = cap;
    }
    return d_slot_[slot];
}

void GpuVectorState::apply_op(const GateOp& op, size_t cb, size_t ce)
{
    if (ce <= cb) return;
    if (prec_ == PREC_C128)
        launch_op<double2>(op, static_cast<double2*>(col_ptr(cb)), dim(), ce - cb, int(nr_bits_), stream_, *this, stats_);
    else
        launch_op<float2>(op, static_cast<float2*>(col_ptr(cb)), dim(), ce - cb, int(nr_bits_), stream_, *this, stats_);
}

// collect_conditional_ranges (qustate.rs:100-127): runs of equal control
// within each column
std::vector<GpuVectorState::CondRange> GpuVectorState::conditional_ranges(const uint8_t* control) const
{
    std::vector<CondRange> ranges;
    size_t off = 0;
    for (size_t icol = 0; icol < counts_.size(); ++icol) {
        const size_t count = counts_[icol];
        if (count == 0) continue;
        size_t begin = off;
        bool prev = control[off] != 0;
        for (size_t ib = off + 1; ib < off + count; ++ib) {
            const bool cb = control[ib] != 0;
            if (cb != prev) {
                ranges.push_back({icol, ib - begin, prev});
                begin = ib;
                prev = !prev;
            }
        }
        if (begin < off + count) ranges.push_back({icol, off + count - begin, prev});
        off += count;
    }
    return ranges;
}

void GpuVectorState::check_conditional(size_t ncontrol, size_t nr_affected, const size_t* qbits, size_t nbits,
                                       const std::string& desc) const
{
    if (ncontrol != nr_shots_) throw err_invalid_nr_control_bits(ncontrol, nr_shots_, desc);
    if (nr_affected != nbits) throw err_invalid_nr_bits(nbits, nr_affected, desc);
    for (size_t k = 0; k < nbits; ++k) check_qbit(qbits[k]);
}

bool GpuVectorState::conditional_enqueue(ItemQueue& q, const uint8_t* control, size_t ncontrol, const cdouble* mat,
                                         size_t nr_affected, const size_t* qbits, size_t nbits,
                                         const std::string& desc)
{
    check_conditional(ncontrol, nr_affected, qbits, nbits, desc);
    flush();
    const std::vector<CondRange> ranges = conditional_ranges(control);
    bool any = false;
    for (const CondRange& r : ranges) any = any || r.apply;
    bool same_layout = ranges.size() == counts_.size();
    for (size_t r = 0; same_layout && r < ranges.size(); ++r) same_layout = ranges[r].col == r;
    // a column set that changes (runs of one column disagree) is copied
    // column by column: left to apply_conditional_gate
    if (!same_layout) return false;
    if (basis_ && !any) return true;
    GateOp op = analyze_gate(mat, nbits, qbits, nr_bits_);
    ColOp co;
    if (op.kind != OpKind::Identity && !col_op_from(op, co)) return false;
    materialize_enqueue(q);
    if (op.kind == OpKind::Identity || !any) return true;
    const uint32_t oi = q.add_op(co);
    const uint64_t nwork = dim() >> co.npos;
    for (size_t r = 0; r < ranges.size(); ++r)
        if (ranges[r].apply) q.add_op_item(col_ptr(r), nwork, oi);
    return true;
}

void GpuVectorState::apply_conditional_gate(const uint8_t* control, size_t ncontrol, const cdouble* mat,
                                            size_t nr_affected, const size_t* qbits, size_t nbits,
                                            const std::string& desc)
{
    check_conditional(ncontrol, nr_affected, qbits, nbits, desc);
    flush();
    {
        ItemQueue& q = iq();
        if (conditional_enqueue(q, control, ncontrol, mat, nr_affected, qbits, nbits, desc)) {
            q.launch(stream_);
            return;
        }
    }
    const std::vector<CondRange> ranges = conditional_ranges(control);
    bool any = false;
    for (const CondRange& r : ranges) any = any || r.apply;
    bool same_layout = ranges.size() == counts_.size();
    for (size_t r = 0; same_layout && r < ranges.size(); ++r) same_layout = ranges[r].col == r;
    if (basis_ && same_layout && !any) return;
    materialize();
    GateOp op = analyze_gate(mat, nbits, qbits, nr_bits_);
    if (!same_layout) {
        // new column set: one column per range, copied from its source column
        const size_t nr = ranges.size();
        void* old = amp_;
        const size_t old_cols = amp_cols_;
        void* fresh = dev_alloc(nr * dim() * elem_bytes());
        for (size_t r = 0; r < nr; ++r)
            HIP_OK(hipMemcpyAsync(static_cast<char*>(fresh) + r * dim() * elem_bytes(), col_ptr(ranges[r].col),
                                  dim() * elem_bytes(), hipMemcpyDeviceToDevice, stream_));
        dev_release(old, old_cols * dim() * elem_bytes());
        amp_ = fresh;
        amp_cols_ = nr;
    }
    counts_.clear();
    for (const CondRange& r : ranges) counts_.push_back(r.len);
    // gates of the column kernels; anything wider by the per-gate kernels
    ColOp co;
    if (op.kind != OpKind::Identity && col_op_from(op, co)) {
        ItemQueue& q = iq();
        const uint32_t oi = q.add_op(co);
        const uint64_t nwork = dim() >> co.npos;
        for (size_t r = 0; r < ranges.size(); ++r)
            if (ranges[r].apply) q.add_op_item(col_ptr(r), nwork, oi);
        q.launch(stream_);
        return;
    }
    for (size_t r = 0; r < ranges.size();) {
        if (!ranges[r].apply) {
            ++r;
            continue;
        }
        size_t e = r;
        while (e < ranges.size() && ranges[e].apply) ++e;
        apply_op(op, r, e);
        r = e;
    }
}

// ---------------------------------------------------------- measurement

void GpuVectorState::check_basis(int basis)
{
    if (basis != BASIS_Z && basis != BASIS_X && basis != BASIS_Y)
        throw Error(ERR_INVALID_ARGUMENT, "Unknown measurement basis " + std::to_string(basis));
}

// Measured in the basis the compact columns were collapsed in, a basis
// vector gives its own bit; in another basis either outcome with
// probability 1/2.
std::vector<double> GpuVectorState::basis_w0(size_t qbit, int basis) const
{
    const int p = int(nr_bits_ - 1 - qbit);
    std::vector<double> w0(counts_.size());
    for (size_t c = 0; c < counts_.size(); ++c)
        w0[c] = basis != basis_rot_ ? 0.5 : ((basis_idx_[c] >> p) & 1) ? 0.0 : 1.0;
    return w0;
}

size_t GpuVectorState::marginal_enqueue(ItemQueue& q, size_t qbit, int basis)
{
    check_qbit(qbit);
    check_basis(basis);
    flush();
    if (basis_) return SIZE_MAX;
    const int p = int(nr_bits_ - 1 - qbit);
    size_t first = SIZE_MAX;
    for (size_t c = 0; c < counts_.size(); ++c) {
        const size_t it = q.add_marginal(MargItem{col_ptr(c), dim() / 2, p, basis});
        if (c == 0) first = it;
    }
    return first;
}

std::vector<double> GpuVectorState::marginal_collect(const ItemQueue& q, size_t first, size_t qbit, int basis) const
{
    if (first == SIZE_MAX) return basis_w0(qbit, basis);
    std::vector<double> w0(counts_.size());
    for (size_t c = 0; c < counts_.size(); ++c) w0[c] = q.w0(first + c);
    return w0;
}

std::vector<double> GpuVectorState::marginal_w0(size_t qbit, int basis)
{
    ItemQueue& q = iq();
    const size_t first = marginal_enqueue(q, qbit, basis);
    if (first == SIZE_MAX) return basis_w0(qbit, basis);
    q.launch(stream_);
    HIP_OK(hipStreamSynchronize(stream_));
    return marginal_collect(q, first, qbit, basis);
}

// ------------------------------------------- allocation-free marginal w0

// The scratch of marginal_w0_noalloc: kNormBlocks partial sums per column
// (columns at stride kNormBlocks), then one w0 per column.  516 KiB; each GPU
// the library runs on holds its own copy until the process ends.  Every call
// overwrites the entries it reads, so nothing needs clearing between calls.
static constexpr size_t kNoAllocMaxCols = 256;
static constexpr size_t kNoAllocW0 = kNoAllocMaxCols * size_t(kNormBlocks);
static __device__ __align__(256) double scratch[kNoAllocW0 + kNoAllocMaxCols];

// w0 of each column: the partials summed one after another from 0.0, the
// order w0_from_partials uses on the host.
__global__ void __launch_bounds__(kBlock) k_sum_partials(const double* __restrict__ part, unsigned nblocks,
                                                         unsigned ncols, double* __restrict__ out)
{
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    double s = 0.0;
    for (unsigned b = 0; b < nblocks; ++b) s += part[size_t(c) * nblocks + b];
    out[c] = s;
}

// The device address of `scratch` on each device, looked up once per device
// with that device current.  The array's name is not a device address in
// host code; only these pointers are passed to kernels and copies.
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

static double* noalloc_scratch(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) throw Error(ERR_HIP, "Device index out of range: " + std::to_string(dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            throw Error(ERR_HIP, std::string("hipGetSymbolAddress of the marginal scratch failed: ") +
                                     hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

void GpuVectorState::marginal_w0_noalloc(size_t qbit, double* w0s, size_t len)
{
    check_qbit(qbit);
    const size_t nc = counts_.size();
    if (len < nc) throw err_not_enough_space(len, nc);
    if (!queue_.empty())
        throw Error(ERR_INVALID_ARGUMENT, "Gates are pending on the state: flush it before the allocation-free marginal");
    if (basis_) {
        const std::vector<double> w = basis_w0(qbit, BASIS_Z);
        std::copy(w.begin(), w.end(), w0s);
        return;
    }
    if (nc > kNoAllocMaxCols)
        throw Error(ERR_INVALID_ARGUMENT, "The allocation-free marginal takes at most " +
                                              std::to_string(kNoAllocMaxCols) + " columns; the state has " +
                                              std::to_string(nc));
    int dev = -1;
    HIP_OK(hipGetDevice(&dev));
    if (dev != device_)
        throw Error(ERR_INVALID_ARGUMENT, "The state lives on device " + std::to_string(device_) +
                                              ", but device " + std::to_string(dev) + " is current");
    double* part = noalloc_scratch(dev);
    double* w = part + kNoAllocW0;
    const int p = int(nr_bits_ - 1 - qbit);
    launch_marginal_cols(prec_ == PREC_C64, col_ptr(0), dim(), uint32_t(nc), dim() / 2, p, BASIS_Z, part, stream_);
    hipLaunchKernelGGL(k_sum_partials, dim3(unsigned((nc + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream_, part,
                       kNormBlocks, unsigned(nc), w);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(w0s, w, nc * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
}

// The bits written and the collapse of measure_into (vectorstate.rs:237-329)
// for draws n0s already made: runs of column c with index < n0s[c] read 0.
void GpuVectorState::collapse_enqueue(ItemQueue& q, size_t qbit, size_t cbit, const size_t* n0s,
                                      const std::vector<double>& w0s, uint64_t* res, size_t res_len, bool collapse,
                                      int basis)
{
    check_qbit(qbit);
    check_basis(basis);
    if (res_len < nr_shots_) throw err_not_enough_space(res_len, nr_shots_);
    const size_t nc = counts_.size();
    const uint64_t one_mask = uint64_t(1) << cbit, zero_mask = ~one_mask;
    size_t rs = 0, new_nr = 0;
    for (size_t c = 0; c < nc; ++c) {
        const size_t n0 = std::min(n0s[c], counts_[c]);
        for (size_t r = rs; r < rs + n0; ++r) res[r] &= zero_mask;
        for (size_t r = rs + n0; r < rs + counts_[c]; ++r) res[r] |= one_mask;
        rs += counts_[c];
        new_nr += (n0 == 0 || n0 == counts_[c]) ? 1 : 2;
    }
    if (!collapse) return;
    const int p = int(nr_bits_ - 1 - qbit);
    if (basis_ && new_nr == nc && basis == basis_rot_) return;   // a basis column collapses onto itself
    materialize_enqueue(q);
    const uint64_t nwork = dim() / 2;
    auto add = [&](const void* src, void* d0, void* d1, double s0, double s1) {
        q.add_collapse(CollItem{src, d0, d1, nwork, s0, s1, p, basis, 0});
    };
    std::vector<size_t> ncounts;
    if (new_nr == nc) {
        for (size_t c = 0; c < nc; ++c) {
            const size_t n0 = std::min(n0s[c], counts_[c]);
            void* col = col_ptr(c);
            if (n0 == counts_[c]) add(col, col, nullptr, 1.0 / std::sqrt(w0s[c]), 0.0);
            else add(col, nullptr, col, 0.0, 1.0 / std::sqrt(1.0 - w0s[c]));
            ncounts.push_back(counts_[c]);
        }
    } else {
        void* fresh = dev_alloc(new_nr * dim() * elem_bytes());
        auto fcol = [&](size_t k) { return static_cast<char*>(fresh) + k * dim() * elem_bytes(); };
        size_t ni = 0;
        for (size_t c = 0; c < nc; ++c) {
            const size_t n0 = std::min(n0s[c], counts_[c]), count = counts_[c];
            const double w0 = w0s[c];
            if (n0 == count) {
                add(col_ptr(c), fcol(ni), nullptr, 1.0 / std::sqrt(w0), 0.0);
                ncounts.push_back(count);
                ni += 1;
            } else if (n0 == 0) {
                add(col_ptr(c), nullptr, fcol(ni), 0.0, 1.0 / std::sqrt(1.0 - w0));
                ncounts.push_back(count);
                ni += 1;
            } else {
                add(col_ptr(c), fcol(ni), fcol(ni + 1), 1.0 / std::sqrt(w0), 1.0 / std::sqrt(1.0 - w0));
                ncounts.push_back(n0);
                ncounts.push_back(count - n0);
                ni += 2;
            }
        }
        // the old columns are read by the queued collapses
        q.release_after(amp_, amp_cols_ * dim() * elem_bytes(), stream_);
        amp_ = fresh;
        amp_cols_ = new_nr;
    }
    counts_ = ncounts;
}

void GpuVectorState::measure_with_n0(size_t qbit, size_t cbit, const size_t* n0s, const std::vector<double>& w0s,
                                     uint64_t* res, size_t res_len, bool collapse, int basis)
{
    ItemQueue& q = iq();
    collapse_enqueue(q, qbit, cbit, n0s, w0s, res, res_len, collapse, basis);
    q.launch(stream_);
}

void GpuVectorState::measure_into(size_t qbit, size_t cbit, uint64_t* res, size_t res_len, RngSource& rng,
                                  bool collapse, int basis)
{
    check_qbit(qbit);
    if (res_len < nr_shots_) throw err_not_enough_space(res_len, nr_shots_);
    std::vector<double> w0 = marginal_w0(qbit, basis);
    // n0 ~ Binomial(count, min(w0, 1)) drawn as a count of Bernoulli trials,
    // one uniform per run, columns in order (vectorstate.rs:264-275).
    std::vector<size_t> n0s(counts_.size());
    for (size_t c = 0; c < counts_.size(); ++c) {
        const double p = w0[c] < 1.0 ? w0[c] : 1.0;
        size_t n0 = 0;
        for (size_t r = 0; r < counts_[c]; ++r)
            if (rng.uniform() < p) ++n0;
        n0s[c] = n0;
    }
    measure_with_n0(qbit, cbit, n0s.data(), w0, res, res_len, collapse, basis);
}

// Levels of the canonical tree of one column: sizes shrink by 16 until <= 16.
static TreeDesc tree_shape(uint64_t dim, uint64_t* entries)
{
    TreeDesc td{};
    td.nlev = 1;
    td.size[0] = dim;
    uint64_t total = 0;
    while (td.size[td.nlev - 1] > uint64_t(kTreeFanout)) {
        td.size[td.nlev] = (td.size[td.nlev - 1] + kTreeFanout - 1) / kTreeFanout;
        total += td.size[td.nlev];
        td.nlev += 1;
    }
    *entries = total;
    return td;
}

// The sampling of measure_all_into_helper (vectorstate.rs:106-161): the
// tree levels of every column and one walk per run, queued.
size_t GpuVectorState::sample_enqueue(ItemQueue& q, const double* u)
{
    flush();
    if (basis_ && basis_rot_ != 0) materialize_enqueue(q);   // rotated basis columns are dense in Z
    if (basis_) return SIZE_MAX;
    const size_t nc = counts_.size();
    uint64_t per_col = 0;
    TreeDesc td = tree_shape(dim(), &per_col);
    const size_t need = std::max<size_t>(per_col * nc, 1);
    if (need > tree_cap_) {
        dev_release(d_tree_, tree_cap_ * sizeof(double));
        d_tree_ = nullptr;
        tree_cap_ = 0;
        d_tree_ = static_cast<double*>(dev_alloc(need * sizeof(double)));
        tree_cap_ = need;
    }
    {   // column 0's level pointers; column c's are per_col entries further
        uint64_t o = 0;
        for (int l = 1; l < td.nlev; ++l) {
            td.lev[l] = d_tree_ + o;
            o += td.size[l];
        }
    }
    std::vector<LevelItem> levels;
    levels.reserve(size_t(td.nlev - 1) * nc);
    for (int l = 1; l < td.nlev; ++l)
        for (size_t c = 0; c < nc; ++c)
            levels.push_back(LevelItem{l == 1 ? static_cast<const void*>(col_ptr(c))
                                              : static_cast<const void*>(td.lev[l - 1] + c * per_col),
                                       const_cast<double*>(td.lev[l]) + c * per_col, td.size[l - 1], 0});
    std::vector<uint64_t> first(nc + 1);
    uint64_t run = 0;
    for (size_t c = 0; c < nc; ++c) {
        first[c] = run;
        run += counts_[c];
    }
    first[nc] = run;
    SampleJob job{};
    job.a0 = col_ptr(0);
    job.col_stride = dim();
    job.td = td;
    job.per_col = per_col;
    job.ncols = int32_t(nc);
    return q.add_sample(job, levels, first.data(), u, size_t(run));
}

// Grouping and bit mapping of measure_all_into_helper (vectorstate.rs:106-161)
// on the sampled indices; outcomes grouped per column in ascending index
// order.
void GpuVectorState::sample_finish(const ItemQueue& q, size_t job, const size_t* cbits, size_t ncbits, uint64_t* res,
                                   size_t res_len, bool collapse, int rot)
{
    check_basis(rot);
    if (res_len < nr_shots_) throw err_not_enough_space(res_len, nr_shots_);
    if (ncbits != nr_bits_) throw err_invalid_nr_measurement_bits(ncbits, nr_bits_);
    const size_t nc = counts_.size();
    std::vector<uint64_t> uidx, ucnt, samples;
    const uint64_t* hidx = nullptr;
    if (!basis_) {
        if (job == SIZE_MAX) throw Error(ERR_INVALID_ARGUMENT, "measure_all: no sampling was queued for dense columns");
        hidx = q.indices(job);
        // all-zero or NaN weights make WeightedIndex::new fail (vectorstate.rs:123-125)
        const double* tops = q.tops(job);
        for (size_t c = 0; c < nc; ++c)
            if (counts_[c] != 0 && !(tops[c] > 0.0 && std::isfinite(tops[c])))
                throw Error(ERR_INVALID_WEIGHTS, "Cannot sample from a state with zero or invalid norm");
    }
    size_t run = 0;
    for (size_t c = 0; c < nc; ++c) {
        const size_t count = counts_[c];
        if (count == 0) continue;
        if (basis_) {
            uidx.push_back(basis_idx_[c]);
            ucnt.push_back(count);
            run += count;
            continue;
        }
        samples.assign(hidx + run, hidx + run + count);
        std::sort(samples.begin(), samples.end());
        for (size_t r = 0; r < count;) {
            size_t e = r;
            while (e < count && samples[e] == samples[r]) ++e;
            uidx.push_back(samples[r]);
            ucnt.push_back(e - r);
            r = e;
        }
        run += count;
    }
    uint64_t mask = 0;
    for (size_t k = 0; k < ncbits; ++k) mask |= uint64_t(1) << cbits[k];
    mask = ~mask;
    size_t off = 0;
    for (size_t k = 0; k < uidx.size(); ++k) {
        const uint64_t perm = shuffle_bits(reverse_bits(uidx[k], nr_bits_), cbits, ncbits);
        for (size_t r = off; r < off + ucnt[k]; ++r) res[r] = (res[r] & mask) | perm;
        off += ucnt[k];
    }
    if (collapse) {
        basis_ = true;
        basis_rot_ = rot;
        basis_idx_ = uidx;
        counts_.assign(ucnt.begin(), ucnt.end());
    }
}

void GpuVectorState::measure_all_uniforms(const size_t* cbits, size_t ncbits, const double* u01, uint64_t* res,
                                          size_t res_len, bool collapse, int rot)
{
    check_basis(rot);
    if (res_len < nr_shots_) throw err_not_enough_space(res_len, nr_shots_);
    if (ncbits != nr_bits_) throw err_invalid_nr_measurement_bits(ncbits, nr_bits_);
    ItemQueue& q = iq();
    const size_t job = sample_enqueue(q, u01);
    if (!q.empty()) {
        q.launch(stream_);
        HIP_OK(hipStreamSynchronize(stream_));
    }
    sample_finish(q, job, cbits, ncbits, res, res_len, collapse, rot);
}

void GpuVectorState::measure_all_into(const size_t* cbits, size_t ncbits, uint64_t* res, size_t res_len,
                                      RngSource& rng, bool collapse, int rot)
{
    if (res_len < nr_shots_) throw err_not_enough_space(res_len, nr_shots_);
    if (ncbits != nr_bits_) throw err_invalid_nr_measurement_bits(ncbits, nr_bits_);
    std::vector<double> u(nr_shots_);
    for (size_t r = 0; r < nr_shots_; ++r) u[r] = rng.uniform();
    measure_all_uniforms(cbits, ncbits, u.data(), res, res_len, collapse, rot);
}

void GpuVectorState::reset(size_t bit, RngSource& rng)
{
    check_qbit(bit);
    std::vector<uint64_t> m(nr_shots_, 0);
    measure_into(bit, 0, m.data(), m.size(), rng, true);
    std::vector<uint8_t> control(nr_shots_);
    for (size_t r = 0; r < nr_shots_; ++r) control[r] = m[r] != 0;
    const cdouble x[4] = {0.0, 1.0, 1.0, 0.0};
    apply_conditional_gate(control.data(), control.size(), x, 1, &bit, 1, "X");
}

void GpuVectorState::reset_all() { set_zero_state(); }

void GpuVectorState::read_amplitudes(size_t col, cdouble* out)
{
    flush();
    if (col >= counts_.size()) throw Error(ERR_INVALID_ARGUMENT, "Column index out of range");
    if (basis_ && basis_rot_ == 0) {
        std::fill(out, out + dim(), cdouble(0, 0));
        out[basis_idx_[col]] = cdouble(1, 0);
        return;
    }
    materialize();
    if (prec_ == PREC_C128) {
        HIP_OK(hipMemcpyAsync(out, col_ptr(col), dim() * sizeof(cdouble), hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    } else {
        std::vector<float> f(2 * dim());
        HIP_OK(hipMemcpyAsync(f.data(), col_ptr(col), f.size() * sizeof(float), hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        for (uint64_t i = 0; i < dim(); ++i) out[i] = cdouble(f[2 * i], f[2 * i + 1]);
    }
}


// ------------------------------------------------------------ profiling

void GpuVectorState::set_graph(bool on)
{
    graph_on_ = on;
    if (!on) drop_graphs();
}

// Every replay is launched inside flush(), which synchronises the stream
// before it returns, so no replay is in flight here.  The tables go back to
// the device block cache, guarded by an event on the stream.
void GpuVectorState::drop_graphs()
{
    for (FlushGraph& fg : fgraphs_) {
        if (fg.exec) (void)hipGraphExecDestroy(fg.exec);
        dev_release(fg.table, fg.cap);
    }
    fgraphs_.clear();
}

size_t GpuVectorState::graph_count() const
{
    size_t n = 0;
    for (const FlushGraph& fg : fgraphs_) n += fg.exec != nullptr;
    return n;
}

GpuVectorState::FlushGraph* GpuVectorState::find_graph(const std::vector<uint64_t>& sig)
{
    for (FlushGraph& fg : fgraphs_)
        if (fg.exec && fg.sig == sig) {
            fg.last_use = ++graph_clock_;
            return &fg;
        }
    return nullptr;
}

GpuVectorState::FlushGraph& GpuVectorState::new_graph_slot(size_t bytes)
{
    FlushGraph* slot = nullptr;
    if (fgraphs_.size() < kMaxFlushGraphs) {
        fgraphs_.emplace_back();
        slot = &fgraphs_.back();
    } else {
        slot = &fgraphs_[0];
        for (FlushGraph& fg : fgraphs_)
            if (fg.last_use < slot->last_use) slot = &fg;
    }
    if (slot->exec) (void)hipGraphExecDestroy(slot->exec);
    slot->exec = nullptr;
    slot->sig.clear();
    if (bytes > slot->cap) {
        dev_release(slot->table, slot->cap);
        slot->table = nullptr;
        slot->cap = 0;
        size_t cap = size_t(1) << 16;
        while (cap < bytes) cap <<= 1;
        slot->table = dev_alloc(cap);
        slot->cap = cap;
    }
    slot->last_use = ++graph_clock_;
    return *slot;
}

void GpuVectorState::set_profiling(bool on)
{
    profiling_ = on;
    prof_.clear();
    ev_pending_.clear();
    ev_used_ = 0;
}

int GpuVectorState::prof_begin()
{
    if (!profiling_) return -1;
    if (ev_used_ == ev_pool_.size()) {
        hipEvent_t a, b;
        HIP_OK(hipEventCreate(&a));
        HIP_OK(hipEventCreate(&b));
        ev_pool_.emplace_back(a, b);
    }
    const int slot = int(ev_used_++);
    HIP_OK(hipEventRecord(ev_pool_[slot].first, stream_));
    return slot;
}

void GpuVectorState::prof_end(int slot, const char* name, double bytes)
{
    if (slot < 0) return;
    HIP_OK(hipEventRecord(ev_pool_[slot].second, stream_));
    ev_pending_.push_back(PendingEvent{slot, name, bytes});
}

void GpuVectorState::set_recording(bool on)
{
    recording_ = on;
    if (on) set_profiling(true);
}

void GpuVectorState::add_sample(const float* feat, double tiles, int slot)
{
    SweepSample s;
    std::memcpy(s.feat, feat, sizeof(s.feat));
    s.tiles = tiles;
    s.slot = slot;
    samples_.push_back(s);
}

std::vector<SweepSample> GpuVectorState::take_samples(size_t max)
{
    const std::ptrdiff_t n = std::ptrdiff_t(std::min(max, samples_done_));
    std::vector<SweepSample> out(samples_.begin(), samples_.begin() + n);
    samples_.erase(samples_.begin(), samples_.begin() + n);
    samples_done_ -= size_t(n);
    return out;
}

void GpuVectorState::prof_collect()
{
    if (ev_pending_.empty()) return;
    HIP_OK(hipStreamSynchronize(stream_));
    slot_ms_.assign(ev_used_, 0.0f);
    for (const PendingEvent& pe : ev_pending_) {
        float ms = 0.0f;
        HIP_OK(hipEventElapsedTime(&ms, ev_pool_[pe.pair].first, ev_pool_[pe.pair].second));
        slot_ms_[size_t(pe.pair)] = ms;
        KernelProfile* kp = nullptr;
        for (KernelProfile& k : prof_)
            if (k.name == pe.name) kp = &k;
        if (!kp) {
            prof_.push_back(KernelProfile{pe.name, 0, 0.0, 0.0});
            kp = &prof_.back();
        }
        kp->launches += 1;
        kp->ms += ms;
        kp->bytes += pe.bytes;
    }
    for (size_t i = samples_done_; i < samples_.size(); ++i) {
        SweepSample& sm = samples_[i];
        if (sm.slot >= 0 && size_t(sm.slot) < slot_ms_.size()) sm.ms = slot_ms_[size_t(sm.slot)];
        sm.slot = -1;
    }
    samples_done_ = samples_.size();
    ev_pending_.clear();
    ev_used_ = 0;
}

std::vector<KernelProfile> GpuVectorState::profile() const { return prof_; }

// ------------------------------------------------------------ comparison

template <typename VA, typename VB>
__global__ void __launch_bounds__(kBlock) k_compare(const VA* __restrict__ a, const VB* __restrict__ b, uint64_t N,
                                                    double* __restrict__ out)
{
    double ore = 0, oim = 0, mx = 0, na = 0, nb = 0;
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += step) {
        const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
        ore += ar * br + ai * bi;   // conj(a) * b
        oim += ar * bi - ai * br;
        na += ar * ar + ai * ai;
        nb += br * br + bi * bi;
        const double dr = ar - br, di = ai - bi;
        mx = fmax(mx, sqrt(dr * dr + di * di));
    }
    const double s0 = block_sum(ore);
    __syncthreads();
    const double s1 = block_sum(oim);
    __syncthreads();
    const double s2 = block_sum(na);
    __syncthreads();
    const double s3 = block_sum(nb);
    __shared__ double red[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0;
        for (int k = 0; k < int(blockDim.x) / 64; ++k) m = fmax(m, red[k]);
        double* o = out + 5 * blockIdx.x;
        o[0] = s0;
        o[1] = s1;
        o[2] = m;
        o[3] = s2;
        o[4] = s3;
    }
}

StateComparison GpuVectorState::compare(GpuVectorState& a, size_t col_a, GpuVectorState& b, size_t col_b)
{
    if (a.nr_bits_ != b.nr_bits_) throw Error(ERR_INVALID_ARGUMENT, "States differ in qubit count");
    a.flush();
    b.flush();
    a.materialize();
    b.materialize();
    if (col_a >= a.counts_.size() || col_b >= b.counts_.size())
        throw Error(ERR_INVALID_ARGUMENT, "Column index out of range");
    HIP_OK(hipStreamSynchronize(b.stream_));
    const unsigned nb = 1024;
    double* part = a.partial_buf(5 * nb);
    const uint64_t N = a.dim();
    const void* pa = a.col_ptr(col_a);
    const void* pb = b.col_ptr(col_b);
    const bool a64 = a.prec_ == PREC_C128, b64 = b.prec_ == PREC_C128;
    if (a64 && b64)
        hipLaunchKernelGGL((k_compare<double2, double2>), dim3(nb), dim3(kBlock), 0, a.stream_,
                           static_cast<const double2*>(pa), static_cast<const double2*>(pb), N, part);
    else if (a64)
        hipLaunchKernelGGL((k_compare<double2, float2>), dim3(nb), dim3(kBlock), 0, a.stream_,
                           static_cast<const double2*>(pa), static_cast<const float2*>(pb), N, part);
    else if (b64)
        hipLaunchKernelGGL((k_compare<float2, double2>), dim3(nb), dim3(kBlock), 0, a.stream_,
                           static_cast<const float2*>(pa), static_cast<const double2*>(pb), N, part);
    else
        hipLaunchKernelGGL((k_compare<float2, float2>), dim3(nb), dim3(kBlock), 0, a.stream_,
                           static_cast<const float2*>(pa), static_cast<const float2*>(pb), N, part);
    HIP_OK(hipGetLastError());
    std::vector<double> h(5 * nb);
    HIP_OK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, a.stream_));
    HIP_OK(hipStreamSynchronize(a.stream_));
    StateComparison r;
    for (unsigned k = 0; k < nb; ++k) {
        r.overlap_re += h[5 * k];
        r.overlap_im += h[5 * k + 1];
        r.max_abs = std::max(r.max_abs, h[5 * k + 2]);
        r.norm_a += h[5 * k + 3];
        r.norm_b += h[5 * k + 4];
    }
    return r;
}

// Device-to-device copy of `bytes` (hipMemcpyAsync on a fresh stream),
// best of `reps` timed with HIP events: the read+write streaming rate a
// full pass over HBM can reach on this part.
double measure_copy_ms(size_t bytes, int reps)
{
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double best = 1e300;
    auto release = [&] {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
    };
    try {
        HIP_OK(hipMalloc(&a, bytes));
        HIP_OK(hipMalloc(&b, bytes));
        HIP_OK(hipStreamCreate(&st));
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        HIP_OK(hipMemsetAsync(a, 0, bytes, st));
        HIP_OK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, st));   // first touch
        for (int r = 0; r < reps; ++r) {
            HIP_OK(hipEventRecord(e0, st));
            HIP_OK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, st));
            HIP_OK(hipEventRecord(e1, st));
            HIP_OK(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, double(ms));
        }
    } catch (...) {
        release();
        throw;
    }
    release();
    return best;
}

}  // namespace q1t
