// Gate-stream planner and multi-gate tile sweeps.
//
// A sweep streams the state once through LDS in tiles of 2^M amplitudes.
// The tile's M local bits are the L lowest index bits (so every tile is made
// of 2^L-element contiguous runs: coalesced HBM traffic) plus up to M-L other
// bits chosen by the planner.  Inside a tile every thread keeps 2^R
// amplitudes in registers; a "phase" fixes which R local bits live in
// registers, so one-qubit updates on those bits are pure register work and a
// phase change is one LDS exchange.  Controls and diagonal gates need no
// local bits: controls outside the tile are a per-tile predicate, a diagonal
// factor on an outer bit is a per-tile constant.
#include "sweep.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <set>

#include "kernels.h"

namespace q1t {

#define SW_HIP_OK(expr)                                                                            \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw Error(e_ == hipErrorOutOfMemory ? ERR_OUT_OF_MEMORY : ERR_HIP,                    \
                        std::string("HIP error in " #expr ": ") + hipGetErrorString(e_));          \
    } while (0)

struct SweepArgs {
    uint64_t ntiles;
    PosList tset;        // global positions of the tile bits (ascending)
    int tpos[16];        // tpos[j] = global position of local bit j
    const SwPhase* phases;
    const SwOp* ops;
    const float* blocks;   // real-embedded block matrices (fused segments)
    int nphases;
};

// LDS slot of local element l: XOR-fold higher bits into the low SW bits so
// strided phase exchanges spread over the banks (an involution).
template <int SW, int M> __device__ __forceinline__ int swz(int l)
{
    int f = 0;
#pragma unroll
    for (int s = SW; s < M; s += SW) f ^= l >> s;
    return l ^ (f & ((1 << SW) - 1));
}

template <typename V, int NR, int J>
__device__ __forceinline__ void dense_pairs(V (&v)[NR], const uint32_t (&lidx)[NR], uint32_t cl, V m0, V m1, V m2, V m3)
{
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if ((k >> J) & 1) continue;
        const int k1 = k | (1 << J);
        if ((lidx[k] & cl) == cl) {
            const V x = v[k], y = v[k1];
            v[k] = cadd(cmul(m0, x), cmul(m1, y));
            v[k1] = cadd(cmul(m2, x), cmul(m3, y));
        }
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Dense K-qubit block Y = U X on the complex64 tile in LDS with FP32 MFMA
// (v_mfma_f32_16x16x4f32, FP32 accumulation).  X is the tile viewed as a
// D x C matrix (rows: the K block bits, columns: the other M-K local bits);
// the complex product is the real GEMM [Yr; Yi] = [[Ur, -Ui], [Ui, Ur]] [Xr; Xi]
// with the 2D x 2D real matrix A in global memory.  Each wave owns 16-column
// chunks, so the in-place update needs no barrier until the block is done.
// Operand layouts: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D[i][j] in lane j + 16(i/4), element i%4.
template <int K, int M, int SW>
__device__ __forceinline__ void block_mfma(float2* lds, const float* __restrict__ A, const SwPhase& ph, int tid)
{
    constexpr int D = 1 << K, D2 = 2 * D, C = 1 << (M - K);
    constexpr int NS = D2 / 4, NRT = D2 / 16;
    const int lane = tid & 63, wave = tid >> 6, nwaves = int(blockDim.x) >> 6;
    const int li = lane & 15, lg = lane >> 4;
    float a[NRT][NS];
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int s = 0; s < NS; ++s) a[rt][s] = A[(16 * rt + li) * D2 + 4 * s + lg];
    uint32_t rbit[K];
#pragma unroll
    for (int j = 0; j < K; ++j) rbit[j] = 1u << ph.rpos[j];
    // local index bits of block row r (row bit K-1-j is local bit rpos[j])
    auto rowpart = [&](int r) {
        uint32_t p = 0;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if ((r >> (K - 1 - j)) & 1) p |= rbit[j];
        return p;
    };
    for (int chunk = wave; chunk < C / 16; chunk += nwaves) {
        const int c = chunk * 16 + li;
        uint32_t cp = 0;
#pragma unroll
        for (int b = 0; b < M - K; ++b)
            if ((c >> b) & 1) cp |= 1u << ph.tbit[b];
        float bv[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int rr = 4 * s + lg;
            const float2 v = lds[swz<SW, M>(int(rowpart(rr & (D - 1)) | cp))];
            bv[s] = rr < D ? v.x : v.y;
        }
        f32x4 acc[NRT];
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt) {
            acc[rt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < NS; ++s) acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][s], bv[s], acc[rt], 0, 0, 0);
        }
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = 16 * rt + 4 * lg + v;
                float* dst = reinterpret_cast<float*>(&lds[swz<SW, M>(int(rowpart(rr & (D - 1)) | cp))]);
                dst[rr < D ? 0 : 1] = acc[rt][v];
            }
    }
}

template <typename V, int M, int R, int SW>
__global__ void __launch_bounds__(1 << (M - R))
__attribute__((amdgpu_waves_per_eu(2)))
k_sweep(V* __restrict__ amp, uint64_t col_stride, SweepArgs sa)
{
    constexpr int NT = 1 << (M - R), NR = 1 << R;
    using Rl = typename RealOf<V>::type;
    __shared__ V lds[1 << M];
    const int tid = threadIdx.x;
    V* a = amp + uint64_t(blockIdx.y) * col_stride;

    // global offset of this thread's part of the natural local order
    uint64_t g_tid = 0;
#pragma unroll
    for (int b = 0; b < M - R; ++b)
        if ((tid >> b) & 1) g_tid |= uint64_t(1) << sa.tpos[b];
    uint64_t g_k[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        uint64_t gk = 0;
#pragma unroll
        for (int b = 0; b < R; ++b)
            if ((k >> b) & 1) gk |= uint64_t(1) << sa.tpos[M - R + b];
        g_k[k] = gk;
    }

    for (uint64_t tile = blockIdx.x; tile < sa.ntiles; tile += gridDim.x) {
        const uint64_t base = deposit(tile, sa.tset);
#pragma unroll
        for (int k = 0; k < NR; ++k) lds[swz<SW, M>(tid + k * NT)] = a[base | g_tid | g_k[k]];
        __syncthreads();

        for (int p = 0; p < sa.nphases; ++p) {
            const SwPhase& ph = sa.phases[p];
            if constexpr (sizeof(V) == 8) {
                if (ph.kind == 1) {
                    // fused dense block: K = reg_mask, matrix at blocks + op_begin
                    const float* A = sa.blocks + ph.op_begin;
                    // the planner emits K = 3 or 4 only
                    if (ph.reg_mask == 3) block_mfma<3, M, SW>(reinterpret_cast<float2*>(lds), A, ph, tid);
                    else block_mfma<4, M, SW>(reinterpret_cast<float2*>(lds), A, ph, tid);
                    __syncthreads();
                    continue;
                }
            }
            // register bits ph.rpos[0..R-1], thread bits ph.tbit[0..M-R-1]
            uint32_t tpart = 0;
#pragma unroll
            for (int b = 0; b < M - R; ++b)
                if ((tid >> b) & 1) tpart |= 1u << ph.tbit[b];
            uint32_t rbit[R];
#pragma unroll
            for (int b = 0; b < R; ++b) rbit[b] = 1u << ph.rpos[b];
            uint32_t lidx[NR];
            V v[NR];
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                uint32_t l = tpart;
#pragma unroll
                for (int b = 0; b < R; ++b)
                    if ((k >> b) & 1) l |= rbit[b];
                lidx[k] = l;
                v[k] = lds[swz<SW, M>(int(l))];
            }
            for (uint32_t oi = ph.op_begin; oi < ph.op_end; ++oi) {
                const SwOp& op = sa.ops[oi];
                if ((base & op.co) != op.co) continue;
                if (op.type == SW_DENSE1) {
                    const V m0 = mkc<V>(Rl(op.m[0]), Rl(op.m[1])), m1 = mkc<V>(Rl(op.m[2]), Rl(op.m[3]));
                    const V m2 = mkc<V>(Rl(op.m[4]), Rl(op.m[5])), m3 = mkc<V>(Rl(op.m[6]), Rl(op.m[7]));
                    switch (op.reg) {
                    case 0: dense_pairs<V, NR, 0>(v, lidx, op.cl, m0, m1, m2, m3); break;
                    case 1: dense_pairs<V, NR, 1>(v, lidx, op.cl, m0, m1, m2, m3); break;
                    case 2: dense_pairs<V, NR, 2>(v, lidx, op.cl, m0, m1, m2, m3); break;
                    case 3: dense_pairs<V, NR, 3>(v, lidx, op.cl, m0, m1, m2, m3); break;
                    default:
                        if constexpr (R > 4) dense_pairs<V, NR, (R > 4 ? 4 : 0)>(v, lidx, op.cl, m0, m1, m2, m3);
                        break;
                    }
                } else if (op.type == SW_PHASE) {
                    const V ph1 = mkc<V>(Rl(op.m[0]), Rl(op.m[1]));
#pragma unroll
                    for (int k = 0; k < NR; ++k)
                        if ((lidx[k] & op.cl) == op.cl) v[k] = cmul(ph1, v[k]);
                } else {  // SW_DIAG1
                    const V d0 = mkc<V>(Rl(op.m[0]), Rl(op.m[1])), d1 = mkc<V>(Rl(op.m[2]), Rl(op.m[3]));
                    if (op.dlocal >= 0) {
#pragma unroll
                        for (int k = 0; k < NR; ++k)
                            if ((lidx[k] & op.cl) == op.cl) v[k] = cmul(((lidx[k] >> op.dlocal) & 1) ? d1 : d0, v[k]);
                    } else {
                        const V d = ((base >> op.dglobal) & 1) ? d1 : d0;
#pragma unroll
                        for (int k = 0; k < NR; ++k)
                            if ((lidx[k] & op.cl) == op.cl) v[k] = cmul(d, v[k]);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < NR; ++k) lds[swz<SW, M>(int(lidx[k]))] = v[k];
            __syncthreads();
        }

#pragma unroll
        for (int k = 0; k < NR; ++k) a[base | g_tid | g_k[k]] = lds[swz<SW, M>(tid + k * NT)];
        __syncthreads();
    }
}

template <typename V, int M, int R, int SW>
static void launch_sweep(GpuVectorState& st, const PlanStep& step, const SwPhase* dphases, int nphases,
                         const SwOp* dops, const float* dblocks, size_t cb, size_t ce)
{
    const int n = int(st.nr_bits());
    SweepArgs sa{};
    sa.ntiles = uint64_t(1) << (n - M);
    sa.tset.cnt = M;
    for (int j = 0; j < M; ++j) {
        sa.tset.s[j] = step.tile[j];
        sa.tpos[j] = step.tile[j];
    }
    sa.phases = dphases;
    sa.ops = dops;
    sa.blocks = dblocks;
    sa.nphases = nphases;
    const uint64_t dim = uint64_t(1) << n;
    const unsigned gx = unsigned(std::min<uint64_t>(sa.ntiles, 2048));
    for (size_t c0 = cb; c0 < ce; c0 += 65535) {
        const unsigned ny = unsigned(std::min<size_t>(65535, ce - c0));
        const int ps = st.prof_begin();
        hipLaunchKernelGGL((k_sweep<V, M, R, SW>), dim3(gx, ny), dim3(1 << (M - R)), 0, st.stream(),
                           static_cast<V*>(st.col_ptr(c0)), dim, sa);
        SW_HIP_OK(hipGetLastError());
        st.prof_end(ps, sizeof(V) == 16 ? "k_sweep_c128" : "k_sweep_c64", 2.0 * double(dim) * sizeof(V) * ny);
    }
}

void run_planned(GpuVectorState& st, const std::vector<PendingGate>& queue, size_t col_begin, size_t col_end,
                 int fusion_k, bool tiling, RunStats& stats)
{
    const int n = int(st.nr_bits());
    const TileGeom g = tile_geometry(st.precision());
    if (!tiling) {
        for (const PendingGate& pg : queue) st.apply_op(pg.op, col_begin, col_end);
        return;
    }
    // dense-block fusion runs on the complex64 state (FP32 MFMA)
    // blocks of 3 or 4 local qubits: the MFMA stage is instantiated for those
    const int K = (st.precision() == PREC_C64 && fusion_k > 0) ? std::min(4, std::max(3, fusion_k)) : 0;
    const std::vector<PlanStep> steps = plan_gates(queue, n, g, K);
    // all sweep programs of this flush in one upload: [phases][ops] per sweep, then the block table
    std::vector<std::vector<SwPhase>> phases(steps.size());
    std::vector<std::vector<SwOp>> ops(steps.size());
    std::vector<float> blocks;
    std::vector<size_t> poff(steps.size()), ooff(steps.size());
    size_t total = 0;
    int nblocks = 0;
    for (size_t s = 0; s < steps.size(); ++s) {
        if (!steps[s].sweep) continue;
        if (K) build_fused_program(queue, steps[s], g.M, g.R, K, phases[s], ops[s], blocks, &nblocks);
        else build_program(queue, steps[s], g.M, g.R, phases[s], ops[s]);
        poff[s] = total;
        total += (phases[s].size() * sizeof(SwPhase) + 255) / 256 * 256;
        ooff[s] = total;
        total += (ops[s].size() * sizeof(SwOp) + 255) / 256 * 256;
    }
    const size_t boff = total;
    total += blocks.size() * sizeof(float);
    char* dprog = nullptr;
    if (total) {
        std::vector<char> host(total);
        for (size_t s = 0; s < steps.size(); ++s) {
            if (!steps[s].sweep) continue;
            std::memcpy(host.data() + poff[s], phases[s].data(), phases[s].size() * sizeof(SwPhase));
            std::memcpy(host.data() + ooff[s], ops[s].data(), ops[s].size() * sizeof(SwOp));
        }
        if (!blocks.empty()) std::memcpy(host.data() + boff, blocks.data(), blocks.size() * sizeof(float));
        dprog = static_cast<char*>(st.scratch(GpuVectorState::SLOT_PROGRAM, total));
        SW_HIP_OK(hipMemcpy(dprog, host.data(), total, hipMemcpyHostToDevice));
    }
    const float* dblocks = reinterpret_cast<const float*>(dprog + boff);
    const double S = double(uint64_t(1) << n) * double(st.elem_bytes()) * double(col_end - col_begin);
    for (size_t s = 0; s < steps.size(); ++s) {
        const PlanStep& step = steps[s];
        if (!step.sweep) {
            for (size_t i : step.gates) st.apply_op(queue[i].op, col_begin, col_end);
            continue;
        }
        const SwPhase* dp = reinterpret_cast<const SwPhase*>(dprog + poff[s]);
        const SwOp* dop = reinterpret_cast<const SwOp*>(dprog + ooff[s]);
        if (st.precision() == PREC_C128)
            launch_sweep<double2, 12, 4, 4>(st, step, dp, int(phases[s].size()), dop, dblocks, col_begin, col_end);
        else
            launch_sweep<float2, 13, 5, 5>(st, step, dp, int(phases[s].size()), dop, dblocks, col_begin, col_end);
        stats.launches += 1;
        stats.sweeps += 1;
        stats.gates += step.gates.size();
        stats.bytes += 2.0 * S;
    }
    stats.blocks += uint64_t(nblocks);
}

}  // namespace q1t
