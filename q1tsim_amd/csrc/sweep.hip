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

TileGeom tile_geometry(Precision p)
{
    // 64 KiB of LDS per tile either way: two workgroups per CU.
    if (p == PREC_C128) return TileGeom{12, 4, 4};
    return TileGeom{13, 5, 5};
}

struct SweepArgs {
    uint64_t ntiles;
    PosList tset;        // global positions of the tile bits (ascending)
    int tpos[16];        // tpos[j] = global position of local bit j
    const SwPhase* phases;
    const SwOp* ops;
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

template <typename V, int M, int R, int SW>
__global__ void __launch_bounds__(1 << (M - R)) k_sweep(V* __restrict__ amp, uint64_t col_stride, SweepArgs sa)
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

// ------------------------------------------------------------- planning

static std::vector<int> mask_list(uint64_t m)
{
    std::vector<int> r;
    for (int b = 0; b < 64; ++b)
        if ((m >> b) & 1) r.push_back(b);
    return r;
}

bool sweepable(const GateOp& op, std::vector<int>* local_targets)
{
    if (local_targets) local_targets->clear();
    switch (op.kind) {
    case OpKind::Identity:
    case OpKind::PhaseOnes:
        return true;
    case OpKind::Diag:
        return op.targets.size() == 1;
    case OpKind::Dense:
        if (op.targets.size() != 1) return false;
        if (local_targets) local_targets->push_back(op.targets[0]);
        return true;
    default:
        return false;
    }
}

static uint64_t touched_bits(const GateOp& op)
{
    uint64_t m = op.ctrl_mask;
    for (int t : op.targets) m |= uint64_t(1) << t;
    return m;
}

std::vector<PlanStep> plan_gates(const std::vector<PendingGate>& q, int n, const TileGeom& g)
{
    std::vector<PlanStep> steps;
    std::vector<size_t> rem;
    for (size_t i = 0; i < q.size(); ++i)
        if (q[i].op.kind != OpKind::Identity) rem.push_back(i);
    const bool tiles_ok = n >= g.M + 2;
    std::vector<int> lt;
    while (!rem.empty()) {
        if (!tiles_ok) {
            for (size_t i : rem) steps.push_back(PlanStep{false, {}, {i}});
            break;
        }
        uint64_t tile = (uint64_t(1) << g.L) - 1;
        int tcount = g.L;
        uint64_t blocked = 0;
        std::vector<size_t> sel, keep;
        for (size_t i : rem) {
            const GateOp& op = q[i].op;
            const uint64_t bits = touched_bits(op);
            if (bits & blocked) {
                blocked |= bits;
                keep.push_back(i);
                continue;
            }
            if (!sweepable(op, &lt)) {
                blocked |= bits;
                keep.push_back(i);
                continue;
            }
            uint64_t need = 0;
            for (int t : lt)
                if (!((tile >> t) & 1)) need |= uint64_t(1) << t;
            const int nn = __builtin_popcountll(need);
            if (tcount + nn <= g.M) {
                tile |= need;
                tcount += nn;
                sel.push_back(i);
            } else {
                blocked |= bits;
                keep.push_back(i);
            }
        }
        if (sel.empty()) {
            // the first remaining gate cannot run inside a tile
            steps.push_back(PlanStep{false, {}, {rem.front()}});
        } else if (sel.size() == 1 && q[sel[0]].op.kind != OpKind::Dense) {
            // a lone diagonal gate touches only part of the state: its own kernel is cheaper
            steps.push_back(PlanStep{false, {}, sel});
        } else {
            // fill the tile up to M bits with the lowest unused bits
            for (int b = 0; b < n && tcount < g.M; ++b)
                if (!((tile >> b) & 1)) {
                    tile |= uint64_t(1) << b;
                    ++tcount;
                }
            PlanStep st;
            st.sweep = true;
            st.tile = mask_list(tile);
            st.gates = sel;
            steps.push_back(st);
        }
        // drop the consumed gates, keeping the original order of the rest
        const std::set<size_t> done(steps.back().gates.begin(), steps.back().gates.end());
        std::vector<size_t> next;
        for (size_t i : rem)
            if (!done.count(i)) next.push_back(i);
        rem.swap(next);
    }
    return steps;
}

// Build the register phases of one sweep: list scheduling that keeps the
// per-qubit order (a gate waits for every earlier gate that shares a bit)
// and packs one-qubit updates on up to R local bits into each phase.
static void build_program(const std::vector<PendingGate>& q, const PlanStep& step, int M, int R,
                          std::vector<SwPhase>& phases, std::vector<SwOp>& ops)
{
    int local_of[64];
    for (int b = 0; b < 64; ++b) local_of[b] = -1;
    uint64_t tilemask = 0;
    for (size_t j = 0; j < step.tile.size(); ++j) {
        local_of[step.tile[j]] = int(j);
        tilemask |= uint64_t(1) << step.tile[j];
    }
    auto split = [&](uint64_t m, uint32_t& cl, uint64_t& co) {
        cl = 0;
        co = 0;
        for (int b = 0; b < 64; ++b)
            if ((m >> b) & 1) {
                if (local_of[b] >= 0) cl |= 1u << local_of[b];
                else co |= uint64_t(1) << b;
            }
    };
    std::vector<size_t> pending = step.gates;
    while (!pending.empty()) {
        uint32_t regs = 0;
        int nregs = 0;
        uint64_t blocked = 0;
        std::vector<size_t> chosen, rest;
        for (size_t i : pending) {
            const GateOp& op = q[i].op;
            const uint64_t bits = touched_bits(op);
            if (bits & blocked) {
                blocked |= bits;
                rest.push_back(i);
                continue;
            }
            if (op.kind == OpKind::Dense) {
                const int lb = local_of[op.targets[0]];
                if ((regs >> lb) & 1) {
                    chosen.push_back(i);
                } else if (nregs < R) {
                    regs |= 1u << lb;
                    ++nregs;
                    chosen.push_back(i);
                } else {
                    blocked |= bits;
                    rest.push_back(i);
                }
            } else {
                chosen.push_back(i);
            }
        }
        for (int b = 0; b < M && nregs < R; ++b)
            if (!((regs >> b) & 1)) {
                regs |= 1u << b;
                ++nregs;
            }
        SwPhase ph;
        std::memset(&ph, 0, sizeof(ph));
        ph.op_begin = uint32_t(ops.size());
        ph.reg_mask = regs;
        {
            int nr = 0, nt = 0;
            for (int b = 0; b < M; ++b) {
                if ((regs >> b) & 1) ph.rpos[nr++] = uint8_t(b);
                else ph.tbit[nt++] = uint8_t(b);
            }
        }
        for (size_t i : chosen) {
            const GateOp& op = q[i].op;
            SwOp o;
            std::memset(&o, 0, sizeof(o));
            o.dlocal = -1;
            o.dglobal = -1;
            if (op.kind == OpKind::Dense) {
                o.type = SW_DENSE1;
                const int lb = local_of[op.targets[0]];
                int idx = 0;
                for (int b = 0; b < lb; ++b) idx += (regs >> b) & 1;
                o.reg = idx;
                split(op.ctrl_mask, o.cl, o.co);
                for (int k = 0; k < 4; ++k) {
                    o.m[2 * k] = op.m[k].real();
                    o.m[2 * k + 1] = op.m[k].imag();
                }
            } else if (op.kind == OpKind::PhaseOnes) {
                o.type = SW_PHASE;
                split(op.ctrl_mask, o.cl, o.co);
                o.m[0] = op.m[0].real();
                o.m[1] = op.m[0].imag();
            } else {  // Diag on one target
                o.type = SW_DIAG1;
                split(op.ctrl_mask, o.cl, o.co);
                const int t = op.targets[0];
                if (local_of[t] >= 0) o.dlocal = local_of[t];
                else o.dglobal = t;
                o.m[0] = op.m[0].real();
                o.m[1] = op.m[0].imag();
                o.m[2] = op.m[1].real();
                o.m[3] = op.m[1].imag();
            }
            ops.push_back(o);
        }
        ph.op_end = uint32_t(ops.size());
        phases.push_back(ph);
        pending.swap(rest);
    }
    (void)tilemask;
}

template <typename V, int M, int R, int SW>
static void launch_sweep(GpuVectorState& st, const PlanStep& step, const SwPhase* dphases, int nphases,
                         const SwOp* dops, size_t cb, size_t ce)
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
    (void)fusion_k;
    const int n = int(st.nr_bits());
    const TileGeom g = tile_geometry(st.precision());
    if (!tiling) {
        for (const PendingGate& pg : queue) st.apply_op(pg.op, col_begin, col_end);
        return;
    }
    const std::vector<PlanStep> steps = plan_gates(queue, n, g);
    // all sweep programs of this flush in one upload: [phases][ops] per sweep
    std::vector<std::vector<SwPhase>> phases(steps.size());
    std::vector<std::vector<SwOp>> ops(steps.size());
    std::vector<size_t> poff(steps.size()), ooff(steps.size());
    size_t total = 0;
    for (size_t s = 0; s < steps.size(); ++s) {
        if (!steps[s].sweep) continue;
        build_program(queue, steps[s], g.M, g.R, phases[s], ops[s]);
        poff[s] = total;
        total += (phases[s].size() * sizeof(SwPhase) + 255) / 256 * 256;
        ooff[s] = total;
        total += (ops[s].size() * sizeof(SwOp) + 255) / 256 * 256;
    }
    char* dprog = nullptr;
    if (total) {
        std::vector<char> host(total);
        for (size_t s = 0; s < steps.size(); ++s) {
            if (!steps[s].sweep) continue;
            std::memcpy(host.data() + poff[s], phases[s].data(), phases[s].size() * sizeof(SwPhase));
            std::memcpy(host.data() + ooff[s], ops[s].data(), ops[s].size() * sizeof(SwOp));
        }
        dprog = static_cast<char*>(st.scratch(GpuVectorState::SLOT_PROGRAM, total));
        SW_HIP_OK(hipMemcpy(dprog, host.data(), total, hipMemcpyHostToDevice));
    }
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
            launch_sweep<double2, 12, 4, 4>(st, step, dp, int(phases[s].size()), dop, col_begin, col_end);
        else
            launch_sweep<float2, 13, 5, 5>(st, step, dp, int(phases[s].size()), dop, col_begin, col_end);
        stats.launches += 1;
        stats.sweeps += 1;
        stats.gates += step.gates.size();
        stats.bytes += 2.0 * S;
    }
}

}  // namespace q1t
