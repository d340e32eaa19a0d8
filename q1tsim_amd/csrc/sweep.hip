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
#include <array>
#include <atomic>
#include <chrono>
#include <exception>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstring>
#include <map>
#include <set>

#include "costmodel.h"
#include "jit.h"
#include "kernels.h"
#include "sweep_dev.h"

namespace q1t {

#define SW_HIP_OK(expr)                                                                            \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw Error(e_ == hipErrorOutOfMemory ? ERR_OUT_OF_MEMORY : ERR_HIP,                    \
                        std::string("HIP error in " #expr ": ") + hipGetErrorString(e_));          \
    } while (0)

template <typename V, int M, int R, int SW, int BLK>
__global__ void __launch_bounds__(1 << (M - R))
__attribute__((amdgpu_waves_per_eu(sizeof(V) == 8 && M == 12 ? (R == 3 ? 4 : 3) : 2)))
k_sweep(V* __restrict__ amp, uint64_t col_stride, SweepArgs sa)
{
    __shared__ V lds[1 << M];
    V* a = amp + uint64_t(blockIdx.y) * col_stride;
    for (uint64_t tile = blockIdx.x; tile < sa.ntiles; tile += gridDim.x) sweep_tile<V, M, R, SW, BLK>(a, sa, tile, lds);
}

// Sweeps of many states in one launch: item i covers flattened tiles
// [first_tile[i], first_tile[i+1]) of one state, its columns one after the
// other (column c at amp + c * col_stride, 2^log_ntiles tiles each), with the
// state's program.
struct SweepItem {
    void* amp;
    uint64_t first_tile;
    uint64_t col_stride;
    uint32_t ncols, log_ntiles;
    SweepArgs sa;
};

// The complex64 12-bit tiles with 3 register bits and no fused blocks fit
// 64 VGPRs (8 waves per SIMD, 4 workgroups per CU by LDS) without spilling.
template <typename V, int M, int R, int SW, int BLK>
__global__ void __launch_bounds__(1 << (M - R))
__attribute__((amdgpu_waves_per_eu(BLK == 0 && sizeof(V) == 8 && M == 12 && R == 3 ? 8 : 2)))
k_sweep_batched(const SweepItem* __restrict__ items, int nitems, uint64_t total_tiles, uint32_t chunk)
{
    __shared__ V lds[1 << M];
    // grid-stride over runs of `chunk` consecutive tiles: a workgroup keeps
    // one item (and its program, in the scalar cache) for a whole run, while
    // neighbouring workgroups still read neighbouring tiles
    const uint64_t nchunks = (total_tiles + chunk - 1) / chunk;
    for (uint64_t j = blockIdx.x; j < nchunks; j += gridDim.x) {
        const uint64_t t0 = j * chunk;
        const uint64_t t1 = t0 + chunk < total_tiles ? t0 + chunk : total_tiles;
        int lo = 0, hi = nitems - 1;   // last item with first_tile <= t0
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].first_tile <= t0) lo = mid;
            else hi = mid - 1;
        }
        for (uint64_t t = t0; t < t1; ++t) {
            while (lo + 1 < nitems && items[lo + 1].first_tile <= t) ++lo;
            const SweepItem& it = items[lo];
            const uint64_t lt = t - it.first_tile;
            V* a = static_cast<V*>(it.amp) + (lt >> it.log_ntiles) * it.col_stride;
            sweep_tile<V, M, R, SW, BLK>(a, it.sa, lt & ((uint64_t(1) << it.log_ntiles) - 1), lds);
        }
    }
}

// One batched complex64 sweep round on the kernel holding the round's
// widest block (blk as block_level).
template <int M, int R, int SW>
static void launch_batched_c64(int blk, unsigned gx, hipStream_t stream, const SweepItem* items, int nitems,
                               uint64_t total_tiles, uint32_t chunk)
{
    const dim3 grid(gx), block(1u << (M - R));
    if (blk == 2) hipLaunchKernelGGL((k_sweep_batched<float2, M, R, SW, 2>), grid, block, 0, stream, items, nitems, total_tiles, chunk);
    else if (blk == 1) hipLaunchKernelGGL((k_sweep_batched<float2, M, R, SW, 1>), grid, block, 0, stream, items, nitems, total_tiles, chunk);
    else hipLaunchKernelGGL((k_sweep_batched<float2, M, R, SW, 0>), grid, block, 0, stream, items, nitems, total_tiles, chunk);
}

template <typename V, int M, int R, int SW>
static int launch_sweep(GpuVectorState& st, const PlanStep& step, const SwPhase* dphases, int nphases,
                         const SwOp* dops, const float* dblocks, size_t cb, size_t ce, int blk,
                         hipFunction_t jit = nullptr)
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
    int first = -1;
    for (size_t c0 = cb; c0 < ce; c0 += 65535) {
        const unsigned ny = unsigned(std::min<size_t>(65535, ce - c0));
        const int ps = st.prof_begin();
        if (first < 0) first = ps;
        if (jit) {
            // the kernel specialised to this program: same grid, block and arguments
            V* amp = static_cast<V*>(st.col_ptr(c0));
            uint64_t stride = dim;
            void* args[] = {&amp, &stride, &sa};
            SW_HIP_OK(hipModuleLaunchKernel(jit, gx, ny, 1, 1u << (M - R), 1, 1, 0, st.stream(), args, nullptr));
            st.prof_end(ps, sizeof(V) == 16 ? "q1t_jit_sweep_c128" : "q1t_jit_sweep_c64",
                        2.0 * double(dim) * sizeof(V) * ny);
            continue;
        }
        V* amp = static_cast<V*>(st.col_ptr(c0));
        if constexpr (sizeof(V) == 8) {
            if (blk == 2)
                hipLaunchKernelGGL((k_sweep<V, M, R, SW, 2>), dim3(gx, ny), dim3(1 << (M - R)), 0, st.stream(), amp, dim, sa);
            else if (blk == 1)
                hipLaunchKernelGGL((k_sweep<V, M, R, SW, 1>), dim3(gx, ny), dim3(1 << (M - R)), 0, st.stream(), amp, dim, sa);
            else
                hipLaunchKernelGGL((k_sweep<V, M, R, SW, 0>), dim3(gx, ny), dim3(1 << (M - R)), 0, st.stream(), amp, dim, sa);
        } else {
            hipLaunchKernelGGL((k_sweep<V, M, R, SW, 0>), dim3(gx, ny), dim3(1 << (M - R)), 0, st.stream(), amp, dim, sa);
        }
        SW_HIP_OK(hipGetLastError());
        st.prof_end(ps, sizeof(V) == 16 ? "k_sweep_c128" : "k_sweep_c64", 2.0 * double(dim) * sizeof(V) * ny);
    }
    return first;
}

// The widest fused dense block of a program, as the kernels' BLK: 0 none,
// 1 blocks of 3 and 4 qubits, 2 wider ones.
static int block_level(const std::vector<SwPhase>& phases)
{
    int lv = 0;
    for (const SwPhase& ph : phases)
        if (ph.kind == 1) lv = std::max(lv, ph.reg_mask > 4 ? 2 : 1);
    return lv;
}

// One sweep's program: phases, ops and its own block table (block phases
// index it through op_begin until the table is merged into the flush's).
struct SweepProgram {
    std::vector<SwPhase> phases;
    std::vector<SwOp> ops;
    std::vector<float> blocks;
    int nblocks = 0;
};

// kind 0: per-gate register phases; 3 .. 6: dense blocks of that width
static SweepProgram build_candidate(const std::vector<PendingGate>& q, const PlanStep& step, const TileGeom& g,
                                    int kind, bool search, bool factors)
{
    SweepProgram p;
    if (kind == 0) build_program(q, step, g.M, g.R, p.phases, p.ops, search, nullptr, factors);
    else build_fused_program(q, step, g.M, g.R, kind, p.phases, p.ops, p.blocks, &p.nblocks, search, factors);
    return p;
}

// Wall-time estimate of a gate applied on its own: one read and one write
// of the columns at the measured streaming rate.
static double single_gate_us(double state_bytes) { return 2.0 * state_bytes / 4.9e12 * 1e6; }

// Pick the steps and sweep programs of a flush.  Plain (complex128, or
// complex64 without fusion): per-gate register phases on the greedy tiles.
// Fused complex64: greedy fixed-width blocks (PLAN_GREEDY), or the cheaper of
// two tilings -- the per-gate tiling and the fusion tiling, each sweep of
// either with the cheapest of {register phases, 3-qubit blocks, 4-qubit
// blocks, blocks of the configured width when it is 5 or 6} -- as ranked by
// the learned cost model (PLAN_LEARNED), or a seeded random choice among the
// same candidates (PLAN_EXPLORE).
// A state's last plan (choose_plan): its steps and each sweep's program kind.
struct PlanCacheEntry {
    std::vector<PlanStep> steps;
    std::vector<int> kind;
};

static void choose_plan(GpuVectorState& st, const std::vector<PendingGate>& queue, int fusion_k, double state_bytes,
                        double tiles, std::vector<PlanStep>& steps, std::vector<SweepProgram>& progs)
{
    const int n = int(st.nr_bits());
    const TileGeom g = st.geometry();
    const int prec = int(st.precision());
    const bool search = st.plan_search(state_bytes);
    const bool factors = st.tile_factors();   // specialised kernels only
    auto build_all = [&](const std::vector<PlanStep>& ss, int kind) {
        std::vector<SweepProgram> out(ss.size());
        for (size_t s = 0; s < ss.size(); ++s)
            if (ss[s].sweep) out[s] = build_candidate(queue, ss[s], g, kind, search, factors);
        return out;
    };
    const bool fused = st.precision() == PREC_C64 && fusion_k > 0;
    // A flush with the previous flush's gate structure (a repeated circuit,
    // new angles allowed) takes that flush's steps and per-sweep program
    // kinds and builds only those programs: the tile searches (and the
    // learned planner's candidates) are most of the planning time.  Kept for
    // states the tile search runs on.
    std::vector<uint64_t> key;
    auto cache_hit = [&](std::vector<int>& kind) {
        const std::shared_ptr<void>& cached = st.plan_cache_steps();
        if (!cached || key != st.plan_cache_key()) return false;
        const PlanCacheEntry& e = *static_cast<const PlanCacheEntry*>(cached.get());
        steps = e.steps;
        kind = e.kind;
        return true;
    };
    auto cache_store = [&](const std::vector<int>& kind) {
        st.plan_cache_key() = key;
        st.plan_cache_steps() = std::shared_ptr<void>(new PlanCacheEntry{steps, kind},
                                                      [](void* p) { delete static_cast<PlanCacheEntry*>(p); });
    };
    if (!fused) {
        if (search) {
            key = plan_key(queue, n, g, 0, search);
            key.insert(key.begin(), 1);   // the per-gate planner
            std::vector<int> kind;
            if (!cache_hit(kind)) {
                steps = plan_gates(queue, n, g, 0, search);
                cache_store(std::vector<int>(steps.size(), 0));
            }
        } else {
            steps = plan_gates(queue, n, g, 0, search);
        }
        progs = build_all(steps, 0);
        return;
    }
    // blocks of 3 to 6 local qubits: block_mfma holds 3 and 4, block_mfma_wide 5 and 6
    const int Kg = std::min(6, std::max(3, fusion_k));
    CostModel* cm = st.cost_model();
    const int mode = st.plan_mode();
    if (mode == PLAN_GREEDY || (mode == PLAN_LEARNED && (!cm || !cm->trained()))) {
        if (search) {
            key = plan_key(queue, n, g, Kg, search);
            key.insert(key.begin(), 3);   // greedy fusion
            std::vector<int> kind;
            if (!cache_hit(kind)) {
                steps = plan_gates(queue, n, g, Kg, search);
                cache_store(std::vector<int>(steps.size(), Kg));
            }
        } else {
            steps = plan_gates(queue, n, g, Kg, search);
        }
        progs = build_all(steps, Kg);
        return;
    }
    const bool cacheable = search && mode == PLAN_LEARNED;
    if (cacheable) {
        // both tilings' inputs, the block width, the model, the kernels the
        // predictions are for and the bytes swept (single gates are costed
        // by them)
        key = plan_key(queue, n, g, 0, search);
        const std::vector<uint64_t> k2 = plan_key(queue, n, g, Kg, search);
        key.insert(key.end(), k2.begin(), k2.end());
        key.insert(key.begin(), {2, uint64_t(Kg), reinterpret_cast<uintptr_t>(cm), cm->version(),
                                 uint64_t(st.jit_enabled()), uint64_t(factors), uint64_t(prec),
                                 uint64_t(state_bytes)});
        std::vector<int> kind;
        if (cache_hit(kind)) {
            progs.assign(steps.size(), SweepProgram{});
            for (size_t s = 0; s < steps.size(); ++s)
                if (steps[s].sweep) progs[s] = build_candidate(queue, steps[s], g, kind[s], search, factors);
            return;
        }
    }
    // candidates: tiling A (per-gate tiles) and tiling B (fusion tiles), each
    // sweep of either with one program per block width (0: register phases,
    // 3, 4 and, above 4, fusion_k; in tiling A gates whose controls lie
    // outside the tile stay register ops beside the blocks)
    // (host planning runs on parallel threads: both tilings, then every
    // candidate program; a flush's planning would otherwise rival its kernels
    // on states of a few GiB)
    std::vector<PlanStep> tl[2];
    {
        std::exception_ptr e0, e1;
        std::thread other([&] {
            try {
                tl[1] = plan_gates(queue, n, g, Kg, search);
            } catch (...) {
                e1 = std::current_exception();
            }
        });
        try {
            tl[0] = plan_gates(queue, n, g, 0, search);
        } catch (...) {
            e0 = std::current_exception();
        }
        other.join();
        if (e0) std::rethrow_exception(e0);
        if (e1) std::rethrow_exception(e1);
    }
    std::vector<int> kinds = {0, 3, 4};
    if (Kg > 4) kinds.push_back(Kg);
    const int nk = int(kinds.size());
    std::vector<std::vector<SweepProgram>> cand[2];
    std::vector<int> pick[2];
    std::vector<std::array<size_t, 3>> work;   // (tiling, sweep, candidate)
    for (int t = 0; t < 2; ++t) {
        cand[t].assign(tl[t].size(), std::vector<SweepProgram>(kinds.size()));
        pick[t].assign(tl[t].size(), 0);
        for (size_t s = 0; s < tl[t].size(); ++s)
            if (tl[t][s].sweep)
                for (int c = 0; c < nk; ++c) work.push_back({size_t(t), s, size_t(c)});
    }
    {
        std::atomic<size_t> next{0};
        std::exception_ptr err;
        std::mutex err_mu;
        auto run = [&] {
            try {
                for (size_t i = next++; i < work.size(); i = next++) {
                    const auto& w = work[i];
                    cand[w[0]][w[1]][w[2]] = build_candidate(queue, tl[w[0]][w[1]], g, kinds[w[2]], search, factors);
                }
            } catch (...) {
                std::lock_guard<std::mutex> lk(err_mu);
                if (!err) err = std::current_exception();
            }
        };
        const size_t hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
        const size_t nth = std::min(hw, std::max<size_t>(1, work.size() / 4));
        std::vector<std::thread> pool;
        for (size_t i = 1; i < nth; ++i) pool.emplace_back(run);
        run();
        for (auto& th : pool) th.join();
        if (err) std::rethrow_exception(err);
    }
    int use = 1;
    if (mode == PLAN_EXPLORE) {
        Xoshiro256& rng = st.plan_rng();
        use = int(rng.next() >> 63);
        for (int t = 0; t < 2; ++t)
            for (size_t s = 0; s < tl[t].size(); ++s) pick[t][s] = int(rng.next() % uint64_t(nk));
    } else {
        std::vector<float> feat;
        struct Who {
            int t;
            size_t s;
            int c;
        };
        std::vector<Who> who;
        float f[kSweepFeatures];
        for (int t = 0; t < 2; ++t)
            for (size_t s = 0; s < tl[t].size(); ++s)
                if (tl[t][s].sweep)
                    for (int c = 0; c < nk; ++c) {
                        sweep_features(cand[t][s][c].phases, cand[t][s][c].ops, tl[t][s], prec, g, st.jit_enabled(), f);
                        feat.insert(feat.end(), f, f + kSweepFeatures);
                        who.push_back(Who{t, s, c});
                    }
        std::vector<float> y(who.size());
        cm->predict(feat.data(), who.size(), y.data());
        const double us_per_unit = tiles / 1024.0;   // the model predicts log(us per 1024 tiles)
        std::vector<std::vector<double>> pred[2];
        for (int t = 0; t < 2; ++t) pred[t].assign(tl[t].size(), std::vector<double>(nk, 0.0));
        for (size_t i = 0; i < who.size(); ++i)
            pred[who[i].t][who[i].s][who[i].c] = std::exp(double(y[i])) * us_per_unit;
        // per sweep: blocks replace the register program only when predicted at
        // least 10 % cheaper (the fit's error is ~7 % in time, and a wrong pick
        // of blocks costs more than a missed one); a tiling costs the sum of
        // its chosen programs' predictions
        constexpr double kBlockMargin = 0.9;
        double cost[2] = {0.0, 0.0};
        for (int t = 0; t < 2; ++t)
            for (size_t s = 0; s < tl[t].size(); ++s) {
                if (!tl[t][s].sweep) {
                    cost[t] += single_gate_us(state_bytes) * double(tl[t][s].gates.size());
                    continue;
                }
                int c_best = 0;
                for (int c = 1; c < nk; ++c)
                    if (pred[t][s][c] < kBlockMargin * pred[t][s][c_best] ||
                        (c_best != 0 && pred[t][s][c] < pred[t][s][c_best]))
                        c_best = c;
                pick[t][s] = c_best;
                cost[t] += pred[t][s][c_best];
            }
        use = cost[0] < cost[1] ? 0 : 1;
    }
    steps = std::move(tl[use]);
    progs.resize(steps.size());
    std::vector<int> kind(steps.size(), 0);
    for (size_t s = 0; s < steps.size(); ++s)
        if (steps[s].sweep) {
            progs[s] = std::move(cand[use][s][pick[use][s]]);
            kind[s] = kinds[pick[use][s]];
        }
    if (cacheable) cache_store(kind);
}

std::vector<std::string> plan_jit_sources(const std::vector<PendingGate>& q, int n, Precision prec, int fusion_k,
                                          bool search, bool prefetch, bool factors)
{
    const TileGeom g = tile_geometry_specialised(prec);
    const int K = prec == PREC_C64 && fusion_k > 0 ? std::min(6, std::max(3, fusion_k)) : 0;
    const std::vector<PlanStep> steps = plan_gates(q, n, g, K, search);
    std::vector<std::string> out;
    size_t nblk_floats = 0;
    for (const PlanStep& step : steps) {
        if (!step.sweep) continue;
        SweepProgram p = build_candidate(q, step, g, K, search, factors);
        for (SwPhase& ph : p.phases)
            if (ph.kind == 1) ph.op_begin += uint32_t(nblk_floats);
        nblk_floats += p.blocks.size();
        out.push_back(jit_sweep_source(p.phases, p.ops, prec == PREC_C64, g.M, g.R, sweep_swizzle_bits(g.M), prefetch));
    }
    return out;
}

void run_planned(GpuVectorState& st, const std::vector<PendingGate>& queue, size_t col_begin, size_t col_end,
                 int fusion_k, bool tiling, RunStats& stats)
{
    const int n = int(st.nr_bits());
    const TileGeom g = st.geometry();
    if (!tiling) {
        for (const PendingGate& pg : queue) st.apply_op(pg.op, col_begin, col_end);
        return;
    }
    const double S = double(uint64_t(1) << n) * double(st.elem_bytes()) * double(col_end - col_begin);
    const double tiles = double(col_end - col_begin) * double(uint64_t(1) << std::max(0, n - g.M));
    std::vector<PlanStep> steps;
    std::vector<SweepProgram> progs;
    choose_plan(st, queue, fusion_k, S, tiles, steps, progs);

    // all sweep programs of this flush in one upload: [phases][ops] per sweep, then the block table
    std::vector<size_t> poff(steps.size()), ooff(steps.size()), boff_s(steps.size());
    size_t total = 0, nblk_floats = 0;
    int nblocks = 0;
    for (size_t s = 0; s < steps.size(); ++s) {
        if (!steps[s].sweep) continue;
        poff[s] = total;
        total += (progs[s].phases.size() * sizeof(SwPhase) + 255) / 256 * 256;
        ooff[s] = total;
        total += (progs[s].ops.size() * sizeof(SwOp) + 255) / 256 * 256;
        boff_s[s] = nblk_floats;
        nblk_floats += progs[s].blocks.size();
        nblocks += progs[s].nblocks;
    }
    const size_t boff = total;
    total += nblk_floats * sizeof(float);
    std::vector<char> host(total);
    for (size_t s = 0; s < steps.size(); ++s) {
        if (!steps[s].sweep) continue;
        // block phases address the merged table
        for (SwPhase& ph : progs[s].phases)
            if (ph.kind == 1) ph.op_begin += uint32_t(boff_s[s]);
        std::memcpy(host.data() + poff[s], progs[s].phases.data(), progs[s].phases.size() * sizeof(SwPhase));
        std::memcpy(host.data() + ooff[s], progs[s].ops.data(), progs[s].ops.size() * sizeof(SwOp));
        if (!progs[s].blocks.empty())
            std::memcpy(host.data() + boff + boff_s[s] * sizeof(float), progs[s].blocks.data(),
                        progs[s].blocks.size() * sizeof(float));
    }
    // kernels specialised to each sweep's program (compiled on first use)
    std::vector<hipFunction_t> jitfn(steps.size(), nullptr);
    if (st.jit_enabled()) {
        std::vector<std::string> src(steps.size());
        std::vector<std::string> all;
        for (size_t s = 0; s < steps.size(); ++s)
            if (steps[s].sweep) {
                src[s] = jit_sweep_source(progs[s].phases, progs[s].ops, st.precision() == PREC_C64, g.M, g.R,
                                          sweep_swizzle_bits(g.M), st.jit_prefetch());
                all.push_back(src[s]);
            }
        JitCache& jc = JitCache::get();
        jc.compile_all(all);
        for (size_t s = 0; s < steps.size(); ++s)
            if (steps[s].sweep) jitfn[s] = jc.function(src[s], false);
    }
    // ordered after earlier work on this (non-blocking) stream, which may
    // still read the table; the host copy must outlive the transfer
    auto upload = [&](char* dst) {
        if (!total) return;
        SW_HIP_OK(hipMemcpyAsync(dst, host.data(), total, hipMemcpyHostToDevice, st.stream()));
        SW_HIP_OK(hipStreamSynchronize(st.stream()));
    };
    auto launch_all = [&](char* dprog) {
        const float* dblocks = reinterpret_cast<const float*>(dprog + boff);
        for (size_t s = 0; s < steps.size(); ++s) {
            const PlanStep& step = steps[s];
            if (!step.sweep) {
                for (size_t i : step.gates) st.apply_op(queue[i].op, col_begin, col_end);
                continue;
            }
            const SwPhase* dp = reinterpret_cast<const SwPhase*>(dprog + poff[s]);
            const SwOp* dop = reinterpret_cast<const SwOp*>(dprog + ooff[s]);
            const int np = int(progs[s].phases.size());
            const int blk = block_level(progs[s].phases);
            int slot;
            // the compiled tile shapes (M local bits, R register bits): see tile_shape_supported()
            const hipFunction_t jf = jitfn[s];
            if (!jf)
                for (const SwPhase& ph : progs[s].phases)
                    if (ph.kind == 2) throw Error(ERR_HIP, "A sweep program with tile factors needs its specialised kernel");
            const int shape = g.M * 16 + g.R;
            if (st.precision() == PREC_C128) {
                switch (shape) {
                case 12 * 16 + 3: slot = launch_sweep<double2, 12, 3, 4>(st, step, dp, np, dop, dblocks, col_begin, col_end, 0, jf); break;
                case 12 * 16 + 4: slot = launch_sweep<double2, 12, 4, 4>(st, step, dp, np, dop, dblocks, col_begin, col_end, 0, jf); break;
                default: throw Error(ERR_INVALID_ARGUMENT, "No sweep kernel for this complex128 tile shape");
                }
            } else {
                switch (shape) {
                case 13 * 16 + 4: slot = launch_sweep<float2, 13, 4, 5>(st, step, dp, np, dop, dblocks, col_begin, col_end, blk, jf); break;
                case 12 * 16 + 3: slot = launch_sweep<float2, 12, 3, 4>(st, step, dp, np, dop, dblocks, col_begin, col_end, blk, jf); break;
                case 12 * 16 + 4: slot = launch_sweep<float2, 12, 4, 4>(st, step, dp, np, dop, dblocks, col_begin, col_end, blk, jf); break;
                default: throw Error(ERR_INVALID_ARGUMENT, "No sweep kernel for this complex64 tile shape");
                }
            }
            if (st.recording() && col_end - col_begin <= 65535) {
                float f[kSweepFeatures];
                sweep_features(progs[s].phases, progs[s].ops, step, int(st.precision()), g, jf != nullptr, f);
                st.add_sample(f, tiles, slot);
            }
            stats.launches += 1;
            stats.sweeps += 1;
            stats.gates += step.gates.size();
            stats.bytes += 2.0 * S;
            stats.phases += progs[s].phases.size();
            for (const SwPhase& ph : progs[s].phases)
                if (ph.kind == 1)
                    stats.mfma_flops += std::ldexp(double(col_end - col_begin), n + int(ph.reg_mask) + 3);
        }
        stats.blocks += uint64_t(nblocks);
    };

    bool graph = st.graph_enabled();
    std::vector<uint64_t> sig;
    if (graph) {
        // the launch signature: everything a captured launch holds by value
        auto put = [&](uint64_t x) { sig.push_back(x); };
        put(reinterpret_cast<uintptr_t>(st.col_ptr(col_begin)));
        put(col_begin);
        put(col_end);
        put(uint64_t(n));
        put(uint64_t(st.precision()));
        put(uint64_t(g.M));
        put(uint64_t(g.R));
        put(total);
        put(boff);
        put(steps.size());
        for (size_t s = 0; s < steps.size() && graph; ++s) {
            put(steps[s].sweep ? 1 : 0);
            if (steps[s].sweep) {
                // a specialised kernel holds the program's structure in its code
                put(reinterpret_cast<uintptr_t>(jitfn[s]));
                for (int b : steps[s].tile) put(uint64_t(b));
                put(poff[s]);
                put(ooff[s]);
                put(progs[s].phases.size());
                put(uint64_t(block_level(progs[s].phases)));
                continue;
            }
            for (size_t i : steps[s].gates) {
                const GateOp& op = queue[i].op;
                // a kernel of 3+ targets uploads its matrix and synchronises
                if (op.kind == OpKind::Dense && op.targets.size() >= 3) graph = false;
                put(uint64_t(op.kind));
                put(op.ctrl_mask);
                put(op.targets.size());
                for (int t : op.targets) put(uint64_t(t));
                put(op.m.size());
                for (const cdouble& c : op.m) {
                    double re = c.real(), im = c.imag();
                    uint64_t u;
                    std::memcpy(&u, &re, 8);
                    put(u);
                    std::memcpy(&u, &im, 8);
                    put(u);
                }
            }
        }
    }
    if (!graph) {
        char* dprog = total ? static_cast<char*>(st.scratch(GpuVectorState::SLOT_PROGRAM, total)) : nullptr;
        upload(dprog);
        launch_all(dprog);
        return;
    }

    // graph path: replay while the signature repeats, else capture anew
    if (GpuVectorState::FlushGraph* hit = st.find_graph(sig)) {
        upload(static_cast<char*>(hit->table));
        SW_HIP_OK(hipGraphLaunch(hit->exec, st.stream()));
        const RunStats& d = hit->delta;
        stats.bytes += d.bytes;
        stats.launches += d.launches;
        stats.sweeps += d.sweeps;
        stats.gates += d.gates;
        stats.blocks += d.blocks;
        stats.phases += d.phases;
        stats.mfma_flops += d.mfma_flops;
        st.note_graph(false);
        return;
    }
    GpuVectorState::FlushGraph& fg = st.new_graph_slot(total);
    char* table = static_cast<char*>(fg.table);
    upload(table);
    const RunStats before = stats;
    SW_HIP_OK(hipStreamBeginCapture(st.stream(), hipStreamCaptureModeThreadLocal));
    try {
        launch_all(table);
    } catch (...) {
        hipGraph_t part = nullptr;
        (void)hipStreamEndCapture(st.stream(), &part);
        if (part) (void)hipGraphDestroy(part);
        throw;
    }
    hipGraph_t gr = nullptr;
    SW_HIP_OK(hipStreamEndCapture(st.stream(), &gr));
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    SW_HIP_OK(ie);
    fg.exec = exec;
    fg.sig = std::move(sig);
    fg.delta = RunStats{};
    fg.delta.bytes = stats.bytes - before.bytes;
    fg.delta.launches = stats.launches - before.launches;
    fg.delta.sweeps = stats.sweeps - before.sweeps;
    fg.delta.gates = stats.gates - before.gates;
    fg.delta.blocks = stats.blocks - before.blocks;
    fg.delta.mfma_flops = stats.mfma_flops - before.mfma_flops;
    fg.delta.phases = stats.phases - before.phases;
    st.note_graph(true);
    SW_HIP_OK(hipGraphLaunch(exec, st.stream()));
}

// ------------------------------------------------------------- batches

void* DeviceBuffer::get(size_t bytes, hipStream_t stream)
{
    if (bytes > cap) {
        SW_HIP_OK(hipStreamSynchronize(stream));
        if (ptr) SW_HIP_OK(hipFree(ptr));
        ptr = nullptr;
        cap = 0;
        const size_t c = std::max<size_t>(bytes, size_t(1) << 20);
        SW_HIP_OK(hipMalloc(&ptr, c));
        cap = c;
    }
    return ptr;
}

DeviceBuffer::~DeviceBuffer()
{
    if (ptr) (void)hipFree(ptr);
}

void* PinnedBuffer::get(size_t bytes)
{
    if (bytes > cap) {
        if (ptr) SW_HIP_OK(hipHostFree(ptr));
        ptr = nullptr;
        cap = 0;
        const size_t c = std::max<size_t>(bytes, size_t(1) << 20) * 3 / 2;
        SW_HIP_OK(hipHostMalloc(&ptr, c, hipHostMallocDefault));
        cap = c;
    }
    return ptr;
}

PinnedBuffer::~PinnedBuffer()
{
    if (ptr) (void)hipHostFree(ptr);
}

// A gate of a small state, applied in LDS.
struct SmallOp {
    int32_t kind;      // 0: dense on nt targets, 1: diagonal on nt targets, 2: phase on cmask, 3: swap t0/t1
    int32_t nt;
    int32_t t[3];      // target index bits, t[0] = local MSB
    int32_t npos;      // deposit positions (targets and controls, ascending)
    int32_t pos[14];
    uint64_t cmask;
    uint32_t moff;     // offset of the matrix / diagonal / phase in the value table
    uint32_t pad;
};

struct SmallItem {
    void* amp;
    int32_t n;
    int32_t op_begin, op_end;
    int32_t pad;
};

__device__ __forceinline__ uint32_t deposit_small(uint32_t j, const SmallOp& op)
{
#pragma unroll
    for (int k = 0; k < 14; ++k) {
        if (k < op.npos) {
            const int s = op.pos[k];
            j = ((j >> s) << (s + 1)) | (j & ((1u << s) - 1));
        }
    }
    return j;
}

// Whole states in LDS, one workgroup per (state, column) item; gates one
// after the other with a barrier in between.
template <typename V>
__global__ void __launch_bounds__(256) k_small_batch(const SmallItem* __restrict__ items, int nitems,
                                                     const SmallOp* __restrict__ ops, const double2* __restrict__ vals)
{
    extern __shared__ unsigned char smem_small[];
    V* s = reinterpret_cast<V*>(smem_small);
    using Rl = typename RealOf<V>::type;
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const SmallItem& item = items[it];
        V* a = static_cast<V*>(item.amp);
        const uint32_t N = 1u << item.n;
        for (uint32_t i = tid; i < N; i += bd) s[i] = a[i];
        __syncthreads();
        for (int o = item.op_begin; o < item.op_end; ++o) {
            const SmallOp& op = ops[o];
            const double2* mv = vals + op.moff;
            if (op.kind == 0) {
                const int nt = op.nt, D = 1 << nt;
                const uint32_t nwork = N >> op.npos;
                for (uint32_t w = tid; w < nwork; w += bd) {
                    const uint32_t base = deposit_small(w, op) | uint32_t(op.cmask);
                    V x[8], y[8];
                    uint32_t ix[8];
#pragma unroll
                    for (int l = 0; l < 8; ++l) {
                        if (l < D) {
                            uint32_t i = base;
#pragma unroll
                            for (int q = 0; q < 3; ++q)
                                if (q < nt && ((l >> (nt - 1 - q)) & 1)) i |= 1u << op.t[q];
                            ix[l] = i;
                            x[l] = s[i];
                        }
                    }
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        if (r < D) {
                            V acc = mkc<V>(0, 0);
#pragma unroll
                            for (int c = 0; c < 8; ++c)
                                if (c < D) {
                                    const double2 m = mv[r * D + c];
                                    acc = cmac(acc, mkc<V>(Rl(m.x), Rl(m.y)), x[c]);
                                }
                            y[r] = acc;
                        }
                    }
#pragma unroll
                    for (int r = 0; r < 8; ++r)
                        if (r < D) s[ix[r]] = y[r];
                }
            } else if (op.kind == 1) {
                const uint32_t nwork = N >> op.npos;
                for (uint32_t w = tid; w < nwork; w += bd) {
                    const uint32_t i = deposit_small(w, op) | uint32_t(op.cmask);
                    int l = 0;
#pragma unroll
                    for (int q = 0; q < 3; ++q)
                        if (q < op.nt) l = (l << 1) | int((i >> op.t[q]) & 1);
                    const double2 d = mv[l];
                    s[i] = cmul(mkc<V>(Rl(d.x), Rl(d.y)), s[i]);
                }
            } else if (op.kind == 2) {
                const uint32_t nwork = N >> op.npos;
                const double2 d = mv[0];
                const V ph = mkc<V>(Rl(d.x), Rl(d.y));
                for (uint32_t w = tid; w < nwork; w += bd) {
                    const uint32_t i = deposit_small(w, op) | uint32_t(op.cmask);
                    s[i] = cmul(ph, s[i]);
                }
            } else {
                const uint32_t nwork = N >> op.npos;
                const uint32_t b0 = 1u << op.t[0], b1 = 1u << op.t[1];
                for (uint32_t w = tid; w < nwork; w += bd) {
                    const uint32_t i = deposit_small(w, op) | uint32_t(op.cmask);
                    const V x = s[i | b1];
                    s[i | b1] = s[i | b0];
                    s[i | b0] = x;
                }
            }
            __syncthreads();
        }
        for (uint32_t i = tid; i < N; i += bd) a[i] = s[i];
        __syncthreads();
    }
}

static SmallOp make_small_op(const GateOp& g, std::vector<double2>& vals)
{
    SmallOp o;
    std::memset(&o, 0, sizeof(o));
    std::vector<int> pos;
    for (int b = 0; b < 64; ++b)
        if ((g.ctrl_mask >> b) & 1) pos.push_back(b);
    o.cmask = g.ctrl_mask;
    o.moff = uint32_t(vals.size());
    auto push = [&](const cdouble& c) { vals.push_back(double2{c.real(), c.imag()}); };
    switch (g.kind) {
    case OpKind::Dense:
        o.kind = 0;
        o.nt = int(g.targets.size());
        for (int q = 0; q < o.nt; ++q) {
            o.t[q] = g.targets[q];
            pos.push_back(g.targets[q]);
        }
        for (const cdouble& c : g.m) push(c);
        break;
    case OpKind::Diag:
        o.kind = 1;
        o.nt = int(g.targets.size());
        for (int q = 0; q < o.nt; ++q) o.t[q] = g.targets[q];
        for (const cdouble& c : g.m) push(c);
        break;
    case OpKind::PhaseOnes:
        o.kind = 2;
        push(g.m[0]);
        break;
    case OpKind::Swap2:
        o.kind = 3;
        o.nt = 2;
        o.t[0] = g.targets[0];
        o.t[1] = g.targets[1];
        pos.push_back(g.targets[0]);
        pos.push_back(g.targets[1]);
        break;
    default:
        break;
    }
    std::sort(pos.begin(), pos.end());
    o.npos = int(pos.size());
    for (size_t k = 0; k < pos.size(); ++k) o.pos[k] = pos[k];
    return o;
}

void run_planned_batch(const std::vector<GpuVectorState*>& states, hipStream_t stream, DeviceBuffer& prog,
                       PinnedBuffer& stage, ItemQueue& iq, RunStats& agg, unsigned tile_chunk, double* plan_s,
                       double* pack_s)
{
    const auto t_plan = std::chrono::steady_clock::now();
    if (tile_chunk == 0) throw Error(ERR_INVALID_ARGUMENT, "tile chunk must be at least 1");
    std::vector<GpuVectorState*> live;
    for (GpuVectorState* s : states)
        if (!s->pending().empty()) {
            s->prepare_batched_flush(iq);
            live.push_back(s);
        }
    if (!iq.empty()) iq.launch(stream);   // compact columns written out before the sweeps
    if (live.empty()) return;
    const Precision prec = live[0]->precision();
    const TileGeom g = live[0]->geometry();
    const int K = (prec == PREC_C64 && live[0]->fusion() > 0) ? std::min(6, std::max(3, live[0]->fusion())) : 0;
    for (GpuVectorState* s : live)
        if (s->precision() != prec || s->geometry().M != g.M || s->geometry().R != g.R)
            throw Error(ERR_INVALID_ARGUMENT, "A batch needs one precision and tile shape");

    // 1. plans and programs (host), small-state op lists: each state on its
    // own, in parallel over host threads; then merged into one upload
    struct Plan {
        std::vector<PlanStep> steps;
        std::vector<std::vector<SwPhase>> ph;
        std::vector<std::vector<SwOp>> ops;
        std::vector<float> blocks;
        std::vector<SmallOp> sops;
        std::vector<double2> svals;
        std::vector<size_t> poff, ooff;
        std::vector<int> nph;
        std::vector<char> blk;    // per step: block_level of its program
        bool small = false;
        int sop_begin = 0, sop_end = 0;
        RunStats st;
        std::string error;
    };
    std::vector<Plan> plans(live.size());
    auto plan_one = [&](size_t i) {
        GpuVectorState& st = *live[i];
        const std::vector<PendingGate>& q = st.pending();
        Plan& p = plans[i];
        const int n = int(st.nr_bits());
        const double S = double(uint64_t(1) << n) * double(st.elem_bytes()) * double(st.nr_cols());
        bool lds_ok = n < g.M + 2;
        for (const PendingGate& pg : q)
            if (pg.op.kind == OpKind::Dense && pg.op.targets.size() > 3) lds_ok = false;
        if (lds_ok) {
            p.small = true;
            for (const PendingGate& pg : q)
                if (pg.op.kind != OpKind::Identity) p.sops.push_back(make_small_op(pg.op, p.svals));
            p.st.sweeps = 1;
            p.st.launches = 1;
            p.st.gates = q.size();
            p.st.bytes = 2.0 * S;
            return;
        }
        const bool search = live[i]->plan_search(S);
        p.steps = plan_gates(q, n, g, K, search);
        p.ph.resize(p.steps.size());
        p.ops.resize(p.steps.size());
        int nblocks = 0;
        for (size_t s = 0; s < p.steps.size(); ++s) {
            if (!p.steps[s].sweep) continue;
            if (K) build_fused_program(q, p.steps[s], g.M, g.R, K, p.ph[s], p.ops[s], p.blocks, &nblocks, search);
            else build_program(q, p.steps[s], g.M, g.R, p.ph[s], p.ops[s], search);
            p.st.sweeps += 1;
            p.st.launches += 1;
            p.st.gates += p.steps[s].gates.size();
            p.st.bytes += 2.0 * S;
            p.st.phases += p.ph[s].size();
        }
        p.st.blocks = uint64_t(nblocks);
    };
    {
        const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const size_t nth = live.size() < 8 ? 1 : std::min(hw, live.size());
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t i = next++; i < live.size(); i = next++) {
                try {
                    plan_one(i);
                } catch (const std::exception& e) {
                    plans[i].error = e.what();
                }
            }
        };
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nth; ++t) pool.emplace_back(work);
        work();
        for (std::thread& t : pool) t.join();
        for (const Plan& p : plans)
            if (!p.error.empty()) throw Error(ERR_INVALID_ARGUMENT, "Batch planning failed: " + p.error);
    }
    if (plan_s) *plan_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_plan).count();
    const auto t_pack = std::chrono::steady_clock::now();
    // 2. one table in pinned memory, uploaded in one copy: every state's
    // phases and ops, the block matrices, the small-state ops and items, and
    // per round one sweep item per state (all its columns).  Offsets first
    // (serial arithmetic), then the states fill their parts in parallel.
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    size_t off = 0, max_steps = 0, nblock_floats = 0, nsops = 0, nsvals = 0;
    int max_small_n = 0;
    std::vector<size_t> bbase(live.size(), 0), sbase(live.size(), 0), vbase(live.size(), 0);
    for (size_t i = 0; i < live.size(); ++i) {
        Plan& p = plans[i];
        if (p.small) {
            sbase[i] = nsops;
            vbase[i] = nsvals;
            p.sop_begin = int(nsops);
            nsops += p.sops.size();
            nsvals += p.svals.size();
            p.sop_end = int(nsops);
            max_small_n = std::max(max_small_n, int(live[i]->nr_bits()));
            continue;
        }
        bbase[i] = nblock_floats;
        nblock_floats += p.blocks.size();
        p.poff.assign(p.steps.size(), 0);
        p.ooff.assign(p.steps.size(), 0);
        p.nph.assign(p.steps.size(), 0);
        p.blk.assign(p.steps.size(), 0);
        for (size_t s = 0; s < p.steps.size(); ++s) {
            if (!p.steps[s].sweep) continue;
            p.blk[s] = char(block_level(p.ph[s]));
            p.poff[s] = off;
            off = up(off + p.ph[s].size() * sizeof(SwPhase));
            p.ooff[s] = off;
            off = up(off + p.ops[s].size() * sizeof(SwOp));
            p.nph[s] = int(p.ph[s].size());
        }
        max_steps = std::max(max_steps, p.steps.size());
    }
    const size_t boff = off;
    off = up(off + nblock_floats * sizeof(float));
    const size_t soff = off;
    off = up(off + nsops * sizeof(SmallOp));
    const size_t voff = off;
    off = up(off + nsvals * sizeof(double2));
    size_t nsitems = 0;
    for (size_t i = 0; i < live.size(); ++i)
        if (plans[i].small && plans[i].sop_end != plans[i].sop_begin) nsitems += live[i]->nr_cols();
    const size_t sioff = off;
    off = up(off + nsitems * sizeof(SmallItem));
    // sweep items: round-major; item (r, i) at item_at[r][i], tiles from
    // round_first[r][i]
    std::vector<std::pair<size_t, size_t>> round_items(max_steps);   // (first item, count)
    std::vector<uint64_t> round_tiles(max_steps, 0);
    std::vector<char> round_blk(max_steps, 0);
    std::vector<std::vector<uint64_t>> item_idx(max_steps), item_tile(max_steps);
    size_t wi = 0;
    for (size_t r = 0; r < max_steps; ++r) {
        round_items[r].first = wi;
        item_idx[r].assign(live.size(), SIZE_MAX);
        item_tile[r].assign(live.size(), 0);
        uint64_t tiles = 0;
        for (size_t i = 0; i < live.size(); ++i) {
            const Plan& p = plans[i];
            if (p.small || r >= p.steps.size() || !p.steps[r].sweep) continue;
            round_blk[r] = std::max(round_blk[r], p.blk[r]);
            item_idx[r][i] = wi++;
            item_tile[r][i] = tiles;
            tiles += uint64_t(live[i]->nr_cols()) << (live[i]->nr_bits() - size_t(g.M));
        }
        round_items[r].second = wi - round_items[r].first;
        round_tiles[r] = tiles;
    }
    const size_t items_off = off;
    const size_t total = up(items_off + wi * sizeof(SweepItem));
    char* dbase = static_cast<char*>(prog.get(total, stream));
    char* host = static_cast<char*>(stage.get(total));
    {
        // small-state items in state order
        std::vector<size_t> si_first(live.size(), 0);
        size_t k = 0;
        for (size_t i = 0; i < live.size(); ++i) {
            si_first[i] = k;
            if (plans[i].small && plans[i].sop_end != plans[i].sop_begin) k += live[i]->nr_cols();
        }
        std::atomic<size_t> next{0};
        auto fill = [&] {
            for (size_t i = next++; i < live.size(); i = next++) {
                Plan& p = plans[i];
                GpuVectorState& st = *live[i];
                if (p.small) {
                    SmallOp* so = reinterpret_cast<SmallOp*>(host + soff) + sbase[i];
                    for (size_t j = 0; j < p.sops.size(); ++j) {
                        so[j] = p.sops[j];
                        so[j].moff += uint32_t(vbase[i]);
                    }
                    std::memcpy(reinterpret_cast<double2*>(host + voff) + vbase[i], p.svals.data(),
                                p.svals.size() * sizeof(double2));
                    if (p.sop_end != p.sop_begin) {
                        SmallItem* it = reinterpret_cast<SmallItem*>(host + sioff) + si_first[i];
                        for (size_t c = 0; c < st.nr_cols(); ++c)
                            it[c] = SmallItem{st.col_ptr(c), int(st.nr_bits()), p.sop_begin, p.sop_end, 0};
                    }
                    continue;
                }
                std::memcpy(reinterpret_cast<float*>(host + boff) + bbase[i], p.blocks.data(),
                            p.blocks.size() * sizeof(float));
                const int n = int(st.nr_bits());
                for (size_t s = 0; s < p.steps.size(); ++s) {
                    if (!p.steps[s].sweep) continue;
                    SwPhase* ph = reinterpret_cast<SwPhase*>(host + p.poff[s]);
                    for (size_t j = 0; j < p.ph[s].size(); ++j) {
                        ph[j] = p.ph[s][j];
                        if (ph[j].kind == 1) ph[j].op_begin += uint32_t(bbase[i]);
                    }
                    std::memcpy(host + p.ooff[s], p.ops[s].data(), p.ops[s].size() * sizeof(SwOp));
                    SweepItem it;
                    std::memset(&it, 0, sizeof(it));
                    it.amp = st.col_ptr(0);
                    it.col_stride = uint64_t(1) << n;
                    it.ncols = uint32_t(st.nr_cols());
                    it.log_ntiles = uint32_t(n - g.M);
                    it.first_tile = item_tile[s][i];
                    it.sa.ntiles = uint64_t(1) << (n - g.M);
                    it.sa.tset.cnt = g.M;
                    for (int j = 0; j < g.M; ++j) {
                        it.sa.tset.s[j] = p.steps[s].tile[j];
                        it.sa.tpos[j] = p.steps[s].tile[j];
                    }
                    it.sa.phases = reinterpret_cast<const SwPhase*>(dbase + p.poff[s]);
                    it.sa.ops = reinterpret_cast<const SwOp*>(dbase + p.ooff[s]);
                    it.sa.blocks = reinterpret_cast<const float*>(dbase + boff);
                    it.sa.nphases = p.nph[s];
                    std::memcpy(host + items_off + item_idx[s][i] * sizeof(SweepItem), &it, sizeof(it));
                }
            }
        };
        const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const size_t nth = live.size() < 64 ? 1 : std::min(hw, live.size() / 32);
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nth; ++t) pool.emplace_back(fill);
        fill();
        for (std::thread& t : pool) t.join();
    }
    SW_HIP_OK(hipMemcpyAsync(dbase, host, total, hipMemcpyHostToDevice, stream));
    if (pack_s) *pack_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_pack).count();

    // 2. small states: one launch
    if (nsitems) {
        const size_t lds = size_t(live[0]->elem_bytes()) << max_small_n;
        if (prec == PREC_C128) {
            SW_HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small_batch<double2>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
            hipLaunchKernelGGL(k_small_batch<double2>, dim3(unsigned(std::min<size_t>(nsitems, 65535))),
                               dim3(256), lds, stream, reinterpret_cast<const SmallItem*>(dbase + sioff),
                               int(nsitems), reinterpret_cast<const SmallOp*>(dbase + soff),
                               reinterpret_cast<const double2*>(dbase + voff));
        } else {
            SW_HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small_batch<float2>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
            hipLaunchKernelGGL(k_small_batch<float2>, dim3(unsigned(std::min<size_t>(nsitems, 65535))),
                               dim3(256), lds, stream, reinterpret_cast<const SmallItem*>(dbase + sioff),
                               int(nsitems), reinterpret_cast<const SmallOp*>(dbase + soff),
                               reinterpret_cast<const double2*>(dbase + voff));
        }
        SW_HIP_OK(hipGetLastError());
        agg.launches += 1;
    }

    // 3. rounds: single-gate steps of each state, then one batched sweep launch
    for (size_t r = 0; r < max_steps; ++r) {
        for (size_t i = 0; i < live.size(); ++i) {
            const Plan& p = plans[i];
            if (r >= p.steps.size() || p.steps[r].sweep) continue;
            for (size_t gi : p.steps[r].gates) live[i]->apply_op(live[i]->pending()[gi].op, 0, live[i]->nr_cols());
        }
        if (!round_items[r].second) continue;
        const SweepItem* ditems = reinterpret_cast<const SweepItem*>(dbase + items_off) + round_items[r].first;
        const int nit = int(round_items[r].second);
        const uint64_t nchunks = (round_tiles[r] + tile_chunk - 1) / tile_chunk;
        const unsigned gx = unsigned(std::min<uint64_t>(nchunks, 4096));
        if (prec == PREC_C128) {
            if (g.R == 3) hipLaunchKernelGGL((k_sweep_batched<double2, 12, 3, 4, 0>), dim3(gx), dim3(512), 0, stream, ditems, nit, round_tiles[r], tile_chunk);
            else hipLaunchKernelGGL((k_sweep_batched<double2, 12, 4, 4, 0>), dim3(gx), dim3(256), 0, stream, ditems, nit, round_tiles[r], tile_chunk);
        } else if (g.M == 13) {
            launch_batched_c64<13, 4, 5>(round_blk[r], gx, stream, ditems, nit, round_tiles[r], tile_chunk);
        } else if (g.R == 3) {
            launch_batched_c64<12, 3, 4>(round_blk[r], gx, stream, ditems, nit, round_tiles[r], tile_chunk);
        } else {
            launch_batched_c64<12, 4, 4>(round_blk[r], gx, stream, ditems, nit, round_tiles[r], tile_chunk);
        }
        SW_HIP_OK(hipGetLastError());
        agg.launches += 1;
    }
    SW_HIP_OK(hipStreamSynchronize(stream));
    for (size_t i = 0; i < live.size(); ++i) {
        agg.sweeps += plans[i].st.sweeps;
        agg.gates += plans[i].st.gates;
        agg.bytes += plans[i].st.bytes;
        agg.phases += plans[i].st.phases;
        agg.blocks += plans[i].st.blocks;
        live[i]->finish_batched_flush(plans[i].st);
    }
}

}  // namespace q1t
