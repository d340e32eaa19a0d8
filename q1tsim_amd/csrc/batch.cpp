// Batched circuit execution (lockstep rounds over many circuits).
#include "batch.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <mutex>
#include <thread>

namespace q1t {

Batch::Batch(Precision prec, int fusion_k, size_t mem_budget)
    : prec_(prec), fusion_k_(fusion_k), geom_(tile_geometry(prec)), mem_budget_(mem_budget)
{
}

Batch::~Batch()
{
    if (iq_) {
        try {
            ItemQueuePool::give(std::move(iq_));
        } catch (...) {
        }
    }
}

void Batch::set_config(const TileGeom& g, int fusion_k)
{
    if (!tile_shape_supported(prec_, g.M, g.R) || g.L < 1 || g.L > g.M)
        throw Error(ERR_INVALID_ARGUMENT, "Unsupported tile geometry for this precision");
    if (fusion_k < 0 || fusion_k > 6) throw Error(ERR_INVALID_ARGUMENT, "fusion_k must be 0..6");
    geom_ = g;
    fusion_k_ = fusion_k;
    for (Circuit* c : circuits_) c->set_backend(prec_, fusion_k_, true);
}

void Batch::set_tile_chunk(unsigned c)
{
    if (c < 1 || c > 1024) throw Error(ERR_INVALID_ARGUMENT, "tile chunk must be 1..1024");
    tile_chunk_ = c;
}

std::vector<BatchConfig> Batch::autotune(size_t nr_shots, size_t sample)
{
    std::vector<BatchConfig> cand;
    if (prec_ == PREC_C128) {
        cand.push_back({TileGeom{12, 4, 4}, 0, 0.0});
        cand.push_back({TileGeom{12, 3, 4}, 0, 0.0});
    } else {
        for (int k : {0, 4}) {
            cand.push_back({TileGeom{12, 4, 4}, k, 0.0});
            cand.push_back({TileGeom{13, 4, 5}, k, 0.0});
        }
        cand.push_back({TileGeom{12, 3, 4}, 0, 0.0});
    }
    sample = std::min(sample, circuits_.size());
    if (!sample) return cand;
    std::vector<Circuit*> part(circuits_.begin(), circuits_.begin() + std::ptrdiff_t(sample));
    std::vector<Xoshiro256> saved;
    for (Circuit* c : part) saved.push_back(c->rng_state());
    size_t best = 0;
    for (size_t i = 0; i < cand.size(); ++i) {
        geom_ = cand[i].geom;
        fusion_k_ = cand[i].fusion_k;
        for (size_t k = 0; k < part.size(); ++k) {
            part[k]->set_backend(prec_, fusion_k_, true);
            part[k]->set_rng_state(saved[k]);
        }
        const auto t0 = std::chrono::steady_clock::now();
        run(part, nr_shots, false);
        cand[i].seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (cand[i].seconds < cand[best].seconds) best = i;
    }
    for (size_t k = 0; k < part.size(); ++k) part[k]->set_rng_state(saved[k]);
    set_config(cand[best].geom, cand[best].fusion_k);
    return cand;
}

void Batch::add(Circuit* c)
{
    if (!c) throw Error(ERR_INVALID_ARGUMENT, "Null circuit");
    c->set_backend(prec_, fusion_k_, true);
    circuits_.push_back(c);
}

// Upper bound on the device bytes of a circuit's state.  A collapsing
// single-qubit measurement or a reset may split every column in two
// (vectorstate.rs:237-329, 402-408); a conditional gate splits columns only
// when some run disagrees with the rest of its column on a control bit
// (vectorstate.rs:193-227), which cannot happen for control bits written by
// an earlier collapsing measurement (each column then holds one outcome); a
// collapsing measure_all leaves basis columns that take device memory only
// if gates follow.
size_t Batch::resident_bytes(const Circuit& c, size_t nr_shots) const
{
    const std::vector<CircuitOp>& ops = c.ops();
    size_t cols = 1, peak = 1;
    std::vector<char> det(c.nr_cbits(), 0);
    for (size_t i = 0; i < ops.size(); ++i) {
        const CircuitOp& op = ops[i];
        switch (op.type) {
        case OpType::Measure:
            cols = std::min(nr_shots, cols * 2);
            det[op.cbit] = 1;
            break;
        case OpType::Peek:
            det[op.cbit] = 0;
            break;
        case OpType::Reset:
            cols = std::min(nr_shots, cols * 2);
            break;
        case OpType::CondGate: {
            bool all = true;
            for (size_t b : op.control) all = all && det[b];
            if (!all) cols = std::min(nr_shots, cols * 2);
            break;
        }
        case OpType::MeasureAll: {
            // The basis columns become dense (up to nr_shots of them) when
            // gates touch them: this op's own X/Y restoring gates, a later
            // gate or conditional gate, a reset (its conditional X), or a
            // later X/Y-basis measurement (its basis change), up to the next
            // reset_all.
            bool dense = op.basis != 'Z';
            for (size_t j = i + 1; j < ops.size() && !dense && ops[j].type != OpType::ResetAll; ++j) {
                const CircuitOp& o = ops[j];
                dense = o.type == OpType::Gate || o.type == OpType::CondGate || o.type == OpType::Reset ||
                        ((o.type == OpType::Measure || o.type == OpType::Peek || o.type == OpType::MeasureAll ||
                          o.type == OpType::PeekAll) && o.basis != 'Z');
            }
            if (dense) cols = nr_shots;
            for (size_t b : op.cbits) det[b] = 1;
            break;
        }
        case OpType::PeekAll:
            for (size_t b : op.cbits) det[b] = 0;
            break;
        case OpType::ResetAll:
            cols = 1;
            break;
        default:
            break;
        }
        peak = std::max(peak, cols);
    }
    const size_t elem = prec_ == PREC_C128 ? 16 : 8;
    return DeviceCache::size_class(peak * (size_t(1) << c.nr_qbits()) * elem);
}

void Batch::execute(size_t nr_shots, bool keep_states)
{
    const auto t0 = std::chrono::steady_clock::now();
    stats_ = BatchStats{};
    run(circuits_, nr_shots, keep_states);
    stats_.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Queue the gate ops of every circuit up to its next non-gate op, on host
// threads (each circuit builds its own matrices and analyses its own gates).
static std::vector<char> queue_all(const std::vector<Circuit*>& active)
{
    std::vector<char> more(active.size(), 0);
    const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nth = active.size() < 64 ? 1 : std::min(hw, active.size() / 32);
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex mu;
    auto work = [&] {
        try {
            for (size_t k = next++; k < active.size(); k = next++) more[k] = active[k]->queue_gates() ? 1 : 0;
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu);
            if (!err) err = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nth; ++t) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
    if (err) std::rethrow_exception(err);
    return more;
}

// The non-gate ops at the circuits' cursors, until every circuit is at a
// gate op or at its end.  Each pass takes one op of every circuit:
// one-qubit measurements, Z-basis measure_all and conditional gates of all
// circuits as items of one queue (one launch per kind, one synchronisation,
// then every circuit's draws from its own random stream and one launch of
// the collapses); any other op circuit by circuit.
void Batch::sync_ops(std::vector<Circuit*>& active)
{
    ItemQueue& q = *iq_;
    auto tick = std::chrono::steady_clock::now();
    auto lap = [&](double& acc) {
        const auto now = std::chrono::steady_clock::now();
        acc += std::chrono::duration<double>(now - tick).count();
        tick = now;
    };
    for (;;) {
        std::vector<Circuit*> meas, all, solo;
        bool any = false;
        for (Circuit* c : active) {
            if (c->at_end() || c->next_is_gate()) continue;
            any = true;
            if (c->next_is_measure()) meas.push_back(c);
            else if (c->next_is_z_measure_all()) all.push_back(c);
            else if (!(c->next_is_conditional() && c->conditional_enqueue(q))) solo.push_back(c);
        }
        if (!any) break;
        for (Circuit* c : meas) c->measure_enqueue(q);
        for (Circuit* c : all) c->measure_all_enqueue(q);
        const uint64_t l0 = q.launches();
        if (!q.empty()) q.launch(stream_->s);
        if (!meas.empty() || !all.empty()) {
            if (hipStreamSynchronize(stream_->s) != hipSuccess) throw Error(ERR_HIP, "measurement launch failed");
            for (Circuit* c : meas) c->measure_complete(q);
            for (Circuit* c : all) c->measure_all_complete(q);
            if (!q.empty()) q.launch(stream_->s);
        }
        stats_.column_launches += q.launches() - l0;
        stats_.sync_ops += meas.size() + all.size();
        lap(stats_.measure_s);
        for (Circuit* c : solo) {
            c->run_next_op();
            stats_.sync_ops += 1;
            stats_.solo_ops += 1;
        }
        lap(stats_.other_s);
    }
}

void Batch::run(const std::vector<Circuit*>& list, size_t nr_shots, bool keep_states)
{
    if (!stream_) stream_ = std::make_shared<SharedStream>();
    if (!iq_) iq_ = ItemQueuePool::take(prec_ == PREC_C64);
    size_t budget = mem_budget_;
    if (!budget) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) throw Error(ERR_HIP, "hipMemGetInfo failed");
        // blocks held by the device block cache are free for the batch
        budget = (free_b + DeviceCache::get().cached_bytes()) / 4 * 3;
    }
    size_t next = 0;
    while (next < list.size()) {
        // one wave: as many circuits, in order, as the budget holds (at least one)
        std::vector<Circuit*> wave;
        size_t used = 0;
        while (next < list.size()) {
            const size_t b = resident_bytes(*list[next], nr_shots);
            if (!wave.empty() && used + b > budget) break;
            wave.push_back(list[next]);
            used += b;
            ++next;
        }
        stats_.waves += 1;
        auto tick = std::chrono::steady_clock::now();
        auto lap = [&](double& acc) {
            const auto now = std::chrono::steady_clock::now();
            acc += std::chrono::duration<double>(now - tick).count();
            tick = now;
        };
        // states start as compact |0..0> columns: written out by the first
        // flush, all of them in one launch
        for (Circuit* c : wave) {
            c->start(nr_shots, stream_, true);
            c->state()->set_geometry(geom_);
        }
        lap(stats_.start_s);
        std::vector<Circuit*> active = wave;
        std::vector<GpuVectorState*> states;
        while (!active.empty()) {
            const std::vector<char> more = queue_all(active);
            states.clear();
            for (Circuit* c : active) states.push_back(c->state());
            lap(stats_.queue_s);
            const uint64_t l0 = iq_->launches();
            run_planned_batch(states, stream_->s, prog_, stage_, *iq_, stats_.gates, tile_chunk_, &stats_.plan_s,
                              &stats_.pack_s);
            stats_.column_launches += iq_->launches() - l0;
            stats_.rounds += 1;
            lap(stats_.flush_s);
            std::vector<Circuit*> still;
            for (size_t k = 0; k < active.size(); ++k) {
                if (more[k]) {
                    still.push_back(active[k]);
                } else {
                    active[k]->finish();
                    if (!keep_states) active[k]->release_state();
                }
            }
            sync_ops(still);
            tick = std::chrono::steady_clock::now();
            active.swap(still);
        }
    }
}

}  // namespace q1t
