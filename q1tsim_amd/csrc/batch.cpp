// Batched circuit execution (lockstep rounds over many circuits).
#include "batch.h"

#include <algorithm>
#include <chrono>

namespace q1t {

Batch::Batch(Precision prec, int fusion_k, size_t mem_budget)
    : prec_(prec), fusion_k_(fusion_k), mem_budget_(mem_budget)
{
}

void Batch::add(Circuit* c)
{
    if (!c) throw Error(ERR_INVALID_ARGUMENT, "Null circuit");
    c->set_backend(prec_, fusion_k_, true);
    circuits_.push_back(c);
}

// Upper bound on the device bytes of a circuit's state: every collapsing
// single-qubit measurement, reset or conditional gate may split each column
// in two (vectorstate.rs:193-227, 237-329, 402-408); a collapsing
// measure_all leaves up to one column per run.
size_t Batch::resident_bytes(const Circuit& c, size_t nr_shots) const
{
    size_t cols = 1;
    for (const CircuitOp& op : c.ops()) {
        switch (op.type) {
        case OpType::Measure:
        case OpType::Reset:
        case OpType::CondGate:
            cols = std::min(nr_shots, cols * 2);
            break;
        case OpType::MeasureAll:
            cols = nr_shots;
            break;
        default:
            break;
        }
    }
    const size_t elem = prec_ == PREC_C128 ? 16 : 8;
    return std::max<size_t>(cols, 1) * (size_t(1) << c.nr_qbits()) * elem;
}

void Batch::execute(size_t nr_shots, bool keep_states)
{
    const auto t0 = std::chrono::steady_clock::now();
    stats_ = BatchStats{};
    if (!stream_) stream_ = std::make_shared<SharedStream>();
    size_t budget = mem_budget_;
    if (!budget) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) throw Error(ERR_HIP, "hipMemGetInfo failed");
        budget = free_b / 4 * 3;
    }
    size_t next = 0;
    while (next < circuits_.size()) {
        // one wave: as many circuits, in order, as the budget holds (at least one)
        std::vector<Circuit*> wave;
        size_t used = 0;
        while (next < circuits_.size()) {
            const size_t b = resident_bytes(*circuits_[next], nr_shots);
            if (!wave.empty() && used + b > budget) break;
            wave.push_back(circuits_[next]);
            used += b;
            ++next;
        }
        stats_.waves += 1;
        for (Circuit* c : wave) c->start(nr_shots, stream_);
        std::vector<Circuit*> active = wave;
        std::vector<GpuVectorState*> states;
        std::vector<char> more;
        while (!active.empty()) {
            states.clear();
            more.assign(active.size(), 0);
            for (size_t k = 0; k < active.size(); ++k) {
                more[k] = active[k]->queue_gates() ? 1 : 0;
                states.push_back(active[k]->state());
            }
            run_planned_batch(states, stream_->s, prog_, stats_.gates);
            stats_.rounds += 1;
            std::vector<Circuit*> still;
            for (size_t k = 0; k < active.size(); ++k) {
                if (more[k]) {
                    active[k]->run_next_op();
                    stats_.sync_ops += 1;
                    still.push_back(active[k]);
                } else {
                    active[k]->finish();
                    if (!keep_states) active[k]->release_state();
                }
            }
            active.swap(still);
        }
    }
    stats_.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace q1t
