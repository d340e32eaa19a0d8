// Circuit: op list + interpreter, mirroring q1tsim's Circuit (src/circuit.rs)
// with the GPU statevector as its quantum state.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "gvs.h"

namespace q1t {

// Parameter::{Direct, FFIRef} (gates/parameter.rs:24-57): read at apply time.
struct Param {
    double value = 0.0;
    const double* ptr = nullptr;
    double get() const { return ptr ? *ptr : value; }
};

enum class OpType { Gate, CondGate, Measure, MeasureAll, Peek, PeekAll, Reset, ResetAll, Barrier };

struct CircuitOp {
    OpType type;
    const GateSpec* gate = nullptr;
    std::vector<Param> params;
    std::vector<size_t> bits;       // qubits of a gate / barrier
    std::vector<size_t> control;    // classical control bits (ConditionalGate)
    uint64_t target = 0;
    size_t qbit = 0, cbit = 0;
    std::vector<size_t> cbits;      // MeasureAll / PeekAll
    char basis = 'Z';
};

class Circuit {
public:
    Circuit(size_t nr_qbits, size_t nr_cbits) : nr_qbits_(nr_qbits), nr_cbits_(nr_cbits) {}

    size_t nr_qbits() const { return nr_qbits_; }
    size_t nr_cbits() const { return nr_cbits_; }
    const std::vector<uint64_t>* cstate() const { return has_cstate_ ? &c_state_ : nullptr; }

    void add_gate(const std::string& name, const size_t* qbits, size_t n, const Param* params, size_t np);
    void add_conditional_gate(const size_t* control, size_t nc, uint64_t target, const std::string& name,
                              const size_t* qbits, size_t n, const Param* params, size_t np);
    void measure_basis(size_t qbit, size_t cbit, char basis, bool collapse);
    void measure_all_basis(const size_t* cbits, size_t n, char basis, bool collapse);
    void reset(size_t qbit);
    void reset_all();
    void barrier(const size_t* qbits, size_t n);

    void execute(size_t nr_shots);
    void reexecute();

    // Resumable execution (batches): `start` makes the fresh state that
    // `execute` makes, on `stream` (`compact`: |0..0> kept as a compact
    // basis column until gates need it); `queue_gates` queues the gate ops
    // from the cursor on, up to the next op that needs their result;
    // `run_next_op` applies that op; `finish` flushes and keeps the stats.
    // `release_state` drops the state (the classical state stays).
    void start(size_t nr_shots, std::shared_ptr<SharedStream> stream, bool compact = false);
    bool queue_gates();
    void run_next_op();
    void finish();
    void release_state() { state_.reset(); }
    // The op at the cursor as items of a queue that many circuits share
    // (batches).  A measure or peek of one qubit (any basis) runs in two
    // steps around the queue's launch and a synchronisation:
    // measure_enqueue queues the marginals, measure_complete draws this
    // circuit's outcomes, queues the collapses and advances.  A Z-basis
    // measure_all / peek_all likewise (the uniforms drawn when queued).  A
    // conditional gate is queued and advances in one step, unless it needs
    // the per-gate kernels (then nothing is done and false returned).
    static int basis_code(char basis);   // 'Z' / 'X' / 'Y' -> 0 / 1 / 2
    bool at_end() const { return cursor_ >= ops_.size(); }
    // a gate or barrier: left to the next queue_gates
    bool next_is_gate() const
    {
        return cursor_ < ops_.size() && (ops_[cursor_].type == OpType::Gate || ops_[cursor_].type == OpType::Barrier);
    }
    bool next_is_measure() const;
    void measure_enqueue(ItemQueue& q);
    void measure_complete(ItemQueue& q);
    bool next_is_z_measure_all() const;
    void measure_all_enqueue(ItemQueue& q);
    void measure_all_complete(const ItemQueue& q);
    bool next_is_conditional() const;
    bool conditional_enqueue(ItemQueue& q);
    size_t nr_ops() const { return ops_.size(); }
    const std::vector<CircuitOp>& ops() const { return ops_; }
    Precision precision() const { return prec_; }
    int fusion() const { return fusion_k_; }

    std::vector<std::pair<std::string, size_t>> histogram_string() const;
    std::vector<size_t> histogram_vec() const;

    // build-added
    void set_backend(Precision prec, int fusion_k, bool tiling);
    // flush graphs of the state (GpuVectorState::set_graph); a reexecute
    // keeps the state, so its flushes can replay the previous execute's
    void set_jit(int mode)
    {
        jit_ = mode;
        if (state_) state_->set_jit(mode);
    }
    void set_graph(bool on)
    {
        graph_ = on;
        if (state_) state_->set_graph(on);
    }
    void seed(uint64_t s) { rng_.reseed(s); }
    Xoshiro256 rng_state() const { return rng_; }
    void set_rng_state(const Xoshiro256& r) { rng_ = r; }
    void set_initial_amplitudes(std::vector<cdouble> amps);
    void set_initial_random(uint64_t seed);
    GpuVectorState* state() { return state_.get(); }
    const GpuVectorState* state() const { return state_.get(); }
    const RunStats& last_stats() const { return stats_; }

private:
    size_t nr_qbits_, nr_cbits_;
    std::vector<CircuitOp> ops_;
    std::unique_ptr<GpuVectorState> state_;
    std::vector<uint64_t> c_state_;
    bool has_cstate_ = false;
    Xoshiro256 rng_{0x1f67a51423cd2615ULL};
    Precision prec_ = PREC_C128;
    int fusion_k_ = 0;
    bool tiling_ = true;
    bool graph_ = false;
    int jit_ = 0;
    int init_mode_ = 0;            // 0: |0..0>, 1: amplitudes, 2: random normal
    std::vector<cdouble> init_amps_;
    uint64_t init_seed_ = 0;
    RunStats stats_;
    size_t cursor_ = 0;
    size_t q_item_ = SIZE_MAX;     // queue item / job of the op at the cursor

    void do_execute();
    void run_op(const CircuitOp& op);
    std::vector<uint8_t> condition(const CircuitOp& op) const;
    void apply_named(GpuVectorState& st, const char* name, size_t q);
    void apply_named_all(GpuVectorState& st, const char* name);
};

}  // namespace q1t
