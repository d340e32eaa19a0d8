// Circuit builder and op interpreter (src/circuit.rs:161-554, 562-841).
#include "circuit.h"

#include <algorithm>
#include <map>

namespace q1t {

static void check_qbits(const size_t* qbits, size_t n, size_t nr_qbits)
{
    for (size_t k = 0; k < n; ++k)
        if (qbits[k] >= nr_qbits) throw err_invalid_qbit(qbits[k]);
}

// add_gate (circuit.rs:161-173) with the name/arity checks of ffi.rs:335-367
void Circuit::add_gate(const std::string& name, const size_t* qbits, size_t n, const Param* params, size_t np)
{
    const GateSpec* g = find_gate(name);
    if (!g) throw err_unknown_gate(name);
    if (size_t(g->nparams) != np) throw err_invalid_nr_arguments(np, size_t(g->nparams), g->desc);
    check_qbits(qbits, n, nr_qbits_);
    CircuitOp op{OpType::Gate};
    op.gate = g;
    op.params.assign(params, params + np);
    op.bits.assign(qbits, qbits + n);
    ops_.push_back(std::move(op));
}

// add_conditional_gate (circuit.rs:182-200)
void Circuit::add_conditional_gate(const size_t* control, size_t nc, uint64_t target, const std::string& name,
                                   const size_t* qbits, size_t n, const Param* params, size_t np)
{
    const GateSpec* g = find_gate(name);
    if (!g) throw err_unknown_gate(name);
    if (size_t(g->nparams) != np) throw err_invalid_nr_arguments(np, size_t(g->nparams), g->desc);
    for (size_t k = 0; k < nc; ++k)
        if (control[k] >= nr_cbits_) throw err_invalid_cbit(control[k]);
    check_qbits(qbits, n, nr_qbits_);
    CircuitOp op{OpType::CondGate};
    op.gate = g;
    op.params.assign(params, params + np);
    op.bits.assign(qbits, qbits + n);
    op.control.assign(control, control + nc);
    op.target = target;
    ops_.push_back(std::move(op));
}

// measure_basis / peek_basis (circuit.rs:206-222, 299-315)
void Circuit::measure_basis(size_t qbit, size_t cbit, char basis, bool collapse)
{
    if (qbit >= nr_qbits_) throw err_invalid_qbit(qbit);
    if (cbit >= nr_cbits_) throw err_invalid_cbit(cbit);
    CircuitOp op{collapse ? OpType::Measure : OpType::Peek};
    op.qbit = qbit;
    op.cbit = cbit;
    op.basis = basis;
    ops_.push_back(std::move(op));
}

// measure_all_basis / peek_all_basis (circuit.rs:268-280, 372-384)
void Circuit::measure_all_basis(const size_t* cbits, size_t n, char basis, bool collapse)
{
    for (size_t k = 0; k < n; ++k)
        if (cbits[k] >= nr_cbits_) throw err_invalid_cbit(cbits[k]);
    CircuitOp op{collapse ? OpType::MeasureAll : OpType::PeekAll};
    op.cbits.assign(cbits, cbits + n);
    op.basis = basis;
    ops_.push_back(std::move(op));
}

void Circuit::reset(size_t qbit)
{
    if (qbit >= nr_qbits_) throw err_invalid_qbit(qbit);
    CircuitOp op{OpType::Reset};
    op.qbit = qbit;
    ops_.push_back(std::move(op));
}

void Circuit::reset_all() { ops_.push_back(CircuitOp{OpType::ResetAll}); }

void Circuit::barrier(const size_t* qbits, size_t n)
{
    check_qbits(qbits, n, nr_qbits_);
    CircuitOp op{OpType::Barrier};
    op.bits.assign(qbits, qbits + n);
    ops_.push_back(std::move(op));
}

void Circuit::set_backend(Precision prec, int fusion_k, bool tiling)
{
    prec_ = prec;
    fusion_k_ = fusion_k;
    tiling_ = tiling;
}

void Circuit::set_initial_amplitudes(std::vector<cdouble> amps)
{
    if (amps.size() != (size_t(1) << nr_qbits_))
        throw Error(ERR_INVALID_ARGUMENT, "Initial amplitude vector must have 2^nr_qbits entries");
    init_amps_ = std::move(amps);
    init_mode_ = 1;
}

void Circuit::set_initial_random(uint64_t seed)
{
    init_seed_ = seed;
    init_mode_ = 2;
}

// execute / execute_with (circuit.rs:562-600): fresh state and zeroed c_state
void Circuit::execute(size_t nr_shots)
{
    start(nr_shots, nullptr);
    do_execute();
}

void Circuit::start(size_t nr_shots, std::shared_ptr<SharedStream> stream, bool compact)
{
    state_.reset();
    state_ = std::make_unique<GpuVectorState>(nr_qbits_, nr_shots, prec_, std::move(stream),
                                              compact && init_mode_ == 0);
    state_->set_fusion(fusion_k_);
    state_->set_tiling(tiling_);
    state_->set_graph(graph_);
    state_->set_jit(jit_);
    if (init_mode_ == 1) state_->set_amplitudes(init_amps_.data());
    if (init_mode_ == 2) state_->set_random_normal(init_seed_);
    c_state_.assign(nr_shots, 0);
    has_cstate_ = true;
    cursor_ = 0;
}

// reexecute (circuit.rs:618-641): continue from the previous end state
void Circuit::reexecute()
{
    if (!has_cstate_ || !state_) throw err_not_executed();
    do_execute();
}

void Circuit::apply_named(GpuVectorState& st, const char* name, size_t q)
{
    const GateSpec* g = find_gate(name);
    std::vector<cdouble> m = gate_matrix(*g, nullptr);
    st.apply_gate(m.data(), 1, &q, 1, g->desc);
}

void Circuit::apply_named_all(GpuVectorState& st, const char* name)
{
    const GateSpec* g = find_gate(name);
    std::vector<cdouble> m = gate_matrix(*g, nullptr);
    st.apply_unary_gate_all(m.data(), g->desc);
}

// do_execute_with (circuit.rs:643-762): the op loop, gates queued on the
// state until an op needs their result
void Circuit::do_execute()
{
    cursor_ = 0;
    while (queue_gates()) run_next_op();
    finish();
}

bool Circuit::queue_gates()
{
    GpuVectorState& st = *state_;
    std::vector<double> pv;
    for (; cursor_ < ops_.size(); ++cursor_) {
        const CircuitOp& op = ops_[cursor_];
        if (op.type == OpType::Barrier) continue;
        if (op.type != OpType::Gate) return true;
        pv.clear();
        for (const Param& p : op.params) pv.push_back(p.get());
        std::vector<cdouble> m = gate_matrix(*op.gate, pv.data());
        st.apply_gate(m.data(), gate_nbits(*op.gate), op.bits.data(), op.bits.size(), op.gate->desc);
    }
    return false;
}

void Circuit::run_next_op()
{
    if (cursor_ >= ops_.size()) return;
    run_op(ops_[cursor_]);
    ++cursor_;
}

int Circuit::basis_code(char basis) { return basis == 'X' ? 1 : basis == 'Y' ? 2 : 0; }

bool Circuit::next_is_measure() const
{
    if (cursor_ >= ops_.size()) return false;
    const CircuitOp& op = ops_[cursor_];
    return op.type == OpType::Measure || op.type == OpType::Peek;
}

void Circuit::measure_enqueue(ItemQueue& q)
{
    const CircuitOp& op = ops_[cursor_];
    q_item_ = state_->marginal_enqueue(q, op.qbit, basis_code(op.basis));
}

// The second half of measure_into (vectorstate.rs:264-275): a Binomial draw
// per column as a count of runs with u < min(w0, 1), then the collapse.
void Circuit::measure_complete(ItemQueue& q)
{
    const CircuitOp& op = ops_[cursor_];
    GpuVectorState& st = *state_;
    if (c_state_.size() < st.nr_shots()) throw err_not_enough_space(c_state_.size(), st.nr_shots());
    std::vector<double> w0 = st.marginal_collect(q, q_item_, op.qbit, basis_code(op.basis));
    RngSource rng = rng_.source();
    const std::vector<size_t>& counts = st.counts();
    std::vector<size_t> n0s(counts.size());
    for (size_t c = 0; c < counts.size(); ++c) {
        const double p = w0[c] < 1.0 ? w0[c] : 1.0;
        size_t n0 = 0;
        for (size_t r = 0; r < counts[c]; ++r)
            if (rng.uniform() < p) ++n0;
        n0s[c] = n0;
    }
    st.collapse_enqueue(q, op.qbit, op.cbit, n0s.data(), w0, c_state_.data(), c_state_.size(),
                        op.type == OpType::Measure, basis_code(op.basis));
    q_item_ = SIZE_MAX;
    ++cursor_;
}

bool Circuit::next_is_z_measure_all() const
{
    if (cursor_ >= ops_.size()) return false;
    const CircuitOp& op = ops_[cursor_];
    return (op.type == OpType::MeasureAll || op.type == OpType::PeekAll) && op.basis == 'Z';
}

// measure_all_into (vectorstate.rs:340-344): nr_shots uniforms from this
// circuit's stream, then the sampling
void Circuit::measure_all_enqueue(ItemQueue& q)
{
    GpuVectorState& st = *state_;
    const CircuitOp& op = ops_[cursor_];
    if (c_state_.size() < st.nr_shots()) throw err_not_enough_space(c_state_.size(), st.nr_shots());
    if (op.cbits.size() != st.nr_bits()) throw err_invalid_nr_measurement_bits(op.cbits.size(), st.nr_bits());
    RngSource rng = rng_.source();
    std::vector<double> u(st.nr_shots());
    for (size_t r = 0; r < u.size(); ++r) u[r] = rng.uniform();
    q_item_ = st.sample_enqueue(q, u.data());
}

void Circuit::measure_all_complete(const ItemQueue& q)
{
    const CircuitOp& op = ops_[cursor_];
    state_->sample_finish(q, q_item_, op.cbits.data(), op.cbits.size(), c_state_.data(), c_state_.size(),
                          op.type == OpType::MeasureAll, 0);
    q_item_ = SIZE_MAX;
    ++cursor_;
}

bool Circuit::next_is_conditional() const
{
    return cursor_ < ops_.size() && ops_[cursor_].type == OpType::CondGate;
}

std::vector<uint8_t> Circuit::condition(const CircuitOp& op) const
{
    // control word: bit i of the word is c_state bit control[i] (circuit.rs:655-665)
    std::vector<uint8_t> apply(c_state_.size());
    for (size_t r = 0; r < c_state_.size(); ++r) {
        uint64_t w = 0;
        for (size_t i = 0; i < op.control.size(); ++i) w |= ((c_state_[r] >> op.control[i]) & 1) << i;
        apply[r] = w == op.target;
    }
    return apply;
}

bool Circuit::conditional_enqueue(ItemQueue& q)
{
    const CircuitOp& op = ops_[cursor_];
    const std::vector<uint8_t> apply = condition(op);
    std::vector<double> pv;
    for (const Param& p : op.params) pv.push_back(p.get());
    std::vector<cdouble> m = gate_matrix(*op.gate, pv.data());
    if (!state_->conditional_enqueue(q, apply.data(), apply.size(), m.data(), gate_nbits(*op.gate), op.bits.data(),
                                     op.bits.size(), op.gate->desc))
        return false;
    ++cursor_;
    return true;
}

void Circuit::finish()
{
    state_->flush();
    stats_ = state_->last_stats();
}

void Circuit::run_op(const CircuitOp& op)
{
    GpuVectorState& st = *state_;
    RngSource rng = rng_.source();
    std::vector<double> pv;
    switch (op.type) {
    case OpType::Gate: {
        pv.clear();
        for (const Param& p : op.params) pv.push_back(p.get());
        std::vector<cdouble> m = gate_matrix(*op.gate, pv.data());
        st.apply_gate(m.data(), gate_nbits(*op.gate), op.bits.data(), op.bits.size(), op.gate->desc);
        break;
    }
    case OpType::CondGate: {
        const std::vector<uint8_t> apply = condition(op);
        pv.clear();
        for (const Param& p : op.params) pv.push_back(p.get());
        std::vector<cdouble> m = gate_matrix(*op.gate, pv.data());
        st.apply_conditional_gate(apply.data(), apply.size(), m.data(), gate_nbits(*op.gate), op.bits.data(),
                                  op.bits.size(), op.gate->desc);
        break;
    }
    case OpType::Measure:
    case OpType::Peek: {
        // X: H, measure, H; Y: Sdg, H, measure, H, S (circuit.rs:669-718) --
        // the basis change is folded into the marginal and collapse kernels
        const bool collapse = op.type == OpType::Measure;
        st.measure_into(op.qbit, op.cbit, c_state_.data(), c_state_.size(), rng, collapse, basis_code(op.basis));
        break;
    }
    case OpType::MeasureAll:
    case OpType::PeekAll: {
        // X: H on all, measure_all, H on all; Y: Sdg, H on all, measure_all,
        // H, S on all (circuit.rs:686-748).  After a collapse the columns are
        // basis vectors, and the rotation back is kept with them in closed
        // form (no passes); a peek rotates the state back.
        const bool collapse = op.type == OpType::MeasureAll;
        if (op.basis == 'X') apply_named_all(st, "h");
        if (op.basis == 'Y') {
            apply_named_all(st, "sdg");
            apply_named_all(st, "h");
        }
        st.measure_all_into(op.cbits.data(), op.cbits.size(), c_state_.data(), c_state_.size(), rng, collapse,
                            collapse ? basis_code(op.basis) : 0);
        if (!collapse && op.basis == 'X') apply_named_all(st, "h");
        if (!collapse && op.basis == 'Y') {
            apply_named_all(st, "h");
            apply_named_all(st, "s");
        }
        break;
    }
    case OpType::Reset:
        st.reset(op.qbit, rng);
        break;
    case OpType::ResetAll:
        st.reset_all();
        break;
    case OpType::Barrier:
        break;
    }
}

// histogram_string (circuit.rs:823-841): binary keys of width nr_cbits, cbit 0 last
std::vector<std::pair<std::string, size_t>> Circuit::histogram_string() const
{
    if (!has_cstate_) throw err_not_executed();
    std::map<uint64_t, size_t> h;
    for (uint64_t k : c_state_) h[k] += 1;
    std::vector<std::pair<std::string, size_t>> out;
    for (const auto& kv : h) {
        std::string s;
        uint64_t k = kv.first;
        size_t width = nr_cbits_;
        size_t nb = 0;
        for (uint64_t t = k; t; t >>= 1) ++nb;
        // {:0width$b} prints at least one digit: key 0 with width 0 is "0"
        size_t len = std::max(width, std::max<size_t>(nb, 1));
        s.assign(len, '0');
        for (size_t b = 0; b < len; ++b)
            if ((k >> b) & 1) s[len - 1 - b] = '1';
        out.emplace_back(s, kv.second);
    }
    return out;
}

// histogram_vec (circuit.rs:800-815)
std::vector<size_t> Circuit::histogram_vec() const
{
    if (!has_cstate_) throw err_not_executed();
    std::vector<size_t> res(size_t(1) << nr_cbits_, 0);
    for (uint64_t k : c_state_) {
        if (k >= res.size()) throw Error(ERR_INVALID_ARGUMENT, "classical state exceeds nr_cbits");
        res[k] += 1;
    }
    return res;
}

}  // namespace q1t
