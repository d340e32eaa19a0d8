// extern "C" layer: the ffi.rs-compatible circuit ABI (include/q1tsim.h) and
// the state-level GpuVectorState ABI (include/q1tsim_gvs.h).  No C++
// exception or HIP failure crosses this boundary: every failure becomes an
// error code / RESULT_ERROR with a message.
#include <cstdlib>
#include <cstring>

#include "q1tsim.h"
#include "batch.h"
#include "circuit.h"
#include "costmodel.h"
#include "jit.h"
#include "sweep.h"

using namespace q1t;

struct circuit {
    Circuit c;
    circuit(size_t q, size_t b) : c(q, b) {}
};

struct q1t_gvs {
    GpuVectorState s;
    q1t_gvs(size_t n, size_t shots, Precision p) : s(n, shots, p) {}
};

struct q1t_xoshiro {
    Xoshiro256 x;
};

struct q1t_batch {
    Batch b;
    q1t_batch(Precision p, int k, size_t budget) : b(p, k, budget) {}
};

static thread_local std::string g_last_error;

extern "C" const char* q1t_last_error(void) { return g_last_error.c_str(); }

template <typename F> static int guard(F&& f)
{
    try {
        f();
        return 0;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "Out of host memory";
        return ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return ERR_INVALID_ARGUMENT;
    }
}

// ------------------------------------------------------------- CResult

static char* dup_cstr(const std::string& s)
{
    char* p = static_cast<char*>(std::malloc(s.size() + 1));
    std::memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

static result_t res_empty() { return result_t{nullptr, 0, 0, RESULT_EMPTY}; }
static result_t res_error(const std::string& msg) { return result_t{dup_cstr(msg), 0, 0, RESULT_ERROR}; }

template <typename F> static result_t guard_result(F&& f)
{
    int rc = guard(f);
    if (rc) return res_error(g_last_error);
    return res_empty();
}

extern "C" void result_free(result_t res)
{
    switch (res.restype) {
    case RESULT_ERROR:
    case RESULT_STRING:
        std::free(res.data);
        break;
    case RESULT_HISTOGRAM: {
        histelem_t* e = static_cast<histelem_t*>(res.data);
        for (size_t k = 0; k < res.length; ++k) std::free(const_cast<char*>(e[k].key));
        std::free(e);
        break;
    }
    case RESULT_CSTATE:
        std::free(res.data);
        break;
    default:
        break;
    }
}

// ------------------------------------------------------------- circuit

extern "C" circuit_t* circuit_new(size_t nr_qbits, size_t nr_cbits)
{
    try {
        return new circuit(nr_qbits, nr_cbits);
    } catch (...) {
        return nullptr;
    }
}

extern "C" void circuit_free(circuit_t* ptr) { delete ptr; }

extern "C" size_t circuit_nr_qbits(const circuit_t* ptr)
{
    if (!ptr) {
        g_last_error = "Pointer to circuit is NULL";
        return 0;
    }
    return ptr->c.nr_qbits();
}

extern "C" size_t circuit_nr_cbits(const circuit_t* ptr)
{
    if (!ptr) {
        g_last_error = "Pointer to circuit is NULL";
        return 0;
    }
    return ptr->c.nr_cbits();
}

extern "C" result_t circuit_cstate(const circuit_t* ptr)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    const std::vector<uint64_t>* cs = ptr->c.cstate();
    if (!cs) return res_error("Circuit has not been run yet");
    uint64_t* d = static_cast<uint64_t*>(std::malloc(std::max<size_t>(1, cs->size()) * sizeof(uint64_t)));
    std::memcpy(d, cs->data(), cs->size() * sizeof(uint64_t));
    return result_t{d, cs->size(), cs->size(), RESULT_CSTATE};
}

static std::vector<Param> to_params(const parameter_t* p, size_t n)
{
    std::vector<Param> v;
    if (!p) return v;
    for (size_t k = 0; k < n; ++k) v.push_back(Param{p[k].value, p[k].value_ptr});
    return v;
}

extern "C" result_t circuit_add_gate(circuit_t* ptr, const char* gate, const size_t* qbits, size_t nr_qbits,
                                     const parameter_t* params, size_t nr_params)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    if (!qbits) return res_error("Pointer to bit indices is NULL");
    if (!gate) return res_error("Invalid gate name");
    return guard_result([&] {
        std::vector<Param> pv = to_params(params, nr_params);
        ptr->c.add_gate(gate, qbits, nr_qbits, pv.data(), pv.size());
    });
}

extern "C" result_t circuit_add_conditional_gate(circuit_t* ptr, const size_t* control, size_t nr_control,
                                                 uint64_t target, const char* gate, const size_t* qbits,
                                                 size_t nr_qbits, const parameter_t* params, size_t nr_params)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    if (!control) return res_error("Pointer to control bit indices is NULL");
    if (!qbits) return res_error("Pointer to bit indices is NULL");
    if (!gate) return res_error("Invalid gate name");
    return guard_result([&] {
        std::vector<Param> pv = to_params(params, nr_params);
        ptr->c.add_conditional_gate(control, nr_control, target, gate, qbits, nr_qbits, pv.data(), pv.size());
    });
}

extern "C" result_t circuit_reset(circuit_t* ptr, size_t qbit)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    return guard_result([&] { ptr->c.reset(qbit); });
}

extern "C" result_t circuit_reset_all(circuit_t* ptr)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    return guard_result([&] { ptr->c.reset_all(); });
}

static bool parse_basis(char dir, char* out)
{
    switch (dir) {
    case 'x': case 'X': *out = 'X'; return true;
    case 'y': case 'Y': *out = 'Y'; return true;
    case 'z': case 'Z': *out = 'Z'; return true;
    default: return false;
    }
}

extern "C" result_t circuit_measure(circuit_t* ptr, size_t qbit, size_t cbit, char dir, uint8_t collapse)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    char b;
    if (!parse_basis(dir, &b)) return res_error(std::string("Invalid measurement basis '") + dir + "'");
    return guard_result([&] { ptr->c.measure_basis(qbit, cbit, b, collapse != 0); });
}

extern "C" result_t circuit_measure_all(circuit_t* ptr, const size_t* cbits, size_t nr_cbits, char dir,
                                        uint8_t collapse)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    if (!cbits) return res_error("Pointer to measurement bit indices is NULL");
    char b;
    if (!parse_basis(dir, &b)) return res_error(std::string("Invalid measurement basis '") + dir + "'");
    return guard_result([&] { ptr->c.measure_all_basis(cbits, nr_cbits, b, collapse != 0); });
}

extern "C" result_t circuit_execute(circuit_t* ptr, size_t nr_shots)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    return guard_result([&] { ptr->c.execute(nr_shots); });
}

extern "C" result_t circuit_reexecute(circuit_t* ptr)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    return guard_result([&] { ptr->c.reexecute(); });
}

extern "C" result_t circuit_histogram(const circuit_t* ptr)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    std::vector<std::pair<std::string, size_t>> h;
    int rc = guard([&] { h = ptr->c.histogram_string(); });
    if (rc) return res_error(g_last_error);
    histelem_t* e = static_cast<histelem_t*>(std::malloc(std::max<size_t>(1, h.size()) * sizeof(histelem_t)));
    for (size_t k = 0; k < h.size(); ++k) e[k] = histelem_t{dup_cstr(h[k].first), h[k].second};
    return result_t{e, h.size(), h.size(), RESULT_HISTOGRAM};
}

static result_t export_not_provided(const circuit_t* ptr, const char* what)
{
    if (!ptr) return res_error("Pointer to circuit is NULL");
    return res_error(std::string(what) + " export is not provided by the GPU statevector backend");
}

extern "C" result_t circuit_latex(const circuit_t* ptr) { return export_not_provided(ptr, "LaTeX"); }
extern "C" result_t circuit_open_qasm(const circuit_t* ptr) { return export_not_provided(ptr, "OpenQasm"); }
extern "C" result_t circuit_c_qasm(const circuit_t* ptr) { return export_not_provided(ptr, "cQasm"); }

// ------------------------------------------------- circuit extensions

#define NEED(p)                                                                                    \
    if (!(p)) {                                                                                    \
        g_last_error = "Null pointer argument";                                                    \
        return ERR_INVALID_ARGUMENT;                                                               \
    }

extern "C" int q1t_circuit_set_backend(circuit_t* ptr, int precision, int fusion_k, int tiling)
{
    NEED(ptr);
    return guard([&] {
        if (precision != PREC_C128 && precision != PREC_C64) throw Error(ERR_INVALID_ARGUMENT, "Unknown precision");
        if (fusion_k < 0 || fusion_k > 6) throw Error(ERR_INVALID_ARGUMENT, "fusion_k must be 0..6");
        ptr->c.set_backend(Precision(precision), fusion_k, tiling != 0);
    });
}

extern "C" int q1t_circuit_set_graph(circuit_t* ptr, int on)
{
    NEED(ptr);
    return guard([&] {
        if (on != 0 && on != 1) throw Error(ERR_INVALID_ARGUMENT, "graph mode must be 0 or 1");
        ptr->c.set_graph(on != 0);
    });
}

extern "C" int q1t_circuit_set_jit(circuit_t* ptr, int on)
{
    NEED(ptr);
    return guard([&] {
        if (on < 0 || on > 3) throw Error(ERR_INVALID_ARGUMENT, "jit mode must be 0..3");
        ptr->c.set_jit(on);
    });
}

// Plan a gate-only circuit on the host and compile the kernels specialised
// to its sweeps (no GPU needed); with out_dir, each code object and source
// is written there as sweep_<i>.co / sweep_<i>.hip.
extern "C" int q1t_circuit_jit_compile(const circuit_t* ptr, int precision, int fusion_k, int mode,
                                       const char* out_dir, size_t* nkernels, double* seconds)
{
    NEED(ptr);
    NEED(nkernels);
    NEED(seconds);
    return guard([&] {
        if (precision != PREC_C128 && precision != PREC_C64) throw Error(ERR_INVALID_ARGUMENT, "unknown precision");
        if (fusion_k < 0 || fusion_k > 6) throw Error(ERR_INVALID_ARGUMENT, "fusion_k must be 0..6");
        if (mode < 1 || mode > 3) throw Error(ERR_INVALID_ARGUMENT, "jit mode must be 1, 2 or 3");
        const size_t n = ptr->c.nr_qbits();
        std::vector<PendingGate> q;
        for (const CircuitOp& op : ptr->c.ops()) {
            if (op.type != OpType::Gate) throw Error(ERR_INVALID_ARGUMENT, "jit compile: the circuit holds non-gate ops");
            std::vector<double> pv;
            for (const Param& p : op.params) pv.push_back(p.get());
            const std::vector<cdouble> m = gate_matrix(*op.gate, pv.data());
            PendingGate pg;
            pg.op = analyze_gate(m.data(), op.bits.size(), op.bits.data(), n);
            pg.matrix = m;
            for (size_t b : op.bits) pg.qbits.push_back(int(n - 1 - b));
            q.push_back(std::move(pg));
        }
        const std::vector<std::string> srcs =
            plan_jit_sources(q, int(n), Precision(precision), fusion_k, true, mode == 2, mode != 3);
        double total = 0.0;
        for (size_t i = 0; i < srcs.size(); ++i) {
            double sec = 0.0;
            const std::vector<char> code = jit_compile(srcs[i], &sec);
            total += sec;
            if (out_dir) {
                const std::string base = std::string(out_dir) + "/sweep_" + std::to_string(i);
                FILE* f = std::fopen((base + ".co").c_str(), "wb");
                if (!f) throw Error(ERR_INVALID_ARGUMENT, "cannot write " + base + ".co");
                std::fwrite(code.data(), 1, code.size(), f);
                std::fclose(f);
                f = std::fopen((base + ".hip").c_str(), "wb");
                if (!f) throw Error(ERR_INVALID_ARGUMENT, "cannot write " + base + ".hip");
                std::fwrite(srcs[i].data(), 1, srcs[i].size(), f);
                std::fclose(f);
            }
        }
        *nkernels = srcs.size();
        *seconds = total;
    });
}

extern "C" int q1t_jit_stats(uint64_t* compiled, uint64_t* hits, double* compile_s)
{
    NEED(compiled);
    NEED(hits);
    NEED(compile_s);
    return guard([&] {
        const JitStats s = JitCache::get().stats();
        *compiled = s.compiled;
        *hits = s.hits;
        *compile_s = s.compile_s;
    });
}

extern "C" int q1t_circuit_graph_stats(circuit_t* ptr, uint64_t* captures, uint64_t* replays)
{
    NEED(ptr);
    NEED(captures);
    NEED(replays);
    GpuVectorState* st = ptr->c.state();
    *captures = st ? st->graph_captures() : 0;
    *replays = st ? st->graph_replays() : 0;
    return 0;
}

extern "C" int q1t_circuit_seed(circuit_t* ptr, uint64_t seed)
{
    NEED(ptr);
    ptr->c.seed(seed);
    return 0;
}

extern "C" int q1t_circuit_set_initial_amplitudes(circuit_t* ptr, const double* amps, size_t len)
{
    NEED(ptr);
    NEED(amps);
    return guard([&] {
        std::vector<cdouble> a(len);
        for (size_t i = 0; i < len; ++i) a[i] = cdouble(amps[2 * i], amps[2 * i + 1]);
        ptr->c.set_initial_amplitudes(std::move(a));
    });
}

extern "C" int q1t_circuit_set_initial_random(circuit_t* ptr, uint64_t seed)
{
    NEED(ptr);
    ptr->c.set_initial_random(seed);
    return 0;
}

extern "C" int q1t_circuit_barrier(circuit_t* ptr, const size_t* qbits, size_t n)
{
    NEED(ptr);
    NEED(qbits);
    return guard([&] { ptr->c.barrier(qbits, n); });
}

extern "C" int q1t_circuit_plan_order(const circuit_t* ptr, int precision, int search, size_t* out, size_t cap,
                                      size_t* count)
{
    NEED(ptr);
    NEED(count);
    return guard([&] {
        if (precision != PREC_C128 && precision != PREC_C64) throw Error(ERR_INVALID_ARGUMENT, "unknown precision");
        const size_t n = ptr->c.nr_qbits();
        std::vector<PendingGate> q;
        for (const CircuitOp& op : ptr->c.ops()) {
            if (op.type != OpType::Gate) throw Error(ERR_INVALID_ARGUMENT, "plan order: the circuit holds non-gate ops");
            std::vector<double> pv;
            for (const Param& p : op.params) pv.push_back(p.get());
            const std::vector<cdouble> m = gate_matrix(*op.gate, pv.data());
            PendingGate pg;
            pg.op = analyze_gate(m.data(), op.bits.size(), op.bits.data(), n);
            q.push_back(std::move(pg));
        }
        const std::vector<size_t> order = plan_order(q, int(n), tile_geometry(Precision(precision)), search != 0);
        *count = order.size();
        for (size_t i = 0; i < order.size() && i < cap && out; ++i) out[i] = order[i];
    });
}

extern "C" int q1t_circuit_nr_columns(circuit_t* ptr, size_t* out)
{
    NEED(ptr);
    NEED(out);
    return guard([&] {
        if (!ptr->c.state()) throw err_not_executed();
        *out = ptr->c.state()->nr_cols();
    });
}

extern "C" int q1t_circuit_counts(circuit_t* ptr, size_t* out, size_t len)
{
    NEED(ptr);
    NEED(out);
    return guard([&] {
        if (!ptr->c.state()) throw err_not_executed();
        const auto& c = ptr->c.state()->counts();
        if (len < c.size()) throw err_not_enough_space(len, c.size());
        std::copy(c.begin(), c.end(), out);
    });
}

extern "C" int q1t_circuit_read_amplitudes(circuit_t* ptr, size_t col, double* out, size_t len)
{
    NEED(ptr);
    NEED(out);
    return guard([&] {
        GpuVectorState* s = ptr->c.state();
        if (!s) throw err_not_executed();
        const size_t N = size_t(1) << s->nr_bits();
        if (len < N) throw err_not_enough_space(len, N);
        s->read_amplitudes(col, reinterpret_cast<cdouble*>(out));
    });
}

extern "C" int q1t_circuit_last_stats(const circuit_t* ptr, q1t_stats* out)
{
    NEED(ptr);
    NEED(out);
    const RunStats& r = ptr->c.last_stats();
    *out = q1t_stats{r.gate_seconds, r.bytes, r.launches, r.sweeps, r.gates, r.blocks, r.phases, r.mfma_flops};
    return 0;
}

extern "C" int q1t_circuit_histogram_vec(const circuit_t* ptr, size_t* out, size_t len)
{
    NEED(ptr);
    NEED(out);
    return guard([&] {
        std::vector<size_t> h = ptr->c.histogram_vec();
        if (len < h.size()) throw err_not_enough_space(len, h.size());
        std::copy(h.begin(), h.end(), out);
    });
}

// ------------------------------------------------- batches

extern "C" int q1t_batch_new(int precision, int fusion_k, size_t mem_budget, q1t_batch** out)
{
    NEED(out);
    *out = nullptr;
    return guard([&] {
        if (precision != PREC_C128 && precision != PREC_C64) throw Error(ERR_INVALID_ARGUMENT, "Unknown precision");
        if (fusion_k < 0 || fusion_k > 6) throw Error(ERR_INVALID_ARGUMENT, "fusion_k must be 0..6");
        *out = new q1t_batch(Precision(precision), fusion_k, mem_budget);
    });
}

extern "C" void q1t_batch_free(q1t_batch* b) { delete b; }

extern "C" int q1t_batch_add(q1t_batch* b, circuit_t* c)
{
    NEED(b);
    NEED(c);
    return guard([&] { b->b.add(&c->c); });
}

extern "C" int q1t_batch_execute(q1t_batch* b, size_t nr_shots, int keep_states)
{
    NEED(b);
    return guard([&] { b->b.execute(nr_shots, keep_states != 0); });
}

extern "C" int q1t_batch_stats_get(const q1t_batch* b, q1t_batch_stats* out)
{
    NEED(b);
    NEED(out);
    const BatchStats& s = b->b.stats();
    const RunStats& r = s.gates;
    out->gates = q1t_stats{r.gate_seconds, r.bytes, r.launches, r.sweeps, r.gates, r.blocks, r.phases, r.mfma_flops};
    out->waves = s.waves;
    out->rounds = s.rounds;
    out->sync_ops = s.sync_ops;
    out->seconds = s.seconds;
    out->start_s = s.start_s;
    out->queue_s = s.queue_s;
    out->flush_s = s.flush_s;
    out->measure_s = s.measure_s;
    out->other_s = s.other_s;
    out->column_launches = s.column_launches;
    out->solo_ops = s.solo_ops;
    out->plan_s = s.plan_s;
    out->pack_s = s.pack_s;
    return 0;
}

extern "C" int q1t_batch_autotune(q1t_batch* b, size_t nr_shots, size_t sample, q1t_batch_config* out, size_t cap,
                                  size_t* count)
{
    NEED(b);
    NEED(count);
    return guard([&] {
        std::vector<BatchConfig> v = b->b.autotune(nr_shots, sample);
        *count = v.size();
        for (size_t i = 0; i < v.size() && i < cap && out; ++i)
            out[i] = q1t_batch_config{v[i].geom.M, v[i].geom.R, v[i].geom.L, v[i].fusion_k, v[i].seconds};
    });
}

extern "C" int q1t_batch_set_config(q1t_batch* b, const q1t_batch_config* cfg)
{
    NEED(b);
    NEED(cfg);
    return guard([&] { b->b.set_config(TileGeom{cfg->M, cfg->R, cfg->L}, cfg->fusion_k); });
}

extern "C" int q1t_batch_set_tile_chunk(q1t_batch* b, unsigned chunk)
{
    NEED(b);
    return guard([&] { b->b.set_tile_chunk(chunk); });
}

extern "C" int q1t_batch_get_config(const q1t_batch* b, q1t_batch_config* out)
{
    NEED(b);
    NEED(out);
    const BatchConfig c = b->b.config();
    *out = q1t_batch_config{c.geom.M, c.geom.R, c.geom.L, c.fusion_k, 0.0};
    return 0;
}

// ------------------------------------------------- state-level ABI

static std::vector<cdouble> to_cd(const double* m, size_t n)
{
    std::vector<cdouble> v(n);
    for (size_t k = 0; k < n; ++k) v[k] = cdouble(m[2 * k], m[2 * k + 1]);
    return v;
}

extern "C" int q1t_gvs_new(size_t nr_bits, size_t nr_shots, int precision, q1t_gvs** out)
{
    NEED(out);
    *out = nullptr;
    return guard([&] {
        if (precision != PREC_C128 && precision != PREC_C64) throw Error(ERR_INVALID_ARGUMENT, "Unknown precision");
        *out = new q1t_gvs(nr_bits, nr_shots, Precision(precision));
    });
}

extern "C" void q1t_gvs_free(q1t_gvs* s) { delete s; }

extern "C" int q1t_gvs_from_qubit_coefs(q1t_gvs* s, const double* coefs, size_t ncoefs)
{
    NEED(s);
    NEED(coefs);
    return guard([&] {
        if (ncoefs != 2 * s->s.nr_bits()) throw Error(ERR_INVALID_ARGUMENT, "Expected 2*nr_bits coefficients");
        std::vector<cdouble> c = to_cd(coefs, ncoefs);
        s->s.set_product_state(c.data());
    });
}

extern "C" int q1t_gvs_set_amplitudes(q1t_gvs* s, const double* amps, size_t len)
{
    NEED(s);
    NEED(amps);
    return guard([&] {
        if (len != (size_t(1) << s->s.nr_bits())) throw Error(ERR_INVALID_ARGUMENT, "Expected 2^nr_bits amplitudes");
        s->s.set_amplitudes(reinterpret_cast<const cdouble*>(amps));
    });
}

extern "C" int q1t_gvs_set_random_normal(q1t_gvs* s, uint64_t seed)
{
    NEED(s);
    return guard([&] { s->s.set_random_normal(seed); });
}

extern "C" size_t q1t_gvs_nr_bits(const q1t_gvs* s) { return s ? s->s.nr_bits() : 0; }
extern "C" size_t q1t_gvs_nr_cols(const q1t_gvs* s) { return s ? s->s.nr_cols() : 0; }

extern "C" int q1t_gvs_counts(const q1t_gvs* s, size_t* out, size_t len)
{
    NEED(s);
    NEED(out);
    return guard([&] {
        const auto& c = s->s.counts();
        if (len < c.size()) throw err_not_enough_space(len, c.size());
        std::copy(c.begin(), c.end(), out);
    });
}

extern "C" int q1t_gvs_apply_gate(q1t_gvs* s, const double* mat, size_t nr_affected, const size_t* bits, size_t nbits,
                                  const char* desc)
{
    NEED(s);
    NEED(mat);
    NEED(bits);
    return guard([&] {
        if (nr_affected > 12) throw Error(ERR_NOT_SUPPORTED, "Gates on more than 12 qubits are not supported");
        const size_t d = size_t(1) << nr_affected;
        std::vector<cdouble> m = to_cd(mat, d * d);
        s->s.apply_gate(m.data(), nr_affected, bits, nbits, desc ? desc : "gate");
    });
}

extern "C" int q1t_gvs_apply_unary_gate_all(q1t_gvs* s, const double* mat2, const char* desc)
{
    NEED(s);
    NEED(mat2);
    return guard([&] {
        std::vector<cdouble> m = to_cd(mat2, 4);
        s->s.apply_unary_gate_all(m.data(), desc ? desc : "gate");
    });
}

extern "C" int q1t_gvs_apply_conditional_gate(q1t_gvs* s, const uint8_t* control, size_t ncontrol,
                                              const double* mat, size_t nr_affected, const size_t* bits,
                                              size_t nbits, const char* desc)
{
    NEED(s);
    NEED(control);
    NEED(mat);
    NEED(bits);
    return guard([&] {
        if (nr_affected > 12) throw Error(ERR_NOT_SUPPORTED, "Gates on more than 12 qubits are not supported");
        const size_t d = size_t(1) << nr_affected;
        std::vector<cdouble> m = to_cd(mat, d * d);
        s->s.apply_conditional_gate(control, ncontrol, m.data(), nr_affected, bits, nbits, desc ? desc : "gate");
    });
}

extern "C" int q1t_gvs_marginal_w0(q1t_gvs* s, size_t qbit, double* w0s, size_t len)
{
    NEED(s);
    NEED(w0s);
    return guard([&] {
        std::vector<double> w = s->s.marginal_w0(qbit);
        if (len < w.size()) throw err_not_enough_space(len, w.size());
        std::copy(w.begin(), w.end(), w0s);
    });
}

extern "C" int q1t_gvs_marginal_w0_noalloc(q1t_gvs* s, size_t qbit, double* w0s, size_t len)
{
    NEED(s);
    NEED(w0s);
    return guard([&] { s->s.marginal_w0_noalloc(qbit, w0s, len); });
}

extern "C" int q1t_gvs_measure_with_n0(q1t_gvs* s, size_t qbit, size_t cbit, const size_t* n0s, const double* w0s,
                                       size_t ncols, uint64_t* res, size_t res_len, int collapse)
{
    NEED(s);
    NEED(n0s);
    NEED(w0s);
    NEED(res);
    return guard([&] {
        if (ncols != s->s.nr_cols()) throw Error(ERR_INVALID_ARGUMENT, "n0s/w0s must have one entry per column");
        std::vector<double> w(w0s, w0s + ncols);
        s->s.measure_with_n0(qbit, cbit, n0s, w, res, res_len, collapse != 0);
    });
}

static RngSource rng_of(const q1t_rng* r)
{
    if (!r || !r->next_u64) throw Error(ERR_INVALID_ARGUMENT, "A random source is required");
    return RngSource{r->ctx, r->next_u64};
}

extern "C" int q1t_gvs_measure_into(q1t_gvs* s, size_t qbit, size_t cbit, uint64_t* res, size_t res_len,
                                    const q1t_rng* rng)
{
    NEED(s);
    NEED(res);
    return guard([&] {
        RngSource r = rng_of(rng);
        s->s.measure_into(qbit, cbit, res, res_len, r, true);
    });
}

extern "C" int q1t_gvs_measure_basis_into(q1t_gvs* s, size_t qbit, size_t cbit, char basis, uint64_t* res,
                                          size_t res_len, const q1t_rng* rng, int collapse)
{
    NEED(s);
    NEED(res);
    return guard([&] {
        const int b = basis == 'Z' || basis == 'z' ? 0 : basis == 'X' || basis == 'x' ? 1 : basis == 'Y' || basis == 'y' ? 2 : -1;
        if (b < 0) throw Error(ERR_INVALID_ARGUMENT, std::string("Unknown measurement basis '") + basis + "'");
        RngSource r = rng_of(rng);
        s->s.measure_into(qbit, cbit, res, res_len, r, collapse != 0, b);
    });
}

extern "C" int q1t_gvs_peek_into(q1t_gvs* s, size_t qbit, size_t cbit, uint64_t* res, size_t res_len,
                                 const q1t_rng* rng)
{
    NEED(s);
    NEED(res);
    return guard([&] {
        RngSource r = rng_of(rng);
        s->s.measure_into(qbit, cbit, res, res_len, r, false);
    });
}

extern "C" int q1t_gvs_measure_all_into(q1t_gvs* s, const size_t* cbits, size_t ncbits, uint64_t* res,
                                        size_t res_len, const q1t_rng* rng)
{
    NEED(s);
    NEED(cbits);
    NEED(res);
    return guard([&] {
        RngSource r = rng_of(rng);
        s->s.measure_all_into(cbits, ncbits, res, res_len, r, true);
    });
}

extern "C" int q1t_gvs_peek_all_into(q1t_gvs* s, const size_t* cbits, size_t ncbits, uint64_t* res, size_t res_len,
                                     const q1t_rng* rng)
{
    NEED(s);
    NEED(cbits);
    NEED(res);
    return guard([&] {
        RngSource r = rng_of(rng);
        s->s.measure_all_into(cbits, ncbits, res, res_len, r, false);
    });
}

extern "C" int q1t_gvs_measure_all_uniforms(q1t_gvs* s, const size_t* cbits, size_t ncbits, const double* u01,
                                            uint64_t* res, size_t res_len, int collapse)
{
    NEED(s);
    NEED(cbits);
    NEED(u01);
    NEED(res);
    return guard([&] { s->s.measure_all_uniforms(cbits, ncbits, u01, res, res_len, collapse != 0); });
}

extern "C" int q1t_gvs_reset(q1t_gvs* s, size_t bit, const q1t_rng* rng)
{
    NEED(s);
    return guard([&] {
        RngSource r = rng_of(rng);
        s->s.reset(bit, r);
    });
}

extern "C" int q1t_gvs_reset_all(q1t_gvs* s)
{
    NEED(s);
    return guard([&] { s->s.reset_all(); });
}

extern "C" int q1t_gvs_read_amplitudes(q1t_gvs* s, size_t col, double* out, size_t len)
{
    NEED(s);
    NEED(out);
    return guard([&] {
        const size_t N = size_t(1) << s->s.nr_bits();
        if (len < N) throw err_not_enough_space(len, N);
        s->s.read_amplitudes(col, reinterpret_cast<cdouble*>(out));
    });
}

extern "C" int q1t_gvs_flush(q1t_gvs* s)
{
    NEED(s);
    return guard([&] { s->s.flush(); });
}

extern "C" int q1t_gvs_set_plan_search(q1t_gvs* s, int mode)
{
    NEED(s);
    return guard([&] {
        if (mode < -1 || mode > 1) throw Error(ERR_INVALID_ARGUMENT, "plan search mode must be -1, 0 or 1");
        s->s.set_plan_search(mode);
    });
}

extern "C" int q1t_gvs_set_jit(q1t_gvs* s, int on)
{
    NEED(s);
    return guard([&] {
        if (on < 0 || on > 3) throw Error(ERR_INVALID_ARGUMENT, "jit mode must be 0..3");
        s->s.set_jit(on);
    });
}

extern "C" int q1t_gvs_set_graph(q1t_gvs* s, int on)
{
    NEED(s);
    return guard([&] {
        if (on != 0 && on != 1) throw Error(ERR_INVALID_ARGUMENT, "graph mode must be 0 or 1");
        s->s.set_graph(on != 0);
    });
}

extern "C" int q1t_gvs_graph_stats(q1t_gvs* s, uint64_t* captures, uint64_t* replays)
{
    NEED(s);
    NEED(captures);
    NEED(replays);
    *captures = s->s.graph_captures();
    *replays = s->s.graph_replays();
    return 0;
}

extern "C" int q1t_gvs_set_plan(q1t_gvs* s, int fusion_k, int tiling)
{
    NEED(s);
    return guard([&] {
        if (fusion_k < 0 || fusion_k > 6) throw Error(ERR_INVALID_ARGUMENT, "fusion_k must be 0..6");
        s->s.set_fusion(fusion_k);
        s->s.set_tiling(tiling != 0);
    });
}

extern "C" int q1t_gvs_set_geometry(q1t_gvs* s, int M, int R, int L)
{
    NEED(s);
    return guard([&] { s->s.set_geometry(TileGeom{M, R, L}); });
}

extern "C" int q1t_gvs_last_stats(const q1t_gvs* s, q1t_stats* out)
{
    NEED(s);
    NEED(out);
    const RunStats& r = s->s.last_stats();
    *out = q1t_stats{r.gate_seconds, r.bytes, r.launches, r.sweeps, r.gates, r.blocks, r.phases, r.mfma_flops};
    return 0;
}

extern "C" q1t_xoshiro* q1t_xoshiro_new(uint64_t seed)
{
    q1t_xoshiro* x = new q1t_xoshiro;
    x->x.reseed(seed);
    return x;
}

extern "C" void q1t_xoshiro_free(q1t_xoshiro* x) { delete x; }

extern "C" uint64_t q1t_xoshiro_next(void* ctx) { return static_cast<q1t_xoshiro*>(ctx)->x.next(); }

extern "C" int q1t_gvs_set_profiling(q1t_gvs* s, int on)
{
    NEED(s);
    return guard([&] { s->s.set_profiling(on != 0); });
}

extern "C" int q1t_device_cache_limit(size_t bytes)
{
    return guard([&] { DeviceCache::get().set_cap(bytes); });
}

extern "C" int q1t_device_cache_bytes(size_t* out)
{
    NEED(out);
    return guard([&] { *out = DeviceCache::get().cached_bytes(); });
}

extern "C" int q1t_copy_bandwidth(size_t bytes, int reps, double* best_ms)
{
    NEED(best_ms);
    return guard([&] {
        if (bytes == 0 || reps < 1) throw Error(ERR_INVALID_ARGUMENT, "bytes and reps must be positive");
        *best_ms = measure_copy_ms(bytes, reps);
    });
}

extern "C" int q1t_gvs_kernel_profile(const q1t_gvs* s, q1t_kernel_profile* out, size_t cap, size_t* count)
{
    NEED(s);
    NEED(count);
    return guard([&] {
        std::vector<KernelProfile> p = s->s.profile();
        *count = p.size();
        for (size_t k = 0; k < p.size() && k < cap && out; ++k) {
            std::memset(out[k].name, 0, sizeof(out[k].name));
            std::strncpy(out[k].name, p[k].name.c_str(), sizeof(out[k].name) - 1);
            out[k].launches = p[k].launches;
            out[k].total_ms = p[k].ms;
            out[k].bytes = p[k].bytes;
        }
    });
}

extern "C" int q1t_gvs_compare(q1t_gvs* a, size_t col_a, q1t_gvs* b, size_t col_b, double* out5)
{
    NEED(a);
    NEED(b);
    NEED(out5);
    return guard([&] {
        StateComparison c = GpuVectorState::compare(a->s, col_a, b->s, col_b);
        out5[0] = c.overlap_re;
        out5[1] = c.overlap_im;
        out5[2] = c.max_abs;
        out5[3] = c.norm_a;
        out5[4] = c.norm_b;
    });
}

extern "C" int q1t_gate_matrix(const char* name, const double* params, size_t nparams, double* out, size_t cap,
                               size_t* nbits)
{
    NEED(name);
    NEED(out);
    NEED(nbits);
    return guard([&] {
        const GateSpec* g = find_gate(name);
        if (!g) throw err_unknown_gate(name);
        if (size_t(g->nparams) != nparams) throw err_invalid_nr_arguments(nparams, size_t(g->nparams), g->desc);
        if (nparams && !params) throw Error(ERR_INVALID_ARGUMENT, "Null parameter array");
        std::vector<cdouble> m = gate_matrix(*g, params);
        if (cap < 2 * m.size()) throw err_not_enough_space(cap, 2 * m.size());
        for (size_t k = 0; k < m.size(); ++k) {
            out[2 * k] = m[k].real();
            out[2 * k + 1] = m[k].imag();
        }
        *nbits = gate_nbits(*g);
    });
}

// ------------------------------------------------- learned cost model

struct q1t_cost_model {
    CostModel m;
};

static_assert(kCmParams == Q1T_CM_PARAMS, "parameter count of the C ABI");
static_assert(kCmFeat == Q1T_CM_FEATURES && kSweepFeatures == Q1T_CM_FEATURES, "feature count of the C ABI");
static_assert(sizeof(SweepSample::feat) == kSweepFeatures * sizeof(float), "recorded sweep features");

extern "C" int q1t_cm_new(uint64_t seed, q1t_cost_model** out)
{
    NEED(out);
    *out = nullptr;
    return guard([&] {
        q1t_cost_model* cm = new q1t_cost_model();
        cm->m.init(seed);
        *out = cm;
    });
}

extern "C" void q1t_cm_free(q1t_cost_model* cm) { delete cm; }

extern "C" int q1t_cm_get_params(const q1t_cost_model* cm, float* out, size_t len)
{
    NEED(cm);
    NEED(out);
    return guard([&] {
        if (len < size_t(kCmParams)) throw err_not_enough_space(len, kCmParams);
        cm->m.get_params(out);
    });
}

extern "C" int q1t_cm_set_params(q1t_cost_model* cm, const float* p, size_t len)
{
    NEED(cm);
    NEED(p);
    return guard([&] {
        if (len != size_t(kCmParams)) throw Error(ERR_INVALID_ARGUMENT, "Expected " + std::to_string(kCmParams) + " parameters");
        cm->m.set_params(p);
    });
}

extern "C" int q1t_cm_set_norm(q1t_cost_model* cm, const float* mu, const float* sigma, size_t nfeat)
{
    NEED(cm);
    NEED(mu);
    NEED(sigma);
    return guard([&] {
        if (nfeat != size_t(kCmFeat)) throw Error(ERR_INVALID_ARGUMENT, "Expected 16 features");
        cm->m.set_norm(mu, sigma);
    });
}

extern "C" int q1t_cm_get_norm(const q1t_cost_model* cm, float* mu, float* sigma, size_t nfeat)
{
    NEED(cm);
    NEED(mu);
    NEED(sigma);
    return guard([&] {
        if (nfeat < size_t(kCmFeat)) throw err_not_enough_space(nfeat, kCmFeat);
        cm->m.get_norm(mu, sigma);
    });
}

extern "C" int q1t_cm_predict(q1t_cost_model* cm, const float* x, size_t nrows, float* y)
{
    NEED(cm);
    NEED(x);
    NEED(y);
    return guard([&] { cm->m.predict(x, nrows, y); });
}

extern "C" int q1t_cm_loss_grad(q1t_cost_model* cm, const float* x, const float* t, size_t nrows, float* grad,
                                size_t len, double* loss)
{
    NEED(cm);
    NEED(x);
    NEED(t);
    NEED(grad);
    NEED(loss);
    return guard([&] {
        if (len < size_t(kCmParams)) throw err_not_enough_space(len, kCmParams);
        *loss = cm->m.loss_and_grad(x, t, nrows, grad);
    });
}

extern "C" int q1t_cm_adam_step(q1t_cost_model* cm, const float* x, const float* t, size_t nrows, float lr,
                                double* loss)
{
    NEED(cm);
    NEED(x);
    NEED(t);
    NEED(loss);
    return guard([&] {
        AdamConfig c;
        c.lr = lr;
        *loss = cm->m.adam_step(x, t, nrows, c);
    });
}

extern "C" int q1t_cm_train(q1t_cost_model* cm, const float* x, const float* t, size_t nrows, int epochs,
                            size_t batch, float lr, uint64_t seed, double* loss)
{
    NEED(cm);
    NEED(x);
    NEED(t);
    NEED(loss);
    return guard([&] {
        if (epochs < 0) throw Error(ERR_INVALID_ARGUMENT, "epochs must be >= 0");
        AdamConfig c;
        c.lr = lr;
        *loss = cm->m.train(x, t, nrows, epochs, batch, c, seed);
    });
}

extern "C" int q1t_cm_set_trained(q1t_cost_model* cm, int trained)
{
    NEED(cm);
    cm->m.mark_trained(trained != 0);
    return 0;
}

extern "C" int q1t_gvs_set_cost_model(q1t_gvs* s, q1t_cost_model* cm, int mode, uint64_t seed)
{
    NEED(s);
    return guard([&] {
        if (mode < PLAN_GREEDY || mode > PLAN_EXPLORE) throw Error(ERR_INVALID_ARGUMENT, "Unknown planner mode");
        s->s.set_cost_model(cm ? &cm->m : nullptr, mode, seed);
    });
}

extern "C" int q1t_gvs_set_recording(q1t_gvs* s, int on)
{
    NEED(s);
    return guard([&] { s->s.set_recording(on != 0); });
}

extern "C" int q1t_gvs_take_samples(q1t_gvs* s, float* feat, double* tiles, double* ms, size_t cap, size_t* count)
{
    NEED(s);
    NEED(feat);
    NEED(tiles);
    NEED(ms);
    NEED(count);
    return guard([&] {
        std::vector<SweepSample> v = s->s.take_samples(cap);
        const size_t n = v.size();
        for (size_t i = 0; i < n; ++i) {
            std::memcpy(feat + i * kCmFeat, v[i].feat, kCmFeat * sizeof(float));
            tiles[i] = v[i].tiles;
            ms[i] = v[i].ms;
        }
        *count = n;
    });
}
