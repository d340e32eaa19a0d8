// GpuVectorState: the MI355X statevector behind q1tsim's QuState interface
// (src/qustate.rs:5-89), replacing VectorState (src/vectorstate.rs).
#pragma once

#include <hip/hip_runtime.h>

#include <deque>
#include <map>
#include <memory>
#include <mutex>

#include "common.h"
#include "gates.h"
#include "items.h"

namespace q1t {

enum Precision { PREC_C128 = 0, PREC_C64 = 1 };

// Per-state counters of the work the last flush did (for bandwidth reports).
struct RunStats {
    double gate_seconds = 0.0;    // wall time of the last flush (host clock)
    double bytes = 0.0;           // algorithmic HBM bytes of the launched sweeps
    uint64_t launches = 0;        // kernels launched
    uint64_t sweeps = 0;          // full-state passes
    uint64_t gates = 0;           // gates applied
    uint64_t blocks = 0;          // fused dense blocks applied per column
    uint64_t phases = 0;          // in-tile phases over all sweeps
    double mfma_flops = 0.0;      // FP32 MFMA flops of the fused blocks: 2^(n+k+3) per k-qubit block and column
};

// Tile geometry of the sweeps: M local bits per tile, R of them in registers
// per phase (2^(M-R) threads), L lowest index bits always in the tile (runs
// of 2^L contiguous amplitudes per HBM access).
struct TileGeom {
    int M, R, L;
};

// Kernel time measured with HIP events on the state's stream.
struct KernelProfile {
    std::string name;
    uint64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;   // algorithmic HBM bytes of those launches
};

// <a|b>, max |a_i - b_i| and the two norms, accumulated in FP64.
struct StateComparison {
    double overlap_re = 0, overlap_im = 0, max_abs = 0, norm_a = 0, norm_b = 0;
};

class CostModel;

// How the fused complex64 planner picks among equivalent sweep programs.
enum PlanMode {
    PLAN_GREEDY = 0,    // fixed-width greedy fusion (blocks of the configured width)
    PLAN_LEARNED = 1,   // candidate programs ranked by the learned cost model
    PLAN_EXPLORE = 2    // seeded random choice among the candidates (data collection)
};

// One recorded sweep: program features, tiles processed, measured time.
struct SweepSample {
    float feat[20];       // sweep_features()
    double tiles = 0.0;   // tiles over all columns of the launch
    int slot = -1;        // profiling slot of the launch until its time is read
    double ms = 0.0;
};

struct PendingGate {
    GateOp op;
    std::vector<cdouble> matrix;  // full FP64 matrix (fused path composes these)
    std::vector<int> qbits;       // index-bit positions of the listed qubits
};

// Process-wide cache of freed device blocks.  Requests are rounded up to a
// size class -- a power of two up to 1 GiB, a multiple of 256 MiB above --
// and blocks are reused by class (the states of a batch come and go with
// many shapes, 2^n amplitudes times a few columns, but few classes).  A
// released block carries an event recorded on its stream: the same stream
// reuses it at once (stream order), another stream only once the event has
// completed.  Cached bytes are capped (default: 7/8 of device memory; a
// batch's states come and go by the hundreds of GB): a release over the cap
// frees the oldest cached blocks first.  On allocation failure the cache is
// emptied and the allocation retried.
class DeviceCache {
public:
    static DeviceCache& get();
    // bytes a request of `bytes` takes
    static size_t size_class(size_t bytes);
    void* alloc(size_t bytes, hipStream_t s);
    // `bytes` as passed to alloc
    void release(void* p, size_t bytes, hipStream_t s);
    void trim();
    // `s` has been synchronised and is about to be destroyed
    void forget_stream(hipStream_t s);
    // cap on cached bytes (0: the default), evicting down to it now
    void set_cap(size_t bytes);
    size_t cached_bytes();

private:
    static size_t default_cap();
    void evict_to(size_t limit);   // caller holds mu_
    // frees cached blocks until `need` bytes are freed (other size classes
    // than `cls` first, oldest first); returns the bytes freed
    size_t evict_for(size_t need, size_t cls);
    struct Entry {
        void* p;
        size_t bytes;     // the size class
        hipStream_t s;    // null with ev: idle, reusable on any stream
        hipEvent_t ev;
        uint64_t seq;     // release order
    };
    std::map<size_t, std::deque<Entry>> free_;   // by size class, oldest first
    uint64_t seq_ = 0;
    size_t cached_ = 0, cap_ = 0;
    std::mutex mu_;
};

// A HIP stream (default priority, non-blocking) shared by the states of a
// batch; destroyed with its last user.
struct SharedStream {
    hipStream_t s = nullptr;
    SharedStream();
    ~SharedStream();
    SharedStream(const SharedStream&) = delete;
    SharedStream& operator=(const SharedStream&) = delete;
};

// Best-of-`reps` time (ms) of a device-to-device copy of `bytes`.
double measure_copy_ms(size_t bytes, int reps);

class GpuVectorState {
public:
    // `shared`: run on a stream shared with other states (batches) instead
    // of a stream of its own; `compact_zero`: start as the compact basis
    // column |0..0> (set_zero_state_compact).
    GpuVectorState(size_t nr_bits, size_t nr_shots, Precision prec, std::shared_ptr<SharedStream> shared = nullptr,
                   bool compact_zero = false);
    ~GpuVectorState();
    GpuVectorState(const GpuVectorState&) = delete;
    GpuVectorState& operator=(const GpuVectorState&) = delete;

    size_t nr_bits() const { return nr_bits_; }
    size_t nr_shots() const { return nr_shots_; }
    size_t nr_cols() const { return counts_.size(); }
    const std::vector<size_t>& counts() const { return counts_; }
    Precision precision() const { return prec_; }
    bool is_basis() const { return basis_; }

    // initial states
    void set_zero_state();                                   // VectorState::new
    void set_product_state(const cdouble* coefs);            // from_qubit_coefs
    void set_amplitudes(const cdouble* amps);                // build-added upload
    void set_random_normal(uint64_t seed);                   // N(0,1) re/im, normalised

    // QuState
    void apply_gate(const cdouble* mat, size_t nr_affected, const size_t* qbits, size_t nbits,
                    const std::string& desc);
    void apply_unary_gate_all(const cdouble* mat2, const std::string& desc);
    void apply_conditional_gate(const uint8_t* control, size_t ncontrol, const cdouble* mat,
                                size_t nr_affected, const size_t* qbits, size_t nbits, const std::string& desc);
    // basis: 0 = Z, 1 = X, 2 = Y (the X/Y-basis change of circuit.rs:669-748
    // folded into the marginal and collapse passes)
    std::vector<double> marginal_w0(size_t qbit, int basis = 0);
    // marginal_w0 without any device allocation: partial sums in a fixed
    // file-scope device array (at most kNoAllocMaxCols columns), summed on
    // the device in the host's order (bit-identical results).  Needs no
    // pending gates.  Calls that overlap on one GPU race on that array.
    void marginal_w0_noalloc(size_t qbit, double* w0s, size_t len);
    void measure_with_n0(size_t qbit, size_t cbit, const size_t* n0s, const std::vector<double>& w0s,
                         uint64_t* res, size_t res_len, bool collapse, int basis = 0);
    void measure_into(size_t qbit, size_t cbit, uint64_t* res, size_t res_len, RngSource& rng, bool collapse,
                      int basis = 0);
    // rot: the basis a collapsing measure_all was taken in (0 = Z, 1 = X,
    // 2 = Y; the caller rotated the state before): the collapsed columns are
    // kept as basis vectors rotated back by H^n (X) or S^n H^n (Y), written
    // out in closed form only when dense amplitudes are needed
    void measure_all_uniforms(const size_t* cbits, size_t ncbits, const double* u01, uint64_t* res,
                              size_t res_len, bool collapse, int rot = 0);
    void measure_all_into(const size_t* cbits, size_t ncbits, uint64_t* res, size_t res_len, RngSource& rng,
                          bool collapse, int rot = 0);
    void reset(size_t bit, RngSource& rng);
    void reset_all();

    // Column work as items of an ItemQueue (batches queue many states' work
    // and launch it together; the operations above run a queue of their own
    // with the same kernels).  After the queue's launch() and a stream
    // synchronisation the results are read back.
    //  - marginal: the outcome-0 weight of qbit per column; returns the first
    //    item (one per column, in order), SIZE_MAX when the state knows the
    //    weights without the device (compact basis columns)
    size_t marginal_enqueue(ItemQueue& q, size_t qbit, int basis);
    std::vector<double> marginal_collect(const ItemQueue& q, size_t first, size_t qbit, int basis) const;
    //  - measure_with_n0 with its collapses (and the release of replaced
    //    columns) queued
    void collapse_enqueue(ItemQueue& q, size_t qbit, size_t cbit, const size_t* n0s, const std::vector<double>& w0s,
                          uint64_t* res, size_t res_len, bool collapse, int basis);
    //  - measure_all: the tree levels and walks of nr_shots runs with
    //    uniforms u; SIZE_MAX for compact basis columns (nothing queued)
    size_t sample_enqueue(ItemQueue& q, const double* u);
    void sample_finish(const ItemQueue& q, size_t job, const size_t* cbits, size_t ncbits, uint64_t* res,
                       size_t res_len, bool collapse, int rot);
    //  - apply_conditional_gate; false (nothing done) when the gate needs the
    //    per-gate kernels or the column set changes
    bool conditional_enqueue(ItemQueue& q, const uint8_t* control, size_t ncontrol, const cdouble* mat,
                             size_t nr_affected, const size_t* qbits, size_t nbits, const std::string& desc);
    //  - compact basis columns written out as dense columns
    void materialize_enqueue(ItemQueue& q);
    // the state starts as the compact basis column |0..0> (no device memory
    // until gates need it)
    void set_zero_state_compact();

    // build-added
    void read_amplitudes(size_t col, cdouble* out);
    void flush();
    void synchronize();
    // batched flushes (many states, one launch per sweep round): the queue is
    // taken over by the batch runner, which applies it and reports stats
    // here; compact columns are written out through `q` first
    std::vector<PendingGate>& pending() { return queue_; }
    void prepare_batched_flush(ItemQueue& q);
    void finish_batched_flush(const RunStats& s);
    const RunStats& last_stats() const { return stats_; }
    void set_fusion(int max_block_bits) { fusion_k_ = max_block_bits; }
    int fusion() const { return fusion_k_; }
    void set_tiling(bool on) { tiling_ = on; }
    void set_geometry(const TileGeom& g);
    const TileGeom& geometry() const { return geom_; }
    hipStream_t stream() const { return stream_; }
    void* device_amplitudes() { return amp_; }
    size_t elem_bytes() const { return prec_ == PREC_C128 ? 16 : 8; }
    void* col_ptr(size_t c) const;
    // Device scratch buffers kept between calls, one per use; growing one
    // synchronises the stream first.
    enum ScratchSlot { SLOT_MATRIX = 0, SLOT_PROGRAM = 1, SLOT_FUSED = 2, SLOT_COUNT = 3 };
    void* scratch(int slot, size_t bytes);
    void* scratch_matrix(size_t bytes) { return scratch(SLOT_MATRIX, bytes); }
    void apply_op(const GateOp& op, size_t col_begin, size_t col_end);

    // planner: learned cost model (not owned) and mode; recording of sweeps
    void set_cost_model(CostModel* cm, int mode, uint64_t seed)
    {
        cost_model_ = cm;
        plan_mode_ = mode;
        plan_rng_.reseed(seed);
    }
    CostModel* cost_model() const { return cost_model_; }
    int plan_mode() const { return plan_mode_; }
    // local search of tiles and register phases: -1 by state size, 0 off, 1 on
    void set_plan_search(int mode) { plan_search_ = mode; }
    bool plan_search(double state_bytes) const
    {
        return plan_search_ < 0 ? state_bytes >= 512.0 * 1024.0 * 1024.0 : plan_search_ != 0;
    }
    Xoshiro256& plan_rng() { return plan_rng_; }
    void set_recording(bool on);
    bool recording() const { return recording_; }
    void add_sample(const float* feat, double tiles, int slot);
    // up to `max` finished samples, oldest first (the rest stay)
    std::vector<SweepSample> take_samples(size_t max);
    size_t sample_count() const { return samples_done_; }

    // Flush graphs (off by default): the launches of a planned flush are
    // captured into one hipGraph, which later flushes replay while their
    // launch signature (tiles, program sizes, single-gate kernels and their
    // arguments, columns, buffer addresses) is unchanged.  Sweep programs,
    // which hold the gate matrices, live in the graph's own device table and
    // are uploaded before every replay, so changed gate parameters reach a
    // replay.  Not used while profiling or recording (events per launch), or
    // for a flush with a single-gate kernel that uploads its matrix.  A state
    // keeps up to kMaxFlushGraphs graphs, one per distinct signature (the
    // flushes between the measurements of a circuit differ), and evicts the
    // least recently used one.
    struct FlushGraph {
        std::vector<uint64_t> sig;
        hipGraphExec_t exec = nullptr;
        void* table = nullptr;
        size_t cap = 0;
        RunStats delta;                  // what one replay adds to the flush's stats
        uint64_t last_use = 0;
    };
    static constexpr size_t kMaxFlushGraphs = 8;
    // Sweep kernels specialised at run time to each sweep program (off by
    // default): compiled with hipRTC on the first flush that needs them,
    // cached per process by program structure.
    // With specialised kernels and no geometry set explicitly, the tile
    // shape is the one measured best for them (tile_geometry_specialised).
    // mode 0: interpreter kernels; 1: specialised kernels; 2: specialised
    // kernels with the next tile prefetched into registers.
    void set_jit(int mode);
    bool jit_enabled() const { return jit_mode_ != 0; }
    bool jit_prefetch() const { return jit_mode_ == 2; }
    // diagonal ops on bits outside the tile folded into one factor per tile
    // (specialised kernels, modes 1 and 2)
    bool tile_factors() const { return jit_mode_ == 1 || jit_mode_ == 2; }

    void set_graph(bool on);
    bool graph_enabled() const { return graph_on_ && !profiling_ && !recording_; }
    // the graph captured for this signature, or null
    FlushGraph* find_graph(const std::vector<uint64_t>& sig);
    // an empty slot whose device table holds at least `bytes` (evicts the
    // least recently used graph when all slots are taken)
    FlushGraph& new_graph_slot(size_t bytes);
    void note_graph(bool captured) { (captured ? graph_captures_ : graph_replays_) += 1; }
    uint64_t graph_captures() const { return graph_captures_; }
    uint64_t graph_replays() const { return graph_replays_; }
    size_t graph_count() const;
    void drop_graphs();

    // kernel profiling (off by default): every launch of a flush is bracketed
    // by HIP events; the times are gathered after the flush synchronises.
    void set_profiling(bool on);
    int prof_begin();
    void prof_end(int slot, const char* name, double bytes);
    std::vector<KernelProfile> profile() const;
    static StateComparison compare(GpuVectorState& a, size_t col_a, GpuVectorState& b, size_t col_b);

private:
    size_t nr_bits_, nr_shots_;
    Precision prec_;
    std::vector<size_t> counts_;
    bool basis_ = false;                  // columns are basis vectors |basis_idx_[c]>
    int basis_rot_ = 0;                   // ... rotated by H^n (1) or S^n H^n (2)
    std::vector<uint64_t> basis_idx_;
    void* amp_ = nullptr;                 // column-major: column c at amp_ + c * 2^n elements
    size_t amp_cols_ = 0;                 // allocated columns
    hipStream_t stream_ = nullptr;
    int device_ = 0;                      // the device current at construction
    std::shared_ptr<SharedStream> shared_;   // set: stream_ belongs to it
    std::vector<PendingGate> queue_;
    RunStats stats_;
    int fusion_k_ = 0;                    // 0: plain per-gate arithmetic
    bool tiling_ = true;
    TileGeom geom_{12, 4, 4};

    // scratch
    double* d_partial_ = nullptr;
    size_t partial_cap_ = 0;
    double* d_tree_ = nullptr;
    size_t tree_cap_ = 0;
    void* d_small_ = nullptr;
    size_t small_cap_ = 0;
    void* d_slot_[SLOT_COUNT] = {nullptr, nullptr, nullptr};
    size_t slot_cap_[SLOT_COUNT] = {0, 0, 0};
    std::unique_ptr<ItemQueue> iq_;       // the single-state operations' queue
    ItemQueue& iq();

    CostModel* cost_model_ = nullptr;
    int plan_mode_ = PLAN_GREEDY;
    int plan_search_ = -1;
    Xoshiro256 plan_rng_{1};
    bool recording_ = false;
    std::vector<SweepSample> samples_;
    size_t samples_done_ = 0;             // samples whose time is known

    int jit_mode_ = 0;
    bool geom_set_ = false;               // set_geometry was called
    bool graph_on_ = false;
    std::vector<FlushGraph> fgraphs_;

public:
    // the last per-gate plan of this state and its key (sweep.hip)
    std::vector<uint64_t>& plan_cache_key() { return plan_key_; }
    std::shared_ptr<void>& plan_cache_steps() { return plan_steps_; }

private:
    std::vector<uint64_t> plan_key_;
    std::shared_ptr<void> plan_steps_;
    uint64_t graph_clock_ = 0, graph_captures_ = 0, graph_replays_ = 0;

    bool profiling_ = false;
    std::vector<float> slot_ms_;          // per profiling slot, last collection
    struct PendingEvent {
        int pair;
        std::string name;
        double bytes;
    };
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool_;
    size_t ev_used_ = 0;
    std::vector<PendingEvent> ev_pending_;
    std::vector<KernelProfile> prof_;
    void prof_collect();

    uint64_t dim() const { return uint64_t(1) << nr_bits_; }
    void check_qbit(size_t q) const;
    void alloc_cols(size_t ncols);
    std::vector<double> basis_w0(size_t qbit, int basis) const;
    static void check_basis(int basis);
    struct CondRange {
        size_t col, len;
        bool apply;
    };
    std::vector<CondRange> conditional_ranges(const uint8_t* control) const;
    void check_conditional(size_t ncontrol, size_t nr_affected, const size_t* qbits, size_t nbits,
                           const std::string& desc) const;
    void* dev_alloc(size_t bytes);
    void dev_release(void* p, size_t bytes);
    void* small_buf(size_t bytes);
    double* partial_buf(size_t n);
    void materialize();
    void apply_queue_plain(size_t col_begin, size_t col_end);
};

}  // namespace q1t
