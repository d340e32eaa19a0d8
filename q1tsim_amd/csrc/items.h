// Column work of many states gathered into one launch per kind.
//
// Measurements, collapses, conditional gates and basis-column write-outs act
// on whole columns of a state, and a batch of circuits needs the same kind of
// work on thousands of columns of different states in the same round.  An
// ItemQueue collects that work as items (one per column, or per state for
// the sampling walk), uploads every item table in one copy, launches one
// kernel per kind in dependency order and copies every result back in one
// copy.  The single-state paths run the same kernels through a queue of
// their own, so a state gives the same bits alone and in a batch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "common.h"
#include "gates.h"

namespace q1t {

// Sampling tree of one column: level 0 is the column's amplitudes (weights
// |a|^2), level l > 0 holds the left folds of 16 consecutive entries of
// level l - 1; lev[l] points at column 0's level l, column c's is per_col
// entries further.
struct TreeDesc {
    int nlev;
    uint64_t size[16];
    const double* lev[16];   // lev[0] unused
};

// One column's outcome-0 weight of one qubit: kNormBlocks partial sums on
// the device, summed in order into one value.
struct MargItem {
    const void* amp;
    uint64_t nwork;   // dim / 2
    int32_t p;        // index bit of the qubit
    int32_t basis;    // 0 Z, 1 X, 2 Y
};

// Collapse of one column onto outcome 0 (dst0, scale s0) and/or 1 (dst1, s1)
// of bit p; src may alias either destination.
struct CollItem {
    const void* src;
    void* dst0;
    void* dst1;
    uint64_t nwork;
    double s0, s1;
    int32_t p, basis;
    uint64_t chunk0;   // first work chunk (filled by the queue)
};

// A one-column gate (conditional gates): kind as ColOp.
struct ColOp {
    int32_t kind;      // 0 dense 1 target, 1 dense 2 targets, 2 diagonal, 3 phase on the controls, 4 swap
    int32_t nt;
    int32_t t[3];      // target bits, t[0] = local MSB
    int32_t npos;      // deposit positions (targets and controls, ascending)
    int32_t pos[16];
    uint64_t cmask;
    double m[32];      // FP64 entries (re, im): dense 2^nt x 2^nt, diagonal 2^nt, phase 1
};

struct OpItem {
    void* amp;
    uint64_t nwork;
    uint64_t chunk0;
    uint32_t op;       // index into the queue's op table
    uint32_t pad;
};

// A dense column written from a compact basis column: |k> (rot 0), H^n|k>
// (rot 1) or S^n H^n|k> (rot 2), amplitude magnitude `mag`.
struct BasisItem {
    void* amp;
    uint64_t dim;
    uint64_t k;
    uint64_t chunk0;
    double mag;
    int32_t rot, pad;
};

// One tree level of one column: out[g] = left fold of in[16g .. 16g+15].
struct LevelItem {
    const void* in;   // level 1: amplitudes; higher levels: doubles
    double* out;
    uint64_t size;
    uint64_t chunk0;
};

// The sampling walk of all runs of one state: run r of the state (global
// run run0 + r) reads uniform u[run0 + r] and writes its index to
// idx[run0 + r]; its column is the last c with first[first_off + c] <= r.
struct SampleJob {
    const void* a0;
    uint64_t col_stride;
    TreeDesc td;
    uint64_t per_col;
    uint64_t first_off;
    uint64_t run0;
    uint64_t col0;     // global index of the state's first column (top totals)
    int32_t ncols, pad;
};

class ItemQueue {
public:
    explicit ItemQueue(bool c64);
    ~ItemQueue();
    ItemQueue(const ItemQueue&) = delete;
    ItemQueue& operator=(const ItemQueue&) = delete;

    bool c64() const { return c64_; }
    // Returns the item index; its weight is w0(index) once the stream has
    // synchronised after launch().
    size_t add_marginal(const MargItem& m);
    void add_collapse(const CollItem& c);
    // a ColOp applied to columns: add the op once, then one item per column
    uint32_t add_op(const ColOp& op);
    void add_op_item(void* amp, uint64_t nwork, uint32_t op);
    void add_basis(const BasisItem& b);
    // Sampling of one state: its level items, its job and its uniforms
    // (nruns = first.back()); returns the job index.  Indices land at
    // indices(job) and per-column top totals at tops(job) after launch().
    size_t add_sample(const SampleJob& job, const std::vector<LevelItem>& levels, const uint64_t* first,
                      const double* u, size_t nruns);
    // device blocks to release (on their stream) once launch() has enqueued
    // the kernels that read them
    void release_after(void* p, size_t bytes, hipStream_t s);

    bool empty() const;
    // Upload, launch everything in the order basis columns, collapses,
    // gates, marginals, tree levels, sampling, then copy the results back;
    // then perform the deferred releases and clear the inputs (results stay
    // readable until the next launch()).
    void launch(hipStream_t s);
    uint64_t launches() const { return launches_; }
    // synchronise and forget the last stream and its upload event (before
    // the queue moves to another owner); pending inputs are dropped
    void detach();

    double w0(size_t item) const;
    const uint64_t* indices(size_t job) const;
    const double* tops(size_t job) const;

private:
    bool c64_;
    std::vector<MargItem> marg_;
    std::vector<CollItem> coll_;
    std::vector<ColOp> ops_;
    std::vector<OpItem> opit_;
    std::vector<BasisItem> basis_;
    std::vector<LevelItem> levels_[16];
    std::vector<SampleJob> jobs_;
    std::vector<uint64_t> first_;   // per job: ncols + 1 entries
    std::vector<uint64_t> run0_;    // per job, plus the total
    std::vector<double> u_;
    uint64_t cols_ = 0;
    struct Release {
        void* p;
        size_t bytes;
        hipStream_t s;
    };
    std::vector<Release> rel_;
    uint64_t coll_chunks_ = 0, op_chunks_ = 0, basis_chunks_ = 0, level_chunks_[16] = {};

    // device and pinned host buffers: inputs, then outputs
    char* dbuf_ = nullptr;
    size_t dcap_ = 0;
    char* hin_ = nullptr;
    size_t hin_cap_ = 0;
    char* hout_ = nullptr;
    size_t hout_cap_ = 0;
    hipEvent_t in_done_ = nullptr;   // the last input upload has completed
    hipStream_t last_stream_ = nullptr;
    size_t out_w0_ = 0, out_idx_ = 0, out_top_ = 0;   // offsets in hout_
    std::vector<uint64_t> res_run0_, res_col0_;         // of the last launch
    uint64_t launches_ = 0;
};

// Queues kept for reuse (their pinned and device buffers survive): states
// come and go by the thousand, a queue's buffers would not.
class ItemQueuePool {
public:
    static std::unique_ptr<ItemQueue> take(bool c64);
    // the queue's last stream must be idle or synchronisable (it is
    // synchronised here and forgotten)
    static void give(std::unique_ptr<ItemQueue> q);
};

// Partial sums of the outcome-0 weight of bit p for `ncols` columns at
// stride col_stride into partial[c * kNormBlocks + b] (the arithmetic of
// the queue's marginal items; used by the allocation-free marginal).
void launch_marginal_cols(bool c64, const void* amp, uint64_t col_stride, uint32_t ncols, uint64_t nwork, int p,
                          int basis, double* partial, hipStream_t s);

// Maximum ColOp arity: dense on 1-2 targets, diagonal on up to 3, phases
// and swaps; anything else is applied by the per-gate kernels.
bool col_op_from(const GateOp& g, ColOp& out);

}  // namespace q1t
