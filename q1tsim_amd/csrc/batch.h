// Batches of circuits on one GPU: many independent circuits (uneven qubit
// counts, branch columns after mid-circuit measurements) executed in
// lockstep on one stream, the gate runs of all of them flushed together
// (one launch per sweep round; small states whole in LDS).
#pragma once

#include <memory>
#include <vector>

#include "circuit.h"
#include "sweep.h"

namespace q1t {

struct BatchStats {
    RunStats gates;            // summed over all flushes of all circuits
    uint64_t waves = 0;        // groups of circuits resident at once
    uint64_t rounds = 0;       // lockstep rounds (gate flush + sync ops)
    uint64_t sync_ops = 0;     // measurements, conditional gates, resets
    double seconds = 0.0;      // wall time of execute (host clock)
    // where that time went: state set-up, gate queueing (matrices, analysis),
    // batched flushes (planning + kernels), batched measurements, other ops
    double start_s = 0.0, queue_s = 0.0, flush_s = 0.0, measure_s = 0.0, other_s = 0.0;
    uint64_t column_launches = 0;   // kernels and copies of the batched column work
    uint64_t solo_ops = 0;          // ops run circuit by circuit
    double plan_s = 0.0;            // host planning inside the flushes
    double pack_s = 0.0;            // host table building and upload inside the flushes
};

// Default consecutive tiles per workgroup step of a batched sweep (4096
// circuits of 10-22 qubits: 16 beat 1, 4 and 64 on both precisions).
constexpr unsigned kDefaultTileChunk = 16;

// One configuration the autotuner timed.
struct BatchConfig {
    TileGeom geom;
    int fusion_k;
    double seconds;   // wall time of the sample under this configuration
};

class Batch {
public:
    // `mem_budget`: device bytes the resident states of one wave may take
    // (upper bound over their branch columns); 0 = 3/4 of free memory.
    Batch(Precision prec, int fusion_k, size_t mem_budget);
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    // The circuit must stay alive until execute returns; its backend
    // precision and fusion are set to the batch's.
    void add(Circuit* c);
    size_t size() const { return circuits_.size(); }
    // Run every circuit as Circuit::execute(nr_shots) would (same seeds, same
    // classical results); `keep_states` keeps each end state on the device.
    void execute(size_t nr_shots, bool keep_states);
    const BatchStats& stats() const { return stats_; }

    // Autotuner: execute the first `sample` circuits under every compiled
    // tile shape of the precision and, for complex64, fusion widths {0, 4}
    // (the 12/3 complex64 shape with fusion 0 only);
    // keep the fastest for execute.  Each sampled circuit's random stream is
    // restored afterwards, so execute still gives the results of
    // Circuit::execute.  Returns every configuration timed.
    std::vector<BatchConfig> autotune(size_t nr_shots, size_t sample);
    void set_config(const TileGeom& g, int fusion_k);
    BatchConfig config() const { return BatchConfig{geom_, fusion_k_, 0.0}; }
    // Consecutive tiles a workgroup of a batched sweep takes at a time (runs
    // of one state's column share its program in the scalar cache).
    void set_tile_chunk(unsigned c);
    unsigned tile_chunk() const { return tile_chunk_; }

private:
    Precision prec_;
    int fusion_k_;
    TileGeom geom_;
    unsigned tile_chunk_ = kDefaultTileChunk;
    size_t mem_budget_;
    std::vector<Circuit*> circuits_;
    std::shared_ptr<SharedStream> stream_;
    DeviceBuffer prog_;
    PinnedBuffer stage_;
    BatchStats stats_;
    std::unique_ptr<ItemQueue> iq_;   // the column work of a round, all circuits

    size_t resident_bytes(const Circuit& c, size_t nr_shots) const;
    void run(const std::vector<Circuit*>& circuits, size_t nr_shots, bool keep_states);
    void sync_ops(std::vector<Circuit*>& active);
};

}  // namespace q1t
