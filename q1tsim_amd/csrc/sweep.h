// Gate-stream planner and multi-gate sweeps (LDS/register tiles, fused blocks).
#pragma once

#include <string>
#include <vector>

#include "gvs.h"
#include "sweep_types.h"

namespace q1t {

// A planned step of the gate stream.
struct PlanStep {
    bool sweep = false;
    std::vector<int> tile;          // sweep: global bit positions of the tile (ascending)
    std::vector<size_t> gates;      // indices into the queue, in application order
};

// Default tile geometry per precision (interpreter kernels), and with the
// kernels specialised to each sweep program.
TileGeom tile_geometry(Precision p);
TileGeom tile_geometry_specialised(Precision p);
// Is (M, R) one of the compiled tile shapes for this precision?
bool tile_shape_supported(Precision p, int M, int R);

// Is a gate applicable inside a tile (targets that must be local, if any)?
bool sweepable(const GateOp& op, std::vector<int>* local_targets);

// Plan: runs of gates whose non-diagonal targets fit M tile bits (the L low
// bits always included); a gate may run before earlier gates it commutes
// with.  With fuse_k > 0 the non-diagonal gates need all their bits in the
// tile.  `search`: refine each greedy tile (and, in the programs below, each
// phase's register bits) by local search -- host time that pays once a
// sweep moves hundreds of MB (GpuVectorState::plan_search).
std::vector<PlanStep> plan_gates(const std::vector<PendingGate>& q, int n, const TileGeom& g, int fuse_k,
                                 bool search = true);

// Everything plan_gates reads of the queue and its settings: equal keys give
// equal plans (a flush repeating the previous flush's gate structure, with
// any angles, reuses its steps).
std::vector<uint64_t> plan_key(const std::vector<PendingGate>& q, int n, const TileGeom& g, int fuse_k, bool search);

// Register phases of one sweep (per-gate arithmetic).  `order`, if given,
// receives the queue indices of the sweep's gates in the order the kernel
// applies them (per phase: its register ops, then its write-back gates).
// `factors`: diagonal ops on bits outside the tile become one tile factor
// (a kind-2 phase; the specialised kernels apply it, the interpreter kernels
// do not take such programs).
void build_program(const std::vector<PendingGate>& q, const PlanStep& step, int M, int R,
                   std::vector<SwPhase>& phases, std::vector<SwOp>& ops, bool search = true,
                   std::vector<size_t>* order = nullptr, bool factors = false);

// The order in which the plain (per-gate) plan of the specialised kernels
// applies the queue's gates (identity gates, which never run, are left out;
// tile factors last in their sweep).
std::vector<size_t> plan_order(const std::vector<PendingGate>& q, int n, const TileGeom& g, bool search);

// Fused program of one sweep: dense blocks of at most K local qubits
// (composed in FP64, real-embedded, rounded to FP32 into `blocks`) and
// register phases for the rest.
void build_fused_program(const std::vector<PendingGate>& q, const PlanStep& step, int M, int R, int K,
                         std::vector<SwPhase>& phases, std::vector<SwOp>& ops, std::vector<float>& blocks,
                         int* nblocks, bool search = true, bool factors = false);

// Features of a sweep program for the learned cost model (20 floats):
//  0 register phases, 1 blocks of 3 qubits, 2 blocks of 4 qubits (a block of
//  k > 4 qubits counts as 4^(k-4) four-qubit blocks), 3-6 one-qubit ops by
//  body (general, real/RX/Y, swap, phase/diagonal), 7-9 ops with thread,
//  register and outer controls, 10 precision, 11 M, 12 R, 13 phases holding
//  a local bit >= 8 in registers, 14 tile bits at positions >= 20 (long HBM
//  strides), 15 gates, 16 widest block, 17 tile bits at positions >= 26,
//  18 ops folded into the tile factors, 19 kernel specialised to the
//  program (1) or the interpreter (0).
constexpr int kSweepFeatures = 20;
void sweep_features(const std::vector<SwPhase>& phases, const std::vector<SwOp>& ops, const PlanStep& step,
                    int precision, const TileGeom& g, bool specialised, float* out);

// Device buffer that grows on demand (synchronising `stream` first).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t cap = 0;
    void* get(size_t bytes, hipStream_t stream);
    ~DeviceBuffer();
};

// Pinned host buffer that grows on demand; the caller makes sure no copy
// from it is in flight when it grows or is rewritten.
struct PinnedBuffer {
    void* ptr = nullptr;
    size_t cap = 0;
    void* get(size_t bytes);
    ~PinnedBuffer();
};

// Flush the pending gates of many states together (all of one precision,
// tile geometry and fusion setting, sharing `stream`): states at least
// M+2 qubits wide are planned into sweeps and the r-th sweep of every state
// runs in one launch over all their tiles; smaller states run whole in LDS,
// one workgroup per column, all in one launch.
// The table goes up from `stage` (the call ends with the stream
// synchronised, so the next call may rewrite it).  `plan_s` and `pack_s`
// (optional) accumulate the host planning and table-building time.
void run_planned_batch(const std::vector<GpuVectorState*>& states, hipStream_t stream, DeviceBuffer& prog,
                       PinnedBuffer& stage, ItemQueue& iq, RunStats& agg, unsigned tile_chunk,
                       double* plan_s = nullptr, double* pack_s = nullptr);

// Sources of the kernels specialised to the sweeps of the plan a flush of
// `q` would run (plain programs; with fusion_k > 0 greedy fixed-width
// fusion), block offsets as in the flush's merged table; `factors`: diagonal
// ops on bits outside the tile folded into tile factors.  Host only.
std::vector<std::string> plan_jit_sources(const std::vector<PendingGate>& q, int n, Precision prec, int fusion_k,
                                          bool search, bool prefetch = false, bool factors = true);

// Apply the pending gates to columns [col_begin, col_end) of the state with
// the planned schedule: runs of gates whose qubits fit one tile are applied
// in one HBM sweep; `fusion_k` > 0 composes runs into dense blocks of at most
// that many qubits first (complex64 path).
void run_planned(GpuVectorState& st, const std::vector<PendingGate>& queue, size_t col_begin, size_t col_end,
                 int fusion_k, bool tiling, RunStats& stats);

}  // namespace q1t
