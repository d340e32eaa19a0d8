// Gate-stream planner and multi-gate sweeps (LDS/register tiles, fused blocks).
#pragma once

#include <vector>

#include "gvs.h"

namespace q1t {

// One in-tile operation of a sweep program (read by every thread of a
// workgroup; uniform).
enum SwOpType : int32_t { SW_DENSE1 = 0, SW_PHASE = 1, SW_DIAG1 = 2, SW_BLOCK = 3 };

// Body of a SW_DENSE1 op, read off its 2x2 matrix (SwOp::aux).
enum SwDenseForm : uint32_t { SWF_GENERAL = 0, SWF_REAL = 1, SWF_SWAP = 2, SWF_RX = 3, SWF_Y = 4 };
// Body of a SW_PHASE op (SwOp::aux): general, times i, times -i, times -1.
enum SwPhaseForm : uint32_t { SWP_GENERAL = 0, SWP_I = 1, SWP_MINUS_I = 2, SWP_NEG = 3 };

// Controls (and the ones-mask of a phase) are split by where their bits live
// during the op's phase: `cr` over register indices (uniform per register
// slot), `ct` over thread-index bits (one lane test per op), `co` over global
// bits outside the tile (uniform per tile).
struct SwOp {
    int32_t type;
    int32_t reg;       // SW_DENSE1: index of the target among the phase's register bits
    int32_t dreg;      // SW_DIAG1: register index of the target, or -1
    int32_t dtid;      // SW_DIAG1: thread-index bit of the target, or -1
    int32_t dglobal;   // SW_DIAG1: global bit of the target when it is outside the tile, or -1
    uint32_t cr;       // control (or phase) mask over register indices
    uint32_t ct;       // control (or phase) mask over thread-index bits
    uint32_t aux;      // SW_DENSE1 / SW_PHASE: body form; SW_BLOCK: offset of the block matrix
    uint64_t co;       // control (or phase) mask over global bits outside the tile
    double m[8];       // SW_DENSE1: m00, m01, m10, m11 (re, im); SW_PHASE: m[0..1]; SW_DIAG1: d0, d1
    float mf[8];       // m rounded to FP32 once (the complex64 sweeps read these)
};

struct SwPhase {
    uint32_t op_begin, op_end;
    uint32_t reg_mask;   // local tile bits held in registers during this phase
    uint32_t kind;       // 0: register phase, 1: fused dense block on LDS
    uint8_t rpos[8];     // local positions of the register bits (ascending)
    uint8_t tbit[16];    // local positions of the thread bits (ascending)
    // register phases: an affine map over GF(2) of the local index, applied
    // when the phase writes the tile back to LDS -- amplitude l moves to
    // XOR over the set bits of l of their columns, XOR wxor.  wcol holds the
    // columns of the thread bits (tbit order), then of the register bits
    // (rpos order).  SWAP, X and singly controlled X on local bits are such
    // maps and cost no op of their own.
    uint32_t wmap;       // 0: identity (plain write-back)
    uint32_t wxor;
    uint16_t wcol[16];
};

// A planned step of the gate stream.
struct PlanStep {
    bool sweep = false;
    std::vector<int> tile;          // sweep: global bit positions of the tile (ascending)
    std::vector<size_t> gates;      // indices into the queue, in application order
};

// Default tile geometry per precision.
TileGeom tile_geometry(Precision p);
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

// Register phases of one sweep (per-gate arithmetic).  `order`, if given,
// receives the queue indices of the sweep's gates in the order the kernel
// applies them (per phase: its register ops, then its write-back gates).
void build_program(const std::vector<PendingGate>& q, const PlanStep& step, int M, int R,
                   std::vector<SwPhase>& phases, std::vector<SwOp>& ops, bool search = true,
                   std::vector<size_t>* order = nullptr);

// The order in which the plain (per-gate) plan applies the queue's gates
// (identity gates, which never run, are left out).
std::vector<size_t> plan_order(const std::vector<PendingGate>& q, int n, const TileGeom& g, bool search);

// Fused program of one sweep: dense blocks of at most K local qubits
// (composed in FP64, real-embedded, rounded to FP32 into `blocks`) and
// register phases for the rest.
void build_fused_program(const std::vector<PendingGate>& q, const PlanStep& step, int M, int R, int K,
                         std::vector<SwPhase>& phases, std::vector<SwOp>& ops, std::vector<float>& blocks,
                         int* nblocks, bool search = true);

// Features of a sweep program for the learned cost model (16 floats):
// register phases, blocks of 3 and of 4 qubits (a block of k > 4 qubits
// counts as 4^(k-4) four-qubit blocks), one-qubit ops by body
// (general, real/RX/Y, swap, phase/diagonal), ops with thread, register and
// outer controls, precision, M, R, phases holding a local bit >= 8 in
// registers, tile bits at positions >= 20 (long HBM strides), gates.
constexpr int kSweepFeatures = 16;
void sweep_features(const std::vector<SwPhase>& phases, const std::vector<SwOp>& ops, const PlanStep& step,
                    int precision, const TileGeom& g, float* out);

// Device buffer that grows on demand (synchronising `stream` first).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t cap = 0;
    void* get(size_t bytes, hipStream_t stream);
    ~DeviceBuffer();
};

// Flush the pending gates of many states together (all of one precision,
// tile geometry and fusion setting, sharing `stream`): states at least
// M+2 qubits wide are planned into sweeps and the r-th sweep of every state
// runs in one launch over all their tiles; smaller states run whole in LDS,
// one workgroup per column, all in one launch.
void run_planned_batch(const std::vector<GpuVectorState*>& states, hipStream_t stream, DeviceBuffer& prog,
                       RunStats& agg, unsigned tile_chunk);

// Apply the pending gates to columns [col_begin, col_end) of the state with
// the planned schedule: runs of gates whose qubits fit one tile are applied
// in one HBM sweep; `fusion_k` > 0 composes runs into dense blocks of at most
// that many qubits first (complex64 path).
void run_planned(GpuVectorState& st, const std::vector<PendingGate>& queue, size_t col_begin, size_t col_end,
                 int fusion_k, bool tiling, RunStats& stats);

}  // namespace q1t
