// Sweep program records (host and device): the in-tile ops and phases a
// tile sweep executes.  Plain C++ with fixed-width types only, so the
// sweep kernels specialised at run time can include it as text.
#pragma once

#ifndef __HIPCC_RTC__
#include <cstdint>
#else
// hipRTC: fixed-width types of its built-in runtime header
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint8_t uint8_t;
#endif

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
    uint32_t kind;       // 0: register phase, 1: fused dense block on LDS,
                         // 2: tile factors -- ops [op_begin, op_end) are
                         // diagonal ops on bits outside the tile only; the
                         // tile is scaled by the product of their factors
                         // as it is stored (the program's last phase;
                         // programs of the specialised kernels only)
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
    // register phases: LDS slot term of register slot k, swz of the local
    // index made of k's bits at the register positions (the slot of a
    // thread's element k is swz(thread part) ^ rslot[k]); set by the planner
    // for the kernel's swizzle (sweep_swizzle_bits)
    uint32_t rslot[16];
};

// LDS swizzle of local element l in a 2^M tile: XOR-fold of the bits above
// SW into the low SW bits (an involution, XOR-linear in l).
constexpr int swizzle_slot(int SW, int M, int l)
{
    int f = 0;
    for (int s = SW; s < M; s += SW) f ^= l >> s;
    return l ^ (f & ((1 << SW) - 1));
}

// Swizzle width of the sweep kernels for 2^M tiles (12 -> 4, 13 -> 5).
constexpr int sweep_swizzle_bits(int M) { return M - 8; }

}  // namespace q1t
