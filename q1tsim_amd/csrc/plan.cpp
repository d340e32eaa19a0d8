// Gate-stream planning for the tile sweeps (host side).
//
// plan_gates   : splits the gate queue into HBM sweeps (runs of gates whose
//                qubits fit one tile) and single-gate kernels.
// build_program: per sweep, register phases for per-gate arithmetic.
// build_fused_program: per sweep, dense k-qubit blocks composed in FP64 and
//                rounded once to FP32 (applied by MFMA), plus register phases
//                for the diagonal gates that reach outside the tile.
#include <algorithm>
#include <cstring>
#include <set>

#include "sweep.h"

namespace q1t {

TileGeom tile_geometry(Precision p)
{
    // 2^12-amplitude tiles in runs of 16 contiguous amplitudes.  complex128:
    // 8 amplitudes per thread, 512 threads (117 VGPRs: 4 waves per SIMD);
    // complex64: 16 per thread, 256 threads.  Measured on the 30-qubit
    // layered circuit and QFT, these beat the other compiled shapes
    // (tile_shape_supported); the batch autotuner still tries them all.
    return p == PREC_C128 ? TileGeom{12, 3, 4} : TileGeom{12, 4, 4};
}

// Measured with the specialised kernels on the 30-qubit layered circuit:
// complex128 M=12 R=4 126.9 ms vs R=3 140.1 ms; complex64 R=4 80.7 ms vs R=3
// 88.6 ms, M=13 107.5 ms (one more sweep).
TileGeom tile_geometry_specialised(Precision p)
{
    (void)p;
    return TileGeom{12, 4, 4};
}

bool tile_shape_supported(Precision p, int M, int R)
{
    if (p == PREC_C128) return M == 12 && (R == 4 || R == 3);
    return (M == 12 && (R == 4 || R == 3)) || (M == 13 && R == 4);
}

static std::vector<int> mask_list(uint64_t m)
{
    std::vector<int> r;
    for (int b = 0; b < 64; ++b)
        if ((m >> b) & 1) r.push_back(b);
    return r;
}

static uint64_t touched_bits(const GateOp& op)
{
    uint64_t m = op.ctrl_mask;
    for (int t : op.targets) m |= uint64_t(1) << t;
    return m;
}

static bool is_diagonal(const GateOp& op) { return op.kind == OpKind::Diag || op.kind == OpKind::PhaseOnes; }

// Bits a gate acts on non-diagonally: the targets of dense and swap gates.
// Controls and the bits of diagonal gates act diagonally; two gates whose
// shared bits are diagonal in both commute.
static uint64_t nondiag_bits(const GateOp& op)
{
    uint64_t m = 0;
    if (op.kind == OpKind::Dense || op.kind == OpKind::Swap2)
        for (int t : op.targets) m |= uint64_t(1) << t;
    return m;
}

namespace {

// The gates a pass skips, as far as later gates are concerned: a later gate
// may run before them only if it commutes with every one of them.
struct Blocker {
    uint64_t any = 0;   // bits the skipped gates touch
    uint64_t nd = 0;    // bits they act on non-diagonally
    bool blocks(uint64_t touched, uint64_t nondiag) const { return (touched & nd) || (nondiag & any); }
    void add(uint64_t touched, uint64_t nondiag)
    {
        any |= touched;
        nd |= nondiag;
    }
};

}  // namespace

bool sweepable(const GateOp& op, std::vector<int>* local_targets)
{
    if (local_targets) local_targets->clear();
    switch (op.kind) {
    case OpKind::Identity:
    case OpKind::PhaseOnes:
        return true;
    case OpKind::Diag:
        return op.targets.size() == 1;
    case OpKind::Dense:
        if (op.targets.size() != 1) return false;
        if (local_targets) local_targets->push_back(op.targets[0]);
        return true;
    case OpKind::Swap2:
        // a permutation of two local bits at a phase's write-back
        if (op.ctrl_mask) return false;
        if (local_targets) *local_targets = op.targets;
        return true;
    default:
        return false;
    }
}

// Which bits a gate needs inside the tile, or false if it cannot join a sweep.
// Plain: non-diagonal targets (controls and diagonal factors outside the tile
// are per-tile predicates / constants).  Fused (fuse_k > 0): every bit of a
// non-diagonal gate, so it can be absorbed into a block; diagonal one-target
// gates and phases may still reach outside.
static bool tile_need(const GateOp& op, int fuse_k, uint64_t* need)
{
    *need = 0;
    if (fuse_k > 0) {
        const uint64_t bits = touched_bits(op);
        if (__builtin_popcountll(bits) > fuse_k) return false;
        if (op.kind == OpKind::PhaseOnes || (op.kind == OpKind::Diag && op.targets.size() == 1)) return true;
        *need = bits;
        return true;
    }
    std::vector<int> lt;
    if (!sweepable(op, &lt)) return false;
    for (int t : lt) *need |= uint64_t(1) << t;
    return true;
}

namespace {

struct GateBits {
    uint64_t touched;   // controls and targets
    uint64_t nd;        // bits acted on non-diagonally
    uint64_t need;      // bits that must be local (tile_need)
    bool ok;            // can join a sweep
};

// Gates of `rem` (queue order) a sweep over the fixed tile `tile` takes: a
// gate whose needed bits are local and that commutes with every earlier
// skipped gate.  The scan stops once every bit is blocked.
size_t absorb_fixed(const std::vector<GateBits>& gb, const std::vector<size_t>& rem, uint64_t tile, uint64_t all,
                    std::vector<size_t>* sel)
{
    Blocker blk;
    size_t count = 0;
    for (size_t i : rem) {
        const GateBits& b = gb[i];
        if (b.ok && !blk.blocks(b.touched, b.nd) && !(b.need & ~tile)) {
            ++count;
            if (sel) sel->push_back(i);
        } else {
            blk.add(b.touched, b.nd);
            if (blk.nd == all) break;
        }
    }
    return count;
}

}  // namespace

std::vector<uint64_t> plan_key(const std::vector<PendingGate>& q, int n, const TileGeom& g, int fuse_k, bool search)
{
    std::vector<uint64_t> k;
    k.reserve(4 * q.size() + 4);
    k.push_back(uint64_t(n));
    k.push_back(uint64_t(g.M) | uint64_t(g.R) << 8 | uint64_t(g.L) << 16);
    k.push_back(uint64_t(fuse_k));
    k.push_back(search ? 1 : 0);
    for (const PendingGate& pg : q) {
        uint64_t need = 0;
        const bool ok = tile_need(pg.op, fuse_k, &need);
        k.push_back(touched_bits(pg.op));
        k.push_back(nondiag_bits(pg.op));
        k.push_back(need);
        k.push_back(uint64_t(pg.op.kind) | uint64_t(ok) << 8 | uint64_t(is_diagonal(pg.op)) << 9);
    }
    return k;
}

std::vector<PlanStep> plan_gates(const std::vector<PendingGate>& q, int n, const TileGeom& g, int fuse_k, bool search)
{
    std::vector<PlanStep> steps;
    std::vector<size_t> rem;
    std::vector<GateBits> gb(q.size());
    for (size_t i = 0; i < q.size(); ++i) {
        gb[i].touched = touched_bits(q[i].op);
        gb[i].nd = nondiag_bits(q[i].op);
        gb[i].ok = tile_need(q[i].op, fuse_k, &gb[i].need);
        if (q[i].op.kind != OpKind::Identity) rem.push_back(i);
    }
    const bool tiles_ok = n >= g.M + 2;
    const uint64_t all = n >= 64 ? ~uint64_t(0) : (uint64_t(1) << n) - 1;
    const uint64_t fixed = (uint64_t(1) << g.L) - 1;   // the contiguous low bits of every tile
    while (!rem.empty()) {
        if (!tiles_ok) {
            for (size_t i : rem) steps.push_back(PlanStep{false, {}, {i}});
            break;
        }
        // greedy start: the first gates' needed bits, in queue order
        uint64_t tile = fixed;
        int tcount = g.L;
        Blocker blk;
        std::vector<size_t> sel;
        for (size_t i : rem) {
            const GateBits& b = gb[i];
            if (blk.blocks(b.touched, b.nd) || !b.ok) {
                blk.add(b.touched, b.nd);
                continue;
            }
            const uint64_t need = b.need & ~tile;
            const int nn = __builtin_popcountll(need);
            if (tcount + nn <= g.M) {
                tile |= need;
                tcount += nn;
                sel.push_back(i);
            } else {
                blk.add(b.touched, b.nd);
            }
        }
        if (sel.empty()) {
            // the first remaining gate cannot run inside a tile
            steps.push_back(PlanStep{false, {}, {rem.front()}});
        } else if (sel.size() == 1 && (is_diagonal(q[sel[0]].op) || q[sel[0]].op.kind == OpKind::Swap2)) {
            // a lone diagonal or swap gate touches only part of the state: its own kernel is cheaper
            steps.push_back(PlanStep{false, {}, sel});
        } else {
            // fill the tile up to M bits with the lowest unused bits
            for (int b = 0; b < n && tcount < g.M; ++b)
                if (!((tile >> b) & 1)) {
                    tile |= uint64_t(1) << b;
                    ++tcount;
                }
            // local search: exchange one free tile bit for one outside bit
            // while that lets the sweep take more gates (first improvement;
            // on layered brick circuits this about halves the sweeps, as the
            // greedy start spends the tile on the first layer's targets)
            size_t best = absorb_fixed(gb, rem, tile, all, nullptr);
            for (int iter = 0; search && iter < 4 * n; ++iter) {
                bool improved = false;
                for (int a = g.L; a < n && !improved; ++a) {
                    if (!((tile >> a) & 1)) continue;
                    for (int b = 0; b < n; ++b) {
                        if ((tile >> b) & 1) continue;
                        const uint64_t t2 = (tile & ~(uint64_t(1) << a)) | (uint64_t(1) << b);
                        const size_t c = absorb_fixed(gb, rem, t2, all, nullptr);
                        if (c > best) {
                            best = c;
                            tile = t2;
                            improved = true;
                            break;
                        }
                    }
                }
                if (!improved) break;
            }
            sel.clear();
            absorb_fixed(gb, rem, tile, all, &sel);
            PlanStep st;
            st.sweep = true;
            st.tile = mask_list(tile);
            st.gates = sel;
            steps.push_back(st);
        }
        // drop the consumed gates, keeping the original order of the rest
        const std::set<size_t> done(steps.back().gates.begin(), steps.back().gates.end());
        std::vector<size_t> next;
        for (size_t i : rem)
            if (!done.count(i)) next.push_back(i);
        rem.swap(next);
    }
    return steps;
}

namespace {

struct LocalMap {
    int local_of[64];
    explicit LocalMap(const std::vector<int>& tile)
    {
        for (int b = 0; b < 64; ++b) local_of[b] = -1;
        for (size_t j = 0; j < tile.size(); ++j) local_of[tile[j]] = int(j);
    }
    void split(uint64_t m, uint32_t& cl, uint64_t& co) const
    {
        cl = 0;
        co = 0;
        for (int b = 0; b < 64; ++b)
            if ((m >> b) & 1) {
                if (local_of[b] >= 0) cl |= 1u << local_of[b];
                else co |= uint64_t(1) << b;
            }
    }
};

SwPhase make_phase(uint32_t kind, uint32_t regs, int M, uint32_t op_begin)
{
    SwPhase ph;
    std::memset(&ph, 0, sizeof(ph));
    ph.kind = kind;
    ph.reg_mask = regs;
    ph.op_begin = op_begin;
    int nr = 0, nt = 0;
    for (int b = 0; b < M; ++b) {
        if ((regs >> b) & 1) ph.rpos[nr++] = uint8_t(b);
        else ph.tbit[nt++] = uint8_t(b);
    }
    if (kind == 0 && nr <= 4)
        for (int k = 0; k < (1 << nr); ++k) {
            int l = 0;
            for (int b = 0; b < nr; ++b)
                if ((k >> b) & 1) l |= 1 << ph.rpos[b];
            ph.rslot[k] = uint32_t(swizzle_slot(sweep_swizzle_bits(M), M, l));
        }
    return ph;
}

// The cheapest exact body for a 2x2 matrix (row-major m00, m01, m10, m11).
uint32_t dense_form(const cdouble* m)
{
    const cdouble z(0, 0), o(1, 0);
    if (m[0] == z && m[1] == o && m[2] == o && m[3] == z) return SWF_SWAP;
    if (m[0] == z && m[1] == cdouble(0, -1) && m[2] == cdouble(0, 1) && m[3] == z) return SWF_Y;
    if (m[0].imag() == 0 && m[1].imag() == 0 && m[2].imag() == 0 && m[3].imag() == 0) return SWF_REAL;
    if (m[0].imag() == 0 && m[0] == m[3] && m[1].real() == 0 && m[1] == m[2]) return SWF_RX;
    return SWF_GENERAL;
}

// Where local bit b lives in a phase with register bits `regs`: register
// index (number of register bits below b) or thread-index bit (number of
// non-register bits below b).
static void local_slot(uint32_t regs, int b, bool* in_reg, int* idx)
{
    int r = 0, t = 0;
    for (int c = 0; c < b; ++c) ((regs >> c) & 1) ? ++r : ++t;
    *in_reg = (regs >> b) & 1;
    *idx = *in_reg ? r : t;
}

SwOp make_op(const GateOp& op, const LocalMap& lm, uint32_t regs)
{
    SwOp o;
    std::memset(&o, 0, sizeof(o));
    o.dreg = -1;
    o.dtid = -1;
    o.dglobal = -1;
    uint32_t cl = 0;
    lm.split(op.ctrl_mask, cl, o.co);
    for (int b = 0; b < 32; ++b)
        if ((cl >> b) & 1) {
            bool in_reg;
            int idx;
            local_slot(regs, b, &in_reg, &idx);
            if (in_reg) o.cr |= 1u << idx;
            else o.ct |= 1u << idx;
        }
    if (op.kind == OpKind::Dense) {
        o.type = SW_DENSE1;
        bool in_reg;
        local_slot(regs, lm.local_of[op.targets[0]], &in_reg, &o.reg);
        for (int k = 0; k < 4; ++k) {
            o.m[2 * k] = op.m[k].real();
            o.m[2 * k + 1] = op.m[k].imag();
        }
        o.aux = dense_form(op.m.data());
    } else if (op.kind == OpKind::PhaseOnes) {
        o.type = SW_PHASE;
        o.m[0] = op.m[0].real();
        o.m[1] = op.m[0].imag();
        const cdouble p = op.m[0];
        o.aux = p == cdouble(0, 1) ? SWP_I : p == cdouble(0, -1) ? SWP_MINUS_I : p == cdouble(-1, 0) ? SWP_NEG : SWP_GENERAL;
    } else {  // Diag on one target
        o.type = SW_DIAG1;
        const int t = op.targets[0];
        if (lm.local_of[t] >= 0) {
            bool in_reg;
            int idx;
            local_slot(regs, lm.local_of[t], &in_reg, &idx);
            (in_reg ? o.dreg : o.dtid) = idx;
        } else {
            o.dglobal = t;
        }
        o.m[0] = op.m[0].real();
        o.m[1] = op.m[0].imag();
        o.m[2] = op.m[1].real();
        o.m[3] = op.m[1].imag();
    }
    for (int k = 0; k < 8; ++k) o.mf[k] = float(o.m[k]);
    return o;
}

// One register phase by list scheduling over `pending`: a gate waits for
// every earlier gate that shares a bit; one-qubit updates on up to R local
// bits share the phase.  Returns the gates left for later phases.
// A gate the write-back map of a phase can apply: an uncontrolled SWAP, or an
// X with at most one control, all its bits local.
static bool writeback_gate(const GateOp& op, const LocalMap& lm)
{
    if (op.kind == OpKind::Swap2) {
        if (op.ctrl_mask) return false;
    } else if (op.kind == OpKind::Dense && op.targets.size() == 1 && op.m.size() == 4) {
        if (dense_form(op.m.data()) != SWF_SWAP || __builtin_popcountll(op.ctrl_mask) > 1) return false;
    } else {
        return false;
    }
    for (int b = 0; b < 64; ++b)
        if (((touched_bits(op) >> b) & 1) && lm.local_of[b] < 0) return false;
    return true;
}

// How a register phase over `regs` runs a gate: as an op on the registers
// (dense gates with their target in a register, diagonal gates), in the
// write-back map (SWAP, and X / CX whose target is not a register bit), or
// not at all.
enum PhaseRole { ROLE_NONE, ROLE_OP, ROLE_WRITEBACK };
static PhaseRole phase_role(const GateOp& op, const LocalMap& lm, uint32_t regs)
{
    if (op.kind == OpKind::Dense && op.targets.size() == 1 && lm.local_of[op.targets[0]] >= 0 &&
        ((regs >> lm.local_of[op.targets[0]]) & 1))
        return sweepable(op, nullptr) ? ROLE_OP : ROLE_NONE;
    if (writeback_gate(op, lm)) return ROLE_WRITEBACK;
    if (sweepable(op, nullptr) && op.kind != OpKind::Dense && op.kind != OpKind::Swap2) return ROLE_OP;
    return ROLE_NONE;
}

// Gates of `pending` a register phase over register bits `regs` takes:
// write-back gates (SWAP, X, CX on local bits: applied as the phase writes
// the tile back, so later gates on their bits wait for the next phase),
// dense gates whose target is a register bit and diagonal gates (any bits),
// in order; a gate not taken blocks its bits for the rest.
size_t phase_absorb(const std::vector<PendingGate>& q, const std::vector<size_t>& pending, const LocalMap& lm,
                    uint32_t regs, std::vector<size_t>* chosen, std::vector<size_t>* rest)
{
    Blocker blk, after_wb;
    size_t count = 0;
    for (size_t i : pending) {
        const GateOp& op = q[i].op;
        const uint64_t bits = touched_bits(op), nd = nondiag_bits(op);
        const bool free = !blk.blocks(bits, nd) && !after_wb.blocks(bits, nd);
        bool take = false;
        if (free) {
            switch (phase_role(op, lm, regs)) {
            case ROLE_OP: take = true; break;
            case ROLE_WRITEBACK:
                take = true;
                after_wb.add(bits, bits);
                break;
            default: break;
            }
        }
        if (take) {
            ++count;
            if (chosen) chosen->push_back(i);
        } else {
            blk.add(bits, nd);
            if (rest) rest->push_back(i);
        }
    }
    return count;
}

// Diagonal ops all of whose bits lie outside the tile: each multiplies the
// whole tile by one factor (a function of the tile's outer bits), so they
// commute with every op of the sweep.  A sweep collects them and scales the
// tile once, by their product, as it stores it (a SwPhase of kind 2).
struct TileFactors {
    std::vector<SwOp> ops;
    std::vector<size_t> order;   // their queue indices
};

static bool tile_scalar(const SwOp& o)
{
    return o.cr == 0 && o.ct == 0 && (o.type == SW_PHASE || (o.type == SW_DIAG1 && o.dreg < 0 && o.dtid < 0));
}

// The collected factors as the program's last phase (kind 2).
static void append_factors(TileFactors& tf, int M, std::vector<SwPhase>& phases, std::vector<SwOp>& ops,
                           std::vector<size_t>* order)
{
    if (tf.ops.empty()) return;
    SwPhase ph = make_phase(2, 0, M, uint32_t(ops.size()));
    ops.insert(ops.end(), tf.ops.begin(), tf.ops.end());
    ph.op_end = uint32_t(ops.size());
    phases.push_back(ph);
    if (order) order->insert(order->end(), tf.order.begin(), tf.order.end());
}

// One register phase: R local bits held in registers.  Start from the
// first dense targets in queue order (filled up with the lowest local bits),
// then exchange one register bit for another local bit while the phase takes
// more gates: every phase costs a full LDS round trip of the tile.  Tile
// factors go to `tf`; a phase left with no op and no write-back map is not
// emitted.
std::vector<size_t> one_reg_phase(const std::vector<PendingGate>& q, const std::vector<size_t>& pending,
                                  const LocalMap& lm, int M, int R, std::vector<SwPhase>& phases,
                                  std::vector<SwOp>& ops, bool search, std::vector<size_t>* order,
                                  TileFactors* tf)
{
    uint32_t regs = 0;
    int nregs = 0;
    Blocker blk;
    for (size_t i : pending) {
        const GateOp& op = q[i].op;
        const uint64_t bits = touched_bits(op), nd = nondiag_bits(op);
        if (blk.blocks(bits, nd) || !sweepable(op, nullptr)) {
            blk.add(bits, nd);
            continue;
        }
        if (op.kind == OpKind::Dense) {
            const int lb = lm.local_of[op.targets[0]];
            if ((regs >> lb) & 1) continue;
            if (nregs < R) {
                regs |= 1u << lb;
                ++nregs;
            } else {
                blk.add(bits, nd);
            }
        }
    }
    for (int b = 0; b < M && nregs < R; ++b)
        if (!((regs >> b) & 1)) {
            regs |= 1u << b;
            ++nregs;
        }
    size_t best = phase_absorb(q, pending, lm, regs, nullptr, nullptr);
    for (int iter = 0; search && iter < 4 * M; ++iter) {
        bool improved = false;
        for (int a = 0; a < M && !improved; ++a) {
            if (!((regs >> a) & 1)) continue;
            for (int b = 0; b < M; ++b) {
                if ((regs >> b) & 1) continue;
                const uint32_t r2 = (regs & ~(1u << a)) | (1u << b);
                const size_t c = phase_absorb(q, pending, lm, r2, nullptr, nullptr);
                if (c > best) {
                    best = c;
                    regs = r2;
                    improved = true;
                    break;
                }
            }
        }
        if (!improved) break;
    }
    std::vector<size_t> chosen, rest;
    phase_absorb(q, pending, lm, regs, &chosen, &rest);
    // Permutation gates taken as register ops (X, CX and SWAP with their
    // target in a register) that no later register op of the phase touches
    // commute with every later op: they join the write-back map (composed
    // in queue order after the ops) and cost no op.
    std::vector<char> as_op(chosen.size(), 0);
    {
        uint64_t later = 0;   // bits of the later gates that stay register ops
        for (size_t x = chosen.size(); x-- > 0;) {
            const GateOp& op = q[chosen[x]].op;
            if (phase_role(op, lm, regs) != ROLE_OP) continue;
            const uint64_t bits = touched_bits(op);
            if (writeback_gate(op, lm) && !(bits & later)) continue;
            as_op[x] = 1;
            later |= bits;
        }
    }
    SwPhase ph = make_phase(0, regs, M, uint32_t(ops.size()));
    // the write-back map, composed in queue order: f(l) = XOR_j l_j col[j] ^ c
    uint32_t col[16], cst = 0;
    for (int j = 0; j < 16; ++j) col[j] = 1u << j;
    auto is_factor = [&](size_t x) {
        return tf && as_op[x] && tile_scalar(make_op(q[chosen[x]].op, lm, regs));
    };
    if (order) {   // execution order: the register ops, then the write-back gates
        for (size_t x = 0; x < chosen.size(); ++x)
            if (as_op[x] && !is_factor(x)) order->push_back(chosen[x]);
        for (size_t x = 0; x < chosen.size(); ++x)
            if (!as_op[x]) order->push_back(chosen[x]);
    }
    auto apply = [&](auto&& g) {   // f := g o f for a map g on local indices
        for (int j = 0; j < M; ++j) col[j] = g(col[j], false);
        cst = g(cst, true);
    };
    for (size_t ci = 0; ci < chosen.size(); ++ci) {
        const size_t i = chosen[ci];
        const GateOp& op = q[i].op;
        if (as_op[ci]) {
            const SwOp so = make_op(op, lm, regs);
            if (tf && tile_scalar(so)) {
                tf->ops.push_back(so);
                tf->order.push_back(i);
            } else {
                ops.push_back(so);
            }
            continue;
        }
        ph.wmap = 1;
        if (op.kind == OpKind::Swap2) {
            const int a = lm.local_of[op.targets[0]], b = lm.local_of[op.targets[1]];
            apply([&](uint32_t x, bool) { const uint32_t d = ((x >> a) ^ (x >> b)) & 1u; return x ^ (d << a) ^ (d << b); });
        } else {
            const int t = lm.local_of[op.targets[0]];
            if (op.ctrl_mask) {
                const int c = lm.local_of[__builtin_ctzll(op.ctrl_mask)];
                apply([&](uint32_t x, bool) { return x ^ (((x >> c) & 1u) << t); });
            } else {
                // X: a constant, so only the affine part changes
                apply([&](uint32_t x, bool affine) { return affine ? x ^ (1u << t) : x; });
            }
        }
    }
    // columns in the kernel's order: thread bits, then register bits
    {
        int k = 0;
        for (int b = 0; b < M - R; ++b) ph.wcol[k++] = uint16_t(col[ph.tbit[b]]);
        for (int j = 0; j < R; ++j) ph.wcol[k++] = uint16_t(col[ph.rpos[j]]);
    }
    ph.wxor = cst;
    ph.op_end = uint32_t(ops.size());
    if (ph.op_end > ph.op_begin || ph.wmap) phases.push_back(ph);
    return rest;
}

// Apply a gate to every column of the block matrix U (dimension 2^|B|);
// pos[b] = block-index bit of global bit b.
void compose(std::vector<cdouble>& U, int nb, const GateOp& op, const int* pos)
{
    const size_t d = size_t(1) << nb;
    uint64_t cmask = 0;
    for (int b = 0; b < 64; ++b)
        if ((op.ctrl_mask >> b) & 1) cmask |= uint64_t(1) << pos[b];
    std::vector<int> tb;
    for (int t : op.targets) tb.push_back(pos[t]);
    const int kt = int(tb.size());
    std::vector<cdouble> tmp(size_t(1) << kt), out(size_t(1) << kt);
    for (size_t c = 0; c < d; ++c) {
        cdouble* col = nullptr;  // column c is strided: U[r * d + c]
        (void)col;
        switch (op.kind) {
        case OpKind::Identity:
            break;
        case OpKind::PhaseOnes:
            for (size_t r = 0; r < d; ++r)
                if ((r & cmask) == cmask) U[r * d + c] *= op.m[0];
            break;
        case OpKind::Diag:
            for (size_t r = 0; r < d; ++r) {
                if ((r & cmask) != cmask) continue;
                size_t l = 0;
                for (int q = 0; q < kt; ++q) l = (l << 1) | ((r >> tb[q]) & 1);
                U[r * d + c] *= op.m[l];
            }
            break;
        case OpKind::Swap2: {
            const size_t b0 = size_t(1) << tb[0], b1 = size_t(1) << tb[1];
            for (size_t r = 0; r < d; ++r)
                if ((r & cmask) == cmask && (r & b0) && !(r & b1)) std::swap(U[r * d + c], U[(r ^ b0 ^ b1) * d + c]);
            break;
        }
        case OpKind::Dense: {
            size_t tmask = 0;
            for (int t : tb) tmask |= size_t(1) << t;
            const size_t D = size_t(1) << kt;
            for (size_t r = 0; r < d; ++r) {
                if ((r & tmask) || (r & cmask) != cmask) continue;
                for (size_t l = 0; l < D; ++l) {
                    size_t ri = r;
                    for (int q = 0; q < kt; ++q)
                        if ((l >> (kt - 1 - q)) & 1) ri |= size_t(1) << tb[q];
                    tmp[l] = U[ri * d + c];
                }
                for (size_t l = 0; l < D; ++l) {
                    cdouble acc = 0;
                    for (size_t m = 0; m < D; ++m) acc += op.m[l * D + m] * tmp[m];
                    out[l] = acc;
                }
                for (size_t l = 0; l < D; ++l) {
                    size_t ri = r;
                    for (int q = 0; q < kt; ++q)
                        if ((l >> (kt - 1 - q)) & 1) ri |= size_t(1) << tb[q];
                    U[ri * d + c] = out[l];
                }
            }
            break;
        }
        }
    }
}

}  // namespace

void build_program(const std::vector<PendingGate>& q, const PlanStep& step, int M, int R,
                   std::vector<SwPhase>& phases, std::vector<SwOp>& ops, bool search, std::vector<size_t>* order,
                   bool factors)
{
    const LocalMap lm(step.tile);
    std::vector<size_t> pending = step.gates;
    TileFactors tf;
    TileFactors* tfp = factors ? &tf : nullptr;
    while (!pending.empty()) pending = one_reg_phase(q, pending, lm, M, R, phases, ops, search, order, tfp);
    append_factors(tf, M, phases, ops, order);
}

std::vector<size_t> plan_order(const std::vector<PendingGate>& q, int n, const TileGeom& g, bool search)
{
    std::vector<size_t> order;
    for (const PlanStep& st : plan_gates(q, n, g, 0, search)) {
        if (!st.sweep) {
            order.insert(order.end(), st.gates.begin(), st.gates.end());
            continue;
        }
        std::vector<SwPhase> phases;
        std::vector<SwOp> ops;
        build_program(q, st, g.M, g.R, phases, ops, search, &order);
    }
    return order;
}

void build_fused_program(const std::vector<PendingGate>& q, const PlanStep& step, int M, int R, int K,
                         std::vector<SwPhase>& phases, std::vector<SwOp>& ops, std::vector<float>& blocks,
                         int* nblocks, bool search, bool factors)
{
    const LocalMap lm(step.tile);
    std::vector<size_t> pending = step.gates;
    TileFactors tf;
    TileFactors* tfp = factors ? &tf : nullptr;
    while (!pending.empty()) {
        // grow one block: gates whose bits are all in the tile, while the
        // union of their local bits stays within K (order kept per bit)
        std::vector<int> B;            // local bits, B[0] = block MSB
        Blocker blk;
        std::vector<size_t> absorbed, rest;
        for (size_t i : pending) {
            const GateOp& op = q[i].op;
            const uint64_t bits = touched_bits(op), nd = nondiag_bits(op);
            bool ok = !blk.blocks(bits, nd);
            std::vector<int> add;
            if (ok) {
                for (int b = 0; b < 64 && ok; ++b)
                    if ((bits >> b) & 1) {
                        const int lb = lm.local_of[b];
                        if (lb < 0) ok = false;
                        else if (std::find(B.begin(), B.end(), lb) == B.end()) add.push_back(lb);
                    }
                ok = ok && int(B.size() + add.size()) <= K;
            }
            if (!ok) {
                blk.add(bits, nd);
                rest.push_back(i);
                continue;
            }
            B.insert(B.end(), add.begin(), add.end());
            absorbed.push_back(i);
        }
        if (absorbed.empty()) {
            pending = one_reg_phase(q, pending, lm, M, R, phases, ops, search, nullptr, tfp);
            continue;
        }
        if (search && int(B.size()) == K) {
            // local search over the block's K local bits: exchange one for
            // another while the block absorbs more gates (fewer blocks)
            auto absorb_set = [&](uint32_t S, std::vector<size_t>* ab, std::vector<size_t>* rs) {
                Blocker bk;
                size_t cnt = 0;
                for (size_t i : pending) {
                    const GateOp& op = q[i].op;
                    const uint64_t bits = touched_bits(op), nd = nondiag_bits(op);
                    bool ok = !bk.blocks(bits, nd);
                    for (int b = 0; b < 64 && ok; ++b)
                        if ((bits >> b) & 1) {
                            const int lb = lm.local_of[b];
                            ok = lb >= 0 && ((S >> lb) & 1);
                        }
                    if (ok) { ++cnt; if (ab) ab->push_back(i); }
                    else { bk.add(bits, nd); if (rs) rs->push_back(i); }
                }
                return cnt;
            };
            uint32_t S = 0;
            for (int lb : B) S |= 1u << lb;
            size_t best = absorb_set(S, nullptr, nullptr);
            for (int iter = 0; iter < 4 * M; ++iter) {
                bool improved = false;
                for (int a = 0; a < M && !improved; ++a) {
                    if (!((S >> a) & 1)) continue;
                    for (int b = 0; b < M; ++b) {
                        if ((S >> b) & 1) continue;
                        const uint32_t S2 = (S & ~(1u << a)) | (1u << b);
                        const size_t c = absorb_set(S2, nullptr, nullptr);
                        if (c > best) { best = c; S = S2; improved = true; break; }
                    }
                }
                if (!improved) break;
            }
            absorbed.clear(); rest.clear();
            absorb_set(S, &absorbed, &rest);
            B.clear();
            for (int b = 0; b < M; ++b) if ((S >> b) & 1) B.push_back(b);
        }
        // pad to at least 3 bits (the MFMA tile needs 2D >= 16)
        for (int b = 0; b < M && B.size() < 3; ++b)
            if (std::find(B.begin(), B.end(), b) == B.end()) B.push_back(b);
        const int nb = int(B.size());
        const size_t d = size_t(1) << nb;
        int pos[64];
        for (int b = 0; b < 64; ++b) pos[b] = -1;
        for (int j = 0; j < nb; ++j) pos[step.tile[B[j]]] = nb - 1 - j;
        std::vector<cdouble> U(d * d, 0.0);
        for (size_t r = 0; r < d; ++r) U[r * d + r] = 1.0;
        for (size_t i : absorbed) compose(U, nb, q[i].op, pos);
        // real embedding [[Ur, -Ui], [Ui, Ur]], rounded once to FP32
        const uint32_t off = uint32_t(blocks.size());
        const size_t d2 = 2 * d;
        blocks.resize(blocks.size() + d2 * d2);
        float* A = blocks.data() + off;
        for (size_t r = 0; r < d; ++r)
            for (size_t c = 0; c < d; ++c) {
                const cdouble u = U[r * d + c];
                A[r * d2 + c] = float(u.real());
                A[r * d2 + d + c] = float(-u.imag());
                A[(d + r) * d2 + c] = float(u.imag());
                A[(d + r) * d2 + d + c] = float(u.real());
            }
        {
            // MFMA fragment order: lane (li, lg) of row tile rt reads
            // A[16 rt + li][4 s + lg], s = 0 .. NS-1, as one contiguous run
            // (vector loads instead of NS scattered ones)
            const size_t NS = d2 / 4, NRT = d2 / 16;
            std::vector<float> f(d2 * d2);
            for (size_t rt = 0; rt < NRT; ++rt)
                for (size_t lane = 0; lane < 64; ++lane)
                    for (size_t s = 0; s < NS; ++s)
                        f[(rt * 64 + lane) * NS + s] = A[(16 * rt + (lane & 15)) * d2 + 4 * s + (lane >> 4)];
            std::copy(f.begin(), f.end(), A);
        }
        uint32_t regs = 0;
        for (int lb : B) regs |= 1u << lb;
        SwPhase ph = make_phase(1, regs, M, off);
        // block bits MSB first; column bits ascending
        for (int j = 0; j < nb; ++j) ph.rpos[j] = uint8_t(B[j]);
        ph.reg_mask = uint32_t(nb);
        ph.op_end = off;
        phases.push_back(ph);
        if (nblocks) ++*nblocks;
        pending.swap(rest);
    }
    append_factors(tf, M, phases, ops, nullptr);
}

void sweep_features(const std::vector<SwPhase>& phases, const std::vector<SwOp>& ops, const PlanStep& step,
                    int precision, const TileGeom& g, bool specialised, float* f)
{
    for (int k = 0; k < kSweepFeatures; ++k) f[k] = 0.0f;
    for (const SwPhase& ph : phases) {
        int nb;
        if (ph.kind == 1) {
            nb = int(ph.reg_mask);
            // blocks wider than 4 qubits count as 4-qubit blocks of equal work
            if (nb == 3) f[1] += 1.0f;
            else f[2] += float(1 << (2 * (nb - 4)));
            f[16] = std::max(f[16], float(nb));
        } else if (ph.kind == 0) {
            nb = g.R;
            f[0] += 1.0f;
        } else {
            // the tile factors: no LDS exchange, one multiply per amplitude
            f[18] += float(ph.op_end - ph.op_begin);
            continue;
        }
        for (int j = 0; j < nb; ++j)
            if (ph.rpos[j] >= 8) {
                f[13] += 1.0f;
                break;
            }
    }
    for (const SwOp& o : ops) {
        if (o.type == SW_DENSE1) {
            if (o.aux == SWF_GENERAL) f[3] += 1.0f;
            else if (o.aux == SWF_SWAP) f[5] += 1.0f;
            else f[4] += 1.0f;
        } else {
            f[6] += 1.0f;
        }
        if (o.ct) f[7] += 1.0f;
        if (o.cr) f[8] += 1.0f;
        if (o.co) f[9] += 1.0f;
    }
    f[10] = float(precision);
    f[11] = float(g.M);
    f[12] = float(g.R);
    for (int b : step.tile) {
        if (b >= 20) f[14] += 1.0f;
        if (b >= 26) f[17] += 1.0f;
    }
    f[15] = float(step.gates.size());
    f[19] = specialised ? 1.0f : 0.0f;
}

}  // namespace q1t
