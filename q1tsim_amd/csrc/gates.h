// Gate matrices and the structure analysis that picks a kernel for a gate.
#pragma once

#include "common.h"

namespace q1t {

// Gate table entry: the uncontrolled base gate plus control layers, as the
// reference builds CX = C<X>, CCX = C<CX>, CU1 = C<U1> (controlled.rs:23-82,
// 402-552) and the names of ffi.rs:335-366 / composite.rs:289-445.
struct GateSpec {
    const char* name;
    const char* desc;   // Gate::description()
    int base;           // BaseGate
    int nctrl;
    int nparams;
};

enum BaseGate {
    G_I, G_H, G_X, G_Y, G_Z, G_S, G_SDG, G_T, G_TDG, G_RX, G_RY, G_RZ,
    G_U1, G_U2, G_U3, G_V, G_VDG, G_SWAP
};

// Lookup by case-insensitive name; nullptr if unknown.
const GateSpec* find_gate(const std::string& name);

// Number of qubits the gate acts on.
size_t gate_nbits(const GateSpec& g);

// Gate::matrix() in FP64, row-major 2^k x 2^k, first listed qubit = MSB of
// the local index (gates.rs:176-183).
std::vector<cdouble> gate_matrix(const GateSpec& g, const double* params);

// How a gate is applied to the state, found from its matrix alone (the
// backend only sees Gate::matrix(), as behind the reference's QuState trait).
enum class OpKind { Identity, Dense, Diag, PhaseOnes, Swap2 };

struct GateOp {
    OpKind kind = OpKind::Dense;
    std::vector<int> targets;   // index-bit positions p = n-1-q; targets[0] = local MSB
    uint64_t ctrl_mask = 0;     // index bits that must all be 1 (controls)
    std::vector<cdouble> m;     // Dense: 2^k x 2^k; Diag: 2^k entries; PhaseOnes: 1 entry
    int nctrl() const { return __builtin_popcountll(ctrl_mask); }
};

// Reduce a (k-qubit, qubits `qbits`) gate to its controls and target block and
// classify the block.  Throws on repeated qubits.
GateOp analyze_gate(const cdouble* mat, size_t k, const size_t* qbits, size_t nr_bits);

}  // namespace q1t
