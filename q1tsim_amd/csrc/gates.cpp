// Gate matrices (closed forms of each gate's matrix() in the reference) and
// the matrix-structure analysis used to choose a kernel.
#include "gates.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>

namespace q1t {

static const GateSpec GATE_TABLE[] = {
    {"i", "I", G_I, 0, 0},       {"h", "H", G_H, 0, 0},        {"x", "X", G_X, 0, 0},
    {"y", "Y", G_Y, 0, 0},       {"z", "Z", G_Z, 0, 0},        {"s", "S", G_S, 0, 0},
    {"sdg", "S†", G_SDG, 0, 0},  {"t", "T", G_T, 0, 0},        {"tdg", "T†", G_TDG, 0, 0},
    {"rx", "RX", G_RX, 0, 1},    {"ry", "RY", G_RY, 0, 1},     {"rz", "RZ", G_RZ, 0, 1},
    {"u1", "U1", G_U1, 0, 1},    {"u2", "U2", G_U2, 0, 2},     {"u3", "U3", G_U3, 0, 3},
    {"v", "V", G_V, 0, 0},       {"vdg", "V†", G_VDG, 0, 0},   {"swap", "Swap", G_SWAP, 0, 0},
    {"cx", "CX", G_X, 1, 0},     {"cy", "CY", G_Y, 1, 0},      {"cz", "CZ", G_Z, 1, 0},
    {"ch", "CH", G_H, 1, 0},     {"crx", "CRX", G_RX, 1, 1},   {"cry", "CRY", G_RY, 1, 1},
    {"crz", "CRZ", G_RZ, 1, 1},  {"cs", "CS", G_S, 1, 0},      {"csdg", "CS†", G_SDG, 1, 0},
    {"ct", "CT", G_T, 1, 0},     {"ctdg", "CT†", G_TDG, 1, 0}, {"cu1", "CU1", G_U1, 1, 1},
    {"cu2", "CU2", G_U2, 1, 2},  {"cu3", "CU3", G_U3, 1, 3},   {"cv", "CV", G_V, 1, 0},
    {"cvdg", "CV†", G_VDG, 1, 0}, {"ccx", "CCX", G_X, 2, 0},   {"ccz", "CCZ", G_Z, 2, 0},
    {"ccrx", "CCRX", G_RX, 2, 1}, {"ccry", "CCRY", G_RY, 2, 1}, {"ccrz", "CCRZ", G_RZ, 2, 1},
};

const GateSpec* find_gate(const std::string& name)
{
    std::string lower(name);
    std::transform(lower.begin(), lower.end(), lower.begin(), [](unsigned char c) { return std::tolower(c); });
    for (const GateSpec& g : GATE_TABLE)
        if (lower == g.name) return &g;
    return nullptr;
}

size_t gate_nbits(const GateSpec& g) { return (g.base == G_SWAP ? 2 : 1) + size_t(g.nctrl); }

static cdouble polar(double r, double th) { return cdouble(r * std::cos(th), r * std::sin(th)); }

std::vector<cdouble> gate_matrix(const GateSpec& g, const double* p)
{
    const double hs = 0.70710678118654752440;
    const cdouble z(0, 0), o(1, 0), i(0, 1);
    std::vector<cdouble> m;
    switch (g.base) {
    case G_I: m = {o, z, z, o}; break;
    case G_H: m = {hs, hs, hs, -hs}; break;
    case G_X: m = {z, o, o, z}; break;
    case G_Y: m = {z, -i, i, z}; break;
    case G_Z: m = {o, z, z, -o}; break;
    case G_S: m = {o, z, z, i}; break;
    case G_SDG: m = {o, z, z, -i}; break;
    case G_T: m = {o, z, z, cdouble(hs, hs)}; break;
    case G_TDG: m = {o, z, z, cdouble(hs, -hs)}; break;
    case G_RX: {
        double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
        m = {c, cdouble(0, -s), cdouble(0, -s), c};
        break;
    }
    case G_RY: {
        double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
        m = {c, -s, s, c};
        break;
    }
    case G_RZ: {
        cdouble e = polar(1.0, 0.5 * p[0]);
        m = {std::conj(e), z, z, e};
        break;
    }
    case G_U1: m = {o, z, z, polar(1.0, p[0])}; break;
    case G_U2:
        m = {cdouble(hs, 0), -polar(hs, p[1]), polar(hs, p[0]), polar(hs, p[0] + p[1])};
        break;
    case G_U3: {
        double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
        m = {cdouble(c, 0), -polar(s, p[2]), polar(s, p[1]), polar(c, p[1] + p[2])};
        break;
    }
    case G_V: m = {cdouble(0.5, 0.5), cdouble(0.5, -0.5), cdouble(0.5, -0.5), cdouble(0.5, 0.5)}; break;
    case G_VDG: m = {cdouble(0.5, -0.5), cdouble(0.5, 0.5), cdouble(0.5, 0.5), cdouble(0.5, -0.5)}; break;
    case G_SWAP:
        m.assign(16, z);
        m[0] = o;
        m[1 * 4 + 2] = o;
        m[2 * 4 + 1] = o;
        m[15] = o;
        break;
    }
    // control layers: eye(2d) with the gate in the lower-right block
    size_t d = (g.base == G_SWAP) ? 4 : 2;
    for (int c = 0; c < g.nctrl; ++c) {
        size_t nd = 2 * d;
        std::vector<cdouble> big(nd * nd, z);
        for (size_t k = 0; k < d; ++k) big[k * nd + k] = o;
        for (size_t r = 0; r < d; ++r)
            for (size_t q = 0; q < d; ++q) big[(d + r) * nd + d + q] = m[r * d + q];
        m.swap(big);
        d = nd;
    }
    return m;
}

// Is local bit `b` a control of the d x d matrix m: no mixing across the bit
// and identity on its 0-subspace?
static bool is_control(const std::vector<cdouble>& m, size_t d, int b)
{
    for (size_t r = 0; r < d; ++r)
        for (size_t c = 0; c < d; ++c) {
            bool rb = (r >> b) & 1, cb = (c >> b) & 1;
            const cdouble& v = m[r * d + c];
            if (rb != cb) {
                if (v != cdouble(0, 0)) return false;
            } else if (!rb) {
                if (v != (r == c ? cdouble(1, 0) : cdouble(0, 0))) return false;
            }
        }
    return true;
}

// Keep the bit-b = 1 block of m.
static std::vector<cdouble> take_one_block(const std::vector<cdouble>& m, size_t d, int b)
{
    size_t h = d / 2;
    std::vector<cdouble> r(h * h);
    auto expand = [b](size_t x) {
        size_t lo = x & ((size_t(1) << b) - 1);
        return ((x >> b) << (b + 1)) | (size_t(1) << b) | lo;
    };
    for (size_t i = 0; i < h; ++i)
        for (size_t j = 0; j < h; ++j) r[i * h + j] = m[expand(i) * d + expand(j)];
    return r;
}

GateOp analyze_gate(const cdouble* mat, size_t k, const size_t* qbits, size_t nr_bits)
{
    GateOp op;
    for (size_t a = 0; a < k; ++a)
        for (size_t b = a + 1; b < k; ++b)
            if (qbits[a] == qbits[b]) throw Error(ERR_INVALID_ARGUMENT, "Repeated qubit index in gate");
    size_t d = size_t(1) << k;
    std::vector<cdouble> m(mat, mat + d * d);
    // qubit listed j sits at local bit k-1-j and at index bit nr_bits-1-q
    std::vector<int> tgt;
    for (size_t j = 0; j < k; ++j) tgt.push_back(int(nr_bits - 1 - qbits[j]));
    bool changed = true;
    while (changed && tgt.size() > 1) {
        changed = false;
        size_t kk = tgt.size();
        for (size_t j = 0; j < kk && tgt.size() > 1; ++j) {
            int b = int(kk - 1 - j);
            if (is_control(m, d, b)) {
                op.ctrl_mask |= uint64_t(1) << tgt[j];
                m = take_one_block(m, d, b);
                d /= 2;
                tgt.erase(tgt.begin() + j);
                changed = true;
                break;
            }
        }
    }
    op.targets = tgt;
    size_t kt = tgt.size();
    bool diag = true, ident = true;
    for (size_t r = 0; r < d; ++r)
        for (size_t c = 0; c < d; ++c) {
            const cdouble& v = m[r * d + c];
            if (r != c && v != cdouble(0, 0)) diag = false;
            if (v != (r == c ? cdouble(1, 0) : cdouble(0, 0))) ident = false;
        }
    if (ident) {
        op.kind = OpKind::Identity;
        return op;
    }
    if (diag) {
        if (kt == 1 && m[0] == cdouble(1, 0)) {
            op.kind = OpKind::PhaseOnes;
            op.ctrl_mask |= uint64_t(1) << tgt[0];
            op.targets.clear();
            op.m = {m[3]};
            return op;
        }
        if (kt <= 3) {
            op.kind = OpKind::Diag;
            for (size_t r = 0; r < d; ++r) op.m.push_back(m[r * d + r]);
            return op;
        }
    }
    if (kt == 2) {
        const cdouble o(1, 0), z(0, 0);
        const cdouble sw[16] = {o, z, z, z, z, z, o, z, z, o, z, z, z, z, z, o};
        if (std::equal(m.begin(), m.end(), sw)) {
            op.kind = OpKind::Swap2;
            return op;
        }
    }
    op.kind = OpKind::Dense;
    op.m = m;
    return op;
}

}  // namespace q1t
