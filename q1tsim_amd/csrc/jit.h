// Sweep kernels specialised at run time to one sweep program.
//
// The interpreter kernel (k_sweep) reads each op's type, body form, register
// slot and controls from the program table and dispatches on them per op; the
// register array it carries through that dispatch costs copies at every op.
// A specialised kernel has the program's structure as straight-line code
// (same op bodies, every field except the gate matrices a compile-time
// constant) and reads only the matrices from the uploaded table, so the same
// kernel serves every program of the same structure, whatever the angles.
// Kernels are compiled with hipRTC from the device headers embedded in the
// library and cached per process by their source text.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "sweep_types.h"

namespace q1t {

// Name of the specialised kernel: q1t_jit_sweep_c128 / q1t_jit_sweep_c64.
const char* jit_kernel_name(bool c64);

// Source of the specialised kernel (jit_kernel_name) for one program: the same
// arguments as k_sweep<V, M, R, SW, *> (amplitudes, column stride,
// SweepArgs), block phases addressing the flush's block table at their
// op_begin.
// `prefetch`: software-pipelined tiles -- a thread's share of the next tile
// is loaded into registers while the phases of the current one run.
std::string jit_sweep_source(const std::vector<SwPhase>& phases, const std::vector<SwOp>& ops, bool c64, int M,
                             int R, int SW, bool prefetch = false);

// Compile a source with hipRTC for gfx950 (no device needed); returns the
// code object, or throws with the compiler log.
std::vector<char> jit_compile(const std::string& src, double* seconds = nullptr);

struct JitStats {
    uint64_t compiled = 0;     // sources compiled
    uint64_t hits = 0;         // launches of an already loaded kernel
    double compile_s = 0.0;    // host seconds spent in hipRTC
};

class JitCache {
public:
    static JitCache& get();
    // The loaded kernel for this source on the current device, compiling it
    // first when `compile` is set; null when it is not compiled and `compile`
    // is not set.
    hipFunction_t function(const std::string& src, bool compile);
    // Compile every source not yet compiled, on parallel host threads.
    void compile_all(const std::vector<std::string>& srcs);
    bool has(const std::string& src);
    JitStats stats();

private:
    JitCache() = default;
    struct Impl;
    Impl* impl();
};

}  // namespace q1t
