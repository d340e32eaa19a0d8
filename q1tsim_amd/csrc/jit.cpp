// Run-time specialised sweep kernels: source generation, hipRTC compile and a
// per-process cache of loaded kernels.
#include "jit.h"

#include <hip/hiprtc.h>

#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <thread>
#include <unordered_map>

#include "common.h"
#include "jit_text.h"   // kJitHeaderNames / kJitHeaderTexts, written by build.py

namespace q1t {

// LDS slot of local element l (the device swz<SW, M>).
static int swz_host(int SW, int M, int l) { return swizzle_slot(SW, M, l); }

static std::string hex64(uint64_t v)
{
    std::ostringstream o;
    o << "0x" << std::hex << v << "ull";
    return o.str();
}

// Straight-line code of one register phase: the reads, every op with its
// structure as constants (the same bodies tile_phases dispatches to), the
// write-back.
static void emit_register_phase(std::ostringstream& o, const SwPhase& ph, const std::vector<SwOp>& ops, int M, int R,
                                int SW)
{
    const int NR = 1 << R;
    o << "    {\n        uint32_t tpart = uint32_t(tid);\n";
    for (int j = 0; j < R; ++j) {
        const uint32_t low = (1u << ph.rpos[j]) - 1u;
        o << "        tpart = (tpart & " << low << "u) | ((tpart & " << (~low) << "u) << 1);\n";
    }
    // a phase without ops (a pure write-back map) holds its values in
    // scalars: an array that is only copied is not promoted to registers
    const bool noops = ph.op_begin == ph.op_end;
    const std::string vname = noops ? "v_" : "v[";
    const std::string vend = noops ? "" : "]";
    o << "        const int st = swz<SW, M>(int(tpart));\n";
    if (!noops) o << "        V v[NR];\n";
    int srp[16];
    for (int k = 0; k < NR; ++k) {
        uint32_t l = 0;
        for (int b = 0; b < R; ++b)
            if ((k >> b) & 1) l |= 1u << ph.rpos[b];
        srp[k] = swz_host(SW, M, int(l));
        o << "        " << (noops ? "const V " : "") << vname << k << vend << " = lds[st ^ " << srp[k] << "];\n";
    }
    o << "        __builtin_amdgcn_s_setprio(0);\n";
    for (uint32_t oi = ph.op_begin; oi < ph.op_end; ++oi) {
        const SwOp& op = ops[oi];
        o << "        {\n";
        std::string ind = "            ";
        if (op.co) {
            o << "            if ((base & " << hex64(op.co) << ") == " << hex64(op.co) << ") {\n";
            ind += "    ";
        }
        if (op.ct) o << ind << "const bool tok = (uint32_t(tid) & " << op.ct << "u) == " << op.ct << "u;\n";
        else o << ind << "const bool tok = true;\n";
        o << ind << "const SwOp op = load_uniform(ops, " << oi << "u);\n";
        const std::string TC = op.ct ? "true" : "false";
        if (op.type == SW_DENSE1) {
            const uint32_t form = op.aux;
            const bool single_rc = op.cr != 0 && (op.cr & (op.cr - 1)) == 0;
            if (form == SWF_SWAP && single_rc) {
                const int cb = __builtin_ctz(op.cr);
                o << ind << "const V z = mkc<V>(0, 0);\n";
                o << ind << "dense_pairs<V, NR, " << op.reg << ", SWF_SWAP, " << TC << ", false, " << cb
                  << ">(v, 0u, tok, z, z, z, z);\n";
            } else {
                const std::string RC = op.cr ? "true" : "false";
                if (form == SWF_SWAP || form == SWF_Y) o << ind << "const V m0 = mkc<V>(0, 0), m1 = m0, m2 = m0, m3 = m0;\n";
                else
                    o << ind
                      << "const V m0 = opm<V>(op, 0), m1 = opm<V>(op, 1), m2 = opm<V>(op, 2), m3 = opm<V>(op, 3);\n";
                o << ind << "dense_pairs<V, NR, " << op.reg << ", " << form << ", " << TC << ", " << RC << ">(v, "
                  << op.cr << "u, tok, m0, m1, m2, m3);\n";
            }
        } else if (op.type == SW_PHASE) {
            o << ind << "phase_op<V, NR>(" << op.aux << "u, v, " << op.cr << "u, " << op.ct
              << "u, tok, opm<V>(op, 0));\n";
        } else {   // SW_DIAG1
            o << ind << "const V d0 = opm<V>(op, 0), d1 = opm<V>(op, 1);\n";
            if (op.dreg >= 0) {
                o << ind << "scale_regs<V, NR>(v, " << op.cr << "u, " << op.ct
                  << "u, tok, [&](int k, V x) { return cmul(sel(((k >> " << op.dreg
                  << ") & 1) != 0, d1, d0), x); });\n";
            } else {
                if (op.dtid >= 0) o << ind << "const bool one = ((tid >> " << op.dtid << ") & 1) != 0;\n";
                else o << ind << "const bool one = ((base >> " << op.dglobal << ") & 1) != 0;\n";
                o << ind << "const V d = sel(one, d1, d0);\n";
                o << ind << "scale_regs<V, NR>(v, " << op.cr << "u, " << op.ct
                  << "u, tok, [&](int, V x) { return cmul(d, x); });\n";
            }
        }
        if (op.co) o << "            }\n";
        o << "        }\n";
    }
    // the phase's LDS exchange (write-back, barrier, the next phase's reads)
    // at priority 1: between the tile I/O (2) and the arithmetic (0)
    o << "        __builtin_amdgcn_s_setprio(1);\n";
    if (ph.wmap == 0) {
        for (int k = 0; k < NR; ++k) o << "        lds[st ^ " << srp[k] << "] = " << vname << k << vend << ";\n";
    } else {
        o << "        uint32_t img = " << ph.wxor << "u;\n";
        for (int b = 0; b < M - R; ++b)
            if (ph.wcol[b]) o << "        if ((tid >> " << b << ") & 1) img ^= " << ph.wcol[b] << "u;\n";
        o << "        const int pst = swz<SW, M>(int(img));\n        __syncthreads();\n";
        for (int k = 0; k < NR; ++k) {
            uint32_t l = 0;
            for (int b = 0; b < R; ++b)
                if ((k >> b) & 1) l ^= ph.wcol[M - R + b];
            o << "        lds[pst ^ " << swz_host(SW, M, int(l)) << "] = " << vname << k << vend << ";\n";
        }
    }
    o << "        __syncthreads();\n    }\n";
}

static void emit_block_phase(std::ostringstream& o, const SwPhase& ph, int M)
{
    const int K = int(ph.reg_mask);
    o << "    {\n        SwPhase ph = {};\n";
    for (int j = 0; j < K; ++j) o << "        ph.rpos[" << j << "] = " << int(ph.rpos[j]) << ";\n";
    for (int b = 0; b < M - K; ++b) o << "        ph.tbit[" << b << "] = " << int(ph.tbit[b]) << ";\n";
    o << "        float2* l2 = reinterpret_cast<float2*>(lds);\n";
    if (K <= 4) o << "        block_mfma<" << K << ", M, SW>(l2, blocks + " << ph.op_begin << "u, ph, tid);\n";
    else o << "        block_mfma_wide<" << K << ", M, SW>(l2, blocks + " << ph.op_begin << "u, ph, tid);\n";
    o << "        __syncthreads();\n    }\n";
}

const char* jit_kernel_name(bool c64) { return c64 ? "q1t_jit_sweep_c64" : "q1t_jit_sweep_c128"; }

std::string jit_sweep_source(const std::vector<SwPhase>& phases, const std::vector<SwOp>& ops, bool c64, int M,
                             int R, int SW, bool prefetch)
{
    std::ostringstream o;
    const int NT = 1 << (M - R);
    // occupancy as the interpreter kernels: two 64 KiB workgroups per CU for
    // M = 12 (four waves per SIMD at R = 3), one for M = 13
    const int waves = M == 12 ? (R == 3 ? 4 : 2) : 2;
    o << "#include \"sweep_dev.h\"\nusing namespace q1t;\n";
    o << "extern \"C\" __global__ void __launch_bounds__(" << NT << ") __attribute__((amdgpu_waves_per_eu(" << waves
      << ")))\n" << jit_kernel_name(c64) << "(" << (c64 ? "float2" : "double2")
      << "* __restrict__ amp, uint64_t col_stride, SweepArgs sa)\n{\n";
    o << "    typedef " << (c64 ? "float2" : "double2") << " V;\n";
    o << "    constexpr int M = " << M << ", R = " << R << ", SW = " << SW << ";\n";
    o << "    constexpr int NT = 1 << (M - R), NR = 1 << R;\n";
    o << "    __shared__ V lds[1 << M];\n";
    o << "    V* a = amp + uint64_t(blockIdx.y) * col_stride;\n";
    o << "    TileAddr<M, R> ta;\n    tile_addr<M, R>(sa, ta);\n";
    if (prefetch) {
        // the first tile's share in registers before the loop; every
        // iteration stores the prefetched share to LDS, then issues the loads
        // of the next tile (in flight while the phases run)
        // (scalars, not an array: a loop-carried array of vector structs
        // is not promoted to registers)
        const int NR = 1 << R;
        for (int k = 0; k < NR; ++k) o << "    V pf" << k << ";\n";
        o << "    if (uint64_t(blockIdx.x) < sa.ntiles) {\n";
        o << "        const uint64_t b0 = deposit(uint64_t(blockIdx.x), sa.tset);\n";
        for (int k = 0; k < NR; ++k) o << "        pf" << k << " = a[b0 | ta.g_tid | ta.g_k[" << k << "]];\n";
        o << "    }\n";
    }
    // A wave moving its tile between HBM and LDS (the store of one tile, the
    // loads of the next) runs at a raised priority, the phases at the base
    // one: with two workgroups per CU, the one at its tile boundary gets its
    // memory instructions issued ahead of the other's arithmetic, which keeps
    // HBM busier (30-qubit layered circuit: c128 135 -> 128 ms, c64 83 -> 80 ms,
    // and another 1 % with the LDS exchanges at an intermediate priority; the
    // interpreter kernels measured no gain and are left as they are).
    o << "    __builtin_amdgcn_s_setprio(2);\n";
    o << "    for (uint64_t tile = blockIdx.x; tile < sa.ntiles; tile += gridDim.x) {\n";
    // The thread index and the table pointers are made opaque per tile:
    // otherwise every phase's LDS addresses (functions of the thread index
    // alone) and every op's matrix (tile invariant) are hoisted out of the
    // tile loop and held in registers across it, which spills.
    o << "    int tid = int(threadIdx.x);\n    asm volatile(\"\" : \"+v\"(tid));\n";
    o << "    const SwOp* ops = sa.ops;\n    asm volatile(\"\" : \"+s\"(ops));\n";
    o << "    const float* blocks = sa.blocks;\n    asm volatile(\"\" : \"+s\"(blocks));\n";
    o << "    const uint64_t base = deposit(tile, sa.tset);\n";
    if (prefetch) {
        const int NR = 1 << R;
        for (int k = 0; k < NR; ++k) o << "    lds[swz<SW, M>(tid + " << k << " * NT)] = pf" << k << ";\n";
        o << "    __syncthreads();\n";
        o << "    if (tile + gridDim.x < sa.ntiles) {\n";
        o << "        const uint64_t nb = deposit(tile + gridDim.x, sa.tset);\n";
        for (int k = 0; k < NR; ++k) o << "        pf" << k << " = a[nb | ta.g_tid | ta.g_k[" << k << "]];\n";
        o << "    }\n";
    } else {
        o << "#pragma unroll\n    for (int k = 0; k < NR; ++k) lds[swz<SW, M>(tid + k * NT)] = load_stream(a + (base | ta.g_tid | ta.g_k[k]));\n";
        o << "    __syncthreads();\n";
    }
    o << "    __builtin_amdgcn_s_setprio(0);\n";
    bool scaled = false;
    for (const SwPhase& ph : phases) {
        if (ph.kind == 1) {
            emit_block_phase(o, ph, M);
        } else if (ph.kind == 2) {
            // the tile factor: the product of tile_factor() in program order
            scaled = true;
            o << "    V fac = mkc<V>(1, 0);\n";
            for (uint32_t oi = ph.op_begin; oi < ph.op_end; ++oi) {
                const SwOp& op = ops[oi];
                o << "    {\n";
                if (op.co) o << "    if ((base & " << hex64(op.co) << ") == " << hex64(op.co) << ") {\n";
                o << "        const SwOp op = load_uniform(ops, " << oi << "u);\n";
                if (op.type == SW_PHASE) o << "        fac = cmul(fac, opm<V>(op, 0));\n";
                else
                    o << "        fac = cmul(fac, sel(((base >> " << op.dglobal
                      << ") & 1) != 0, opm<V>(op, 1), opm<V>(op, 0)));\n";
                if (op.co) o << "    }\n";
                o << "    }\n";
            }
        } else {
            emit_register_phase(o, ph, ops, M, R, SW);
        }
    }
    o << "    __builtin_amdgcn_s_setprio(2);\n";
    if (scaled)
        o << "#pragma unroll\n    for (int k = 0; k < NR; ++k) store_stream(a + (base | ta.g_tid | ta.g_k[k]), cmul(fac, lds[swz<SW, M>(tid + k * NT)]));\n";
    else
        o << "#pragma unroll\n    for (int k = 0; k < NR; ++k) store_stream(a + (base | ta.g_tid | ta.g_k[k]), lds[swz<SW, M>(tid + k * NT)]);\n";
    o << "    __syncthreads();\n    }\n}\n";
    return o.str();
}

#define RTC_OK(expr)                                                                               \
    do {                                                                                           \
        hiprtcResult r_ = (expr);                                                                  \
        if (r_ != HIPRTC_SUCCESS)                                                                  \
            throw Error(ERR_HIP, std::string("hipRTC error in " #expr ": ") + hiprtcGetErrorString(r_)); \
    } while (0)

std::vector<char> jit_compile(const std::string& src, double* seconds)
{
    const auto t0 = std::chrono::steady_clock::now();
    hiprtcProgram prog;
    const int nh = int(sizeof(kJitHeaderNames) / sizeof(kJitHeaderNames[0]));
    RTC_OK(hiprtcCreateProgram(&prog, src.c_str(), "q1t_jit_sweep.hip", nh, kJitHeaderTexts, kJitHeaderNames));
    // the options the library's sweep kernels are built with (build.py):
    // packed FP32 off, no contraction beyond the spelled-out fused
    // multiply-adds, no sinking of common code out of switch cases, uniform
    // regions left unstructurized
    const char* opts[] = {"--offload-arch=gfx950",
                          "-O3",
                          "-std=c++17",
                          "-ffp-contract=off",
                          "-Xclang",
                          "-target-feature",
                          "-Xclang",
                          "-packed-fp32-ops",
                          "-mllvm",
                          "-simplifycfg-sink-common=false",
                          "-mllvm",
                          "-structurizecfg-skip-uniform-regions"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, int(sizeof(opts) / sizeof(opts[0])), opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t ls = 0;
        std::string log;
        if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
            log.resize(ls);
            (void)hiprtcGetProgramLog(prog, &log[0]);
        }
        (void)hiprtcDestroyProgram(&prog);
        throw Error(ERR_HIP, std::string("hipRTC compile of a sweep kernel failed: ") + hiprtcGetErrorString(rc) +
                                 "\n" + log.substr(0, 4000));
    }
    size_t n = 0;
    RTC_OK(hiprtcGetCodeSize(prog, &n));
    std::vector<char> code(n);
    RTC_OK(hiprtcGetCode(prog, code.data()));
    RTC_OK(hiprtcDestroyProgram(&prog));
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return code;
}

struct JitCache::Impl {
    std::mutex mu;
    struct Entry {
        std::vector<char> code;                 // empty: not compiled
        std::map<int, hipModule_t> mod;         // per device
        std::map<int, hipFunction_t> fn;
    };
    std::unordered_map<std::string, std::shared_ptr<Entry>> map;
    JitStats st;
};

JitCache& JitCache::get()
{
    // never destroyed: loaded modules stay valid until the process ends
    static JitCache* c = new JitCache();
    return *c;
}

JitCache::Impl* JitCache::impl()
{
    static Impl* p = new Impl();
    return p;
}

bool JitCache::has(const std::string& src)
{
    Impl* I = impl();
    std::lock_guard<std::mutex> lk(I->mu);
    auto it = I->map.find(src);
    return it != I->map.end() && !it->second->code.empty();
}

JitStats JitCache::stats()
{
    Impl* I = impl();
    std::lock_guard<std::mutex> lk(I->mu);
    return I->st;
}

void JitCache::compile_all(const std::vector<std::string>& srcs)
{
    Impl* I = impl();
    std::vector<const std::string*> todo;
    {
        std::lock_guard<std::mutex> lk(I->mu);
        std::map<std::string, int> seen;
        for (const std::string& s : srcs) {
            auto it = I->map.find(s);
            if ((it == I->map.end() || it->second->code.empty()) && !seen.count(s)) {
                seen[s] = 1;
                todo.push_back(&s);
            }
        }
    }
    if (todo.empty()) return;
    std::vector<std::vector<char>> codes(todo.size());
    std::vector<double> secs(todo.size(), 0.0);
    std::vector<std::string> errs(todo.size());
    const size_t nth = std::min<size_t>(todo.size(), std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    std::mutex qm;
    size_t next = 0;
    for (size_t t = 0; t < nth; ++t)
        pool.emplace_back([&] {
            for (;;) {
                size_t i;
                {
                    std::lock_guard<std::mutex> lk(qm);
                    if (next >= todo.size()) return;
                    i = next++;
                }
                try {
                    codes[i] = jit_compile(*todo[i], &secs[i]);
                } catch (const std::exception& e) {
                    errs[i] = e.what();
                }
            }
        });
    for (auto& th : pool) th.join();
    for (const std::string& e : errs)
        if (!e.empty()) throw Error(ERR_HIP, e);
    std::lock_guard<std::mutex> lk(I->mu);
    for (size_t i = 0; i < todo.size(); ++i) {
        auto& e = I->map[*todo[i]];
        if (!e) e = std::make_shared<Impl::Entry>();
        if (e->code.empty()) {
            e->code = std::move(codes[i]);
            I->st.compiled += 1;
            I->st.compile_s += secs[i];
        }
    }
}

hipFunction_t JitCache::function(const std::string& src, bool compile)
{
    Impl* I = impl();
    if (compile && !has(src)) compile_all({src});
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw Error(ERR_HIP, "hipGetDevice failed");
    std::lock_guard<std::mutex> lk(I->mu);
    auto it = I->map.find(src);
    if (it == I->map.end() || it->second->code.empty()) return nullptr;
    Impl::Entry& e = *it->second;
    auto f = e.fn.find(dev);
    if (f != e.fn.end()) {
        I->st.hits += 1;
        return f->second;
    }
    hipModule_t m = nullptr;
    hipError_t he = hipModuleLoadData(&m, e.code.data());
    if (he != hipSuccess) {
        (void)hipGetLastError();
        throw Error(ERR_HIP, std::string("hipModuleLoadData of a sweep kernel failed: ") + hipGetErrorString(he));
    }
    hipFunction_t fn = nullptr;
    const char* name = src.find(jit_kernel_name(true)) != std::string::npos ? jit_kernel_name(true) : jit_kernel_name(false);
    he = hipModuleGetFunction(&fn, m, name);
    if (he != hipSuccess) {
        (void)hipGetLastError();
        throw Error(ERR_HIP, std::string("hipModuleGetFunction failed: ") + hipGetErrorString(he));
    }
    e.mod[dev] = m;
    e.fn[dev] = fn;
    return fn;
}

}  // namespace q1t
