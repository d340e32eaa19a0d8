// Device side of the tile sweeps: the in-tile op bodies, the fused MFMA
// blocks and the tile load / store.  Shared by the compiled interpreter
// kernels (sweep.hip) and the sweep kernels specialised at run time
// (jit.cpp), which include this header as text.
#pragma once

#include "kernels.h"
#include "sweep_types.h"

namespace q1t {

struct SweepArgs {
    uint64_t ntiles;
    PosList tset;        // global positions of the tile bits (ascending)
    int tpos[16];        // tpos[j] = global position of local bit j
    const SwPhase* phases;
    const SwOp* ops;
    const float* blocks;   // real-embedded block matrices (fused segments)
    int nphases;
};

// p[i] read through the constant address space: wave-uniform scalar loads
// even where the kernel writes global memory (stores, DMA) before the read.
template <typename T> __device__ __forceinline__ T load_uniform(const T* p, uint32_t i)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) T*)(p))[i];
#else
    return p[i];
#endif
}

// Complex entry k of an op's values in the state's precision (complex64
// reads the copy rounded on the host: no conversions in the op loop).
template <typename V> __device__ __forceinline__ V opm(const SwOp& op, int k);
template <> __device__ __forceinline__ double2 opm<double2>(const SwOp& op, int k)
{
    return mkc<double2>(op.m[2 * k], op.m[2 * k + 1]);
}
template <> __device__ __forceinline__ float2 opm<float2>(const SwOp& op, int k)
{
    return mkc<float2>(op.mf[2 * k], op.mf[2 * k + 1]);
}

// Makes every field the op bodies read live at this point, so the op's
// scalar loads issue together and the op waits once for them (else the mask
// test's load is waited for first and the rest loaded after the branch).
template <typename V> __device__ __forceinline__ void pin_op(const SwOp& op)
{
    asm volatile("" ::"s"(op.type), "s"(op.reg), "s"(op.dreg), "s"(op.dtid), "s"(op.dglobal), "s"(op.cr),
                 "s"(op.ct), "s"(op.aux), "s"(op.co));
    if constexpr (sizeof(V) == 8)
        asm volatile("" ::"s"(op.mf[0]), "s"(op.mf[1]), "s"(op.mf[2]), "s"(op.mf[3]), "s"(op.mf[4]), "s"(op.mf[5]),
                     "s"(op.mf[6]), "s"(op.mf[7]));
    else
        asm volatile("" ::"s"(op.m[0]), "s"(op.m[1]), "s"(op.m[2]), "s"(op.m[3]), "s"(op.m[4]), "s"(op.m[5]),
                     "s"(op.m[6]), "s"(op.m[7]));
}

// LDS slot of local element l: XOR-fold higher bits into the low SW bits so
// strided phase exchanges spread over the banks (an involution).
template <int SW, int M> __device__ __forceinline__ int swz(int l)
{
    return swizzle_slot(SW, M, l);
}

// LDS slot of the thread's tile element tid + k * 2^(M-R) is
// swz(tid) ^ swz(k * 2^(M-R)) (swz is XOR-linear; the second term is a
// constant).  The lane index passes through an empty asm statement, so
// swz(tid) is recomputed at each tile load and store (a few ALU ops) instead
// of the compiler hoisting all 2^R slots out of the tile loop, where they
// would hold 2^R registers for the kernel's lifetime and lower its occupancy.
template <int SW, int M> __device__ __forceinline__ int tile_slot_base()
{
    int t = int(threadIdx.x);
    asm volatile("" : "+v"(t));
    return swz<SW, M>(t);
}

// Tile I/O with the non-temporal hint: every amplitude of a sweep is read
// once and written once, so it need not stay in the caches.
template <typename V> __device__ __forceinline__ V load_stream(const V* p)
{
    typedef typename RealOf<V>::type Rl;
    typedef Rl Rl2 __attribute__((ext_vector_type(2)));
    const Rl2 x = __builtin_nontemporal_load(reinterpret_cast<const Rl2*>(p));
    return mkc<V>(x.x, x.y);
}
template <typename V> __device__ __forceinline__ void store_stream(V* p, V v)
{
    typedef typename RealOf<V>::type Rl;
    typedef Rl Rl2 __attribute__((ext_vector_type(2)));
    Rl2 x;
    x.x = v.x;
    x.y = v.y;
    __builtin_nontemporal_store(x, reinterpret_cast<Rl2*>(p));
}

// c ? a : b per component (a conditional on whole vector structs selects
// between their addresses, which keeps register arrays out of registers).
template <typename V> __device__ __forceinline__ V sel(bool c, V a, V b)
{
    return mkc<V>(c ? a.x : b.x, c ? a.y : b.y);
}

// One-qubit update of the register pairs (k, k | 2^J) whose register
// controls `cr` are set; with TC the lane's thread controls decide (`tok`).
// FORM picks the body from the matrix structure, as the reference's per-gate
// apply_mat_slice bodies do: general complex, real (H, RY), swap (X, CX),
// [[a, ib], [ib, a]] (RX) and [[0, -i], [i, 0]] (Y).
// CB >= 0: the one register control is register bit CB, known at compile
// time (straight-line code; a run-time mask costs a branch per pair).
template <typename V, int NR, int J, int FORM, bool TC, bool RC, int CB = -1>
__device__ __forceinline__ void dense_pairs(V (&v)[NR], uint32_t cr, bool tok, V m0, V m1, V m2, V m3)
{
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if ((k >> J) & 1) continue;
        if (CB >= 0 && !((k >> CB) & 1)) continue;
        if (RC && (uint32_t(k) & cr) != cr) continue;
        const int k1 = k | (1 << J);
        const V x = v[k], y = v[k1];
        V nx, ny;
        if constexpr (FORM == SWF_GENERAL) {
            nx = cmac(cmul(m0, x), m1, y);
            ny = cmac(cmul(m2, x), m3, y);
        } else if constexpr (FORM == SWF_REAL) {
            nx = mkc<V>(fmad(m0.x, x.x, m1.x * y.x), fmad(m0.x, x.y, m1.x * y.y));
            ny = mkc<V>(fmad(m2.x, x.x, m3.x * y.x), fmad(m2.x, x.y, m3.x * y.y));
        } else if constexpr (FORM == SWF_SWAP) {
            nx = y;
            ny = x;
        } else if constexpr (FORM == SWF_RX) {
            const auto a = m0.x, b = m1.y;
            nx = mkc<V>(fmad(a, x.x, -(b * y.y)), fmad(a, x.y, b * y.x));
            ny = mkc<V>(fmad(a, y.x, -(b * x.y)), fmad(a, y.y, b * x.x));
        } else {  // SWF_Y
            nx = mkc<V>(y.y, -y.x);
            ny = mkc<V>(-x.y, x.x);
        }
        if constexpr (TC) {
            v[k] = sel(tok, nx, x);
            v[k1] = sel(tok, ny, y);
        } else {
            v[k] = nx;
            v[k1] = ny;
        }
    }
}

template <typename V, int NR, int R, int FORM, bool TC, bool RC, int CB = -1>
__device__ __forceinline__ void dense_reg(int reg, V (&v)[NR], uint32_t cr, bool tok, V m0, V m1, V m2, V m3)
{
    switch (reg) {
    case 0: dense_pairs<V, NR, 0, FORM, TC, RC, CB>(v, cr, tok, m0, m1, m2, m3); break;
    case 1: dense_pairs<V, NR, 1, FORM, TC, RC, CB>(v, cr, tok, m0, m1, m2, m3); break;
    case 2: dense_pairs<V, NR, 2, FORM, TC, RC, CB>(v, cr, tok, m0, m1, m2, m3); break;
    case 3: dense_pairs<V, NR, 3, FORM, TC, RC, CB>(v, cr, tok, m0, m1, m2, m3); break;
    default:
        if constexpr (R > 4) dense_pairs<V, NR, (R > 4 ? 4 : 0), FORM, TC, RC, CB>(v, cr, tok, m0, m1, m2, m3);
        break;
    }
}

// Swap-form ops (X, CX, Toffoli targets) with exactly one register control:
// the control's register index picked at compile time.
template <typename V, int NR, int R, bool TC>
__device__ __forceinline__ void swap_one_rc(int cb, int reg, V (&v)[NR], bool tok)
{
    const V z = mkc<V>(0, 0);
    switch (cb) {
    case 0: dense_reg<V, NR, R, SWF_SWAP, TC, false, 0>(reg, v, 0u, tok, z, z, z, z); break;
    case 1: dense_reg<V, NR, R, SWF_SWAP, TC, false, 1>(reg, v, 0u, tok, z, z, z, z); break;
    case 2: dense_reg<V, NR, R, SWF_SWAP, TC, false, 2>(reg, v, 0u, tok, z, z, z, z); break;
    default:
        if constexpr (R > 3) dense_reg<V, NR, R, SWF_SWAP, TC, false, (R > 3 ? 3 : 0)>(reg, v, 0u, tok, z, z, z, z);
        break;
    }
}

// Controls by where they live: thread bits (TC, a per-lane predicate) and
// register bits (RC, a per-slot predicate); each combination has its own
// straight-line body, so the common uncontrolled op carries no selects.
template <typename V, int NR, int R, int FORM>
__device__ __forceinline__ void dense_ctl(int reg, V (&v)[NR], uint32_t cr, uint32_t ct, bool tok, V m0, V m1, V m2,
                                          V m3)
{
    if constexpr (FORM == SWF_SWAP) {
        if (cr != 0 && (cr & (cr - 1)) == 0) {
            const int cb = __builtin_ctz(cr);
            if (ct) swap_one_rc<V, NR, R, true>(cb, reg, v, tok);
            else swap_one_rc<V, NR, R, false>(cb, reg, v, tok);
            return;
        }
    }
    if (ct) {
        if (cr) dense_reg<V, NR, R, FORM, true, true>(reg, v, cr, tok, m0, m1, m2, m3);
        else dense_reg<V, NR, R, FORM, true, false>(reg, v, cr, tok, m0, m1, m2, m3);
    } else {
        if (cr) dense_reg<V, NR, R, FORM, false, true>(reg, v, cr, tok, m0, m1, m2, m3);
        else dense_reg<V, NR, R, FORM, false, false>(reg, v, cr, tok, m0, m1, m2, m3);
    }
}

template <typename V, int NR, int R>
__device__ __forceinline__ void dense_form_dispatch(uint32_t form, int reg, V (&v)[NR], uint32_t cr, uint32_t ct,
                                                    bool tok, V m0, V m1, V m2, V m3)
{
    switch (form) {
    case SWF_REAL: dense_ctl<V, NR, R, SWF_REAL>(reg, v, cr, ct, tok, m0, m1, m2, m3); break;
    case SWF_SWAP: dense_ctl<V, NR, R, SWF_SWAP>(reg, v, cr, ct, tok, m0, m1, m2, m3); break;
    case SWF_RX: dense_ctl<V, NR, R, SWF_RX>(reg, v, cr, ct, tok, m0, m1, m2, m3); break;
    case SWF_Y: dense_ctl<V, NR, R, SWF_Y>(reg, v, cr, ct, tok, m0, m1, m2, m3); break;
    default: dense_ctl<V, NR, R, SWF_GENERAL>(reg, v, cr, ct, tok, m0, m1, m2, m3); break;
    }
}

// Multiply the registers whose register controls `cr` are set (and, with TC,
// whose lane passes `tok`) by the value f(k, v[k]) gives.
template <typename V, int NR, bool TC, bool RC, typename F, int CB = -1>
__device__ __forceinline__ void scale_regs_t(V (&v)[NR], uint32_t cr, bool tok, F f)
{
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (CB >= 0 && !((k >> CB) & 1)) continue;
        if (RC && (uint32_t(k) & cr) != cr) continue;
        const V nv = f(k, v[k]);
        if constexpr (TC) v[k] = sel(tok, nv, v[k]);
        else v[k] = nv;
    }
}

// The registers with register bit CB set (CB known at compile time).
template <typename V, int NR, typename F>
__device__ __forceinline__ void scale_regs_one_rc(int cb, V (&v)[NR], uint32_t ct, bool tok, F f)
{
#define Q1T_SCALE_CB(C)                                                     \
    if (ct) scale_regs_t<V, NR, true, false, F, C>(v, 0u, tok, f);          \
    else scale_regs_t<V, NR, false, false, F, C>(v, 0u, tok, f)
    switch (cb) {
    case 0: Q1T_SCALE_CB(0); break;
    case 1: Q1T_SCALE_CB(1); break;
    case 2: Q1T_SCALE_CB(2); break;
    default:
        if constexpr (NR > 8) Q1T_SCALE_CB((NR > 8 ? 3 : 0));
        break;
    }
#undef Q1T_SCALE_CB
}

template <typename V, int NR, typename F>
__device__ __forceinline__ void scale_regs(V (&v)[NR], uint32_t cr, uint32_t ct, bool tok, F f)
{
    if (cr != 0 && (cr & (cr - 1)) == 0) {
        scale_regs_one_rc<V, NR>(__builtin_ctz(cr), v, ct, tok, f);
        return;
    }
    if (ct) {
        if (cr) scale_regs_t<V, NR, true, true>(v, cr, tok, f);
        else scale_regs_t<V, NR, true, false>(v, cr, tok, f);
    } else {
        if (cr) scale_regs_t<V, NR, false, true>(v, cr, tok, f);
        else scale_regs_t<V, NR, false, false>(v, cr, tok, f);
    }
}

template <typename V, int NR>
__device__ __forceinline__ void phase_op(uint32_t form, V (&v)[NR], uint32_t cr, uint32_t ct, bool tok, V ph1)
{
    switch (form) {
    case SWP_I: scale_regs<V, NR>(v, cr, ct, tok, [](int, V x) { return mkc<V>(-x.y, x.x); }); break;
    case SWP_MINUS_I: scale_regs<V, NR>(v, cr, ct, tok, [](int, V x) { return mkc<V>(x.y, -x.x); }); break;
    case SWP_NEG: scale_regs<V, NR>(v, cr, ct, tok, [](int, V x) { return mkc<V>(-x.x, -x.y); }); break;
    default:
        if (ct) {
            // thread controls: one per-lane factor (1 or the phase), then an
            // unconditional product, instead of a select per register
            const V f = sel(tok, ph1, mkc<V>(1, 0));
            scale_regs<V, NR>(v, cr, 0u, tok, [&](int, V x) { return cmul(f, x); });
        } else {
            scale_regs<V, NR>(v, cr, 0u, tok, [&](int, V x) { return cmul(ph1, x); });
        }
        break;
    }
}

template <typename V, int NR>
__device__ __forceinline__ void diag_op(const SwOp& op, V (&v)[NR], uint32_t cr, uint32_t ct, bool tok, uint64_t base,
                                        int tid, V d0, V d1)
{
    if (op.dreg >= 0) {
        const int dr = op.dreg;
        scale_regs<V, NR>(v, cr, ct, tok, [&](int k, V x) { return cmul(sel(((k >> dr) & 1) != 0, d1, d0), x); });
    } else {
        const bool one = op.dtid >= 0 ? ((tid >> op.dtid) & 1) : ((base >> op.dglobal) & 1);
        const V d = sel(one, d1, d0);
        scale_regs<V, NR>(v, cr, ct, tok, [&](int, V x) { return cmul(d, x); });
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Dense K-qubit block Y = U X on the complex64 tile in LDS with FP32 MFMA
// (v_mfma_f32_16x16x4f32, FP32 accumulation).  X is the tile viewed as a
// D x C matrix (rows: the K block bits, columns: the other M-K local bits);
// the complex product is the real GEMM [Yr; Yi] = [[Ur, -Ui], [Ui, Ur]] [Xr; Xi]
// with the 2D x 2D real matrix A in global memory.  Each wave owns 16-column
// chunks, so the in-place update needs no barrier until the block is done.
// Operand layouts: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D[i][j] in lane j + 16(i/4), element i%4.
template <int K, int M, int SW>
__device__ __forceinline__ void block_mfma(float2* lds, const float* __restrict__ A, const SwPhase& ph, int tid)
{
    constexpr int D = 1 << K, D2 = 2 * D, C = 1 << (M - K);
    constexpr int NS = D2 / 4, NRT = D2 / 16, NH = NS / 2;
    const int lane = tid & 63, wave = tid >> 6, nwaves = int(blockDim.x) >> 6;
    const int li = lane & 15, lg = lane >> 4;
    // A fragments: the host stores blocks of 3 and 4 qubits in fragment
    // order, lane's NS values of row tile rt contiguous (16-byte loads)
    float a[NRT][NS];
    const float4* Af = reinterpret_cast<const float4*>(A);
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int q = 0; q < NS / 4; ++q) {
            const float4 t = Af[(rt * 64 + lane) * (NS / 4) + q];
            a[rt][4 * q] = t.x;
            a[rt][4 * q + 1] = t.y;
            a[rt][4 * q + 2] = t.z;
            a[rt][4 * q + 3] = t.w;
        }
    uint32_t rbit[K];
#pragma unroll
    for (int j = 0; j < K; ++j) rbit[j] = 1u << ph.rpos[j];
    // local index bits of block row r (row bit K-1-j is local bit rpos[j])
    auto rowpart = [&](int r) {
        uint32_t p = 0;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if ((r >> (K - 1 - j)) & 1) p |= rbit[j];
        return p;
    };
    auto colpart = [&](int c) {
        uint32_t p = 0;
#pragma unroll
        for (int b = 0; b < M - K; ++b)
            if ((c >> b) & 1) p |= 1u << ph.tbit[b];
        return p;
    };
    // B fragment of one 16-column chunk: MFMA step s reads block row 4s+lg;
    // rows D.. are the imaginary parts of rows 0.., so step s and step s+NH
    // share one complex amplitude (one LDS read for both)
    auto load_b = [&](uint32_t cp, float (&bv)[NS]) {
#pragma unroll
        for (int s = 0; s < NH; ++s) {
            const float2 v = lds[swz<SW, M>(int(rowpart(4 * s + lg) | cp))];
            bv[s] = v.x;
            bv[s + NH] = v.y;
        }
    };
    auto store = [&](uint32_t cp, const f32x4 (&acc)[NRT]) {
        if constexpr (NRT == 2) {
            // rows 4lg+v (real) and 16+4lg+v (imaginary) of one amplitude
#pragma unroll
            for (int v = 0; v < 4; ++v)
                lds[swz<SW, M>(int(rowpart(4 * lg + v) | cp))] = make_float2(acc[0][v], acc[1][v]);
        } else {
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rr = 16 * rt + 4 * lg + v;
                    float* dst = reinterpret_cast<float*>(&lds[swz<SW, M>(int(rowpart(rr & (D - 1)) | cp))]);
                    dst[rr < D ? 0 : 1] = acc[rt][v];
                }
        }
    };
    // two chunks per pass: 2 NRT independent accumulator chains
    for (int chunk = wave; chunk < C / 16; chunk += 2 * nwaves) {
        const int chunk2 = chunk + nwaves;
        const bool two = chunk2 < C / 16;
        const uint32_t cp0 = colpart(chunk * 16 + li), cp1 = colpart(chunk2 * 16 + li);
        float b0[NS], b1[NS];
        load_b(cp0, b0);
        if (two) load_b(cp1, b1);
        f32x4 acc0[NRT], acc1[NRT];
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt) {
            acc0[rt] = f32x4{0.f, 0.f, 0.f, 0.f};
            acc1[rt] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
        if (two) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt) {
                    acc0[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][s], b0[s], acc0[rt], 0, 0, 0);
                    acc1[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][s], b1[s], acc1[rt], 0, 0, 0);
                }
            store(cp0, acc0);
            store(cp1, acc1);
        } else {
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int rt = 0; rt < NRT; ++rt)
                    acc0[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][s], b0[s], acc0[rt], 0, 0, 0);
            store(cp0, acc0);
        }
    }
}

// The same block product for K = 5 or 6 (2D = 64 or 128 rows): the 2D x 2D
// real matrix (16 or 64 KiB) is too large to keep per lane, so each slice of
// 8 MFMA steps reads its A values from global memory (L1/L2-resident: every
// tile of the sweep uses the same matrix; fragment order makes them two
// 16-byte loads per row tile); B fragments and accumulators stay in registers.
template <int K, int M, int SW>
__device__ __forceinline__ void block_mfma_wide(float2* lds, const float* __restrict__ A, const SwPhase& ph, int tid)
{
    constexpr int D = 1 << K, D2 = 2 * D, C = 1 << (M - K);
    constexpr int NS = D2 / 4, NRT = D2 / 16;
    static_assert(C >= 16, "at least one 16-column chunk per tile");
    const int lane = tid & 63, wave = tid >> 6, nwaves = int(blockDim.x) >> 6;
    const int li = lane & 15, lg = lane >> 4;
    uint32_t rbit[K];
#pragma unroll
    for (int j = 0; j < K; ++j) rbit[j] = 1u << ph.rpos[j];
    auto rowpart = [&](int r) {
        uint32_t p = 0;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if ((r >> (K - 1 - j)) & 1) p |= rbit[j];
        return p;
    };
    for (int chunk = wave; chunk < C / 16; chunk += nwaves) {
        const int c = chunk * 16 + li;
        uint32_t cp = 0;
#pragma unroll
        for (int b = 0; b < M - K; ++b)
            if ((c >> b) & 1) cp |= 1u << ph.tbit[b];
        // the K dimension in slices of 8 MFMA steps: 8 B fragments and NRT x 8
        // A fragments in flight at a time
        f32x4 acc[NRT];
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt) acc[rt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int s0 = 0; s0 < NS; s0 += 8) {
            float bv[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int rr = 4 * (s0 + s) + lg;
                const float2 v = lds[swz<SW, M>(int(rowpart(rr & (D - 1)) | cp))];
                bv[s] = rr < D ? v.x : v.y;
            }
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt) {
                // fragment order: this lane's 8 values of the slice are contiguous
                const float4* af = reinterpret_cast<const float4*>(A + (rt * 64 + lane) * NS + s0);
                const float4 f0 = af[0], f1 = af[1];
                const float as[8] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w};
#pragma unroll
                for (int s = 0; s < 8; ++s)
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(as[s], bv[s], acc[rt], 0, 0, 0);
            }
        }
#pragma unroll
        for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = 16 * rt + 4 * lg + v;
                float* dst = reinterpret_cast<float*>(&lds[swz<SW, M>(int(rowpart(rr & (D - 1)) | cp))]);
                dst[rr < D ? 0 : 1] = acc[rt][v];
            }
    }
}

// One tile of a sweep: HBM -> LDS, the program's phases, LDS -> HBM.
// Global offsets of a thread's share of a tile in natural local order:
// element tid + k * 2^(M-R) sits at tile base | g_tid | g_k[k].
template <int M, int R> struct TileAddr {
    uint64_t g_tid;
    uint64_t g_k[1 << R];
};

template <int M, int R> __device__ __forceinline__ void tile_addr(const SweepArgs& sa, TileAddr<M, R>& ta)
{
    const int tid = threadIdx.x;
    ta.g_tid = 0;
#pragma unroll
    for (int b = 0; b < M - R; ++b)
        if ((tid >> b) & 1) ta.g_tid |= uint64_t(1) << sa.tpos[b];
#pragma unroll
    for (int k = 0; k < (1 << R); ++k) {
        uint64_t gk = 0;
#pragma unroll
        for (int b = 0; b < R; ++b)
            if ((k >> b) & 1) gk |= uint64_t(1) << sa.tpos[M - R + b];
        ta.g_k[k] = gk;
    }
}

// The specialised kernels' tile factors (a program's kind-2 phase; the
// interpreter kernels' programs have none):
// the product, in program order, of the factors of the tile-factor ops
// [begin, end) for the tile at `base`: a phase-on-ones op contributes its
// phase when all its (outer) bits are set, a one-target diagonal on an outer
// bit d0 or d1 by that bit, each only when its outer controls are set.
template <typename V>
__device__ __forceinline__ V tile_factor(const SwOp* ops, uint32_t begin, uint32_t end, uint64_t base)
{
    V f = mkc<V>(1, 0);
    for (uint32_t oi = begin; oi < end; ++oi) {
        const SwOp op = load_uniform(ops, oi);
        if ((base & op.co) != op.co) continue;
        if (op.type == SW_PHASE) f = cmul(f, opm<V>(op, 0));
        else f = cmul(f, sel(((base >> op.dglobal) & 1) != 0, opm<V>(op, 1), opm<V>(op, 0)));
    }
    return f;
}

// The program's phases on a tile already in LDS, then LDS -> HBM.  BLK is
// the widest fused dense block the kernel holds: 0 none (programs without
// blocks: the MFMA code's registers would halve the kernel's occupancy),
// 1 the 3- and 4-qubit blocks, 2 also the 5- and 6-qubit ones.
template <typename V, int M, int R, int SW, int BLK>
__device__ __forceinline__ void tile_phases(V* __restrict__ a, const SweepArgs& sa, uint64_t base, V* lds,
                                            const TileAddr<M, R>& ta)
{
    constexpr int NT = 1 << (M - R), NR = 1 << R;
    using Rl = typename RealOf<V>::type;
    const int tid = threadIdx.x;
    {
        for (int p = 0; p < sa.nphases; ++p) {
            // the program is read through the constant address space: scalar
            // loads, whatever else the kernel writes (including the DMA)
            const SwPhase ph = load_uniform(sa.phases, uint32_t(p));
            if constexpr (sizeof(V) == 8 && BLK > 0) {
                if (ph.kind == 1) {
                    // fused dense block: K = reg_mask, matrix at blocks + op_begin
                    const float* A = sa.blocks + ph.op_begin;
                    // the planner emits K = 3 .. 6, and K > 4 only to BLK 2 kernels
                    float2* l2 = reinterpret_cast<float2*>(lds);
                    if constexpr (BLK == 2) {
                        switch (ph.reg_mask) {
                        case 3: block_mfma<3, M, SW>(l2, A, ph, tid); break;
                        case 4: block_mfma<4, M, SW>(l2, A, ph, tid); break;
                        case 5: block_mfma_wide<5, M, SW>(l2, A, ph, tid); break;
                        default: block_mfma_wide<6, M, SW>(l2, A, ph, tid); break;
                        }
                    } else {
                        if (ph.reg_mask == 3) block_mfma<3, M, SW>(l2, A, ph, tid);
                        else block_mfma<4, M, SW>(l2, A, ph, tid);
                    }
                    __syncthreads();
                    continue;
                }
            }
            // register bits ph.rpos[0..R-1] (ascending), thread bits the
            // others: the thread's part of a local index is tid with a zero
            // inserted at each register position, lowest first
            uint32_t tpart = uint32_t(tid);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint32_t low = (1u << ph.rpos[j]) - 1u;
                tpart = (tpart & low) | ((tpart & ~low) << 1);
            }
            // the swizzle is XOR-linear and the thread and register parts
            // have disjoint bits: slot k's LDS index = swz(tpart) ^ swz(rp[k]),
            // the second term wave-uniform
            const int st = swz<SW, M>(int(tpart));
            static_assert(SW == sweep_swizzle_bits(M), "the planner's register slot terms use this swizzle");
            int srp[NR];
#pragma unroll
            for (int k = 0; k < NR; ++k) srp[k] = int(ph.rslot[k]);
            V v[NR];
#pragma unroll
            for (int k = 0; k < NR; ++k) v[k] = lds[st ^ srp[k]];
            for (uint32_t oi = ph.op_begin; oi < ph.op_end; ++oi) {
                const SwOp op = load_uniform(sa.ops, oi);
                pin_op<V>(op);
                if ((base & op.co) != op.co) continue;
                const uint32_t cr = op.cr, ct = op.ct;
                const bool tok = (uint32_t(tid) & ct) == ct;
                if (op.type == SW_DENSE1) {
                    const V m0 = opm<V>(op, 0), m1 = opm<V>(op, 1), m2 = opm<V>(op, 2), m3 = opm<V>(op, 3);
                    dense_form_dispatch<V, NR, R>(op.aux, op.reg, v, cr, ct, tok, m0, m1, m2, m3);
                } else if (op.type == SW_PHASE) {
                    phase_op<V, NR>(op.aux, v, cr, ct, tok, opm<V>(op, 0));
                } else {  // SW_DIAG1
                    const V d0 = opm<V>(op, 0), d1 = opm<V>(op, 1);
                    diag_op<V, NR>(op, v, cr, ct, tok, base, tid, d0, d1);
                }
            }
            if (ph.wmap == 0) {
#pragma unroll
                for (int k = 0; k < NR; ++k) lds[st ^ srp[k]] = v[k];
            } else {
                // SWAP / X / CX gates of local bits: slot k goes back to the
                // affine image of its index.  The linear part splits into a
                // per-thread term (the thread bits' columns) and a uniform
                // term (the register bits' columns), like the read index;
                // other threads' slots are overwritten, so every read of the
                // phase must be done first.
                uint32_t img = ph.wxor;
#pragma unroll
                for (int b = 0; b < M - R; ++b)
                    if ((tid >> b) & 1) img ^= ph.wcol[b];
                const int pst = swz<SW, M>(int(img));
                __syncthreads();
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    uint32_t l = 0;
#pragma unroll
                    for (int b = 0; b < R; ++b)
                        if ((k >> b) & 1) l ^= ph.wcol[M - R + b];
                    lds[pst ^ swz<SW, M>(int(l))] = v[k];
                }
            }
            __syncthreads();
        }

        {
            const int s0 = tile_slot_base<SW, M>();
#pragma unroll
            for (int k = 0; k < NR; ++k) store_stream(a + (base | ta.g_tid | ta.g_k[k]), lds[s0 ^ swz<SW, M>(k * NT)]);
        }
        __syncthreads();
    }
}

// One tile of a sweep: HBM -> LDS, the program's phases, LDS -> HBM.
template <typename V, int M, int R, int SW, int BLK>
__device__ __forceinline__ void sweep_tile(V* __restrict__ a, const SweepArgs& sa, uint64_t tile, V* lds)
{
    constexpr int NT = 1 << (M - R), NR = 1 << R;
    TileAddr<M, R> ta;
    tile_addr<M, R>(sa, ta);
    const uint64_t base = deposit(tile, sa.tset);
    const int s0 = tile_slot_base<SW, M>();
#pragma unroll
    for (int k = 0; k < NR; ++k) lds[s0 ^ swz<SW, M>(k * NT)] = load_stream(a + (base | ta.g_tid | ta.g_k[k]));
    __syncthreads();
    tile_phases<V, M, R, SW, BLK>(a, sa, base, lds, ta);
}

}  // namespace q1t
