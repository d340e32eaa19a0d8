// Device helpers and the per-gate / measurement kernels (gfx950, wave64).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>

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

constexpr int kBlock = 256;       // threads per workgroup (4 waves)
constexpr int kTreeFanout = 16;   // F of the canonical sampling tree
constexpr unsigned kNormBlocks = 256;   // partial sums per column of a one-qubit marginal

// Sorted (ascending) index-bit positions at which zero bits are deposited.
struct PosList {
    int cnt;
    int s[16];
};

struct Targets {
    int cnt;
    int t[12];   // t[0] = most significant local bit
};

template <typename V> struct RealOf;
template <> struct RealOf<double2> { using type = double; };
template <> struct RealOf<float2> { using type = float; };

template <typename V> __host__ __device__ __forceinline__ V mkc(typename RealOf<V>::type re, typename RealOf<V>::type im)
{
    V v;
    v.x = re;
    v.y = im;
    return v;
}
// a*b + c rounded once.  The complex products below spell out their fused
// multiply-adds, so their results do not depend on how the compiler would
// contract the surrounding code (the sweep kernels are built in two shapes,
// interpreter and specialised, that must agree bit for bit).
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename V> __device__ __forceinline__ V cmul(V a, V b)
{
    return mkc<V>(fmad(a.x, b.x, -(a.y * b.y)), fmad(a.x, b.y, a.y * b.x));
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return mkc<V>(a.x + b.x, a.y + b.y); }
// acc + a*b
template <typename V> __device__ __forceinline__ V cmac(V acc, V a, V b)
{
    return mkc<V>(fmad(a.x, b.x, fmad(-a.y, b.y, acc.x)), fmad(a.x, b.y, fmad(a.y, b.x, acc.y)));
}

// Deposit zero bits of j at the positions in p (ascending): maps a work index
// to a state index whose bits at those positions are 0.
// The loop is unrolled over the fixed capacity so the position array is
// indexed with constants (kernel-argument arrays indexed dynamically would
// be copied to scratch memory).
__device__ __forceinline__ uint64_t deposit(uint64_t j, const PosList& p)
{
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (k < p.cnt) {
            const int s = p.s[k];
            const uint64_t lo = j & ((uint64_t(1) << s) - 1);
            j = ((j >> s) << (s + 1)) | lo;
        }
    }
    return j;
}

// Canonical sampling weight |a|^2 = fma(re, re, im*im) in FP64 (the same
// operation order as the CPU oracle's orc_norm_sqr).
template <typename V> __device__ __forceinline__ double weight(V v)
{
    const double re = double(v.x), im = double(v.y);
    return __fma_rn(re, re, __dmul_rn(im, im));
}

}  // namespace q1t
