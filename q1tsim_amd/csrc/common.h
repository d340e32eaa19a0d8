// Shared definitions for the q1tsim MI355X statevector backend.
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace q1t {

using cdouble = std::complex<double>;

// Error kinds; the values are the public Q1T_ERR_* codes of q1tsim_gvs.h and
// mirror the variants of the reference's Error enum (src/error.rs:135-173).
enum ErrCode : int {
    OK = 0,
    ERR_INVALID_NR_BITS = 1,
    ERR_INVALID_QBIT = 2,
    ERR_INVALID_CBIT = 3,
    ERR_INVALID_NR_CONTROL_BITS = 4,
    ERR_INVALID_NR_MEASUREMENT_BITS = 5,
    ERR_NOT_ENOUGH_SPACE = 6,
    ERR_NOT_EXECUTED = 7,
    ERR_INVALID_WEIGHTS = 8,
    ERR_UNKNOWN_GATE = 9,
    ERR_INVALID_NR_ARGUMENTS = 10,
    ERR_HIP = 11,
    ERR_OUT_OF_MEMORY = 12,
    ERR_INVALID_ARGUMENT = 13,
    ERR_NOT_SUPPORTED = 14,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

// Display strings of src/error.rs:192-256.
inline Error err_invalid_nr_bits(size_t actual, size_t expected, const std::string& desc)
{
    return Error(ERR_INVALID_NR_BITS, "Expected " + std::to_string(expected) + " bits for \"" + desc +
                                          "\", got " + std::to_string(actual));
}
inline Error err_invalid_qbit(size_t bit)
{
    return Error(ERR_INVALID_QBIT, "Invalid index " + std::to_string(bit) + " for a quantum bit");
}
inline Error err_invalid_cbit(size_t bit)
{
    return Error(ERR_INVALID_CBIT, "Invalid index " + std::to_string(bit) + " for a classical bit");
}
inline Error err_invalid_nr_control_bits(size_t actual, size_t expected, const std::string& desc)
{
    return Error(ERR_INVALID_NR_CONTROL_BITS,
                 "The number of runs is " + std::to_string(expected) + ", but received " +
                     std::to_string(actual) + " control bits for controlled " + desc + " operation");
}
inline Error err_invalid_nr_measurement_bits(size_t actual, size_t expected)
{
    return Error(ERR_INVALID_NR_MEASUREMENT_BITS, "Expected " + std::to_string(expected) +
                                                      " measurement bits, but got " + std::to_string(actual));
}
inline Error err_not_enough_space(size_t actual, size_t expected)
{
    return Error(ERR_NOT_ENOUGH_SPACE, "Not enough space to store " + std::to_string(expected) +
                                           " measurement results in array of length " + std::to_string(actual));
}
inline Error err_not_executed() { return Error(ERR_NOT_EXECUTED, "The circuit has not been executed yet"); }
inline Error err_unknown_gate(const std::string& name)
{
    return Error(ERR_UNKNOWN_GATE, "Unknown gate \"" + name + "\"");
}
inline Error err_invalid_nr_arguments(size_t actual, size_t expected, const std::string& name)
{
    return Error(ERR_INVALID_NR_ARGUMENTS, "Expected " + std::to_string(expected) + " arguments to \"" + name +
                                               "\" gate, got " + std::to_string(actual));
}

// Random source: raw 64-bit draws from the caller (a Rust `R: Rng` through a
// trampoline, or the built-in xoshiro256**).
struct RngSource {
    void* ctx;
    uint64_t (*next_u64)(void* ctx);
    uint64_t next() { return next_u64(ctx); }
    // rand 0.7 Uniform<f64> construction: 52 random mantissa bits in [0,1).
    double uniform() { return double(next() >> 12) * (1.0 / 4503599627370496.0); }
};

// xoshiro256** (Blackman & Vigna, public domain algorithm) seeded by SplitMix64.
struct Xoshiro256 {
    uint64_t s[4];
    explicit Xoshiro256(uint64_t seed = 0) { reseed(seed); }
    void reseed(uint64_t seed)
    {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) {
            uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    static uint64_t trampoline(void* ctx) { return static_cast<Xoshiro256*>(ctx)->next(); }
    RngSource source() { return RngSource{this, &Xoshiro256::trampoline}; }
};

// support.rs:50-60
inline uint64_t reverse_bits(uint64_t idx, size_t nr_bits)
{
    uint64_t res = 0, sidx = idx;
    for (size_t i = 0; i < nr_bits; ++i) {
        res |= (sidx & 1) << (nr_bits - 1 - i);
        sidx >>= 1;
    }
    return res;
}

// support.rs:65-75
inline uint64_t shuffle_bits(uint64_t idx, const size_t* bits, size_t nbits)
{
    uint64_t res = 0, sidx = idx;
    for (size_t k = 0; k < nbits; ++k) {
        res |= (sidx & 1) << bits[k];
        sidx >>= 1;
    }
    return res;
}

}  // namespace q1t
