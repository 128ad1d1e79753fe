// The reference's 24-bit CRCs (src/crc.rs) on device: MSB-first register,
// initial value 0, no reflection, no final XOR, remainder of M(x) * x^24 mod
// g(x) with g = 0x864CFB (CRC-24A) or 0x800063 (CRC-24B).
//
// A workgroup computes one CRC over L bits by giving each of its T threads
// one W = ceil(L / T)-bit chunk of the message left-padded with zeros to
// T * W bits (leading zeros leave the remainder unchanged), then folding the
// T chunk remainders pairwise: crc(A || B) = crc(A) * x^len(B) + crc(B)
// mod g, with x^(W * 2^l) mod g prepared once per fold level.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace syn {

constexpr uint32_t CRC24A_POLY = 0x864CFBu, CRC24B_POLY = 0x800063u, CRC24_MASK = 0xFFFFFFu;

__device__ __forceinline__ uint32_t crc24_step(uint32_t reg, uint32_t bit, uint32_t poly)
{
    const uint32_t fb = ((reg >> 23) ^ bit) & 1u;
    reg = (reg << 1) & CRC24_MASK;
    return fb ? reg ^ poly : reg;
}

// a * b mod g for 24-bit residues (carry-less, Horner over a's bits)
__device__ __forceinline__ uint32_t crc24_mulmod(uint32_t a, uint32_t b, uint32_t poly)
{
    uint32_t r = 0;
    for (int i = 23; i >= 0; --i) {
        const uint32_t top = (r >> 23) & 1u;
        r = (r << 1) & CRC24_MASK;
        if (top) r ^= poly;
        if ((a >> i) & 1u) r ^= b;
    }
    return r;
}

// x^n mod g (square and multiply)
__device__ __forceinline__ uint32_t crc24_xpow(long long n, uint32_t poly)
{
    uint32_t result = 1, base = 2;  // x^0, x^1
    while (n > 0) {
        if (n & 1) result = crc24_mulmod(result, base, poly);
        base = crc24_mulmod(base, base, poly);
        n >>= 1;
    }
    return result;
}

// CRC of bits [0, len) of `bit(i)` by the first `nt` threads of the
// workgroup (nt a power of two, blockDim.x when 0); `red` holds nt words of
// LDS.  Every thread returns the remainder.  Call from all threads of the
// workgroup.
template <class Bit>
__device__ uint32_t group_crc24(const Bit &bit, long long len, uint32_t poly, uint32_t *red, int nt_use = 0)
{
    const int t = threadIdx.x, nt = nt_use > 0 ? nt_use : (int)blockDim.x;
    const bool part = t < nt;
    const long long w = (len + nt - 1) / nt;
    const long long pad = w * nt - len;
    uint32_t reg = 0;
    const long long b0 = t * w - pad;
    if (part)
        for (long long i = 0; i < w; ++i) {
            const long long b = b0 + i;
            reg = crc24_step(reg, b >= 0 ? bit(b) : 0u, poly);
        }
    if (part) red[t] = reg;
    uint32_t xl = crc24_xpow(w, poly);
    __syncthreads();
    for (int stride = 1; stride < nt; stride *= 2) {
        if (part && t % (2 * stride) == 0) red[t] = crc24_mulmod(red[t], xl, poly) ^ red[t + stride];
        xl = crc24_mulmod(xl, xl, poly);
        __syncthreads();
    }
    const uint32_t r = red[0];
    __syncthreads();
    return r;
}

// the 24 bits at [at, at + 24) of `bit`, MSB first
template <class Bit>
__device__ __forceinline__ uint32_t received24(const Bit &bit, long long at)
{
    uint32_t v = 0;
    for (int i = 0; i < 24; ++i) v = (v << 1) | bit(at + i);
    return v;
}

}  // namespace syn
