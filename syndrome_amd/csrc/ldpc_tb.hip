// Transport-block checks on device: per-code-block CRC-24B, reassembly of
// the transport block from the decoded code blocks, and its CRC-24A
// (DlSchDecoder::decode after the code-block loop, src/transport_block.rs:
// 963-1044; CRC per src/crc.rs).
//
// CRCs here are the reference's: MSB-first register, initial value 0, no
// reflection, no final XOR, remainder of M(x) * x^24 mod g(x) with
// g = 0x864CFB (24A) or 0x800063 (24B).  A workgroup computes one CRC over
// L bits by giving each of its T threads one W = ceil(L / T)-bit chunk of
// the message left-padded with zeros to T * W bits (leading zeros leave the
// remainder unchanged), then folding the T chunk remainders pairwise:
// crc(A || B) = crc(A) * x^len(B) + crc(B) mod g, with x^(W * 2^l) mod g
// prepared once per fold level.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "ldpc_common.h"

namespace syn {

namespace {

constexpr int TB_THREADS = 256;
constexpr uint32_t CRC24A = 0x864CFBu, CRC24B = 0x800063u, MASK24 = 0xFFFFFFu;

__device__ __forceinline__ uint32_t crc_step(uint32_t reg, uint32_t bit, uint32_t poly)
{
    const uint32_t fb = ((reg >> 23) ^ bit) & 1u;
    reg = (reg << 1) & MASK24;
    return fb ? reg ^ poly : reg;
}

// a * b mod g for 24-bit residues (carry-less, Horner over a's bits)
__device__ __forceinline__ uint32_t mulmod(uint32_t a, uint32_t b, uint32_t poly)
{
    uint32_t r = 0;
    for (int i = 23; i >= 0; --i) {
        const uint32_t top = (r >> 23) & 1u;
        r = (r << 1) & MASK24;
        if (top) r ^= poly;
        if ((a >> i) & 1u) r ^= b;
    }
    return r;
}

// Where bit t of a TB's reassembled payload stream lives: block t / P at
// offset t % P (every block contributes P bits, transport_block.rs:566).
struct Stream {
    const uint8_t *hard;
    const long long *offs;  // per code block of the plan
    int first, p;

    __device__ __forceinline__ uint32_t bit(long long t) const
    {
        const int ci = (int)(t / p);
        return hard[offs[first + ci] + (t - (long long)ci * p)] & 1u;
    }
};

// CRC over bits [0, len) of `s` by the whole workgroup; every thread returns
// the remainder.  `red` is TB_THREADS words of LDS.
__device__ uint32_t group_crc(const Stream &s, long long len, uint32_t poly, uint32_t *red)
{
    const int t = threadIdx.x;
    const long long w = (len + TB_THREADS - 1) / TB_THREADS;
    const long long pad = w * TB_THREADS - len;
    uint32_t reg = 0;
    long long b0 = t * w - pad;
    for (long long i = 0; i < w; ++i) {
        const long long b = b0 + i;
        reg = crc_step(reg, b >= 0 ? s.bit(b) : 0u, poly);
    }
    red[t] = reg;
    // x^w mod g, then squared per fold level
    uint32_t xl = 1;
    for (long long i = 0; i < w; ++i) xl = crc_step(xl, 0u, poly);
    __syncthreads();
    for (int stride = 1; stride < TB_THREADS; stride *= 2) {
        if (t % (2 * stride) == 0) red[t] = mulmod(red[t], xl, poly) ^ red[t + stride];
        xl = mulmod(xl, xl, poly);
        __syncthreads();
    }
    const uint32_t r = red[0];
    __syncthreads();
    return r;
}

// the 24 bits at [at, at + 24) of `s`, MSB first
__device__ uint32_t received24(const Stream &s, long long at)
{
    uint32_t v = 0;
    for (int i = 0; i < 24; ++i) v = (v << 1) | s.bit(at + i);
    return v;
}

__global__ void __launch_bounds__(TB_THREADS) tb_check(const uint8_t *__restrict__ hard,
                                                       const long long *__restrict__ offs,
                                                       const SynTb *__restrict__ tbs, int n_tb,
                                                       uint8_t *__restrict__ tb_out, uint8_t *__restrict__ tb_ok,
                                                       uint8_t *__restrict__ cb_ok)
{
    __shared__ uint32_t red[TB_THREADS];
    __shared__ int s_all;
    for (int ti = blockIdx.x; ti < n_tb; ti += gridDim.x) {
        const SynTb tb = tbs[ti];
        const bool seg = tb.num_cb > 1;
        const int p = seg ? tb.k_prime - 24 : tb.k_prime;  // payload bits per block
        if (threadIdx.x == 0) s_all = 1;
        __syncthreads();
        // CRC-24B of every block: its first K'-24 bits against its last 24
        for (int ci = 0; ci < tb.num_cb; ++ci) {
            bool ok = true;
            if (seg) {
                const Stream cb{hard, offs, tb.first_cb + ci, tb.k_prime};
                const uint32_t r = group_crc(cb, tb.k_prime - 24, CRC24B, red);
                ok = r == received24(cb, tb.k_prime - 24);
            }
            if (threadIdx.x == 0) {
                cb_ok[tb.first_cb + ci] = ok ? 1 : 0;
                if (!ok) s_all = 0;
            }
        }
        // CRC-24A of the reassembled transport block
        const Stream s{hard, offs, tb.first_cb, p};
        const uint32_t r = group_crc(s, tb.tb_size, CRC24A, red);
        const bool ok = r == received24(s, tb.tb_size) && s_all;
        if (ok)  // the payload is delivered only when every check passed (transport_block.rs:1036-1040)
            for (long long t = threadIdx.x; t < tb.tb_size; t += TB_THREADS) tb_out[tb.out_off + t] = (uint8_t)s.bit(t);
        if (threadIdx.x == 0) tb_ok[ti] = ok ? 1 : 0;
        __syncthreads();
    }
}

}  // namespace

int launch_tb_check(const uint8_t *hard, const long long *offs, const SynTb *tbs, int n_tb, uint8_t *tb_out,
                    uint8_t *tb_ok, uint8_t *cb_ok, hipStream_t stream)
{
    if (n_tb <= 0) return SYN_OK;
    const int grid = n_tb < 65535 ? n_tb : 65535;
    hipLaunchKernelGGL(tb_check, dim3(grid), dim3(TB_THREADS), 0, stream, hard, offs, tbs, n_tb, tb_out, tb_ok, cb_ok);
    return hipGetLastError() == hipSuccess ? SYN_OK : SYN_DEVICE_ERROR;
}

}  // namespace syn
