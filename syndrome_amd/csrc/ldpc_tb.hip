// Transport-block checks on device: per-code-block CRC-24B, reassembly of
// the transport block from the decoded code blocks, and its CRC-24A
// (DlSchDecoder::decode after the code-block loop, src/transport_block.rs:
// 963-1044; CRC per src/crc.rs).
//
// CRCs per ldpc_crc.h.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "ldpc_common.h"
#include "ldpc_crc.h"

namespace syn {

namespace {

constexpr int TB_THREADS = 256;

// Where bit t of a TB's reassembled payload stream lives: block t / P at
// offset t % P (every block contributes P bits, transport_block.rs:566).
struct Stream {
    const uint8_t *hard;
    const long long *offs;  // per code block of the plan
    int first, p;

    __device__ __forceinline__ uint32_t operator()(long long t) const
    {
        const int ci = (int)(t / p);
        return hard[offs[first + ci] + (t - (long long)ci * p)] & 1u;
    }
};

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
                const uint32_t r = group_crc24(cb, tb.k_prime - 24, CRC24B_POLY, red);
                ok = r == received24(cb, tb.k_prime - 24);
            }
            if (threadIdx.x == 0) {
                cb_ok[tb.first_cb + ci] = ok ? 1 : 0;
                if (!ok) s_all = 0;
            }
        }
        // CRC-24A of the reassembled transport block
        const Stream s{hard, offs, tb.first_cb, p};
        const uint32_t r = group_crc24(s, tb.tb_size, CRC24A_POLY, red);
        const bool ok = r == received24(s, tb.tb_size) && s_all;
        if (ok)  // the payload is delivered only when every check passed (transport_block.rs:1036-1040)
            for (long long t = threadIdx.x; t < tb.tb_size; t += TB_THREADS) tb_out[tb.out_off + t] = (uint8_t)s(t);
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
