// Work-queue decode (f32), the queue reset and the device rate de-matcher.
#include "ldpc_persist.h"

namespace syn {

const void *persist_kernel_f32() { return (const void *)persist_decode<false>; }

int launch_reset(int *head, hipStream_t stream)
{
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, stream, head);
    return hipGetLastError() == hipSuccess ? SYN_OK : SYN_DEVICE_ERROR;
}

int launch_dematch(const float *e_llr, float *out, const SynDematchCb *cbs, int count, int max_n, bool combine,
                   int rv, hipStream_t stream)
{
    if (count <= 0) return SYN_OK;
    const int bx = (max_n + 255) / 256 < 32 ? (max_n + 255) / 256 : 32;
    const int by = count < 65535 ? count : 65535;
    hipLaunchKernelGGL(dematch, dim3(bx, by), dim3(256), 0, stream, e_llr, out, cbs, count, combine ? 1 : 0, rv);
    return hipGetLastError() == hipSuccess ? SYN_OK : SYN_DEVICE_ERROR;
}

}  // namespace syn
