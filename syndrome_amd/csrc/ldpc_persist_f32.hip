// Work-queue decode (f32) for launches that hold codes with Z > 256, and the
// queue reset.
#include "ldpc_persist.h"

namespace syn {

const void *persist_kernel_f32_big() { return (const void *)persist_decode<false, 0>; }

int launch_reset(int *head, hipStream_t stream)
{
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, stream, head);
    return hipGetLastError() == hipSuccess ? SYN_OK : SYN_DEVICE_ERROR;
}

}  // namespace syn
