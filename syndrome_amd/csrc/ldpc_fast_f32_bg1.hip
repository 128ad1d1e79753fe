// Register-resident f32 decode kernels for BG1: the run-time-Z kernel and
// the compile-time-Z specialisation for Z=384.
#include "ldpc_fast.h"

namespace syn {

const void *fast_kernel_f32_bg1(int z)
{
    if (z == 384) return (const void *)fast_f32<SYN_KIND_BG1, 384>;
    return (const void *)fast_f32<SYN_KIND_BG1, 0>;
}

}  // namespace syn
