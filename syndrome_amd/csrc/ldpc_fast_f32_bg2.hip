// Register-resident f32 decode kernels for BG2: the run-time-Z kernel and
// the compile-time-Z specialisation for Z=208.
#include "ldpc_fast.h"

namespace syn {

const void *fast_kernel_f32_bg2(int z)
{
    if (z == 208) return (const void *)fast_f32<SYN_KIND_BG2, 208>;
    return (const void *)fast_f32<SYN_KIND_BG2, 0>;
}

}  // namespace syn
