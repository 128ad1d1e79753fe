// Register-resident i8 decode kernels for BG1: the run-time-Z kernel and
// the compile-time-Z specialisation for Z=384.
#include "ldpc_fast.h"

namespace syn {

const void *fast_kernel_i8_bg1(int z)
{
    if (z == 384) return (const void *)fast_i8<SYN_KIND_BG1, 384>;
    return (const void *)fast_i8<SYN_KIND_BG1, 0>;
}

}  // namespace syn
