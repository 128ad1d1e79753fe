// Register-resident i8 decode kernels for BG2: the run-time-Z kernel and
// the compile-time-Z specialisation for Z=208.
#include "ldpc_fast.h"

namespace syn {

const void *fast_kernel_i8_bg2(int z)
{
    if (z == 208) return (const void *)fast_i8<SYN_KIND_BG2, 208>;
    return (const void *)fast_i8<SYN_KIND_BG2, 0>;
}

}  // namespace syn
