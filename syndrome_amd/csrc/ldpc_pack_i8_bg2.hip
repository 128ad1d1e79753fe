// Pair-packed int8 decode kernels for BG2: the run-time-Z kernel and the
// compile-time-Z specialisation for Z=208.
#include "ldpc_pack.h"

namespace syn {

const void *pack_kernel_i8_bg2(int z)
{
    if (z == 208) return (const void *)pack_i8<SYN_KIND_BG2, 208>;
    return (const void *)pack_i8<SYN_KIND_BG2, 0>;
}

}  // namespace syn
