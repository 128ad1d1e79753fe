// Register-resident f32 decode of BG1 Z=384 with the extension columns'
// posteriors in HBM (ldpc_fastx.h): two codewords, 12 waves, per CU.
#include "ldpc_fastx.h"

namespace syn {

const void *fastx_kernel_f32_bg1(int z)
{
    if (z == 384) return (const void *)fastx_f32<SYN_KIND_BG1, 384>;
    return nullptr;
}

const void *persistx_kernel_f32_bg1(int z)
{
    if (z == 384) return (const void *)persistx_f32<SYN_KIND_BG1, 384>;
    return nullptr;
}

// LDS bytes and lanes of one workgroup, and codewords per workgroup
int fastx_shape_f32_bg1(int z, int *lanes, int *cws)
{
    *lanes = FASTX_CPW * z;
    *cws = FASTX_CPW;
    if (z == 384) return fastx_group_lds_bytes<SYN_KIND_BG1, 384>();
    return 0;
}

}  // namespace syn
