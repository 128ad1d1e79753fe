// Register-resident f32 decode of BG1 with the extension columns'
// posteriors in HBM (ldpc_fastx.h): two codewords per CU, for the lifting
// sizes above 256, per-config and work-queue kernels.
#include "ldpc_fastx.h"

namespace syn {

const void *fastx_kernel_f32_bg1(int z)
{
    switch (z) {
    case 384: return (const void *)fastx_f32<SYN_KIND_BG1, 384>;
    case 352: return (const void *)fastx_f32<SYN_KIND_BG1, 352>;
    case 320: return (const void *)fastx_f32<SYN_KIND_BG1, 320>;
    case 288: return (const void *)fastx_f32<SYN_KIND_BG1, 288>;
    default: return nullptr;
    }
}

// The work-queue kernels: Z=384 alone, and all four sizes in one kernel for
// plans that hold the others too.  Each further size in one kernel costs
// registers for all of them (two sizes: 87 VGPR spills, against 74 for
// one), so a plan of Z=384 blocks only keeps the one-size kernel.
const void *persistx_kernel_f32_bg1(int z)
{
    switch (z) {
    case 384: return (const void *)persistx_f32<SYN_KIND_BG1, 384>;
    case 352:
    case 320:
    case 288: return (const void *)persistx_f32<SYN_KIND_BG1, 384, 352, 320, 288>;
    default: return nullptr;
    }
}

int persistx_lanes_f32_bg1() { return fastx_max_lanes<384>(); }

// LDS bytes of one workgroup at lifting size z, its lanes (per-config
// launch) and codewords per workgroup
int fastx_shape_f32_bg1(int z, int *lanes, int *cws)
{
    *cws = FASTX_CPW;
    *lanes = FASTX_CPW * fastx_lanes(z);
    switch (z) {
    case 384: return fastx_group_lds_bytes<SYN_KIND_BG1, 384>();
    case 352: return fastx_group_lds_bytes<SYN_KIND_BG1, 352>();
    case 320: return fastx_group_lds_bytes<SYN_KIND_BG1, 320>();
    case 288: return fastx_group_lds_bytes<SYN_KIND_BG1, 288>();
    default: return 0;
    }
}

}  // namespace syn
