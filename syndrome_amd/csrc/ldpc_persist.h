// Work-queue (persistent) decode launch.
#pragma once
#include "ldpc_fast.h"

namespace syn {

// Zero the queue head in the launch's own stream (a fill kernel, not a
// memset node, so the sequence can be captured in a HIP graph).
static __global__ void reset_counter(int *head) { *head = 0; }

// Work items of a plan's launch that holds codes with Z > 256 (512-lane
// workgroups) and of its launch for Z <= 256 (256-lane workgroups): the
// f32 kernels carry compile-time-Z bodies for the largest lifting sizes of
// each part, where the work of a batch with Z spread over the 51 sizes
// concentrates (work goes with E*Z).
using ZBig = ZSet<384, 352, 320, 288>;
using ZSmall = ZSet<256, 240, 224, 208>;

template <bool I8, int PART>
__global__ void __launch_bounds__(512) persist_decode(const SynCode *__restrict__ codes,
                                                       const SynItem *__restrict__ items, int n_items, int *head,
                                                       const int *__restrict__ ids,
                                                       const long long *__restrict__ offs,
                                                       const int *__restrict__ nfill, SynLaunch a)
{
    using Z = std::conditional_t<I8, ZSet<>, std::conditional_t<PART == 0, ZBig, ZSmall>>;
    persist_body<I8, Z, Z>(codes, items, n_items, head, ids, offs, nfill, a);
}

}  // namespace syn
