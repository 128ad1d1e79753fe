// Work-queue (persistent) decode launch.
#pragma once
#include "ldpc_fast.h"

namespace syn {

// Zero the queue head in the launch's own stream (a fill kernel, not a
// memset node, so the sequence can be captured in a HIP graph).
static __global__ void reset_counter(int *head) { *head = 0; }

template <bool I8>
__global__ void __launch_bounds__(512) persist_decode(const SynCode *__restrict__ codes,
                                                       const SynItem *__restrict__ items, int n_items, int *head,
                                                       const int *__restrict__ ids,
                                                       const long long *__restrict__ offs,
                                                       const int *__restrict__ nfill, SynLaunch a)
{
    persist_body<I8>(codes, items, n_items, head, ids, offs, nfill, a);
}

}  // namespace syn
