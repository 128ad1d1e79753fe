// Work-queue (persistent) decode launch.
#pragma once
#include "ldpc_pack.h"

namespace syn {

// Persistent work-queue decoder: a grid sized to the resident workgroups
// pulls items off a device counter until the queue is drained (every
// workgroup reaches the exit test with the same counter, so all of them
// leave), decoding each item with the register-resident body for its base
// graph at its lifting size: a compile-time-Z body for the sizes in Z1
// (BG1) and Z2 (BG2), the run-time-Z body for the rest.  Items of
// different base graphs, lifting sizes and filler counts share one launch.
// The int8 kernel decodes BG1 items with the register-lean one-codeword
// body and BG2 items marked SYN_ITEM_PACKED pair-packed (pack_group): such an
// item holds up to two codewords per Z lanes.
template <bool I8, class Z1, class Z2>
__device__ __forceinline__ void persist_body(const SynCode *__restrict__ codes, const SynItem *__restrict__ items,
                                             int n_items, int *head, const int *__restrict__ ids,
                                             const long long *__restrict__ offs, const int *__restrict__ nfill,
                                             const SynLaunch &a)
{
    __shared__ int s_item;
    // the item's per-codeword element offsets and filler counts, looked up
    // once per item
    __shared__ long long s_off[SYN_ITEM_MAX_CWS];
    __shared__ int s_nf[SYN_ITEM_MAX_CWS];
    if (threadIdx.x == 0) s_item = atomicAdd(head, 1);
    for (;;) {
        __syncthreads();  // s_item is written
        const int w = s_item;
        if (w >= n_items) return;
        const SynItem it = items[w];
        const SynCode *code = codes + it.code;
        if ((int)threadIdx.x < it.count) {
            const int id = ids[it.first + threadIdx.x];
            s_off[threadIdx.x] = offs[id];
            s_nf[threadIdx.x] = nfill ? nfill[id] : (int)a.n_filler;
        }
        __syncthreads();
        GroupIO io{it.count, 0, ids + it.first, offs, nfill, nullptr, nullptr};
        io.t_off = s_off;
        io.t_nf = s_nf;
        // the next item is claimed once this one's iterations are over, so
        // the atomic's round trip overlaps the output stores (claiming it
        // any earlier would let one workgroup hold two items at the end of
        // the queue while others idle)
        int next = 0;
        const auto claim = [&]() {
            if (threadIdx.x == 0) next = atomicAdd(head, 1);
        };
        const int kind = it.kind & ~SYN_ITEM_PACKED;
        if (kind == SYN_KIND_BG1) {
            item_group<SYN_KIND_BG1, I8>(Z1{}, code, it.count, io, a, claim);
        } else if (I8 && (it.kind & SYN_ITEM_PACKED)) {
            if constexpr (I8) pack_group<SYN_KIND_BG2, 0>(code, code->z, (it.count + 1) & ~1, io, a, claim);
        } else {
            item_group<SYN_KIND_BG2, I8>(Z2{}, code, it.count, io, a, claim);
        }
        __syncthreads();  // this item is finished with LDS, s_item and the tables
        if (threadIdx.x == 0) s_item = next;
    }
}


// Zero the queue head in the launch's own stream (a fill kernel, not a
// memset node, so the sequence can be captured in a HIP graph).
static __global__ void reset_counter(int *head) { *head = 0; }

// Work items of a plan's launch that holds codes with Z > 256 (512-lane
// workgroups) and of its launches for Z <= 256: the f32 kernels carry
// compile-time-Z bodies for the largest lifting sizes of each part, where
// the work of a batch with Z spread over the 51 sizes concentrates (work
// goes with E*Z).  The Z <= 256 codes of the two base graphs go to two
// kernels (PART 1: BG1, PART 2: BG2), each with its graph's 13 bodies, in
// translation units of their own: one kernel with all 26 takes the
// compiler several times as long as two with 13 each.
using ZBig = ZSet<384, 352, 320, 288>;
using ZSmall = ZSet<256, 240, 224, 208, 192, 176, 160, 144, 128, 120, 112, 104, 96>;

// The int8 kernel holds three waves per SIMD (168 VGPRs), the budget of
// its pair-packed BG2 body.
template <bool I8, int PART>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(I8 ? 3 : 1))) persist_decode(const SynCode *__restrict__ codes,
                                                       const SynItem *__restrict__ items, int n_items, int *head,
                                                       const int *__restrict__ ids,
                                                       const long long *__restrict__ offs,
                                                       const int *__restrict__ nfill, SynLaunch a)
{
    using Z1 = std::conditional_t<I8 || PART == 2, ZSet<>, std::conditional_t<PART == 0, ZBig, ZSmall>>;
    using Z2 = std::conditional_t<I8 || PART == 1, ZSet<>, std::conditional_t<PART == 0, ZBig, ZSmall>>;
    persist_body<I8, Z1, Z2>(codes, items, n_items, head, ids, offs, nfill, a);
}

}  // namespace syn
