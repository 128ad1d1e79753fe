// Resident serving decoder (f32): a grid of one workgroup per CU stays on
// the device while work flows and pulls queued items one at a time.
//
// The host publishes items (SynWork: one code, up to SYN_WORK_MAX code
// blocks named by their server slots) into a ring in coherent pinned host
// memory and then advances ctl->tail.  A workgroup whose last item is
// finished claims the next index by compare-and-swap on a device counter,
// but only an index below the tail it has seen, so nothing is ever claimed
// before it exists and every claimed item is decoded.  Waiting workgroups
// read a device-memory copy of the tail; one of them at a time (a lock)
// refreshes the copy from host memory.  Each item is
// decoded with the register-resident body for its base graph (the
// compile-time-Z instances for BG1 Z=384 and BG2 Z=208; BG1 Z=384 items of
// two blocks with the two-codeword kernel of ldpc_fastx.h, its extension
// columns in a scratch area of the workgroup's own), its inputs read
// straight from each block's source pointer (HBM or pinned host memory),
// its hard decisions and iteration counts written straight into pinned host
// memory; then every wave waits for its own stores to complete, the
// workgroup meets at a barrier, and lane 0 stores the item's completion word
// done[w % cap] = w + 1 with a system-scope release (resident_protocol.h).
//
// A workgroup waiting for work leaves once no workgroup has claimed an item
// for `idle_ticks` of the device wall clock (counted from the grid's start
// at the earliest, so the whole grid stays up while anything flows and goes
// down together), once the kernel is older than `life_ticks`, or when the
// host sets ctl->stop: every wait ends, and the host relaunches the grid
// when work is published after it has left.
#include "ldpc_crc.h"
#include "ldpc_fastx.h"
#include "resident_protocol.h"

namespace syn {

// The protocol's atomics on this device (resident_protocol.h): device-memory
// words at agent scope, host-memory words at system scope.
struct DevOps {
    __device__ unsigned load_relaxed(const unsigned *p) const
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ long long load_relaxed(const long long *p) const
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ unsigned load_acquire(const unsigned *p) const
    {
        return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ unsigned load_acquire_system(const unsigned *p) const
    {
        return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __device__ unsigned load_relaxed_system(const unsigned *p) const
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __device__ void store_relaxed(unsigned *p, unsigned v) const
    {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ void store_relaxed(long long *p, long long v) const
    {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ void store_release(unsigned *p, unsigned v) const
    {
        __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ void store_release_system(unsigned *p, unsigned v) const
    {
        __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __device__ bool try_lock(unsigned *p) const
    {
        unsigned unlocked = 0;
        return __hip_atomic_compare_exchange_strong(p, &unlocked, 1u, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ void unlock(unsigned *p) const { __hip_atomic_store(p, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ bool cas_relaxed(unsigned *p, unsigned expect, unsigned want) const
    {
        return __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ void fence_acquire() const { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
    __device__ void fence_release_system() const { __builtin_amdgcn_fence(__ATOMIC_RELEASE, ""); }
    __device__ void fence_seq_cst() const { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent"); }
    __device__ unsigned atomic_xor(unsigned *p, unsigned v) const { return atomicXor(p, v); }
    __device__ unsigned atomic_or(unsigned *p, unsigned v) const { return atomicOr(p, v); }
    __device__ unsigned atomic_add(unsigned *p, unsigned v) const { return atomicAdd(p, v); }
    __device__ unsigned atomic_exch(unsigned *p, unsigned v) const { return atomicExch(p, v); }
    __device__ long long clock() const { return (long long)wall_clock64(); }
    // sleeps of 512 cycles, n of them
    __device__ void backoff(int n) const
    {
        for (int k = 0; k < n; ++k) __builtin_amdgcn_s_sleep(8);
    }
};

// Before the workgroup barrier that precedes lane 0's publication, every
// wave waits until its own vector stores have completed (s_waitcnt
// vmcnt(0): on gfx950 vmcnt counts stores as well as loads).  A system-scope
// release is a write-back of this XCD's L2 (for data it caches) followed by
// that same wait, and it covers only the releasing wave's stores; the
// results published here go to coherent host memory, which the L2 does not
// hold, so each wave's completed stores plus lane 0's release after the
// barrier make every wave's results visible before the completion word.
// (A full system-scope release fence in every wave, eight L2 write-backs per
// item, measured 10 % worse p99 latency in serving.)
__device__ __forceinline__ void release_all_waves() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)

// A word of a record in host memory, read at system scope so that no GPU
// cache can return an older copy of it.
template <class T>
__device__ __forceinline__ T host_load(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// One decoded code block's share of its transport block's checks, from its
// posterior still in LDS (hard bit i = posterior[i] < 0), by the whole
// workgroup:
//   * its CRC-24B over its first P bits against the next 24 (when the
//     transport block has more than one block);
//   * the CRC-24A remainder of its slice [ci*P, ci*P + L) of the payload
//     stream (L = its bits below tb_size), shifted by x^(tb_size - slice end)
//     so that the XOR of all blocks' shares is the stream's remainder;
//   * the received CRC-24A bits that fall in its P bits;
//   * its payload bits written to the transport block's output.
// Lane 0 records its CRC-24B outcome and folds these into the transport
// block's device state (tb_account, resident_protocol.h); the block that
// completes the count decides the transport block and publishes ok[] and
// done[] (every wave fences the block's payload writes at system scope
// before its count increment).
// The workgroup's CRC reductions run on its first 512 lanes, in a spare
// part of the dynamic LDS (`red`, 512 words: see resident_decode).
constexpr int FOLD_LANES = 512;

__device__ __noinline__ void tb_fold(const SynResidentTb &tbs, int ti, int ci, int slot, const float *post,
                                     uint32_t *red)
{
    __shared__ unsigned s_id;
    __shared__ int s_c, s_p, s_a;
    __shared__ unsigned char *s_out;
    if (threadIdx.x == 0) {
        const SynTbWork *tw = tbs.work + ti;
        // host-written record: system-scope loads, past every cache
        s_id = host_load(&tw->id);
        s_c = host_load(&tw->num_cb);
        s_p = host_load(&tw->p);
        s_a = host_load(&tw->tb_size);
        s_out = host_load(&tw->out);
    }
    __syncthreads();
    const int c = s_c, p = s_p, a = s_a;
    auto bit = [&](long long i) -> uint32_t { return post[i] < 0.f ? 1u : 0u; };
    bool cb_ok = true;
    if (c > 1) cb_ok = group_crc24(bit, p, CRC24B_POLY, red, FOLD_LANES) == received24(bit, p);
    const long long s0 = (long long)ci * p;
    const long long len = a - s0 < 0 ? 0 : (a - s0 > p ? p : a - s0);
    const uint32_t part = len > 0 ? group_crc24(bit, len, CRC24A_POLY, red, FOLD_LANES) : 0u;
    for (long long u = threadIdx.x; u < len; u += blockDim.x) s_out[s0 + u] = (unsigned char)bit(u);
    release_all_waves();  // this block's payload bits are visible before it is counted
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t share = crc24_mulmod(part, crc24_xpow(a - (s0 + len), CRC24A_POLY), CRC24A_POLY);
        uint32_t rx = 0;
        for (int i = 0; i < 24; ++i) {
            const long long t = (long long)a + i;
            if (t >= s0 && t < s0 + p) rx |= bit(t - s0) << (23 - i);
        }
        tbs.cbok[slot] = cb_ok ? 1u : 0u;  // released with the block's count
        tb_account(DevOps{}, tbs.state + ti, share, rx, cb_ok, c, s_id, tbs.ok + ti, tbs.done + ti);
    }
    __syncthreads();
}

// Lanes of the resident workgroup: two BG1 Z=384 codewords (12 waves, three
// per SIMD, so 168 VGPRs); the other codes' items use the first 512.
constexpr int RESIDENT_LANES = FASTX_CPW * 384;
// Where `red` sits in the dynamic LDS: the second codeword region's first
// spill slots, free once a two-codeword item's decode is over, and past the
// end of every one-codeword-kind item's layout (at most 97 KB).
constexpr int RED_OFF = fastx_region_bytes<SYN_KIND_BG1, 384>() + fast_pad(384, 4) +
                        spill_offset<384, kX<Graph<SYN_KIND_BG1>>.core>(0, 0);
static_assert(RED_OFF + 4 * FOLD_LANES <= fastx_group_lds_bytes<SYN_KIND_BG1, 384>(), "red in the LDS");

__global__ void __launch_bounds__(RESIDENT_LANES) __attribute__((amdgpu_waves_per_eu(3, 3)))
resident_decode(const SynCode *__restrict__ codes, const SynWork *__restrict__ ring, unsigned cap,
                SynResidentCtl *ctl, unsigned *done, ResidentQueue *q, long long *last_claim,
                const long long *__restrict__ offs, SynLaunch a, long long idle_ticks, long long life_ticks,
                SynResidentTb tbs, float *xscratch)
{
    __shared__ unsigned s_w;
    __shared__ int s_have, s_code, s_count, s_aligned, s_any_tb;
    __shared__ int s_slots[SYN_WORK_MAX], s_nf[SYN_WORK_MAX], s_tb[SYN_WORK_MAX], s_tbci[SYN_WORK_MAX];
    __shared__ const void *s_src[SYN_WORK_MAX];
    const long long t0 = wall_clock64();
    for (;;) {
        __syncthreads();  // the previous item is finished with LDS and the s_* words
        if (threadIdx.x == 0) {
            unsigned w = 0;
            // Idle workgroups back off exponentially between polls: each
            // poll's acquire invalidates this XCD's L2 (each XCD has its own),
            // and idle workgroups polling without pause kept evicting the
            // decoding workgroups' data, which slowed a lone BG1 Z=384 block
            // from 43 to about 100 us.  The item's ring entry is read with
            // system-scope loads (host_load), which no cache serves.
            const int have = resident_claim(DevOps{}, q, ctl, last_claim, t0, idle_ticks, life_ticks, w) ? 1 : 0;
            s_w = w;
            s_have = have;
            if (have) {
                const SynWork *wk = ring + w % cap;
                s_code = host_load(&wk->code);
                s_count = host_load(&wk->count);
                s_aligned = host_load(&wk->aligned);
            }
        }
        __syncthreads();
        if (!s_have) return;
        const unsigned w = s_w;
        if (threadIdx.x == 0) s_any_tb = 0;
        __syncthreads();
        if ((int)threadIdx.x < s_count) {
            const SynWork *wk = ring + w % cap;
            s_slots[threadIdx.x] = host_load(&wk->slot[threadIdx.x]);
            s_nf[threadIdx.x] = host_load(&wk->nfill[threadIdx.x]);
            s_src[threadIdx.x] = host_load(&wk->src[threadIdx.x]);
            s_tb[threadIdx.x] = host_load(&wk->tb[threadIdx.x]);
            s_tbci[threadIdx.x] = host_load(&wk->tb_ci[threadIdx.x]);
            if (s_tb[threadIdx.x] >= 0) s_any_tb = 1;
        }
        __syncthreads();
        SynLaunch la = a;
        if (s_aligned) la.flags |= SYN_F_ALIGNED;
        const SynCode *code = codes + s_code;
        GroupIO io{s_count, 0, s_slots, offs, nullptr, s_src, s_nf};
        // transport-block code blocks are folded from LDS (tb_fold) and
        // delivered per transport block: their hard decisions need not cross
        // to host memory
        if (s_any_tb)
            for (int j = 0; j < s_count; ++j)
                if (s_tb[j] >= 0) io.no_hard |= 1u << j;
        extern __shared__ __align__(16) unsigned char smem[];
        const bool pair = (la.flags & SYN_F_XPAIR) && code->kind == SYN_KIND_BG1 && code->z == 384;
        if (pair) {
            // up to two blocks: the two-codeword kernel, its extension
            // columns in this workgroup's scratch (two codewords' worth)
            constexpr int N = 68 * 384;
            const int cwl = (int)threadIdx.x >= 384 ? 1 : 0;
            XSlot sl;
            sl.live = cwl < s_count;
            sl.live_o = (1 - cwl) < s_count;
            const int j = sl.live ? cwl : 0;
            sl.id = s_slots[j];
            sl.nf = s_nf[j];
            sl.in = (const float *)s_src[j];
            sl.xb = xscratch + ((size_t)blockIdx.x * FASTX_CPW + cwl) * N;
            sl.post = nullptr;
            sl.hard = ((io.no_hard >> j) & 1u) ? nullptr : la.hard + offs[s_slots[j]];
            fastx_group<SYN_KIND_BG1, 384>(la, sl, NoHook{}, [&]() { return s_slots[cwl]; });
        } else if (code->kind == SYN_KIND_BG1) {
            fast_group<SYN_KIND_BG1, 0, false, true>(code, code->z, s_count, io, la);
        } else {
            if (code->z == 208) fast_group<SYN_KIND_BG2, 208, false, true, false>(code, 208, s_count, io, la);
            else fast_group<SYN_KIND_BG2, 0, false, true>(code, code->z, s_count, io, la);
        }
        if (s_any_tb) {
            __syncthreads();  // the decode is finished with the posterior in LDS
            const int pad = fast_pad(code->z, 4);
            const int n = code->n_b * code->z;
            uint32_t *red = (uint32_t *)(smem + RED_OFF);
            for (int j = 0; j < s_count; ++j)
                if (s_tb[j] >= 0) {
                    // a two-codeword item's posteriors: the core columns of
                    // region j (the payload and its CRC lie below 22Z)
                    const float *post = pair ? (const float *)(smem + j * fastx_region_bytes<SYN_KIND_BG1, 384>() + pad)
                                             : (const float *)(smem + pad) + (size_t)j * n;
                    tb_fold(tbs, s_tb[j], s_tbci[j], s_slots[j], post, red);
                }
        }
        // Every wave releases its own hard decisions and counts (system
        // scope) before the barrier; lane 0 then publishes the item's
        // completion word with a system-scope release.
        release_all_waves();
        __syncthreads();
        if (threadIdx.x == 0) item_complete(DevOps{}, done, cap, w);
    }
}

const void *resident_kernel() { return (const void *)resident_decode; }
int resident_lanes() { return RESIDENT_LANES; }
// LDS of a two-codeword item, and the extension scratch per workgroup (floats)
int resident_pair_lds() { return fastx_group_lds_bytes<SYN_KIND_BG1, 384>(); }
size_t resident_scratch_floats() { return (size_t)FASTX_CPW * 68 * 384; }

}  // namespace syn
