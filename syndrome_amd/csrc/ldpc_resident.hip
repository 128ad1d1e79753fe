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
// compile-time-Z instances for BG1 Z=384 and BG2 Z=208), its inputs read
// straight from each block's source pointer (HBM or pinned host memory),
// its hard decisions and iteration counts written straight into pinned host
// memory; then a system-scope fence and the item's completion word
// done[w % cap] = w + 1.
//
// A workgroup waiting for work leaves once no workgroup has claimed an item
// for `idle_ticks` of the device wall clock (counted from the grid's start
// at the earliest, so the whole grid stays up while anything flows and goes
// down together), once the kernel is older than `life_ticks`, or when the
// host sets ctl->stop: every wait ends, and the host relaunches the grid
// when work is published after it has left.
#include "ldpc_crc.h"
#include "ldpc_fast.h"

namespace syn {

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
// Lane 0 folds these into the transport block's device state; the block
// that completes the count decides the transport block and publishes
// ok[] and done[] (the payload writes of every block are fenced at system
// scope before that block's count increment).
__device__ __noinline__ void tb_fold(const SynResidentTb &tbs, int ti, int ci, const float *post)
{
    __shared__ uint32_t red[512];
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
    if (c > 1) cb_ok = group_crc24(bit, p, CRC24B_POLY, red) == received24(bit, p);
    const long long s0 = (long long)ci * p;
    const long long len = a - s0 < 0 ? 0 : (a - s0 > p ? p : a - s0);
    const uint32_t part = len > 0 ? group_crc24(bit, len, CRC24A_POLY, red) : 0u;
    for (long long u = threadIdx.x; u < len; u += blockDim.x) s_out[s0 + u] = (unsigned char)bit(u);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t share = crc24_mulmod(part, crc24_xpow(a - (s0 + len), CRC24A_POLY), CRC24A_POLY);
        uint32_t rx = 0;
        for (int i = 0; i < 24; ++i) {
            const long long t = (long long)a + i;
            if (t >= s0 && t < s0 + p) rx |= bit(t - s0) << (23 - i);
        }
        SynTbState *st = tbs.state + ti;
        __threadfence_system();  // this block's payload bits are visible before it is counted
        atomicXor(&st->acc, share);
        atomicOr(&st->rx, rx);
        if (!cb_ok) atomicOr(&st->bad, 1u);
        __threadfence();
        if (atomicAdd(&st->count, 1u) == (unsigned)(c - 1)) {
            __threadfence();
            const unsigned acc = atomicOr(&st->acc, 0u), rxv = atomicOr(&st->rx, 0u), bad = atomicOr(&st->bad, 0u);
            atomicExch(&st->acc, 0u);  // ready for the ring index's next transport block
            atomicExch(&st->rx, 0u);
            atomicExch(&st->bad, 0u);
            atomicExch(&st->count, 0u);
            tbs.ok[ti] = (!bad && acc == rxv) ? 1u : 0u;
            __hip_atomic_store(tbs.done + ti, s_id + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    __syncthreads();
}

// Device-side queue words (SYN_RESIDENT_QUEUE_BYTES of device memory): the
// claim counter, alone on its cache line, then the device's copy of the
// host's tail and stop words and the lock that lets one waiting workgroup
// at a time refresh that copy from host memory (so idle workgroups spin on
// device memory, not on PCIe).  Waiters poll the second line; a claim's
// compare-and-swap on the first does not queue behind their traffic.
struct ResidentQueue {
    unsigned head;
    unsigned pad0[31];
    unsigned tail, stop, poll_lock;
    unsigned pad1[29];
};
static_assert(sizeof(ResidentQueue) == SYN_RESIDENT_QUEUE_BYTES, "queue block size");

__global__ void __launch_bounds__(512) resident_decode(const SynCode *__restrict__ codes,
                                                        const SynWork *__restrict__ ring, unsigned cap,
                                                        SynResidentCtl *ctl, unsigned *done, ResidentQueue *q,
                                                        long long *last_claim, const long long *__restrict__ offs,
                                                        SynLaunch a, long long idle_ticks, long long life_ticks,
                                                        SynResidentTb tbs)
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
            int have = 0;
            int backoff = 1;  // sleeps of 512 cycles between idle polls, doubling to 64
            // Idle workgroups back off exponentially between polls: each
            // poll's acquire invalidates this XCD's L2 (each XCD has its own),
            // and idle workgroups polling without pause kept evicting the
            // decoding workgroups' data, which slowed a lone BG1 Z=384 block
            // from 43 to about 100 us.  The item's ring entry is read with
            // system-scope loads (host_load), which no cache serves.
            for (;;) {
                const unsigned h = __hip_atomic_load(&q->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                unsigned t = __hip_atomic_load(&q->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                if ((int)(t - h) <= 0) {
                    unsigned unlocked = 0;
                    if (__hip_atomic_compare_exchange_strong(&q->poll_lock, &unlocked, 1u, __ATOMIC_ACQUIRE,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        t = __hip_atomic_load(&ctl->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
                        const unsigned stop = __hip_atomic_load(&ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                        // the lock makes this the only writer: a plain store keeps the
                        // copy monotonic modulo 2^32
                        __hip_atomic_store(&q->tail, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (stop) __hip_atomic_store(&q->stop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(&q->poll_lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                if ((int)(t - h) > 0) {
                    unsigned expect = h;
                    if (__hip_atomic_compare_exchange_strong(&q->head, &expect, h + 1, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        // the item's inputs (HBM or host memory) may have been
                        // rewritten since this XCD's L2 last held them: one
                        // agent-scope acquire invalidates those lines
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                        w = h;
                        have = 1;
                        __hip_atomic_store(last_claim, (long long)wall_clock64(), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    continue;
                }
                // idle = no claim by any workgroup since this grid started
                const long long now = wall_clock64();
                const long long lc = __hip_atomic_load(last_claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const long long since = now - (lc > t0 ? lc : t0);
                if (__hip_atomic_load(&q->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || since > idle_ticks ||
                    now - t0 > life_ticks)
                    break;
                for (int k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(8);
                backoff = backoff < 64 ? 2 * backoff : 64;
            }
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
        if (code->kind == SYN_KIND_BG1) {
            if (code->z == 384) fast_group<SYN_KIND_BG1, 384, false, true, false>(code, 384, s_count, io, la);
            else fast_group<SYN_KIND_BG1, 0, false, true>(code, code->z, s_count, io, la);
        } else {
            if (code->z == 208) fast_group<SYN_KIND_BG2, 208, false, true, false>(code, 208, s_count, io, la);
            else fast_group<SYN_KIND_BG2, 0, false, true>(code, code->z, s_count, io, la);
        }
        if (s_any_tb) {
            __syncthreads();  // the decode is finished with the posterior in LDS
            const int pad = fast_pad(code->z, 4);
            const int n = code->n_b * code->z;
            extern __shared__ __align__(16) unsigned char smem[];
            for (int j = 0; j < s_count; ++j)
                if (s_tb[j] >= 0) tb_fold(tbs, s_tb[j], s_tbci[j], (const float *)(smem + pad) + (size_t)j * n);
        }
        // Every lane's hard decisions and counts happen before the barrier;
        // the system-scope release by lane 0 then publishes all of them
        // with the item's completion word.
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(done + w % cap, w + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

const void *resident_kernel() { return (const void *)resident_decode; }

}  // namespace syn
