// The resident serving decoder's device-side protocol: how a workgroup
// claims the next queued item, how a decoded transport-block code block is
// counted into its transport block, and how an item's completion is
// published.  Written once over an `Ops` policy that supplies the atomics,
// fences and clock at the scopes named here: ldpc_resident.hip runs it with
// the device's scoped atomics, and the CPU thread model of the protocol
// (tests/model/resident_model.cpp) runs the same code with std::atomic_ref.
//
// Memory placement: the queue block (ResidentQueue) and the transport-block
// fold states live in device memory; the ring of items, the control block
// (tail, stop), the completion words and the transport-block records live
// in coherent host memory, written by the host dispatcher.
//
// Ordering (the AMDGPU memory model's scoped release/acquire; x86-TSO on the
// host side):
//  * item publish: the host writes ring entry w, then stores ctl->tail with
//    release (system).  The workgroup holding poll_lock loads ctl->tail with
//    acquire (system) and stores its device copy q->tail with release
//    (agent); a claimer loads q->tail with acquire (agent), so the host's
//    ring writes happen before the claimer's reads of entry w.
//  * claim: a compare-and-swap on q->head takes index w only below a tail the
//    claimer has seen, so every claimed index is published, and each index
//    is claimed once.  After the claim one acquire fence (agent) drops stale
//    cached copies of the item's inputs.
//  * completion: every wave of the workgroup waits for its own result stores
//    to complete before the workgroup barrier (ldpc_resident.hip,
//    release_all_waves); lane 0 then stores done[w % cap] = w + 1 with
//    release (system).  A release waits only for the releasing wave's
//    stores, so without the per-wave wait another wave's results could
//    reach the host after the completion word.
//  * transport blocks: every wave's payload stores complete (as above)
//    before lane 0 counts the block, and lane 0's release fence (system)
//    precedes its first read-modify-write; the fold's XOR /
//    OR / count are agent-scope read-modify-writes on one state, separated by
//    seq_cst fences, so the block whose increment completes the count reads
//    every other block's contribution (their fences synchronise through the
//    count's read-modify-write chain).  It resets the state, stores ok, and
//    stores the transport block's done word with release (system).
//  * A transport-block record is rewritten (a new submission at that index)
//    only after the host has seen its done word, and the reset of its fold
//    state precedes that word, so a new transport block starts from zero.
#pragma once

#include "ldpc_common.h"

#ifndef SYN_HD
#define SYN_HD __host__ __device__
#endif

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

namespace syn {

// Claim the next item for the calling workgroup (lane 0), or return false
// when the grid should leave: the stop word is set, no workgroup has claimed
// anything for idle_ticks of the clock (counted from t0, the grid's start,
// at the earliest), or the grid is older than life_ticks.  Idle waiters back
// off exponentially between polls (ops.backoff(1..64)).
template <class Ops>
SYN_HD bool resident_claim(const Ops &ops, ResidentQueue *q, SynResidentCtl *ctl, long long *last_claim, long long t0,
                           long long idle_ticks, long long life_ticks, unsigned &w)
{
    int backoff = 1;
    for (;;) {
        const unsigned h = ops.load_relaxed(&q->head);
        unsigned t = ops.load_acquire(&q->tail);
        if ((int)(t - h) <= 0 && ops.try_lock(&q->poll_lock)) {
            t = ops.load_acquire_system(&ctl->tail);
            const unsigned stop = ops.load_relaxed_system(&ctl->stop);
            // the lock makes this the only writer: the copy stays monotonic
            // modulo 2^32
            ops.store_release(&q->tail, t);
            if (stop) ops.store_relaxed(&q->stop, 1u);
            ops.unlock(&q->poll_lock);
        }
        if ((int)(t - h) > 0) {
            if (ops.cas_relaxed(&q->head, h, h + 1)) {
                // the item's inputs (HBM or host memory) may have been
                // rewritten since this cache last held them
                ops.fence_acquire();
                w = h;
                ops.store_relaxed(last_claim, ops.clock());
                return true;
            }
            continue;
        }
        const long long now = ops.clock();
        const long long lc = ops.load_relaxed(last_claim);
        const long long since = now - (lc > t0 ? lc : t0);
        if (ops.load_relaxed(&q->stop) || since > idle_ticks || now - t0 > life_ticks) return false;
        ops.backoff(backoff);
        backoff = backoff < 64 ? 2 * backoff : 64;
    }
}

// Count one decoded code block into its transport block's fold state (lane
// 0, after every wave's payload stores have completed):
// share = its slice's CRC-24A remainder contribution, rx = the received
// CRC-24A bits it holds, cb_ok = its CRC-24B passed, c = the transport
// block's code blocks, id = its submission number.  The block that completes
// the count decides the transport block: ok = every check passed, then the
// done word id + 1.
template <class Ops>
SYN_HD void tb_account(const Ops &ops, SynTbState *st, unsigned share, unsigned rx, bool cb_ok, int c, unsigned id,
                       unsigned *ok, unsigned *done)
{
    ops.fence_release_system();
    ops.atomic_xor(&st->acc, share);
    ops.atomic_or(&st->rx, rx);
    if (!cb_ok) ops.atomic_or(&st->bad, 1u);
    ops.fence_seq_cst();
    if (ops.atomic_add(&st->count, 1u) == (unsigned)(c - 1)) {
        ops.fence_seq_cst();
        const unsigned acc = ops.atomic_or(&st->acc, 0u), rxv = ops.atomic_or(&st->rx, 0u);
        const unsigned bad = ops.atomic_or(&st->bad, 0u);
        // ready for the ring index's next transport block
        ops.atomic_exch(&st->acc, 0u);
        ops.atomic_exch(&st->rx, 0u);
        ops.atomic_exch(&st->bad, 0u);
        ops.atomic_exch(&st->count, 0u);
        *ok = (!bad && acc == rxv) ? 1u : 0u;
        ops.store_release_system(done, id + 1u);
    }
}

// Publish item w's completion (lane 0, after the workgroup barrier that
// follows the completion of every wave's result stores).
template <class Ops>
SYN_HD void item_complete(const Ops &ops, unsigned *done, unsigned cap, unsigned w)
{
    ops.store_release_system(done + w % cap, w + 1u);
}

}  // namespace syn
