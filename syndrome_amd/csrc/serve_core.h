// The serving loop's host-side core, free of HIP calls: slot admission, the
// submit / receive / release protocol (the reference pipeline's,
// src/ldpc_pipeline.rs:499-576), transport-block records, and the resident
// decoder's dispatcher (publish items into the ring, scan completions,
// decide when the grid must be relaunched, and the failure path).
//
// ServeCore<Backend> leaves the device to `Backend`:
//   int32_t launch(ServeCore &)   start a resident grid (status)
//   int query()                   1 once the last grid has left, 0 while it runs, -1 on error
//   void wait()                   block until the last grid has left
// ldpc_serve.hip instantiates it with HIP (pinned host memory, the resident
// kernel); tests/model/resident_model.cpp with host threads that run the
// kernel's protocol (resident_protocol.h) over plain memory.
//
// Completion order: the grid claims items in publish order and holds at
// most `grid` of them at once, so past the (grid + 1)-th unfinished item
// nothing has been claimed; the scan stops there and picks the rest up on a
// later pass.  A ring entry is rewritten only once its item is finished and
// popped (at most cap - 1 items outstanding), and every claimed item is
// finished before its grid leaves.
#pragma once

#include <limits.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <deque>
#include <initializer_list>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/syndrome_ldpc.h"
#include "ldpc_common.h"

namespace syn {

constexpr int SERVE_NMAX = 68 * 384;  // decoder inputs per slot: the largest code

inline double now_ns()
{
    return (double)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// The server lock.  Its critical sections last well under a microsecond,
// while a thread that blocks in the kernel on a contended std::mutex takes
// tens of microseconds to be woken: a waiter retries for a while before it
// blocks.  That wake-up was most of a code block's host-side latency.
class ServerMutex {
  public:
    void lock()
    {
        for (int i = 0; i < SPINS; ++i)
            if (m_.try_lock()) return;
        m_.lock();
    }
    bool try_lock() { return m_.try_lock(); }
    void unlock() { m_.unlock(); }

  private:
    static constexpr int SPINS = 4000;
    std::mutex m_;
};

template <class Backend>
struct ServeCore {
    Backend be;
    float beta = 0.5f;
    int iterations = 8;
    int max_batch = 256;
    int slots = 0;
    // host-visible arenas: decoder-input staging (host view, device view),
    // hard decisions and iteration counts per slot, the per-block CRC-24B
    // flags of transport-block code blocks
    float *h_llr = nullptr, *h_llr_dev = nullptr;
    uint8_t *h_hard = nullptr;
    int *h_it = nullptr;
    unsigned *h_cbok = nullptr;
    std::vector<SynCode> codes;  // bg * 51 + z index
    struct Slot {
        const float *src = nullptr;  // device-accessible input
        int code = -1, n = 0, nfill = 0;
        int tb = -1, tb_ci = 0;      // transport block (resident mode) and index in it
        uint64_t tag = 0;
        double t_submit = 0, t_done = 0;
        int state = 0;  // 0 free, 1 acquired, 2 submitted, 3 on gpu, 4 done, 5 delivered
    };
    std::vector<Slot> st;
    std::vector<int> free_list;
    std::deque<int> submitted, done;
    ServerMutex mu;
    // Sizes of submitted, done, rs.tb_done_q and rs.tb_outstanding, stored
    // (sync_counts, under mu) after every change: a thread with nothing to
    // take reads them instead of taking the lock.  A count may be stale;
    // a non-zero one only sends the reader to the lock, where it checks.
    std::atomic<int> n_submitted{0}, n_done{0}, n_tb_done{0}, n_tb_out{0};
    std::atomic<bool> stop{false};
    std::atomic<long long> batches{0}, decoded{0};
    // Sticky: a failed launch puts the server in a failed state, reported by
    // every later submit and by the blocks it could not decode.
    std::atomic<int32_t> error{SYNDROME_OK};

    bool resident = false;
    struct Outstanding {
        unsigned id;
        int n;
        int slot[SYN_WORK_MAX];
        bool fin;  // seen complete (popped once everything before it is)
    };
    struct TbRec {
        std::vector<int> slots;
        uint64_t tag = 0;
        size_t tb_size = 0;
        double t_submit = 0, t_done = 0;
        int state = 0;        // 0 free, 2 submitted, 4 done, 5 delivered
        bool failed = false;  // some block was never decoded (failed launch)
    };
    struct Resident {
        unsigned cap = 0;
        SynWork *ring = nullptr;      // coherent host memory
        SynResidentCtl *ctl = nullptr;
        unsigned *done = nullptr;
        bool running = false;
        unsigned published = 0;
        std::deque<Outstanding> outstanding;  // in publish (= claim) order
        std::vector<int> cws;  // code blocks per item, per code
        int grid = 0;
        unsigned tb_cap = 0, tb_next = 0;
        SynTbWork *tbw = nullptr;
        unsigned *tb_done = nullptr, *tb_ok = nullptr;
        std::vector<TbRec> tbr;
        std::vector<int> tb_free, tb_outstanding;
        std::deque<int> tb_done_q;
    } rs;

    void sync_counts()
    {
        n_submitted.store((int)submitted.size(), std::memory_order_release);
        n_done.store((int)done.size(), std::memory_order_release);
        n_tb_done.store((int)rs.tb_done_q.size(), std::memory_order_release);
        n_tb_out.store((int)rs.tb_outstanding.size(), std::memory_order_release);
    }

    void init_slots(int n)
    {
        slots = n;
        st.assign(n, Slot{});
        free_list.clear();
        for (int i = n - 1; i >= 0; --i) free_list.push_back(i);
    }

    void init_tb_ring(unsigned cap)
    {
        rs.tb_cap = cap;
        rs.tbr.assign(cap, TbRec{});
        rs.tb_free.clear();
        for (int i = (int)cap - 1; i >= 0; --i) rs.tb_free.push_back(i);
    }

    static int code_index(int bg, long long z)
    {
        static constexpr int VALID_Z[51] = {2,   3,   4,   5,   6,   7,   8,   9,   10,  11,  12,  13,  14,
                                            15,  16,  18,  20,  22,  24,  26,  28,  30,  32,  36,  40,  44,
                                            48,  52,  56,  60,  64,  72,  80,  88,  96,  104, 112, 120, 128,
                                            144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384};
        if (bg != 0 && bg != 1) return -1;
        for (int i = 0; i < 51; ++i)
            if (VALID_Z[i] == z) return bg * 51 + i;
        return -1;
    }

    // Marks slots[0..n) with state `mark` if each is in range, in one of the
    // states `want` (and, with no_tb, not part of a transport block) and
    // listed once (a repeat finds its own mark); on any failure every mark
    // is undone and false returned.  Called with the server lock held.
    bool claim_slots(const uint32_t *sl, size_t n, std::initializer_list<int> want, int mark, bool no_tb)
    {
        std::vector<int> prev(n);
        size_t i = 0;
        for (; i < n; ++i) {
            const uint32_t s = sl[i];
            if (s >= (uint32_t)slots) break;
            const int stt = st[s].state;
            bool ok = false;
            for (int w : want) ok = ok || stt == w;
            if (!ok || (no_tb && st[s].tb >= 0)) break;
            prev[i] = stt;
            st[s].state = mark;
        }
        if (i == n) return true;
        while (i-- > 0) st[sl[i]].state = prev[i];
        return false;
    }

    // -- the caller-facing protocol -----------------------------------------

    void take_slot(int slot, int code, int n_filler)
    {
        Slot &s = st[slot];
        s.code = code;
        s.n = codes[code].n_b * codes[code].z;
        s.nfill = n_filler;
        s.state = 1;
        s.src = h_llr_dev + (size_t)slot * SERVE_NMAX;
    }

    int32_t acquire(int32_t bg, size_t z, size_t n_filler, float **out_llr, uint32_t *out_slot)
    {
        const int code = z <= 384 ? code_index(bg, (long long)z) : -1;
        if (code < 0) return SYNDROME_INVALID_PARAM;
        std::lock_guard<ServerMutex> lk(mu);
        if (free_list.empty()) return SYNDROME_BUSY;
        const int slot = free_list.back();
        free_list.pop_back();
        take_slot(slot, code, (int)std::min(n_filler, (size_t)INT_MAX));
        *out_llr = h_llr + (size_t)slot * SERVE_NMAX;
        *out_slot = (uint32_t)slot;
        return SYNDROME_OK;
    }

    int32_t acquire_many(const int32_t *bg, const int32_t *z, const int32_t *n_filler, size_t n, float **out_llr,
                         uint32_t *out_slots, size_t *out_acquired)
    {
        *out_acquired = 0;
        std::vector<int> code(n);
        for (size_t i = 0; i < n; ++i) {
            code[i] = z[i] > 0 && z[i] <= 384 ? code_index(bg[i], z[i]) : -1;
            if (code[i] < 0 || (n_filler && n_filler[i] < 0)) return SYNDROME_INVALID_PARAM;
        }
        std::lock_guard<ServerMutex> lk(mu);
        size_t k = 0;
        for (; k < n && !free_list.empty(); ++k) {
            const int slot = free_list.back();
            free_list.pop_back();
            take_slot(slot, code[k], n_filler ? n_filler[k] : 0);
            if (out_llr) out_llr[k] = h_llr + (size_t)slot * SERVE_NMAX;
            out_slots[k] = (uint32_t)slot;
        }
        *out_acquired = k;
        return k == 0 && n ? SYNDROME_BUSY : SYNDROME_OK;
    }

    int32_t submit(uint32_t slot, const float *src, uint64_t tag)
    {
        if (int32_t e = error.load()) return e;
        std::lock_guard<ServerMutex> lk(mu);
        if (slot >= (uint32_t)slots || st[slot].state != 1) return SYNDROME_INVALID_PARAM;
        Slot &sl = st[slot];
        if (src) sl.src = src;
        sl.tag = tag;
        sl.t_submit = now_ns();
        sl.state = 2;
        submitted.push_back((int)slot);
        sync_counts();
        return SYNDROME_OK;
    }

    int32_t submit_many(const uint32_t *sl, const float *const *srcs, const uint64_t *tags, size_t n)
    {
        if (int32_t e = error.load()) return e;
        std::lock_guard<ServerMutex> lk(mu);
        if (!claim_slots(sl, n, {1}, -2, false)) return SYNDROME_INVALID_PARAM;
        const double t = now_ns();
        for (size_t i = 0; i < n; ++i) {
            Slot &s = st[sl[i]];
            if (srcs && srcs[i]) s.src = srcs[i];
            s.tag = tags ? tags[i] : 0;
            s.t_submit = t;
            s.state = 2;
            submitted.push_back((int)sl[i]);
        }
        sync_counts();
        return SYNDROME_OK;
    }

    void fill_served(int slot, SyndromeServed &o)
    {
        Slot &sl = st[slot];
        sl.state = 5;
        o.slot = (uint32_t)slot;
        o.tag = sl.tag;
        o.iterations = h_it[slot];
        o.hard = h_hard + (size_t)slot * SERVE_NMAX;
        o.n = (size_t)sl.n;
        o.submit_ns = sl.t_submit;
        o.done_ns = sl.t_done;
    }

    int32_t try_recv(SyndromeServed *out)
    {
        if (n_done.load(std::memory_order_acquire) == 0) return SYNDROME_BUSY;
        std::lock_guard<ServerMutex> lk(mu);
        if (done.empty()) return SYNDROME_BUSY;
        const int slot = done.front();
        done.pop_front();
        sync_counts();
        fill_served(slot, *out);
        return out->iterations < 0 ? SYNDROME_DEVICE_ERROR : SYNDROME_OK;
    }

    int32_t try_recv_many(SyndromeServed *out, size_t cap, size_t *out_n)
    {
        *out_n = 0;
        if (n_done.load(std::memory_order_acquire) == 0) return SYNDROME_BUSY;
        std::lock_guard<ServerMutex> lk(mu);
        size_t k = 0;
        bool failed = false;
        for (; k < cap && !done.empty(); ++k) {
            const int slot = done.front();
            done.pop_front();
            fill_served(slot, out[k]);
            failed = failed || out[k].iterations < 0;
        }
        sync_counts();
        *out_n = k;
        if (!k) return SYNDROME_BUSY;
        return failed ? SYNDROME_DEVICE_ERROR : SYNDROME_OK;
    }

    int32_t release(uint32_t slot)
    {
        std::lock_guard<ServerMutex> lk(mu);
        if (slot >= (uint32_t)slots || (st[slot].state != 5 && st[slot].state != 1) || st[slot].tb >= 0)
            return SYNDROME_INVALID_PARAM;
        st[slot].state = 0;
        free_list.push_back((int)slot);
        return SYNDROME_OK;
    }

    int32_t release_many(const uint32_t *sl, size_t n)
    {
        std::lock_guard<ServerMutex> lk(mu);
        if (!claim_slots(sl, n, {1, 5}, -3, true)) return SYNDROME_INVALID_PARAM;
        for (size_t i = 0; i < n; ++i) {
            st[sl[i]].state = 0;
            free_list.push_back((int)sl[i]);
        }
        return SYNDROME_OK;
    }

    int32_t submit_tb(const uint32_t *sl, const float *const *srcs, size_t num_cb, size_t tb_size, uint8_t *tb_out,
                      uint64_t tag)
    {
        if (!resident || num_cb == 0 || num_cb > (size_t)slots || tb_size == 0 || tb_size > (size_t)INT_MAX)
            return SYNDROME_INVALID_PARAM;
        if (int32_t e = error.load()) return e;
        std::lock_guard<ServerMutex> lk(mu);
        const Slot &f = st[sl[0] < (uint32_t)slots ? sl[0] : 0];
        for (size_t i = 0; i < num_cb; ++i) {
            if (sl[i] >= (uint32_t)slots) return SYNDROME_INVALID_PARAM;
            const Slot &x = st[sl[i]];
            if (x.state != 1 || x.code != f.code || x.nfill != f.nfill) return SYNDROME_INVALID_PARAM;
            for (size_t j = 0; j < i; ++j)
                if (sl[j] == sl[i]) return SYNDROME_INVALID_PARAM;
        }
        const SynCode &c = codes[f.code];
        const long long kp = (long long)(c.n_b - c.m_b) * c.z - f.nfill;
        const long long p = num_cb > 1 ? kp - 24 : kp;
        if (p < 1 || (long long)num_cb * p < (long long)tb_size + 24) return SYNDROME_INVALID_PARAM;
        if (rs.tb_free.empty()) return SYNDROME_BUSY;
        const int ti = rs.tb_free.back();
        rs.tb_free.pop_back();
        const unsigned id = rs.tb_next++;
        rs.tbw[ti] = SynTbWork{id, (int)num_cb, (int)p, (int)tb_size, tb_out};
        TbRec &r = rs.tbr[ti];
        r.slots.assign(sl, sl + num_cb);
        r.tag = tag;
        r.tb_size = tb_size;
        r.t_submit = now_ns();
        r.state = 2;
        r.failed = false;
        for (size_t i = 0; i < num_cb; ++i) {
            Slot &x = st[sl[i]];
            if (srcs && srcs[i]) x.src = srcs[i];
            x.tb = ti;
            x.tb_ci = (int)i;
            x.tag = tag;
            x.t_submit = r.t_submit;
            x.state = 2;
            submitted.push_back((int)sl[i]);
        }
        rs.tb_outstanding.push_back(ti);
        sync_counts();
        return SYNDROME_OK;
    }

    int32_t try_recv_tb(SyndromeServedTb *out)
    {
        if (n_tb_done.load(std::memory_order_acquire) == 0) return SYNDROME_BUSY;
        std::lock_guard<ServerMutex> lk(mu);
        if (rs.tb_done_q.empty()) return SYNDROME_BUSY;
        const int ti = rs.tb_done_q.front();
        rs.tb_done_q.pop_front();
        sync_counts();
        TbRec &r = rs.tbr[ti];
        r.state = 5;
        int its = 0;
        for (int sl : r.slots) its = std::max(its, h_it[sl]);
        out->tb = (uint32_t)ti;
        out->tag = r.tag;
        out->crc_ok = !r.failed && __atomic_load_n(rs.tb_ok + ti, __ATOMIC_ACQUIRE) ? 1 : 0;
        out->max_iterations = r.failed ? -1 : its;
        out->tb_size = r.tb_size;
        out->num_cb = r.slots.size();
        out->submit_ns = r.t_submit;
        out->done_ns = r.t_done;
        return r.failed ? SYNDROME_DEVICE_ERROR : SYNDROME_OK;
    }

    // CRC-24B outcome of each code block of a received transport block
    // (DecodeReport::cb_crc, transport_block.rs:140-149): 1 = passed, or the
    // transport block has a single block (no CRC-24B); 0 = failed or never
    // decoded.
    int32_t tb_block_crc(uint32_t tb, uint8_t *out, size_t cap, size_t *out_n)
    {
        std::lock_guard<ServerMutex> lk(mu);
        if (!resident || tb >= rs.tb_cap || rs.tbr[tb].state != 5) return SYNDROME_INVALID_PARAM;
        const TbRec &r = rs.tbr[tb];
        *out_n = r.slots.size();
        if (cap < r.slots.size()) return SYNDROME_BUFFER_TOO_SMALL;
        for (size_t i = 0; i < r.slots.size(); ++i)
            out[i] = !r.failed && __atomic_load_n(h_cbok + r.slots[i], __ATOMIC_ACQUIRE) ? 1 : 0;
        return SYNDROME_OK;
    }

    int32_t release_tb(uint32_t tb)
    {
        std::lock_guard<ServerMutex> lk(mu);
        if (!resident || tb >= rs.tb_cap || rs.tbr[tb].state != 5) return SYNDROME_INVALID_PARAM;
        TbRec &r = rs.tbr[tb];
        for (int sl : r.slots) {
            st[sl].tb = -1;
            st[sl].state = 0;
            free_list.push_back(sl);
        }
        r.slots.clear();
        r.state = 0;
        rs.tb_free.push_back((int)tb);
        return SYNDROME_OK;
    }

    // -- the resident dispatcher ---------------------------------------------

    // A block that will not be decoded: reported with iterations -1, through
    // its transport block when it has one.  Called with the lock held.
    void fail_block(int sl, double t)
    {
        st[sl].state = 4;
        st[sl].t_done = t;
        h_it[sl] = -1;
        if (st[sl].tb >= 0) rs.tbr[st[sl].tb].failed = true;
        else done.push_back(sl);
    }

    // The failure path of a launch: the server turns failed (sticky), and no
    // grid runs again.  Items published but unfinished are marked dead in
    // the ring (count 0) and dropped; they and every block still waiting for
    // publication are delivered as failed, and so are the transport blocks
    // they belong to.
    void fail_resident(int32_t rc)
    {
        std::lock_guard<ServerMutex> lk(mu);
        error.store(rc);
        const double t = now_ns();
        for (const Outstanding &o : rs.outstanding) {
            if (o.fin) continue;
            __atomic_store_n(&rs.ring[o.id % rs.cap].count, 0, __ATOMIC_RELAXED);
            for (int j = 0; j < o.n; ++j) fail_block(o.slot[j], t);
        }
        rs.outstanding.clear();
        while (!submitted.empty()) {
            fail_block(submitted.front(), t);
            submitted.pop_front();
        }
        tb_check(t);
        sync_counts();
    }

    // Transport blocks that are decided (or failed) and whose blocks are all
    // seen complete go to the delivery queue.  Called with the lock held.
    bool tb_check(double t)
    {
        bool progressed = false;
        for (size_t k = 0; k < rs.tb_outstanding.size();) {
            const int ti = rs.tb_outstanding[k];
            TbRec &r = rs.tbr[ti];
            bool blocks_done = true;  // and every block's item seen complete, so release_tb can free them
            for (int sl : r.slots) blocks_done = blocks_done && st[sl].state == 4;
            if (blocks_done && (r.failed || __atomic_load_n(rs.tb_done + ti, __ATOMIC_ACQUIRE) == rs.tbw[ti].id + 1u)) {
                r.state = 4;
                r.t_done = t;
                rs.tb_done_q.push_back(ti);
                rs.tb_outstanding[k] = rs.tb_outstanding.back();
                rs.tb_outstanding.pop_back();
                progressed = true;
            } else {
                ++k;
            }
        }
        return progressed;
    }

    // Completions: each outstanding item's done word, in publish order,
    // stopping after the (grid + 1)-th unfinished one; finished items are
    // popped from the front.
    bool scan_outstanding()
    {
        bool progressed = false;
        int open_seen = 0;
        double t = 0;
        std::unique_lock<ServerMutex> lk(mu, std::defer_lock);
        for (size_t k = 0; k < rs.outstanding.size() && open_seen <= rs.grid; ++k) {
            Outstanding &o = rs.outstanding[k];
            if (o.fin) continue;
            if (__atomic_load_n(rs.done + o.id % rs.cap, __ATOMIC_ACQUIRE) != o.id + 1u) {
                ++open_seen;
                continue;
            }
            if (!lk.owns_lock()) {
                t = now_ns();
                lk.lock();
            }
            for (int j = 0; j < o.n; ++j) {
                st[o.slot[j]].state = 4;
                st[o.slot[j]].t_done = t;
                if (st[o.slot[j]].tb < 0) done.push_back(o.slot[j]);  // TB blocks are delivered per TB
            }
            decoded += o.n;
            o.fin = true;
            progressed = true;
        }
        if (lk.owns_lock()) sync_counts();
        while (!rs.outstanding.empty() && rs.outstanding.front().fin) rs.outstanding.pop_front();
        return progressed;
    }

    // Pending submissions grouped by code into items (no waiting: an item
    // holds what is pending now) and published; what does not fit in the
    // ring goes back to the front of the queue.
    bool publish_pending(std::vector<int> &pend, std::vector<std::vector<int>> &by_code)
    {
        pend.clear();
        if (n_submitted.load(std::memory_order_acquire) > 0) {
            std::lock_guard<ServerMutex> lk(mu);
            while (!submitted.empty()) {
                pend.push_back(submitted.front());
                submitted.pop_front();
            }
            sync_counts();
        }
        if (pend.empty()) return false;
        for (int sl : pend) by_code[st[sl].code].push_back(sl);
        std::vector<int> back;
        for (int sl : pend) {
            std::vector<int> &v = by_code[st[sl].code];
            if (v.empty()) continue;
            const int per = rs.cws[st[sl].code];
            size_t j = 0;
            for (; j < v.size(); j += (size_t)per) {
                if (rs.outstanding.size() + 1 >= rs.cap) break;
                const int cnt = (int)std::min(v.size() - j, (size_t)per);
                SynWork &wk = rs.ring[rs.published % rs.cap];
                Outstanding o{rs.published, cnt, {}, false};
                int aligned = 1;
                for (int q = 0; q < cnt; ++q) {
                    const Slot &x = st[v[j + q]];
                    wk.slot[q] = v[j + q];
                    wk.nfill[q] = x.nfill;
                    wk.src[q] = x.src;
                    wk.tb[q] = x.tb;
                    wk.tb_ci[q] = x.tb_ci;
                    aligned &= ((uintptr_t)x.src & 15u) == 0;
                    o.slot[q] = v[j + q];
                }
                wk.code = st[v[j]].code;
                wk.count = cnt;
                wk.aligned = aligned;
                rs.outstanding.push_back(o);
                ++rs.published;
            }
            for (; j < v.size(); ++j) back.push_back(v[j]);
            v.clear();
        }
        {
            std::lock_guard<ServerMutex> lk(mu);
            for (int sl : pend) st[sl].state = 3;
            for (auto it = back.rbegin(); it != back.rend(); ++it) {
                st[*it].state = 2;
                submitted.push_front(*it);
            }
            sync_counts();
        }
        __atomic_store_n(&rs.ctl->tail, rs.published, __ATOMIC_RELEASE);
        return true;
    }

    // The resident dispatcher loop: completions, transport blocks, pending
    // submissions, and a relaunch whenever the grid has left while items are
    // outstanding.  Once the grid is seen finished, every item it claimed has
    // its done word set, so items still outstanding after the scan that
    // follows were never claimed.
    void run_resident()
    {
        std::vector<int> pend;
        std::vector<std::vector<int>> by_code(codes.size());
        while (!stop.load()) {
            bool progressed = false;
            if (rs.running) {
                const int q = be.query();
                if (q != 0) rs.running = false;
                if (q < 0) {
                    fail_resident(SYNDROME_DEVICE_ERROR);
                    progressed = true;
                }
            }
            progressed = scan_outstanding() || progressed;
            if (n_tb_out.load(std::memory_order_acquire) > 0) {
                std::lock_guard<ServerMutex> lk(mu);
                progressed = tb_check(now_ns()) || progressed;
                sync_counts();
            }
            if (error.load() == SYNDROME_OK) {
                progressed = publish_pending(pend, by_code) || progressed;
                if (!rs.running && !rs.outstanding.empty()) {
                    if (int32_t rc = be.launch(*this)) fail_resident(rc);
                    else {
                        rs.running = true;
                        batches++;
                    }
                    progressed = true;
                }
            } else if (n_submitted.load(std::memory_order_acquire) > 0) {
                fail_resident(error.load());  // submitted before the failure was seen
                progressed = true;
            }
            if (!progressed) std::this_thread::yield();
        }
        __atomic_store_n(&rs.ctl->stop, 1u, __ATOMIC_RELEASE);
        if (rs.running) be.wait();
    }
};

}  // namespace syn
