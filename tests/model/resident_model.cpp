// CPU thread model of the resident serving decoder's protocol (test
// infrastructure, built with -fsanitize=thread by tests/test_resident_model.py).
//
// It runs the product's own code on both sides of the queue:
//  * the host side is ServeCore (syndrome_amd/csrc/serve_core.h), the
//    server's dispatcher and caller-facing protocol, unchanged;
//  * each "workgroup" of a resident grid is a host thread running the
//    kernel's protocol steps from resident_protocol.h (resident_claim,
//    tb_account, item_complete) in the kernel's order, with the decode
//    replaced by a function of the block's input, and with random delays
//    between every protocol step.
// Client threads acquire slots, submit code blocks and transport blocks,
// receive, check and release them while grids come and go (a short idle
// limit makes every lull end the grid, so relaunches race with
// publication).  The model stops at its first failure: a result delivered
// twice, a wrong result, a transport block with the wrong decision, per-block
// CRC flags or payload, or nothing delivered for too long.  With
// --fail-launch, one relaunch fails, and every block and transport block
// still pending must then be delivered once, as failed, and submissions
// refused.
//
// Fidelity: the device's fences (tb_account's release / seq_cst fences
// around relaxed read-modify-writes) are modelled by the same fences plus
// acquire-release read-modify-writes, which ThreadSanitizer can follow; the
// scoped atomics become plain C++ atomics on the same words.
#define SYN_HD
#include <stdio.h>
#include <stdlib.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <map>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "resident_protocol.h"
#include "serve_core.h"

namespace {

using Clock = std::chrono::steady_clock;

[[noreturn]] void fail(const std::string &msg)
{
    fprintf(stderr, "MODEL FAILURE: %s\n", msg.c_str());
    fflush(stderr);
    abort();
}

thread_local std::mt19937_64 tl_rng{12345};

// The protocol's atomics over plain memory, with a random delay before each.
struct HostOps {
    static void jitter()
    {
        const unsigned r = (unsigned)(tl_rng() & 63);
        if (r < 6) std::this_thread::yield();
        else if (r == 6) std::this_thread::sleep_for(std::chrono::microseconds(tl_rng() % 50));
    }
    template <class T>
    static std::atomic_ref<T> ref(const T *p)
    {
        return std::atomic_ref<T>(*const_cast<T *>(p));
    }
    unsigned load_relaxed(const unsigned *p) const { jitter(); return ref(p).load(std::memory_order_relaxed); }
    long long load_relaxed(const long long *p) const { jitter(); return ref(p).load(std::memory_order_relaxed); }
    unsigned load_acquire(const unsigned *p) const { jitter(); return ref(p).load(std::memory_order_acquire); }
    unsigned load_acquire_system(const unsigned *p) const { jitter(); return ref(p).load(std::memory_order_acquire); }
    unsigned load_relaxed_system(const unsigned *p) const { jitter(); return ref(p).load(std::memory_order_relaxed); }
    void store_relaxed(unsigned *p, unsigned v) const { jitter(); ref(p).store(v, std::memory_order_relaxed); }
    void store_relaxed(long long *p, long long v) const { jitter(); ref(p).store(v, std::memory_order_relaxed); }
    void store_release(unsigned *p, unsigned v) const { jitter(); ref(p).store(v, std::memory_order_release); }
    void store_release_system(unsigned *p, unsigned v) const { jitter(); ref(p).store(v, std::memory_order_release); }
    bool try_lock(unsigned *p) const
    {
        jitter();
        unsigned e = 0;
        return ref(p).compare_exchange_strong(e, 1u, std::memory_order_acquire, std::memory_order_relaxed);
    }
    void unlock(unsigned *p) const { jitter(); ref(p).store(0u, std::memory_order_release); }
    bool cas_relaxed(unsigned *p, unsigned expect, unsigned want) const
    {
        jitter();
        return ref(p).compare_exchange_strong(expect, want, std::memory_order_relaxed, std::memory_order_relaxed);
    }
    void fence_acquire() const { std::atomic_thread_fence(std::memory_order_acquire); }
    void fence_release_system() const { std::atomic_thread_fence(std::memory_order_release); }
    void fence_seq_cst() const { std::atomic_thread_fence(std::memory_order_seq_cst); }
    unsigned atomic_xor(unsigned *p, unsigned v) const { jitter(); return ref(p).fetch_xor(v, std::memory_order_acq_rel); }
    unsigned atomic_or(unsigned *p, unsigned v) const { jitter(); return ref(p).fetch_or(v, std::memory_order_acq_rel); }
    unsigned atomic_add(unsigned *p, unsigned v) const { jitter(); return ref(p).fetch_add(v, std::memory_order_acq_rel); }
    unsigned atomic_exch(unsigned *p, unsigned v) const { jitter(); return ref(p).exchange(v, std::memory_order_acq_rel); }
    long long clock() const
    {
        return (long long)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
    }
    void backoff(int n) const { std::this_thread::sleep_for(std::chrono::microseconds(n)); }
};

// What a block's "decode" yields: a function of the first words of its input.
inline uint8_t hard_of(unsigned token, int i) { return (uint8_t)((token * 2654435761u >> (8 * i)) & 0xff); }
inline int iters_of(unsigned token) { return 1 + (int)(token % 7); }
inline unsigned share_of(unsigned tb_id, int ci) { return ((tb_id + 1) * 0x9E3779B1u ^ (ci * 0x85EBCA77u)) & 0xffffffu; }
inline uint8_t payload_of(unsigned tb_id, int ci) { return (uint8_t)(1 + ((tb_id * 31 + ci) % 250)); }

// Input words a client writes into a block's decoder input (floats hold
// small integers exactly): [0] token, [1] CRC-24B outcome, [2] received
// CRC-24A bits it carries (transport blocks only).
constexpr int IN_TOKEN = 0, IN_CBOK = 1, IN_RX = 2;

struct Model;

// One resident grid: G threads, each a workgroup running the kernel's loop.
struct ModelGrid {
    Model *m = nullptr;
    std::vector<std::thread> wg;
    std::atomic<int> live{0};
    int launches = 0;
    int fail_at = -1;  // the launch number that fails (--fail-launch)
    template <class Core>
    int32_t launch(Core &);
    int query()
    {
        if (live.load(std::memory_order_acquire) != 0) return 0;
        for (auto &t : wg) t.join();
        wg.clear();
        return 1;
    }
    void wait()
    {
        for (auto &t : wg) t.join();
        wg.clear();
    }
};

using Core = syn::ServeCore<ModelGrid>;

struct Model {
    Core core;
    // "device memory"
    ResidentQueue q{};
    long long last_claim = 0;
    std::vector<SynTbState> tbstate;
    long long idle_ns = 300000, life_ns = 60000000000LL;
    int grid = 4;
    // host memory the model allocates for the core
    std::vector<float> llr;
    std::vector<uint8_t> hard;
    std::vector<int> iters;
    std::vector<unsigned> cbok, done, tb_done, tb_ok;
    std::vector<SynWork> ring;
    SynResidentCtl ctl{};
    std::vector<SynTbWork> tbw;

    void wg_main()
    {
        HostOps ops;
        const long long t0 = ops.clock();
        for (;;) {
            unsigned w = 0;
            if (!syn::resident_claim(ops, &q, &ctl, &last_claim, t0, idle_ns, life_ns, w)) break;
            const SynWork *wk = core.rs.ring + w % core.rs.cap;
            const int count = HostOps::ref(&wk->count).load(std::memory_order_relaxed);
            for (int j = 0; j < count; ++j) {
                const int slot = HostOps::ref(&wk->slot[j]).load(std::memory_order_relaxed);
                const float *src = (const float *)HostOps::ref(&wk->src[j]).load(std::memory_order_relaxed);
                const int tb = HostOps::ref(&wk->tb[j]).load(std::memory_order_relaxed);
                const int ci = HostOps::ref(&wk->tb_ci[j]).load(std::memory_order_relaxed);
                const unsigned token = (unsigned)src[IN_TOKEN];
                HostOps::jitter();
                core.h_it[slot] = iters_of(token);
                if (tb < 0) {
                    for (int i = 0; i < 4; ++i) core.h_hard[(size_t)slot * syn::SERVE_NMAX + i] = hard_of(token, i);
                    continue;
                }
                // the transport block's share of this block (tb_fold)
                const SynTbWork *tw = core.rs.tbw + tb;
                const unsigned id = HostOps::ref(&tw->id).load(std::memory_order_relaxed);
                const int c = HostOps::ref(&tw->num_cb).load(std::memory_order_relaxed);
                unsigned char *out = HostOps::ref(&tw->out).load(std::memory_order_relaxed);
                const bool cb_ok = src[IN_CBOK] != 0.f;
                out[ci] = payload_of(id, ci);
                core.h_cbok[slot] = cb_ok ? 1u : 0u;
                syn::tb_account(ops, &tbstate[tb], share_of(id, ci), (unsigned)src[IN_RX], cb_ok, c, id,
                                core.rs.tb_ok + tb, core.rs.tb_done + tb);
            }
            ops.fence_release_system();  // every wave's stores complete before the barrier
            syn::item_complete(ops, core.rs.done, core.rs.cap, w);
        }
    }
};

template <class CoreT>
int32_t ModelGrid::launch(CoreT &)
{
    if (launches++ == fail_at) return SYNDROME_DEVICE_ERROR;
    live.store(m->grid, std::memory_order_release);
    const int L = launches;
    for (int i = 0; i < m->grid; ++i)
        wg.emplace_back([this, i, L] {
            tl_rng.seed(0x51ed27u * (unsigned)(L * 131 + i));
            m->wg_main();
            live.fetch_sub(1, std::memory_order_acq_rel);
        });
    return SYNDROME_OK;
}

struct Expect {
    std::atomic<int> delivered{0};
    bool is_tb = false;
    unsigned token = 0;                 // code block
    int num_cb = 0;                     // transport block
    bool ok = false;
    std::vector<uint8_t> cb_ok;
    std::vector<uint8_t> out;           // its payload destination (one byte per code block)
};

}  // namespace

int main(int argc, char **argv)
{
    double seconds = 3.0;
    unsigned seed = 1;
    bool fail_launch = false;
    int clients = 3;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--seconds" && i + 1 < argc) seconds = atof(argv[++i]);
        else if (a == "--seed" && i + 1 < argc) seed = (unsigned)atoi(argv[++i]);
        else if (a == "--fail-launch") fail_launch = true;
        else if (a == "--clients" && i + 1 < argc) clients = atoi(argv[++i]);
    }
    Model M;
    Core &core = M.core;
    core.be.m = &M;
    if (fail_launch) core.be.fail_at = 3 + (int)(seed % 5);
    // three codes are enough for the protocol: BG1 Z=384, BG2 Z=208, BG2 Z=52
    core.codes.resize(102);
    for (int c = 0; c < 102; ++c) {
        core.codes[c].n_b = c < 51 ? 68 : 52;
        core.codes[c].m_b = c < 51 ? 46 : 42;
        core.codes[c].z = 2;
    }
    const int code_ids[3] = {Core::code_index(0, 384), Core::code_index(1, 208), Core::code_index(1, 52)};
    core.codes[code_ids[0]].z = 384;
    core.codes[code_ids[1]].z = 208;
    core.codes[code_ids[2]].z = 52;
    const int SLOTS = 48;
    M.llr.assign((size_t)SLOTS * syn::SERVE_NMAX, 0.f);
    M.hard.assign((size_t)SLOTS * syn::SERVE_NMAX, 0);
    M.iters.assign(SLOTS, 0);
    M.cbok.assign(SLOTS, 0);
    core.h_llr = core.h_llr_dev = M.llr.data();
    core.h_hard = M.hard.data();
    core.h_it = M.iters.data();
    core.h_cbok = M.cbok.data();
    core.init_slots(SLOTS);
    core.resident = true;
    core.rs.cap = 64;  // small: the ring wraps often
    M.ring.assign(core.rs.cap, SynWork{});
    M.done.assign(core.rs.cap, 0);
    core.rs.ring = M.ring.data();
    core.rs.ctl = &M.ctl;
    core.rs.done = M.done.data();
    core.rs.grid = M.grid;
    core.rs.cws.assign(core.codes.size(), 3);  // up to 3 blocks per item
    const unsigned TBC = 16;  // transport-block records: reused often
    M.tbw.assign(TBC, SynTbWork{});
    M.tb_done.assign(TBC, 0);
    M.tb_ok.assign(TBC, 0);
    M.tbstate.assign(TBC, SynTbState{});
    core.rs.tbw = M.tbw.data();
    core.rs.tb_done = M.tb_done.data();
    core.rs.tb_ok = M.tb_ok.data();
    core.init_tb_ring(TBC);

    std::thread dispatcher([&] {
        tl_rng.seed(seed * 7 + 3);
        core.run_resident();
    });

    std::mutex emu;
    std::map<uint64_t, Expect> expect;  // by tag
    std::atomic<uint64_t> next_tag{1};
    std::atomic<long long> delivered{0}, failed_seen{0}, refused{0};
    std::atomic<long long> last_progress{(long long)Clock::now().time_since_epoch().count()};
    const auto t_end = Clock::now() + std::chrono::duration<double>(seconds);
    std::atomic<bool> submitting{true};

    auto receive = [&](int) {
        SyndromeServed r;
        for (;;) {
            const int32_t rc = core.try_recv(&r);
            if (rc == SYNDROME_BUSY) break;
            std::lock_guard<std::mutex> lk(emu);
            auto it = expect.find(r.tag);
            if (it == expect.end() || it->second.is_tb) fail("code block with an unknown tag " + std::to_string(r.tag));
            Expect &e = it->second;
            if (e.delivered.fetch_add(1) != 0) fail("code block delivered twice, tag " + std::to_string(r.tag));
            if (rc == SYNDROME_OK) {
                if (r.iterations != iters_of(e.token)) fail("wrong iteration count");
                for (int i = 0; i < 4; ++i)
                    if (r.hard[i] != hard_of(e.token, i)) fail("wrong hard decisions, tag " + std::to_string(r.tag));
            } else if (rc == SYNDROME_DEVICE_ERROR && r.iterations < 0) {
                failed_seen++;
            } else {
                fail("try_recv returned " + std::to_string(rc));
            }
            if (core.release(r.slot) != SYNDROME_OK) fail("release refused");
            delivered++;
            last_progress = (long long)Clock::now().time_since_epoch().count();
        }
        SyndromeServedTb t;
        for (;;) {
            const int32_t rc = core.try_recv_tb(&t);
            if (rc == SYNDROME_BUSY) break;
            std::lock_guard<std::mutex> lk(emu);
            auto it = expect.find(t.tag);
            if (it == expect.end() || !it->second.is_tb) fail("transport block with an unknown tag");
            Expect &e = it->second;
            if (e.delivered.fetch_add(1) != 0) fail("transport block delivered twice, tag " + std::to_string(t.tag));
            std::vector<uint8_t> cbf(e.num_cb);
            size_t n = 0;
            if (core.tb_block_crc(t.tb, cbf.data(), cbf.size(), &n) != SYNDROME_OK || (int)n != e.num_cb)
                fail("tb_block_crc refused");
            if (rc == SYNDROME_OK) {
                if ((t.crc_ok != 0) != e.ok) fail("wrong transport-block decision, tag " + std::to_string(t.tag));
                for (int i = 0; i < e.num_cb; ++i) {
                    if (cbf[i] != e.cb_ok[i]) fail("wrong per-block CRC flag");
                    if (e.out[i] != payload_of((unsigned)(t.tag >> 1), i)) fail("wrong payload byte");
                }
            } else if (rc == SYNDROME_DEVICE_ERROR) {
                if (t.crc_ok) fail("failed transport block reported as passing");
                failed_seen++;
            } else {
                fail("try_recv_tb returned " + std::to_string(rc));
            }
            if (core.release_tb(t.tb) != SYNDROME_OK) fail("release_tb refused");
            delivered++;
            last_progress = (long long)Clock::now().time_since_epoch().count();
        }
    };

    auto client = [&](int ci) {
        tl_rng.seed(seed * 1000 + ci);
        std::mt19937_64 rng(seed * 77 + ci);
        while (submitting.load()) {
            if (Clock::now() >= t_end) break;
            const int kind = (int)(rng() % 3);
            const int code = code_ids[rng() % 3];
            const int bg = code < 51 ? 0 : 1, z = core.codes[code].z;
            if (kind < 2) {  // a code block
                float *buf = nullptr;
                uint32_t slot = 0;
                if (core.acquire(bg, (size_t)z, 0, &buf, &slot) == SYNDROME_OK) {
                    const uint64_t tag = next_tag.fetch_add(1) << 1;
                    const unsigned token = (unsigned)(rng() % 1000000);
                    buf[IN_TOKEN] = (float)token;
                    {
                        std::lock_guard<std::mutex> lk(emu);
                        Expect &e = expect[tag];
                        e.token = token;
                    }
                    const int32_t rc = core.submit(slot, nullptr, tag);
                    if (rc == SYNDROME_DEVICE_ERROR) {
                        std::lock_guard<std::mutex> lk(emu);
                        expect.erase(tag);
                        refused++;
                        (void)core.release(slot);
                    } else if (rc != SYNDROME_OK) fail("submit returned " + std::to_string(rc));
                }
            } else {  // a transport block of 1..5 code blocks
                const int c = 1 + (int)(rng() % 5);
                std::vector<uint32_t> sl;
                for (int i = 0; i < c; ++i) {
                    float *buf = nullptr;
                    uint32_t s = 0;
                    if (core.acquire(bg, (size_t)z, 0, &buf, &s) != SYNDROME_OK) break;
                    sl.push_back(s);
                }
                if ((int)sl.size() < c) {
                    for (uint32_t s : sl) (void)core.release(s);
                    receive(ci);
                    continue;
                }
                Expect e;
                e.is_tb = true;
                e.num_cb = c;
                e.cb_ok.resize(c);
                bool all = true;
                for (int i = 0; i < c; ++i) {
                    e.cb_ok[i] = c == 1 ? 1 : (rng() % 8 != 0);
                    all = all && e.cb_ok[i];
                }
                const bool crc_a = rng() % 6 != 0;  // the CRC-24A result the shares will give
                e.ok = all && crc_a;
                // the server numbers transport blocks in submit order; take
                // the lock across submit so the number is known here
                std::lock_guard<std::mutex> lk(emu);
                const unsigned id = core.rs.tb_next;  // read before submit_tb increments it (single submitter here)
                unsigned rx = 0;
                for (int i = 0; i < c; ++i) rx ^= share_of(id, i);
                if (!crc_a) rx ^= 1u;
                for (int i = 0; i < c; ++i) {
                    float *buf = core.h_llr + (size_t)sl[i] * syn::SERVE_NMAX;
                    buf[IN_TOKEN] = (float)(rng() % 1000000);
                    buf[IN_CBOK] = e.cb_ok[i] ? 1.f : 0.f;
                    buf[IN_RX] = (float)(i == 0 ? rx : 0u);
                }
                const uint64_t t2 = ((uint64_t)id << 1) | 1;  // tag = submission number, marked as a TB
                expect.emplace(std::piecewise_construct, std::forward_as_tuple(t2), std::forward_as_tuple());
                Expect &ee = expect[t2];
                ee.is_tb = true;
                ee.num_cb = c;
                ee.cb_ok = e.cb_ok;
                ee.ok = e.ok;
                ee.out.assign(64, 0);
                const int32_t rc = core.submit_tb(sl.data(), nullptr, (size_t)c, 1, ee.out.data(), t2);
                if (rc == SYNDROME_DEVICE_ERROR || rc == SYNDROME_BUSY) {
                    expect.erase(t2);
                    if (rc == SYNDROME_DEVICE_ERROR) refused++;
                    for (uint32_t s : sl) (void)core.release(s);
                } else if (rc != SYNDROME_OK) fail("submit_tb returned " + std::to_string(rc));
            }
            receive(ci);
        }
    };
    std::vector<std::thread> cl;
    for (int i = 0; i < clients; ++i) cl.emplace_back([&, i] { client(i); });
    for (auto &t : cl) t.join();
    submitting = false;
    // drain
    for (;;) {
        receive(0);
        long long pending = 0;
        {
            std::lock_guard<std::mutex> lk(emu);
            for (auto &kv : expect) pending += kv.second.delivered.load() == 0;
        }
        if (pending == 0) break;
        const long long idle = (long long)Clock::now().time_since_epoch().count() - last_progress.load();
        if (idle > 20000000000LL) fail("nothing delivered for 20 s with " + std::to_string(pending) + " outstanding");
        std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    core.stop = true;
    dispatcher.join();
    M.core.be.wait();
    printf("model ok: seed %u, %lld delivered (%lld failed, %lld submits refused), %d grid launches, sticky error %d\n",
           seed, delivered.load(), failed_seen.load(), refused.load(), core.be.launches, (int)core.error.load());
    if (fail_launch && core.error.load() == SYNDROME_OK && core.be.launches > core.be.fail_at)
        fail("the failed launch was not reported");
    return 0;
}
