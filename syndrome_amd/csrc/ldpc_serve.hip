// Serving loop: continuous batching of code blocks on one GPU.
//
// The caller-facing protocol is the reference pipeline's
// (src/ldpc_pipeline.rs:499-576): acquire a slot, write the code block's
// decoder input into it, submit, poll completions, release.  The protocol,
// admission and the resident dispatcher are the host-side core
// (serve_core.h); this file holds what touches the device.  Behind it:
//
//  * admission by HBM budget: the device input arena is cut into fixed
//    slots sized for the largest code (68 x 384 f32); `budget / slot_bytes`
//    slots exist and acquire refuses with SYNDROME_BUSY when none is free;
//  * a dispatcher thread forms a batch from everything submitted since the
//    last one as soon as one of `streams` launch lanes is free, so requests
//    that arrive while earlier batches run on the GPU are admitted into the
//    next batch and overlap with them on the device;
//  * a batch is two launches on its lane: a staging kernel (decoder inputs
//    from pinned host memory or HBM into the slots, the descriptor into
//    device memory) and the persistent mixed-code decode, which writes hard
//    bits and iteration counts straight into mapped pinned memory;
//    completions are observed by event.
// Each code block's result is exactly what decoding it alone gives.
#include <hip/hip_runtime.h>

#include <limits.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/syndrome_ldpc.h"
#include "code_host.h"
#include "ldpc_common.h"
#include "serve_core.h"

namespace syn {
const void *persist_kernel_f32_big();
const void *resident_kernel();
int resident_lanes();
int resident_pair_lds();
size_t resident_scratch_floats();
int allow_max_dynamic_lds(const void *fn);
}  // namespace syn

struct SyndromeServer;

namespace {

using syn::now_ns;
constexpr int NMAX = syn::SERVE_NMAX;
// Pinned host memory the device reads and writes while kernels run: mapped
// into the device's address space and coherent (never held in device
// caches), so a resident kernel sees the host's writes and the host sees
// the kernel's without a kernel boundary in between.
constexpr unsigned HOST_FLAGS = hipHostMallocMapped | hipHostMallocCoherent;
constexpr unsigned RING_ITEMS = 8192;
constexpr unsigned TB_RING = 4096;                  // transport blocks in flight (resident mode)
constexpr double RESIDENT_IDLE_US = 2000.0;     // grid leaves after this long with nothing claimed
constexpr double RESIDENT_LIFE_S = 600.0;       // and never waits beyond this age
constexpr int VALID_Z[51] = {2,  3,  4,  5,  6,  7,  8,  9,  10,  11,  12,  13,  14,  15,  16,  18,  20,
                             22, 24, 26, 28, 30, 32, 36, 40, 44,  48,  52,  56,  60,  64,  72,  80,  88,
                             96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384};

// Batch descriptor, written by the host into mapped pinned memory:
// items[ni] | ids[nb] | offs[nb] | src[nb] | nfill[nb] | slot[nb] | n[nb].
struct DescLayout {
    unsigned o_items, o_ids, o_offs, o_src, o_nf, o_slot, o_n;
    int ni;
};

constexpr int STAGE_THREADS = 256, STAGE_UNROLL = 8, STAGE_CHUNK = STAGE_THREADS * STAGE_UNROLL;

// One launch stages a batch: block (x, j) copies elements
// [x * STAGE_CHUNK, (x + 1) * STAGE_CHUNK) of entry j's decoder input (mapped
// pinned host memory or HBM) into its HBM slot, every load of a thread issued
// before its stores so that a PCIe round trip is paid once per chunk, not
// once per element; blocks (0, j) also copy the descriptor fields the decode
// reads into device memory, and block (0, 0) resets the work-queue head.
__global__ void __launch_bounds__(STAGE_THREADS) stage(const char *__restrict__ hd, char *__restrict__ dd,
                                                       DescLayout lay, float *__restrict__ dev, int *__restrict__ head)
{
    const int j = blockIdx.y;
    const int n = ((const int *)(hd + lay.o_n))[j];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (j < lay.ni) ((SynItem *)(dd + lay.o_items))[j] = ((const SynItem *)(hd + lay.o_items))[j];
        ((int *)(dd + lay.o_ids))[j] = ((const int *)(hd + lay.o_ids))[j];
        ((long long *)(dd + lay.o_offs))[j] = ((const long long *)(hd + lay.o_offs))[j];
        ((int *)(dd + lay.o_nf))[j] = ((const int *)(hd + lay.o_nf))[j];
        if (j == 0) *head = 0;
    }
    const int i0 = blockIdx.x * STAGE_CHUNK;
    if (i0 >= n) return;
    const float *s = ((const float *const *)(hd + lay.o_src))[j];
    float *d = dev + (long long)((const int *)(hd + lay.o_slot))[j] * NMAX;
    float v[STAGE_UNROLL];
#pragma unroll
    for (int k = 0; k < STAGE_UNROLL; ++k) {
        const int i = i0 + k * STAGE_THREADS + (int)threadIdx.x;
        v[k] = i < n ? s[i] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < STAGE_UNROLL; ++k) {
        const int i = i0 + k * STAGE_THREADS + (int)threadIdx.x;
        if (i < n) d[i] = v[k];
    }
}

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<int> slots;
    // descriptor: host-written mapped copy (h_desc, device view hd_desc) and
    // the device copy the decode reads (d_desc)
    void *h_desc = nullptr, *hd_desc = nullptr, *d_desc = nullptr;
    size_t desc_bytes = 0;
    int *d_head = nullptr;
    int *h_iters = nullptr, *hd_iters = nullptr;  // per batch position, written by the decode
};

int32_t launch_resident(SyndromeServer &s);

// The resident grid on HIP: launched on the server's stream, its end seen
// by an event recorded after it.
struct HipResident {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    template <class Core>
    int32_t launch(Core &c)
    {
        return launch_resident(static_cast<SyndromeServer &>(c));
    }
    int query()
    {
        const hipError_t e = hipEventQuery(ev);
        return e == hipSuccess ? 1 : (e == hipErrorNotReady ? 0 : -1);
    }
    void wait() { (void)hipEventSynchronize(ev); }
};

}  // namespace

struct SyndromeServer : syn::ServeCore<HipResident> {
    int device = 0;
    int lds_limit = 0, cus = 0;
    // decoder inputs in HBM (d_llr); device views of the pinned arenas
    float *d_llr = nullptr;
    uint8_t *h_hard_dev = nullptr;
    int *h_it_dev = nullptr;
    unsigned *h_cbok_dev = nullptr;
    SynCode *d_codes = nullptr;  // 102 codes: bg * 51 + z index
    std::vector<Lane> lanes;
    std::thread worker;
    // resident mode: device views of the ring and records, device-side state
    struct ResidentDev {
        SynWork *ring_dev = nullptr;
        SynResidentCtl *ctl_dev = nullptr;
        unsigned *done_dev = nullptr;
        unsigned *d_head = nullptr;
        long long *d_last = nullptr, *d_offs = nullptr;
        float *d_xscratch = nullptr;  // extension columns of two-codeword items, per workgroup
        bool xpair = false;           // BG1 Z=384 items hold two blocks (the two-codeword kernel)
        size_t lds = 0;
        long long idle_ticks = 0, life_ticks = 0;
        SynTbWork *tbw_dev = nullptr;
        SynTbState *d_tbstate = nullptr;
        unsigned *tb_done_dev = nullptr, *tb_ok_dev = nullptr;
    } rd;

    int32_t launch_batch(Lane &ln, std::vector<int> &batch);
    int32_t init_resident();
    void run();
};

namespace {

template <class F>
int32_t guard(F &&body)
{
    try {
        return body();
    } catch (...) {
        return SYNDROME_PANIC;
    }
}

struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceScope()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

size_t group_bytes(const SynCode &c, int cws)
{
    const size_t n = syn::code_n(c);
    return (size_t)(((c.z * 4) + 15) & ~15) + ((cws * n * 4 + 3) & ~(size_t)3) + 12 * (size_t)cws;
}


// One resident grid on the server's stream (HipResident::launch).
int32_t launch_resident(SyndromeServer &s)
{
    SynLaunch a{};
    a.hard = s.h_hard_dev;
    a.iters = s.h_it_dev;
    a.iterations = s.iterations;
    a.flags = SYN_F_INIT_5G | SYN_F_WRITE_HARD | (s.rd.xpair ? SYN_F_XPAIR : 0u);  // SYN_F_ALIGNED per item
    a.beta = s.beta;
    const SynCode *cp = s.d_codes;
    const SynWork *ring = s.rd.ring_dev;
    unsigned cap = s.rs.cap;
    SynResidentCtl *ctl = s.rd.ctl_dev;
    unsigned *dn = s.rd.done_dev;
    void *queue = s.rd.d_head;  // head | tail copy | stop copy | poll lock
    long long *last = s.rd.d_last;
    const long long *offs = s.rd.d_offs;
    long long idle = s.rd.idle_ticks, life = s.rd.life_ticks;
    SynResidentTb tbs{s.rd.tbw_dev, s.rd.d_tbstate, s.rd.tb_done_dev, s.rd.tb_ok_dev, s.h_cbok_dev};
    float *xs = s.rd.d_xscratch;
    void *args[] = {&cp, &ring, &cap, &ctl, &dn, &queue, &last, &offs, &a, &idle, &life, &tbs, &xs};
    DeviceScope ds(s.device);
    if (hipLaunchKernel(syn::resident_kernel(), dim3(s.rs.grid), dim3(syn::resident_lanes()), args, s.rd.lds,
                        s.be.stream) != hipSuccess ||
        hipEventRecord(s.be.ev, s.be.stream) != hipSuccess)
        return SYNDROME_DEVICE_ERROR;
    return SYNDROME_OK;
}

}  // namespace

// Build the batch's work items (grouped by code, as many codewords per item
// as fit 512 lanes and the LDS, longest first) and issue its two launches.
int32_t SyndromeServer::launch_batch(Lane &ln, std::vector<int> &batch)
{
    const int nb = (int)batch.size();
    std::vector<std::vector<int>> by_code(codes.size());
    for (int j = 0; j < nb; ++j) by_code[st[batch[j]].code].push_back(j);
    std::vector<SynItem> items;
    std::vector<int> ids;
    size_t lds = 0;
    std::vector<std::pair<long long, SynItem>> tmp;
    for (size_t c = 0; c < codes.size(); ++c) {
        if (by_code[c].empty()) continue;
        int k = std::max(1, 512 / codes[c].z);
        while (k > 1 && group_bytes(codes[c], k) > (size_t)lds_limit) --k;
        lds = std::max(lds, group_bytes(codes[c], k));
        for (size_t j = 0; j < by_code[c].size(); j += (size_t)k) {
            const int cnt = (int)std::min(by_code[c].size() - j, (size_t)k);
            SynItem it{(int)c, codes[c].kind, (int)ids.size(), cnt};
            for (int q = 0; q < cnt; ++q) ids.push_back(by_code[c][j + q]);
            tmp.push_back({(long long)codes[c].n_edges * codes[c].z * cnt, it});
        }
    }
    std::stable_sort(tmp.begin(), tmp.end(), [](auto &a, auto &b) { return a.first > b.first; });
    for (auto &x : tmp) items.push_back(x.second);
    const int ni = (int)items.size();
    DescLayout lay;
    lay.ni = ni;
    lay.o_items = 0;
    lay.o_ids = (unsigned)(sizeof(SynItem) * ni);
    lay.o_offs = (lay.o_ids + 4u * nb + 7u) & ~7u;
    lay.o_src = lay.o_offs + 8u * nb;
    lay.o_nf = lay.o_src + 8u * nb;
    lay.o_slot = lay.o_nf + 4u * nb;
    lay.o_n = lay.o_slot + 4u * nb;
    const size_t bytes = lay.o_n + 4u * (size_t)nb;
    if (bytes > ln.desc_bytes) return SYNDROME_INVALID_PARAM;
    char *h = (char *)ln.h_desc;
    memcpy(h + lay.o_items, items.data(), sizeof(SynItem) * ni);
    memcpy(h + lay.o_ids, ids.data(), 4 * (size_t)nb);
    int n_max = 0;
    for (int j = 0; j < nb; ++j) {
        const Slot &s = st[batch[j]];
        ((long long *)(h + lay.o_offs))[j] = (long long)batch[j] * NMAX;
        ((const float **)(h + lay.o_src))[j] = s.src;
        ((int *)(h + lay.o_nf))[j] = s.nfill;
        ((int *)(h + lay.o_slot))[j] = batch[j];
        ((int *)(h + lay.o_n))[j] = s.n;
        n_max = std::max(n_max, s.n);
    }
    char *d = (char *)ln.d_desc;
    const dim3 sgrid((n_max + STAGE_CHUNK - 1) / STAGE_CHUNK, nb);
    hipLaunchKernelGGL(stage, sgrid, dim3(STAGE_THREADS), 0, ln.stream, (const char *)ln.hd_desc, d, lay, d_llr,
                       ln.d_head);
    if (hipGetLastError() != hipSuccess) return SYNDROME_DEVICE_ERROR;
    SynLaunch a{};
    a.in = d_llr;
    a.post = nullptr;
    a.hard = h_hard_dev;  // same slot layout as d_llr: results land in pinned memory
    a.iters = ln.hd_iters;
    a.batch = nb;
    a.iterations = iterations;
    a.flags = SYN_F_INIT_5G | SYN_F_WRITE_HARD | SYN_F_ALIGNED;  // slot offsets are multiples of NMAX
    a.beta = beta;
    const void *fn = syn::persist_kernel_f32_big();
    if (syn::allow_max_dynamic_lds(fn)) return SYNDROME_DEVICE_ERROR;
    const SynCode *cp = d_codes;
    const SynItem *ip = (const SynItem *)(d + lay.o_items);
    int n_items = ni;
    int *head = ln.d_head;
    const int *idp = (const int *)(d + lay.o_ids);
    const long long *op = (const long long *)(d + lay.o_offs);
    const int *nfp = (const int *)(d + lay.o_nf);
    void *args[] = {&cp, &ip, &n_items, &head, &idp, &op, &nfp, &a};
    const int grid = std::max(1, std::min(ni, cus));
    if (hipLaunchKernel(fn, dim3(grid), dim3(512), args, lds, ln.stream) != hipSuccess) return SYNDROME_DEVICE_ERROR;
    if (hipEventRecord(ln.done, ln.stream) != hipSuccess) return SYNDROME_DEVICE_ERROR;
    return SYNDROME_OK;
}


void SyndromeServer::run()
{
    DeviceScope ds(device);
    std::vector<int> batch;
    while (!stop.load()) {
        bool progressed = false;
        for (Lane &ln : lanes) {
            if (ln.busy && hipEventQuery(ln.done) == hipSuccess) {
                const double t = now_ns();
                std::lock_guard<syn::ServerMutex> lk(mu);
                for (size_t k = 0; k < ln.slots.size(); ++k) {
                    const int s = ln.slots[k];
                    h_it[s] = ln.h_iters[k];
                    st[s].state = 4;
                    st[s].t_done = t;
                    done.push_back(s);
                }
                decoded += (long long)ln.slots.size();
                sync_counts();
                ln.slots.clear();
                ln.busy = false;
                progressed = true;
            }
        }
        for (Lane &ln : lanes) {
            if (ln.busy) continue;
            batch.clear();
            if (n_submitted.load(std::memory_order_acquire) > 0) {
                std::lock_guard<syn::ServerMutex> lk(mu);
                while (!submitted.empty() && (int)batch.size() < max_batch) {
                    const int s = submitted.front();
                    submitted.pop_front();
                    st[s].state = 3;
                    batch.push_back(s);
                }
                sync_counts();
            }
            if (batch.empty()) break;
            if (int32_t rc = launch_batch(ln, batch)) {
                std::lock_guard<syn::ServerMutex> lk(mu);
                error.store(rc);
                for (int s : batch) {  // report the failure per code block
                    st[s].state = 4;
                    st[s].t_done = now_ns();
                    h_it[s] = -1;
                    done.push_back(s);
                }
                sync_counts();
                continue;
            }
            ln.slots = batch;
            ln.busy = true;
            batches++;
            progressed = true;
        }
        if (!progressed) std::this_thread::yield();
    }
    for (Lane &ln : lanes)
        if (ln.busy) (void)hipEventSynchronize(ln.done);
}

int32_t SyndromeServer::init_resident()
{
    const void *fn = syn::resident_kernel();
    hipFuncAttributes fa{};
    int khz = 0, per_cu = 0;
    if (syn::allow_max_dynamic_lds(fn) || hipFuncGetAttributes(&fa, fn) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0)
        return SYNDROME_DEVICE_ERROR;
    const size_t room = (size_t)lds_limit - fa.sharedSizeBytes;
    rs.cws.resize(codes.size());
    rd.lds = 0;
    for (size_t c = 0; c < codes.size(); ++c) {
        if (codes[c].kind == SYN_KIND_BG1 && codes[c].z == 384 && (size_t)syn::resident_pair_lds() <= room) {
            // two blocks per item, the two-codeword kernel (ldpc_resident.hip)
            rs.cws[c] = 2;
            rd.lds = std::max(rd.lds, (size_t)syn::resident_pair_lds());
            rd.xpair = true;
            continue;
        }
        int k = std::min(SYN_WORK_MAX, std::max(1, 512 / codes[c].z));
        while (k > 1 && group_bytes(codes[c], k) > room) --k;
        if (group_bytes(codes[c], k) > room) return SYNDROME_INVALID_PARAM;
        rs.cws[c] = k;
        rd.lds = std::max(rd.lds, group_bytes(codes[c], k));
    }
    if (rd.lds > room) return SYNDROME_INVALID_PARAM;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, syn::resident_lanes(), rd.lds) != hipSuccess ||
        per_cu < 1)
        return SYNDROME_DEVICE_ERROR;
    rs.grid = per_cu * cus;
    if (hipMalloc((void **)&rd.d_xscratch, sizeof(float) * syn::resident_scratch_floats() * (size_t)rs.grid) !=
        hipSuccess)
        return SYNDROME_DEVICE_ERROR;
    rd.idle_ticks = (long long)(RESIDENT_IDLE_US * khz / 1000.0);
    rd.life_ticks = (long long)(RESIDENT_LIFE_S * khz * 1000.0);
    rs.cap = RING_ITEMS;
    std::vector<long long> offs(slots);
    for (int i = 0; i < slots; ++i) offs[i] = (long long)i * NMAX;
    if (hipHostMalloc((void **)&rs.ring, sizeof(SynWork) * rs.cap, HOST_FLAGS) != hipSuccess ||
        hipHostMalloc((void **)&rs.ctl, sizeof(SynResidentCtl), HOST_FLAGS) != hipSuccess ||
        hipHostMalloc((void **)&rs.done, 4 * (size_t)rs.cap, HOST_FLAGS) != hipSuccess ||
        hipHostGetDevicePointer((void **)&rd.ring_dev, rs.ring, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&rd.ctl_dev, rs.ctl, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&rd.done_dev, rs.done, 0) != hipSuccess ||
        hipMalloc((void **)&rd.d_head, SYN_RESIDENT_QUEUE_BYTES) != hipSuccess ||
        hipMalloc((void **)&rd.d_last, 64) != hipSuccess ||
        hipMalloc((void **)&rd.d_offs, 8 * (size_t)slots) != hipSuccess ||
        hipMemcpy(rd.d_offs, offs.data(), 8 * (size_t)slots, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(rd.d_head, 0, SYN_RESIDENT_QUEUE_BYTES) != hipSuccess || hipMemset(rd.d_last, 0, 64) != hipSuccess ||
        hipStreamCreateWithFlags(&be.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&be.ev, hipEventDisableTiming) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return SYNDROME_DEVICE_ERROR;
    memset(rs.ring, 0, sizeof(SynWork) * rs.cap);
    memset(rs.ctl, 0, sizeof(SynResidentCtl));
    memset(rs.done, 0, 4 * (size_t)rs.cap);
    rs.tb_cap = TB_RING;
    if (hipHostMalloc((void **)&rs.tbw, sizeof(SynTbWork) * rs.tb_cap, HOST_FLAGS) != hipSuccess ||
        hipHostMalloc((void **)&rs.tb_done, 4 * (size_t)rs.tb_cap, HOST_FLAGS) != hipSuccess ||
        hipHostMalloc((void **)&rs.tb_ok, 4 * (size_t)rs.tb_cap, HOST_FLAGS) != hipSuccess ||
        hipHostGetDevicePointer((void **)&rd.tbw_dev, rs.tbw, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&rd.tb_done_dev, rs.tb_done, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&rd.tb_ok_dev, rs.tb_ok, 0) != hipSuccess ||
        hipMalloc((void **)&rd.d_tbstate, sizeof(SynTbState) * rs.tb_cap) != hipSuccess ||
        hipMemset(rd.d_tbstate, 0, sizeof(SynTbState) * rs.tb_cap) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return SYNDROME_DEVICE_ERROR;
    memset(rs.tbw, 0, sizeof(SynTbWork) * rs.tb_cap);
    memset(rs.tb_done, 0, 4 * (size_t)rs.tb_cap);
    memset(rs.tb_ok, 0, 4 * (size_t)rs.tb_cap);
    init_tb_ring(rs.tb_cap);
    return SYNDROME_OK;
}

extern "C" {

int32_t syndrome_server_create(float offset_beta, size_t iterations, size_t hbm_budget_bytes, int32_t max_batch,
                               int32_t streams, SyndromeServer **out)
{
    return guard([&]() -> int32_t {
        if (!out) return SYNDROME_NULL_POINTER;
        *out = nullptr;
        const size_t slot_bytes = (size_t)NMAX * 4 + 64;
        if (hbm_budget_bytes < slot_bytes || streams < 0 || streams > 8 || max_batch < 1 || max_batch > 65535)
            return SYNDROME_INVALID_PARAM;
        std::unique_ptr<SyndromeServer> s(new SyndromeServer());
        s->beta = offset_beta;
        s->iterations = (int)std::min(iterations, (size_t)INT_MAX);
        s->max_batch = max_batch;
        s->slots = (int)std::min(hbm_budget_bytes / slot_bytes, (size_t)1 << 20);
        if (hipGetDevice(&s->device) != hipSuccess) return SYNDROME_DEVICE_ERROR;
        if (hipDeviceGetAttribute(&s->lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device) != hipSuccess ||
            hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess)
            return SYNDROME_DEVICE_ERROR;
        for (int b = 0; b < 2; ++b)
            for (int zi = 0; zi < 51; ++zi) {
                SynCode c;
                if (syn::build_code_3gpp(b, VALID_Z[zi], &c)) return SYNDROME_INVALID_PARAM;
                s->codes.push_back(c);
            }
        const size_t n = (size_t)s->slots * NMAX;
        if (hipMalloc((void **)&s->d_codes, sizeof(SynCode) * s->codes.size()) != hipSuccess ||
            hipMemcpy(s->d_codes, s->codes.data(), sizeof(SynCode) * s->codes.size(), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMalloc((void **)&s->d_llr, n * 4) != hipSuccess ||
            hipHostMalloc((void **)&s->h_llr, n * 4, HOST_FLAGS) != hipSuccess ||
            hipHostMalloc((void **)&s->h_hard, n, HOST_FLAGS) != hipSuccess ||
            hipHostMalloc((void **)&s->h_it, 4 * (size_t)s->slots, HOST_FLAGS) != hipSuccess ||
            hipHostMalloc((void **)&s->h_cbok, 4 * (size_t)s->slots, HOST_FLAGS) != hipSuccess ||
            hipHostGetDevicePointer((void **)&s->h_cbok_dev, s->h_cbok, 0) != hipSuccess ||
            hipHostGetDevicePointer((void **)&s->h_llr_dev, s->h_llr, 0) != hipSuccess ||
            hipHostGetDevicePointer((void **)&s->h_hard_dev, s->h_hard, 0) != hipSuccess ||
            hipHostGetDevicePointer((void **)&s->h_it_dev, s->h_it, 0) != hipSuccess) {
            syndrome_server_destroy(s.release());
            return SYNDROME_DEVICE_ERROR;
        }
        memset(s->h_it, 0, 4 * (size_t)s->slots);
        memset(s->h_cbok, 0, 4 * (size_t)s->slots);
        s->init_slots(s->slots);
        if (streams == 0) {
            s->resident = true;
            if (int32_t rc = s->init_resident()) {
                syndrome_server_destroy(s.release());
                return rc;
            }
            SyndromeServer *raw = s.get();
            s->worker = std::thread([raw] {
                DeviceScope ds(raw->device);
                raw->run_resident();
            });
            *out = s.release();
            return SYNDROME_OK;
        }
        s->lanes.resize(streams);
        for (Lane &ln : s->lanes) {
            ln.desc_bytes = (sizeof(SynItem) + 4 + 8 + 8 + 4 + 4 + 4) * (size_t)max_batch + 64;
            if (hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&ln.done, hipEventDisableTiming) != hipSuccess ||
                hipHostMalloc(&ln.h_desc, ln.desc_bytes, HOST_FLAGS) != hipSuccess ||
                hipHostGetDevicePointer(&ln.hd_desc, ln.h_desc, 0) != hipSuccess ||
                hipMalloc(&ln.d_desc, ln.desc_bytes) != hipSuccess || hipMalloc((void **)&ln.d_head, 64) != hipSuccess ||
                hipHostMalloc((void **)&ln.h_iters, 4 * (size_t)max_batch, HOST_FLAGS) != hipSuccess ||
                hipHostGetDevicePointer((void **)&ln.hd_iters, ln.h_iters, 0) != hipSuccess) {
                syndrome_server_destroy(s.release());
                return SYNDROME_DEVICE_ERROR;
            }
        }
        SyndromeServer *raw = s.get();
        s->worker = std::thread([raw] { raw->run(); });
        *out = s.release();
        return SYNDROME_OK;
    });
}

void syndrome_server_destroy(SyndromeServer *s)
{
    if (!s) return;
    try {
        s->stop = true;
        if (s->worker.joinable()) s->worker.join();
        DeviceScope ds(s->device);
        for (Lane &ln : s->lanes) {
            if (ln.stream) (void)hipStreamSynchronize(ln.stream);
            if (ln.done) (void)hipEventDestroy(ln.done);
            if (ln.stream) (void)hipStreamDestroy(ln.stream);
            for (void *p : {ln.h_desc, (void *)ln.h_iters})
                if (p) (void)hipHostFree(p);
            for (void *p : {ln.d_desc, (void *)ln.d_head})
                if (p) (void)hipFree(p);
        }
        auto &r = s->rs;
        auto &d = s->rd;
        if (r.ctl) __atomic_store_n(&r.ctl->stop, 1u, __ATOMIC_RELEASE);
        if (s->be.stream) (void)hipStreamSynchronize(s->be.stream);
        if (s->be.ev) (void)hipEventDestroy(s->be.ev);
        if (s->be.stream) (void)hipStreamDestroy(s->be.stream);
        for (void *p : {(void *)r.ring, (void *)r.ctl, (void *)r.done, (void *)r.tbw, (void *)r.tb_done, (void *)r.tb_ok})
            if (p) (void)hipHostFree(p);
        for (void *p : {(void *)d.d_head, (void *)d.d_last, (void *)d.d_offs, (void *)d.d_tbstate, (void *)d.d_xscratch})
            if (p) (void)hipFree(p);
        for (void *p : {(void *)s->d_llr, (void *)s->d_codes})
            if (p) (void)hipFree(p);
        for (void *p : {(void *)s->h_llr, (void *)s->h_hard, (void *)s->h_it, (void *)s->h_cbok})
            if (p) (void)hipHostFree(p);
    } catch (...) {
    }
    delete s;
}

int32_t syndrome_server_acquire(SyndromeServer *s, int32_t base_graph, size_t z, size_t n_filler, float **out_llr,
                                uint32_t *out_slot)
{
    return guard([&]() -> int32_t {
        if (!s || !out_llr || !out_slot) return SYNDROME_NULL_POINTER;
        return s->acquire(base_graph, z, n_filler, out_llr, out_slot);
    });
}

int32_t syndrome_server_submit(SyndromeServer *s, uint32_t slot, uint64_t tag)
{
    return guard([&]() -> int32_t {
        if (!s) return SYNDROME_NULL_POINTER;
        return s->submit(slot, nullptr, tag);
    });
}

int32_t syndrome_server_submit_from(SyndromeServer *s, uint32_t slot, const float *src, uint64_t tag)
{
    return guard([&]() -> int32_t {
        if (!s || !src) return SYNDROME_NULL_POINTER;
        return s->submit(slot, src, tag);
    });
}

int32_t syndrome_server_try_recv(SyndromeServer *s, SyndromeServed *out)
{
    return guard([&]() -> int32_t {
        if (!s || !out) return SYNDROME_NULL_POINTER;
        return s->try_recv(out);
    });
}

int32_t syndrome_server_release(SyndromeServer *s, uint32_t slot)
{
    return guard([&]() -> int32_t {
        if (!s) return SYNDROME_NULL_POINTER;
        return s->release(slot);
    });
}

int32_t syndrome_server_acquire_many(SyndromeServer *s, const int32_t *base_graph, const int32_t *z,
                                     const int32_t *n_filler, size_t n, float **out_llr, uint32_t *out_slots,
                                     size_t *out_acquired)
{
    return guard([&]() -> int32_t {
        if (!s || !out_acquired || (n && (!base_graph || !z || !out_slots))) return SYNDROME_NULL_POINTER;
        return s->acquire_many(base_graph, z, n_filler, n, out_llr, out_slots, out_acquired);
    });
}

int32_t syndrome_server_submit_many(SyndromeServer *s, const uint32_t *slots, const float *const *srcs,
                                    const uint64_t *tags, size_t n)
{
    return guard([&]() -> int32_t {
        if (!s || (n && !slots)) return SYNDROME_NULL_POINTER;
        return s->submit_many(slots, srcs, tags, n);
    });
}

int32_t syndrome_server_try_recv_many(SyndromeServer *s, SyndromeServed *out, size_t cap, size_t *out_n)
{
    return guard([&]() -> int32_t {
        if (!s || !out_n || (cap && !out)) return SYNDROME_NULL_POINTER;
        return s->try_recv_many(out, cap, out_n);
    });
}

int32_t syndrome_server_release_many(SyndromeServer *s, const uint32_t *slots, size_t n)
{
    return guard([&]() -> int32_t {
        if (!s || (n && !slots)) return SYNDROME_NULL_POINTER;
        return s->release_many(slots, n);
    });
}

int32_t syndrome_server_submit_tb(SyndromeServer *s, const uint32_t *slots, const float *const *srcs, size_t num_cb,
                                  size_t tb_size, uint8_t *tb_out, uint64_t tag)
{
    return guard([&]() -> int32_t {
        if (!s || !slots || !tb_out) return SYNDROME_NULL_POINTER;
        return s->submit_tb(slots, srcs, num_cb, tb_size, tb_out, tag);
    });
}

int32_t syndrome_server_try_recv_tb(SyndromeServer *s, SyndromeServedTb *out)
{
    return guard([&]() -> int32_t {
        if (!s || !out) return SYNDROME_NULL_POINTER;
        return s->try_recv_tb(out);
    });
}

int32_t syndrome_server_tb_block_crc(SyndromeServer *s, uint32_t tb, uint8_t *cb_ok, size_t cap, size_t *out_n)
{
    return guard([&]() -> int32_t {
        if (!s || !out_n || (cap && !cb_ok)) return SYNDROME_NULL_POINTER;
        return s->tb_block_crc(tb, cb_ok, cap, out_n);
    });
}

int32_t syndrome_server_release_tb(SyndromeServer *s, uint32_t tb)
{
    return guard([&]() -> int32_t {
        if (!s) return SYNDROME_NULL_POINTER;
        return s->release_tb(tb);
    });
}

int32_t syndrome_server_stats(const SyndromeServer *s, int64_t *out_batches, int64_t *out_decoded, int32_t *out_slots)
{
    return guard([&]() -> int32_t {
        if (!s) return SYNDROME_NULL_POINTER;
        if (out_batches) *out_batches = s->batches.load();
        if (out_decoded) *out_decoded = s->decoded.load();
        if (out_slots) *out_slots = s->slots;
        return s->error.load();
    });
}

void *syndrome_host_alloc(size_t bytes)
{
    void *p = nullptr;
    try {
        if (hipHostMalloc(&p, bytes ? bytes : 1, HOST_FLAGS) != hipSuccess) return nullptr;
    } catch (...) {
        return nullptr;
    }
    return p;
}

void syndrome_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

int32_t syndrome_server_loadgen(SyndromeServer *s, const float *pool, const int64_t *cb_off, const int32_t *cb_bg,
                                const int32_t *cb_z, const int32_t *cb_nfill, const int32_t *cb_info_bits,
                                const int32_t *tb_first_cb, const int32_t *tb_num_cb, size_t n_tb, double tb_rate,
                                double duration_s, uint64_t seed, double *cb_lat_ns, size_t cb_cap, size_t *n_cb,
                                double *tb_lat_ns, size_t tb_cap, size_t *n_tbs, double *info_bps, int64_t *waits)
{
    return guard([&]() -> int32_t {
        if (!s || !pool || !cb_off || !cb_bg || !cb_z || !tb_first_cb || !tb_num_cb || !n_tb) return SYNDROME_NULL_POINTER;
        struct Tb {
            double arrival;
            int pool_tb, left;
            double last_done;
        };
        struct Cb {
            int tb_slot, pool_cb;
            double arrival;
        };
        std::vector<Tb> tbs;
        std::deque<Cb> pending;
        std::vector<Cb> inflight(s->slots);
        constexpr int BATCH = 256;
        std::vector<int32_t> a_bg(BATCH), a_z(BATCH), a_nf(BATCH);
        std::vector<uint32_t> a_slot(BATCH), a_rel(BATCH);
        std::vector<const float *> a_src(BATCH);
        std::vector<uint64_t> a_tag(BATCH);
        std::vector<SyndromeServed> served(BATCH);
        uint64_t rs = seed | 1;
        auto urand = [&]() {
            rs ^= rs << 13;
            rs ^= rs >> 7;
            rs ^= rs << 17;
            return ((rs >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        const double t0 = now_ns(), t_end = t0 + duration_s * 1e9;
        double next = t0;
        size_t ncb = 0, ntb = 0, next_tb = 0;
        long long info = 0, wait_count = 0;
        double last_done = t0;
        long long outstanding = 0;
        for (;;) {
            const double now = now_ns();
            const bool open = now < t_end;
            if (open) {
                if (tb_rate > 0) {
                    while (next <= now && next < t_end) {
                        const int p = (int)(next_tb++ % n_tb);
                        tbs.push_back(Tb{next, p, tb_num_cb[p], 0});
                        for (int c = 0; c < tb_num_cb[p]; ++c) pending.push_back(Cb{(int)tbs.size() - 1, tb_first_cb[p] + c, next});
                        next += -log(urand()) / tb_rate * 1e9;
                    }
                } else {
                    while (pending.size() < 4 * (size_t)s->max_batch) {  // saturating stream
                        const int p = (int)(next_tb++ % n_tb);
                        tbs.push_back(Tb{now, p, tb_num_cb[p], 0});
                        for (int c = 0; c < tb_num_cb[p]; ++c) pending.push_back(Cb{(int)tbs.size() - 1, tb_first_cb[p] + c, now});
                    }
                }
            }
            // what is pending goes in in batches: one server lock per batch
            while (!pending.empty() && (open || tb_rate > 0)) {
                const size_t m = std::min(pending.size(), (size_t)BATCH);
                for (size_t i = 0; i < m; ++i) {
                    const int pc = pending[i].pool_cb;
                    a_bg[i] = cb_bg[pc];
                    a_z[i] = cb_z[pc];
                    a_nf[i] = cb_nfill ? cb_nfill[pc] : 0;
                }
                size_t k = 0;
                const int32_t rc =
                    syndrome_server_acquire_many(s, a_bg.data(), a_z.data(), a_nf.data(), m, nullptr, a_slot.data(), &k);
                if (rc && rc != SYNDROME_BUSY) return rc;
                for (size_t i = 0; i < k; ++i) {
                    inflight[a_slot[i]] = pending[i];
                    a_src[i] = pool + cb_off[pending[i].pool_cb];
                    a_tag[i] = a_slot[i];
                }
                if (k) {
                    if (int32_t r2 = syndrome_server_submit_many(s, a_slot.data(), a_src.data(), a_tag.data(), k))
                        return r2;
                }
                pending.erase(pending.begin(), pending.begin() + (long)k);
                outstanding += (long long)k;
                if (k < m) {
                    ++wait_count;
                    break;
                }
            }
            if (!open && tb_rate <= 0) pending.clear();
            for (;;) {
                size_t k = 0;
                const int32_t rc = syndrome_server_try_recv_many(s, served.data(), served.size(), &k);
                if (rc == SYNDROME_BUSY) break;
                if (rc) return rc;
                for (size_t i = 0; i < k; ++i) {
                    const SyndromeServed &r = served[i];
                    const Cb &cb = inflight[r.slot];
                    if (ncb < cb_cap && cb_lat_ns) cb_lat_ns[ncb] = r.done_ns - cb.arrival;
                    ++ncb;
                    if (cb_info_bits) info += cb_info_bits[cb.pool_cb];
                    Tb &tb = tbs[cb.tb_slot];
                    tb.last_done = std::max(tb.last_done, r.done_ns);
                    last_done = std::max(last_done, r.done_ns);
                    if (--tb.left == 0) {
                        if (ntb < tb_cap && tb_lat_ns) tb_lat_ns[ntb] = tb.last_done - tb.arrival;
                        ++ntb;
                    }
                    a_rel[i] = r.slot;
                }
                if (int32_t r2 = syndrome_server_release_many(s, a_rel.data(), k)) return r2;
                outstanding -= (long long)k;
            }
            if (!open && pending.empty() && outstanding == 0) break;
            if (now > t_end + 30e9) return SYNDROME_DEVICE_ERROR;  // drained for 30 s without finishing
        }
        if (n_cb) *n_cb = ncb;
        if (n_tbs) *n_tbs = ntb;
        if (info_bps) *info_bps = (double)info / ((last_done - t0) * 1e-9);
        if (waits) *waits = wait_count;
        return SYNDROME_OK;
    });
}

// Transport-block load generator over the resident server: as
// syndrome_server_loadgen, but each arrival is one transport block (its
// blocks acquired together, submitted with syndrome_server_submit_tb with
// inputs read from the pool, payloads written to tb_out + tb_out_off[p]),
// and latency is transport-block completion (CRCs decided on the device)
// minus scheduled arrival.
int32_t syndrome_server_loadgen_tb(SyndromeServer *s, const float *pool, const int64_t *cb_off,
                                   const int32_t *cb_bg, const int32_t *cb_z, const int32_t *cb_nfill,
                                   const int32_t *tb_first_cb, const int32_t *tb_num_cb, const int32_t *tb_size,
                                   uint8_t *tb_out, const int64_t *tb_out_off, size_t n_tb, double tb_rate,
                                   double duration_s, uint64_t seed, double *tb_lat_ns, uint8_t *tb_lat_ok,
                                   size_t lat_cap, size_t *n_tbs, size_t *n_cbs, double *info_bps,
                                   double *goodput_bps, int64_t *crc_fails, int64_t *waits)
{
    return guard([&]() -> int32_t {
        if (!s || !pool || !cb_off || !cb_bg || !cb_z || !tb_first_cb || !tb_num_cb || !tb_size || !tb_out ||
            !tb_out_off || !n_tb)
            return SYNDROME_NULL_POINTER;
        if (!s->resident) return SYNDROME_INVALID_PARAM;
        struct Pending {
            int pool_tb;
            double arrival;
            std::vector<uint32_t> slots;
        };
        std::deque<Pending> pending;
        std::vector<std::pair<double, int>> meta;  // per submission (its tag): arrival, pool transport block
        uint64_t rs = seed | 1;
        auto urand = [&]() {
            rs ^= rs << 13;
            rs ^= rs >> 7;
            rs ^= rs << 17;
            return ((rs >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        const double t0 = now_ns(), t_end = t0 + duration_s * 1e9;
        double next = t0, last_done = t0;
        size_t ntb = 0, ncb = 0, next_tb = 0;
        long long info = 0, good = 0, fails = 0, wait_count = 0, outstanding = 0;
        std::vector<const float *> srcs;
        for (;;) {
            const double now = now_ns();
            const bool open = now < t_end;
            if (open) {
                if (tb_rate > 0) {
                    while (next <= now && next < t_end) {
                        pending.push_back(Pending{(int)(next_tb++ % n_tb), next, {}});
                        next += -log(urand()) / tb_rate * 1e9;
                    }
                } else {
                    while (pending.size() < 64) pending.push_back(Pending{(int)(next_tb++ % n_tb), now, {}});
                }
            }
            while (!pending.empty() && (open || tb_rate > 0)) {
                Pending &pd = pending.front();
                const int p = pd.pool_tb, c = tb_num_cb[p], f = tb_first_cb[p];
                bool busy = false;
                if ((int)pd.slots.size() < c) {  // a transport block holds its slots as it gets them
                    const int have = (int)pd.slots.size(), want = c - have;
                    pd.slots.resize(c);
                    size_t got = 0;
                    const int32_t rc = syndrome_server_acquire_many(s, cb_bg + f + have, cb_z + f + have,
                                                                    cb_nfill ? cb_nfill + f + have : nullptr,
                                                                    (size_t)want, nullptr, pd.slots.data() + have, &got);
                    pd.slots.resize(have + got);
                    if (rc && rc != SYNDROME_BUSY) return rc;
                    busy = (int)got < want;
                }
                if (!busy) {
                    srcs.resize(c);
                    for (int i = 0; i < c; ++i) srcs[i] = pool + cb_off[f + i];
                    const int32_t rc = syndrome_server_submit_tb(s, pd.slots.data(), srcs.data(), (size_t)c,
                                                                 (size_t)tb_size[p], tb_out + tb_out_off[p],
                                                                 (uint64_t)meta.size());
                    if (rc == SYNDROME_BUSY) busy = true;
                    else if (rc) return rc;
                }
                if (busy) {
                    ++wait_count;
                    break;
                }
                meta.push_back({pd.arrival, p});
                pending.pop_front();
                ++outstanding;
            }
            if (!open && tb_rate <= 0) {
                for (Pending &pd : pending)
                    for (uint32_t sl : pd.slots) (void)syndrome_server_release(s, sl);
                pending.clear();
            }
            SyndromeServedTb r;
            for (;;) {
                const int32_t rc = syndrome_server_try_recv_tb(s, &r);
                if (rc == SYNDROME_BUSY) break;
                if (rc != SYNDROME_OK) return rc;  // a transport block the server could not decode
                const auto &m = meta[r.tag];
                if (ntb < lat_cap && tb_lat_ns) tb_lat_ns[ntb] = r.done_ns - m.first;
                if (ntb < lat_cap && tb_lat_ok) tb_lat_ok[ntb] = r.crc_ok ? 1 : 0;
                ++ntb;
                ncb += r.num_cb;
                info += tb_size[m.second];
                if (r.crc_ok) good += tb_size[m.second];
                if (!r.crc_ok) ++fails;
                last_done = std::max(last_done, r.done_ns);
                --outstanding;
                (void)syndrome_server_release_tb(s, r.tb);
            }
            if (!open && pending.empty() && outstanding == 0) break;
            if (now > t_end + 30e9) return SYNDROME_DEVICE_ERROR;  // drained for 30 s without finishing
        }
        if (n_tbs) *n_tbs = ntb;
        if (n_cbs) *n_cbs = ncb;
        if (info_bps) *info_bps = (double)info / ((last_done - t0) * 1e-9);
        if (goodput_bps) *goodput_bps = (double)good / ((last_done - t0) * 1e-9);
        if (crc_fails) *crc_fails = fails;
        if (waits) *waits = wait_count;
        return SYNDROME_OK;
    });
}

}  // extern "C"
