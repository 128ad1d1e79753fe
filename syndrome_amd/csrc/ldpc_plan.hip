// Work-queue batch plans: many code blocks of different 3GPP base graphs,
// lifting sizes and filler counts decoded in one persistent launch, plus the
// device rate de-matcher that feeds them and the launch-shape tuner.
#include <hip/hip_runtime.h>

#include <limits.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/syndrome_ldpc.h"
#include "code_host.h"
#include "ldpc_common.h"

namespace syn {
const void *persist_kernel_f32_big();
const void *persist_kernel_f32_small();
const void *persist_kernel_f32_small_bg2();
const void *persist_kernel_i8();
const void *persistx_kernel_f32_bg1(int z);
int persistx_lanes_f32_bg1();
int fastx_shape_f32_bg1(int z, int *lanes, int *cws);
int launch_reset(int *head, hipStream_t stream);
int allow_max_dynamic_lds(const void *fn);
void dematch_derive(SynDematchCb &cb);
void dematch_chunks(const SynDematchCb *cbs, int count, std::vector<SynDematchChunk> &out);
int launch_dematch(const float *e_llr, float *out, const SynDematchCb *cbs, const SynDematchChunk *chunks,
                   int n_chunks, bool combine, int rv, hipStream_t stream);
int launch_tb_check(const uint8_t *hard, const long long *offs, const SynTb *tbs, int n_tb, uint8_t *tb_out,
                    uint8_t *tb_ok, uint8_t *cb_ok, hipStream_t stream);
}  // namespace syn

struct SyndromeLdpcPlan {
    int device = 0;
    bool i8 = false;
    float beta = 0.f;
    int lds_limit = 0;
    int cus = 0;
    // per code block
    std::vector<int> bg, z, nfill;
    std::vector<long long> offs;
    bool offs_aligned = true;  // every offset a multiple of 4 elements
    // unique codes
    std::vector<SynCode> codes;
    std::vector<int> cb_code;
    // current launch shape; a plan whose largest Z exceeds SPLIT_LANES
    // decodes its codes of Z <= SPLIT_LANES in a second launch (the *_s
    // fields) of SPLIT_LANES-lane workgroups
    int block = 0, grid = 0, n_items = 0;
    size_t lds = 0;
    int block_s = 0, grid_s = 0, n_items_s = 0;
    size_t lds_s = 0;
    // the kernel of each launch: f32 codes with Z > 256 in a launch take the
    // kernel whose compile-time-Z bodies are the large sizes
    const void *fn_main = nullptr, *fn_small = nullptr;
    // f32 BG2 codes of Z <= 256: a launch of the kernel with their
    // compile-time-Z bodies (the Z <= 256 launch, or the main launch when
    // the plan has no split, keeps the BG1 codes)
    int block_2 = 0, grid_2 = 0, n_items_2 = 0;
    size_t lds_2 = 0;
    const void *fn_2 = nullptr;
    // f32 BG1 codes of Z > 256: a third launch of the two-codeword kernel that
    // keeps the extension columns in the posterior buffer (ldpc_fastx.h),
    // for runs that write the posterior back; other runs decode them in the
    // main launch, one codeword per workgroup: the *_mx list is the main
    // list with them (the same as the main list when the plan has none)
    int block_x = 0, grid_x = 0, n_items_x = 0;
    size_t lds_x = 0;
    // whether the extension launch pays when it runs after the others (a
    // run under stream capture); a forked run always uses it
    bool x_alone = false;
    int grid_mx = 0, n_items_mx = 0;
    size_t lds_mx = 0;
    const void *fn_x = nullptr, *fn_mx = nullptr;
    std::vector<int> cws_of_code;
    // device buffers
    SynCode *d_codes = nullptr;
    SynItem *d_items = nullptr, *d_items_s = nullptr, *d_items_x = nullptr, *d_items_mx = nullptr, *d_items_2 = nullptr;
    int *d_ids = nullptr, *d_nfill = nullptr, *d_head = nullptr;
    long long *d_offs = nullptr;
    SynDematchCb *d_dm = nullptr;
    SynDematchChunk *d_dm_chunks = nullptr;
    int n_dm_chunks = 0;
    SynTb *d_tbs = nullptr;
    int n_tbs = 0;
    // queue counters: RUN_RING runs, four launches each, one 128-byte line
    // per counter; run k uses ring entry k % RUN_RING
    unsigned run_seq = 0;
    // device tables replaced by tune / set_dematch / set_tbs: kept until the
    // plan is destroyed, so a HIP graph captured before the change still
    // replays valid pointers (with the launch shape it was captured with)
    std::vector<void *> retired;
    // Stream and events of an eager run's fork: some launches go on a side
    // stream that waits for the caller's stream and that it waits for in
    // turn, so their workgroups fill CUs that another launch's drained
    // queue leaves idle.
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::mutex mu;
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

template <class T>
int32_t upload(SyndromeLdpcPlan *p, T **dst, const std::vector<T> &src)
{
    if (*dst) p->retired.push_back((void *)*dst);
    *dst = nullptr;
    if (src.empty()) return SYNDROME_OK;
    if (hipMalloc((void **)dst, sizeof(T) * src.size()) != hipSuccess) return SYNDROME_DEVICE_ERROR;
    if (hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice) != hipSuccess)
        return SYNDROME_DEVICE_ERROR;
    return SYNDROME_OK;
}

// Lanes of the small codes' workgroups in a split plan, the spacing (ints,
// one 128-byte line) of the queue counters, and how many runs of one plan
// may be in flight at once, each with its own pair of counters.
constexpr int SPLIT_LANES = 256;
constexpr int HEAD_STRIDE = 32;
constexpr unsigned RUN_RING = 16;
constexpr int LAUNCHES = 4;
constexpr int X_MIN_ITEMS_PER_CU = 4;

// A code of the two-codeword extension-in-HBM kernel's launch.
bool x_code(const SynCode &c, bool i8) { return !i8 && c.kind == SYN_KIND_BG1 && syn::persistx_kernel_f32_bg1(c.z); }

// An int8 BG2 item normally runs pair-packed (ldpc_pack.h): its codewords
// go two to a lane, in 32-bit LDS words that interleave a pair.
bool can_pack(const SynCode &c, bool i8) { return i8 && c.kind == SYN_KIND_BG2; }

size_t group_bytes(const SynCode &c, int cws, bool i8, bool packed = false)
{
    const size_t n = syn::code_n(c);
    if (packed) {
        const size_t pairs = (size_t)(cws + 1) / 2;
        return (size_t)(((c.z * 4) + 15) & ~15) + pairs * n * 4 + 3 * sizeof(int) * 2 * pairs;
    }
    const size_t esz = i8 ? 2 : 4;
    const size_t pad = (size_t)(((c.z * (int)esz) + 15) & ~15);
    return pad + ((cws * n * esz + 3) & ~(size_t)3) + 3 * sizeof(int) * (size_t)cws;
}

// Items for workgroups of `block` lanes: each code's blocks in groups of as
// many codewords as fit the lanes and the LDS, longest work first.  When the
// plan's largest Z exceeds SPLIT_LANES, codes of Z <= SPLIT_LANES form a
// second launch of SPLIT_LANES-lane workgroups with their own (smaller) LDS:
// more of those share a CU (the int8 kernel fits 12 waves per CU, one
// 512-lane workgroup only 8), and their waves spread evenly over the SIMDs.
int32_t build_items(SyndromeLdpcPlan *p, int block)
{
    if (block <= 0 || block > 512 || block % 64) return SYNDROME_INVALID_PARAM;
    const size_t nc = p->codes.size();
    int max_z = 0;
    for (const SynCode &code : p->codes) max_z = std::max(max_z, code.z);
    const bool split = max_z > SPLIT_LANES && block > SPLIT_LANES;
    std::vector<int> cws(nc), part(nc, 0), cws_mx(nc, 0);
    size_t lds[5] = {0, 0, 0, 0, 0};
    // the dynamic LDS an item may take: the device's per-workgroup limit less
    // the largest static LDS of the kernels this plan may launch (their item
    // tables)
    size_t lds_cap = (size_t)p->lds_limit, static_max = 0;
    for (const void *fn : {syn::persist_kernel_i8(), syn::persist_kernel_f32_big(), syn::persist_kernel_f32_small(),
                           syn::persist_kernel_f32_small_bg2()}) {
        hipFuncAttributes fa{};
        if (hipFuncGetAttributes(&fa, fn) != hipSuccess) return SYNDROME_DEVICE_ERROR;
        static_max = std::max(static_max, (size_t)fa.sharedSizeBytes);
    }
    lds_cap -= std::min(lds_cap, static_max);
    // The small codes' launch reserves its largest item's LDS for every
    // workgroup, so an item there takes no more than lets as many
    // SPLIT_LANES-lane workgroups share a CU as the kernel's registers allow;
    // a code that cannot stay under that goes to the main launch.
    size_t small_cap = lds_cap;
    if (split) {
        const void *fn_s = p->i8 ? syn::persist_kernel_i8() : syn::persist_kernel_f32_small();
        int wgs = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&wgs, fn_s, SPLIT_LANES, 0) != hipSuccess || wgs < 1) wgs = 1;
        small_cap = std::min(lds_cap, (size_t)p->lds_limit / (size_t)wgs - std::min((size_t)p->lds_limit / wgs, static_max));
    }
    const int max_k = SYN_ITEM_MAX_CWS;
    auto fit = [&](const SynCode &code, bool pk, int lanes, size_t cap) {
        int k = std::min(max_k, std::max(1, lanes / code.z) * (pk ? 2 : 1));
        while (k > 1 && group_bytes(code, k, p->i8, pk) > cap) --k;
        return k;
    };
    auto fits = [&](const SynCode &code, bool pk, int lanes, size_t cap) {
        return group_bytes(code, fit(code, pk, lanes, cap), p->i8, pk) <= cap;
    };
    // A forked run (plan_launch) overlaps the extension launch with the
    // others.  A run in sequence has it after them on the stream: when its
    // codes are a small part of the plan (a few items per size), its tail
    // would leave most CUs idle behind the main launch, so such a run uses
    // it only where they hold at least half the plan's work (E*Z per block)
    // or a few items per CU.
    size_t x_items = 0;
    double x_work = 0.0, all_work = 0.0;
    {
        std::vector<size_t> per_code(nc, 0);
        for (int cc : p->cb_code) ++per_code[cc];
        for (size_t c = 0; c < nc; ++c) {
            const double w = (double)p->codes[c].n_edges * p->codes[c].z * (double)per_code[c];
            all_work += w;
            if (x_code(p->codes[c], p->i8)) {
                x_items += (per_code[c] + 1) / 2;
                x_work += w;
            }
        }
    }
    p->x_alone = 2.0 * x_work >= all_work || x_items >= (size_t)X_MIN_ITEMS_PER_CU * (size_t)std::max(1, p->cus);
    std::vector<char> packed(nc, 0);
    for (size_t c = 0; c < nc; ++c) {
        const SynCode &code = p->codes[c];
        bool pk = can_pack(code, p->i8);
        part[c] = split && code.z <= SPLIT_LANES ? 1 : 0;
        if (part[c] && !fits(code, pk, SPLIT_LANES, small_cap)) {
            // one pair of a large Z overfills the small launch's share of a
            // CU: such a code runs one codeword per lane there, or else in
            // the main launch
            if (pk && fits(code, false, SPLIT_LANES, small_cap)) pk = false;
            else part[c] = 0;
        }
        const int lanes = part[c] ? SPLIT_LANES : block;
        if (code.z > lanes) return SYNDROME_INVALID_PARAM;
        const int k = fit(code, pk, lanes, part[c] ? small_cap : lds_cap);
        if (group_bytes(code, k, p->i8, pk) > lds_cap) return SYNDROME_INVALID_PARAM;
        cws[c] = k;
        packed[c] = pk;
        int xl = 0, xc = 1;
        const size_t xb = x_code(code, p->i8) ? (size_t)syn::fastx_shape_f32_bg1(code.z, &xl, &xc) : 0;
        if (x_code(code, p->i8) && xb + 64 <= (size_t)p->lds_limit) {
            // the extension-in-HBM launch; cws_mx: its main-launch form
            part[c] = 2;
            cws_mx[c] = k;
            cws[c] = xc;
            lds[2] = std::max(lds[2], xb);
            p->block_x = syn::persistx_lanes_f32_bg1();
            (void)xl;
            continue;
        }
        // f32 BG2 codes of Z <= 256, from the Z <= 256 launch or from an
        // unsplit main launch: the BG2 kernel's launch, of the same lanes
        if (!p->i8 && code.kind == SYN_KIND_BG2 && code.z <= SPLIT_LANES && (part[c] == 1 || !split)) part[c] = 4;
        lds[part[c]] = std::max(lds[part[c]], group_bytes(code, k, p->i8, pk));
    }
    size_t lds_mx = lds[0];
    for (size_t c = 0; c < nc; ++c)
        if (part[c] == 2) lds_mx = std::max(lds_mx, group_bytes(p->codes[c], cws_mx[c], false));
    std::vector<std::vector<int>> members(nc);
    for (size_t i = 0; i < p->cb_code.size(); ++i) members[p->cb_code[i]].push_back((int)i);
    struct Item {
        SynItem it;
        long long cost;
    };
    // items[3]: the main list with the extension-launch codes in main-launch
    // form (its *_mx list); items[4]: the f32 BG2 codes of Z <= 256
    std::vector<Item> items[5];
    std::vector<int> ids;
    auto add = [&](size_t c, size_t j0, int k, int q, bool pk) {
        for (size_t j = 0; j < members[c].size(); j += (size_t)k) {
            const int cnt = (int)std::min(members[c].size() - j, (size_t)k);
            Item x;
            x.it = SynItem{(int)c, p->codes[c].kind | (pk ? SYN_ITEM_PACKED : 0), (int)(j0 + j), cnt};
            x.cost = (long long)p->codes[c].n_edges * p->codes[c].z * ((cnt + 0LL));
            items[q].push_back(x);
        }
    };
    for (size_t c = 0; c < nc; ++c) {
        const size_t j0 = ids.size();
        for (int m : members[c]) ids.push_back(m);
        add(c, j0, cws[c], part[c], packed[c]);
        if (part[c] == 2) add(c, j0, cws_mx[c], 3, false);
        else if (part[c] == 0) add(c, j0, cws[c], 3, packed[c]);
    }
    std::vector<SynItem> flat[5];
    for (int q = 0; q < 5; ++q) {
        // longest-processing-time first: the queue hands out big items early
        std::stable_sort(items[q].begin(), items[q].end(), [](const Item &a, const Item &b) { return a.cost > b.cost; });
        for (auto &x : items[q]) flat[q].push_back(x.it);
    }
    const bool has_x = !flat[2].empty();
    if (int32_t rc = upload(p, &p->d_items, flat[0])) return rc;
    if (int32_t rc = upload(p, &p->d_items_s, flat[1])) return rc;
    if (int32_t rc = upload(p, &p->d_items_x, flat[2])) return rc;
    if (int32_t rc = upload(p, &p->d_items_mx, has_x ? flat[3] : std::vector<SynItem>())) return rc;
    if (int32_t rc = upload(p, &p->d_items_2, flat[4])) return rc;
    if (int32_t rc = upload(p, &p->d_ids, ids)) return rc;
    int main_z = 0, mx_z = 0;  // largest Z in the main launch, without and with the extension-launch codes
    for (size_t c = 0; c < nc; ++c) {
        if (!part[c]) main_z = std::max(main_z, p->codes[c].z);
        if (part[c] == 0 || part[c] == 2) mx_z = std::max(mx_z, p->codes[c].z);
    }
    const void *fn = p->i8 ? syn::persist_kernel_i8()
                           : (main_z > SPLIT_LANES ? syn::persist_kernel_f32_big() : syn::persist_kernel_f32_small());
    const void *fn_s = p->i8 ? syn::persist_kernel_i8() : syn::persist_kernel_f32_small();
    if (syn::allow_max_dynamic_lds(fn) || syn::allow_max_dynamic_lds(fn_s)) return SYNDROME_DEVICE_ERROR;
    p->fn_main = fn;
    p->fn_small = fn_s;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds[0]) != hipSuccess || per_cu < 1)
        per_cu = 1;
    p->block = block;
    p->lds = lds[0];
    p->n_items = (int)flat[0].size();
    p->grid = std::max(1, std::min(p->n_items, per_cu * p->cus));
    p->block_s = SPLIT_LANES;
    p->lds_s = lds[1];
    p->n_items_s = (int)flat[1].size();
    p->grid_s = 0;
    if (p->n_items_s) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn_s, SPLIT_LANES, lds[1]) != hipSuccess || per_cu < 1)
            per_cu = 1;
        p->grid_s = std::max(1, std::min(p->n_items_s, per_cu * p->cus));
    }
    p->block_2 = split ? SPLIT_LANES : block;
    p->lds_2 = lds[4];
    p->n_items_2 = (int)flat[4].size();
    p->grid_2 = 0;
    p->fn_2 = syn::persist_kernel_f32_small_bg2();
    if (p->n_items_2) {
        if (syn::allow_max_dynamic_lds(p->fn_2)) return SYNDROME_DEVICE_ERROR;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, p->fn_2, p->block_2, lds[4]) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        p->grid_2 = std::max(1, std::min(p->n_items_2, per_cu * p->cus));
    }
    p->cws_of_code = cws;
    p->n_items_x = (int)flat[2].size();
    p->lds_x = lds[2];
    p->grid_x = 0;
    p->fn_x = nullptr;
    if (has_x) {
        const void *fx = nullptr;
        // the Z=384 kernel, or the kernel of all four sizes when another is here
        for (size_t c = 0; c < nc; ++c)
            if (part[c] == 2 && (!fx || p->codes[c].z != 384)) fx = syn::persistx_kernel_f32_bg1(p->codes[c].z);
        if (syn::allow_max_dynamic_lds(fx)) return SYNDROME_DEVICE_ERROR;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fx, p->block_x, lds[2]) != hipSuccess || per_cu < 1)
            per_cu = 1;
        p->fn_x = fx;
        p->grid_x = std::max(1, std::min(p->n_items_x, per_cu * p->cus));
        const void *fm = mx_z > SPLIT_LANES ? syn::persist_kernel_f32_big() : syn::persist_kernel_f32_small();
        if (syn::allow_max_dynamic_lds(fm)) return SYNDROME_DEVICE_ERROR;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fm, block, lds_mx) != hipSuccess || per_cu < 1)
            per_cu = 1;
        p->fn_mx = fm;
        p->n_items_mx = (int)flat[3].size();
        p->lds_mx = lds_mx;
        p->grid_mx = std::max(1, std::min(p->n_items_mx, per_cu * p->cus));
    } else {
        p->fn_mx = p->fn_main;
        p->n_items_mx = p->n_items;
        p->lds_mx = p->lds;
        p->grid_mx = p->grid;
    }
    return SYNDROME_OK;
}

int32_t plan_launch(SyndromeLdpcPlan *p, const void *in, void *post, uint8_t *hard, int32_t *iters, size_t iterations,
                    float scale, int16_t app_clamp, uint32_t flags, hipStream_t stream)
{
    SynLaunch a{};
    a.in = in;
    a.post = post;
    a.hard = hard;
    a.iters = iters;
    a.batch = (int)p->bg.size();
    a.iterations = (int)std::min(iterations, (size_t)INT_MAX);
    a.flags = flags & (SYN_F_INIT_5G | SYN_F_WRITE_POST | SYN_F_WRITE_HARD | SYN_F_QUANTIZE_IN);
    {
        const auto al = [](const void *q) { return ((uintptr_t)q & 15u) == 0; };
        if (p->offs_aligned && al(in) && (!(a.flags & SYN_F_WRITE_POST) || al(post)) &&
            (!(a.flags & SYN_F_WRITE_HARD) || al(hard)))
            a.flags |= SYN_F_ALIGNED;  // every 3GPP N = 68Z or 52Z is a multiple of 4
    }
    a.beta = p->beta;
    a.scale = scale;
    a.app_clamp = app_clamp < 0 ? 0 : app_clamp;
    if (p->i8) {
        const float b = p->beta * scale;
        a.beta_q = !(b - b == 0.f) ? 127 : (int)std::min(127.f, std::max(0.f, roundf(b)));
    }
    const SynCode *codes = p->d_codes;
    const int *ids = p->d_ids;
    const long long *offs = p->d_offs;
    const int *nf = p->d_nfill;
    // this run's counters: runs in flight at once (on different streams, or
    // a graph replay beside an eager run) never share one
    int *heads = p->d_head + (size_t)(p->run_seq++ % RUN_RING) * LAUNCHES * HEAD_STRIDE;
    // Each launch is a persistent grid whose workgroups leave once its queue
    // is drained, and none waits for another, so launches may run in any
    // overlap.  An eager run forks them over the caller's stream and the
    // plan's side stream (below).  A run under stream capture keeps them in
    // sequence on the caller's stream: a side stream joined into one
    // caller's capture would capture another caller's eager run.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const bool can_fork = p->ev_fork && hipStreamIsCapturing(stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone;
    // the extension-in-HBM launch needs the posterior buffer and an
    // iteration; without them, or where it does not pay in sequence, its
    // codes go in the main launch
    const bool use_x = p->n_items_x > 0 && (a.flags & SYN_F_WRITE_POST) && a.iterations >= 1 && (can_fork || p->x_alone);
    struct L {
        const void *fn;
        const SynItem *items;
        int n, grid, block;
        size_t lds;
    } ls[LAUNCHES] = {
        use_x ? L{p->fn_main, p->d_items, p->n_items, p->grid, p->block, p->lds}
              : L{p->fn_mx, p->n_items_x ? p->d_items_mx : p->d_items, p->n_items_mx, p->grid_mx, p->block, p->lds_mx},
        L{p->fn_small, p->d_items_s, p->n_items_s, p->grid_s, p->block_s, p->lds_s},
        L{p->fn_x, p->d_items_x, use_x ? p->n_items_x : 0, p->grid_x, p->block_x, p->lds_x},
        L{p->fn_2, p->d_items_2, p->n_items_2, p->grid_2, p->block_2, p->lds_2}};
    // the main launch, the small codes', the extension launch, the small BG2
    // codes' (each with its own counter)
    const auto launch = [&](int q, hipStream_t s) -> int32_t {
        const SynItem *items = ls[q].items;
        int n_items = ls[q].n;
        int *head = heads + q * HEAD_STRIDE;
        if (syn::launch_reset(head, s)) return SYNDROME_DEVICE_ERROR;
        void *args[] = {&codes, &items, &n_items, &head, &ids, &offs, &nf, &a};
        if (hipLaunchKernel(ls[q].fn, dim3(ls[q].grid), dim3(ls[q].block), args, ls[q].lds, s) != hipSuccess)
            return SYNDROME_DEVICE_ERROR;
        return SYNDROME_OK;
    };
    int n_act = 0;
    for (int q = 0; q < LAUNCHES; ++q) n_act += ls[q].n > 0;
    // Two streams: the caller's and the side stream, each running its
    // launches in order.  With the extension launch, [main, then the small
    // BG1 codes'] runs beside [extension, then the small BG2 codes']: the
    // two launches of whole-CU workgroups start together with the longest
    // items, and each stream's small-code launch fills its CUs as they
    // drain.  Without it, the main launch runs beside the small codes'
    // launches.  (Uneven workload, f32 decode-only: 9.22 Gbit/s in
    // sequence, 10.03 so; int8: 8.33 in sequence, 9.5 beside.)
    const bool side_of[LAUNCHES] = {false, !use_x, true, true};
    const bool fork = can_fork && n_act > 1;
    if (!fork) {
        for (int q : {0, 2, 1, 3})
            if (ls[q].n)
                if (int32_t rc = launch(q, stream)) return rc;
        return SYNDROME_OK;
    }
    if (hipEventRecord(p->ev_fork, stream) != hipSuccess || hipStreamWaitEvent(p->side, p->ev_fork, 0) != hipSuccess)
        return SYNDROME_DEVICE_ERROR;
    int32_t rc = SYNDROME_OK;
    for (int q : {0, 2, 1, 3})
        if (ls[q].n && (rc = launch(q, side_of[q] ? p->side : stream))) break;
    if (hipEventRecord(p->ev_join, p->side) != hipSuccess || hipStreamWaitEvent(stream, p->ev_join, 0) != hipSuccess)
        rc = SYNDROME_DEVICE_ERROR;
    return rc;
}

}  // namespace

extern "C" {

int32_t syndrome_ldpc_plan_create(const int32_t *bg, const int32_t *z, const int32_t *n_filler, const int64_t *offsets,
                                  size_t count, float offset_beta, int32_t int8, int32_t block,
                                  SyndromeLdpcPlan **out)
{
    return guard([&]() -> int32_t {
        if (!out) return SYNDROME_NULL_POINTER;
        *out = nullptr;
        if (count && (!bg || !z || !offsets)) return SYNDROME_NULL_POINTER;
        if (count == 0 || count > (size_t)INT_MAX / 2) return SYNDROME_INVALID_PARAM;
        std::unique_ptr<SyndromeLdpcPlan> p(new SyndromeLdpcPlan());
        p->i8 = int8 != 0;
        p->beta = offset_beta;
        if (hipGetDevice(&p->device) != hipSuccess) return SYNDROME_DEVICE_ERROR;
        if (hipDeviceGetAttribute(&p->lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device) != hipSuccess ||
            hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess)
            return SYNDROME_DEVICE_ERROR;
        std::map<std::pair<int, int>, int> index;
        int max_z = 0;
        for (size_t i = 0; i < count; ++i) {
            if (bg[i] != 0 && bg[i] != 1) return SYNDROME_INVALID_PARAM;
            if (offsets[i] < 0 || (n_filler && n_filler[i] < 0)) return SYNDROME_INVALID_PARAM;
            auto key = std::make_pair((int)bg[i], (int)z[i]);
            auto f = index.find(key);
            if (f == index.end()) {
                SynCode c;
                if (syn::build_code_3gpp(bg[i], z[i], &c)) return SYNDROME_INVALID_PARAM;
                f = index.emplace(key, (int)p->codes.size()).first;
                p->codes.push_back(c);
            }
            p->cb_code.push_back(f->second);
            p->bg.push_back(bg[i]);
            p->z.push_back(z[i]);
            p->nfill.push_back(n_filler ? n_filler[i] : 0);
            p->offs.push_back(offsets[i]);
            p->offs_aligned = p->offs_aligned && (offsets[i] % 4 == 0);
            max_z = std::max(max_z, (int)z[i]);
        }
        DeviceScope ds(p->device);
        if (int32_t rc = upload(p.get(), &p->d_codes, p->codes)) return rc;
        if (int32_t rc = upload(p.get(), &p->d_offs, p->offs)) return rc;
        if (int32_t rc = upload(p.get(), &p->d_nfill, p->nfill)) return rc;
        if (hipMalloc((void **)&p->d_head, RUN_RING * LAUNCHES * HEAD_STRIDE * sizeof(int)) != hipSuccess)
            return SYNDROME_DEVICE_ERROR;
        if (hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess)
            return SYNDROME_DEVICE_ERROR;
        const int min_block = ((max_z + 63) / 64) * 64;
        const int b = block > 0 ? block : std::max(min_block, 512);
        if (int32_t rc = build_items(p.get(), std::max(b, min_block))) return rc;
        *out = p.release();
        return SYNDROME_OK;
    });
}

void syndrome_ldpc_plan_destroy(SyndromeLdpcPlan *p)
{
    if (!p) return;
    try {
        DeviceScope ds(p->device);
        for (void *q : {(void *)p->d_codes, (void *)p->d_items, (void *)p->d_items_s, (void *)p->d_items_x,
                        (void *)p->d_items_mx, (void *)p->d_items_2, (void *)p->d_ids, (void *)p->d_nfill,
                        (void *)p->d_head, (void *)p->d_offs, (void *)p->d_dm, (void *)p->d_dm_chunks,
                        (void *)p->d_tbs})
            if (q) (void)hipFree(q);
        for (void *q : p->retired) (void)hipFree(q);
        if (p->side) (void)hipStreamDestroy(p->side);
        if (p->ev_join) (void)hipEventDestroy(p->ev_join);
        if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    } catch (...) {
    }
    delete p;
}

int32_t syndrome_ldpc_plan_run(SyndromeLdpcPlan *p, const void *d_in, void *d_post, uint8_t *d_hard, int32_t *d_iters,
                               size_t iterations, float scale, int16_t app_clamp, uint32_t flags, void *stream)
{
    return guard([&]() -> int32_t {
        if (!p || !d_in || !d_iters) return SYNDROME_NULL_POINTER;
        if ((flags & SYNDROME_BATCH_WRITE_POST) && !d_post) return SYNDROME_NULL_POINTER;
        if ((flags & SYNDROME_BATCH_WRITE_HARD) && !d_hard) return SYNDROME_NULL_POINTER;
        if ((flags & SYNDROME_BATCH_QUANTIZE) && !p->i8) return SYNDROME_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        DeviceScope ds(p->device);
        return plan_launch(p, d_in, d_post, d_hard, d_iters, iterations, scale, app_clamp, flags, (hipStream_t)stream);
    });
}

int32_t syndrome_ldpc_plan_shape(const SyndromeLdpcPlan *p, int32_t *block, int32_t *grid, int32_t *items,
                                 int32_t *lds_bytes)
{
    return guard([&]() -> int32_t {
        if (!p) return SYNDROME_NULL_POINTER;
        // the main launch's shape, or, when the main launch has no items, the
        // extension launch's or the small BG2 codes'
        const bool x = p->n_items == 0 && p->n_items_x > 0;
        const bool b2 = p->n_items == 0 && !x && p->n_items_2 > 0;
        if (block) *block = x ? p->block_x : b2 ? p->block_2 : p->block;
        if (grid) *grid = x ? p->grid_x : b2 ? p->grid_2 : p->grid;
        if (items) *items = p->n_items + p->n_items_s + p->n_items_x + p->n_items_2;  // every launch's items
        if (lds_bytes) *lds_bytes = (int32_t)(x ? p->lds_x : b2 ? p->lds_2 : p->lds);
        return SYNDROME_OK;
    });
}

// Launch-shape tuner: time the plan at each candidate workgroup size (the
// lanes that set how many codewords of each code share a workgroup and
// the LDS of the launch) on the caller's data and keep the fastest.
int32_t syndrome_ldpc_plan_tune(SyndromeLdpcPlan *p, const void *d_in, void *d_post, uint8_t *d_hard,
                                int32_t *d_iters, size_t iterations, float scale, int16_t app_clamp, uint32_t flags,
                                int32_t reps, int32_t *out_block)
{
    return guard([&]() -> int32_t {
        if (!p || !d_in || !d_iters) return SYNDROME_NULL_POINTER;
        std::lock_guard<std::mutex> lk(p->mu);
        DeviceScope ds(p->device);
        int max_z = 0;
        for (int zz : p->z) max_z = std::max(max_z, zz);
        const int min_block = ((max_z + 63) / 64) * 64;
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return SYNDROME_DEVICE_ERROR;
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        float best_ms = 1e30f;
        int best = p->block;
        const int nrep = reps > 0 ? reps : 2;
        for (int b = min_block; b <= 512; b += 64) {
            if (build_items(p, b)) continue;
            if (plan_launch(p, d_in, d_post, d_hard, d_iters, iterations, scale, app_clamp, flags, s)) continue;
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < nrep; ++r)
                (void)plan_launch(p, d_in, d_post, d_hard, d_iters, iterations, scale, app_clamp, flags, s);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best_ms) {
                best_ms = ms;
                best = b;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        if (int32_t rc = build_items(p, best)) return rc;
        if (out_block) *out_block = best;
        return SYNDROME_OK;
    });
}

int32_t syndrome_ldpc_plan_set_dematch(SyndromeLdpcPlan *p, const int64_t *e_offsets, const int32_t *e_lengths,
                                       const int32_t *qm, const int32_t *rv)
{
    return guard([&]() -> int32_t {
        if (!p || !e_offsets || !e_lengths) return SYNDROME_NULL_POINTER;
        std::vector<SynDematchCb> v(p->bg.size());
        for (size_t i = 0; i < v.size(); ++i) {
            const int q = qm ? qm[i] : 1, r = rv ? rv[i] : 0;
            if (q <= 0 || e_lengths[i] < 0 || e_lengths[i] % q || r < 0 || r > 3 || e_offsets[i] < 0)
                return SYNDROME_INVALID_PARAM;
            v[i] = SynDematchCb{e_offsets[i], p->offs[i], e_lengths[i], q, p->z[i], p->bg[i], r, p->nfill[i]};
            syn::dematch_derive(v[i]);
        }
        std::vector<SynDematchChunk> ch;
        syn::dematch_chunks(v.data(), (int)v.size(), ch);
        if (ch.size() > (size_t)INT_MAX) return SYNDROME_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        DeviceScope ds(p->device);
        if (int32_t rc = upload(p, &p->d_dm, v)) return rc;
        if (int32_t rc = upload(p, &p->d_dm_chunks, ch)) return rc;
        p->n_dm_chunks = (int)ch.size();
        return SYNDROME_OK;
    });
}

int32_t syndrome_ldpc_plan_dematch(SyndromeLdpcPlan *p, const float *d_e_llr, float *d_llr_out, void *stream)
{
    return guard([&]() -> int32_t {
        if (!p || !d_e_llr || !d_llr_out) return SYNDROME_NULL_POINTER;
        if (!p->d_dm) return SYNDROME_INVALID_PARAM;
        DeviceScope ds(p->device);
        return syn::launch_dematch(d_e_llr, d_llr_out, p->d_dm, p->d_dm_chunks, p->n_dm_chunks, false, -1,
                                   (hipStream_t)stream)
                   ? SYNDROME_DEVICE_ERROR
                   : SYNDROME_OK;
    });
}

int32_t syndrome_ldpc_plan_set_tbs(SyndromeLdpcPlan *p, size_t n_tb, const int32_t *tb_first_cb,
                                   const int32_t *tb_num_cb, const int32_t *tb_size, const int64_t *tb_out_off)
{
    return guard([&]() -> int32_t {
        if (!p || (n_tb && (!tb_first_cb || !tb_num_cb || !tb_size || !tb_out_off))) return SYNDROME_NULL_POINTER;
        if (n_tb > (size_t)INT_MAX) return SYNDROME_INVALID_PARAM;
        std::vector<SynTb> v(n_tb);
        const long long count = (long long)p->bg.size();
        for (size_t t = 0; t < n_tb; ++t) {
            const long long f = tb_first_cb[t], c = tb_num_cb[t];
            if (f < 0 || c < 1 || f + c > count || tb_size[t] < 1 || tb_out_off[t] < 0) return SYNDROME_INVALID_PARAM;
            // K' of every block: K - n_filler (all blocks of a TB share it, segmentation.rs:298-303)
            const auto kprime = [&](long long i) {
                return (p->bg[i] == 0 ? 22 : 10) * p->z[i] - p->nfill[i];
            };
            const int kp = kprime(f);
            for (long long i = f; i < f + c; ++i)
                if (kprime(i) != kp) return SYNDROME_INVALID_PARAM;
            const long long pay = c > 1 ? kp - 24 : kp;
            if (pay < 1 || c * pay < (long long)tb_size[t] + 24) return SYNDROME_INVALID_PARAM;
            v[t] = SynTb{tb_out_off[t], (int)f, (int)c, kp, tb_size[t]};
        }
        std::lock_guard<std::mutex> lk(p->mu);
        DeviceScope ds(p->device);
        if (int32_t rc = upload(p, &p->d_tbs, v)) return rc;
        p->n_tbs = (int)n_tb;
        return SYNDROME_OK;
    });
}

int32_t syndrome_ldpc_plan_check_tbs(SyndromeLdpcPlan *p, const uint8_t *d_hard, uint8_t *d_tb_out, uint8_t *d_tb_ok,
                                     uint8_t *d_cb_ok, void *stream)
{
    return guard([&]() -> int32_t {
        if (!p || !d_hard || !d_tb_out || !d_tb_ok || !d_cb_ok) return SYNDROME_NULL_POINTER;
        if (!p->d_tbs) return SYNDROME_INVALID_PARAM;
        DeviceScope ds(p->device);
        return syn::launch_tb_check(d_hard, p->d_offs, p->d_tbs, p->n_tbs, d_tb_out, d_tb_ok, d_cb_ok,
                                    (hipStream_t)stream)
                   ? SYNDROME_DEVICE_ERROR
                   : SYNDROME_OK;
    });
}

int32_t syndrome_ldpc_plan_harq_combine(SyndromeLdpcPlan *p, const float *d_e_llr, float *d_acc, int32_t rv,
                                        void *stream)
{
    return guard([&]() -> int32_t {
        if (!p || !d_e_llr || !d_acc) return SYNDROME_NULL_POINTER;
        if (!p->d_dm || rv > 3) return SYNDROME_INVALID_PARAM;  // rv < 0: each block's own
        DeviceScope ds(p->device);
        return syn::launch_dematch(d_e_llr, d_acc, p->d_dm, p->d_dm_chunks, p->n_dm_chunks, true, rv < 0 ? -1 : rv,
                                   (hipStream_t)stream)
                   ? SYNDROME_DEVICE_ERROR
                   : SYNDROME_OK;
    });
}

}  // extern "C"
