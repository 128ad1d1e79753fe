// C ABI of the MI355X LOMS decoder (declarations: include/syndrome_ldpc.h).
//
// Part 1 mirrors capi/src/lib.rs symbol for symbol: null checks first, then
// the Rust validation order (qc_ldpc.rs:711-732, 980-1004), status codes
// from capi/src/lib.rs:74-105, and no exception ever crosses the boundary
// (the `guard` of capi/src/lib.rs:117-119).
#include <hip/hip_runtime.h>

#include <limits.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/syndrome_ldpc.h"
#include "code_host.h"
#include "ldpc_common.h"

namespace syn {
// One host-buffer call's device staging and its stream.
struct Stage {
    hipStream_t stream = nullptr;
    void *d = nullptr;
    size_t bytes = 0;
};
int launch_decode(int kind, bool i8, bool fast, const SynCode *d_code, int z, const SynLaunch &a, int block,
                  size_t lds, hipStream_t stream);
int launch_quantize(const float *in, int16_t *out, size_t n, float scale, hipStream_t stream);
int launch_quantize_i8(const float *in, int8_t *out, size_t n, float scale, hipStream_t stream);
int launch_dequantize_i8(const int8_t *in, float *out, size_t n, float scale, hipStream_t stream);
WideShape wide_shape(const SynWideCode &w, bool i8, int lds_limit);
int launch_wide(bool i8, const SynWideCode &w, const SynLaunch &a, const WideShape &s, char *work,
                hipStream_t stream);
size_t encode_lds_bytes(const SynCode &c, const SynEncCode &e, const void *info, const void *cw);
int launch_encode(const SynCode *d_code, const SynEncCode *d_enc, int z, int k, int n, size_t lds,
                  const uint8_t *info, uint8_t *cw, long long count, hipStream_t stream);
}

struct SyndromeLdpcDecoder {
    SynCode code;
    float beta = 0.f;
    int device = 0;
    int lds_limit = 0;
    SynCode *d_code = nullptr;
    // a raw code outside the inline tables (code.kind == SYN_KIND_WIDE):
    // host tables, and the header whose table pointers address d_wide
    std::unique_ptr<syn::WideCode> wide;
    SynWideCode wide_dev{};
    void *d_wide = nullptr;
    // host-buffer entry points: a pool of staging areas, each with its own
    // stream, so concurrent calls on one handle run side by side (the
    // reference's decoder is Send + Sync, qc_ldpc.rs:480-484)
    std::mutex pool_mu;
    std::condition_variable pool_cv;
    std::vector<std::unique_ptr<syn::Stage>> stages;  // every staging area made
    std::vector<syn::Stage *> free_stages;
    // lazily built encoder for traffic generation
    std::once_flag enc_once;
    std::unique_ptr<syn::Encoder> enc;
    // its device form (circulant terms of the core inverse), built on the
    // first device encode
    std::once_flag denc_once;
    SynEncCode *d_enc = nullptr;
    std::unique_ptr<SynEncCode> h_enc;
    int32_t denc_status = SYNDROME_OK;
};

struct SyndromeLdpcEncoder {
    SynCode code;
    std::unique_ptr<syn::Encoder> enc;
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

// Restores the caller's current HIP device on scope exit.
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

void free_decoder(SyndromeLdpcDecoder *d)
{
    DeviceScope ds(d->device);
    if (d->d_code) (void)hipFree(d->d_code);
    if (d->d_wide) (void)hipFree(d->d_wide);
    for (auto &st : d->stages) {
        if (st->d) (void)hipFree(st->d);
        if (st->stream) (void)hipStreamDestroy(st->stream);
    }
    if (d->d_enc) (void)hipFree(d->d_enc);
}

struct DecoderOwner {
    void operator()(SyndromeLdpcDecoder *d) const
    {
        free_decoder(d);
        delete d;
    }
};

// The wide form's tables in one device allocation: layer offsets, columns,
// shifts, then the dup flags.
int32_t upload_wide(SyndromeLdpcDecoder *d)
{
    const syn::WideCode &w = *d->wide;
    const size_t b_off = w.layer_off.size() * 4, b_cols = w.cols.size() * 4, b_dup = w.dup.size();
    std::vector<char> h(b_off + 2 * b_cols + b_dup + 16);
    memcpy(h.data(), w.layer_off.data(), b_off);
    if (b_cols) {
        memcpy(h.data() + b_off, w.cols.data(), b_cols);
        memcpy(h.data() + b_off + b_cols, w.shifts.data(), b_cols);
    }
    if (b_dup) memcpy(h.data() + b_off + 2 * b_cols, w.dup.data(), b_dup);
    if (hipMalloc(&d->d_wide, h.size()) != hipSuccess) return SYNDROME_DEVICE_ERROR;
    if (hipMemcpy(d->d_wide, h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess) return SYNDROME_DEVICE_ERROR;
    char *base = (char *)d->d_wide;
    d->wide_dev = w.hdr;
    d->wide_dev.layer_off = (const int *)base;
    d->wide_dev.cols = (const int *)(base + b_off);
    d->wide_dev.shifts = (const int *)(base + b_off + b_cols);
    d->wide_dev.dup = (const unsigned char *)(base + b_off + 2 * b_cols);
    return SYNDROME_OK;
}

int32_t make_decoder(const SynCode &code, float beta, SyndromeLdpcDecoder **out,
                     std::unique_ptr<syn::WideCode> wide = nullptr)
{
    std::unique_ptr<SyndromeLdpcDecoder, DecoderOwner> d(new SyndromeLdpcDecoder());
    d->code = code;
    d->beta = beta;
    d->wide = std::move(wide);
    if (hipGetDevice(&d->device) != hipSuccess) return SYNDROME_DEVICE_ERROR;
    if (hipDeviceGetAttribute(&d->lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, d->device) != hipSuccess)
        return SYNDROME_DEVICE_ERROR;
    if (hipMalloc((void **)&d->d_code, sizeof(SynCode)) != hipSuccess) return SYNDROME_DEVICE_ERROR;
    if (hipMemcpy(d->d_code, &d->code, sizeof(SynCode), hipMemcpyHostToDevice) != hipSuccess)
        return SYNDROME_DEVICE_ERROR;
    if (d->code.kind == SYN_KIND_WIDE)
        if (int32_t rc = upload_wide(d.get())) return rc;
    *out = d.release();
    return SYNDROME_OK;
}

struct Shape {
    int cws = 0, block = 0;
    size_t lds = 0;
};

bool use_fast(const SynCode &c, bool plain)
{
    return !plain && (c.kind == SYN_KIND_BG1 || c.kind == SYN_KIND_BG2 || c.kind >= SYN_KIND_WIFI0);
}

// Codewords per workgroup: as many whole codewords as fit the lane budget
// (1024 lanes for the plain kernels; 256 for the register-resident ones,
// up to one codeword of any size) and the LDS of one workgroup.  Small
// register-resident workgroups let several share a CU with every SIMD
// evenly loaded: int8 BG2 Z=208 as four 4-wave workgroups per CU instead
// of two 7-wave ones went from 9.1 to 10.8 Gbit/s, 802.11 Z=27 with 9
// codewords per workgroup instead of 18 from 16.8 to 21.7.
int32_t launch_shape(const SyndromeLdpcDecoder *d, bool i8, bool plain, Shape *s)
{
    const SynCode &c = d->code;
    if (c.kind == SYN_KIND_WIDE) {
        const WideShape w = syn::wide_shape(d->wide_dev, i8, d->lds_limit);
        s->cws = 1;
        s->block = w.block;
        s->lds = w.lds;
        return SYNDROME_OK;
    }
    const size_t n = syn::code_n(c);
    if (i8 && use_fast(c, plain) && c.kind == SYN_KIND_BG2) {
        // pair-packed kernel (ldpc_pack.h): codeword pairs, 32-bit words of
        // two interleaved i16 posteriors, up to 256 lanes
        const size_t pad = (size_t)(((c.z * 4) + 15) & ~15);
        auto bytes = [&](int k) { return pad + (size_t)k * n * 4 + 3 * sizeof(int) * (size_t)(2 * k); };
        int k = std::max(1, 256 / c.z);
        while (k > 1 && bytes(k) > (size_t)d->lds_limit) --k;
        if (bytes(k) > (size_t)d->lds_limit) return SYNDROME_INVALID_PARAM;
        s->cws = 2 * k;
        s->block = ((k * c.z + 63) / 64) * 64;
        s->lds = bytes(k);
        return SYNDROME_OK;
    }
    const size_t elem = i8 ? 2 : 4;
    const int max_lanes = use_fast(c, plain) ? 256 : 1024;
    // fast kernels keep a Z*size pad in front of the posterior (ldpc_fast.h)
    const size_t pad = use_fast(c, plain) ? (size_t)(((c.z * (int)elem) + 15) & ~15) : 0;
    auto bytes = [&](int k) { return pad + ((k * n * elem + 3) & ~(size_t)3) + 3 * sizeof(int) * (size_t)k; };
    int k = std::max(1, max_lanes / c.z);
    while (k > 1 && bytes(k) > (size_t)d->lds_limit) --k;
    if (bytes(k) > (size_t)d->lds_limit) return SYNDROME_INVALID_PARAM;
    s->cws = k;
    s->block = ((k * c.z + 63) / 64) * 64;
    s->lds = bytes(k);
    return SYNDROME_OK;
}

int32_t run_batch(const SyndromeLdpcDecoder *d, bool i8, const void *in, void *post, void *edge, uint8_t *hard,
                  int32_t *iters, size_t batch, size_t iterations, size_t n_filler, float scale, int16_t app_clamp,
                  uint32_t flags, hipStream_t stream, float *trace = nullptr, char *work = nullptr)
{
    if (batch == 0) return SYNDROME_OK;
    if (batch > (size_t)INT_MAX / 2) return SYNDROME_INVALID_PARAM;
    if ((flags & SYNDROME_BATCH_INIT_5G) && !syn::code_5g_ok(d->code)) return SYNDROME_INVALID_PARAM;
    // a traced decode runs on the plain kernel (same bits; the trace store
    // stays out of the register-resident kernels' register budget)
    const bool plain = (flags & SYNDROME_BATCH_PLAIN) != 0 || (trace && !i8);
    if (!in || !iters) return SYNDROME_NULL_POINTER;
    if ((flags & SYNDROME_BATCH_WRITE_POST) && !post) return SYNDROME_NULL_POINTER;
    if ((flags & SYNDROME_BATCH_WRITE_HARD) && !hard) return SYNDROME_NULL_POINTER;
    const bool fast = use_fast(d->code, plain);
    if ((!fast || (flags & SYNDROME_BATCH_WRITE_EDGE)) && !edge) return SYNDROME_NULL_POINTER;
    Shape s;
    if (int32_t rc = launch_shape(d, i8, plain, &s)) return rc;
    SynLaunch a{};
    a.in = in;
    a.post = post;
    a.edge = edge;
    a.hard = hard;
    a.iters = iters;
    a.batch = (int)batch;
    a.iterations = (int)std::min(iterations, (size_t)INT_MAX);
    a.cws = s.cws;
    a.flags = flags & SYN_F_PUBLIC;
    {
        const auto al = [](const void *q) { return ((uintptr_t)q & 15u) == 0; };
        if (syn::code_n(d->code) % 4 == 0 && al(in) && (!(flags & SYNDROME_BATCH_WRITE_POST) || al(post)) &&
            (!(flags & SYNDROME_BATCH_WRITE_HARD) || al(hard)))
            a.flags |= SYN_F_ALIGNED;  // contiguous codewords: offsets b * N
    }
    if (trace && !i8) {
        a.trace = trace;
        a.flags |= SYN_F_TRACE;
    }
    a.n_filler = (long long)std::min(n_filler, (size_t)LLONG_MAX);
    a.beta = d->beta;
    a.scale = scale;
    a.app_clamp = app_clamp < 0 ? 0 : app_clamp;
    if (i8) {
        // QuantParams::beta_q (quantize.rs:298-304): round half away from
        // zero, clamp to [0, 127], non-finite -> 127.
        const float b = d->beta * scale;
        a.beta_q = !(b - b == 0.f) ? 127 : (int)std::min(127.f, std::max(0.f, roundf(b)));
    }
    DeviceScope ds(d->device);
    if (d->code.kind != SYN_KIND_WIDE)
        return syn::launch_decode(d->code.kind, i8, fast, d->d_code, d->code.z, a, s.block, s.lds, stream);
    // the wide form: its workspace is the caller's (staged decodes) or
    // stream-ordered memory of this call's own
    const WideShape w = syn::wide_shape(d->wide_dev, i8, d->lds_limit);
    const size_t ws = w.work_stride * batch;
    const bool own = ws && !work;
    if (own && hipMallocAsync((void **)&work, ws, stream) != hipSuccess) return SYNDROME_DEVICE_ERROR;
    const int rc = syn::launch_wide(i8, d->wide_dev, a, w, work, stream);
    if (own && hipFreeAsync(work, stream) != hipSuccess) return SYNDROME_DEVICE_ERROR;
    return rc ? SYNDROME_DEVICE_ERROR : SYNDROME_OK;
}

// Staging areas of a handle: up to STAGES_MAX, made on demand; a call takes
// a free one (waiting when all are busy), grows it if needed, and returns
// it.
constexpr size_t STAGES_MAX = 8;

int32_t take_stage(SyndromeLdpcDecoder *d, size_t bytes, syn::Stage **out)
{
    syn::Stage *st = nullptr;
    {
        std::unique_lock<std::mutex> lk(d->pool_mu);
        d->pool_cv.wait(lk, [&] { return !d->free_stages.empty() || d->stages.size() < STAGES_MAX; });
        if (!d->free_stages.empty()) {
            st = d->free_stages.back();
            d->free_stages.pop_back();
        } else {
            d->stages.emplace_back(new syn::Stage());
            st = d->stages.back().get();
        }
    }
    if (!st->stream && hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking) != hipSuccess) {
        st->stream = nullptr;
    } else if (st->bytes < bytes) {
        if (st->d) (void)hipFree(st->d);
        st->d = nullptr;
        st->bytes = 0;
        if (hipMalloc(&st->d, bytes) == hipSuccess) st->bytes = bytes;
        else st->d = nullptr;
    }
    if (!st->stream || st->bytes < bytes) {
        std::lock_guard<std::mutex> lk(d->pool_mu);
        d->free_stages.push_back(st);
        d->pool_cv.notify_one();
        return SYNDROME_DEVICE_ERROR;
    }
    *out = st;
    return SYNDROME_OK;
}

void give_stage(SyndromeLdpcDecoder *d, syn::Stage *st)
{
    std::lock_guard<std::mutex> lk(d->pool_mu);
    d->free_stages.push_back(st);
    d->pool_cv.notify_one();
}

// Where one codeword's staging lives in a device buffer: posterior (elem
// bytes per variable), R messages, hard decisions, iteration count, the
// workspace of a wide-form code, then the optional trace; each part
// 256-byte aligned.
struct StageLayout {
    size_t off_edge, off_hard, off_iter, off_work, off_trace, total;
};

StageLayout stage_layout(const SyndromeLdpcDecoder *d, bool i8, size_t trace_floats)
{
    const SynCode &c = d->code;
    const size_t n = syn::code_n(c), ez = syn::code_ez(c);
    const size_t pe = i8 ? 2 : 4, me = i8 ? 1 : 4;
    const size_t work = c.kind == SYN_KIND_WIDE ? syn::wide_shape(d->wide_dev, i8, d->lds_limit).work_stride : 0;
    StageLayout L;
    L.off_edge = (n * pe + 255) & ~(size_t)255;
    L.off_hard = L.off_edge + ((ez * me + 255) & ~(size_t)255);
    L.off_iter = L.off_hard + ((n + 255) & ~(size_t)255);
    L.off_work = L.off_iter + 256;
    L.off_trace = L.off_work + work;
    L.total = L.off_trace + trace_floats * 4;
    return L;
}

// One codeword through the GPU from host buffers, staged at `base` (device
// memory laid out by stage_layout) on `stream`: the path behind the
// reference-shaped entry points.  `elem` = 4 (f32 posterior, f32 messages)
// or 2 (i16 posterior, i8 messages).  Returns once the results are back in
// the caller's buffers.
int32_t host_decode_at(SyndromeLdpcDecoder *d, char *base, hipStream_t stream, bool i8, void *llr, void *edge_r,
                       uint8_t *hard, size_t iterations, size_t n_filler, bool init5g, float scale,
                       int16_t app_clamp, size_t *out_used, float *trace)
{
    const SynCode &c = d->code;
    const size_t n = syn::code_n(c), ez = syn::code_ez(c);
    const size_t pe = i8 ? 2 : 4, me = i8 ? 1 : 4;
    const StageLayout L = stage_layout(d, i8, trace ? iterations * n : 0);
    const uint32_t flags = (init5g ? SYNDROME_BATCH_INIT_5G : 0u) | SYNDROME_BATCH_WRITE_POST |
                           SYNDROME_BATCH_WRITE_HARD | SYNDROME_BATCH_WRITE_EDGE;
    if (hipMemcpyAsync(base, llr, n * pe, hipMemcpyHostToDevice, stream) != hipSuccess) return SYNDROME_DEVICE_ERROR;
    int32_t rc = run_batch(d, i8, base, base, base + L.off_edge, (uint8_t *)(base + L.off_hard),
                           (int32_t *)(base + L.off_iter), 1, iterations, n_filler, scale, app_clamp, flags, stream,
                           trace ? (float *)(base + L.off_trace) : nullptr, base + L.off_work);
    if (rc) return rc;
    int32_t used = 0;
    if (hipMemcpyAsync(llr, base, n * pe, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipMemcpyAsync(edge_r, base + L.off_edge, ez * me, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipMemcpyAsync(hard, base + L.off_hard, n, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipMemcpyAsync(&used, base + L.off_iter, 4, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return SYNDROME_DEVICE_ERROR;
    if (trace && used > 0 &&
        (hipMemcpyAsync(trace, base + L.off_trace, (size_t)used * n * 4, hipMemcpyDeviceToHost, stream) !=
             hipSuccess ||
         hipStreamSynchronize(stream) != hipSuccess))
        return SYNDROME_DEVICE_ERROR;
    if (out_used) *out_used = (size_t)used;
    return SYNDROME_OK;
}

// The handle's staging pool: each call takes a staging area of its own
// (allocated on first use and grown), so calls on one handle overlap.
int32_t host_decode(SyndromeLdpcDecoder *d, bool i8, void *llr, void *edge_r, size_t edge_elems, uint8_t *hard,
                    size_t iterations, size_t n_filler, bool init5g, float scale, int16_t app_clamp,
                    size_t *out_used, float *trace = nullptr)
{
    (void)edge_elems;
    const StageLayout L = stage_layout(d, i8, trace ? iterations * syn::code_n(d->code) : 0);
    DeviceScope ds(d->device);
    syn::Stage *st = nullptr;
    if (int32_t rc = take_stage(d, L.total, &st)) return rc;
    const int32_t rc = host_decode_at(d, (char *)st->d, st->stream, i8, llr, edge_r, hard, iterations, n_filler,
                                      init5g, scale, app_clamp, out_used, trace);
    give_stage(d, st);
    return rc;
}

// ---------------------------------------------------------------------------
// The allocation-free decode_5g: staging in a file-scope device array
// ---------------------------------------------------------------------------

// Staging of syndrome_ldpc_decode_5g_static: one array per GPU, held until
// the process ends.  In host code `scratch` names a host-side stand-in, not
// the array's device address: host code only ever uses the address
// hipGetSymbolAddress returns for the current device (scratch_on).
static __device__ __align__(256) unsigned char scratch[SYNDROME_STATIC_SCRATCH_BYTES];

constexpr int SCRATCH_DEVICES = 64;
struct ScratchSlot {
    std::once_flag once;
    void *ptr = nullptr;
    int32_t status = SYNDROME_OK;
    std::mutex mu;  // one call at a time on this GPU's array
};
ScratchSlot g_scratch[SCRATCH_DEVICES];

// This device's copy of `scratch`, looked up once per device with that
// device current (each GPU holds its own copy at its own address).
ScratchSlot *scratch_on(int device)
{
    if (device < 0 || device >= SCRATCH_DEVICES) return nullptr;
    ScratchSlot &sl = g_scratch[device];
    std::call_once(sl.once, [&] {
        DeviceScope ds(device);
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            sl.status = SYNDROME_DEVICE_ERROR;
            return;
        }
        sl.ptr = p;
    });
    return &sl;
}

// init_5g_llr on host memory (src/qc_ldpc.rs:711-732, 1505-1514): applied
// when a later validation step fails, so the caller's buffer is left as the
// reference leaves it (the prologue runs before the decode's own checks),
// and on a code the prologue is not defined on.  Returns false where the
// reference panics: n_b < m_b (the filler range is out of bounds, nothing
// written) or N < 2Z (the fillers are written, then the punctured range is
// out of bounds).
template <class T>
bool init_5g_host(const SynCode &c, T *buf, size_t n_filler, T filler)
{
    if (c.n_b < c.m_b) return false;
    const size_t k = syn::code_k(c);
    const size_t kp = n_filler >= k ? 0 : k - n_filler;
    for (size_t i = kp; i < k; ++i) buf[i] = filler;
    if (syn::code_n(c) < 2 * (size_t)c.z) return false;
    for (size_t i = 0; i < (size_t)(2 * c.z); ++i) buf[i] = 0;
    return true;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Part 1
// ---------------------------------------------------------------------------

int32_t syndrome_ldpc_decoder_create(int32_t base_graph, size_t z, float offset_beta, SyndromeLdpcDecoder **out)
{
    return guard([&]() -> int32_t {
        if (!out) return SYNDROME_NULL_POINTER;
        *out = nullptr;
        if (base_graph != 0 && base_graph != 1) return SYNDROME_INVALID_PARAM;
        SynCode c;
        if (syn::build_code_3gpp(base_graph, z > (size_t)INT_MAX ? -1 : (long long)z, &c)) return SYNDROME_INVALID_PARAM;
        return make_decoder(c, offset_beta, out);
    });
}

void syndrome_ldpc_decoder_destroy(SyndromeLdpcDecoder *d)
{
    if (!d) return;
    try {
        free_decoder(d);
    } catch (...) {
    }
    delete d;
}

size_t syndrome_ldpc_decoder_variable_node_count(const SyndromeLdpcDecoder *d) { return syn::code_n(d->code); }
size_t syndrome_ldpc_decoder_info_bit_count(const SyndromeLdpcDecoder *d) { return syn::code_k(d->code); }
size_t syndrome_ldpc_decoder_required_edge_buffer(const SyndromeLdpcDecoder *d) { return syn::code_ez(d->code); }
size_t syndrome_ldpc_decoder_required_layer_buffer(const SyndromeLdpcDecoder *d)
{
    return syn::code_layer_buf(d->code);
}

int32_t syndrome_ldpc_decode_5g(SyndromeLdpcDecoder *d, float *llr, size_t llr_len, size_t n_filler, float *edge_r,
                                size_t edge_r_len, float *layer_scratch, size_t layer_scratch_len, uint8_t *hard,
                                size_t hard_len, size_t iterations, size_t *out_used)
{
    return guard([&]() -> int32_t {
        if (!d || !llr || !edge_r || !layer_scratch || !hard) return SYNDROME_NULL_POINTER;
        const SynCode &c = d->code;
        const size_t n = syn::code_n(c);
        if (llr_len != n) return SYNDROME_INVALID_PARAM;  // init_5g_llr
        if (!syn::code_5g_ok(c)) {
            (void)init_5g_host<float>(c, llr, n_filler, 1000000.0f);
            return SYNDROME_PANIC;
        }
        int32_t rc = SYNDROME_OK;
        if (hard_len != n) rc = SYNDROME_INVALID_PARAM;
        else if (edge_r_len < syn::code_ez(c)) rc = SYNDROME_BUFFER_TOO_SMALL;
        else if (layer_scratch_len < syn::code_layer_buf(c)) rc = SYNDROME_BUFFER_TOO_SMALL;
        if (rc) {
            init_5g_host<float>(c, llr, n_filler, 1000000.0f);
            return rc;
        }
        return host_decode(d, false, llr, edge_r, edge_r_len, hard, iterations, n_filler, true, 0.f, 0, out_used);
    });
}

int32_t syndrome_ldpc_decode_5g_static(SyndromeLdpcDecoder *d, float *llr, size_t llr_len, size_t n_filler,
                                       float *edge_r, size_t edge_r_len, float *layer_scratch,
                                       size_t layer_scratch_len, uint8_t *hard, size_t hard_len, size_t iterations,
                                       size_t *out_used, void *stream)
{
    return guard([&]() -> int32_t {
        if (!d || !llr || !edge_r || !layer_scratch || !hard) return SYNDROME_NULL_POINTER;
        const SynCode &c = d->code;
        const size_t n = syn::code_n(c);
        if (llr_len != n) return SYNDROME_INVALID_PARAM;  // init_5g_llr
        if (!syn::code_5g_ok(c)) {
            (void)init_5g_host<float>(c, llr, n_filler, 1000000.0f);
            return SYNDROME_PANIC;
        }
        int32_t rc = SYNDROME_OK;
        if (hard_len != n) rc = SYNDROME_INVALID_PARAM;
        else if (edge_r_len < syn::code_ez(c)) rc = SYNDROME_BUFFER_TOO_SMALL;
        else if (layer_scratch_len < syn::code_layer_buf(c)) rc = SYNDROME_BUFFER_TOO_SMALL;
        if (rc) {
            init_5g_host<float>(c, llr, n_filler, 1000000.0f);
            return rc;
        }
        // the codeword's staging must fit the array
        if (stage_layout(d, false, 0).total > SYNDROME_STATIC_SCRATCH_BYTES) return SYNDROME_INVALID_PARAM;
        ScratchSlot *sl = scratch_on(d->device);
        if (!sl) return SYNDROME_DEVICE_ERROR;
        if (sl->status) return sl->status;
        std::lock_guard<std::mutex> lk(sl->mu);
        DeviceScope ds(d->device);
        return host_decode_at(d, (char *)sl->ptr, (hipStream_t)stream, false, llr, edge_r, hard, iterations, n_filler,
                              true, 0.f, 0, out_used, nullptr);
    });
}

int32_t syndrome_ldpc_encoder_create(int32_t base_graph, size_t z, SyndromeLdpcEncoder **out)
{
    return guard([&]() -> int32_t {
        if (!out) return SYNDROME_NULL_POINTER;
        *out = nullptr;
        if (base_graph != 0 && base_graph != 1) return SYNDROME_INVALID_PARAM;
        std::unique_ptr<SyndromeLdpcEncoder> e(new SyndromeLdpcEncoder());
        if (syn::build_code_3gpp(base_graph, z > (size_t)INT_MAX ? -1 : (long long)z, &e->code))
            return SYNDROME_INVALID_PARAM;
        e->enc.reset(new syn::Encoder(e->code));
        if (!e->enc->ok()) return SYNDROME_INVALID_PARAM;
        *out = e.release();
        return SYNDROME_OK;
    });
}

void syndrome_ldpc_encoder_destroy(SyndromeLdpcEncoder *e) { delete e; }
size_t syndrome_ldpc_encoder_info_bit_count(const SyndromeLdpcEncoder *e) { return syn::code_k(e->code); }
size_t syndrome_ldpc_encoder_codeword_bit_count(const SyndromeLdpcEncoder *e) { return syn::code_n(e->code); }

int32_t syndrome_ldpc_encode_5g(const SyndromeLdpcEncoder *e, const uint8_t *info, size_t info_len, size_t n_filler,
                                uint8_t *codeword, size_t codeword_len)
{
    return guard([&]() -> int32_t {
        if (!e || !info || !codeword) return SYNDROME_NULL_POINTER;
        // QcLdpcEncoder::encode_5g / encode (qc_ldpc.rs:2547-2606)
        const size_t k = syn::code_k(e->code), n = syn::code_n(e->code);
        if (n_filler > k) return SYNDROME_INVALID_PARAM;
        if (info_len != k - n_filler) return SYNDROME_BUFFER_TOO_SMALL;
        if (codeword_len != n) return SYNDROME_BUFFER_TOO_SMALL;
        std::vector<uint8_t> full(k, 0);
        memcpy(full.data(), info, info_len);
        e->enc->encode(full.data(), codeword);
        return SYNDROME_OK;
    });
}

// ---------------------------------------------------------------------------
// Part 2
// ---------------------------------------------------------------------------

int32_t syndrome_ldpc_decoder_create_raw(const uint8_t *rows, const uint8_t *cols, const int16_t *shifts, size_t n,
                                         size_t m_b, size_t n_b, size_t z, float offset_beta,
                                         SyndromeLdpcDecoder **out)
{
    return guard([&]() -> int32_t {
        if (!out) return SYNDROME_NULL_POINTER;
        *out = nullptr;
        if (n > 0 && (!rows || !cols || !shifts)) return SYNDROME_NULL_POINTER;
        if (z > SYN_Z_MAX || m_b > (size_t)INT_MAX || n_b > (size_t)INT_MAX) return SYNDROME_INVALID_PARAM;
        std::vector<int> r(n), c(n), s(n);
        for (size_t i = 0; i < n; ++i) {
            r[i] = rows[i];
            c[i] = cols[i];
            s[i] = shifts[i];
        }
        SynCode code;
        std::unique_ptr<syn::WideCode> wide(new syn::WideCode());
        if (syn::build_code_raw(r.data(), c.data(), s.data(), n, (long long)m_b, (long long)n_b, (long long)z, &code,
                                wide.get()))
            return SYNDROME_INVALID_PARAM;
        if (code.kind != SYN_KIND_WIDE) wide.reset();
        return make_decoder(code, offset_beta, out, std::move(wide));
    });
}

int32_t syndrome_ldpc_decoder_create_wifi(size_t z, size_t rate_num, size_t rate_den, float offset_beta,
                                          SyndromeLdpcDecoder **out)
{
    return guard([&]() -> int32_t {
        if (!out) return SYNDROME_NULL_POINTER;
        *out = nullptr;
        SynCode code;
        if (z > SYN_Z_MAX || rate_num > 16 || rate_den > 16 ||
            syn::build_code_wifi((long long)z, (long long)rate_num, (long long)rate_den, &code))
            return SYNDROME_INVALID_PARAM;
        return make_decoder(code, offset_beta, out);
    });
}

size_t syndrome_ldpc_decoder_check_node_count(const SyndromeLdpcDecoder *d) { return (size_t)d->code.m_b * d->code.z; }
float syndrome_ldpc_decoder_offset_beta(const SyndromeLdpcDecoder *d) { return d->beta; }

int32_t syndrome_ldpc_decode(SyndromeLdpcDecoder *d, float *llr, size_t llr_len, float *edge_r, size_t edge_r_len,
                             float *layer_scratch, size_t layer_scratch_len, uint8_t *hard, size_t hard_len,
                             size_t iterations, size_t *out_used)
{
    return guard([&]() -> int32_t {
        if (!d || !llr || !edge_r || !layer_scratch || !hard) return SYNDROME_NULL_POINTER;
        const SynCode &c = d->code;
        const size_t n = syn::code_n(c);
        if (llr_len != n || hard_len != n) return SYNDROME_INVALID_PARAM;  // qc_ldpc.rs:980-1004
        if (edge_r_len < syn::code_ez(c) || layer_scratch_len < syn::code_layer_buf(c)) return SYNDROME_BUFFER_TOO_SMALL;
        return host_decode(d, false, llr, edge_r, edge_r_len, hard, iterations, 0, false, 0.f, 0, out_used);
    });
}

int32_t syndrome_ldpc_decode_traced(SyndromeLdpcDecoder *d, float *llr, size_t llr_len, float *edge_r,
                                    size_t edge_r_len, float *layer_scratch, size_t layer_scratch_len, uint8_t *hard,
                                    size_t hard_len, size_t iterations, float *trace, size_t trace_len,
                                    size_t *out_used)
{
    return guard([&]() -> int32_t {
        if (!d || !llr || !edge_r || !layer_scratch || !hard || (!trace && iterations)) return SYNDROME_NULL_POINTER;
        const SynCode &c = d->code;
        const size_t n = syn::code_n(c);
        if (llr_len != n || hard_len != n) return SYNDROME_INVALID_PARAM;  // qc_ldpc.rs:980-1004
        if (edge_r_len < syn::code_ez(c) || layer_scratch_len < syn::code_layer_buf(c)) return SYNDROME_BUFFER_TOO_SMALL;
        if (iterations > (size_t)INT_MAX || (iterations && trace_len / iterations < n))
            return SYNDROME_BUFFER_TOO_SMALL;
        return host_decode(d, false, llr, edge_r, edge_r_len, hard, iterations, 0, false, 0.f, 0, out_used,
                           iterations ? trace : nullptr);
    });
}

int32_t syndrome_ldpc_decode_i8(SyndromeLdpcDecoder *d, int16_t *app, size_t app_len, int8_t *edge_r, size_t edge_r_len,
                                int8_t *layer_scratch, size_t layer_scratch_len, uint8_t *hard, size_t hard_len,
                                size_t iterations, float scale, int16_t app_clamp, size_t *out_used)
{
    return guard([&]() -> int32_t {
        if (!d || !app || !edge_r || !layer_scratch || !hard) return SYNDROME_NULL_POINTER;
        const SynCode &c = d->code;
        const size_t n = syn::code_n(c);
        if (app_len != n || hard_len != n) return SYNDROME_INVALID_PARAM;  // qc_ldpc.rs:1544-1568
        if (edge_r_len < syn::code_ez(c) || layer_scratch_len < syn::code_layer_buf(c)) return SYNDROME_BUFFER_TOO_SMALL;
        return host_decode(d, true, app, edge_r, edge_r_len, hard, iterations, 0, false, scale, app_clamp, out_used);
    });
}

int32_t syndrome_ldpc_decode_5g_i8(SyndromeLdpcDecoder *d, int16_t *app, size_t app_len, size_t n_filler, int8_t *edge_r,
                                   size_t edge_r_len, int8_t *layer_scratch, size_t layer_scratch_len, uint8_t *hard,
                                   size_t hard_len, size_t iterations, float scale, int16_t app_clamp,
                                   size_t *out_used)
{
    return guard([&]() -> int32_t {
        if (!d || !app || !edge_r || !layer_scratch || !hard) return SYNDROME_NULL_POINTER;
        const SynCode &c = d->code;
        const size_t n = syn::code_n(c);
        if (app_len != n) return SYNDROME_INVALID_PARAM;  // qc_ldpc.rs:1499-1503
        if (!syn::code_5g_ok(c)) {
            (void)init_5g_host<int16_t>(c, app, n_filler, app_clamp < 0 ? (int16_t)0 : app_clamp);
            return SYNDROME_PANIC;
        }
        int32_t rc = SYNDROME_OK;
        if (hard_len != n) rc = SYNDROME_INVALID_PARAM;
        else if (edge_r_len < syn::code_ez(c)) rc = SYNDROME_BUFFER_TOO_SMALL;
        else if (layer_scratch_len < syn::code_layer_buf(c)) rc = SYNDROME_BUFFER_TOO_SMALL;
        if (rc) {
            init_5g_host<int16_t>(c, app, n_filler, app_clamp < 0 ? (int16_t)0 : app_clamp);
            return rc;
        }
        return host_decode(d, true, app, edge_r, edge_r_len, hard, iterations, n_filler, true, scale, app_clamp,
                           out_used);
    });
}

int32_t syndrome_ldpc_decode_batch_f32(const SyndromeLdpcDecoder *d, const float *in, float *post, float *edge,
                                       uint8_t *hard, int32_t *iters, size_t batch, size_t iterations, size_t n_filler,
                                       uint32_t flags, void *stream)
{
    return guard([&]() -> int32_t {
        if (!d) return SYNDROME_NULL_POINTER;
        if (flags & SYNDROME_BATCH_QUANTIZE) return SYNDROME_INVALID_PARAM;
        return run_batch(d, false, in, post, edge, hard, iters, batch, iterations, n_filler, 0.f, 0, flags,
                         (hipStream_t)stream);
    });
}

int32_t syndrome_ldpc_decode_batch_i8(const SyndromeLdpcDecoder *d, const void *in, int16_t *post, int8_t *edge,
                                      uint8_t *hard, int32_t *iters, size_t batch, size_t iterations, size_t n_filler,
                                      float scale, int16_t app_clamp, uint32_t flags, void *stream)
{
    return guard([&]() -> int32_t {
        if (!d) return SYNDROME_NULL_POINTER;
        return run_batch(d, true, in, post, edge, hard, iters, batch, iterations, n_filler, scale, app_clamp, flags,
                         (hipStream_t)stream);
    });
}

int32_t syndrome_ldpc_launch_shape(const SyndromeLdpcDecoder *d, int32_t int8, int32_t plain, int32_t *cws,
                                   int32_t *threads, int32_t *lds)
{
    return guard([&]() -> int32_t {
        if (!d || !cws || !threads || !lds) return SYNDROME_NULL_POINTER;
        Shape s;
        if (int32_t rc = launch_shape(d, int8 != 0, plain != 0, &s)) return rc;
        *cws = s.cws;
        *threads = s.block;
        *lds = (int32_t)s.lds;
        return SYNDROME_OK;
    });
}

int32_t syndrome_ldpc_quantize_i16_device(const float *d_llr, int16_t *d_out, size_t n, float scale, void *stream)
{
    return guard([&]() -> int32_t {
        if ((!d_llr || !d_out) && n) return SYNDROME_NULL_POINTER;
        return syn::launch_quantize(d_llr, d_out, n, scale, (hipStream_t)stream);
    });
}

int32_t syndrome_ldpc_quantize_i8_device(const float *d_llr, int8_t *d_out, size_t n, float scale, void *stream)
{
    return guard([&]() -> int32_t {
        if ((!d_llr || !d_out) && n) return SYNDROME_NULL_POINTER;
        return syn::launch_quantize_i8(d_llr, d_out, n, scale, (hipStream_t)stream);
    });
}

int32_t syndrome_ldpc_dequantize_i8_device(const int8_t *d_q, float *d_out, size_t n, float scale, void *stream)
{
    return guard([&]() -> int32_t {
        if ((!d_q || !d_out) && n) return SYNDROME_NULL_POINTER;
        return syn::launch_dequantize_i8(d_q, d_out, n, scale, (hipStream_t)stream);
    });
}

int32_t syndrome_awgn_transmit(float ebno_db, float code_rate, uint64_t seed, const uint8_t *bits, size_t n,
                               float *llr_out)
{
    return guard([&]() -> int32_t {
        if ((!bits || !llr_out) && n) return SYNDROME_NULL_POINTER;
        syn::awgn_transmit(ebno_db, code_rate, seed, bits, n, llr_out);
        return SYNDROME_OK;
    });
}

int32_t syndrome_ldpc_encode_batch(const SyndromeLdpcDecoder *d, const uint8_t *info, uint8_t *cws, size_t count,
                                   int32_t threads)
{
    return guard([&]() -> int32_t {
        if (!d || ((!info || !cws) && count)) return SYNDROME_NULL_POINTER;
        SyndromeLdpcDecoder *md = const_cast<SyndromeLdpcDecoder *>(d);
        if (d->code.kind == SYN_KIND_WIDE) return SYNDROME_INVALID_PARAM;  // no encoder for the wide form
        std::call_once(md->enc_once, [&] { md->enc.reset(new syn::Encoder(md->code)); });
        if (!md->enc || !md->enc->ok()) return SYNDROME_INVALID_PARAM;
        const size_t k = syn::code_k(d->code), n = syn::code_n(d->code);
        int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
        nt = (int)std::min<size_t>((size_t)nt, std::max<size_t>(count, 1));
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                for (size_t i = (size_t)t; i < count; i += (size_t)nt) md->enc->encode(info + i * k, cws + i * n);
            });
        for (auto &th : pool) th.join();
        return SYNDROME_OK;
    });
}

int32_t syndrome_ldpc_encode_batch_device(const SyndromeLdpcDecoder *d, const uint8_t *d_info, uint8_t *d_codewords,
                                          size_t count, void *stream)
{
    return guard([&]() -> int32_t {
        if (!d || ((!d_info || !d_codewords) && count)) return SYNDROME_NULL_POINTER;
        if (d->code.kind == SYN_KIND_WIDE) return SYNDROME_INVALID_PARAM;  // no encoder for the wide form
        if (count > (size_t)LLONG_MAX / syn::code_n(d->code)) return SYNDROME_INVALID_PARAM;
        SyndromeLdpcDecoder *md = const_cast<SyndromeLdpcDecoder *>(d);
        DeviceScope ds(d->device);
        std::call_once(md->denc_once, [&] {
            std::call_once(md->enc_once, [&] { md->enc.reset(new syn::Encoder(md->code)); });
            std::vector<int> off, blk, sh;
            if (!md->enc || !md->enc->core_terms(off, blk, sh) || blk.size() > SYN_ENC_MAX_TERMS) {
                md->denc_status = SYNDROME_INVALID_PARAM;
                return;
            }
            std::unique_ptr<SynEncCode> h(new SynEncCode());
            h->core = md->enc->core();
            h->n_terms = (int)blk.size();
            for (size_t i = 0; i < off.size(); ++i) h->term_off[i] = off[i];
            for (size_t t = 0; t < blk.size(); ++t) {
                h->term_blk[t] = (short)blk[t];
                h->term_shift[t] = (short)sh[t];
            }
            SynEncCode *p = nullptr;
            if (hipMalloc((void **)&p, sizeof(SynEncCode)) != hipSuccess) {
                md->denc_status = SYNDROME_DEVICE_ERROR;
                return;
            }
            if (hipMemcpy(p, h.get(), sizeof(SynEncCode), hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipFree(p);
                md->denc_status = SYNDROME_DEVICE_ERROR;
                return;
            }
            md->d_enc = p;
            md->h_enc = std::move(h);
        });
        if (md->denc_status != SYNDROME_OK) return md->denc_status;
        if (!count) return SYNDROME_OK;
        const SynCode &c = d->code;
        const size_t lds = syn::encode_lds_bytes(c, *md->h_enc, d_info, d_codewords);
        return syn::launch_encode(d->d_code, md->d_enc, c.z, (int)syn::code_k(c), (int)syn::code_n(c), lds, d_info,
                                  d_codewords, (long long)count, (hipStream_t)stream)
                   ? SYNDROME_DEVICE_ERROR
                   : SYNDROME_OK;
    });
}

namespace {

// Encode -> rate match (rv 0, Qm 1) -> AWGN for every block; info bits
// either drawn from the seed (written to info when given) or read from
// info_in.
int32_t traffic(const int32_t *bg, const int32_t *z, const int32_t *n_filler, const int32_t *e_len,
                const float *ebno_db, const float *rate, uint64_t seed, size_t count, float *e_llr,
                const int64_t *e_off, uint8_t *info, const uint8_t *info_in, const int64_t *info_off, int32_t threads)
{
    return guard([&]() -> int32_t {
        if (count && (!bg || !z || !e_len || !ebno_db || !rate || !e_llr || !e_off)) return SYNDROME_NULL_POINTER;
        if ((info || info_in) && !info_off) return SYNDROME_NULL_POINTER;
        // one encoder per distinct code, built up front
        std::map<std::pair<int, int>, std::unique_ptr<syn::Encoder>> enc;
        std::map<std::pair<int, int>, SynCode> codes;
        for (size_t i = 0; i < count; ++i) {
            auto key = std::make_pair((int)bg[i], (int)z[i]);
            if (codes.count(key)) continue;
            SynCode c;
            if (syn::build_code_3gpp(bg[i], z[i], &c)) return SYNDROME_INVALID_PARAM;
            codes[key] = c;
            enc[key].reset(new syn::Encoder(c));
            if (!enc[key]->ok()) return SYNDROME_INVALID_PARAM;
        }
        for (size_t i = 0; i < count; ++i) {
            const size_t k = syn::code_k(codes[std::make_pair((int)bg[i], (int)z[i])]);
            if (e_len[i] < 0 || (n_filler && (n_filler[i] < 0 || (size_t)n_filler[i] > k))) return SYNDROME_INVALID_PARAM;
        }
        int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
        nt = (int)std::min<size_t>((size_t)nt, std::max<size_t>(count, 1));
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                std::vector<uint8_t> inf, cw, e;
                for (size_t i = (size_t)t; i < count; i += (size_t)nt) {
                    auto key = std::make_pair((int)bg[i], (int)z[i]);
                    const SynCode &c = codes[key];
                    const size_t k = syn::code_k(c), n = syn::code_n(c);
                    const size_t nf = n_filler ? (size_t)n_filler[i] : 0;
                    inf.assign(k, 0);
                    if (info_in) {
                        for (size_t b = 0; b < k - nf; ++b) inf[b] = info_in[info_off[i] + (int64_t)b] & 1u;
                    } else {
                        uint64_t s = (seed + 0x9E3779B97F4A7C15ULL * (i + 1)) | 1;
                        for (size_t b = 0; b < k - nf; ++b) {  // xorshift64, a middle bit per info bit
                            s ^= s << 13;
                            s ^= s >> 7;
                            s ^= s << 17;
                            inf[b] = (uint8_t)((s >> 27) & 1);
                        }
                        if (info) memcpy(info + info_off[i], inf.data(), k - nf);
                    }
                    cw.resize(n);
                    enc[key]->encode(inf.data(), cw.data());
                    e.resize((size_t)e_len[i]);
                    syn::rate_match(cw.data(), n, e.data(), e.size(), 0, 1, bg[i], z[i], nf);
                    syn::awgn_transmit(ebno_db[i], rate[i], seed ^ (0xD1B54A32D192ED03ULL * (i + 1)), e.data(),
                                       e.size(), e_llr + e_off[i]);
                }
            });
        for (auto &th : pool) th.join();
        return SYNDROME_OK;
    });
}

}  // namespace

int32_t syndrome_traffic_blocks(const int32_t *bg, const int32_t *z, const int32_t *n_filler, const int32_t *e_len,
                                const float *ebno_db, const float *rate, uint64_t seed, size_t count, float *e_llr,
                                const int64_t *e_off, uint8_t *info, const int64_t *info_off, int32_t threads)
{
    return traffic(bg, z, n_filler, e_len, ebno_db, rate, seed, count, e_llr, e_off, info, nullptr, info_off, threads);
}

int32_t syndrome_traffic_blocks_info(const int32_t *bg, const int32_t *z, const int32_t *n_filler,
                                     const int32_t *e_len, const float *ebno_db, const float *rate, uint64_t seed,
                                     size_t count, const uint8_t *info, const int64_t *info_off, float *e_llr,
                                     const int64_t *e_off, int32_t threads)
{
    if (count && !info) return SYNDROME_NULL_POINTER;
    return traffic(bg, z, n_filler, e_len, ebno_db, rate, seed, count, e_llr, e_off, nullptr, info, info_off, threads);
}

}  // extern "C"
