// Plain (runtime-structured) decode kernels, the quantizer, and the launch
// dispatch for every decode kernel.
#include "ldpc_device.h"

#include <mutex>
#include <set>
#include <utility>

namespace syn {

// ---------------------------------------------------------------------------
// plain kernels: runtime structure, R uncompressed in HBM
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(1024) plain_f32(const SynCode *__restrict__ code, SynLaunch a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int Z = code->z, m_b = code->m_b, n_b = code->n_b, N = n_b * Z;
    const int EZ = code->n_edges * Z, cws = a.cws, cw0 = blockIdx.x * cws;
    float *P0 = (float *)smem;
    CwFlags fl = cw_flags(P0 + cws * N, cws);
    const int t = threadIdx.x;
    const GroupIO io = contiguous_group(cw0, cws, a.batch);
    load_input<false>(a, Z, n_b, m_b, N, cws, io, P0);
    float *edge = (float *)a.edge;
    for (int i = t; i < cws * EZ; i += blockDim.x)
        if (cw0 + i / EZ < a.batch) edge[(size_t)cw0 * EZ + i] = 0.f;  // qc_ldpc.rs:1006-1009
    flags_init(fl, cws, io);
    __syncthreads();

    const bool lane_ok = t < cws * Z;
    const int cl = lane_ok ? t / Z : 0;
    const int z = lane_ok ? t - cl * Z : 0;
    float *P = P0 + cl * N;
    float *R = edge + (size_t)(cw0 + cl) * EZ;
    const float beta = a.beta;

    for (int it = 0; it < a.iterations; ++it) {
        if (all_done(fl, cws)) break;
        const bool act = lane_ok && !fl.done[cl];
        if (t < cws && !fl.done[t]) fl.used[t] = it + 1;
        for (int l = 0; l < m_b; ++l) {
            const int lb = code->layer_off[l], le = code->layer_off[l + 1];
            if (act) {
                // Q-build + pass 1 (qc_ldpc.rs:1072-1094, 1155-1178)
                float m1 = FLT_MAX, m2 = FLT_MAX;
                uint32_t sx = 0;
                for (int e = lb; e < le; ++e) {
                    const int v = code->cols[e] * Z + wrap_add(z, code->shifts[e], Z);
                    const float q = P[v] - R[(size_t)e * Z + z];
                    const uint32_t qb = __float_as_uint(q);
                    const float aa = __uint_as_float(qb & 0x7fffffffu);
                    sx ^= qb & 0x80000000u;
                    if (aa <= m1) { m2 = m1; m1 = aa; }
                    else if (aa < m2) m2 = aa;
                }
                if (m1 == FLT_MAX) m1 = 0.f;
                if (m2 == FLT_MAX) m2 = m1;
                // pass 2 (qc_ldpc.rs:1181-1204); q is rebuilt from the same
                // posterior and message, so it is the same value.
                for (int e = lb; e < le; ++e) {
                    const int v = code->cols[e] * Z + wrap_add(z, code->shifts[e], Z);
                    const float Lv = P[v];
                    const float ro = R[(size_t)e * Z + z];
                    const uint32_t qb = __float_as_uint(Lv - ro);
                    const float aa = __uint_as_float(qb & 0x7fffffffu);
                    const uint32_t sg = sx ^ (qb & 0x80000000u);
                    const float me = (aa == m1) ? m2 : m1;
                    float mag = me - beta;
                    mag = (mag < 0.f) ? 0.f : mag;
                    const float rn = __uint_as_float(__float_as_uint(mag) | sg);
                    R[(size_t)e * Z + z] = rn;
                    P[v] = Lv + (rn - ro);
                }
            }
            __syncthreads();
        }
        trace_store(a, P, N, Z, z, act, io, cl, it);
        // end-of-iteration syndrome (check_syndrome_f32, qc_ldpc.rs:1231-1252)
        if (act) {
            int fail = 0;
            for (int l = 0; l < m_b && !fail; ++l) {
                unsigned p = 0;
                for (int e = code->layer_off[l]; e < code->layer_off[l + 1]; ++e)
                    p ^= P[code->cols[e] * Z + wrap_add(z, code->shifts[e], Z)] < 0.f ? 1u : 0u;
                fail = (int)p;
            }
            if (fail) fl.fail[cl] = 1;
        }
        __syncthreads();
        flags_settle(fl, cws);
        __syncthreads();
    }
    store_output<false>(a, N, io, P0);
    store_iters(a, fl, io);
}

__global__ void __launch_bounds__(1024) plain_i8(const SynCode *__restrict__ code, SynLaunch a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int Z = code->z, m_b = code->m_b, n_b = code->n_b, N = n_b * Z;
    const int EZ = code->n_edges * Z, cws = a.cws, cw0 = blockIdx.x * cws;
    int16_t *P0 = (int16_t *)smem;
    CwFlags fl = cw_flags(P0 + ((cws * N + 1) & ~1), cws);
    const int t = threadIdx.x;
    const GroupIO io = contiguous_group(cw0, cws, a.batch);
    load_input<true>(a, Z, n_b, m_b, N, cws, io, P0);
    int8_t *edge = (int8_t *)a.edge;
    for (int i = t; i < cws * EZ; i += blockDim.x)
        if (cw0 + i / EZ < a.batch) edge[(size_t)cw0 * EZ + i] = 0;  // qc_ldpc.rs:1579-1581
    flags_init(fl, cws, io);
    __syncthreads();

    const bool lane_ok = t < cws * Z;
    const int cl = lane_ok ? t / Z : 0;
    const int z = lane_ok ? t - cl * Z : 0;
    int16_t *P = P0 + cl * N;
    int8_t *R = edge + (size_t)(cw0 + cl) * EZ;
    const int bq = a.beta_q, C = a.app_clamp;

    for (int it = 0; it < a.iterations; ++it) {
        if (all_done(fl, cws)) break;
        const bool act = lane_ok && !fl.done[cl];
        if (t < cws && !fl.done[t]) fl.used[t] = it + 1;
        for (int l = 0; l < m_b; ++l) {
            const int lb = code->layer_off[l], le = code->layer_off[l + 1];
            if (act) {
                // Q-build + pass 1 (qc_ldpc.rs:1619-1690)
                int m1 = 127, m2 = 127;
                unsigned sx = 0;
                for (int e = lb; e < le; ++e) {
                    const int v = code->cols[e] * Z + wrap_add(z, code->shifts[e], Z);
                    int q = (int)P[v] - (int)R[(size_t)e * Z + z];
                    q = q < -127 ? -127 : (q > 127 ? 127 : q);
                    const int aa = q < 0 ? -q : q;
                    sx ^= q < 0 ? 1u : 0u;
                    const int mx = m1 > aa ? m1 : aa;
                    m2 = m2 < mx ? m2 : mx;
                    m1 = m1 < aa ? m1 : aa;
                }
                if (le - lb < 2) {  // qc_ldpc.rs:1692-1706
                    if (le == lb) m1 = 0;
                    m2 = m1;
                }
                // pass 2 (qc_ldpc.rs:1708-1744)
                for (int e = lb; e < le; ++e) {
                    const int v = code->cols[e] * Z + wrap_add(z, code->shifts[e], Z);
                    const int av = P[v];
                    const int ro = R[(size_t)e * Z + z];
                    int q = av - ro;
                    q = q < -127 ? -127 : (q > 127 ? 127 : q);
                    const int aa = q < 0 ? -q : q;
                    const unsigned sg = sx ^ (q < 0 ? 1u : 0u);
                    const int me = (aa == m1) ? m2 : m1;
                    int mag = me - bq;
                    mag = mag < 0 ? 0 : mag;
                    const int rn = sg ? -mag : mag;
                    R[(size_t)e * Z + z] = (int8_t)rn;
                    int nv = av + (rn - ro);  // saturating i16 add then clamp == clamp
                    nv = nv < -C ? -C : (nv > C ? C : nv);
                    P[v] = (int16_t)nv;
                }
            }
            __syncthreads();
        }
        if (act) {  // check_syndrome_i16 (qc_ldpc.rs:1766-1787)
            int fail = 0;
            for (int l = 0; l < m_b && !fail; ++l) {
                unsigned p = 0;
                for (int e = code->layer_off[l]; e < code->layer_off[l + 1]; ++e)
                    p ^= P[code->cols[e] * Z + wrap_add(z, code->shifts[e], Z)] < 0 ? 1u : 0u;
                fail = (int)p;
            }
            if (fail) fl.fail[cl] = 1;
        }
        __syncthreads();
        flags_settle(fl, cws);
        __syncthreads();
    }
    store_output<true>(a, N, io, P0);
    store_iters(a, fl, io);
}

// quantize_llr_i16 (quantize.rs:331-341) and quantize_llr (quantize.rs:148-154)
// as standalone passes: v = l * scale; NaN -> 0 (the i16 form's test, and
// Rust's saturating `as i32` for the i8 form); round half away from zero;
// clamp to +-127 (which also takes +-inf to +-127).
template <class T>
__global__ void __launch_bounds__(256) quantize_pass(const float *__restrict__ in, T *__restrict__ out, size_t n,
                                                     float scale)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float f = in[i] * scale;
        int v = 0;
        if (f == f) {
            f = roundf(f);
            f = f < -127.f ? -127.f : (f > 127.f ? 127.f : f);
            v = (int)f;
        }
        out[i] = (T)v;
    }
}

// dequantize_llr (quantize.rs:163-169): q * (1 / scale), the reciprocal
// computed once on the host, as the reference computes it once per call.
__global__ void __launch_bounds__(256) dequantize_i8(const int8_t *__restrict__ in, float *__restrict__ out, size_t n,
                                                     float inv)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i] * inv;
}

const void *fast_kernel_f32_bg1(int z);
const void *fast_kernel_f32_bg2(int z);
const void *fast_kernel_i8_bg1(int z);
const void *fast_kernel_wifi(int kind, bool i8);
const void *pack_kernel_i8_bg2(int z);
const void *fastx_kernel_f32_bg1(int z);
int fastx_shape_f32_bg1(int z, int *lanes, int *cws);

// The kernel a decode of (kind, format, Z) runs: plain, or the
// register-resident one for the base graph (its Z-specialised instance when
// one was compiled for this Z).
const void *decode_kernel(int kind, bool i8, bool fast, int z)
{
    if (!fast) return i8 ? (const void *)plain_i8 : (const void *)plain_f32;
    if (kind == SYN_KIND_BG1) return i8 ? fast_kernel_i8_bg1(z) : fast_kernel_f32_bg1(z);
    if (kind == SYN_KIND_BG2) return i8 ? pack_kernel_i8_bg2(z) : fast_kernel_f32_bg2(z);
    if (kind >= SYN_KIND_WIFI0) return fast_kernel_wifi(kind, i8);
    return nullptr;
}

// Dynamic LDS is a per-function attribute, global to the process.  Raising
// it once, per function and device, to everything the device allows besides
// the function's static LDS means that no launcher ever has to lower it, so
// concurrent launchers with different footprints cannot undercut each other.
int allow_max_dynamic_lds(const void *fn)
{
    static std::mutex mu;
    static std::set<std::pair<const void *, int>> raised;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SYN_DEVICE_ERROR;
    std::lock_guard<std::mutex> lk(mu);
    if (raised.count({fn, dev})) return SYN_OK;
    int limit = 0;
    hipFuncAttributes fa{};
    if (hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        hipFuncGetAttributes(&fa, fn) != hipSuccess)
        return SYN_DEVICE_ERROR;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, limit - (int)fa.sharedSizeBytes) !=
        hipSuccess)
        return SYN_DEVICE_ERROR;
    raised.insert({fn, dev});
    return SYN_OK;
}

int launch_decode(int kind, bool i8, bool fast, const SynCode *d_code, int z, const SynLaunch &a, int block,
                  size_t lds, hipStream_t stream)
{
    int grid = (a.batch + a.cws - 1) / a.cws;
    if (grid <= 0) return SYN_OK;
    const void *fn = decode_kernel(kind, i8, fast, z);
    if (!fn) return SYN_INVALID_PARAM;
    // BG1 f32 with one codeword per workgroup, the posterior written back
    // and at least one iteration: the kernel that keeps the extension
    // columns in that posterior buffer (ldpc_fastx.h), two per CU
    if (fast && !i8 && kind == SYN_KIND_BG1 && a.cws == 1 && (a.flags & SYN_F_WRITE_POST) && a.post &&
        a.iterations >= 1 && !(a.flags & SYN_F_TRACE)) {
        int lanes = 0, per = 1, dev = 0, lim = 0;
        const void *fx = fastx_kernel_f32_bg1(z);
        const size_t xl = fx ? (size_t)fastx_shape_f32_bg1(z, &lanes, &per) : 0;
        // (a device with less LDS per workgroup keeps the all-LDS kernel)
        if (fx && hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
            xl <= (size_t)lim) {
            lds = xl;
            fn = fx;
            block = lanes;
            grid = (a.batch + per - 1) / per;
        }
    }
    if (int rc = allow_max_dynamic_lds(fn)) return rc;
    SynLaunch arg = a;
    void *args[] = {(void *)&d_code, (void *)&arg};
    if (hipLaunchKernel(fn, dim3(grid), dim3(block), args, lds, stream) != hipSuccess) return SYN_DEVICE_ERROR;
    return SYN_OK;
}

int launch_quantize(const float *in, int16_t *out, size_t n, float scale, hipStream_t stream)
{
    if (n == 0) return SYN_OK;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(quantize_pass<int16_t>, dim3((unsigned)blocks), dim3(256), 0, stream, in, out, n, scale);
    return hipGetLastError() == hipSuccess ? SYN_OK : SYN_DEVICE_ERROR;
}

int launch_quantize_i8(const float *in, int8_t *out, size_t n, float scale, hipStream_t stream)
{
    if (n == 0) return SYN_OK;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(quantize_pass<int8_t>, dim3((unsigned)blocks), dim3(256), 0, stream, in, out, n, scale);
    return hipGetLastError() == hipSuccess ? SYN_OK : SYN_DEVICE_ERROR;
}

int launch_dequantize_i8(const int8_t *in, float *out, size_t n, float scale, hipStream_t stream)
{
    if (n == 0) return SYN_OK;
    const float inv = 1.0f / scale;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(dequantize_i8, dim3((unsigned)blocks), dim3(256), 0, stream, in, out, n, inv);
    return hipGetLastError() == hipSuccess ? SYN_OK : SYN_DEVICE_ERROR;
}

}  // namespace syn
