// Device-side building blocks shared by the LOMS QC-LDPC kernels for gfx950
// (MI355X).  Kernel families:
//
// Two kernel families, both bit-exact with the reference's layered offset
// min-sum decode (src/qc_ldpc.rs:962-1252 f32, 1530-1787 int8), plus the
// wide kernels of ldpc_wide.hip for raw codes outside the inline tables:
//
//  * plain_*  -- runtime-structured: any inline-table code, the layer loop
//                reads the code tables at run time, the check-to-variable
//                messages R live uncompressed in HBM (the edge_r buffer),
//                the posterior of each codeword is staged in LDS.
//  * fast_*   -- specialised per 3GPP base graph (BG1 / BG2) at compile
//                time: the whole iteration is unrolled, column indices are
//                immediates, only the Z-dependent shifts are loaded (scalar
//                loads).  R never leaves the chip: each lane owns check row
//                (layer, z) for every layer and keeps that row's messages in
//                registers in the compressed min-sum form
//                (|R| of the minimum edge, |R| of every other edge, the
//                argmin edge and one sign bit per edge).  The posterior
//                lives in LDS.  HBM sees only the channel input and the
//                outputs.
//
// Thread mapping (both families): a workgroup decodes `cws` codewords of
// the same code; lane t handles codeword t / Z, lifting index z = t % Z.
// All lanes of a workgroup walk the layers in the reference's strict order
// with a workgroup barrier between dependent layers; z positions of one
// layer touch disjoint posteriors (in these codes each variable appears at
// most once per layer), so no two lanes ever write the same LDS word between
// barriers.
//
// Exactness: every f32 operation is the reference's single IEEE operation
// in the reference's order (q = L - R_old; L' = L + (R_new - R_old)); the
// build uses -ffp-contract=off and keeps denormals.  The min/second-min/
// sign reduction does not round, so its order is free.
#pragma once
#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>
#include <utility>

#include "ldpc_common.h"

#define SYN_TABLE_QUAL static constexpr
#include "code_tables.inc"

namespace syn {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int wrap_add(int z, int sh, int Z)
{
    int s = z + sh;
    return s >= Z ? s - Z : s;
}

// Hide a value from loop-invariant code motion: the posterior addresses of
// all 316 edges depend only on z and the shifts, and hoisting them out of
// the iteration loop would hold them in registers the message state needs.
__device__ __forceinline__ int opaque_v(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ int opaque_s(int v)
{
    asm volatile("" : "+s"(v));
    return v;
}

// clamp(x, lo, hi) with lo <= hi: the backend emits one v_med3.
__device__ __forceinline__ uint32_t clampu(uint32_t x, uint32_t lo, uint32_t hi)
{
    const uint32_t t = x > lo ? x : lo;
    return t < hi ? t : hi;
}

__device__ __forceinline__ int clampi(int x, int lo, int hi)
{
    const int t = x > lo ? x : lo;
    return t < hi ? t : hi;
}

// The median of three, which is clamp(x, lo, hi) whenever lo <= hi: one
// v_med3 where the backend would otherwise emit a max and a min (it cannot
// see that the running minimum never exceeds the running second minimum).
__device__ __forceinline__ uint32_t med3u(uint32_t x, uint32_t lo, uint32_t hi)
{
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
    return r;
}

__device__ __forceinline__ int med3i(int x, int lo, int hi)
{
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
    return r;
}

// (a ^ b) + c in one VALU operation.  Written out because the compiler
// rewrites "(m ^ n) + c - n" (a sign applied to a magnitude, then added) as
// two bit-field extracts, an xor and an add3, one operation more.
__device__ __forceinline__ int xad(int a, int b, int c)
{
    int r;
    asm("v_xad_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

template <int I, int N>
struct SFor {
    template <class F>
    __device__ __forceinline__ static void run(F &&f)
    {
        if constexpr (I < N) {
            f(std::integral_constant<int, I>{});
            SFor<I + 1, N>::run(f);
        }
    }
};

// Where a workgroup's codewords live: `count` codewords, ids base + j (or
// ids[j]), element offsets id * N (or offs[id]), filler counts the launch's
// (or nfill[id]); inputs at the launch's `in` + offset.  A resident-server
// item instead brings, per position j, its own input pointer srcs[j] and
// filler count nfill_pos[j].  The contiguous form serves the plain batch
// launches; the indirect forms serve work-queue items over a packed arena
// of mixed codes and the resident server.
struct GroupIO {
    int count;
    int base;
    const int *ids;
    const long long *offs;
    const int *nfill;
    const void *const *srcs;
    const int *nfill_pos;
    uint32_t no_hard = 0;  // bit j: codeword j's hard decisions are not written (its consumer reads LDS)
    // Optional LDS copies of offs[id(j)] and nfill[id(j)] for j < count.
    // Without them every input and output chunk of the indirect form looks
    // its codeword up through ids, then offs (a chain of dependent global
    // loads that the output stores, which may alias them, force again on
    // every chunk).
    const long long *t_off = nullptr;
    const int *t_nf = nullptr;

    __device__ __forceinline__ int id(int j) const { return ids ? ids[j] : base + j; }
    __device__ __forceinline__ long long off(int j, int N) const
    {
        if (t_off) return t_off[j];
        return offs ? offs[id(j)] : (long long)id(j) * N;
    }
    __device__ __forceinline__ long long n_filler(int j, long long dflt) const
    {
        if (t_nf) return t_nf[j];
        return nfill_pos ? nfill_pos[j] : (nfill ? nfill[id(j)] : dflt);
    }
    // first input byte of codeword j (elements of esz bytes)
    __device__ __forceinline__ const char *src(int j, int N, const void *in, int esz) const
    {
        return srcs ? (const char *)srcs[j] : (const char *)in + off(j, N) * esz;
    }
};

__device__ __forceinline__ GroupIO contiguous_group(int cw0, int cws, int batch)
{
    const int n = batch - cw0;
    return GroupIO{n < cws ? n : cws, cw0, nullptr, nullptr, nullptr, nullptr, nullptr};
}

// Per-element input prologue: the 5G layout (init_5g_llr,
// qc_ldpc.rs:711-732: the 2Z punctured positions read 0, the fillers
// K'..K read the saturating 1e6), for the int8 path quantisation
// (quantize_llr_i16, quantize.rs:331-341: v*s, NaN -> 0, round half away
// from zero, clamp to +-127), the 5G layout in the int8 form (decode_5g_i8,
// qc_ldpc.rs:1505-1514) and the entry clamp to +-C (qc_ldpc.rs:1582-1584).
__device__ __forceinline__ float prologue_f32(float v, int pos, int Z2, int kp, int K, bool init5g)
{
    if (init5g) {
        if (pos < Z2) v = 0.f;
        else if (pos >= kp && pos < K) v = 1000000.0f;
    }
    return v;
}

__device__ __forceinline__ int quantize_one(float x, float scale)
{
    float f = x * scale;
    if (f != f) return 0;
    f = roundf(f);  // ties away from zero, as f32::round
    f = f < -127.f ? -127.f : (f > 127.f ? 127.f : f);
    return (int)f;
}

__device__ __forceinline__ int prologue_i8(int v, int pos, int Z2, int kp, int K, bool init5g, int C)
{
    if (init5g) {
        if (pos < Z2) v = 0;
        else if (pos >= kp && pos < K) v = C;
    }
    return v < -C ? -C : (v > C ? C : v);
}

// Load the channel input of the workgroup's codewords into LDS with the
// prologue above.  `cws` codeword slots are filled; slots past io.count get
// zeros.  With SYN_F_ALIGNED (every element offset and N a multiple of 4,
// base pointers 16-byte aligned) each lane moves 4 elements per access and
// keeps 8 accesses in flight before it writes any of them to LDS, so a
// workgroup pays a handful of HBM round trips, not one per element.
template <bool I8>
__device__ __forceinline__ void load_input(const SynLaunch &a, int Z, int n_b, int m_b, int N, int cws,
                                           const GroupIO &io, void *lds)
{
    const int K = (n_b - m_b) * Z, Z2 = 2 * Z;
    const bool init5g = (a.flags & SYN_F_INIT_5G) != 0;
    const bool qin = I8 && (a.flags & SYN_F_QUANTIZE_IN) != 0;
    const int C = a.app_clamp;
    if ((a.flags & SYN_F_ALIGNED) && (N & 3) == 0) {
        constexpr int U = 8;
        const int n4 = N >> 2, total4 = cws * n4, step = (int)blockDim.x;
        for (int b = threadIdx.x; b < total4; b += U * step) {
            float4 vf[U];
            uint2 vh[U];
            int kp[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const int i4 = b + k * step;
                const int cl = i4 / n4;
                vf[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                vh[k] = make_uint2(0u, 0u);
                kp[k] = K;
                if (i4 < total4 && cl < io.count) {
                    const long long nf = io.n_filler(cl, a.n_filler);
                    kp[k] = nf >= K ? 0 : K - (int)nf;
                    if (!I8 || qin)
                        vf[k] = *(const float4 *)(io.src(cl, N, a.in, 4) + 16LL * (i4 - cl * n4));
                    else
                        vh[k] = *(const uint2 *)(io.src(cl, N, a.in, 2) + 8LL * (i4 - cl * n4));
                }
            }
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const int i4 = b + k * step;
                if (i4 < total4) {
                    const int cl = i4 / n4, p = 4 * (i4 - cl * n4);
                    const bool live = cl < io.count;
                    if constexpr (!I8) {
                        float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (live) {
                            o.x = prologue_f32(vf[k].x, p, Z2, kp[k], K, init5g);
                            o.y = prologue_f32(vf[k].y, p + 1, Z2, kp[k], K, init5g);
                            o.z = prologue_f32(vf[k].z, p + 2, Z2, kp[k], K, init5g);
                            o.w = prologue_f32(vf[k].w, p + 3, Z2, kp[k], K, init5g);
                        }
                        *(float4 *)((float *)lds + 4LL * i4) = o;
                    } else {
                        int e[4] = {0, 0, 0, 0};
                        if (live) {
                            if (qin) {
                                e[0] = quantize_one(vf[k].x, a.scale);
                                e[1] = quantize_one(vf[k].y, a.scale);
                                e[2] = quantize_one(vf[k].z, a.scale);
                                e[3] = quantize_one(vf[k].w, a.scale);
                            } else {
                                e[0] = (int16_t)(vh[k].x & 0xffffu);
                                e[1] = (int16_t)(vh[k].x >> 16);
                                e[2] = (int16_t)(vh[k].y & 0xffffu);
                                e[3] = (int16_t)(vh[k].y >> 16);
                            }
#pragma unroll
                            for (int q = 0; q < 4; ++q) e[q] = prologue_i8(e[q], p + q, Z2, kp[k], K, init5g, C);
                        }
                        *(uint2 *)((int16_t *)lds + 4LL * i4) =
                            make_uint2(((uint32_t)e[0] & 0xffffu) | ((uint32_t)e[1] << 16),
                                       ((uint32_t)e[2] & 0xffffu) | ((uint32_t)e[3] << 16));
                    }
                }
            }
        }
        return;
    }
    const int total = cws * N;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int cl = i / N, pos = i - cl * N;
        const bool live = cl < io.count;
        const long long nf = live ? io.n_filler(cl, a.n_filler) : 0;
        const int kp = nf >= K ? 0 : K - (int)nf;
        if constexpr (!I8) {
            float v = 0.f;
            if (live) v = prologue_f32(((const float *)io.src(cl, N, a.in, 4))[pos], pos, Z2, kp, K, init5g);
            ((float *)lds)[i] = v;
        } else {
            int v = 0;
            if (live) {
                v = qin ? quantize_one(((const float *)io.src(cl, N, a.in, 4))[pos], a.scale)
                        : ((const int16_t *)io.src(cl, N, a.in, 2))[pos];
                v = prologue_i8(v, pos, Z2, kp, K, init5g, C);
            }
            ((int16_t *)lds)[i] = (int16_t)v;
        }
    }
}

// Write the posterior and/or hard decisions of the live codewords back;
// with SYN_F_ALIGNED 4 elements per lane and access (hard decisions as one
// 32-bit word).
template <bool I8>
__device__ __forceinline__ void store_output(const SynLaunch &a, int N, const GroupIO &io, const void *lds)
{
    const bool post = (a.flags & SYN_F_WRITE_POST) != 0, hard = (a.flags & SYN_F_WRITE_HARD) != 0;
    if ((a.flags & SYN_F_ALIGNED) && (N & 3) == 0) {
        const int n4 = N >> 2, total4 = io.count * n4;
        for (int i4 = threadIdx.x; i4 < total4; i4 += blockDim.x) {
            const int cl = i4 / n4;
            const long long g = io.off(cl, N) + 4LL * (i4 - cl * n4);
            uint32_t h;
            if constexpr (!I8) {
                const float4 v = *(const float4 *)((const float *)lds + 4LL * i4);
                if (post) *(float4 *)((float *)a.post + g) = v;
                h = (v.x < 0.f ? 1u : 0u) | (v.y < 0.f ? 1u << 8 : 0u) | (v.z < 0.f ? 1u << 16 : 0u) |
                    (v.w < 0.f ? 1u << 24 : 0u);
            } else {
                const uint2 v = *(const uint2 *)((const int16_t *)lds + 4LL * i4);
                if (post) *(uint2 *)((int16_t *)a.post + g) = v;
                h = ((v.x >> 15) & 1u) | (((v.x >> 31) & 1u) << 8) | (((v.y >> 15) & 1u) << 16) |
                    (((v.y >> 31) & 1u) << 24);
            }
            if (hard && !((io.no_hard >> cl) & 1u)) *(uint32_t *)(a.hard + g) = h;
        }
        return;
    }
    const int total = io.count * N;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int cl = i / N, pos = i - cl * N;
        const long long g = io.off(cl, N) + pos;
        if constexpr (!I8) {
            const float v = ((const float *)lds)[i];
            if (post) ((float *)a.post)[g] = v;
            if (hard && !((io.no_hard >> cl) & 1u)) a.hard[g] = v < 0.f ? 1 : 0;
        } else {
            const int16_t v = ((const int16_t *)lds)[i];
            if (post) ((int16_t *)a.post)[g] = v;
            if (hard && !((io.no_hard >> cl) & 1u)) a.hard[g] = v < 0 ? 1 : 0;
        }
    }
}

// Per-codeword bookkeeping in LDS after the posterior: fail flag, done flag,
// iterations used.
struct CwFlags {
    int *fail, *done, *used;
};

__device__ __forceinline__ CwFlags cw_flags(void *base, int cws)
{
    int *p = (int *)base;
    return {p, p + cws, p + 2 * cws};
}

__device__ __forceinline__ void flags_init(CwFlags f, int cws, const GroupIO &io)
{
    if ((int)threadIdx.x < cws) {
        f.fail[threadIdx.x] = 0;
        f.done[threadIdx.x] = ((int)threadIdx.x >= io.count) ? 1 : 0;
        f.used[threadIdx.x] = 0;
    }
}

// decode_layered_offset_min_sum_traced (qc_ldpc.rs:853-874, the
// on_iteration call at :1208): after pass it + 1, before its syndrome check,
// the posterior of every codeword still decoding goes to its trace row.
// P = the lane's codeword posterior (f32, column-major as in HBM).  Used by
// the plain kernel, which every traced decode runs on.
__device__ __forceinline__ void trace_store(const SynLaunch &a, const float *P, int N, int Z, int z, bool act,
                                            const GroupIO &io, int cl, int it)
{
    if (!(a.flags & SYN_F_TRACE) || !act || cl >= io.count) return;
    float *dst = a.trace + ((size_t)io.id(cl) * (size_t)a.iterations + (size_t)it) * (size_t)N;
    for (int v = z; v < N; v += Z) dst[v] = P[v];
}

__device__ __forceinline__ void store_iters(const SynLaunch &a, CwFlags f, const GroupIO &io)
{
    if ((int)threadIdx.x < io.count) a.iters[io.id(threadIdx.x)] = f.used[threadIdx.x];
}

__device__ __forceinline__ bool all_done(CwFlags f, int cws)
{
    int d = 1;
    for (int c = 0; c < cws; ++c) d &= f.done[c];
    return d != 0;
}

// After the end-of-iteration syndrome pass: a codeword whose lanes raised
// no failure has a zero syndrome and stops (qc_ldpc.rs:1211 / 1748).
__device__ __forceinline__ void flags_settle(CwFlags f, int cws)
{
    if ((int)threadIdx.x < cws) {
        if (!f.done[threadIdx.x] && !f.fail[threadIdx.x]) f.done[threadIdx.x] = 1;
        f.fail[threadIdx.x] = 0;
    }
}


}  // namespace syn
