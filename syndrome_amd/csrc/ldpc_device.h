// Device-side building blocks shared by the LOMS QC-LDPC kernels for gfx950
// (MI355X).  Kernel families:
//
// Two kernel families, both bit-exact with the reference's layered offset
// min-sum decode (src/qc_ldpc.rs:962-1252 f32, 1530-1787 int8):
//
//  * plain_*  -- runtime-structured: any (raw-edge) code, the layer loop
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
// layer touch disjoint posteriors (each variable appears at most once per
// layer), so no two lanes ever write the same LDS word between barriers.
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
// (or nfill[id]).  The contiguous form serves the plain batch launches; the
// indirect form serves work-queue items over a packed arena of mixed codes.
struct GroupIO {
    int count;
    int base;
    const int *ids;
    const long long *offs;
    const int *nfill;

    __device__ __forceinline__ int id(int j) const { return ids ? ids[j] : base + j; }
    __device__ __forceinline__ long long off(int j, int N) const
    {
        return offs ? offs[id(j)] : (long long)id(j) * N;
    }
    __device__ __forceinline__ long long n_filler(int j, long long dflt) const { return nfill ? nfill[id(j)] : dflt; }
};

__device__ __forceinline__ GroupIO contiguous_group(int cw0, int cws, int batch)
{
    const int n = batch - cw0;
    return GroupIO{n < cws ? n : cws, cw0, nullptr, nullptr, nullptr};
}

// Load the channel input of the workgroup's codewords into LDS, applying the
// 5G prologue (init_5g_llr, qc_ldpc.rs:711-732 / decode_5g_i8, 1505-1514),
// quantisation (quantize_llr_i16, quantize.rs:331-341) and the int8 path's
// entry clamp (qc_ldpc.rs:1582-1584) on the way.  `cws` codeword slots are
// filled; slots past io.count get zeros.
template <bool I8>
__device__ __forceinline__ void load_input(const SynLaunch &a, int Z, int n_b, int m_b, int N, int cws,
                                           const GroupIO &io, void *lds)
{
    const int total = cws * N;
    const int K = (n_b - m_b) * Z;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int cl = i / N, pos = i - cl * N;
        const bool live = cl < io.count;
        const long long g = live ? io.off(cl, N) + pos : 0;
        const long long nf = live ? io.n_filler(cl, a.n_filler) : 0;
        const long long kp = nf >= K ? 0 : K - nf;
        if constexpr (!I8) {
            float v = 0.f;
            if (live) {
                v = ((const float *)a.in)[g];
                if (a.flags & SYN_F_INIT_5G) {
                    if (pos < 2 * Z) v = 0.f;
                    else if (pos >= kp && pos < K) v = 1000000.0f;
                }
            }
            ((float *)lds)[i] = v;
        } else {
            int v = 0;
            if (live) {
                if (a.flags & SYN_F_QUANTIZE_IN) {
                    float f = ((const float *)a.in)[g] * a.scale;
                    if (f != f) v = 0;
                    else {
                        f = roundf(f);  // ties away from zero, as f32::round
                        f = f < -127.f ? -127.f : (f > 127.f ? 127.f : f);
                        v = (int)f;
                    }
                } else {
                    v = ((const int16_t *)a.in)[g];
                }
                if (a.flags & SYN_F_INIT_5G) {
                    if (pos < 2 * Z) v = 0;
                    else if (pos >= kp && pos < K) v = a.app_clamp;
                }
                v = v < -a.app_clamp ? -a.app_clamp : (v > a.app_clamp ? a.app_clamp : v);
            }
            ((int16_t *)lds)[i] = (int16_t)v;
        }
    }
}

template <bool I8>
__device__ __forceinline__ void store_output(const SynLaunch &a, int N, const GroupIO &io, const void *lds)
{
    const int total = io.count * N;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int cl = i / N, pos = i - cl * N;
        const long long g = io.off(cl, N) + pos;
        if constexpr (!I8) {
            const float v = ((const float *)lds)[i];
            if (a.flags & SYN_F_WRITE_POST) ((float *)a.post)[g] = v;
            if (a.flags & SYN_F_WRITE_HARD) a.hard[g] = v < 0.f ? 1 : 0;
        } else {
            const int16_t v = ((const int16_t *)lds)[i];
            if (a.flags & SYN_F_WRITE_POST) ((int16_t *)a.post)[g] = v;
            if (a.flags & SYN_F_WRITE_HARD) a.hard[g] = v < 0 ? 1 : 0;
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
