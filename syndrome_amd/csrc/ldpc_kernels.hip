// LOMS QC-LDPC decode kernels for gfx950 (MI355X).
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

// Load the channel input of the workgroup's codewords into LDS, applying the
// 5G prologue (init_5g_llr, qc_ldpc.rs:711-732 / decode_5g_i8, 1505-1514),
// quantisation (quantize_llr_i16, quantize.rs:331-341) and the int8 path's
// entry clamp (qc_ldpc.rs:1582-1584) on the way.
template <bool I8>
__device__ __forceinline__ void load_input(const SynLaunch &a, int Z, int n_b, int m_b, int N, int cw0,
                                           void *lds)
{
    const int total = a.cws * N;
    const int K = (n_b - m_b) * Z;
    const long long kp = a.n_filler >= K ? 0 : K - a.n_filler;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int cl = i / N, pos = i - cl * N, cw = cw0 + cl;
        if constexpr (!I8) {
            float v = 0.f;
            if (cw < a.batch) {
                v = ((const float *)a.in)[(size_t)cw * N + pos];
                if (a.flags & SYN_F_INIT_5G) {
                    if (pos < 2 * Z) v = 0.f;
                    else if (pos >= kp && pos < K) v = 1000000.0f;
                }
            }
            ((float *)lds)[i] = v;
        } else {
            int v = 0;
            if (cw < a.batch) {
                if (a.flags & SYN_F_QUANTIZE_IN) {
                    float f = ((const float *)a.in)[(size_t)cw * N + pos] * a.scale;
                    if (f != f) v = 0;
                    else {
                        f = roundf(f);  // ties away from zero, as f32::round
                        f = f < -127.f ? -127.f : (f > 127.f ? 127.f : f);
                        v = (int)f;
                    }
                } else {
                    v = ((const int16_t *)a.in)[(size_t)cw * N + pos];
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
__device__ __forceinline__ void store_output(const SynLaunch &a, int N, int cw0, const void *lds)
{
    const int total = a.cws * N;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int cl = i / N, pos = i - cl * N, cw = cw0 + cl;
        if (cw >= a.batch) continue;
        const size_t g = (size_t)cw * N + pos;
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

__device__ __forceinline__ void flags_init(CwFlags f, int cws, int cw0, int batch)
{
    if ((int)threadIdx.x < cws) {
        f.fail[threadIdx.x] = 0;
        f.done[threadIdx.x] = (cw0 + (int)threadIdx.x >= batch) ? 1 : 0;
        f.used[threadIdx.x] = 0;
    }
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
    load_input<false>(a, Z, n_b, m_b, N, cw0, P0);
    float *edge = (float *)a.edge;
    for (int i = t; i < cws * EZ; i += blockDim.x)
        if (cw0 + i / EZ < a.batch) edge[(size_t)cw0 * EZ + i] = 0.f;  // qc_ldpc.rs:1006-1009
    flags_init(fl, cws, cw0, a.batch);
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
    store_output<false>(a, N, cw0, P0);
    if (t < cws && cw0 + t < a.batch) a.iters[cw0 + t] = fl.used[t];
}

__global__ void __launch_bounds__(1024) plain_i8(const SynCode *__restrict__ code, SynLaunch a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int Z = code->z, m_b = code->m_b, n_b = code->n_b, N = n_b * Z;
    const int EZ = code->n_edges * Z, cws = a.cws, cw0 = blockIdx.x * cws;
    int16_t *P0 = (int16_t *)smem;
    CwFlags fl = cw_flags(P0 + ((cws * N + 1) & ~1), cws);
    const int t = threadIdx.x;
    load_input<true>(a, Z, n_b, m_b, N, cw0, P0);
    int8_t *edge = (int8_t *)a.edge;
    for (int i = t; i < cws * EZ; i += blockDim.x)
        if (cw0 + i / EZ < a.batch) edge[(size_t)cw0 * EZ + i] = 0;  // qc_ldpc.rs:1579-1581
    flags_init(fl, cws, cw0, a.batch);
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
    store_output<true>(a, N, cw0, P0);
    if (t < cws && cw0 + t < a.batch) a.iters[cw0 + t] = fl.used[t];
}

// ---------------------------------------------------------------------------
// fast kernels: compile-time base graph, compressed R in registers
// ---------------------------------------------------------------------------

template <int KIND>
struct Graph;

template <>
struct Graph<SYN_KIND_BG1> {
    static constexpr int ROWS = SYN_BG1_ROWS, COLS = SYN_BG1_COLS, EDGES = SYN_BG1_EDGES;
    static constexpr int deg(int l) { return SYN_BG1_ROW_DEG[l]; }
    static constexpr int col(int e) { return SYN_BG1_EDGE_COL[e]; }
};

template <>
struct Graph<SYN_KIND_BG2> {
    static constexpr int ROWS = SYN_BG2_ROWS, COLS = SYN_BG2_COLS, EDGES = SYN_BG2_EDGES;
    static constexpr int deg(int l) { return SYN_BG2_ROW_DEG[l]; }
    static constexpr int col(int e) { return SYN_BG2_EDGE_COL[e]; }
};

template <class G>
constexpr int layer_begin(int l)
{
    int s = 0;
    for (int i = 0; i < l; ++i) s += G::deg(i);
    return s;
}

// A barrier is needed after layer l unless layer l+1 touches none of the
// columns touched since the last barrier: layers over disjoint variables
// commute exactly, so running them without a barrier between them changes
// nothing.
template <class G>
constexpr bool barrier_after(int l)
{
    if (l + 1 >= G::ROWS) return true;
    // Greedy grouping from layer 0: a layer opens a new group (and needs a
    // barrier before it) iff it shares a column with the current group.
    bool used[SYN_MAX_COLS] = {};
    for (int i = 0; i <= l + 1; ++i) {
        const int b = layer_begin<G>(i);
        bool clash = false;
        for (int e = 0; e < G::deg(i); ++e) clash = clash || used[G::col(b + e)];
        if (i == l + 1) return clash;
        if (clash)
            for (int c = 0; c < SYN_MAX_COLS; ++c) used[c] = false;
        for (int e = 0; e < G::deg(i); ++e) used[G::col(b + e)] = true;
    }
    return true;
}

// Per-layer sign word: bits 0..deg-1 hold the sign of R on each edge, bits
// deg..deg+4 the argmin edge (31 = none).  Low-degree layers share a
// register: words are packed greedily, in layer order, into 32-bit words.
template <class G>
constexpr int sg_bits(int l)
{
    return G::deg(l) + 5;
}

template <class G>
constexpr int sg_pos(int l, bool want_word)
{
    int w = 0, off = 0;
    for (int i = 0; i <= l; ++i) {
        if (off + sg_bits<G>(i) > 32) {
            ++w;
            off = 0;
        }
        if (i == l) return want_word ? w : off;
        off += sg_bits<G>(i);
    }
    return 0;
}

template <class G>
constexpr int sg_words()
{
    return sg_pos<G>(G::ROWS - 1, true) + 1;
}

template <class G>
struct StateF32 {
    float g1[G::ROWS];            // |R| on every edge but the argmin: max(min1 - beta, 0)
    float g2[G::ROWS];            // |R| on the argmin edge:          max(min2 - beta, 0)
    uint32_t sg[sg_words<G>()];   // packed sign words
};

template <class G>
struct StateI8 {
    uint32_t mg[(G::ROWS + 1) / 2];  // per layer 16 bits: |R| others (bits 0..7), argmin (8..15)
    uint32_t sg[sg_words<G>()];
};

template <class G, int L>
__device__ __forceinline__ uint32_t sg_get(const uint32_t *sg)
{
    constexpr int W = sg_pos<G>(L, true), O = sg_pos<G>(L, false);
    return sg[W] >> O;
}

template <class G, int L>
__device__ __forceinline__ void sg_put(uint32_t *sg, uint32_t v)
{
    constexpr int W = sg_pos<G>(L, true), O = sg_pos<G>(L, false);
    constexpr uint32_t M = ((1u << sg_bits<G>(L)) - 1u) << O;
    sg[W] = (sg[W] & ~M) | ((v << O) & M);
}

// R of edge e rebuilt from a layer's compressed state (s = its sign word).
// The reference picks min2 for an edge whose |q| equals min1 and min1
// otherwise; an edge other than the tracked argmin can equal min1 only when
// min1 is tied, and a tie forces min2 == min1 in pass 1, so "is the argmin"
// selects the same value.
template <int D>
__device__ __forceinline__ float r_f32(float g1, float g2, uint32_t s, int e)
{
    const float m = (((s >> D) & 31u) == (uint32_t)e) ? g2 : g1;
    return __uint_as_float(__float_as_uint(m) | ((s << (31 - e)) & 0x80000000u));
}

template <int D>
__device__ __forceinline__ int r_i8(uint32_t mg, uint32_t s, int e)
{
    const int m = (((s >> D) & 31u) == (uint32_t)e) ? (int)((mg >> 8) & 0xffu) : (int)(mg & 0xffu);
    return ((s >> e) & 1u) ? -m : m;
}

template <class G, int L>
__device__ __forceinline__ void fast_layer_f32(float *__restrict__ P, int z, int Z, const int *__restrict__ shifts,
                                               StateF32<G> &st, float beta, bool act)
{
    constexpr int D = G::deg(L), B = layer_begin<G>(L);
    static_assert(D >= 2 && D <= SYN_MAX_DEG, "row degree");
    float Lv[D];
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        Lv[e] = P[G::col(B + e) * Z + wrap_add(z, shifts[B + e], Z)];
    });
    const float o1 = st.g1[L], o2 = st.g2[L];
    const uint32_t os = sg_get<G, L>(st.sg);
    float m1 = FLT_MAX, m2 = FLT_MAX;
    uint32_t qs = 0, am = 31;
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        const float q = Lv[e] - r_f32<D>(o1, o2, os, e);  // Q-build
        const uint32_t qb = __float_as_uint(q);
        const float aa = __uint_as_float(qb & 0x7fffffffu);
        qs |= (qb >> 31) << e;
        const bool le = aa <= m1;  // pass 1
        m2 = le ? m1 : (aa < m2 ? aa : m2);
        m1 = le ? aa : m1;
        am = le ? (uint32_t)e : am;
    });
    if (m1 == FLT_MAX) m1 = 0.f;
    if (m2 == FLT_MAX) m2 = m1;
    float g1 = m1 - beta;
    g1 = g1 < 0.f ? 0.f : g1;
    float g2 = m2 - beta;
    g2 = g2 < 0.f ? 0.f : g2;
    // sign of R_e = XOR of all q signs but e's own
    const uint32_t ns = (qs ^ ((__popc(qs) & 1u) ? ((1u << D) - 1u) : 0u)) | (am << D);
    if (act) {
        SFor<0, D>::run([&](auto ec) {
            constexpr int e = decltype(ec)::value;
            const float rn = r_f32<D>(g1, g2, ns, e);
            P[G::col(B + e) * Z + wrap_add(z, shifts[B + e], Z)] = Lv[e] + (rn - r_f32<D>(o1, o2, os, e));
        });
        st.g1[L] = g1;
        st.g2[L] = g2;
        sg_put<G, L>(st.sg, ns);
    }
}

template <class G, int L>
__device__ __forceinline__ uint32_t mg_get(const StateI8<G> &st)
{
    return (st.mg[L / 2] >> (16 * (L % 2))) & 0xffffu;
}

template <class G, int L>
__device__ __forceinline__ void mg_put(StateI8<G> &st, uint32_t v)
{
    constexpr uint32_t M = 0xffffu << (16 * (L % 2));
    st.mg[L / 2] = (st.mg[L / 2] & ~M) | (v << (16 * (L % 2)));
}

template <class G, int L>
__device__ __forceinline__ void fast_layer_i8(int16_t *__restrict__ P, int z, int Z, const int *__restrict__ shifts,
                                              StateI8<G> &st, int bq, int C, bool act)
{
    constexpr int D = G::deg(L), B = layer_begin<G>(L);
    static_assert(D >= 2 && D <= SYN_MAX_DEG, "row degree");
    int Av[D];
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        Av[e] = P[G::col(B + e) * Z + wrap_add(z, shifts[B + e], Z)];
    });
    const uint32_t om = mg_get<G, L>(st), os = sg_get<G, L>(st.sg);
    int m1 = 127, m2 = 127;
    uint32_t qs = 0, am = 31;
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        int q = Av[e] - r_i8<D>(om, os, e);  // Q-build, clamped to the message range
        q = q < -127 ? -127 : (q > 127 ? 127 : q);
        const int aa = q < 0 ? -q : q;
        qs |= (q < 0 ? 1u : 0u) << e;
        const int mx = m1 > aa ? m1 : aa;  // pass 1: sorted insertion
        m2 = m2 < mx ? m2 : mx;
        am = (aa <= m1) ? (uint32_t)e : am;
        m1 = m1 < aa ? m1 : aa;
    });
    int g1 = m1 - bq;
    g1 = g1 < 0 ? 0 : g1;
    int g2 = m2 - bq;
    g2 = g2 < 0 ? 0 : g2;
    const uint32_t ns = (qs ^ ((__popc(qs) & 1u) ? ((1u << D) - 1u) : 0u)) | (am << D);
    const uint32_t nm = (uint32_t)g1 | ((uint32_t)g2 << 8);
    if (act) {
        SFor<0, D>::run([&](auto ec) {
            constexpr int e = decltype(ec)::value;
            int nv = Av[e] + (r_i8<D>(nm, ns, e) - r_i8<D>(om, os, e));
            nv = nv < -C ? -C : (nv > C ? C : nv);
            P[G::col(B + e) * Z + wrap_add(z, shifts[B + e], Z)] = (int16_t)nv;
        });
        mg_put<G, L>(st, nm);
        sg_put<G, L>(st.sg, ns);
    }
}

template <class G, int L, class T>
__device__ __forceinline__ uint32_t row_parity(const T *__restrict__ P, int z, int Z, const int *__restrict__ shifts)
{
    constexpr int D = G::deg(L), B = layer_begin<G>(L);
    uint32_t p = 0;
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        p ^= (P[G::col(B + e) * Z + wrap_add(z, shifts[B + e], Z)] < (T)0) ? 1u : 0u;
    });
    return p;
}

// Syndrome of the lane's check rows, all layers, stopping early once this
// lane or another lane of its codeword has found a failed check.
template <class G, int L, class T>
__device__ __forceinline__ void check_rows(const T *__restrict__ P, int z, int Z, const int *__restrict__ shifts,
                                           uint32_t &fail, volatile int *cw_fail)
{
    if constexpr (L < G::ROWS) {
        fail |= row_parity<G, L>(P, z, Z, shifts);
        if constexpr (L == 0 || L == 3 || L == 9 || L == 19 || L == 31) {
            if (fail) {
                *cw_fail = 1;
                return;
            }
            if (*cw_fail) return;
        }
        check_rows<G, L + 1>(P, z, Z, shifts, fail, cw_fail);
    }
}

template <int KIND>
__global__ void __launch_bounds__(512) fast_f32(const SynCode *__restrict__ code, SynLaunch a)
{
    using G = Graph<KIND>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int Z = code->z, N = G::COLS * Z, cws = a.cws, cw0 = blockIdx.x * cws;
    const int EZ = G::EDGES * Z;
    float *P0 = (float *)smem;
    CwFlags fl = cw_flags(P0 + cws * N, cws);
    const int t = threadIdx.x;
    const int *__restrict__ shifts = code->shifts;
    load_input<false>(a, Z, G::COLS, G::ROWS, N, cw0, P0);
    flags_init(fl, cws, cw0, a.batch);
    __syncthreads();

    const bool lane_ok = t < cws * Z;
    const int cl = lane_ok ? t / Z : 0;
    const int z = lane_ok ? t - cl * Z : 0;
    float *P = P0 + cl * N;
    const float beta = a.beta;

    StateF32<G> st;  // R = +0.0 everywhere (qc_ldpc.rs:1006-1009)
    SFor<0, G::ROWS>::run([&](auto lc) {
        constexpr int L = decltype(lc)::value;
        st.g1[L] = 0.f;
        st.g2[L] = 0.f;
    });
    SFor<0, sg_words<G>()>::run([&](auto wc) { st.sg[decltype(wc)::value] = 0u; });
    SFor<0, G::ROWS>::run([&](auto lc) { sg_put<G, decltype(lc)::value>(st.sg, 31u << G::deg(decltype(lc)::value)); });

    for (int it = 0; it < a.iterations; ++it) {
        if (all_done(fl, cws)) break;
        const bool act = lane_ok && !fl.done[cl];
        if (t < cws && !fl.done[t]) fl.used[t] = it + 1;
        const int zi = opaque_v(z);
        const int *__restrict__ sh = shifts + opaque_s(0);
        SFor<0, G::ROWS>::run([&](auto lc) {
            constexpr int L = decltype(lc)::value;
            fast_layer_f32<G, L>(P, zi, Z, sh, st, beta, act);
            if constexpr (barrier_after<G>(L)) __syncthreads();
        });
        if (act) {
            uint32_t fail = 0;
            check_rows<G, 0>(P, zi, Z, sh, fail, fl.fail + cl);
            if (fail) fl.fail[cl] = 1;
        }
        __syncthreads();
        flags_settle(fl, cws);
        __syncthreads();
    }
    store_output<false>(a, N, cw0, P0);
    if (t < cws && cw0 + t < a.batch) a.iters[cw0 + t] = fl.used[t];
    if ((a.flags & SYN_F_WRITE_EDGE) && lane_ok && cw0 + cl < a.batch) {
        float *E = (float *)a.edge + (size_t)(cw0 + cl) * EZ;
        SFor<0, G::ROWS>::run([&](auto lc) {
            constexpr int L = decltype(lc)::value;
            constexpr int D = G::deg(L), B = layer_begin<G>(L);
            SFor<0, D>::run([&](auto ec) {
                constexpr int e = decltype(ec)::value;
                E[(size_t)(B + e) * Z + z] = r_f32<D>(st.g1[L], st.g2[L], sg_get<G, L>(st.sg), e);
            });
        });
    }
}

template <int KIND>
__global__ void __launch_bounds__(512) fast_i8(const SynCode *__restrict__ code, SynLaunch a)
{
    using G = Graph<KIND>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int Z = code->z, N = G::COLS * Z, cws = a.cws, cw0 = blockIdx.x * cws;
    const int EZ = G::EDGES * Z;
    int16_t *P0 = (int16_t *)smem;
    CwFlags fl = cw_flags(P0 + ((cws * N + 1) & ~1), cws);
    const int t = threadIdx.x;
    const int *__restrict__ shifts = code->shifts;
    load_input<true>(a, Z, G::COLS, G::ROWS, N, cw0, P0);
    flags_init(fl, cws, cw0, a.batch);
    __syncthreads();

    const bool lane_ok = t < cws * Z;
    const int cl = lane_ok ? t / Z : 0;
    const int z = lane_ok ? t - cl * Z : 0;
    int16_t *P = P0 + cl * N;
    const int bq = a.beta_q, C = a.app_clamp;

    StateI8<G> st;  // R = 0 everywhere (qc_ldpc.rs:1579-1581)
    SFor<0, (G::ROWS + 1) / 2>::run([&](auto wc) { st.mg[decltype(wc)::value] = 0u; });
    SFor<0, sg_words<G>()>::run([&](auto wc) { st.sg[decltype(wc)::value] = 0u; });
    SFor<0, G::ROWS>::run([&](auto lc) { sg_put<G, decltype(lc)::value>(st.sg, 31u << G::deg(decltype(lc)::value)); });

    for (int it = 0; it < a.iterations; ++it) {
        if (all_done(fl, cws)) break;
        const bool act = lane_ok && !fl.done[cl];
        if (t < cws && !fl.done[t]) fl.used[t] = it + 1;
        const int zi = opaque_v(z);
        const int *__restrict__ sh = shifts + opaque_s(0);
        SFor<0, G::ROWS>::run([&](auto lc) {
            constexpr int L = decltype(lc)::value;
            fast_layer_i8<G, L>(P, zi, Z, sh, st, bq, C, act);
            if constexpr (barrier_after<G>(L)) __syncthreads();
        });
        if (act) {
            uint32_t fail = 0;
            check_rows<G, 0>(P, zi, Z, sh, fail, fl.fail + cl);
            if (fail) fl.fail[cl] = 1;
        }
        __syncthreads();
        flags_settle(fl, cws);
        __syncthreads();
    }
    store_output<true>(a, N, cw0, P0);
    if (t < cws && cw0 + t < a.batch) a.iters[cw0 + t] = fl.used[t];
    if ((a.flags & SYN_F_WRITE_EDGE) && lane_ok && cw0 + cl < a.batch) {
        int8_t *E = (int8_t *)a.edge + (size_t)(cw0 + cl) * EZ;
        SFor<0, G::ROWS>::run([&](auto lc) {
            constexpr int L = decltype(lc)::value;
            constexpr int D = G::deg(L), B = layer_begin<G>(L);
            SFor<0, D>::run([&](auto ec) {
                constexpr int e = decltype(ec)::value;
                E[(size_t)(B + e) * Z + z] = (int8_t)r_i8<D>(mg_get<G, L>(st), sg_get<G, L>(st.sg), e);
            });
        });
    }
}

// quantize_llr_i16 (quantize.rs:331-341) as a standalone pass.
__global__ void __launch_bounds__(256) quantize_i16(const float *__restrict__ in, int16_t *__restrict__ out, size_t n,
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
        out[i] = (int16_t)v;
    }
}

}  // namespace syn

// ---------------------------------------------------------------------------
// launchers (host)
// ---------------------------------------------------------------------------

namespace syn {

int launch_decode(int kind, bool i8, bool fast, const SynCode *d_code, const SynLaunch &a, int block, size_t lds,
                  hipStream_t stream)
{
    const int grid = (a.batch + a.cws - 1) / a.cws;
    if (grid <= 0) return SYN_OK;
    const void *fn;
    if (!fast) fn = i8 ? (const void *)plain_i8 : (const void *)plain_f32;
    else if (kind == SYN_KIND_BG1) fn = i8 ? (const void *)fast_i8<SYN_KIND_BG1> : (const void *)fast_f32<SYN_KIND_BG1>;
    else if (kind == SYN_KIND_BG2) fn = i8 ? (const void *)fast_i8<SYN_KIND_BG2> : (const void *)fast_f32<SYN_KIND_BG2>;
    else return SYN_INVALID_PARAM;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return SYN_DEVICE_ERROR;
    SynLaunch arg = a;
    void *args[] = {(void *)&d_code, (void *)&arg};
    if (hipLaunchKernel(fn, dim3(grid), dim3(block), args, lds, stream) != hipSuccess) return SYN_DEVICE_ERROR;
    return SYN_OK;
}

int launch_quantize(const float *in, int16_t *out, size_t n, float scale, hipStream_t stream)
{
    if (n == 0) return SYN_OK;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(quantize_i16, dim3((unsigned)blocks), dim3(256), 0, stream, in, out, n, scale);
    return hipGetLastError() == hipSuccess ? SYN_OK : SYN_DEVICE_ERROR;
}

}  // namespace syn
