// ORACLE side (CPU baseline) -- test / benchmark infrastructure, not product.
//
// C++ restatement of the reference's *CPU* decode path, for timing on the
// GPU host's cores beside the GPU decoders:
//   * the AVX2 f32 kernels: Q-build by two contiguous spans
//     (q_build_edge_avx2, src/simd_avx2.rs:464-507) and the 8-lane
//     min1/min2/sign passes with scalar posterior scatter
//     (decode_layer_passes_avx2, src/simd_avx2.rs:78-286), driven by the
//     layer loop of decode_layered_offset_min_sum_dispatch
//     (src/qc_ldpc.rs:962-1221) with check_syndrome_f32 (1231-1252);
//   * the AVX2 int8 kernels: 16-lane Q-build with a widened saturating
//     subtract (q_build_edge_i8_avx2, src/simd_avx2.rs:591-668), 32-lane
//     min1/min2/sign passes and the widened delta update with one scalar
//     group per wrapping edge (decode_layer_passes_i8_avx2 and
//     apply_delta16_i8_avx2, src/simd_avx2.rs:546-577, 683-847), driven by
//     decode_layered_offset_min_sum_i8_dispatch (src/qc_ldpc.rs:1530-1758)
//     with check_syndrome_i16 (1766-1787), fed by quantize_llr_i16
//     (src/quantize.rs:331-341) and the decode_5g_i8 prologue
//     (src/qc_ldpc.rs:1486-1523);
//   * the multi-worker pipeline: a fixed slot pool, one SPSC work ring and
//     one SPSC done ring per worker, least-outstanding dispatch with a
//     rotating tie-break, spinning workers pinned to cores
//     (src/ldpc_pipeline.rs:98-589, src/spsc_queue.rs:171-376); a stream may
//     mix codes (each frame names its code), which is how the transport-block
//     chain feeds it (transport_block.rs:924-983).
// Results are bit-identical to the scalar oracle; tests/test_cpu_baseline.py
// checks the kernels and the pipeline against it.
#include <immintrin.h>
#include <pthread.h>
#include <sched.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cfloat>
#include <chrono>
#include <thread>
#include <vector>

#include "loms_oracle.h"

namespace {

inline int vix(const orc_code_t *c, int e, int zi)
{
    int s = zi + c->shifts[e];
    return c->cols[e] * c->z + (s >= c->z ? s - c->z : s);
}

void qbuild_avx2(int z, int shift, int var_base, int q_base, int base_edge, const float *llr, const float *er, float *q)
{
    const int span1 = z - shift;
    int i = 0;
    for (; i + 8 <= span1; i += 8)
        _mm256_storeu_ps(q + q_base + i, _mm256_sub_ps(_mm256_loadu_ps(llr + var_base + shift + i),
                                                       _mm256_loadu_ps(er + base_edge + i)));
    for (; i < span1; ++i) q[q_base + i] = llr[var_base + shift + i] - er[base_edge + i];
    for (i = 0; i + 8 <= shift; i += 8)
        _mm256_storeu_ps(q + q_base + span1 + i,
                         _mm256_sub_ps(_mm256_loadu_ps(llr + var_base + i), _mm256_loadu_ps(er + base_edge + span1 + i)));
    for (; i < shift; ++i) q[q_base + span1 + i] = llr[var_base + i] - er[base_edge + span1 + i];
}

struct alignas(64) Scratch {
    float m1[ORC_Z_MAX], m2[ORC_Z_MAX];
    uint32_t sx[ORC_Z_MAX];
};

void layer_passes_avx2(const orc_code_t *c, int lb, int deg, float beta, const float *q, float *er, float *llr,
                       Scratch &s)
{
    const int z = c->z, full = z & ~7;
    const __m256 absm = _mm256_castsi256_ps(_mm256_set1_epi32(0x7fffffff));
    const __m256i sgnm = _mm256_set1_epi32((int)0x80000000u);
    const __m256 bv = _mm256_set1_ps(beta), zero = _mm256_setzero_ps();
    for (int i = 0; i < z; ++i) {
        s.m1[i] = FLT_MAX;
        s.m2[i] = FLT_MAX;
        s.sx[i] = 0;
    }
    for (int e = 0; e < deg; ++e) {
        const float *qe = q + e * z;
        for (int i = 0; i < full; i += 8) {
            const __m256 qv = _mm256_loadu_ps(qe + i);
            const __m256 av = _mm256_and_ps(qv, absm);
            __m256i sx = _mm256_load_si256((const __m256i *)(s.sx + i));
            _mm256_store_si256((__m256i *)(s.sx + i), _mm256_xor_si256(sx, _mm256_and_si256(_mm256_castps_si256(qv), sgnm)));
            const __m256 m1 = _mm256_load_ps(s.m1 + i), m2 = _mm256_load_ps(s.m2 + i);
            const __m256 le = _mm256_cmp_ps(av, m1, _CMP_LE_OQ);
            __m256 n2 = _mm256_blendv_ps(m2, m1, le);
            const __m256 n1 = _mm256_blendv_ps(m1, av, le);
            n2 = _mm256_blendv_ps(n2, av, _mm256_andnot_ps(le, _mm256_cmp_ps(av, m2, _CMP_LT_OQ)));
            _mm256_store_ps(s.m1 + i, n1);
            _mm256_store_ps(s.m2 + i, n2);
        }
        for (int i = full; i < z; ++i) {
            uint32_t b;
            memcpy(&b, qe + i, 4);
            const uint32_t ab = b & 0x7fffffffu;
            float a;
            memcpy(&a, &ab, 4);
            s.sx[i] ^= b & 0x80000000u;
            if (a <= s.m1[i]) {
                s.m2[i] = s.m1[i];
                s.m1[i] = a;
            } else if (a < s.m2[i]) s.m2[i] = a;
        }
    }
    for (int i = 0; i < z; ++i) {
        if (s.m1[i] == FLT_MAX) s.m1[i] = 0.f;
        if (s.m2[i] == FLT_MAX) s.m2[i] = s.m1[i];
    }
    for (int e = 0; e < deg; ++e) {
        const int ge = lb + e, base = ge * z;
        const float *qe = q + e * z;
        for (int i = 0; i < full; i += 8) {
            const __m256 qv = _mm256_loadu_ps(qe + i);
            const __m256 av = _mm256_and_ps(qv, absm);
            const __m256i sx = _mm256_load_si256((const __m256i *)(s.sx + i));
            const __m256 es = _mm256_castsi256_ps(_mm256_xor_si256(sx, _mm256_and_si256(_mm256_castps_si256(qv), sgnm)));
            const __m256 m1 = _mm256_load_ps(s.m1 + i), m2 = _mm256_load_ps(s.m2 + i);
            const __m256 mex = _mm256_blendv_ps(m1, m2, _mm256_cmp_ps(av, m1, _CMP_EQ_OQ));
            const __m256 nr = _mm256_or_ps(_mm256_max_ps(_mm256_sub_ps(mex, bv), zero), es);
            const __m256 orv = _mm256_loadu_ps(er + base + i);
            _mm256_storeu_ps(er + base + i, nr);
            alignas(32) float d[8];
            _mm256_store_ps(d, _mm256_sub_ps(nr, orv));
            for (int j = 0; j < 8; ++j) llr[vix(c, ge, i + j)] += d[j];
        }
        for (int i = full; i < z; ++i) {
            uint32_t b;
            memcpy(&b, qe + i, 4);
            const uint32_t ab = b & 0x7fffffffu;
            float a;
            memcpy(&a, &ab, 4);
            const uint32_t sg = s.sx[i] ^ (b & 0x80000000u);
            const float me = (a == s.m1[i]) ? s.m2[i] : s.m1[i];
            float mag = me - beta;
            mag = mag < 0.f ? 0.f : mag;
            uint32_t rb;
            memcpy(&rb, &mag, 4);
            rb |= sg;
            float nr;
            memcpy(&nr, &rb, 4);
            const float o = er[base + i];
            er[base + i] = nr;
            llr[vix(c, ge, i)] += nr - o;
        }
    }
}

// init_5g_llr (src/qc_ldpc.rs:711-732)
void init_5g(const orc_code_t *c, float *llr, long n_filler)
{
    const long k = (long)(c->n_b - c->m_b) * c->z;
    const long kp = n_filler >= k ? 0 : k - n_filler;
    for (long i = kp; i < k; ++i) llr[i] = 1000000.0f;
    for (long i = 0; i < 2L * c->z; ++i) llr[i] = 0.0f;
}

bool syndrome_ok(const orc_code_t *c, const float *llr)
{
    for (int l = 0; l < c->m_b; ++l)
        for (int zi = 0; zi < c->z; ++zi) {
            unsigned p = 0;
            for (int e = c->layer_off[l]; e < c->layer_off[l + 1]; ++e) p ^= llr[vix(c, e, zi)] < 0.f;
            if (p) return false;
        }
    return true;
}

long decode_avx2(const orc_code_t *c, float beta, float *llr, float *er, float *q, uint8_t *hard, long iterations)
{
    const int z = c->z;
    const size_t n = (size_t)c->n_b * z;
    memset(er, 0, sizeof(float) * (size_t)c->n_edges * z);
    Scratch s;
    long used = 0;
    for (long it = 0; it < iterations; ++it) {
        ++used;
        for (int l = 0; l < c->m_b; ++l) {
            const int lb = c->layer_off[l], deg = c->layer_off[l + 1] - lb;
            for (int e = 0; e < deg; ++e)
                qbuild_avx2(z, c->shifts[lb + e], c->cols[lb + e] * z, e * z, (lb + e) * z, llr, er, q);
            layer_passes_avx2(c, lb, deg, beta, q, er, llr, s);
        }
        if (syndrome_ok(c, llr)) break;
    }
    for (size_t i = 0; i < n; ++i) hard[i] = llr[i] < 0.f;
    return used;
}

// ---------------------------------------------------------------------------
// int8 path
// ---------------------------------------------------------------------------

constexpr int MSG_MAX = 127;

// apply_delta16_i8_avx2 (simd_avx2.rs:546-577): 16 widened deltas added to
// the posterior at z positions z_idx.. with a saturating add and the +-C
// clamp; the one group per edge whose indices wrap goes scalar.
inline void apply_delta16(int16_t *app, int var_base, int shift, int z, int z_idx, __m256i delta, __m256i lo,
                          __m256i hi, int16_t clamp)
{
    const int s = z_idx + shift;
    const int start = s >= z ? s - z : s;
    if (start + 16 <= z) {
        __m256i *p = (__m256i *)(app + var_base + start);
        const __m256i v = _mm256_adds_epi16(_mm256_loadu_si256(p), delta);
        _mm256_storeu_si256(p, _mm256_min_epi16(_mm256_max_epi16(v, lo), hi));
    } else {
        alignas(32) int16_t d[16];
        _mm256_store_si256((__m256i *)d, delta);
        for (int j = 0; j < 16; ++j) {
            const int sj = start + j;
            int16_t &slot = app[var_base + (sj >= z ? sj - z : sj)];
            int v = (int)slot + (int)d[j];  // saturating_add then clamp
            v = v < -32768 ? -32768 : (v > 32767 ? 32767 : v);
            slot = (int16_t)(v < -clamp ? -clamp : (v > clamp ? clamp : v));
        }
    }
}

// q_build_edge_i8_avx2 (simd_avx2.rs:591-668)
void qbuild_i8_avx2(int z, int shift, int var_base, int q_base, int base_edge, const int16_t *app, const int8_t *er,
                    int8_t *q)
{
    const __m256i mhi = _mm256_set1_epi16(MSG_MAX), mlo = _mm256_set1_epi16(-MSG_MAX);
    const int full = z & ~15;
    int i = 0;
    for (; i < full; i += 16) {
        const int s = i + shift;
        const int start = s >= z ? s - z : s;
        __m256i a;
        if (start + 16 <= z) {
            a = _mm256_loadu_si256((const __m256i *)(app + var_base + start));
        } else {
            alignas(32) int16_t t[16];
            for (int j = 0; j < 16; ++j) {
                const int sj = start + j;
                t[j] = app[var_base + (sj >= z ? sj - z : sj)];
            }
            a = _mm256_load_si256((const __m256i *)t);
        }
        const __m256i r = _mm256_cvtepi8_epi16(_mm_loadu_si128((const __m128i *)(er + base_edge + i)));
        const __m256i dc = _mm256_min_epi16(_mm256_max_epi16(_mm256_subs_epi16(a, r), mlo), mhi);
        const __m256i pk = _mm256_packs_epi16(dc, dc);
        _mm_storel_epi64((__m128i *)(q + q_base + i), _mm256_castsi256_si128(pk));
        _mm_storel_epi64((__m128i *)(q + q_base + i + 8), _mm256_extracti128_si256(pk, 1));
    }
    for (int j = i; j < z; ++j) {
        const int s = j + shift;
        const int d = (int)app[var_base + (s >= z ? s - z : s)] - (int)er[base_edge + j];
        q[q_base + j] = (int8_t)(d < -MSG_MAX ? -MSG_MAX : (d > MSG_MAX ? MSG_MAX : d));
    }
}

struct alignas(64) ScratchI8 {
    int8_t m1[ORC_Z_MAX], m2[ORC_Z_MAX];
    uint8_t sx[ORC_Z_MAX];
};

inline int8_t sat_abs8(int8_t v) { return v == -128 ? 127 : (int8_t)(v < 0 ? -v : v); }

// decode_layer_passes_i8_avx2 (simd_avx2.rs:683-847), row degree >= 2
void layer_passes_i8_avx2(const orc_code_t *c, int lb, int deg, int8_t beta_q, int16_t clamp, const int8_t *q,
                          int8_t *er, int16_t *app, ScratchI8 &s)
{
    const int z = c->z, full = z & ~31;
    const __m256i sgn = _mm256_set1_epi8((char)0x80), mmax = _mm256_set1_epi8(MSG_MAX), ones = _mm256_set1_epi8(1);
    const __m256i bv = _mm256_set1_epi8(beta_q), zero = _mm256_setzero_si256();
    const __m256i chi = _mm256_set1_epi16(clamp), clo = _mm256_set1_epi16((int16_t)-clamp);
    for (int i = 0; i < full; i += 32) {
        _mm256_storeu_si256((__m256i *)(s.m1 + i), mmax);
        _mm256_storeu_si256((__m256i *)(s.m2 + i), mmax);
        _mm256_storeu_si256((__m256i *)(s.sx + i), zero);
    }
    for (int i = full; i < z; ++i) {
        s.m1[i] = MSG_MAX;
        s.m2[i] = MSG_MAX;
        s.sx[i] = 0;
    }
    for (int e = 0; e < deg; ++e) {
        const int8_t *qe = q + e * z;
        for (int i = 0; i < full; i += 32) {
            const __m256i qv = _mm256_loadu_si256((const __m256i *)(qe + i));
            const __m256i a = _mm256_abs_epi8(qv);
            __m256i *sp = (__m256i *)(s.sx + i);
            _mm256_storeu_si256(sp, _mm256_xor_si256(_mm256_loadu_si256(sp), _mm256_and_si256(qv, sgn)));
            __m256i *p1 = (__m256i *)(s.m1 + i), *p2 = (__m256i *)(s.m2 + i);
            const __m256i m1 = _mm256_loadu_si256(p1), m2 = _mm256_loadu_si256(p2);
            _mm256_storeu_si256(p1, _mm256_min_epi8(m1, a));
            _mm256_storeu_si256(p2, _mm256_min_epi8(m2, _mm256_max_epi8(m1, a)));
        }
        for (int i = full; i < z; ++i) {
            const int8_t qq = qe[i], a = sat_abs8(qq), m1 = s.m1[i];
            s.sx[i] ^= (uint8_t)qq & 0x80;
            s.m1[i] = m1 < a ? m1 : a;
            const int8_t mx = m1 > a ? m1 : a;
            s.m2[i] = s.m2[i] < mx ? s.m2[i] : mx;
        }
    }
    for (int e = 0; e < deg; ++e) {
        const int ge = lb + e, base = ge * z, var_base = c->cols[ge] * z, shift = c->shifts[ge];
        const int8_t *qe = q + e * z;
        int8_t *rp = er + base;
        for (int i = 0; i < full; i += 32) {
            const __m256i qv = _mm256_loadu_si256((const __m256i *)(qe + i));
            const __m256i a = _mm256_abs_epi8(qv);
            const __m256i m1 = _mm256_loadu_si256((const __m256i *)(s.m1 + i));
            const __m256i m2 = _mm256_loadu_si256((const __m256i *)(s.m2 + i));
            const __m256i sx = _mm256_loadu_si256((const __m256i *)(s.sx + i));
            const __m256i mex = _mm256_blendv_epi8(m1, m2, _mm256_cmpeq_epi8(a, m1));
            const __m256i mag = _mm256_max_epi8(_mm256_sub_epi8(mex, bv), zero);
            const __m256i se = _mm256_xor_si256(sx, _mm256_and_si256(qv, sgn));
            const __m256i nr = _mm256_sign_epi8(mag, _mm256_or_si256(se, ones));
            const __m256i orv = _mm256_loadu_si256((const __m256i *)(rp + i));
            _mm256_storeu_si256((__m256i *)(rp + i), nr);
            const __m256i dlo = _mm256_sub_epi16(_mm256_cvtepi8_epi16(_mm256_castsi256_si128(nr)),
                                                 _mm256_cvtepi8_epi16(_mm256_castsi256_si128(orv)));
            const __m256i dhi = _mm256_sub_epi16(_mm256_cvtepi8_epi16(_mm256_extracti128_si256(nr, 1)),
                                                 _mm256_cvtepi8_epi16(_mm256_extracti128_si256(orv, 1)));
            apply_delta16(app, var_base, shift, z, i, dlo, clo, chi, clamp);
            apply_delta16(app, var_base, shift, z, i + 16, dhi, clo, chi, clamp);
        }
        for (int i = full; i < z; ++i) {
            const int8_t qq = qe[i], a = sat_abs8(qq);
            const uint8_t se = s.sx[i] ^ ((uint8_t)qq & 0x80);
            const int8_t mex = (a == s.m1[i]) ? s.m2[i] : s.m1[i];
            int mag = (int)mex - (int)beta_q;
            mag = mag < 0 ? 0 : mag;
            const int neg = ((int8_t)se) >> 7;
            const int8_t nr = (int8_t)((mag ^ neg) - neg);
            const int8_t o = rp[i];
            rp[i] = nr;
            const int st = i + shift;
            int16_t &slot = app[var_base + (st >= z ? st - z : st)];
            int v = (int)slot + ((int)nr - (int)o);
            v = v < -32768 ? -32768 : (v > 32767 ? 32767 : v);
            slot = (int16_t)(v < -clamp ? -clamp : (v > clamp ? clamp : v));
        }
    }
}

bool syndrome_ok_i16(const orc_code_t *c, const int16_t *app)
{
    for (int l = 0; l < c->m_b; ++l)
        for (int zi = 0; zi < c->z; ++zi) {
            unsigned p = 0;
            for (int e = c->layer_off[l]; e < c->layer_off[l + 1]; ++e) p ^= app[vix(c, e, zi)] < 0;
            if (p) return false;
        }
    return true;
}

// decode_layered_offset_min_sum_i8_dispatch (qc_ldpc.rs:1530-1758) with the
// AVX2 kernels; every 3GPP and 802.11 row has degree >= 2 (a code with a
// shorter row is refused: the reference routes those rows to its scalar
// kernel, which the baseline does not need).
long decode_i8_avx2(const orc_code_t *c, int8_t beta_q, int16_t app_clamp, int16_t *app, int8_t *er, int8_t *q,
                    uint8_t *hard, long iterations)
{
    for (int l = 0; l < c->m_b; ++l)
        if (c->layer_off[l + 1] - c->layer_off[l] < 2) return -1;
    const int z = c->z;
    const size_t n = (size_t)c->n_b * z;
    const int16_t C = app_clamp < 0 ? 0 : app_clamp;
    memset(er, 0, (size_t)c->n_edges * z);
    for (size_t i = 0; i < n; ++i) app[i] = app[i] < -C ? (int16_t)-C : (app[i] > C ? C : app[i]);
    ScratchI8 s;
    long used = 0;
    for (long it = 0; it < iterations; ++it) {
        ++used;
        for (int l = 0; l < c->m_b; ++l) {
            const int lb = c->layer_off[l], deg = c->layer_off[l + 1] - lb;
            for (int e = 0; e < deg; ++e)
                qbuild_i8_avx2(z, c->shifts[lb + e], c->cols[lb + e] * z, e * z, (lb + e) * z, app, er, q);
            layer_passes_i8_avx2(c, lb, deg, beta_q, C, q, er, app, s);
        }
        if (syndrome_ok_i16(c, app)) break;
    }
    for (size_t i = 0; i < n; ++i) hard[i] = app[i] < 0;
    return used;
}

// quantize_llr_i16 (quantize.rs:331-341): v = l * s, NaN -> 0, round half
// away from zero, clamp to +-127
void quantize_i16(const float *llr, int16_t *out, size_t n, float scale)
{
    for (size_t i = 0; i < n; ++i) {
        const float v = llr[i] * scale;
        if (v != v) {
            out[i] = 0;
            continue;
        }
        const float r = roundf(v);
        out[i] = (int16_t)(r < -127.f ? -127 : (r > 127.f ? 127 : (int)r));
    }
}

// decode_5g_i8's prologue (qc_ldpc.rs:1505-1514): fillers at the clamp,
// the 2Z punctured positions at 0
void init_5g_i8(const orc_code_t *c, int16_t *app, long n_filler, int16_t app_clamp)
{
    const long k = (long)(c->n_b - c->m_b) * c->z;
    const long kp = n_filler >= k ? 0 : k - n_filler;
    const int16_t f = app_clamp < 0 ? 0 : app_clamp;
    for (long i = kp; i < k; ++i) app[i] = f;
    for (long i = 0; i < 2L * c->z; ++i) app[i] = 0;
}

// SPSC ring of slot indices: one producer, one consumer, cache-line padded
// head and tail, acquire/release ordering (spsc_queue.rs:171-256).
struct Ring {
    static constexpr uint32_t CAP = 64;
    alignas(64) std::atomic<uint32_t> head{0};
    alignas(64) std::atomic<uint32_t> tail{0};
    alignas(64) uint32_t buf[CAP];
    bool push(uint32_t v)
    {
        const uint32_t t = tail.load(std::memory_order_relaxed);
        if (t - head.load(std::memory_order_acquire) == CAP) return false;
        buf[t % CAP] = v;
        tail.store(t + 1, std::memory_order_release);
        return true;
    }
    bool pop(uint32_t &v)
    {
        const uint32_t h = head.load(std::memory_order_relaxed);
        if (h == tail.load(std::memory_order_acquire)) return false;
        v = buf[h % CAP];
        head.store(h + 1, std::memory_order_release);
        return true;
    }
};

void pin(int core)
{
    cpu_set_t set;
    CPU_ZERO(&set);
    CPU_SET(core % CPU_SETSIZE, &set);
    (void)pthread_setaffinity_np(pthread_self(), sizeof(set), &set);
}

}  // namespace

extern "C" {

long cpb_decode_f32_avx2(const orc_code_t *c, float beta, float *llr, float *edge_r, uint8_t *hard, long iterations)
{
    std::vector<float> q((size_t)(c->max_deg > 0 ? c->max_deg : 1) * c->z);
    return decode_avx2(c, beta, llr, edge_r, q.data(), hard, iterations);
}

// decode_layered_offset_min_sum_i8 with the AVX2 kernels: app [N] in/out,
// edge_r [E*Z] out, hard [N] out; -1 for a code with a row of degree < 2.
long cpb_decode_i8_avx2(const orc_code_t *c, int8_t beta_q, int16_t app_clamp, int16_t *app, int8_t *edge_r,
                        uint8_t *hard, long iterations)
{
    std::vector<int8_t> q((size_t)(c->max_deg > 0 ? c->max_deg : 1) * c->z);
    return decode_i8_avx2(c, beta_q, app_clamp, app, edge_r, q.data(), hard, iterations);
}

// A stream of `count` frames through an n_workers pipeline.  Stream position
// i decodes frame order[i]: code codes[frame_code[f]], decoder input of N_f
// f32 LLRs at llr_in + in_off[f] (prepared as decode_5g inputs with
// n_filler[f] fillers when init5g; n_filler may be NULL).  int8 != 0: the
// worker quantizes (quantize_llr_i16 at `scale`), applies the decode_5g_i8
// prologue when init5g, and runs the int8 decode with beta_q / app_clamp;
// otherwise the f32 decode with beta.  Fills iters[f] and the N_f hard
// decisions at hard_out + in_off[f]; returns wall-clock nanoseconds for the
// whole stream.
long long cpb_pipeline_mixed(const orc_code_t *codes, const int *frame_code, const long long *in_off,
                             const long *n_filler, const int *order, long count, const float *llr_in, int int8,
                             float beta, float scale, int8_t beta_q, int16_t app_clamp, long iterations,
                             int n_workers, int init5g, int *iters, uint8_t *hard_out)
{
    size_t n_max = 0, ez_max = 0, q_max = 0;
    for (long i = 0; i < count; ++i) {
        const orc_code_t *c = codes + frame_code[order[i]];
        n_max = std::max(n_max, (size_t)c->n_b * c->z);
        ez_max = std::max(ez_max, (size_t)c->n_edges * c->z);
        q_max = std::max(q_max, (size_t)(c->max_deg > 0 ? c->max_deg : 1) * c->z);
    }
    const int pool = n_workers * 2 > 16 ? n_workers * 2 : 16;
    struct Slot {
        std::vector<float> llr, er, q;
        std::vector<int16_t> app;
        std::vector<int8_t> er8, q8;
        std::vector<uint8_t> hard;
        long used = 0;
        int frame = 0;
    };
    std::vector<Slot> slots(pool);
    for (auto &s : slots) {
        s.llr.resize(n_max);
        s.hard.resize(n_max);
        if (int8) {
            s.app.resize(n_max);
            s.er8.resize(ez_max);
            s.q8.resize(q_max);
        } else {
            s.er.resize(ez_max);
            s.q.resize(q_max);
        }
    }
    std::vector<Ring> work(n_workers), done(n_workers);
    std::atomic<bool> stop{false};
    std::vector<std::thread> th;
    for (int w = 0; w < n_workers; ++w)
        th.emplace_back([&, w] {
            pin(w);
            uint32_t idx;
            for (;;) {
                if (work[w].pop(idx)) {
                    Slot &s = slots[idx];
                    const orc_code_t *c = codes + frame_code[s.frame];
                    const size_t n = (size_t)c->n_b * c->z;
                    const long nf = n_filler ? n_filler[s.frame] : 0;
                    if (int8) {
                        quantize_i16(s.llr.data(), s.app.data(), n, scale);
                        if (init5g) init_5g_i8(c, s.app.data(), nf, app_clamp);
                        s.used = decode_i8_avx2(c, beta_q, app_clamp, s.app.data(), s.er8.data(), s.q8.data(),
                                                s.hard.data(), iterations);
                    } else {
                        if (init5g) init_5g(c, s.llr.data(), nf);
                        s.used = decode_avx2(c, beta, s.llr.data(), s.er.data(), s.q.data(), s.hard.data(),
                                             iterations);
                    }
                    while (!done[w].push(idx)) _mm_pause();
                } else if (stop.load(std::memory_order_acquire)) break;
                else _mm_pause();
            }
        });
    std::vector<uint32_t> free_list;
    for (int i = pool - 1; i >= 0; --i) free_list.push_back((uint32_t)i);
    std::vector<long> outstanding(n_workers, 0);
    unsigned long cursor = 0;
    long submitted = 0, received = 0;
    const auto t0 = std::chrono::steady_clock::now();
    while (received < count) {
        if (submitted < count && !free_list.empty()) {
            const uint32_t idx = free_list.back();
            Slot &s = slots[idx];
            s.frame = order[submitted];
            const orc_code_t *c = codes + frame_code[s.frame];
            memcpy(s.llr.data(), llr_in + in_off[s.frame], sizeof(float) * (size_t)c->n_b * c->z);
            // least outstanding, rotating tie-break (ldpc_pipeline.rs:267-283)
            int best = (int)(cursor % n_workers);
            for (int k = 1; k < n_workers; ++k) {
                const int wi = (int)((cursor + k) % n_workers);
                if (outstanding[wi] < outstanding[best]) best = wi;
            }
            if (work[best].push(idx)) {
                free_list.pop_back();
                outstanding[best]++;
                cursor++;
                submitted++;
            }
        }
        bool got = false;
        for (int w = 0; w < n_workers && !got; ++w) {
            uint32_t idx;
            if (done[w].pop(idx)) {
                outstanding[w]--;
                Slot &s = slots[idx];
                const orc_code_t *c = codes + frame_code[s.frame];
                if (iters) iters[s.frame] = (int)s.used;
                if (hard_out) memcpy(hard_out + in_off[s.frame], s.hard.data(), (size_t)c->n_b * c->z);
                free_list.push_back(idx);
                received++;
                got = true;
            }
        }
        if (!got) _mm_pause();
    }
    const auto t1 = std::chrono::steady_clock::now();
    stop.store(true, std::memory_order_release);
    for (auto &t : th) t.join();
    return (long long)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
}

// One code, frames llr [count][N] in stream order (the f32 pipeline of
// ldpc_pipeline_bench.rs): cpb_pipeline_mixed with frame i at i * N.
long long cpb_pipeline_run(const orc_code_t *c, float beta, const float *llr_in, long count, long iterations,
                           int n_workers, int init5g, long n_filler, int *iters, uint8_t *hard_out)
{
    const long long n = (long long)c->n_b * c->z;
    std::vector<int> fc((size_t)count, 0), order((size_t)count);
    std::vector<long long> off((size_t)count);
    std::vector<long> nf((size_t)count, n_filler);
    for (long i = 0; i < count; ++i) {
        order[i] = (int)i;
        off[i] = i * n;
    }
    return cpb_pipeline_mixed(c, fc.data(), off.data(), nf.data(), order.data(), count, llr_in, 0, beta, 0.f, 0, 0,
                              iterations, n_workers, init5g, iters, hard_out);
}

}  // extern "C"
