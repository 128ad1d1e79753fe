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
//   * the multi-worker pipeline: a fixed slot pool, one SPSC work ring and
//     one SPSC done ring per worker, least-outstanding dispatch with a
//     rotating tie-break, spinning workers pinned to cores
//     (src/ldpc_pipeline.rs:98-589, src/spsc_queue.rs:171-376).
// Results are bit-identical to the scalar oracle (tests check it).
#include <immintrin.h>
#include <pthread.h>
#include <sched.h>
#include <string.h>

#include <atomic>
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

// Run `count` frames (llr [count][n], prepared as decode_5g inputs when
// init5g) through an n_workers pipeline; fills iters [count] and hard
// [count][n]; returns wall-clock nanoseconds for the whole stream.
long long cpb_pipeline_run(const orc_code_t *c, float beta, const float *llr_in, long count, long iterations,
                           int n_workers, int init5g, long n_filler, int *iters, uint8_t *hard_out)
{
    const size_t n = (size_t)c->n_b * c->z, ez = (size_t)c->n_edges * c->z;
    const int pool = n_workers * 2 > 16 ? n_workers * 2 : 16;
    struct Slot {
        std::vector<float> llr, er, q;
        std::vector<uint8_t> hard;
        long used = 0, tag = 0;
    };
    std::vector<Slot> slots(pool);
    for (auto &s : slots) {
        s.llr.resize(n);
        s.er.resize(ez);
        s.q.resize((size_t)c->max_deg * c->z);
        s.hard.resize(n);
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
                    s.used = decode_avx2(c, beta, s.llr.data(), s.er.data(), s.q.data(), s.hard.data(), iterations);
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
            memcpy(s.llr.data(), llr_in + (size_t)submitted * n, n * sizeof(float));
            if (init5g) init_5g(c, s.llr.data(), n_filler);
            s.tag = submitted;
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
                if (iters) iters[s.tag] = (int)s.used;
                if (hard_out) memcpy(hard_out + (size_t)s.tag * n, s.hard.data(), n);
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

}  // extern "C"
