/*
 * ORACLE -- test infrastructure, not product code.  See loms_oracle.h.
 *
 * Scalar restatement of the reference's LOMS decode path.  Built with
 * -O2 -ffp-contract=off and no fast-math so that every f32 operation is the
 * single IEEE-754 round-to-nearest operation the reference performs.
 */
#include "loms_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../syndrome_amd/csrc/code_tables.inc"

/* ------------------------------------------------------------------ */
/* Code construction                                                   */
/* ------------------------------------------------------------------ */

/* LIFTING_SETS / ils_for_z (qc_ldpc.rs:103-125). */
static int ils_for_z(int z)
{
    static const int sets[8][8] = {
        {2, 4, 8, 16, 32, 64, 128, 256}, {3, 6, 12, 24, 48, 96, 192, 384},
        {5, 10, 20, 40, 80, 160, 320, 0}, {7, 14, 28, 56, 112, 224, 0, 0},
        {9, 18, 36, 72, 144, 288, 0, 0},  {11, 22, 44, 88, 176, 352, 0, 0},
        {13, 26, 52, 104, 208, 0, 0, 0},  {15, 30, 60, 120, 240, 0, 0, 0},
    };
    for (int s = 0; s < 8; ++s)
        for (int j = 0; j < 8; ++j)
            if (sets[s][j] == z && z != 0) return s;
    return -1;
}

/* build_layout (qc_ldpc.rs:281-318): group edges by row block, keeping the
 * input order within each row. */
static void build_layout(orc_code_t *c, const int *rows, const int *cols, const int *shifts, int n)
{
    int deg[ORC_MAX_ROWS] = {0};
    int fill[ORC_MAX_ROWS] = {0};
    for (int i = 0; i < n; ++i) deg[rows[i]]++;
    c->layer_off[0] = 0;
    c->max_deg = 0;
    for (int r = 0; r < c->m_b; ++r) {
        c->layer_off[r + 1] = c->layer_off[r] + deg[r];
        if (deg[r] > c->max_deg) c->max_deg = deg[r];
    }
    for (int i = 0; i < n; ++i) {
        int pos = c->layer_off[rows[i]] + fill[rows[i]]++;
        c->cols[pos] = cols[i];
        c->shifts[pos] = shifts[i];
    }
    c->n_edges = n;
}

int orc_code_3gpp(orc_code_t *c, int bg, int z)
{
    int ils = ils_for_z(z);
    if (ils < 0 || (bg != 0 && bg != 1)) return ORC_INVALID_PARAM;
    const unsigned char *deg = bg == 0 ? SYN_BG1_ROW_DEG : SYN_BG2_ROW_DEG;
    const unsigned char *ecol = bg == 0 ? SYN_BG1_EDGE_COL : SYN_BG2_EDGE_COL;
    const short(*ev)[8] = bg == 0 ? SYN_BG1_EDGE_V : SYN_BG2_EDGE_V;
    int m_b = bg == 0 ? SYN_BG1_ROWS : SYN_BG2_ROWS;
    int n = bg == 0 ? SYN_BG1_EDGES : SYN_BG2_EDGES;
    static int rows[ORC_MAX_EDGES], cols[ORC_MAX_EDGES], sh[ORC_MAX_EDGES];
    int e = 0;
    for (int r = 0; r < m_b; ++r)
        for (int j = 0; j < deg[r]; ++j, ++e) {
            rows[e] = r;
            cols[e] = ecol[e];
            sh[e] = ev[e][ils] % z; /* entry_col_shift, qc_ldpc.rs:134-140 */
        }
    memset(c, 0, sizeof(*c));
    c->z = z;
    c->m_b = m_b;
    c->n_b = bg == 0 ? SYN_BG1_COLS : SYN_BG2_COLS;
    build_layout(c, rows, cols, sh, n);
    return ORC_OK;
}

int orc_code_raw(orc_code_t *c, const int *rows, const int *cols, const int *shifts, int n, int m_b,
                 int n_b, int z)
{
    if (z <= 0 || z > ORC_Z_MAX || m_b < 0 || m_b > ORC_MAX_ROWS || n_b < 0 || n_b > ORC_MAX_COLS ||
        n < 0 || n > ORC_MAX_EDGES)
        return ORC_INVALID_PARAM;
    for (int i = 0; i < n; ++i)
        if (rows[i] < 0 || rows[i] >= m_b || cols[i] < 0 || cols[i] >= n_b || shifts[i] < 0 ||
            shifts[i] >= z)
            return ORC_INVALID_PARAM;
    memset(c, 0, sizeof(*c));
    c->z = z;
    c->m_b = m_b;
    c->n_b = n_b;
    build_layout(c, rows, cols, shifts, n);
    return ORC_OK;
}

int orc_code_wifi(orc_code_t *c, int z, int rate_num, int rate_den)
{
    for (int i = 0; i < SYN_WIFI_COUNT; ++i) {
        const syn_wifi_matrix_t *m = &SYN_WIFI_MATRICES[i];
        if (m->z == z && m->rate_num == rate_num && m->rate_den == rate_den) {
            int rows[ORC_MAX_EDGES], cols[ORC_MAX_EDGES], sh[ORC_MAX_EDGES];
            for (int e = 0; e < m->n_edges; ++e) {
                rows[e] = m->edges[e][0];
                cols[e] = m->edges[e][1];
                sh[e] = m->edges[e][2];
            }
            return orc_code_raw(c, rows, cols, sh, m->n_edges, m->row_blocks, m->col_blocks, z);
        }
    }
    return ORC_INVALID_PARAM;
}

size_t orc_n(const orc_code_t *c) { return (size_t)c->n_b * c->z; }
size_t orc_k(const orc_code_t *c) { return (size_t)(c->n_b - c->m_b) * c->z; }
size_t orc_edge_buf(const orc_code_t *c) { return (size_t)c->n_edges * c->z; }
size_t orc_layer_buf(const orc_code_t *c) { return (size_t)c->max_deg * c->z; }

static inline int var_index(const orc_code_t *c, int e, int z_idx)
{
    int s = z_idx + c->shifts[e];
    return c->cols[e] * c->z + (s >= c->z ? s - c->z : s);
}

/* ------------------------------------------------------------------ */
/* f32 decode                                                          */
/* ------------------------------------------------------------------ */

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* check_syndrome_f32 (qc_ldpc.rs:1231-1252). */
static int syndrome_ok_f32(const orc_code_t *c, const float *llr)
{
    for (int l = 0; l < c->m_b; ++l)
        for (int z = 0; z < c->z; ++z) {
            unsigned p = 0;
            for (int e = c->layer_off[l]; e < c->layer_off[l + 1]; ++e) p ^= (llr[var_index(c, e, z)] < 0.0f);
            if (p) return 0;
        }
    return 1;
}

int orc_init_5g_f32(const orc_code_t *c, float *llr, size_t llr_len, size_t n_filler)
{
    /* qc_ldpc.rs:711-732 */
    if (llr_len != orc_n(c)) return ORC_INVALID_PARAM;
    size_t k = orc_k(c);
    size_t kp = n_filler > k ? 0 : k - n_filler;
    for (size_t i = kp; i < k; ++i) llr[i] = 1000000.0f;
    for (size_t i = 0; i < (size_t)(2 * c->z); ++i) llr[i] = 0.0f;
    return ORC_OK;
}

long orc_decode_f32(const orc_code_t *c, float beta, float *llr, size_t llr_len, float *edge_r,
                    size_t edge_len, uint8_t *hard, size_t hard_len, size_t iterations, float *trace)
{
    const size_t n = orc_n(c);
    /* Validation order: qc_ldpc.rs:980-1004. */
    if (llr_len != n) return ORC_INVALID_PARAM;
    if (hard_len != n) return ORC_INVALID_PARAM;
    if (edge_len < orc_edge_buf(c)) return ORC_BUFFER_TOO_SMALL;
    const int z = c->z;
    for (size_t i = 0; i < orc_edge_buf(c); ++i) edge_r[i] = 0.0f; /* 1006-1009 */

    float *q = (float *)malloc(sizeof(float) * (size_t)(c->max_deg > 0 ? c->max_deg : 1) * z);
    float min1[ORC_Z_MAX], min2[ORC_Z_MAX];
    uint32_t sxor[ORC_Z_MAX];
    size_t used = 0;
    for (size_t it = 0; it < iterations; ++it) {
        used++;
        for (int l = 0; l < c->m_b; ++l) {
            const int lb = c->layer_off[l], deg = c->layer_off[l + 1] - lb;
            /* Q-build, qc_ldpc.rs:1072-1094 */
            for (int e = 0; e < deg; ++e)
                for (int zi = 0; zi < z; ++zi)
                    q[e * z + zi] = llr[var_index(c, lb + e, zi)] - edge_r[(size_t)(lb + e) * z + zi];
            /* Pass 1, qc_ldpc.rs:1155-1178 */
            for (int zi = 0; zi < z; ++zi) { min1[zi] = FLT_MAX; min2[zi] = FLT_MAX; sxor[zi] = 0; }
            for (int e = 0; e < deg; ++e)
                for (int zi = 0; zi < z; ++zi) {
                    uint32_t b = f2u(q[e * z + zi]);
                    float a = u2f(b & 0x7FFFFFFFu);
                    sxor[zi] ^= b & 0x80000000u;
                    float m1 = min1[zi];
                    if (a <= m1) { min2[zi] = m1; min1[zi] = a; }
                    else if (a < min2[zi]) min2[zi] = a;
                }
            for (int zi = 0; zi < z; ++zi) {
                if (min1[zi] == FLT_MAX) min1[zi] = 0.0f;
                if (min2[zi] == FLT_MAX) min2[zi] = min1[zi];
            }
            /* Pass 2, qc_ldpc.rs:1181-1204 */
            for (int e = 0; e < deg; ++e)
                for (int zi = 0; zi < z; ++zi) {
                    uint32_t qb = f2u(q[e * z + zi]);
                    float a = u2f(qb & 0x7FFFFFFFu);
                    uint32_t sgn = sxor[zi] ^ (qb & 0x80000000u);
                    float me = (a == min1[zi]) ? min2[zi] : min1[zi];
                    float mag = me - beta;
                    mag = (mag < 0.0f) ? 0.0f : mag;
                    float nr = u2f(f2u(mag) | sgn);
                    size_t ri = (size_t)(lb + e) * z + zi;
                    float orr = edge_r[ri];
                    edge_r[ri] = nr;
                    int v = var_index(c, lb + e, zi);
                    llr[v] += nr - orr;
                }
        }
        if (trace) memcpy(trace + it * n, llr, n * sizeof(float)); /* on_iteration, 1208 */
        if (syndrome_ok_f32(c, llr)) break;                        /* 1211 */
    }
    for (size_t i = 0; i < n; ++i) hard[i] = llr[i] < 0.0f; /* 1216-1218 */
    free(q);
    return (long)used;
}

/* ------------------------------------------------------------------ */
/* fixed point                                                         */
/* ------------------------------------------------------------------ */

/* f32::round: half away from zero (float_ext.rs:20-22 -> roundf). */
int8_t orc_beta_q(float beta, float scale)
{
    float s = beta * scale; /* quantize.rs:298-304 */
    if (!isfinite(s)) return 127;
    float r = roundf(s);
    if (r < 0.0f) r = 0.0f;
    if (r > 127.0f) r = 127.0f;
    return (int8_t)r;
}

static inline int32_t sat_f2i(float r)
{
    /* Rust `as i32` saturates (NaN handled by the callers). */
    if (r >= 2147483647.0f) return INT32_MAX;
    if (r <= -2147483648.0f) return INT32_MIN;
    return (int32_t)r;
}

void orc_quantize_llr_i16(const float *llr, int16_t *out, size_t n, float scale)
{
    for (size_t i = 0; i < n; ++i) { /* quantize.rs:331-341 */
        float v = llr[i] * scale;
        if (isnan(v)) { out[i] = 0; continue; }
        int32_t q = sat_f2i(roundf(v));
        out[i] = (int16_t)(q < -127 ? -127 : q > 127 ? 127 : q);
    }
}

void orc_quantize_llr_i8(const float *llr, int8_t *out, size_t n, float scale)
{
    for (size_t i = 0; i < n; ++i) { /* quantize.rs:148-154; `as i32` maps NaN to 0 */
        float v = roundf(llr[i] * scale);
        int32_t q = isnan(v) ? 0 : sat_f2i(v);
        out[i] = (int8_t)(q < -127 ? -127 : q > 127 ? 127 : q);
    }
}

void orc_dequantize_llr_i8(const int8_t *q, float *out, size_t n, float scale)
{
    const float inv = 1.0f / scale; /* quantize.rs:163-169 */
    for (size_t i = 0; i < n; ++i) out[i] = (float)q[i] * inv;
}

int orc_init_5g_i8(const orc_code_t *c, int16_t *app, size_t app_len, size_t n_filler, int16_t app_clamp)
{
    /* qc_ldpc.rs:1495-1514 */
    if (app_len != orc_n(c)) return ORC_INVALID_PARAM;
    size_t k = orc_k(c);
    size_t kp = n_filler > k ? 0 : k - n_filler;
    int16_t f = app_clamp < 0 ? 0 : app_clamp;
    for (size_t i = kp; i < k; ++i) app[i] = f;
    for (size_t i = 0; i < (size_t)(2 * c->z); ++i) app[i] = 0;
    return ORC_OK;
}

/* check_syndrome_i16 (qc_ldpc.rs:1766-1787). */
static int syndrome_ok_i16(const orc_code_t *c, const int16_t *app)
{
    for (int l = 0; l < c->m_b; ++l)
        for (int z = 0; z < c->z; ++z) {
            unsigned p = 0;
            for (int e = c->layer_off[l]; e < c->layer_off[l + 1]; ++e) p ^= (app[var_index(c, e, z)] < 0);
            if (p) return 0;
        }
    return 1;
}

static inline int16_t clamp16(int32_t v, int32_t lo, int32_t hi) { return (int16_t)(v < lo ? lo : v > hi ? hi : v); }

long orc_decode_i8(const orc_code_t *c, float beta, float scale, int16_t app_clamp_in, int16_t *app,
                   size_t app_len, int8_t *edge_r, size_t edge_len, uint8_t *hard, size_t hard_len,
                   size_t iterations)
{
    const size_t n = orc_n(c);
    if (app_len != n) return ORC_INVALID_PARAM; /* 1544-1568 */
    if (hard_len != n) return ORC_INVALID_PARAM;
    if (edge_len < orc_edge_buf(c)) return ORC_BUFFER_TOO_SMALL;
    const int z = c->z;
    const int8_t bq = orc_beta_q(beta, scale);                      /* 1571 */
    const int32_t C = app_clamp_in < 0 ? 0 : app_clamp_in;          /* 1574 */
    for (size_t i = 0; i < orc_edge_buf(c); ++i) edge_r[i] = 0;     /* 1579-1581 */
    for (size_t i = 0; i < n; ++i) app[i] = clamp16(app[i], -C, C); /* 1582-1584 */

    int8_t *q = (int8_t *)malloc((size_t)(c->max_deg > 0 ? c->max_deg : 1) * z);
    int8_t min1[ORC_Z_MAX], min2[ORC_Z_MAX];
    uint8_t sxor[ORC_Z_MAX];
    size_t used = 0;
    for (size_t it = 0; it < iterations; ++it) {
        used++;
        for (int l = 0; l < c->m_b; ++l) {
            const int lb = c->layer_off[l], deg = c->layer_off[l + 1] - lb;
            for (int e = 0; e < deg; ++e) /* Q-build 1619-1642 */
                for (int zi = 0; zi < z; ++zi) {
                    int32_t d = (int32_t)app[var_index(c, lb + e, zi)] - (int32_t)edge_r[(size_t)(lb + e) * z + zi];
                    q[e * z + zi] = (int8_t)(d < -127 ? -127 : d > 127 ? 127 : d);
                }
            for (int zi = 0; zi < z; ++zi) { min1[zi] = 127; min2[zi] = 127; sxor[zi] = 0; }
            for (int e = 0; e < deg; ++e) /* pass 1, 1670-1690 */
                for (int zi = 0; zi < z; ++zi) {
                    int8_t qq = q[e * z + zi];
                    int8_t a = (int8_t)(qq < 0 ? -qq : qq); /* q >= -127, so no overflow */
                    sxor[zi] ^= (uint8_t)qq & 0x80;
                    int8_t m1 = min1[zi];
                    min1[zi] = m1 < a ? m1 : a;
                    int8_t mx = m1 > a ? m1 : a;
                    min2[zi] = min2[zi] < mx ? min2[zi] : mx;
                }
            if (deg < 2) /* 1692-1706 */
                for (int zi = 0; zi < z; ++zi) {
                    if (deg == 0) min1[zi] = 0;
                    min2[zi] = min1[zi];
                }
            for (int e = 0; e < deg; ++e) /* pass 2, 1708-1744 */
                for (int zi = 0; zi < z; ++zi) {
                    int8_t qq = q[e * z + zi];
                    int8_t a = (int8_t)(qq < 0 ? -qq : qq);
                    uint8_t se = sxor[zi] ^ ((uint8_t)qq & 0x80);
                    int8_t me = (a == min1[zi]) ? min2[zi] : min1[zi];
                    int32_t mag = (int32_t)me - (int32_t)bq;
                    if (mag < 0) mag = 0;
                    int8_t nr = (int8_t)(se ? -mag : mag);
                    size_t ri = (size_t)(lb + e) * z + zi;
                    int8_t orr = edge_r[ri];
                    edge_r[ri] = nr;
                    int32_t delta = (int32_t)nr - (int32_t)orr;
                    int v = var_index(c, lb + e, zi);
                    int32_t s = (int32_t)app[v] + delta; /* saturating_add on i16 */
                    s = s < -32768 ? -32768 : s > 32767 ? 32767 : s;
                    app[v] = clamp16(s, -C, C);
                }
        }
        if (syndrome_ok_i16(c, app)) break; /* 1748 */
    }
    for (size_t i = 0; i < n; ++i) hard[i] = app[i] < 0; /* 1753-1755 */
    free(q);
    return (long)used;
}

/* ------------------------------------------------------------------ */
/* misc                                                                */
/* ------------------------------------------------------------------ */

int orc_syndrome_ok_bits(const orc_code_t *c, const uint8_t *bits)
{
    for (int l = 0; l < c->m_b; ++l)
        for (int z = 0; z < c->z; ++z) {
            unsigned p = 0;
            for (int e = c->layer_off[l]; e < c->layer_off[l + 1]; ++e) p ^= bits[var_index(c, e, z)] & 1;
            if (p) return 0;
        }
    return 1;
}

uint64_t orc_fnv1a_hard(const uint8_t *hard, size_t n)
{
    uint64_t h = 0xcbf29ce484222325ULL;
    for (size_t i = 0; i < n; ++i) {
        h ^= (uint64_t)(hard[i] & 1);
        h *= 0x00000100000001b3ULL;
    }
    return h;
}

void orc_knuth_info_bits(uint8_t *bits, size_t k)
{
    for (size_t i = 0; i < k; ++i) bits[i] = (uint8_t)((((uint64_t)i * 2654435761ULL) >> 31) & 1);
}

/* splitmix64 (channel_sim.rs:61-67) */
static uint64_t splitmix64(uint64_t seed)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z == 0 ? 1 : z;
}

/* next_uniform (channel_sim.rs:77-91) */
static float next_uniform(uint64_t *st)
{
    uint64_t x = *st;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *st = x;
    return u2f(0x3F800000u | (uint32_t)(x >> 41)) - 1.0f;
}

/* next_gaussian (channel_sim.rs:102-106): Box-Muller with f32 libm calls. */
static float next_gaussian(uint64_t *st)
{
    float u1 = fmaxf(next_uniform(st), 1e-38f);
    float u2 = next_uniform(st);
    const float two_pi = 2.0f * 3.14159265358979323846f;
    return sqrtf(-2.0f * logf(u1)) * cosf(two_pi * u2);
}

float orc_awgn_sigma(float ebno_db, float code_rate)
{
    float lin = powf(10.0f, ebno_db / 10.0f); /* channel_sim.rs:139-147 */
    return sqrtf(1.0f / (2.0f * code_rate * lin));
}

void orc_awgn_transmit(float ebno_db, float code_rate, uint64_t seed, const uint8_t *bits, size_t n,
                       float *llr_out)
{
    float sigma = orc_awgn_sigma(ebno_db, code_rate);
    uint64_t st = splitmix64(seed);
    float s2 = sigma * sigma; /* transmit, channel_sim.rs:204-215 */
    float scale = 2.0f / s2;
    for (size_t i = 0; i < n; ++i) {
        float x = bits[i] == 0 ? 1.0f : -1.0f;
        float g = next_gaussian(&st);
        float r = x + g * sigma;
        llr_out[i] = r * scale;
    }
}

/* ------------------------------------------------------------------ */
/* encoder: parity by GF(2) elimination                                */
/* ------------------------------------------------------------------ */

typedef struct {
    orc_code_t code;
    int core_rows;  /* rows [0, core_rows) solved jointly */
    int nbits;      /* core_rows * z */
    int words;
    uint64_t *inv;  /* nbits x nbits inverse of the core parity block, row-major bitsets */
} enc_cache_t;

static enc_cache_t g_enc;

static int has_edge(const orc_code_t *c, int r, int col, int *shift)
{
    for (int e = c->layer_off[r]; e < c->layer_off[r + 1]; ++e)
        if (c->cols[e] == col) { if (shift) *shift = c->shifts[e]; return 1; }
    return 0;
}

/* Smallest core size such that every later row closes with an identity
 * block in its own parity column (the 5G extension structure); otherwise
 * the whole parity part is solved at once. */
static int find_core(const orc_code_t *c)
{
    const int kb = c->n_b - c->m_b;
    for (int core = 1; core <= c->m_b; ++core) {
        int ok = 1;
        for (int r = 0; r < c->m_b && ok; ++r)
            for (int e = c->layer_off[r]; e < c->layer_off[r + 1]; ++e) {
                int col = c->cols[e];
                if (col < kb) continue;
                int in_core = col < kb + core;
                if (r < core) { if (!in_core) { ok = 0; break; } }
                else if (!in_core && !(col == kb + r && c->shifts[e] == 0)) { ok = 0; break; }
            }
        for (int r = core; r < c->m_b && ok; ++r) {
            int s;
            if (!has_edge(c, r, kb + r, &s) || s != 0) ok = 0;
        }
        if (ok) return core;
    }
    return -1;
}

static int build_encoder(const orc_code_t *c)
{
    if (g_enc.inv && memcmp(&g_enc.code, c, sizeof(*c)) == 0) return 0;
    free(g_enc.inv);
    g_enc.inv = NULL;
    const int z = c->z, kb = c->n_b - c->m_b;
    int core = find_core(c);
    if (core < 0) return -1;
    const int nb = core * z, w = (nb + 63) / 64;
    uint64_t *a = (uint64_t *)calloc((size_t)nb * w, 8);
    uint64_t *inv = (uint64_t *)calloc((size_t)nb * w, 8);
    /* row (r, zi) of the core block: bit (col - kb) * z + (zi + s) % z */
    for (int r = 0; r < core; ++r)
        for (int e = c->layer_off[r]; e < c->layer_off[r + 1]; ++e) {
            int col = c->cols[e];
            if (col < kb) continue;
            for (int zi = 0; zi < z; ++zi) {
                int row = r * z + zi;
                int bit = (col - kb) * z + (zi + c->shifts[e]) % z;
                a[(size_t)row * w + bit / 64] ^= 1ULL << (bit % 64);
            }
        }
    for (int i = 0; i < nb; ++i) inv[(size_t)i * w + i / 64] |= 1ULL << (i % 64);
    for (int col = 0; col < nb; ++col) {
        int piv = -1;
        for (int r = col; r < nb; ++r)
            if (a[(size_t)r * w + col / 64] >> (col % 64) & 1) { piv = r; break; }
        if (piv < 0) { free(a); free(inv); return -1; }
        if (piv != col)
            for (int k = 0; k < w; ++k) {
                uint64_t t = a[(size_t)piv * w + k]; a[(size_t)piv * w + k] = a[(size_t)col * w + k]; a[(size_t)col * w + k] = t;
                t = inv[(size_t)piv * w + k]; inv[(size_t)piv * w + k] = inv[(size_t)col * w + k]; inv[(size_t)col * w + k] = t;
            }
        for (int r = 0; r < nb; ++r)
            if (r != col && (a[(size_t)r * w + col / 64] >> (col % 64) & 1))
                for (int k = 0; k < w; ++k) {
                    a[(size_t)r * w + k] ^= a[(size_t)col * w + k];
                    inv[(size_t)r * w + k] ^= inv[(size_t)col * w + k];
                }
    }
    free(a);
    g_enc.code = *c;
    g_enc.core_rows = core;
    g_enc.nbits = nb;
    g_enc.words = w;
    g_enc.inv = inv;
    return 0;
}

int orc_encode(const orc_code_t *c, const uint8_t *info, uint8_t *cw)
{
    if (build_encoder(c) != 0) return -1;
    const int z = c->z, kb = c->n_b - c->m_b, core = g_enc.core_rows, nb = g_enc.nbits, w = g_enc.words;
    const size_t k = orc_k(c), n = orc_n(c);
    memcpy(cw, info, k);
    memset(cw + k, 0, n - k);
    /* syndrome of the systematic part over the core rows */
    uint64_t *s = (uint64_t *)calloc((size_t)w, 8);
    for (int r = 0; r < core; ++r)
        for (int e = c->layer_off[r]; e < c->layer_off[r + 1]; ++e) {
            if (c->cols[e] >= kb) continue;
            for (int zi = 0; zi < z; ++zi)
                if (cw[var_index(c, e, zi)] & 1) s[(r * z + zi) / 64] ^= 1ULL << ((r * z + zi) % 64);
        }
    for (int i = 0; i < nb; ++i) {
        uint64_t acc = 0;
        for (int k2 = 0; k2 < w; ++k2) acc ^= g_enc.inv[(size_t)i * w + k2] & s[k2];
        cw[(size_t)kb * z + i] = (uint8_t)(__builtin_popcountll(acc) & 1);
    }
    free(s);
    /* extension rows: each closes on the identity in its own parity column */
    for (int r = core; r < c->m_b; ++r)
        for (int zi = 0; zi < z; ++zi) {
            unsigned p = 0;
            for (int e = c->layer_off[r]; e < c->layer_off[r + 1]; ++e) {
                if (c->cols[e] == kb + r) continue;
                p ^= cw[var_index(c, e, zi)] & 1;
            }
            cw[(size_t)(kb + r) * z + zi] = (uint8_t)p;
        }
    return orc_syndrome_ok_bits(c, cw) ? 0 : -1;
}

/* ------------------------------------------------------------------ */
/* rate matching                                                       */
/* ------------------------------------------------------------------ */

/* RV_K0_SETS / rv_k0 (rate_matching.rs:72-94) */
static size_t rv_k0(int bg, int rv, int z)
{
    static const int j[2][4] = {{0, 17, 33, 56}, {0, 13, 25, 43}};
    return (size_t)j[bg][rv] * (size_t)z;
}

int orc_rate_match(const uint8_t *cw, size_t cw_len, uint8_t *e_out, size_t e_len, int rv, int qm, int bg, int z,
                   size_t n_filler)
{
    if (rv < 0 || rv >= 4 || qm <= 0 || z <= 0 || e_len % (size_t)qm) return -1;
    const size_t ncb = (size_t)(bg == 0 ? 66 : 50) * z, k = (size_t)(bg == 0 ? 22 : 10) * z;
    const size_t kp = n_filler > k ? 0 : k - n_filler, two_z = 2 * (size_t)z;
    const size_t fs = kp > two_z ? kp - two_z : 0, fe = k > two_z ? k - two_z : 0;
    const size_t k0 = rv_k0(bg, rv, z);
    uint8_t *e = (uint8_t *)calloc(e_len ? e_len : 1, 1); /* slots past the guard stay 0 */
    size_t ks = 0, j = 0;
    while (ks < e_len) {
        size_t pos = (k0 + j) % ncb;
        if (pos >= fs && pos < fe) { ++j; continue; }
        size_t cp = pos + two_z;
        e[ks++] = cp < cw_len ? cw[cp] : 0;
        ++j;
        if (j > 2 * ncb) break;
    }
    /* interleave: write row-major, read column-major (rate_matching.rs:101-113) */
    size_t rows = e_len / (size_t)qm;
    for (int col = 0; col < qm; ++col)
        for (size_t r = 0; r < rows; ++r) e_out[(size_t)col * rows + r] = e[r * (size_t)qm + (size_t)col];
    free(e);
    return 0;
}

int orc_rate_dematch(const float *e_llr, size_t e_len, float *cb, size_t cb_len, int rv, int qm, int bg, int z,
                     size_t n_filler)
{
    if (rv < 0 || rv >= 4 || qm <= 0 || z <= 0 || e_len % (size_t)qm) return -1;
    const size_t ncb = (size_t)(bg == 0 ? 66 : 50) * z, k = (size_t)(bg == 0 ? 22 : 10) * z;
    const size_t kp = n_filler > k ? 0 : k - n_filler, two_z = 2 * (size_t)z;
    const size_t fs = kp > two_z ? kp - two_z : 0, fe = k > two_z ? k - two_z : 0;
    size_t rows = e_len / (size_t)qm;
    float *d = (float *)malloc(sizeof(float) * (e_len ? e_len : 1));
    for (int col = 0; col < qm; ++col) /* deinterleave */
        for (size_t r = 0; r < rows; ++r) d[r * (size_t)qm + (size_t)col] = e_llr[(size_t)col * rows + r];
    const size_t k0 = rv_k0(bg, rv, z);
    size_t ks = 0, j = 0;
    while (ks < e_len) {
        size_t pos = (k0 + j) % ncb;
        if (pos >= fs && pos < fe) { ++j; continue; }
        if (pos < cb_len) cb[pos] += d[ks];
        ++ks;
        ++j;
        if (j > 2 * ncb) break;
    }
    free(d);
    return 0;
}

/* RateMatchCache::rate_dematch_llr_into (rate_matching.rs:531-551 selection
 * table, 675-715 accumulation), the de-matcher behind HarqBuffer::combine
 * (harq.rs:142-158) and so behind prepare_cb_llr: the selection walk of
 * selection_walk (rate_matching.rs:418-458) fills a table of E circular-buffer
 * positions, leaving 0 in the slots after its 2*Ncb guard stops it; the
 * received LLRs are then added in their received (interleaved) order q, each
 * to position table[(q % rows) * qm + q / rows]. */
int orc_rate_dematch_cached(const float *e_llr, size_t e_len, float *cb, size_t cb_len, int rv, int qm, int bg,
                            int z, size_t n_filler)
{
    if (rv < 0 || rv >= 4 || qm <= 0 || z <= 0 || e_len % (size_t)qm) return -1;
    const size_t ncb = (size_t)(bg == 0 ? 66 : 50) * z, k = (size_t)(bg == 0 ? 22 : 10) * z;
    const size_t kp = n_filler > k ? 0 : k - n_filler, two_z = 2 * (size_t)z;
    const size_t fs = kp > two_z ? kp - two_z : 0, fe = k > two_z ? k - two_z : 0;
    size_t *sel = (size_t *)calloc(e_len ? e_len : 1, sizeof(size_t));
    const size_t k0 = rv_k0(bg, rv, z);
    size_t ks = 0, j = 0;
    while (ks < e_len) {
        const size_t pos = (k0 + j) % ncb;
        if (pos >= fs && pos < fe) {
            ++j;
            continue;
        }
        sel[ks++] = pos;
        ++j;
        if (j > 2 * ncb) break;
    }
    const size_t rows = e_len / (size_t)qm;
    for (size_t q = 0; q < e_len; ++q) {
        const size_t idx = sel[(q % rows) * (size_t)qm + q / rows];
        if (idx < cb_len) cb[idx] += e_llr[q];
    }
    free(sel);
    return 0;
}

/* bit_serial_step / Crc24::compute (crc.rs:119-126, 258-300) for a CRC of
 * `len` bits: MSB-first register, initial value 0, no reflection, no final
 * XOR.  The byte-table form of the reference computes the same register. */
uint32_t orc_crc(const uint8_t *bits, size_t n, int len, uint32_t poly)
{
    const uint32_t mask = (len >= 32) ? 0xFFFFFFFFu : ((1u << len) - 1u);
    uint32_t reg = 0;
    poly &= mask;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t fb = ((reg >> (len - 1)) ^ (uint32_t)(bits[i] & 1u)) & 1u;
        reg = (reg << 1) & mask;
        if (fb) reg ^= poly;
    }
    return reg;
}

/* segment (segmentation.rs:361-410): the transport block with its CRC-24A
 * (b bits) cut into c blocks of P = K' - 24 payload bits (K' when c == 1),
 * the tail zero-padded, each block followed by its CRC-24B when c > 1. */
int orc_segment(const uint8_t *tb_with_crc, size_t b, size_t c, size_t k_prime, uint8_t *blocks)
{
    const size_t p = c > 1 ? k_prime - 24 : k_prime;
    if (c == 0 || k_prime < (c > 1 ? 25u : 1u) || c * p < b) return -1;
    for (size_t ci = 0; ci < c; ++ci) {
        uint8_t *blk = blocks + ci * k_prime;
        for (size_t i = 0; i < p; ++i) {
            const size_t t = ci * p + i;
            blk[i] = t < b ? (tb_with_crc[t] & 1u) : 0;
        }
        if (c > 1) {
            const uint32_t r = orc_crc(blk, p, 24, 0x800063u);
            for (int i = 0; i < 24; ++i) blk[p + (size_t)i] = (uint8_t)((r >> (23 - i)) & 1u);
        }
    }
    return 0;
}

/* The tail of DlSchDecoder::decode (transport_block.rs:963-1044): block ci's
 * first K' hard bits (at hard + cb_off[ci]) pass CRC-24B unless c == 1; its
 * first P bits go to all_info[ci*P ...]; the transport block passes when
 * every block passed and CRC-24A of all_info[0, tb_size) equals
 * all_info[tb_size, +24).  tb_out gets the tb_size bits only on a pass.
 * Returns 1 (pass), 0 (fail) or -1 (bad sizes). */
int orc_tb_check(const uint8_t *hard, const int64_t *cb_off, size_t c, size_t k_prime, size_t tb_size,
                 uint8_t *cb_ok, uint8_t *tb_out)
{
    const size_t p = c > 1 ? k_prime - 24 : k_prime;
    if (c == 0 || c * p < tb_size + 24) return -1;
    uint8_t *all = (uint8_t *)malloc(c * p);
    int every = 1;
    for (size_t ci = 0; ci < c; ++ci) {
        const uint8_t *info = hard + cb_off[ci];
        int ok = 1;
        if (c > 1) {
            uint32_t rx = 0;
            for (int i = 0; i < 24; ++i) rx = (rx << 1) | (info[p + (size_t)i] & 1u);
            ok = orc_crc(info, p, 24, 0x800063u) == rx;
        }
        cb_ok[ci] = (uint8_t)ok;
        every &= ok;
        for (size_t i = 0; i < p; ++i) all[ci * p + i] = info[i] & 1u;
    }
    uint32_t rx = 0;
    for (int i = 0; i < 24; ++i) rx = (rx << 1) | all[tb_size + (size_t)i];
    const int pass = every && orc_crc(all, tb_size, 24, 0x864CFBu) == rx;
    if (pass) memcpy(tb_out, all, tb_size);
    free(all);
    return pass;
}
