// Host-side code construction, encoder and channel model.
#include "code_host.h"

#include <math.h>
#include <string.h>

#include <algorithm>

#include "code_tables.inc"

namespace syn {

namespace {

// 3GPP TS 38.212 Table 5.3.2-1: lifting-size sets, iLS = 0..7
// (src/qc_ldpc.rs:103-125).
int lifting_set(long long z)
{
    static const int sets[8][8] = {
        {2, 4, 8, 16, 32, 64, 128, 256}, {3, 6, 12, 24, 48, 96, 192, 384}, {5, 10, 20, 40, 80, 160, 320},
        {7, 14, 28, 56, 112, 224},       {9, 18, 36, 72, 144, 288},        {11, 22, 44, 88, 176, 352},
        {13, 26, 52, 104, 208},          {15, 30, 60, 120, 240},
    };
    for (int s = 0; s < 8; ++s)
        for (int v : sets[s])
            if (v != 0 && v == z) return s;
    return -1;
}

// Group edges by row block, keeping the given order inside a row
// (build_layout, src/qc_ldpc.rs:281-318): per-row offsets into the grouped
// edge list, and the maximum row degree.
void group_rows(const int *rows, const int *cols, const int *shifts, size_t n, int m_b, std::vector<int> &off,
                std::vector<int> &gc, std::vector<int> &gs, int &max_deg)
{
    std::vector<int> deg((size_t)m_b, 0), fill((size_t)m_b, 0);
    for (size_t i = 0; i < n; ++i) deg[(size_t)rows[i]]++;
    off.assign((size_t)m_b + 1, 0);
    max_deg = 0;
    for (int r = 0; r < m_b; ++r) {
        off[(size_t)r + 1] = off[(size_t)r] + deg[(size_t)r];
        max_deg = std::max(max_deg, deg[(size_t)r]);
    }
    gc.assign(n, 0);
    gs.assign(n, 0);
    for (size_t i = 0; i < n; ++i) {
        const size_t p = (size_t)off[(size_t)rows[i]] + (size_t)fill[(size_t)rows[i]]++;
        gc[p] = cols[i];
        gs[p] = shifts[i];
    }
}

void layout(SynCode *c, const int *rows, const int *cols, const int *shifts, int n)
{
    std::vector<int> off, gc, gs;
    group_rows(rows, cols, shifts, (size_t)n, c->m_b, off, gc, gs, c->max_deg);
    std::copy(off.begin(), off.end(), c->layer_off);
    std::copy(gc.begin(), gc.end(), c->cols);
    std::copy(gs.begin(), gs.end(), c->shifts);
    c->n_edges = n;
}

}  // namespace

int build_code_3gpp(int bg, long long z, SynCode *out)
{
    const int ils = lifting_set(z);
    if (ils < 0 || (bg != 0 && bg != 1)) return SYN_INVALID_PARAM;
    const bool one = bg == 0;
    const int m_b = one ? SYN_BG1_ROWS : SYN_BG2_ROWS;
    const int n = one ? SYN_BG1_EDGES : SYN_BG2_EDGES;
    const unsigned char *deg = one ? SYN_BG1_ROW_DEG : SYN_BG2_ROW_DEG;
    const unsigned char *col = one ? SYN_BG1_EDGE_COL : SYN_BG2_EDGE_COL;
    const short(*v)[8] = one ? SYN_BG1_EDGE_V : SYN_BG2_EDGE_V;
    int rows[SYN_MAX_EDGES], cols[SYN_MAX_EDGES], sh[SYN_MAX_EDGES];
    for (int r = 0, e = 0; r < m_b; ++r)
        for (int j = 0; j < deg[r]; ++j, ++e) {
            rows[e] = r;
            cols[e] = col[e];
            sh[e] = (int)(v[e][ils] % z);  // entry_col_shift (src/qc_ldpc.rs:134-140)
        }
    memset(out, 0, sizeof(*out));
    out->kind = one ? SYN_KIND_BG1 : SYN_KIND_BG2;
    out->z = (int)z;
    out->m_b = m_b;
    out->n_b = one ? SYN_BG1_COLS : SYN_BG2_COLS;
    layout(out, rows, cols, sh, n);
    return SYN_OK;
}

int build_code_raw(const int *rows, const int *cols, const int *shifts, size_t n, long long m_b, long long n_b,
                   long long z, SynCode *out, WideCode *wide)
{
    // src/qc_ldpc.rs:418-440
    if (z <= 0 || z > SYN_Z_MAX || m_b < 0 || n_b < 0) return SYN_INVALID_PARAM;
    for (size_t i = 0; i < n; ++i) {
        if (rows[i] < 0 || rows[i] >= m_b) return SYN_INVALID_PARAM;
        if (cols[i] < 0 || cols[i] >= n_b) return SYN_INVALID_PARAM;
        if (shifts[i] < 0 || shifts[i] >= z) return SYN_INVALID_PARAM;
    }
    // 32-bit element indices on the device
    const long long lim = 0x7fffffffLL;
    if (m_b > lim / z || n_b > lim / z || n > (size_t)(lim / z)) return SYN_INVALID_PARAM;
    // rows past the last one named hold no edges: the tables cover rows
    // 0 .. rmax only, whatever m_b and n_b are
    int rmax = -1, cmax = -1;
    for (size_t i = 0; i < n; ++i) rmax = std::max(rmax, rows[i]), cmax = std::max(cmax, cols[i]);
    const int nr = rmax + 1;
    std::vector<int> off, gc, gs;
    int max_deg = 0;
    group_rows(rows, cols, shifts, n, nr, off, gc, gs, max_deg);
    // a layer with a column block twice shares variables between its rows
    std::vector<unsigned char> dup((size_t)nr, 0);
    bool any_dup = false;
    {
        std::vector<int> seen((size_t)cmax + 1, -1);
        for (int r = 0; r < nr; ++r)
            for (int e = off[(size_t)r]; e < off[(size_t)r + 1]; ++e) {
                if (seen[(size_t)gc[(size_t)e]] == (int)r) dup[(size_t)r] = 1, any_dup = true;
                seen[(size_t)gc[(size_t)e]] = (int)r;
            }
    }
    memset(out, 0, sizeof(*out));
    out->z = (int)z;
    out->m_b = (int)m_b;
    out->n_b = (int)n_b;
    out->n_edges = (int)n;
    out->max_deg = max_deg;
    const bool inline_form =
        m_b >= 1 && m_b <= SYN_MAX_ROWS && n_b >= 1 && n_b <= SYN_MAX_COLS && n <= SYN_MAX_EDGES && !any_dup;
    if (inline_form) {
        out->kind = SYN_KIND_RAW;
        std::copy(off.begin(), off.end(), out->layer_off);
        for (long long r = nr; r < m_b; ++r) out->layer_off[r + 1] = (int)n;
        std::copy(gc.begin(), gc.end(), out->cols);
        std::copy(gs.begin(), gs.end(), out->shifts);
        if (wide) *wide = WideCode{};
        return SYN_OK;
    }
    if (!wide) return SYN_INVALID_PARAM;
    out->kind = SYN_KIND_WIDE;
    // the layers that hold edges, in order
    WideCode w;
    w.layer_off.push_back(0);
    for (int r = 0; r < nr; ++r) {
        if (off[(size_t)r + 1] == off[(size_t)r]) continue;
        w.layer_off.push_back(off[(size_t)r + 1]);
        w.dup.push_back(dup[(size_t)r]);
    }
    w.cols = std::move(gc);
    w.shifts = std::move(gs);
    w.hdr.z = (int)z;
    w.hdr.m_b = (int)m_b;
    w.hdr.n_b = (int)n_b;
    w.hdr.n_layers = (int)w.dup.size();
    w.hdr.n_edges = (int)n;
    w.hdr.max_deg = max_deg;
    w.hdr.any_dup = any_dup ? 1 : 0;
    *wide = std::move(w);
    return SYN_OK;
}

int build_code_wifi(long long z, long long rate_num, long long rate_den, SynCode *out)
{
    for (int i = 0; i < SYN_WIFI_COUNT; ++i) {
        const syn_wifi_matrix_t &m = SYN_WIFI_MATRICES[i];
        if (m.z != z || m.rate_num != rate_num || m.rate_den != rate_den) continue;
        int rows[SYN_MAX_EDGES], cols[SYN_MAX_EDGES], sh[SYN_MAX_EDGES];
        for (int e = 0; e < m.n_edges; ++e) {
            rows[e] = m.edges[e][0];
            cols[e] = m.edges[e][1];
            sh[e] = m.edges[e][2];
        }
        const int rc = build_code_raw(rows, cols, sh, (size_t)m.n_edges, m.row_blocks, m.col_blocks, m.z, out);
        if (rc == SYN_OK) out->kind = SYN_KIND_WIFI0 + i;  // the register-resident kernels know this graph
        return rc;
    }
    return SYN_INVALID_PARAM;
}

// ---------------------------------------------------------------------------
// encoder
// ---------------------------------------------------------------------------

namespace {

inline int vidx(const SynCode &c, int e, int z)
{
    int s = z + c.shifts[e];
    if (s >= c.z) s -= c.z;
    return c.cols[e] * c.z + s;
}

// Smallest number of leading block rows that must be solved jointly: every
// later row r has parity support inside the core plus an identity block in
// parity column k_b + r.
int core_rows(const SynCode &c)
{
    const int kb = c.n_b - c.m_b;
    for (int core = 1; core <= c.m_b; ++core) {
        bool ok = true;
        for (int r = 0; r < c.m_b && ok; ++r) {
            bool own = false;
            for (int e = c.layer_off[r]; e < c.layer_off[r + 1]; ++e) {
                const int col = c.cols[e];
                if (col < kb) continue;
                if (col < kb + core) continue;
                if (r >= core && col == kb + r && c.shifts[e] == 0) own = true;
                else ok = false;
            }
            if (r >= core && !own) ok = false;
        }
        if (ok) return core;
    }
    return -1;
}

}  // namespace

Encoder::Encoder(const SynCode &c) : c_(c)
{
    core_ = core_rows(c);
    if (core_ < 0) return;
    const int z = c.z, kb = c.n_b - c.m_b;
    bits_ = core_ * z;
    words_ = (bits_ + 63) / 64;
    std::vector<uint64_t> a((size_t)bits_ * words_, 0);
    inv_.assign((size_t)bits_ * words_, 0);
    auto flip = [&](std::vector<uint64_t> &m, int row, int bit) { m[(size_t)row * words_ + bit / 64] ^= 1ULL << (bit & 63); };
    for (int r = 0; r < core_; ++r)
        for (int e = c.layer_off[r]; e < c.layer_off[r + 1]; ++e) {
            if (c.cols[e] < kb) continue;
            for (int zi = 0; zi < z; ++zi) flip(a, r * z + zi, vidx(c, e, zi) - kb * z);
        }
    for (int i = 0; i < bits_; ++i) flip(inv_, i, i);
    // Gauss-Jordan over GF(2)
    for (int col = 0; col < bits_; ++col) {
        int piv = -1;
        for (int r = col; r < bits_ && piv < 0; ++r)
            if ((a[(size_t)r * words_ + col / 64] >> (col & 63)) & 1) piv = r;
        if (piv < 0) return;
        if (piv != col)
            for (int w = 0; w < words_; ++w) {
                std::swap(a[(size_t)piv * words_ + w], a[(size_t)col * words_ + w]);
                std::swap(inv_[(size_t)piv * words_ + w], inv_[(size_t)col * words_ + w]);
            }
        const uint64_t *pa = &a[(size_t)col * words_];
        const uint64_t *pi = &inv_[(size_t)col * words_];
        for (int r = 0; r < bits_; ++r) {
            if (r == col || !((a[(size_t)r * words_ + col / 64] >> (col & 63)) & 1)) continue;
            uint64_t *ra = &a[(size_t)r * words_];
            uint64_t *ri = &inv_[(size_t)r * words_];
            for (int w = 0; w < words_; ++w) {
                ra[w] ^= pa[w];
                ri[w] ^= pi[w];
            }
        }
    }
    ok_ = true;
}

void Encoder::encode(const uint8_t *info, uint8_t *cw) const
{
    const SynCode &c = c_;
    const int z = c.z, kb = c.n_b - c.m_b;
    const size_t k = code_k(c), n = code_n(c);
    memcpy(cw, info, k);
    memset(cw + k, 0, n - k);
    std::vector<uint64_t> s((size_t)words_, 0);
    for (int r = 0; r < core_; ++r)
        for (int e = c.layer_off[r]; e < c.layer_off[r + 1]; ++e) {
            if (c.cols[e] >= kb) continue;
            for (int zi = 0; zi < z; ++zi)
                if (cw[vidx(c, e, zi)] & 1) s[(r * z + zi) / 64] ^= 1ULL << ((r * z + zi) & 63);
        }
    for (int i = 0; i < bits_; ++i) {
        const uint64_t *row = &inv_[(size_t)i * words_];
        uint64_t acc = 0;
        for (int w = 0; w < words_; ++w) acc ^= row[w] & s[w];
        cw[(size_t)kb * z + i] = (uint8_t)(__builtin_popcountll(acc) & 1);
    }
    for (int r = core_; r < c.m_b; ++r)
        for (int zi = 0; zi < z; ++zi) {
            unsigned p = 0;
            for (int e = c.layer_off[r]; e < c.layer_off[r + 1]; ++e)
                if (c.cols[e] != kb + r) p ^= cw[vidx(c, e, zi)] & 1;
            cw[(size_t)(kb + r) * z + zi] = (uint8_t)p;
        }
}

bool Encoder::core_terms(std::vector<int> &off, std::vector<int> &blk, std::vector<int> &shift) const
{
    if (!ok_) return false;
    const int z = c_.z;
    auto bit = [&](int row, int col) { return (inv_[(size_t)row * words_ + col / 64] >> (col & 63)) & 1; };
    off.assign(1, 0);
    blk.clear();
    shift.clear();
    for (int i = 0; i < core_; ++i) {
        for (int j = 0; j < core_; ++j)
            for (int t = 0; t < z; ++t)
                if (bit(i * z, j * z + t)) {
                    blk.push_back(j);
                    shift.push_back(t);
                }
        off.push_back((int)blk.size());
        // every row of the block row must be the first one rotated
        for (int zi = 1; zi < z; ++zi)
            for (int j = 0; j < core_; ++j)
                for (int t = 0; t < z; ++t) {
                    int col = zi + t;
                    if (col >= z) col -= z;
                    if (bit(i * z + zi, j * z + col) != bit(i * z, j * z + t)) return false;
                }
    }
    return true;
}

// ---------------------------------------------------------------------------
// rate matching
// ---------------------------------------------------------------------------

bool rate_match(const uint8_t *cw, size_t cw_len, uint8_t *e_out, size_t e_len, int rv, int qm, int bg, int z,
                size_t n_filler)
{
    if (rv < 0 || rv > 3 || qm <= 0 || z <= 0 || (bg != 0 && bg != 1) || e_len % (size_t)qm) return false;
    static const int k0_table[2][4] = {{0, 17, 33, 56}, {0, 13, 25, 43}};  // TS 38.212 Table 5.4.2.1-2
    const size_t ncb = (size_t)(bg == 0 ? 66 : 50) * z, k = (size_t)(bg == 0 ? 22 : 10) * z, twoz = 2 * (size_t)z;
    const size_t kp = n_filler > k ? 0 : k - n_filler;
    const size_t fill_lo = kp > twoz ? kp - twoz : 0, fill_hi = k > twoz ? k - twoz : 0;
    std::vector<uint8_t> sel(e_len);  // slots the walk does not reach stay 0
    size_t pos = (size_t)k0_table[bg][rv] * z, taken = 0, steps = 0;
    while (taken < e_len) {
        const bool filler = pos >= fill_lo && pos < fill_hi;
        if (!filler) sel[taken++] = pos + twoz < cw_len ? cw[pos + twoz] : 0;
        pos = pos + 1 == ncb ? 0 : pos + 1;
        ++steps;
        if (!filler && steps > 2 * ncb) break;  // the walk's guard, checked after a selection
    }
    const size_t rows = e_len / (size_t)qm;  // write rows of Qm, read columns
    for (size_t r = 0; r < rows; ++r)
        for (int c = 0; c < qm; ++c) e_out[(size_t)c * rows + r] = sel[r * (size_t)qm + (size_t)c];
    return true;
}

// ---------------------------------------------------------------------------
// channel
// ---------------------------------------------------------------------------

namespace {

// splitmix64 seed expansion (src/channel_sim.rs:61-67).
uint64_t seed_expand(uint64_t seed)
{
    uint64_t x = seed + 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x ? x : 1;
}

// xorshift64 -> f32 in [0, 1) from the top 23 bits (src/channel_sim.rs:77-91).
float uniform01(uint64_t &st)
{
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    const uint32_t bits = 0x3F800000u | (uint32_t)(st >> 41);
    float f;
    memcpy(&f, &bits, 4);
    return f - 1.0f;
}

// Box-Muller in f32 (src/channel_sim.rs:102-106).
float gauss(uint64_t &st)
{
    const float u1 = std::max(uniform01(st), 1e-38f);
    const float u2 = uniform01(st);
    return sqrtf(-2.0f * logf(u1)) * cosf((2.0f * 3.14159265358979323846f) * u2);
}

}  // namespace

void awgn_transmit(float ebno_db, float code_rate, uint64_t seed, const uint8_t *bits, size_t n, float *llr_out)
{
    const float sigma = sqrtf(1.0f / (2.0f * code_rate * powf(10.0f, ebno_db / 10.0f)));
    const float scale = 2.0f / (sigma * sigma);
    uint64_t st = seed_expand(seed);
    for (size_t i = 0; i < n; ++i) {
        const float x = bits[i] ? -1.0f : 1.0f;
        const float g = gauss(st);
        const float r = x + g * sigma;
        llr_out[i] = r * scale;
    }
}

}  // namespace syn
