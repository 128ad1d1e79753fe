// Register-resident LOMS kernels: compile-time base graph (BG1 / BG2) and,
// optionally, compile-time lifting size.
//
// One workgroup decodes `cws` codewords of one (BG, Z).  Lane t owns
// lifting index z = t % Z of codeword t / Z and processes check row
// (layer, z) of every layer, in the reference's strict layer order
// (src/qc_ldpc.rs:1055-1206).  The posterior lives in LDS; the row's
// check-to-variable messages never leave the lane's registers, kept in the
// compressed min-sum form (|R| of the argmin edge, |R| of every other
// edge, the argmin and one sign bit per edge).
//
// ZC > 0 instantiates the kernel for one lifting size: every shift becomes
// an immediate, so a posterior address is one compare, one select between
// two per-lane bases, and the column offset folded into the LDS
// instruction.  ZC == 0 reads the shifts of the decoder's Z at run time.
#pragma once
#include "ldpc_device.h"

namespace syn {

// The IEEE 802.11 matrices (KIND = SYN_KIND_WIFI0 + index): edges in the
// table's row-grouped order, which is from_raw_edges' order
// (src/qc_ldpc.rs:412-455); shifts are given for the matrix's own Z.
template <int KIND>
struct Graph {
    static constexpr int M = KIND - SYN_KIND_WIFI0;
    static_assert(M >= 0 && M < SYN_WIFI_COUNT, "graph kind");
    static constexpr int ROWS = SYN_WIFI_MATRICES[M].row_blocks, COLS = SYN_WIFI_MATRICES[M].col_blocks;
    static constexpr int EDGES = SYN_WIFI_MATRICES[M].n_edges;
    static constexpr int Z = SYN_WIFI_MATRICES[M].z;
    static constexpr int deg(int l)
    {
        int d = 0;
        for (int e = 0; e < EDGES; ++e) d += SYN_WIFI_MATRICES[M].edges[e][0] == l;
        return d;
    }
    static constexpr int col(int e) { return SYN_WIFI_MATRICES[M].edges[e][1]; }
    static constexpr int v(int e, int) { return SYN_WIFI_MATRICES[M].edges[e][2]; }
};

template <>
struct Graph<SYN_KIND_BG1> {
    static constexpr int ROWS = SYN_BG1_ROWS, COLS = SYN_BG1_COLS, EDGES = SYN_BG1_EDGES;
    static constexpr int deg(int l) { return SYN_BG1_ROW_DEG[l]; }
    static constexpr int col(int e) { return SYN_BG1_EDGE_COL[e]; }
    static constexpr int v(int e, int ils) { return SYN_BG1_EDGE_V[e][ils]; }
};

template <>
struct Graph<SYN_KIND_BG2> {
    static constexpr int ROWS = SYN_BG2_ROWS, COLS = SYN_BG2_COLS, EDGES = SYN_BG2_EDGES;
    static constexpr int deg(int l) { return SYN_BG2_ROW_DEG[l]; }
    static constexpr int col(int e) { return SYN_BG2_EDGE_COL[e]; }
    static constexpr int v(int e, int ils) { return SYN_BG2_EDGE_V[e][ils]; }
};

// 3GPP TS 38.212 Table 5.3.2-1 (src/qc_ldpc.rs:103-125).
constexpr int lifting_set(int z)
{
    int a = z;
    while (a % 2 == 0) a /= 2;
    switch (a) {
    case 1: return 0;
    case 3: return 1;
    case 5: return 2;
    case 7: return 3;
    case 9: return 4;
    case 11: return 5;
    case 13: return 6;
    case 15: return 7;
    default: return -1;
    }
}

template <class G>
constexpr int layer_begin(int l)
{
    int s = 0;
    for (int i = 0; i < l; ++i) s += G::deg(i);
    return s;
}

// Edges of column c, and whether edge e's column has no other edge (a
// degree-1 variable: BG1 columns 26..67, BG2 columns 14..51).
template <class G>
constexpr int col_degree(int c)
{
    int d = 0;
    for (int e = 0; e < G::EDGES; ++e) d += G::col(e) == c;
    return d;
}

template <class G>
constexpr bool edge_deg1(int e)
{
    return col_degree<G>(G::col(e)) == 1;
}

template <class G>
constexpr bool any_deg1()
{
    for (int e = 0; e < G::EDGES; ++e)
        if (edge_deg1<G>(e)) return true;
    return false;
}

// A barrier is needed after layer l unless layer l+1 touches none of the
// columns touched since the last barrier: layers over disjoint variables
// commute exactly, so running them without a barrier between them changes
// nothing.  Greedy grouping from layer 0.
template <class G>
constexpr bool barrier_after(int l)
{
    if (l + 1 >= G::ROWS) return true;
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

// Per-layer sign word: bits 0..deg-1 hold the sign of R on each edge, edge e
// at bit deg-1-e (the order a shift-in accumulation produces), bits
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
    float g1[G::ROWS];           // |R| on every edge but the argmin: max(min1 - beta, 0)
    float g2[G::ROWS];           // |R| on the argmin edge:          max(min2 - beta, 0)
    uint32_t sg[sg_words<G>()];  // packed sign words
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

// Bytes of LDS reserved in front of the posterior so that "base - Z*size"
// stays non-negative (lets the column offset ride in the LDS instruction).
__host__ __device__ constexpr int fast_pad(int z, int esz) { return ((z * esz) + 15) & ~15; }

// Per-lane addressing for one iteration: pb = LDS address of element
// (codeword, column 0, z); pbw = pb - Z*size, used where (z + shift) wraps;
// *h = the same plus 64 KiB, for column offsets beyond the 16-bit LDS
// instruction offset.
struct Lane {
    unsigned char *pb, *pbw, *pbh, *pbwh;
    int z;
};

// Which lanes wrap on each edge of a compile-time (graph, Z): bit l of
// m[w][E] is set when lane 64w + l, whose lifting index is
// (64w + l) mod ZC, has z + shift(E) >= ZC.  A wave reads its own row with
// scalar loads, so choosing a lane's LDS base costs one select on a lane
// mask already in SGPRs instead of a per-lane compare and a select.
constexpr int WRAP_WAVES = 8;  // 512-lane workgroups

template <class G, int ZC>
struct WrapMasks {
    unsigned long long m[WRAP_WAVES][G::EDGES];
};

template <class G, int ZC>
constexpr WrapMasks<G, ZC> make_wrap_masks()
{
    WrapMasks<G, ZC> w{};
    for (int e = 0; e < G::EDGES; ++e) {
        const int T = ZC - G::v(e, lifting_set(ZC)) % ZC;  // lanes z >= T wrap
        for (int t = 0; t < 64 * WRAP_WAVES; ++t)
            if (t % ZC >= T) w.m[t / 64][e] |= 1ull << (t % 64);
    }
    return w;
}

template <class G, int ZC>
__constant__ const WrapMasks<G, ZC> kWrap = make_wrap_masks<G, ZC>();

template <class G, int ZC, bool WM = true>
struct Geo {
    int Z;                         // run-time lifting size (ZC == 0)
    const int *__restrict__ sh;    // run-time shifts (ZC == 0)
    const unsigned long long *__restrict__ wm;  // this wave's row of kWrap (ZC > 0 and WM)

    __device__ __forceinline__ int z_size() const
    {
        if constexpr (ZC > 0) return ZC;
        else return Z;
    }

    // LDS address of variable col(E)*Z + (z + shift(E)) mod Z.
    template <int E, int ESZ>
    __device__ __forceinline__ unsigned char *addr(const Lane &ln) const
    {
        constexpr int col = G::col(E);
        if constexpr (ZC > 0) {
            constexpr int s = G::v(E, lifting_set(ZC)) % ZC;
            constexpr int T = ZC - s;  // lanes z >= T wrap
            constexpr int A = (col * ZC + s) * ESZ;
            if constexpr (s == 0) {
                return ((A >= 65536) ? ln.pbh : ln.pb) + (A & 0xffff);
            } else {
                static_assert(T > 0);
                bool wr;
                if constexpr (WM) wr = __builtin_amdgcn_inverse_ballot_w64(wm[E]);
                else wr = ln.z >= T;
                if constexpr (A >= 65536) return (wr ? ln.pbwh : ln.pbh) + (A & 0xffff);
                else return (wr ? ln.pbw : ln.pb) + A;
            }
        } else {
            const int s = sh[E];
            return ((ln.z >= Z - s) ? ln.pbw : ln.pb) + (col * Z + s) * ESZ;
        }
    }
};

// R of edge e rebuilt from a layer's compressed state: s = sign word, am =
// argmin edge.  The reference picks min2 for an edge whose |q| equals min1
// and min1 otherwise; an edge other than the tracked argmin can equal min1
// only when min1 is tied, and a tie forces min2 == min1 in pass 1, so "is
// the argmin" selects the same value.
template <int D>
__device__ __forceinline__ float r_f32(float g1, float g2, uint32_t s, uint32_t am, int e)
{
    const float m = (am == (uint32_t)e) ? g2 : g1;
    return __uint_as_float(__float_as_uint(m) | ((s << (31 - (D - 1 - e))) & 0x80000000u));
}

template <int D>
__device__ __forceinline__ int r_i8(int g1, int g2, uint32_t s, uint32_t am, int e)
{
    const int m = (am == (uint32_t)e) ? g2 : g1;
    const int neg = -(int)((s >> (D - 1 - e)) & 1u);
    return (m ^ neg) - neg;
}

// x + R for the R of r_i8, as (m ^ n) + x - n with n = 0 or -1 from the
// sign bit.  Given the inverted sign bits it returns x - R.
template <int D>
__device__ __forceinline__ int add_r_i8(int x, int g1, int g2, uint32_t s, uint32_t am, int e)
{
    const int m = (am == (uint32_t)e) ? g2 : g1;
    const int n = __builtin_amdgcn_sbfe((int)s, D - 1 - e, 1);  // one v_bfe_i32
    return xad(m, n, x) - n;
}

// One layer of the f32 decode (src/qc_ldpc.rs:1072-1204).  KEEPQ keeps q's
// bits in registers from pass 1 to pass 2, which saves one operation per
// edge in pass 2 for D more live registers.
template <class G, int ZC, int L, bool KEEPQ, bool WM>
__device__ __forceinline__ void layer_f32(const Lane &ln, const Geo<G, ZC, WM> &geo, StateF32<G> &st, float beta,
                                          int glo, int g0, bool act)
{
    constexpr int D = G::deg(L), B = layer_begin<G>(L);
    static_assert(D >= 2 && D <= 27, "row degree: its signs and argmin share one 32-bit word");
    float *adr[D];
    float Lv[D], ro[D];
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        adr[e] = (float *)geo.template addr<B + e, 4>(ln);
        Lv[e] = *adr[e];
    });
    const float o1 = st.g1[L], o2 = st.g2[L];
    const uint32_t os = sg_get<G, L>(st.sg);
    const uint32_t oam = (os >> D) & 31u;
    // Pass 1 on the bit patterns of |q|: for non-negative floats unsigned
    // order is float order and a NaN sorts above +inf, so the integer
    // min / clamp below keep exactly the reference's compare-based
    // min1/min2 (a NaN is ignored there, and here it can never be selected).
    uint32_t m1 = 0x7f7fffffu, m2 = 0x7f7fffffu;  // FLT_MAX
    uint32_t qs = 0, am = 31;
    uint32_t qsg[D];  // q's bits, kept for the sign of R_new (KEEPQ)
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        ro[e] = r_f32<D>(o1, o2, os, oam, e);
        const uint32_t qb = __float_as_uint(Lv[e] - ro[e]);  // Q-build
        if constexpr (KEEPQ) qsg[e] = qb;
        qs = __builtin_amdgcn_alignbit(qs, qb, 31);           // shift in q's sign
        const uint32_t a = qb & 0x7fffffffu;
        if constexpr (!KEEPQ) am = (a <= m1) ? (uint32_t)e : am;
        // min(m2, max(m1, a)) as m1 <= m2: one v_med3 (the register-lean
        // build keeps the compiler's max/min form, which needs fewer VGPRs)
        // (on edge 0, m1 == m2 == FLT_MAX and m2 stays)
        if constexpr (e == 0) {
        } else if constexpr (KEEPQ) {
            m2 = med3u(a, m1, m2);
        } else {
            m2 = clampu(a, m1, m2);
        }
        m1 = a < m1 ? a : m1;
    });
    // mag = me - beta, then (mag < 0) ? 0 : mag, as one signed-integer max
    // of mag's bits against glo: negative floats have negative bit patterns
    // and mag is never -0 (me >= +0), so max(bits, 0) is that select; with
    // a NaN beta, glo = INT_MIN and the NaN passes whatever its sign, as it
    // does the compare.  The sentinel rule (min1 = FLT_MAX -> 0, then
    // min2 = FLT_MAX -> min1) is applied last, to the finished magnitudes:
    // g0 = the magnitude of a min of 0, so the rule's compares run beside
    // the subtract and the max instead of before them.
    const int g1s = max(__float_as_int(__uint_as_float(m1) - beta), glo);
    const int g2s = max(__float_as_int(__uint_as_float(m2) - beta), glo);
    const int g1i = m1 == 0x7f7fffffu ? g0 : g1s;
    const float g1 = __int_as_float(g1i);
    const float g2 = __int_as_float(m2 == 0x7f7fffffu ? g1i : g2s);
    // sign of R_e = XOR of all q signs but e's own
    const uint32_t par = __popc(qs) & 1u;
    const uint32_t ns = qs ^ (par ? ((1u << D) - 1u) : 0u);
    if (act) {
        // R_new's sign is q's own sign XOR the parity of all of them: carry
        // the parity in bit 31 of both magnitudes, XOR q's sign bit in
        const uint32_t g1p = __float_as_uint(g1) | (par << 31), g2p = __float_as_uint(g2) | (par << 31);
        const float m1f = __uint_as_float(m1);
        SFor<0, D>::run([&](auto ec) {
            constexpr int e = decltype(ec)::value;
            float rn;
            if constexpr (KEEPQ) {
                // the argmin is the last edge whose |q| equals min1 (the
                // running "a <= m1" of pass 1 ends there); |q| compares as a
                // float, which for non-negative values is bit equality and
                // never matches a NaN
                const bool at = fabsf(__uint_as_float(qsg[e])) == m1f;
                am = at ? (uint32_t)e : am;
                rn = __uint_as_float((qsg[e] & 0x80000000u) ^ (at ? g2p : g1p));
            } else {
                rn = r_f32<D>(g1, g2, ns, am, e);
            }
            *adr[e] = Lv[e] + (rn - ro[e]);  // pass 2: posterior += R_new - R_old
        });
        st.g1[L] = g1;
        st.g2[L] = g2;
        sg_put<G, L>(st.sg, ns | (am << D));
    }
}

// One layer of the int8 decode (src/qc_ldpc.rs:1619-1745).  KEEPA keeps
// |q| in registers for pass 2, which then finds the argmin instead of
// pass 1 (one operation less per edge, D more live registers).
//
// D1: the degree-1 variables hold their channel value c for the whole
// decode.  Such a variable's only edge is this row's, so (integer
// arithmetic, no clamp reached) its posterior is A = c + R after every
// pass and its Q = A - R_old is c itself; the decode then needs neither its
// Q-build nor its posterior update, and c + R is written once at the end
// (fast_group).  The caller takes this path only when every degree-1 c of
// the workgroup has |c| <= C - 127: then |c + R| <= C for every message R
// (|R| <= 127), so neither the saturating add nor the clamp to [-C, C] of
// qc_ldpc.rs:1738-1742 ever changes the reference's posterior, which
// stays c + R exactly.
template <class G, int ZC, int L, bool KEEPA, bool WM, bool D1 = false>
__device__ __forceinline__ void layer_i8(const Lane &ln, const Geo<G, ZC, WM> &geo, StateI8<G> &st, int bq,
                                         int lo, int hi, bool act)
{
    constexpr int D = G::deg(L), B = layer_begin<G>(L);
    static_assert(D >= 2 && D <= 27, "row degree: its signs and argmin share one 32-bit word");
    int16_t *adr[D];
    int Av[D], qv[D];
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        adr[e] = (int16_t *)geo.template addr<B + e, 2>(ln);
        Av[e] = *adr[e];
    });
    const uint32_t om = mg_get<G, L>(st), os = sg_get<G, L>(st.sg);
    const int o1 = (int)(om & 0xffu), o2 = (int)((om >> 8) & 0xffu);
    const uint32_t oam = (os >> D) & 31u;
    int m1 = 127, m2 = 127;
    uint32_t qs = 0, am = 31;
    int av[D];  // |q| per edge, kept for pass 2's argmin
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        // Q-build: q = A - R_old (R_old's sign from the inverted bits).
        // The reference clamps q to [-127, 127]; the clamp changes no output here, so it is left out: the sign is the same, and with
        // m1 = m2 = 127 at the start an |q| above 127 moves neither min1 nor
        // min2.  It could only change "a == m1" / "a <= m1" when min1 ends
        // at 127, and then min2 == min1, so the argmin does not matter.
        int q;
        if constexpr (D1 && edge_deg1<G>(B + e)) q = Av[e];  // Q = c (D1 above)
        else q = add_r_i8<D>(Av[e], o1, o2, ~os, oam, e);
        qv[e] = q;
        qs = __builtin_amdgcn_alignbit(qs, (uint32_t)q, 31);
        const int a = q < 0 ? -q : q;
        if constexpr (KEEPA) av[e] = a;
        else am = (a <= m1) ? (uint32_t)e : am;  // pass 1: sorted insertion
        // min(m2, max(m1, a)), as m1 <= m2 always; on edge 0, m1 == m2 == 127
        // and m2 stays
        if constexpr (e > 0) m2 = med3i(a, m1, m2);
        m1 = a < m1 ? a : m1;
    });
    int g1 = m1 - bq;
    g1 = g1 < 0 ? 0 : g1;
    int g2 = m2 - bq;
    g2 = g2 < 0 ? 0 : g2;
    const uint32_t ns = qs ^ ((__popc(qs) & 1u) ? ((1u << D) - 1u) : 0u);
    if (act) {
        SFor<0, D>::run([&](auto ec) {
            constexpr int e = decltype(ec)::value;
            // The argmin is the last edge whose |q| equals min1 (where pass
            // 1's sorted insertion "a <= m1" ends); R_new takes g2 there.
            // A + (R_new - R_old) = q + R_new; the saturating add and the
            // clamp after it equal the clamp alone.
            int nv;
            if constexpr (D1 && edge_deg1<G>(B + e)) {
                // no posterior update (D1 above): only the argmin
                if constexpr (KEEPA) am = (av[e] == m1) ? (uint32_t)e : am;
                return;
            } else if constexpr (KEEPA) {
                const bool at = av[e] == m1;
                am = at ? (uint32_t)e : am;
                const int mag = at ? g2 : g1;
                const int neg = __builtin_amdgcn_sbfe((int)ns, D - 1 - e, 1);
                nv = xad(mag, neg, qv[e]) - neg;
            } else {
                nv = add_r_i8<D>(qv[e], g1, g2, ns, am, e);
            }
            // clamp to [lo, hi] = [-C, C] (C >= 0) as one v_med3 on bounds
            // held in VGPRs: the compiler's max/min pair reads them from
            // SGPRs, and one VALU instruction cannot read two SGPRs
            *adr[e] = (int16_t)med3i(nv, lo, hi);
        });
        mg_put<G, L>(st, (uint32_t)g1 | ((uint32_t)g2 << 8));
        sg_put<G, L>(st.sg, ns | (am << D));
    }
}

// The row's R on edge e from the int8 state (the layer's g1, g2, signs and
// argmin).
template <class G, int L, int e>
__device__ __forceinline__ int state_r_i8(const StateI8<G> &st)
{
    constexpr int D = G::deg(L);
    const uint32_t mw = mg_get<G, L>(st), sw = sg_get<G, L>(st.sg);
    return r_i8<D>((int)(mw & 0xffu), (int)((mw >> 8) & 0xffu), sw, (sw >> D) & 31u, e);
}

// Parity of the hard decisions of row L.  With a D1 state (the int8 layer's
// degree-1 bypass), a degree-1 variable's posterior is its stored channel
// value plus the row's own R on that edge.
template <class G, int ZC, int L, class T, bool WM, class S = void>
__device__ __forceinline__ bool row_parity(const Lane &ln, const Geo<G, ZC, WM> &geo, const S *st = nullptr)
{
    constexpr int D = G::deg(L), B = layer_begin<G>(L);
    T v[D];  // every load of the row issued before the first compare
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        v[e] = *(const T *)geo.template addr<B + e, sizeof(T)>(ln);
    });
    bool p = false;
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        if constexpr (!std::is_void_v<S>) {
            if constexpr (edge_deg1<G>(B + e)) {
                p ^= ((int)v[e] + state_r_i8<G, L, e>(*st)) < 0;
                return;
            }
        }
        p ^= v[e] < (T)0;
    });
    return p;
}

template <class G, int ZC, int L0, int L1, class T, bool WM, class S = void>
__device__ __forceinline__ bool rows_parity(const Lane &ln, const Geo<G, ZC, WM> &geo, const S *st = nullptr)
{
    bool p = false;
    // every row of the segment, without a branch between rows
    SFor<L0, L1>::run([&](auto lc) { p = p | row_parity<G, ZC, decltype(lc)::value, T, WM, S>(ln, geo, st); });
    return p;
}

__device__ __forceinline__ bool wave_any(bool v) { return __builtin_amdgcn_ballot_w64(v) != 0; }

// Syndrome of the lane's check rows (check_syndrome_f32 / _i16,
// src/qc_ldpc.rs:1231-1252, 1766-1787) in layer segments.  A lane drops out
// once it or another lane of its codeword has found a failed check; a wave
// skips the remaining segments once none of its lanes is still checking
// (wave-uniform branches, so no exec-mask bookkeeping).  Returns whether
// this lane found a failed check.
template <class G, int ZC, class T, bool WM, class S = void>
__device__ __forceinline__ bool check_rows(const Lane &ln, const Geo<G, ZC, WM> &geo, bool going, volatile int *cw_fail,
                                           const S *st = nullptr)
{
    constexpr int R = G::ROWS;
    constexpr int S0 = R < 1 ? R : 1, S1 = R < 4 ? R : 4, S2 = R < 10 ? R : 10, S3 = R < 20 ? R : 20,
                  S4 = R < 32 ? R : 32, S5 = R;
    bool fail = false;
    auto segment = [&](auto lo, auto hi) {
        constexpr int A = decltype(lo)::value, B = decltype(hi)::value;
        if (!wave_any(going)) return;
        const bool f = going && rows_parity<G, ZC, A, B, T, WM, S>(ln, geo, st);
        if (f) *cw_fail = 1;
        fail = fail || f;
        going = going && !f && !*cw_fail;
    };
    segment(std::integral_constant<int, 0>{}, std::integral_constant<int, S0>{});
    segment(std::integral_constant<int, S0>{}, std::integral_constant<int, S1>{});
    segment(std::integral_constant<int, S1>{}, std::integral_constant<int, S2>{});
    segment(std::integral_constant<int, S2>{}, std::integral_constant<int, S3>{});
    segment(std::integral_constant<int, S3>{}, std::integral_constant<int, S4>{});
    segment(std::integral_constant<int, S4>{}, std::integral_constant<int, S5>{});
    return fail;
}

// Decode one workgroup's codewords: `cws` codeword slots laid out in LDS
// and over the lanes, `io` saying where the live ones come from and go to.
// I8 selects the int8 path.  Waves without a live lane keep only the
// barriers.  D1 (int8 only) lets the workgroup take the degree-1 bypass of
// layer_i8 when every live codeword's degree-1 channel values allow it.
struct NoHook {
    __device__ __forceinline__ void operator()() const {}
};

// `done` runs once the iterations are over, before the outputs are stored
// (a work-queue workgroup claims its next item there).
template <int KIND, int ZC, bool I8, bool KEEPQ = true, bool WM = true, bool D1 = false, class Hook = NoHook>
__device__ __forceinline__ void fast_group(const SynCode *__restrict__ code, int Z, int cws, const GroupIO &io,
                                           const SynLaunch &a, Hook done = {})
{
    using G = Graph<KIND>;
    using State = std::conditional_t<I8, StateI8<G>, StateF32<G>>;
    constexpr int ESZ = I8 ? 2 : 4;
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = G::COLS * Z;
    const int EZ = G::EDGES * Z;
    const int pad = fast_pad(Z, ESZ);
    unsigned char *P0 = smem + pad;
    CwFlags fl = cw_flags(P0 + ((cws * N * ESZ + 3) & ~3), cws);
    const int t = threadIdx.x;
    const int *__restrict__ shifts = code->shifts;
    load_input<I8>(a, Z, G::COLS, G::ROWS, N, cws, io, P0);
    flags_init(fl, cws, io);
    __syncthreads();

    const bool lane_ok = t < cws * Z;
    const bool wave_live = wave_any(lane_ok && (t / Z) < io.count);
    const int cl = lane_ok ? t / Z : 0;
    const int z = lane_ok ? t - cl * Z : 0;
    const int zb0 = pad + (cl * N + z) * ESZ;
    int wave = 0;  // this lane's wave, for its row of kWrap
    if constexpr (ZC > 0 && WM) wave = __builtin_amdgcn_readfirstlane(t) >> 6;

    State st;  // R = 0 everywhere (qc_ldpc.rs:1006-1009 / 1579-1581)
    if constexpr (I8) {
        SFor<0, (G::ROWS + 1) / 2>::run([&](auto wc) { st.mg[decltype(wc)::value] = 0u; });
    } else {
        SFor<0, G::ROWS>::run([&](auto lc) {
            st.g1[decltype(lc)::value] = 0.f;
            st.g2[decltype(lc)::value] = 0.f;
        });
    }
    SFor<0, sg_words<G>()>::run([&](auto wc) { st.sg[decltype(wc)::value] = 0u; });
    SFor<0, G::ROWS>::run([&](auto lc) { sg_put<G, decltype(lc)::value>(st.sg, 31u << G::deg(decltype(lc)::value)); });

    const float beta = a.beta;
    const int glo = beta != beta ? INT_MIN : 0;  // the floor of |R|'s bits (none for a NaN beta)
    const int g0 = max(__float_as_int(0.f - beta), glo);  // |R| where the min is the FLT_MAX sentinel
    const int bq = a.beta_q, C = a.app_clamp;
    // this lane's LDS addressing (the loop below rebuilds it per iteration)
    auto lane_at = [&](unsigned char *pb) {
        return Lane{pb, pb - Z * ESZ, pb + 65536, pb - Z * ESZ + 65536, z};
    };
    auto geo_at = [&]() {
        const unsigned long long *wm = nullptr;
        if constexpr (ZC > 0 && WM) wm = kWrap<G, ZC>.m[wave];
        return Geo<G, ZC, WM>{Z, shifts, wm};
    };
    auto iterate = [&](auto d1c) {
        constexpr bool BYPASS = decltype(d1c)::value;
        for (int it = 0; it < a.iterations; ++it) {
            if (all_done(fl, cws)) break;
            const bool act = lane_ok && !fl.done[cl];
            if (t < cws && !fl.done[t]) fl.used[t] = it + 1;
            // Opaque per-iteration copies keep the 316 address computations
            // from being hoisted out of the loop into registers; the masks
            // re-establish that the bases are small and non-negative.
            unsigned char *pb = smem + (opaque_v(zb0) & 0x3ffff);
            const Lane ln{pb, pb - Z * ESZ, pb + 65536, pb - Z * ESZ + 65536, opaque_v(z) & 0x3ff};
            const unsigned long long *wm = nullptr;
            if constexpr (ZC > 0 && WM) wm = kWrap<G, ZC>.m[opaque_s(wave)];
            const Geo<G, ZC, WM> geo{Z, shifts + opaque_s(0), wm};
            const int clo = opaque_v(-C), chi = opaque_v(C);  // the int8 posterior clamp's bounds, in VGPRs
            SFor<0, G::ROWS>::run([&](auto lc) {
                constexpr int L = decltype(lc)::value;
                if (wave_live) {
                    if constexpr (I8)
                        layer_i8<G, ZC, L, KEEPQ && KIND != SYN_KIND_BG1, WM, BYPASS>(ln, geo, st, bq, clo, chi, act);
                    else layer_f32<G, ZC, L, KEEPQ>(ln, geo, st, beta, glo, g0, act);
                }
                if constexpr (barrier_after<G>(L)) __syncthreads();
            });
            if constexpr (BYPASS) check_rows<G, ZC, int16_t, WM, State>(ln, geo, act, fl.fail + cl, &st);
            else check_rows<G, ZC, std::conditional_t<I8, int16_t, float>>(ln, geo, act, fl.fail + cl);
            __syncthreads();
            flags_settle(fl, cws);
            __syncthreads();
        }
    };
    if constexpr (I8 && D1) {
        // every degree-1 channel value of the live codewords within C - 127
        const bool live = lane_ok && cl < io.count;
        const Lane ln = lane_at(smem + zb0);
        const Geo<G, ZC, WM> geo = geo_at();
        bool ok = true;
        if (live)
            SFor<0, G::ROWS>::run([&](auto lc) {
                constexpr int L = decltype(lc)::value;
                constexpr int D = G::deg(L), B = layer_begin<G>(L);
                SFor<0, D>::run([&](auto ec) {
                    constexpr int e = decltype(ec)::value;
                    if constexpr (edge_deg1<G>(B + e)) {
                        const int c = *(const int16_t *)geo.template addr<B + e, 2>(ln);
                        ok = ok && (c < 0 ? -c : c) <= C - 127;
                    }
                });
            });
        if (__syncthreads_and(ok ? 1 : 0)) {
            iterate(std::true_type{});
            // the degree-1 posteriors: c + R (no clamp reached)
            if (live)
                SFor<0, G::ROWS>::run([&](auto lc) {
                    constexpr int L = decltype(lc)::value;
                    constexpr int D = G::deg(L), B = layer_begin<G>(L);
                    SFor<0, D>::run([&](auto ec) {
                        constexpr int e = decltype(ec)::value;
                        if constexpr (edge_deg1<G>(B + e)) {
                            int16_t *v = (int16_t *)geo.template addr<B + e, 2>(ln);
                            *v = (int16_t)((int)*v + state_r_i8<G, L, e>(st));
                        }
                    });
                });
            __syncthreads();
        } else {
            iterate(std::false_type{});
        }
    } else {
        iterate(std::false_type{});
    }
    done();
    store_output<I8>(a, N, io, P0);
    store_iters(a, fl, io);
    if ((a.flags & SYN_F_WRITE_EDGE) && lane_ok && cl < io.count) {
        const long long ebase = (long long)io.id(cl) * EZ;
        SFor<0, G::ROWS>::run([&](auto lc) {
            constexpr int L = decltype(lc)::value;
            constexpr int D = G::deg(L), B = layer_begin<G>(L);
            const uint32_t sw = sg_get<G, L>(st.sg);
            SFor<0, D>::run([&](auto ec) {
                constexpr int e = decltype(ec)::value;
                const size_t o = (size_t)ebase + (size_t)(B + e) * Z + z;
                if constexpr (I8) {
                    const uint32_t mw = mg_get<G, L>(st);
                    ((int8_t *)a.edge)[o] =
                        (int8_t)r_i8<D>((int)(mw & 0xffu), (int)((mw >> 8) & 0xffu), sw, (sw >> D) & 31u, e);
                } else {
                    ((float *)a.edge)[o] = r_f32<D>(st.g1[L], st.g2[L], sw, (sw >> D) & 31u, e);
                }
            });
        });
    }
}

template <int KIND, int ZC, bool I8>
__device__ __forceinline__ void fast_body(const SynCode *__restrict__ code, const SynLaunch &a)
{
    const int Z = (ZC > 0) ? ZC : code->z;
    const GroupIO io = contiguous_group(blockIdx.x * a.cws, a.cws, a.batch);
    // the per-config int8 kernels carry the degree-1 bypass where the graph
    // has degree-1 columns
    fast_group<KIND, ZC, I8, true, true, I8 && any_deg1<Graph<KIND>>()>(code, Z, a.cws, io, a);
}

template <int KIND, int ZC>
__global__ void __launch_bounds__(512) fast_f32(const SynCode *__restrict__ code, SynLaunch a)
{
    fast_body<KIND, ZC, false>(code, a);
}

template <int KIND, int ZC>
__global__ void __launch_bounds__(512) fast_i8(const SynCode *__restrict__ code, SynLaunch a)
{
    fast_body<KIND, ZC, true>(code, a);
}

}  // namespace syn

namespace syn {

// Compile-time-Z bodies a work-queue kernel carries: the lifting sizes ZS of
// base graph KIND get the register-resident body with immediate shifts and
// wrap masks (fast_group<KIND, Z>); other sizes run the run-time-Z body.
template <int... ZS>
struct ZSet {};

template <int KIND, bool I8, int Z0, int... ZS, class Hook>
__device__ __forceinline__ bool static_group(int z, const SynCode *__restrict__ code, int count, const GroupIO &io,
                                             const SynLaunch &a, Hook done)
{
    if (z == Z0) {
        fast_group<KIND, Z0, I8, !I8>(code, Z0, count, io, a, done);
        return true;
    }
    if constexpr (sizeof...(ZS) > 0) return static_group<KIND, I8, ZS...>(z, code, count, io, a, done);
    else return false;
}

template <int KIND, bool I8, int... ZS, class Hook>
__device__ __forceinline__ void item_group(ZSet<ZS...>, const SynCode *__restrict__ code, int count, const GroupIO &io,
                                           const SynLaunch &a, Hook done)
{
    const int z = code->z;
    if constexpr (sizeof...(ZS) > 0)
        if (static_group<KIND, I8, ZS...>(z, code, count, io, a, done)) return;
    fast_group<KIND, 0, I8, !I8>(code, z, count, io, a, done);
}

}  // namespace syn
