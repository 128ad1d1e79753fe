// Pair-packed int8 LOMS kernel: two codewords per lane, one in each 16-bit
// half of every register.
//
// The int8 decode (src/qc_ldpc.rs:1530-1787) is 16-bit integer arithmetic
// on the posterior and 8-bit on the messages, so two codewords of the same
// code decode in one lane with packed 16-bit instructions (v_pk_add_i16,
// v_pk_sub_i16, v_pk_min_i16, v_pk_max_i16, v_pk_mad_u16, v_pk_ashrrev_i16),
// which halves the VALU instructions per check row.  Lane t owns lifting
// index z = t % Z of codeword pair p = t / Z: codeword 2p in the low halves,
// 2p + 1 in the high halves.  Both codewords use the same shifts, so one
// address serves both, and LDS holds the pair interleaved: the 32-bit word
// at pair p, element col*Z + j is {posterior of codeword 2p, of codeword
// 2p + 1} at that element, so an edge is one 32-bit LDS load and one store.
//
// The check-to-variable messages are kept uncompressed, one byte per edge
// and codeword, four bytes a register: word k holds edges 2k and 2k + 1 of
// both codewords (low half: codeword 2p's R of edge 2k in bits 0..7, of edge
// 2k + 1 in bits 8..15; high half the same for codeword 2p + 1).  Reading a
// message back is then one or two packed shifts instead of the compressed
// form's argmin compare, select and sign extraction.
//
// Exactness, per half, against the reference's scalar int8 kernel:
//  * Q-build q = A - R_old saturates to 16 bits instead of clamping to
//    +-127 (qc_ldpc.rs:1639-1640).  The sign is the same, |q| is computed
//    with a saturating negate, and min1 / min2 start at 127: an |q| above
//    127 moves neither, and "|q| == min1" can then differ only when min1 ends
//    at 127, where min2 == min1 and both choices give the same magnitude.
//  * min_excl is min2 where |q| == min1, else min1 (qc_ldpc.rs:1712-1716),
//    as g2 + (g1 - g2) * min(|q| - min1, 1) in 16-bit arithmetic.
//  * The exclusive sign is the XOR of all q signs of the row (bit 15 of the
//    XOR of the q words) with q's own.
//  * The posterior update is the reference's: A' = clamp(sat16(A + (R_new -
//    R_old)), -C, C) (qc_ldpc.rs:1736-1742), with A re-read from the load.
//  * A codeword that has stopped keeps R and A: for its half, R_new is
//    replaced by R_old before the update, so delta is 0 and A' = A.
#pragma once
#include "ldpc_fast.h"

namespace syn {

typedef short pk_s2 __attribute__((ext_vector_type(2)));
typedef unsigned short pk_u2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ pk_s2 pk_s(uint32_t v) { return __builtin_bit_cast(pk_s2, v); }
__device__ __forceinline__ uint32_t pk_w(pk_s2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ pk_s2 pk_splat(int v) { return pk_s2{(short)v, (short)v}; }

// g2 + dg * min(d, 1) per 16-bit half, as one v_pk_min_u16 and one
// v_pk_mad_u16 (the backend otherwise lowers the pair to a compare and a
// select per half and a repack, five instructions).
__device__ __forceinline__ pk_s2 pk_pick(pk_u2 d, pk_u2 dg, pk_s2 g2, uint32_t one2)
{
    uint32_t s, r;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(s) : "v"(d), "v"(one2));
    asm("v_pk_mad_u16 %0, %1, %2, %3" : "=v"(r) : "v"(dg), "v"(s), "v"(g2));
    return pk_s(r);
}

// R of edge E (both halves) from the message words.
template <int E>
__device__ __forceinline__ pk_s2 pk_r_get(const uint32_t *rw)
{
    const pk_s2 w = pk_s(rw[E / 2]);
    if constexpr (E % 2 == 0) return (pk_s2)(w << 8) >> 8;
    else return w >> 8;
}

// Store R of edge E (both halves; values in [-127, 127]) into its word.
template <int E>
__device__ __forceinline__ void pk_r_put(uint32_t *rw, pk_s2 r)
{
    // v_perm_b32: selector bytes 0..3 pick bytes of the second source, 4..7
    // of the first
    if constexpr (E % 2 == 0) rw[E / 2] = __builtin_amdgcn_perm(pk_w(r), rw[E / 2], 0x03060104u);
    else rw[E / 2] = __builtin_amdgcn_perm(pk_w(r), rw[E / 2], 0x06020400u);
}

// Both edges of word k at once (edges 2k, 2k + 1 of one layer).
template <int E>
__device__ __forceinline__ void pk_r_put2(uint32_t *rw, pk_s2 r0, pk_s2 r1)
{
    static_assert(E % 2 == 0);
    rw[E / 2] = __builtin_amdgcn_perm(pk_w(r1), pk_w(r0), 0x06020400u);
}

// One layer of the pair-packed int8 decode.  dm: the halves of codewords
// that have stopped (0xffff per half); act: at least one half still runs.
// D1: degree-1 variables hold their channel value c for the whole decode
// (the condition and argument are layer_i8's): their Q is c, and their
// posterior c + R is written once at the end.
template <class G, int ZC, int L, bool WM, bool D1>
__device__ __forceinline__ void layer_pk(const Lane &ln, const Geo<G, ZC, WM> &geo, uint32_t *rw, pk_s2 bq,
                                         pk_s2 lo, pk_s2 hi, uint32_t one2, uint32_t dm, bool act)
{
    constexpr int D = G::deg(L), B = layer_begin<G>(L);
    uint32_t *adr[D];
    pk_s2 Av[D], Ro[D], qv[D], av[D];
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        adr[e] = (uint32_t *)geo.template addr<B + e, 4>(ln);
        Av[e] = pk_s(*adr[e]);
    });
    pk_s2 m1 = pk_splat(127), m2 = pk_splat(127);
    uint32_t par = 0;
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        Ro[e] = pk_r_get<B + e>(rw);
        if constexpr (D1 && edge_deg1<G>(B + e)) qv[e] = Av[e];  // Q = c
        else qv[e] = __builtin_elementwise_sub_sat(Av[e], Ro[e]);
        par ^= pk_w(qv[e]);
        av[e] = __builtin_elementwise_max(qv[e], __builtin_elementwise_sub_sat(pk_splat(0), qv[e]));
        if constexpr (e > 0) m2 = __builtin_elementwise_min(m2, __builtin_elementwise_max(m1, av[e]));
        m1 = __builtin_elementwise_min(m1, av[e]);
    });
    if (!act) return;
    const pk_s2 g1 = __builtin_elementwise_max(m1 - bq, pk_splat(0));
    const pk_s2 g2 = __builtin_elementwise_max(m2 - bq, pk_splat(0));
    const pk_u2 dg = (pk_u2)g1 - (pk_u2)g2;
    const pk_s2 pn = pk_s(par) >> 15;  // parity of all q signs, 0 / -1 per half
    pk_s2 rpend = pk_splat(0);
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        // min_excl - beta_q floored: g2 where |q| == min1, else g1
        const pk_s2 mag = pk_pick((pk_u2)(av[e] - m1), dg, g2, one2);
        const pk_s2 n = pn ^ (qv[e] >> 15);
        pk_s2 rn = (mag ^ n) - n;
        rn = pk_s((dm & pk_w(Ro[e])) | (~dm & pk_w(rn)));  // a stopped half keeps R_old
        // messages: two edges of one word in one layer go in together
        constexpr int E = B + e;
        if constexpr (E % 2 == 0 && e + 1 < D) rpend = rn;
        else if constexpr (E % 2 == 1 && e > 0) pk_r_put2<E - 1>(rw, rpend, rn);
        else pk_r_put<E>(rw, rn);
        if constexpr (!(D1 && edge_deg1<G>(E))) {
            pk_s2 a2 = __builtin_elementwise_add_sat(Av[e], rn - Ro[e]);
            a2 = __builtin_elementwise_min(__builtin_elementwise_max(a2, lo), hi);
            *adr[e] = pk_w(a2);
        }
    });
}

// Parity words of rows L0..L1-1, ORed: bit 15 set when a row of the
// low-half codeword fails, bit 31 for the high half.  Under D1 a degree-1
// variable's posterior is its channel value plus the row's R.
template <class G, int ZC, int L0, int L1, bool WM, bool D1>
__device__ __forceinline__ uint32_t rows_parity_pk(const Lane &ln, const Geo<G, ZC, WM> &geo, const uint32_t *rw)
{
    uint32_t f = 0;
    SFor<L0, L1>::run([&](auto lc) {
        constexpr int L = decltype(lc)::value;
        constexpr int D = G::deg(L), B = layer_begin<G>(L);
        uint32_t v[D];
        SFor<0, D>::run([&](auto ec) {
            constexpr int e = decltype(ec)::value;
            v[e] = *(const uint32_t *)geo.template addr<B + e, 4>(ln);
        });
        uint32_t p = 0;
        SFor<0, D>::run([&](auto ec) {
            constexpr int e = decltype(ec)::value;
            if constexpr (D1 && edge_deg1<G>(B + e)) p ^= pk_w(pk_s(v[e]) + pk_r_get<B + e>(rw));
            else p ^= v[e];
        });
        f |= p;
    });
    return f;
}

// Syndrome of the lane's rows for both halves, in the segments of
// check_rows; a half drops out once it, or another lane of its codeword,
// has found a failed check.
template <class G, int ZC, bool WM, bool D1>
__device__ __forceinline__ void check_rows_pk(const Lane &ln, const Geo<G, ZC, WM> &geo, bool go0, bool go1,
                                              volatile int *fail0, volatile int *fail1, const uint32_t *rw)
{
    constexpr int R = G::ROWS;
    constexpr int S0 = R < 1 ? R : 1, S1 = R < 4 ? R : 4, S2 = R < 10 ? R : 10, S3 = R < 20 ? R : 20,
                  S4 = R < 32 ? R : 32, S5 = R;
    auto segment = [&](auto lo, auto hi) {
        constexpr int A = decltype(lo)::value, B = decltype(hi)::value;
        if (!wave_any(go0 || go1)) return;
        const uint32_t f = (go0 || go1) ? rows_parity_pk<G, ZC, A, B, WM, D1>(ln, geo, rw) : 0u;
        const bool f0 = go0 && (f & 0x8000u), f1 = go1 && (f & 0x80000000u);
        if (f0) *fail0 = 1;
        if (f1) *fail1 = 1;
        go0 = go0 && !f0 && !*fail0;
        go1 = go1 && !f1 && !*fail1;
    };
    segment(std::integral_constant<int, 0>{}, std::integral_constant<int, S0>{});
    segment(std::integral_constant<int, S0>{}, std::integral_constant<int, S1>{});
    segment(std::integral_constant<int, S1>{}, std::integral_constant<int, S2>{});
    segment(std::integral_constant<int, S2>{}, std::integral_constant<int, S3>{});
    segment(std::integral_constant<int, S3>{}, std::integral_constant<int, S4>{});
    segment(std::integral_constant<int, S4>{}, std::integral_constant<int, S5>{});
}

// Channel input of `pairs` codeword pairs into LDS, interleaved (word i of
// pair p = {codeword 2p, codeword 2p + 1} at element i), with load_input's
// prologue (5G layout, quantizer, clamp to +-C).  Slots past io.count read
// zeros.  With SYN_F_ALIGNED a lane moves 4 elements of both codewords per
// step, up to 8 steps in flight.
__device__ __forceinline__ void load_input_pk(const SynLaunch &a, int Z, int n_b, int m_b, int N, int pairs,
                                              const GroupIO &io, uint32_t *lds)
{
    const int K = (n_b - m_b) * Z, Z2 = 2 * Z;
    const bool init5g = (a.flags & SYN_F_INIT_5G) != 0;
    const bool qin = (a.flags & SYN_F_QUANTIZE_IN) != 0;
    const int C = a.app_clamp;
    auto elem = [&](int cl, int pos, int raw_q, float raw_f, int kp) {
        const int v = qin ? quantize_one(raw_f, a.scale) : raw_q;
        return (uint32_t)prologue_i8(v, pos, Z2, kp, K, init5g, C) & 0xffffu;
    };
    auto kp_of = [&](int cl) {
        const long long nf = io.n_filler(cl, a.n_filler);
        return nf >= K ? 0 : K - (int)nf;
    };
    if ((a.flags & SYN_F_ALIGNED) && (N & 3) == 0) {
        constexpr int U = 4;
        const int n4 = N >> 2, total4 = pairs * n4, step = (int)blockDim.x;
        for (int b = threadIdx.x; b < total4; b += U * step) {
            float4 vf[U][2];
            uint2 vh[U][2];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const int i4 = b + k * step;
                const int p = i4 / n4;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    vf[k][h] = make_float4(0.f, 0.f, 0.f, 0.f);
                    vh[k][h] = make_uint2(0u, 0u);
                    const int cl = 2 * p + h;
                    if (i4 < total4 && cl < io.count) {
                        if (qin) vf[k][h] = *(const float4 *)(io.src(cl, N, a.in, 4) + 16LL * (i4 - p * n4));
                        else vh[k][h] = *(const uint2 *)(io.src(cl, N, a.in, 2) + 8LL * (i4 - p * n4));
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const int i4 = b + k * step;
                if (i4 >= total4) continue;
                const int p = i4 / n4, pos = 4 * (i4 - p * n4);
                uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int cl = 2 * p + h;
                    if (cl >= io.count) continue;
                    const int kp = kp_of(cl);
                    const float f[4] = {vf[k][h].x, vf[k][h].y, vf[k][h].z, vf[k][h].w};
                    const int r[4] = {(int16_t)(vh[k][h].x & 0xffffu), (int16_t)(vh[k][h].x >> 16),
                                      (int16_t)(vh[k][h].y & 0xffffu), (int16_t)(vh[k][h].y >> 16)};
#pragma unroll
                    for (int q = 0; q < 4; ++q) w[q] |= elem(cl, pos + q, r[q], f[q], kp) << (16 * h);
                }
                *(uint4 *)(lds + 4LL * i4) = make_uint4(w[0], w[1], w[2], w[3]);
            }
        }
        return;
    }
    const int total = pairs * N;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int p = i / N, pos = i - p * N;
        uint32_t w = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int cl = 2 * p + h;
            if (cl >= io.count) continue;
            const float f = qin ? ((const float *)io.src(cl, N, a.in, 4))[pos] : 0.f;
            const int r = qin ? 0 : ((const int16_t *)io.src(cl, N, a.in, 2))[pos];
            w |= elem(cl, pos, r, f, kp_of(cl)) << (16 * h);
        }
        lds[i] = w;
    }
}

// Posterior and hard decisions of the live codewords back to HBM, out of
// the interleaved LDS words.
__device__ __forceinline__ void store_output_pk(const SynLaunch &a, int N, int pairs, const GroupIO &io,
                                                const uint32_t *lds)
{
    const bool post = (a.flags & SYN_F_WRITE_POST) != 0, hard = (a.flags & SYN_F_WRITE_HARD) != 0;
    if ((a.flags & SYN_F_ALIGNED) && (N & 3) == 0) {
        const int n4 = N >> 2, total4 = pairs * n4;
        for (int i4 = threadIdx.x; i4 < total4; i4 += blockDim.x) {
            const int p = i4 / n4, pos = 4 * (i4 - p * n4);
            const uint4 w = *(const uint4 *)(lds + 4LL * i4);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int cl = 2 * p + h;
                if (cl >= io.count) continue;
                const uint32_t sel = h ? 0x07060302u : 0x05040100u;  // the halves of codeword cl
                const uint32_t v0 = __builtin_amdgcn_perm(w.y, w.x, sel), v1 = __builtin_amdgcn_perm(w.w, w.z, sel);
                const long long g = io.off(cl, N) + pos;
                if (post) *(uint2 *)((int16_t *)a.post + g) = make_uint2(v0, v1);
                if (hard && !((io.no_hard >> cl) & 1u))
                    *(uint32_t *)(a.hard + g) = ((v0 >> 15) & 1u) | (((v0 >> 31) & 1u) << 8) |
                                                (((v1 >> 15) & 1u) << 16) | (((v1 >> 31) & 1u) << 24);
            }
        }
        return;
    }
    const int total = pairs * N;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int p = i / N, pos = i - p * N;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int cl = 2 * p + h;
            if (cl >= io.count) continue;
            const int16_t v = (int16_t)(lds[i] >> (16 * h));
            const long long g = io.off(cl, N) + pos;
            if (post) ((int16_t *)a.post)[g] = v;
            if (hard && !((io.no_hard >> cl) & 1u)) a.hard[g] = v < 0 ? 1 : 0;
        }
    }
}

// Decode one workgroup's `cws` codewords (cws / 2 pairs, cws even) of one
// compile-time base graph, pair-packed.  Same contract as fast_group<.., I8>
// (`done` runs before the outputs are stored).
template <int KIND, int ZC, bool WM = true, class Hook = NoHook>
__device__ __forceinline__ void pack_group(const SynCode *__restrict__ code, int Z, int cws, const GroupIO &io,
                                           const SynLaunch &a, Hook done = {})
{
    using G = Graph<KIND>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int pairs = cws / 2;
    const int N = G::COLS * Z;
    const int EZ = G::EDGES * Z;
    const int pad = fast_pad(Z, 4);
    uint32_t *P0 = (uint32_t *)(smem + pad);
    CwFlags fl = cw_flags(P0 + (size_t)pairs * N, cws);
    const int t = threadIdx.x;
    const int *__restrict__ shifts = code->shifts;
    load_input_pk(a, Z, G::COLS, G::ROWS, N, pairs, io, P0);
    flags_init(fl, cws, io);
    __syncthreads();

    const bool lane_ok = t < pairs * Z;
    const int pl = lane_ok ? t / Z : 0;
    const int c0 = 2 * pl, c1 = 2 * pl + 1;
    const bool wave_live = wave_any(lane_ok && c0 < io.count);
    const int z = lane_ok ? t - pl * Z : 0;
    const int zb0 = pad + (pl * N + z) * 4;
    int wave = 0;
    if constexpr (ZC > 0 && WM) wave = __builtin_amdgcn_readfirstlane(t) >> 6;

    uint32_t rw[(G::EDGES + 1) / 2];  // R = 0 everywhere (qc_ldpc.rs:1579-1581)
    SFor<0, (G::EDGES + 1) / 2>::run([&](auto wc) { rw[decltype(wc)::value] = 0u; });

    const int C = a.app_clamp;
    const pk_s2 bq = pk_splat(a.beta_q);
    auto geo_at = [&](int wv) {
        const unsigned long long *wm = nullptr;
        if constexpr (ZC > 0 && WM) wm = kWrap<G, ZC>.m[wv];
        return Geo<G, ZC, WM>{Z, shifts, wm};
    };
    auto iterate = [&](auto d1c) {
        constexpr bool BYPASS = decltype(d1c)::value;
        for (int it = 0; it < a.iterations; ++it) {
            if (all_done(fl, cws)) break;
            const bool a0 = lane_ok && !fl.done[c0], a1 = lane_ok && !fl.done[c1];
            const uint32_t dm = (a0 ? 0u : 0xffffu) | (a1 ? 0u : 0xffff0000u);
            if (t < cws && !fl.done[t]) fl.used[t] = it + 1;
            unsigned char *pb = smem + (opaque_v(zb0) & 0x3ffff);
            const Lane ln{pb, pb - Z * 4, pb + 65536, pb - Z * 4 + 65536, opaque_v(z) & 0x3ff};
            const Geo<G, ZC, WM> geo = geo_at(opaque_s(wave));
            const pk_s2 lo = pk_s(opaque_v((int)pk_w(pk_splat(-C)))), hi = pk_s(opaque_v((int)pk_w(pk_splat(C))));
            const uint32_t one2 = opaque_v(0x00010001);  // {1, 1}, in a VGPR
            SFor<0, G::ROWS>::run([&](auto lc) {
                constexpr int L = decltype(lc)::value;
                if (wave_live) layer_pk<G, ZC, L, WM, BYPASS>(ln, geo, rw, bq, lo, hi, one2, dm, a0 || a1);
                if constexpr (barrier_after<G>(L)) __syncthreads();
            });
            check_rows_pk<G, ZC, WM, BYPASS>(ln, geo, a0, a1, fl.fail + c0, fl.fail + c1, rw);
            __syncthreads();
            flags_settle(fl, cws);
            __syncthreads();
        }
    };
    {
        // degree-1 bypass when every degree-1 channel value of the live
        // codewords is within C - 127 (layer_i8's condition)
        const bool live = lane_ok && c0 < io.count;
        const Lane ln{smem + zb0, smem + zb0 - Z * 4, smem + zb0 + 65536, smem + zb0 - Z * 4 + 65536, z};
        const Geo<G, ZC, WM> geo = geo_at(wave);
        bool ok = true;
        if (live)
            SFor<0, G::ROWS>::run([&](auto lc) {
                constexpr int L = decltype(lc)::value;
                constexpr int D = G::deg(L), B = layer_begin<G>(L);
                SFor<0, D>::run([&](auto ec) {
                    constexpr int e = decltype(ec)::value;
                    if constexpr (edge_deg1<G>(B + e)) {
                        const pk_s2 c = pk_s(*(const uint32_t *)geo.template addr<B + e, 4>(ln));
                        const pk_s2 ac = __builtin_elementwise_max(c, -c);  // |c| <= C, so -c cannot overflow
                        ok = ok && ac.x <= C - 127 && ac.y <= C - 127;
                    }
                });
            });
        if (__syncthreads_and(ok ? 1 : 0)) {
            iterate(std::true_type{});
            if (live)
                SFor<0, G::ROWS>::run([&](auto lc) {
                    constexpr int L = decltype(lc)::value;
                    constexpr int D = G::deg(L), B = layer_begin<G>(L);
                    SFor<0, D>::run([&](auto ec) {
                        constexpr int e = decltype(ec)::value;
                        if constexpr (edge_deg1<G>(B + e)) {
                            uint32_t *v = (uint32_t *)geo.template addr<B + e, 4>(ln);
                            *v = pk_w(pk_s(*v) + pk_r_get<B + e>(rw));
                        }
                    });
                });
            __syncthreads();
        } else {
            iterate(std::false_type{});
        }
    }
    done();
    store_output_pk(a, N, pairs, io, P0);
    store_iters(a, fl, io);
    if ((a.flags & SYN_F_WRITE_EDGE) && lane_ok) {
        SFor<0, G::EDGES>::run([&](auto ec) {
            constexpr int E = decltype(ec)::value;
            const pk_s2 r = pk_r_get<E>(rw);
            if (c0 < io.count) ((int8_t *)a.edge)[(size_t)io.id(c0) * EZ + (size_t)E * Z + z] = (int8_t)r.x;
            if (c1 < io.count) ((int8_t *)a.edge)[(size_t)io.id(c1) * EZ + (size_t)E * Z + z] = (int8_t)r.y;
        });
    }
}

template <int KIND, int ZC>
__global__ void __launch_bounds__(384) __attribute__((amdgpu_waves_per_eu(3))) pack_i8(const SynCode *__restrict__ code, SynLaunch a)
{
    const int Z = (ZC > 0) ? ZC : code->z;
    const GroupIO io = contiguous_group(blockIdx.x * a.cws, a.cws, a.batch);
    pack_group<KIND, ZC>(code, Z, a.cws, io, a);
}

}  // namespace syn
