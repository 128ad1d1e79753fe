// Register-resident f32 decode with the degree-1 (extension) columns'
// posteriors in HBM instead of LDS, two codewords per workgroup, for a
// compile-time (base graph, Z): BG1 Z = 384, 352, 320, 288.
//
// Why: a whole BG1 Z=384 posterior is 104,448 B of LDS, so a CU holds one
// such codeword: 6 waves on 4 SIMDs, two SIMDs carrying two waves and two
// carrying one, which caps VALU issue at 75 % (every layer costs two wave
// passes on the busy SIMDs).  BG1 columns 26..67 are degree-1 variables:
// each is in exactly one check row, at shift 0, so variable (col, z) is
// read and written only by lane z, in its row's layer.  Its posterior is
// private to that lane and can live in HBM (in practice L2) with no
// synchronisation, which leaves the 26 core columns (39,936 B) in LDS.
// Two codewords then fit one workgroup's LDS, and at <= 168 VGPRs per lane
// its 12 waves put three on every SIMD.
//
// Per lane, the extension value of a layer is loaded a few extension layers
// ahead (global loads stay in flight across the layer barriers, which wait
// for LDS only), read by pass 1 like any posterior, and written back by
// pass 2.  Whether it is negative (the reference's `llr < 0`) is shifted
// into a per-lane bit record for the syndrome check and the hard decisions.
// The arithmetic and its order are layer_f32's, so the decode is
// bit-identical to the all-LDS kernel and to the reference
// (src/qc_ldpc.rs:1072-1252).
//
// Users: the per-config batch kernel (fastx_f32), the work queue's launch
// for BG1 Z=384 (persistx_f32, ldpc_plan.hip) and the resident serving
// kernel's BG1 Z=384 items (ldpc_resident.hip).
#pragma once
#include "ldpc_fast.h"

namespace syn {

// The extension part of G, computed once: core = the first degree-1
// column (every column from it on is degree-1); ok = the layout this
// kernel relies on (no core column is degree-1, a row has at most one
// degree-1 column, and it sits at shift 0 for every lifting set, so its
// variable is lane z's own element z); per row, whether it has an
// extension edge, that edge's place in the row and the row's position
// among such rows; per extension edge k, its row and column.
struct XInfo {
    int core = 0, count = 0;
    bool ok = true;
    bool has[SYN_MAX_ROWS] = {};
    int pos[SYN_MAX_ROWS] = {};
    int idx[SYN_MAX_ROWS] = {};
    int row[SYN_MAX_ROWS] = {};
    int col[SYN_MAX_ROWS] = {};
};

template <class G>
constexpr XInfo make_xinfo()
{
    XInfo x{};
    int deg[SYN_MAX_COLS] = {};
    for (int e = 0; e < G::EDGES; ++e) ++deg[G::col(e)];
    int c = G::COLS;
    while (c > 0 && deg[c - 1] == 1) --c;
    x.core = c;
    for (int i = 0; i < c; ++i)
        if (deg[i] < 2) x.ok = false;
    int b = 0;
    for (int l = 0; l < G::ROWS; ++l) {
        const int d = G::deg(l);
        x.idx[l] = x.count;
        for (int e = 0; e < d; ++e) {
            if (G::col(b + e) < c) continue;
            if (x.has[l]) x.ok = false;
            for (int s = 0; s < 8; ++s)
                if (G::v(b + e, s) != 0) x.ok = false;
            x.has[l] = true;
            x.pos[l] = e;
            x.row[x.count] = l;
            x.col[x.count] = G::col(b + e);
            ++x.count;
        }
        b += d;
    }
    return x;
}

template <class G>
inline constexpr XInfo kX = make_xinfo<G>();

// The helpers below are consteval: a constexpr function called in an
// ordinary expression may be left to run time, and the GPU would then walk
// the base-graph tables in every layer.

template <class G>
consteval bool row_has_x(int l)
{
    return kX<G>.has[l];
}

template <class G>
consteval int x_pos(int l)
{
    return kX<G>.pos[l];
}

template <class G>
consteval int x_index(int l)
{
    return kX<G>.idx[l];
}

template <class G>
consteval int x_count()
{
    return kX<G>.count;
}

template <class G>
consteval int x_col_of(int k)
{
    return kX<G>.col[k];
}

// Bit record of the extension posteriors' `< 0`, shifted in per extension
// layer in layer order: extension edge k lands in word k / 32, and after
// the pass of a full iteration its bit is (len - 1 - k % 32), len being the
// number of extension edges that word takes.
template <class G>
consteval int x_words()
{
    return (x_count<G>() + 31) / 32;
}

template <class G>
consteval int x_bit(int k)
{
    const int w = k / 32, len = (x_count<G>() - 32 * w) < 32 ? (x_count<G>() - 32 * w) : 32;
    return len - 1 - (k % 32);
}

// Sign words (the packing of sg_pos: each layer's D sign bits and 5-bit
// argmin in one 32-bit word with its neighbours), kept in shift-in form:
// a layer's new field is shifted in at the bottom of its word, one
// v_lshl_or with no mask constant.  After a full pass, the field of layer
// l sits at the summed width of the word's later layers (sgx_rest); while
// a pass runs, every field still to be read sits at the word's total width
// minus its own (it has moved past every other field of the word exactly
// once since it was written).
template <class G>
consteval int sgx_total(int w)
{
    int t = 0;
    for (int l = 0; l < G::ROWS; ++l)
        if (sg_pos<G>(l, true) == w) t += sg_bits<G>(l);
    return t;
}

template <class G>
consteval int sgx_rest(int l)
{
    int t = 0;
    for (int j = l + 1; j < G::ROWS; ++j)
        if (sg_pos<G>(j, true) == sg_pos<G>(l, true)) t += sg_bits<G>(j);
    return t;
}

template <class G, int L>
__device__ __forceinline__ uint32_t sgx_get(const uint32_t *sg)
{
    constexpr int W = sg_pos<G>(L, true);
    return sg[W] >> (sgx_total<G>(W) - sg_bits<G>(L));
}

template <class G, int L>
__device__ __forceinline__ void sgx_put(uint32_t *sg, uint32_t v)
{
    constexpr int W = sg_pos<G>(L, true);
    // opaque: materialise the word now (the compiler would otherwise keep
    // the fields apart through the pass and pack them at the loop latch)
    sg[W] = (uint32_t)opaque_v((int)((sg[W] << sg_bits<G>(L)) | v));
}

template <class G, int L>
__device__ __forceinline__ uint32_t sgx_get_rest(const uint32_t *sg)
{
    return sg[sg_pos<G>(L, true)] >> sgx_rest<G>(L);
}

// A workgroup barrier that orders LDS only: the extension loads and stores
// are lane-private global accesses, and a full fence would make every
// barrier wait for the loads in flight.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Where the layers' |R| pair (g1, g2) lives: in registers, or for
// FASTX_SPILL layers in an LDS slot of the lane (byte offset from the
// lane's column-0 address).  The registers of a 19-edge layer's working
// set plus all 46 pairs would exceed the 168 VGPRs of three waves per
// SIMD.
constexpr int FASTX_SPILL = 13;

template <class G>
consteval int spill_slot(int l)
{
    return l < FASTX_SPILL ? l : -1;
}

template <int ZC, int CORE>
consteval int spill_offset(int slot, int which)
{
    return CORE * ZC * 4 + 16 + (2 * slot + which) * ZC * 4;
}

__device__ __forceinline__ float *lds_at(const Lane &ln, int A)
{
    return (float *)(((A >= 65536) ? ln.pbh : ln.pb) + (A & 0xffff));
}

// Per-lane view of the codeword's extension posteriors in HBM: buffer
// resources over the codeword's input and posterior (out-of-range offsets
// read 0 and drop stores), and the lane's byte offset z*4.
struct XMem {
    __amdgpu_buffer_rsrc_t src, dst;
    int voff;
};

// The column offset goes in the scalar offset; made opaque at each use, or
// the compiler hoists all 42 column constants out of the iteration loop
// into SGPRs, which then spill.
__device__ __forceinline__ float x_load(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, opaque_s(soff), 0));
}

__device__ __forceinline__ void x_store(__amdgpu_buffer_rsrc_t r, int voff, int soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, opaque_s(soff), 0);
}

// One layer, as layer_f32 (src/qc_ldpc.rs:1072-1204), with the row's
// extension posterior (if any) taken from `xin` and written to HBM.  KQ
// keeps q's bits from pass 1 to pass 2 (one operation less per edge, D
// more live registers); without it pass 1 tracks the argmin, as the
// reference does.
template <class G, int ZC, int L, bool WM, bool KQ>
__device__ __forceinline__ void layer_f32x(const Lane &ln, const Geo<G, ZC, WM> &geo, StateF32<G> &st, float beta,
                                           int glo, int g0, float xin, const XMem &xm, uint32_t *rec)
{
    constexpr int D = G::deg(L), B = layer_begin<G>(L);
    constexpr bool HX = row_has_x<G>(L);
    constexpr int XP = HX ? x_pos<G>(L) : -1;  // the extension edge's place in the row
    constexpr int SP = spill_slot<G>(L);
    static_assert(D >= 2 && D <= 27, "row degree: its signs and argmin share one 32-bit word");
    float *adr[D];
    float Lv[D], ro[D];
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        if constexpr (e == XP) {
            Lv[e] = xin;
        } else {
            adr[e] = (float *)geo.template addr<B + e, 4>(ln);
            Lv[e] = *adr[e];
        }
    });
    float o1, o2;
    if constexpr (SP >= 0) {
        o1 = *lds_at(ln, spill_offset<ZC, kX<G>.core>(SP, 0));
        o2 = *lds_at(ln, spill_offset<ZC, kX<G>.core>(SP, 1));
    } else {
        o1 = st.g1[L];
        o2 = st.g2[L];
    }
    const uint32_t os = sgx_get<G, L>(st.sg);
    const uint32_t oam = (os >> D) & 31u;
    // pass 1 on the bit patterns of |q| (as layer_f32)
    uint32_t m1 = 0x7f7fffffu, m2 = 0x7f7fffffu;
    uint32_t qs = 0, am = 31;
    uint32_t qsg[D];
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        ro[e] = r_f32<D>(o1, o2, os, oam, e);
        const uint32_t qb = __float_as_uint(Lv[e] - ro[e]);  // Q-build
        if constexpr (KQ) qsg[e] = qb;
        qs = __builtin_amdgcn_alignbit(qs, qb, 31);
        const uint32_t a = qb & 0x7fffffffu;
        if constexpr (!KQ) am = (a <= m1) ? (uint32_t)e : am;
        if constexpr (e == 0) {
        } else if constexpr (KQ) {
            m2 = med3u(a, m1, m2);
        } else {
            m2 = clampu(a, m1, m2);
        }
        m1 = a < m1 ? a : m1;
    });
    const int g1s = max(__float_as_int(__uint_as_float(m1) - beta), glo);
    const int g2s = max(__float_as_int(__uint_as_float(m2) - beta), glo);
    const int g1i = m1 == 0x7f7fffffu ? g0 : g1s;
    const float g1 = __int_as_float(g1i);
    const float g2 = __int_as_float(m2 == 0x7f7fffffu ? g1i : g2s);
    const uint32_t par = __popc(qs) & 1u;
    const uint32_t ns = qs ^ (par ? ((1u << D) - 1u) : 0u);
    const uint32_t g1p = __float_as_uint(g1) | (par << 31), g2p = __float_as_uint(g2) | (par << 31);
    const float m1f = __uint_as_float(m1);
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        float rn;
        if constexpr (KQ) {
            const bool at = fabsf(__uint_as_float(qsg[e])) == m1f;
            am = at ? (uint32_t)e : am;
            rn = __uint_as_float((qsg[e] & 0x80000000u) ^ (at ? g2p : g1p));
        } else {
            rn = r_f32<D>(g1, g2, ns, am, e);
        }
        const float v = Lv[e] + (rn - ro[e]);  // pass 2: posterior += R_new - R_old
        if constexpr (e == XP) {
            x_store(xm.dst, xm.voff, x_col_of<G>(x_index<G>(L)) * ZC * 4, v);
            constexpr int k = x_index<G>(L);
            // opaque: shift the bit in now, not as a lane mask held to the latch
            rec[k / 32] = (uint32_t)opaque_v((int)((rec[k / 32] << 1) | (v < 0.f ? 1u : 0u)));
        } else {
            *adr[e] = v;
        }
    });
    if constexpr (SP >= 0) {
        *lds_at(ln, spill_offset<ZC, kX<G>.core>(SP, 0)) = g1;
        *lds_at(ln, spill_offset<ZC, kX<G>.core>(SP, 1)) = g2;
    } else {
        st.g1[L] = g1;
        st.g2[L] = g2;
    }
    sgx_put<G, L>(st.sg, ns | (am << D));
}

// Parity of row L's hard decisions: core columns from LDS, the extension
// column from the bit record.
template <class G, int ZC, int L, bool WM>
__device__ __forceinline__ bool row_parity_x(const Lane &ln, const Geo<G, ZC, WM> &geo, const uint32_t *rec)
{
    constexpr int D = G::deg(L), B = layer_begin<G>(L);
    constexpr bool HX = row_has_x<G>(L);
    constexpr int XP = HX ? x_pos<G>(L) : -1;
    float v[D];
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        if constexpr (e != XP) v[e] = *(const float *)geo.template addr<B + e, 4>(ln);
    });
    bool p = false;
    SFor<0, D>::run([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        if constexpr (e != XP) p ^= v[e] < 0.f;
    });
    if constexpr (HX) {
        constexpr int k = x_index<G>(L);
        p ^= ((rec[k / 32] >> x_bit<G>(k)) & 1u) != 0;
    }
    return p;
}

// Rows L0..L1 with no branch between them, their loads issued in blocks
// of about FASTX_CHECK_LOADS edges: a scheduling barrier after each block
// keeps the compiler from hoisting every load of the segment (up to 66)
// ahead of the first compare, which the 168-VGPR budget cannot hold.
constexpr int FASTX_CHECK_LOADS = 20;

template <class G>
consteval bool check_block_end(int l)
{
    int n = 0;
    for (int i = 0; i <= l; ++i) {
        n += G::deg(i);
        if (n >= FASTX_CHECK_LOADS || i == l) {
            if (i == l) return n >= FASTX_CHECK_LOADS;
            n = 0;
        }
    }
    return false;
}

template <class G, int ZC, int L0, int L1, bool WM>
__device__ __forceinline__ bool rows_parity_x(const Lane &ln, const Geo<G, ZC, WM> &geo, const uint32_t *rec)
{
    bool p = false;
    SFor<L0, L1>::run([&](auto lc) {
        constexpr int L = decltype(lc)::value;
        p = p | row_parity_x<G, ZC, L, WM>(ln, geo, rec);
        if constexpr (check_block_end<G>(L)) __builtin_amdgcn_sched_barrier(0);
    });
    return p;
}

// The syndrome (check_syndrome_f32, src/qc_ldpc.rs:1231-1252) in layer
// segments, as check_rows: a wave stops once any lane of the codeword has
// found a failed check.
template <class G, int ZC, bool WM>
__device__ __forceinline__ void check_rows_x(const Lane &ln, const Geo<G, ZC, WM> &geo, const uint32_t *rec,
                                             volatile int *fail)
{
    constexpr int R = G::ROWS;
    constexpr int S1 = R < 1 ? R : 1, S2 = R < 4 ? R : 4, S3 = R < 10 ? R : 10, S4 = R < 20 ? R : 20,
                  S5 = R < 32 ? R : 32;
    bool going = true;
    auto segment = [&](auto lo, auto hi) {
        constexpr int A = decltype(lo)::value, B = decltype(hi)::value;
        if (!wave_any(going)) return;
        const bool f = going && rows_parity_x<G, ZC, A, B, WM>(ln, geo, rec);
        if (f) *fail = 1;
        going = going && !f && !*fail;
    };
    segment(std::integral_constant<int, 0>{}, std::integral_constant<int, S1>{});
    segment(std::integral_constant<int, S1>{}, std::integral_constant<int, S2>{});
    segment(std::integral_constant<int, S2>{}, std::integral_constant<int, S3>{});
    segment(std::integral_constant<int, S3>{}, std::integral_constant<int, S4>{});
    segment(std::integral_constant<int, S4>{}, std::integral_constant<int, S5>{});
    segment(std::integral_constant<int, S5>{}, std::integral_constant<int, R>{});
}

// Bytes of LDS the kernel takes: the pad, the core posterior, two fail
// flags (in 16 bytes), the spilled |R| pairs.
template <int KIND, int ZC>
__host__ __device__ constexpr int fastx_lds_bytes()
{
    return fast_pad(ZC, 4) + spill_offset<ZC, kX<Graph<KIND>>.core>(FASTX_SPILL, 0);
}

// Extension loads in flight per lane (in extension layers).
// (2 and 6 measured the same or slower on BG1 Z=384; 42 extension layers
// must be a multiple of it)
constexpr int FASTX_PF = 3;
// Layers of degree at least this run pass 1 with the argmin (fewer live
// registers); lighter ones keep q for pass 2.
// (keeping q in every layer spilled 80 VGPRs and ran 8 % slower on BG1
// Z=384; tracking the argmin in every layer ran 2 % slower)
constexpr int FASTX_KQ_BELOW = 12;

// Codewords per workgroup.  Two workgroups of 6 waves would fit a CU by
// registers and LDS (the occupancy query says 2), but measured, the second
// one is not placed beside the first (SQ wave-cycles per second stayed at
// one workgroup's; the same code ran 3.36 ms per launch as 6-wave
// workgroups, 2.51 ms as 12-wave ones); one workgroup of two
// codewords has 12 waves, which can only be spread three on every SIMD.
constexpr int FASTX_CPW = 2;

// Lanes of one codeword's slot: Z rounded up to whole waves (a wave holds
// lanes of one codeword only; lanes z >= Z of the last wave idle).
__host__ __device__ constexpr int fastx_lanes(int z) { return (z + 63) & ~63; }

// LDS of one codeword's region (16-byte multiple) and of a workgroup.
template <int KIND, int ZC>
__host__ __device__ constexpr int fastx_region_bytes()
{
    return (fastx_lds_bytes<KIND, ZC>() + 15) & ~15;
}

template <int KIND, int ZC>
__host__ __device__ constexpr int fastx_group_lds_bytes()
{
    return FASTX_CPW * fastx_region_bytes<KIND, ZC>();
}

// Where a workgroup's two codewords come from and go to: per lane, its own
// codeword's slot (whether it exists, element offset of its input /
// posterior / hard decisions, filler count, index for iterations and
// edge_r) and whether the partner slot exists.
struct XSlot {
    bool live, live_o;
    long long nf;
    int id;
    const float *in;  // the codeword's channel input (N elements)
    float *xb;        // N elements the decode may write: the extension columns live here
    float *post;      // the core posterior goes here when not null (xb, when the posterior is wanted)
    uint8_t *hard;    // hard decisions (N elements) when not null
};

template <class T>
__device__ __forceinline__ T *rfl_ptr(T *p)
{
    const unsigned long long v = (unsigned long long)p;
    return (T *)(((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
                 (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v));
}

// FASTX_CPW codewords per workgroup, the lane's codeword slot t / lanes(Z),
// lifting index z = t % lanes(Z); lanes with z >= Z, and in a wider
// work-queue workgroup whole waves past the slots, take part in the
// barriers and the loop's exit only.  Each codeword needs an N-element
// buffer it may write (XSlot::xb: the posterior buffer of a batch launch)
// to hold its extension columns during the decode, and the launch at least
// one iteration.  The codewords share the layer barriers; a converged one stops
// (its waves skip the layers) while its partner goes on, as in
// qc_ldpc.rs:1211 each stops at its own iteration.  `done` runs once the
// iterations are over, before the outputs are stored (a work-queue
// workgroup claims its next item there).
template <int KIND, int ZC, class Hook, class IdFn>
__device__ __forceinline__ void fastx_group(const SynLaunch &a, const XSlot &sl, Hook done_hook, IdFn id_of)
{
    using G = Graph<KIND>;
    static_assert(kX<G>.ok, "degree-1 columns: one per row, shift 0");
    constexpr int C = kX<G>.core, NX = x_count<G>(), PF = FASTX_PF;
    constexpr int NC = C * ZC, N = G::COLS * ZC, K = (G::COLS - G::ROWS) * ZC;
    static_assert(NX >= PF && NX % PF == 0 && NC % 4 == 0, "shape: the prefetch slots repeat every iteration");
    constexpr int RB = fastx_region_bytes<KIND, ZC>();
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int pad = fast_pad(ZC, 4);
    constexpr int LPC = fastx_lanes(ZC);
    const int t = threadIdx.x;
    static_assert(FASTX_CPW == 2, "two codewords per workgroup");
    // idle: a wave past both slots (it follows slot 0's flags, so that its
    // loop ends with the others)
    const bool idle = t >= FASTX_CPW * LPC;
    const int cwl = t >= LPC && !idle ? 1 : 0;
    const int z = idle ? 0 : t - cwl * LPC;
    const bool lane_ok = !idle && (LPC == ZC || z < ZC);  // a lane with a check row
    // a wave's lanes all belong to one codeword: its slot values are
    // wave-uniform, and read as such (scalar registers; the buffer
    // resources below would otherwise be per-lane values, which the
    // compiler serves with a waterfall loop around every access)
    const bool live = __builtin_amdgcn_readfirstlane(sl.live ? 1 : 0) != 0;
    unsigned char *R0 = smem + cwl * RB;  // this codeword's LDS region
    float *P0 = (float *)(R0 + pad);
    volatile int *fail = (volatile int *)(P0 + NC);
    // the partner codeword's fail flags
    volatile int *fail_o = (volatile int *)((float *)(smem + (1 - cwl) * RB + pad) + NC);
    const bool init5g = (a.flags & SYN_F_INIT_5G) != 0;
    const long long nf = sl.nf;
    const int kp = nf >= K ? 0 : K - (int)nf;
    const float *in = rfl_ptr(sl.in);
    float *xb = rfl_ptr(sl.xb);
    float *post = rfl_ptr(sl.post);
    uint8_t *hard = rfl_ptr(sl.hard);

    // core columns -> LDS, with the 5G prologue (it touches no extension
    // column: the punctured 2Z and the fillers lie below K)
    if (live && lane_ok) {
        if ((a.flags & SYN_F_ALIGNED) != 0) {
            constexpr int U = 4, n4 = NC / 4;
            for (int b = z; b < n4; b += U * ZC) {
                float4 v[U];
#pragma unroll
                for (int k = 0; k < U; ++k)
                    if (b + k * ZC < n4) v[k] = *(const float4 *)(in + 4 * (b + k * ZC));
#pragma unroll
                for (int k = 0; k < U; ++k) {
                    const int i4 = b + k * ZC;
                    if (i4 < n4) {
                        const int p = 4 * i4;
                        float4 o;
                        o.x = prologue_f32(v[k].x, p, 2 * ZC, kp, K, init5g);
                        o.y = prologue_f32(v[k].y, p + 1, 2 * ZC, kp, K, init5g);
                        o.z = prologue_f32(v[k].z, p + 2, 2 * ZC, kp, K, init5g);
                        o.w = prologue_f32(v[k].w, p + 3, 2 * ZC, kp, K, init5g);
                        *(float4 *)(P0 + p) = o;
                    }
                }
            }
        } else {
            for (int i = z; i < NC; i += ZC) P0[i] = prologue_f32(in[i], i, 2 * ZC, kp, K, init5g);
        }
    }
    if (!idle && z == 0) fail[0] = 0;

    const int wave = __builtin_amdgcn_readfirstlane(z) >> 6;  // the wave's place in its codeword
    const int zb0 = (int)(R0 - smem) + pad + z * 4;
    StateF32<G> st;  // R = 0 everywhere (qc_ldpc.rs:1006-1009)
    SFor<0, G::ROWS>::run([&](auto lc) {
        constexpr int L = decltype(lc)::value;
        if constexpr (spill_slot<G>(L) < 0) {
            st.g1[L] = 0.f;
            st.g2[L] = 0.f;
        }
    });
    if (lane_ok)
        SFor<0, FASTX_SPILL>::run([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            P0[spill_offset<ZC, C>(j, 0) / 4 + z] = 0.f;
            P0[spill_offset<ZC, C>(j, 1) / 4 + z] = 0.f;
        });
    SFor<0, sg_words<G>()>::run([&](auto wc) { st.sg[decltype(wc)::value] = 0u; });
    SFor<0, G::ROWS>::run([&](auto lc) {
        constexpr int L = decltype(lc)::value;
        st.sg[sg_pos<G>(L, true)] |= (31u << G::deg(L)) << sgx_rest<G>(L);  // argmin "none", signs 0
    });
    uint32_t rec[x_words<G>()];
    SFor<0, x_words<G>()>::run([&](auto wc) { rec[decltype(wc)::value] = 0u; });
    __syncthreads();

    const float beta = a.beta;
    const int glo = beta != beta ? INT_MIN : 0;
    const int g0 = max(__float_as_int(0.f - beta), glo);

    const __amdgpu_buffer_rsrc_t r_in = __builtin_amdgcn_make_buffer_rsrc((void *)in, 0, N * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t r_post = __builtin_amdgcn_make_buffer_rsrc((void *)xb, 0, N * 4, 0x00020000);
    // the first PF extension values of iteration 0
    float xq[PF] = {};
    if (lane_ok)
        SFor<0, PF>::run([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            xq[k] = x_load(r_in, z * 4, x_col_of<G>(k) * ZC * 4);
        });
    // done: this codeword (wave-uniform) and the partner; a codeword past
    // the batch starts done
    bool done = !live;
    bool done_o = !sl.live_o;
    const bool work = lane_ok;  // a lane that decodes (while its codeword is not done)
    int used = 0;
    for (int it = 0; it < a.iterations; ++it) {
        if (!done) used = it + 1;
        unsigned char *pb = smem + (opaque_v(zb0) & 0x3ffff);
        const Lane ln{pb, pb - ZC * 4, pb + 65536, pb - ZC * 4 + 65536, opaque_v(z) & 0x3ff};
        const unsigned long long *wm = nullptr;
        if constexpr (WRAP_WAVES * 64 >= ZC) wm = kWrap<G, ZC>.m[opaque_s(wave)];
        const Geo<G, ZC, true> geo{ZC, nullptr, wm};
        // iteration 0 reads the extension columns from the input, later
        // ones from the posterior the previous pass wrote
        const XMem xm{it == 0 ? r_in : r_post, r_post, opaque_v(z) * 4};
        SFor<0, G::ROWS>::run([&](auto lc) {
            constexpr int L = decltype(lc)::value;
            if (!done && work) {
                float xv = 0.f;
                if constexpr (row_has_x<G>(L)) {
                    constexpr int k = x_index<G>(L), kn = k + PF;
                    xv = xq[k % PF];
                    if constexpr (kn < NX) xq[k % PF] = x_load(xm.src, xm.voff, x_col_of<G>(kn) * ZC * 4);
                    else xq[k % PF] = x_load(r_post, xm.voff, x_col_of<G>(kn - NX) * ZC * 4);  // next iteration
                }
                layer_f32x<G, ZC, L, true, (G::deg(L) < FASTX_KQ_BELOW)>(ln, geo, st, beta, glo, g0, xv, xm, rec);
            }
            if constexpr (barrier_after<G>(L)) lds_barrier();
        });
        if (!idle && z == 0) fail[(it + 1) & 1] = 0;
        if (!done && work) {
            // fresh opaque copies of the lane's bases and mask row for the
            // check: the compiler would otherwise keep the layers' 316 LDS
            // addresses alive (common subexpressions) until the check reads
            // the same variables, far more registers than the budget
            unsigned char *pbc = smem + (opaque_v(zb0) & 0x3ffff);
            const Lane lnc{pbc, pbc - ZC * 4, pbc + 65536, pbc - ZC * 4 + 65536, opaque_v(z) & 0x3ff};
            const unsigned long long *wmc = nullptr;
            if constexpr (WRAP_WAVES * 64 >= ZC) wmc = kWrap<G, ZC>.m[opaque_s(wave)];
            const Geo<G, ZC, true> geoc{ZC, nullptr, wmc};
            check_rows_x<G, ZC, true>(lnc, geoc, rec, fail + (it & 1));
        }
        lds_barrier();
        // zero syndrome: the codeword stops (qc_ldpc.rs:1211); a done
        // codeword never raises its flag again
        done = done || __builtin_amdgcn_readfirstlane(fail[it & 1]) == 0;
        done_o = done_o || fail_o[it & 1] == 0;
        if (done && done_o) break;
    }
    done_hook();
    __syncthreads();
    if (!live || !lane_ok) return;

    // outputs: the core posterior and hard decisions from LDS, the
    // extension hard decisions from the bit record (its posterior is in
    // place already)
    const bool wpost = post != nullptr, whard = hard != nullptr;
    if ((a.flags & SYN_F_ALIGNED) != 0) {
        for (int i4 = z; i4 < NC / 4; i4 += ZC) {
            const float4 v = *(const float4 *)(P0 + 4 * i4);
            if (wpost) *(float4 *)(post + 4 * i4) = v;
            if (whard)
                *(uint32_t *)(hard + 4 * i4) = (v.x < 0.f ? 1u : 0u) | (v.y < 0.f ? 1u << 8 : 0u) |
                                                       (v.z < 0.f ? 1u << 16 : 0u) | (v.w < 0.f ? 1u << 24 : 0u);
        }
    } else {
        for (int i = z; i < NC; i += ZC) {
            const float v = P0[i];
            if (wpost) post[i] = v;
            if (whard) hard[i] = v < 0.f ? 1 : 0;
        }
    }
    if (whard)
        SFor<0, NX>::run([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            hard[x_col_of<G>(k) * ZC + z] = (uint8_t)((rec[k / 32] >> x_bit<G>(k)) & 1u);
        });
    // the codeword's index, fetched again here rather than held through the
    // decode
    const int id = id_of();
    if (z == 0) a.iters[id] = used;
    if (a.flags & SYN_F_WRITE_EDGE) {
        const long long ebase = (long long)id * G::EDGES * ZC;
        SFor<0, G::ROWS>::run([&](auto lc) {
            constexpr int L = decltype(lc)::value;
            constexpr int D = G::deg(L), B = layer_begin<G>(L);
            const uint32_t sw = sgx_get_rest<G, L>(st.sg);
            SFor<0, D>::run([&](auto ec) {
                constexpr int e = decltype(ec)::value;
                float g1, g2;
                if constexpr (spill_slot<G>(L) >= 0) {
                    g1 = P0[spill_offset<ZC, C>(spill_slot<G>(L), 0) / 4 + z];
                    g2 = P0[spill_offset<ZC, C>(spill_slot<G>(L), 1) / 4 + z];
                } else {
                    g1 = st.g1[L];
                    g2 = st.g2[L];
                }
                ((float *)a.edge)[ebase + (long long)(B + e) * ZC + z] = r_f32<D>(g1, g2, sw, (sw >> D) & 31u, e);
            });
        });
    }
}


// A batch launch's buffers for the codeword at element offset `off`: input
// from `in`, the posterior buffer (required) holds the extension columns
// and takes the whole posterior, hard decisions when asked for.
__device__ __forceinline__ void x_slot_io(const SynLaunch &a, XSlot &sl, long long off)
{
    sl.in = (const float *)a.in + off;
    sl.xb = (float *)a.post + off;
    sl.post = sl.xb;
    sl.hard = (a.flags & SYN_F_WRITE_HARD) ? a.hard + off : nullptr;
}

// The per-config kernel: codewords blockIdx.x * FASTX_CPW + t / Z of a
// contiguous batch.
template <int KIND, int ZC>
__global__ void __launch_bounds__(FASTX_CPW * fastx_lanes(ZC)) __attribute__((amdgpu_waves_per_eu(3, 3)))
fastx_f32(const SynCode *__restrict__ code, SynLaunch a)
{
    constexpr int N = Graph<KIND>::COLS * ZC;
    const int cwl = (int)threadIdx.x >= fastx_lanes(ZC) ? 1 : 0;
    const int cw = blockIdx.x * FASTX_CPW + cwl;
    XSlot sl;
    sl.live = cw < a.batch;
    sl.live_o = blockIdx.x * FASTX_CPW + (1 - cwl) < a.batch;
    sl.id = sl.live ? cw : 0;
    sl.nf = a.n_filler;
    x_slot_io(a, sl, (long long)sl.id * N);
    fastx_group<KIND, ZC>(a, sl, NoHook{}, [&]() { return cw; });
}

// The work-queue form: workgroups pull items of one or two BG1 codewords
// of one lifting size in ZS (a plan's arena offsets and filler counts) off
// a device counter until the queue is drained, as persist_body does.  The
// workgroup has the lanes of the largest size's two slots; with a smaller
// size the waves past its slots idle.
template <int... ZS>
constexpr int fastx_max_lanes()
{
    int m = 0;
    ((m = m > FASTX_CPW * fastx_lanes(ZS) ? m : FASTX_CPW * fastx_lanes(ZS)), ...);
    return m;
}

template <int KIND, int... ZS>
__global__ void __launch_bounds__(fastx_max_lanes<ZS...>()) __attribute__((amdgpu_waves_per_eu(3, 3)))
persistx_f32(const SynCode *__restrict__ codes, const SynItem *__restrict__ items, int n_items, int *head,
             const int *__restrict__ ids, const long long *__restrict__ offs, const int *__restrict__ nfill,
             SynLaunch a)
{
    __shared__ int s_item;
    if (threadIdx.x == 0) s_item = atomicAdd(head, 1);
    for (;;) {
        __syncthreads();  // s_item is written; the previous item is done with LDS
        const int w = s_item;
        if (w >= n_items) return;
        const SynItem it = items[w];
        const int zc = codes[it.code].z;
        int next = 0;
        const auto body = [&](auto zk) {
            constexpr int ZC = decltype(zk)::value;
            const int slot = (int)threadIdx.x / fastx_lanes(ZC);  // 2 and up: idle, following slot 0
            const int cwl = slot == 1 ? 1 : 0;
            XSlot sl;
            sl.live = cwl < it.count;
            sl.live_o = (1 - cwl) < it.count;
            const bool mine = slot < FASTX_CPW && sl.live;
            sl.id = mine ? ids[it.first + cwl] : 0;
            sl.nf = mine ? (nfill ? nfill[sl.id] : a.n_filler) : 0;
            x_slot_io(a, sl, mine ? offs[sl.id] : 0);
            fastx_group<KIND, ZC>(
                a, sl,
                [&]() {
                    if (threadIdx.x == 0) next = atomicAdd(head, 1);
                },
                [&]() { return sl.id; });
        };
        ((zc == ZS ? (body(std::integral_constant<int, ZS>{}), 0) : 0), ...);
        __syncthreads();  // every lane has read this item
        if (threadIdx.x == 0) s_item = next;
    }
}

}  // namespace syn
