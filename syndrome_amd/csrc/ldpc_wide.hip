// Decode kernels for raw-edge codes outside the inline code tables
// (SynWideCode): any number of layers, columns and edges, and layers in which
// a column block appears more than once.
//
// Same arithmetic as the plain kernels (src/qc_ldpc.rs:1072-1204 f32,
// 1619-1744 int8), one codeword per workgroup, lane z = lifting index.  The
// posterior sits in LDS when it fits and in a per-codeword HBM workspace
// otherwise (the code is then larger than one workgroup's LDS); R streams
// through the caller's edge buffer.
//
// A layer with a repeated column block ("dup" layer) needs the reference's
// order exactly: every Q of the layer is built from the posterior as it was
// before the layer (Q-build, :1072-1094, runs before pass 2 writes), and
// pass 2 applies the posterior deltas edge by edge (:1181-1204), so a
// variable two edges reach gets (L + d1) + d2 in edge order.  There the
// lanes keep Q in the workspace and pass 2 runs one edge at a time with a
// workgroup barrier between edges (inside one edge, lanes touch distinct
// variables: the shift is a permutation of 0..Z-1).  Every other layer runs
// as in the plain kernel: its lanes touch disjoint variables, and Q is
// rebuilt in pass 2 from the unchanged posterior.
#include "ldpc_device.h"

namespace syn {

int allow_max_dynamic_lds(const void *fn);

namespace {

template <bool I8>
__global__ void __launch_bounds__(SYN_Z_MAX) wide_decode(SynWideCode w, SynLaunch a, char *work, size_t work_stride,
                                                         size_t q_off, int p_in_lds)
{
    using P_t = std::conditional_t<I8, int16_t, float>;
    using M_t = std::conditional_t<I8, int8_t, float>;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int flag_words[3];
    const int Z = w.z, N = w.n_b * Z, cw = blockIdx.x;
    const size_t EZ = (size_t)w.n_edges * Z;
    char *ws = work + (size_t)cw * work_stride;
    P_t *P = p_in_lds ? (P_t *)smem : (P_t *)ws;
    M_t *Q = (M_t *)(ws + q_off);
    M_t *R = (M_t *)a.edge + (size_t)cw * EZ;
    CwFlags fl = cw_flags(flag_words, 1);
    const GroupIO io = contiguous_group(cw, 1, a.batch);
    const int t = threadIdx.x;
    load_input<I8>(a, Z, w.n_b, w.m_b, N, 1, io, P);
    for (size_t i = t; i < EZ; i += blockDim.x) R[i] = 0;  // qc_ldpc.rs:1006-1009, 1579-1581
    flags_init(fl, 1, io);
    __syncthreads();

    const bool lane_ok = t < Z;
    const int z = lane_ok ? t : 0;
    for (int it = 0; it < a.iterations; ++it) {
        if (all_done(fl, 1)) break;
        const bool act = lane_ok && !fl.done[0];
        if (t == 0 && !fl.done[0]) fl.used[0] = it + 1;
        for (int l = 0; l < w.n_layers; ++l) {
            const int lb = w.layer_off[l], le = w.layer_off[l + 1];
            const bool dup = w.dup[l] != 0;
            // Q-build + pass 1 (qc_ldpc.rs:1072-1094, 1155-1178 / 1619-1690)
            using A_t = std::conditional_t<I8, int, float>;
            A_t m1, m2;
            if constexpr (I8) m1 = m2 = 127;
            else m1 = m2 = FLT_MAX;
            uint32_t sx = 0;
            if (act) {
                for (int e = lb; e < le; ++e) {
                    const int v = w.cols[e] * Z + wrap_add(z, w.shifts[e], Z);
                    if constexpr (!I8) {
                        const float q = P[v] - R[(size_t)e * Z + z];
                        if (dup) Q[(size_t)(e - lb) * Z + z] = q;
                        const uint32_t qb = __float_as_uint(q);
                        const float aa = __uint_as_float(qb & 0x7fffffffu);
                        sx ^= qb & 0x80000000u;
                        if (aa <= m1) { m2 = m1; m1 = aa; }
                        else if (aa < m2) m2 = aa;
                    } else {
                        int q = (int)P[v] - (int)R[(size_t)e * Z + z];
                        q = q < -127 ? -127 : (q > 127 ? 127 : q);
                        if (dup) Q[(size_t)(e - lb) * Z + z] = (int8_t)q;
                        const int aa = q < 0 ? -q : q;
                        sx ^= q < 0 ? 1u : 0u;
                        const int mx = m1 > aa ? m1 : aa;
                        m2 = m2 < mx ? m2 : mx;
                        m1 = m1 < aa ? m1 : aa;
                    }
                }
                if constexpr (!I8) {
                    if (m1 == FLT_MAX) m1 = 0.f;
                    if (m2 == FLT_MAX) m2 = m1;
                } else if (le - lb < 2) {  // qc_ldpc.rs:1692-1706
                    if (le == lb) m1 = 0;
                    m2 = m1;
                }
            }
            // a dup layer's writes wait for every lane's Q-build
            if (dup) __syncthreads();
            // pass 2 (qc_ldpc.rs:1181-1204 / 1708-1744)
            for (int e = lb; e < le; ++e) {
                if (act) {
                    const int v = w.cols[e] * Z + wrap_add(z, w.shifts[e], Z);
                    const size_t ri = (size_t)e * Z + z;
                    if constexpr (!I8) {
                        const float Lv = P[v];
                        const float ro = R[ri];
                        const float q = dup ? Q[(size_t)(e - lb) * Z + z] : Lv - ro;
                        const uint32_t qb = __float_as_uint(q);
                        const float aa = __uint_as_float(qb & 0x7fffffffu);
                        const uint32_t sg = sx ^ (qb & 0x80000000u);
                        const float me = (aa == m1) ? m2 : m1;
                        float mag = me - a.beta;
                        mag = (mag < 0.f) ? 0.f : mag;
                        const float rn = __uint_as_float(__float_as_uint(mag) | sg);
                        R[ri] = rn;
                        P[v] = Lv + (rn - ro);
                    } else {
                        const int av = P[v];
                        const int ro = R[ri];
                        int q;
                        if (dup) {
                            q = Q[(size_t)(e - lb) * Z + z];
                        } else {
                            q = av - ro;
                            q = q < -127 ? -127 : (q > 127 ? 127 : q);
                        }
                        const int aa = q < 0 ? -q : q;
                        const unsigned sg = sx ^ (q < 0 ? 1u : 0u);
                        const int me = (aa == m1) ? m2 : m1;
                        int mag = me - a.beta_q;
                        mag = mag < 0 ? 0 : mag;
                        const int rn = sg ? -mag : mag;
                        R[ri] = (int8_t)rn;
                        // saturating i16 add then clamp == clamp of the exact sum
                        int nv = av + (rn - ro);
                        nv = nv < -a.app_clamp ? -a.app_clamp : (nv > a.app_clamp ? a.app_clamp : nv);
                        P[v] = (int16_t)nv;
                    }
                }
                // edge e's deltas land before edge e + 1 reads the posterior
                if (dup) __syncthreads();
            }
            __syncthreads();
        }
        if constexpr (!I8) trace_store(a, P, N, Z, z, act, io, 0, it);
        // end-of-iteration syndrome (qc_ldpc.rs:1231-1252 / 1766-1787)
        if (act) {
            int fail = 0;
            for (int l = 0; l < w.n_layers && !fail; ++l) {
                unsigned p = 0;
                for (int e = w.layer_off[l]; e < w.layer_off[l + 1]; ++e)
                    p ^= P[w.cols[e] * Z + wrap_add(z, w.shifts[e], Z)] < 0 ? 1u : 0u;
                fail = (int)p;
            }
            if (fail) fl.fail[0] = 1;
        }
        __syncthreads();
        flags_settle(fl, 1);
        __syncthreads();
    }
    store_output<I8>(a, N, io, P);
    store_iters(a, fl, io);
}

}  // namespace

// Per-codeword workspace of a wide decode: the posterior (when it does not
// fit LDS) then, for codes with dup layers, the layer's Q messages.
WideShape wide_shape(const SynWideCode &w, bool i8, int lds_limit)
{
    WideShape s;
    const size_t n = (size_t)w.n_b * w.z, pe = i8 ? 2 : 4, me = i8 ? 1 : 4;
    const size_t p_bytes = (n * pe + 255) & ~(size_t)255;
    // the kernel's own static LDS (the per-codeword flags) stays below 64 bytes
    const size_t lds = (n * pe + 15) & ~(size_t)15;
    s.p_in_lds = lds + 64 <= (size_t)lds_limit;
    s.lds = s.p_in_lds ? lds : 0;
    s.q_off = s.p_in_lds ? 0 : p_bytes;
    const size_t q_bytes = w.any_dup ? (((size_t)w.max_deg * w.z * me + 255) & ~(size_t)255) : 0;
    s.work_stride = s.q_off + q_bytes;
    s.block = ((w.z + 63) / 64) * 64;
    return s;
}

int launch_wide(bool i8, const SynWideCode &w, const SynLaunch &a, const WideShape &s, char *work,
                hipStream_t stream)
{
    if (a.batch <= 0) return SYN_OK;
    const void *fn = i8 ? (const void *)wide_decode<true> : (const void *)wide_decode<false>;
    if (int rc = allow_max_dynamic_lds(fn)) return rc;
    SynWideCode wc = w;
    SynLaunch arg = a;
    size_t stride = s.work_stride, qo = s.q_off;
    int pl = s.p_in_lds ? 1 : 0;
    void *args[] = {(void *)&wc, (void *)&arg, (void *)&work, (void *)&stride, (void *)&qo, (void *)&pl};
    if (hipLaunchKernel(fn, dim3((unsigned)a.batch), dim3((unsigned)s.block), args, s.lds, stream) != hipSuccess)
        return SYN_DEVICE_ERROR;
    return SYN_OK;
}

}  // namespace syn
