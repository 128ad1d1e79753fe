// Rate de-matching on device, writing the decoder's input.
//
// The reference's decode chain: RateMatchCache::rate_dematch_llr_into
// (src/rate_matching.rs:531-551, 675-715) behind HarqBuffer::combine
// (src/harq.rs:142-158), then the placement of prepare_cb_llr
// (src/transport_block.rs:473-496).
//   * Selection k of the circular buffer (length Ncb = n_b*Z, n_b = 66 / 50)
//     is the k-th non-filler position counted from k0 = {0,17,33,56}*Z (BG1)
//     or {0,13,25,43}*Z (BG2), for k <= 2m (m = non-filler positions per
//     cycle); the reference's walk stops there (its 2*Ncb guard) and its
//     table holds position 0 for every later k.
//   * Received LLR q (interleaved order) belongs to selection
//     k = (q % rows) * Qm + q / rows, rows = E / Qm.
//   * Each position adds its LLRs in increasing q, as the reference's
//     "+=" loop over q does: a position has at most three regular
//     selections (r, r+m, r+2m), sorted here by their q; position 0 also
//     takes every selection past the guard, then in one pass over q.
//   * The decoder input is [2Z punctured zeros | circular buffer]; filler
//     and unselected positions keep 0 (decode_5g pins the fillers).
// With `combine` set the circular buffer is a HARQ accumulator: each
// position starts from the value it already holds.  `rv` >= 0 overrides
// every block's redundancy version for this launch.  One thread owns each
// output element, so there are no atomics and the order is fixed.
//
// Work decomposition: a table of chunks of SYN_DEMATCH_CHUNK output
// elements of one code block each (built once per plan); one workgroup per
// chunk, so every launch has as many workgroups as the output needs and no
// empty ones.  A lane owns 4 consecutive elements and writes them with one
// 16-byte store when the block's offset allows.
#include <hip/hip_runtime.h>

#include <vector>

#include "ldpc_common.h"

namespace syn {

namespace {

constexpr int DM_THREADS = 128;
// Workgroups of a launch at most (one per chunk below that; they stride
// over the chunk table above it).
constexpr int DM_GRID = 1 << 20;

__global__ void __launch_bounds__(DM_THREADS) dematch(const float *__restrict__ e_llr, float *__restrict__ out,
                                                       const SynDematchCb *__restrict__ cbs,
                                                       const SynDematchChunk *__restrict__ chunks, int n_chunks,
                                                       int combine, int rv)
{
    for (int ci = blockIdx.x; ci < n_chunks; ci += gridDim.x) {
        const SynDematchChunk ch = chunks[ci];
        const SynDematchCb &cb = cbs[ch.cb];
        const int z = cb.z, n = cb.n, fs = cb.fs, fe = cb.fe, m = cb.m, qm = cb.qm, e = cb.e;
        const int rows = cb.rows, kmax = cb.kmax, bncb = cb.bncb, qsh = cb.qsh;
        const int rvi = (rv >= 0 ? rv : cb.rv) & 3;
        const int k0 = cb.k0[rvi], bk0 = cb.bk0[rvi];
        const long long out_off = cb.out_off;
        const float *el = e_llr + cb.e_off;
        // non-filler positions in [0, x)
        auto below = [&](int x) {
            const int f = x <= fs ? 0 : (x >= fe ? fe - fs : x - fs);
            return x - f;
        };
        // selection k -> received index (k % qm) * rows + k / qm, without a
        // division when qm is a power of two
        auto q_of = [&](int kk) { return qsh >= 0 ? (kk & (qm - 1)) * rows + (kk >> qsh) : (kk % qm) * rows + kk / qm; };
        const int p_end = ch.p0 + SYN_DEMATCH_CHUNK < n ? ch.p0 + SYN_DEMATCH_CHUNK : n;
        // one element: its value before the LLRs, plus its LLRs in q order
        auto elem = [&](int p) {
            const int c = p - 2 * z;
            float v = (combine && c >= 0) ? out[out_off + p] : 0.f;
            if (c >= 0 && m > 0 && qm > 0) {
                const bool sel = !(c >= fs && c < fe);
                int ks[3], nk = 0, r = -1;
                if (sel) {
                    // rank of c among the non-filler positions walked from k0
                    r = c >= k0 ? below(c) - bk0 : (bncb - bk0) + below(c);
                    for (int t = 0; t < 3; ++t)
                        if (r + t * m < kmax) ks[nk++] = r + t * m;
                }
                if (c == 0 && e > 2 * m + 1) {
                    for (int q = 0; q < e; ++q) {
                        const int kk = (q % rows) * qm + q / rows;
                        if (kk > 2 * m || (sel && kk < kmax && (kk - r) % m == 0 && kk >= r)) v += el[q];
                    }
                } else {
                    int qs[3];
                    for (int i = 0; i < nk; ++i) qs[i] = q_of(ks[i]);
                    if (qm > 1)  // with qm == 1, q = k is already increasing
                        for (int i = 1; i < nk; ++i)  // insertion sort of at most three
                            for (int j = i; j > 0 && qs[j - 1] > qs[j]; --j) {
                                const int t = qs[j];
                                qs[j] = qs[j - 1];
                                qs[j - 1] = t;
                            }
                    for (int i = 0; i < nk; ++i) v += el[qs[i]];
                }
            }
            return v;
        };
        // With E <= m a position has at most one selection, k = its rank r,
        // and only when r < E: the general element below reduces to
        // v0 + el[q(r)], v0 = 0 or, when combining, the value already held
        // (the add keeps the reference's "+=", which turns 0 + -0 into +0).
        // The loads are unconditional (index 0 when nothing is selected; E
        // is a multiple of Qm, so q(r) < E for r < E) so that a lane's loads
        // go out together instead of one branch and wait each.
        const bool simple = qm > 0 && m > 0 && e > 0 && e <= m;
        auto elem1 = [&](int p) {
            const int c = p - 2 * z;
            const int r = c >= k0 ? below(c) - bk0 : (bncb - bk0) + below(c);
            const bool take = c >= 0 && !(c >= fs && c < fe) && r < e;
            const float x = el[take ? q_of(r) : 0];
            const float v0 = (combine && c >= 0) ? out[out_off + p] : 0.f;
            return take ? v0 + x : v0;
        };
        // 4 consecutive elements per lane, one 16-byte store when aligned
        const bool vec = (out_off & 3) == 0;
        if (simple && vec && p_end - ch.p0 == SYN_DEMATCH_CHUNK) {
            // a full chunk: every load of the lane's G groups is issued
            // before the first store
            constexpr int G = SYN_DEMATCH_CHUNK / (4 * DM_THREADS);
            float v[G][4];
#pragma unroll
            for (int j = 0; j < G; ++j) {
                const int p = ch.p0 + 4 * ((int)threadIdx.x + j * DM_THREADS);
#pragma unroll
                for (int u = 0; u < 4; ++u) v[j][u] = elem1(p + u);
            }
#pragma unroll
            for (int j = 0; j < G; ++j) {
                const int p = ch.p0 + 4 * ((int)threadIdx.x + j * DM_THREADS);
                *reinterpret_cast<float4 *>(out + out_off + p) = make_float4(v[j][0], v[j][1], v[j][2], v[j][3]);
            }
            continue;
        }
        for (int p = ch.p0 + 4 * (int)threadIdx.x; p < p_end; p += 4 * DM_THREADS) {
            if (simple && vec && p + 3 < p_end) {
                const float v0 = elem1(p), v1 = elem1(p + 1), v2 = elem1(p + 2), v3 = elem1(p + 3);
                *reinterpret_cast<float4 *>(out + out_off + p) = make_float4(v0, v1, v2, v3);
            } else if (vec && p + 3 < p_end) {
                const float v0 = elem(p), v1 = elem(p + 1), v2 = elem(p + 2), v3 = elem(p + 3);
                *reinterpret_cast<float4 *>(out + out_off + p) = make_float4(v0, v1, v2, v3);
            } else {
                for (int j = 0; j < 4 && p + j < p_end; ++j) out[out_off + p + j] = elem(p + j);
            }
        }
    }
}

}  // namespace

// Fills the derived fields of a code block (its circular-buffer geometry,
// TS 38.212 5.4.2.1 as rate_matching.rs:181-260 walks it).
void dematch_derive(SynDematchCb &cb)
{
    const int z = cb.z;
    const int nb = cb.bg == 0 ? 66 : 50, kb = cb.bg == 0 ? 22 : 10;
    const int ncb = nb * z, k = kb * z;
    const int kp = cb.n_filler >= k ? 0 : k - cb.n_filler;
    cb.n = ncb + 2 * z;
    cb.fs = kp >= 2 * z ? kp - 2 * z : 0;
    cb.fe = k >= 2 * z ? k - 2 * z : 0;
    cb.m = ncb - (cb.fe - cb.fs);
    cb.rows = cb.qm > 0 ? cb.e / cb.qm : 0;
    cb.kmax = cb.e < 2 * cb.m + 1 ? cb.e : 2 * cb.m + 1;
    const auto below = [&](int x) { return x - (x <= cb.fs ? 0 : (x >= cb.fe ? cb.fe - cb.fs : x - cb.fs)); };
    cb.bncb = below(ncb);
    cb.qsh = -1;
    for (int b = 0; b < 31; ++b)
        if (cb.qm == (1 << b)) cb.qsh = b;
    static const int j0[2][4] = {{0, 17, 33, 56}, {0, 13, 25, 43}};  // TS 38.212 Table 5.4.2.1-2
    for (int r = 0; r < 4; ++r) {
        cb.k0[r] = j0[cb.bg][r] * z;
        cb.bk0[r] = below(cb.k0[r]);
    }
}

// The chunk table of `count` code blocks with N_i = (bg ? 52 : 68) * z_i.
void dematch_chunks(const SynDematchCb *cbs, int count, std::vector<SynDematchChunk> &out)
{
    out.clear();
    for (int i = 0; i < count; ++i) {
        const int n = (cbs[i].bg == 0 ? 68 : 52) * cbs[i].z;
        for (int p0 = 0; p0 < n; p0 += SYN_DEMATCH_CHUNK) out.push_back(SynDematchChunk{i, p0});
    }
}

int launch_dematch(const float *e_llr, float *out, const SynDematchCb *cbs, const SynDematchChunk *chunks,
                   int n_chunks, bool combine, int rv, hipStream_t stream)
{
    if (n_chunks <= 0) return SYN_OK;
    const unsigned grid = (unsigned)(n_chunks < DM_GRID ? n_chunks : DM_GRID);
    hipLaunchKernelGGL(dematch, dim3(grid), dim3(DM_THREADS), 0, stream, e_llr, out, cbs, chunks, n_chunks,
                       combine ? 1 : 0, rv);
    return hipGetLastError() == hipSuccess ? SYN_OK : SYN_DEVICE_ERROR;
}

}  // namespace syn
