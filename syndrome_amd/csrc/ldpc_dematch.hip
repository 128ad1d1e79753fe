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
// chunk, its lanes on consecutive elements, so every launch has as many
// workgroups as the output needs and no empty ones.
#include <hip/hip_runtime.h>

#include <vector>

#include "ldpc_common.h"

namespace syn {

namespace {

constexpr int DM_THREADS = 256;

__global__ void __launch_bounds__(DM_THREADS) dematch(const float *__restrict__ e_llr, float *__restrict__ out,
                                                       const SynDematchCb *__restrict__ cbs,
                                                       const SynDematchChunk *__restrict__ chunks, int n_chunks,
                                                       int combine, int rv)
{
    for (int ci = blockIdx.x; ci < n_chunks; ci += gridDim.x) {
        const SynDematchChunk ch = chunks[ci];
        const SynDematchCb cb = cbs[ch.cb];
        const int z = cb.z;
        const int nb = cb.bg == 0 ? 66 : 50, kb = cb.bg == 0 ? 22 : 10;
        const int ncb = nb * z, n = ncb + 2 * z, k = kb * z;
        const int kp = cb.n_filler >= k ? 0 : k - cb.n_filler;
        const int fs = kp >= 2 * z ? kp - 2 * z : 0;  // filler range in the circular buffer
        const int fe = k >= 2 * z ? k - 2 * z : 0;
        const int m = ncb - (fe - fs);                 // selectable positions per cycle
        const int j0tab[2][4] = {{0, 17, 33, 56}, {0, 13, 25, 43}};
        const int k0 = j0tab[cb.bg][(rv >= 0 ? rv : cb.rv) & 3] * z;
        const int qm = cb.qm, e = cb.e;
        const int rows = qm > 0 ? e / qm : 0;
        const int kmax = e < 2 * m + 1 ? e : 2 * m + 1;  // selections that follow the walk
        const float *el = e_llr + cb.e_off;
        // non-filler positions in [0, x)
        auto below = [&](int x) {
            const int f = x <= fs ? 0 : (x >= fe ? fe - fs : x - fs);
            return x - f;
        };
        const int bk0 = below(k0), bncb = below(ncb);
        const int p_end = ch.p0 + SYN_DEMATCH_CHUNK < n ? ch.p0 + SYN_DEMATCH_CHUNK : n;
        for (int p = ch.p0 + (int)threadIdx.x; p < p_end; p += DM_THREADS) {
            const int c = p - 2 * z;
            float v = (combine && c >= 0) ? out[cb.out_off + p] : 0.f;
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
                    for (int i = 0; i < nk; ++i) qs[i] = (ks[i] % qm) * rows + ks[i] / qm;
                    for (int i = 1; i < nk; ++i)  // insertion sort of at most three
                        for (int j = i; j > 0 && qs[j - 1] > qs[j]; --j) {
                            const int t = qs[j];
                            qs[j] = qs[j - 1];
                            qs[j - 1] = t;
                        }
                    for (int i = 0; i < nk; ++i) v += el[qs[i]];
                }
            }
            out[cb.out_off + p] = v;
        }
    }
}

}  // namespace

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
    const unsigned grid = (unsigned)(n_chunks < (1 << 20) ? n_chunks : (1 << 20));
    hipLaunchKernelGGL(dematch, dim3(grid), dim3(DM_THREADS), 0, stream, e_llr, out, cbs, chunks, n_chunks,
                       combine ? 1 : 0, rv);
    return hipGetLastError() == hipSuccess ? SYN_OK : SYN_DEVICE_ERROR;
}

}  // namespace syn
