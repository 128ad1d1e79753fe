// Work-queue (persistent) decode launch and the device rate de-matcher.
#pragma once
#include "ldpc_fast.h"

namespace syn {

// Zero the queue head in the launch's own stream (a fill kernel, not a
// memset node, so the sequence can be captured in a HIP graph).
static __global__ void reset_counter(int *head) { *head = 0; }

template <bool I8>
__global__ void __launch_bounds__(512) persist_decode(const SynCode *__restrict__ codes,
                                                       const SynItem *__restrict__ items, int n_items, int *head,
                                                       const int *__restrict__ ids,
                                                       const long long *__restrict__ offs,
                                                       const int *__restrict__ nfill, SynLaunch a)
{
    persist_body<I8>(codes, items, n_items, head, ids, offs, nfill, a);
}

// Rate de-matching of one code block into the decoder's input
// (rate_dematch_llr, src/rate_matching.rs:307-369, followed by the
// placement of prepare_cb_llr, src/transport_block.rs:473-496):
//   * the E received LLRs are de-interleaved (Qm columns, row-major write,
//     column-major read: e_deint[row*Qm + col] = e[col*(E/Qm) + row]);
//   * entry k lands on the k-th non-filler position of the circular buffer
//     (length Ncb = n_b*Z, n_b = 66 / 50) counted from k0 = {0,17,33,56}*Z
//     (BG1) or {0,13,25,43}*Z (BG2), and repeated positions accumulate in
//     increasing k, as the reference's sequential "+=" does;
//   * the decoder input is [2Z punctured zeros | circular buffer]; filler
//     and unselected positions are 0 (decode_5g pins the fillers).
// With `combine` set the circular buffer is a HARQ accumulator
// (HarqBuffer::combine, src/harq.rs:142-158): each position starts from
// the value it already holds and the new transmission's LLRs are added to
// it in increasing k, exactly as the reference's "+=" into `acc`.  `rv`
// >= 0 overrides every block's redundancy version for this launch.
static __global__ void __launch_bounds__(256) dematch(const float *__restrict__ e_llr, float *__restrict__ out,
                                                      const SynDematchCb *__restrict__ cbs, int count, int combine,
                                                      int rv)
{
  for (int cbi = blockIdx.y; cbi < count; cbi += gridDim.y) {
    const SynDematchCb cb = cbs[cbi];
    const int z = cb.z;
    const int nb = cb.bg == 0 ? 66 : 50, kb = cb.bg == 0 ? 22 : 10;
    const int ncb = nb * z, n = ncb + 2 * z, k = kb * z;
    const int kp = cb.n_filler >= k ? 0 : k - cb.n_filler;
    const int fs = kp >= 2 * z ? kp - 2 * z : 0;  // filler range in the circular buffer
    const int fe = k >= 2 * z ? k - 2 * z : 0;
    const int m = ncb - (fe - fs);                 // selectable positions per cycle
    const int j0tab[2][4] = {{0, 17, 33, 56}, {0, 13, 25, 43}};
    const int k0 = j0tab[cb.bg][(rv >= 0 ? rv : cb.rv) & 3] * z;
    const int rows = cb.qm > 0 ? cb.e / cb.qm : 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int c = p - 2 * z;
        float v = (combine && c >= 0) ? out[cb.out_off + p] : 0.f;
        if (c >= 0 && !(c >= fs && c < fe) && m > 0) {
            // rank of c among the non-filler positions walked from k0
            auto below = [&](int x) {  // non-filler positions in [0, x)
                const int f = x <= fs ? 0 : (x >= fe ? fe - fs : x - fs);
                return x - f;
            };
            const int r = c >= k0 ? below(c) - below(k0) : (below(ncb) - below(k0)) + below(c);
            for (int kk = r; kk < cb.e; kk += m) {
                const int row = kk / cb.qm, col = kk - row * cb.qm;
                v += e_llr[cb.e_off + (long long)col * rows + row];
            }
        }
        out[cb.out_off + p] = v;
    }
  }
}

}  // namespace syn
