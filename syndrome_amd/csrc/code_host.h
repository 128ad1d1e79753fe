// Host-side code construction, encoder and channel model.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "ldpc_common.h"

namespace syn {

// QcLdpcParams::new (src/qc_ldpc.rs:335-366): 3GPP BG1 (bg = 0) or BG2 (bg = 1).
int build_code_3gpp(int bg, long long z, SynCode *out);
// The host copy of a code in the wide form (SynWideCode's tables).
struct WideCode {
    std::vector<int> layer_off, cols, shifts;
    std::vector<unsigned char> dup;
    SynWideCode hdr{};  // table pointers unset
};

// QcLdpcParams::from_raw_edges (src/qc_ldpc.rs:412-455) over its whole
// domain: z in 1..=384, every row < m_b, column < n_b, 0 <= shift < z, any
// number of entries, repeated (row, col) pairs allowed.  A code that fits
// SynCode's inline tables (at most SYN_MAX_ROWS rows, SYN_MAX_COLS columns
// and SYN_MAX_EDGES edges, no column twice in a layer) comes back in that
// form and `wide` is cleared; any other goes to `wide` (kind SYN_KIND_WIDE
// in `out`, which then carries the sizes only), or is rejected when `wide`
// is null.  Sizes whose m_b * z, n_b * z or n * z reach 2^31 are rejected.
int build_code_raw(const int *rows, const int *cols, const int *shifts, size_t n, long long m_b, long long n_b,
                   long long z, SynCode *out, WideCode *wide = nullptr);
// wifi_ldpc_matrix (src/wifi_ldpc_tables.rs:455-466).
int build_code_wifi(long long z, long long rate_num, long long rate_den, SynCode *out);

inline size_t code_n(const SynCode &c) { return (size_t)c.n_b * c.z; }
// info_bit_count_5g (src/qc_ldpc.rs:736-739); n_b >= m_b for every code the
// 5G entry points accept (code_5g_ok).
inline size_t code_k(const SynCode &c) { return c.n_b >= c.m_b ? (size_t)(c.n_b - c.m_b) * c.z : 0; }
// init_5g_llr (src/qc_ldpc.rs:711-732) is defined on the code: the filler
// range needs n_b >= m_b and the 2Z punctured positions N >= 2Z.  The
// reference panics on any other code.
inline bool code_5g_ok(const SynCode &c) { return c.n_b >= c.m_b && c.n_b >= 2; }
inline size_t code_ez(const SynCode &c) { return (size_t)c.n_edges * c.z; }
inline size_t code_layer_buf(const SynCode &c) { return (size_t)c.max_deg * c.z; }

// Systematic encoder: codeword = [info (K bits) | parity], H * codeword = 0.
// The parity part of every supported H is invertible, so the codeword is
// unique and equals QcLdpcEncoder::encode (src/qc_ldpc.rs:2592-2614).
// The parity solve splits H's parity columns into a jointly solved core
// (the first `core` block rows, via a precomputed GF(2) inverse) and
// extension rows that each close on an identity block in their own parity
// column (the 3GPP structure); codes without that structure are solved as
// one core.
class Encoder {
public:
    explicit Encoder(const SynCode &c);
    bool ok() const { return ok_; }
    void encode(const uint8_t *info, uint8_t *codeword) const;
    // Block rows solved jointly; rows core() .. m_b-1 close on their own
    // identity block.
    int core() const { return core_; }
    // The core inverse as circulant terms.  The core commutes with the cyclic
    // shift inside every Z-block, so its inverse does too: each Z x Z block
    // (i, j) of the inverse is a sum of permutations P_t, row zi holding a 1
    // at column (zi + t) mod Z.  Terms of block row i are
    // [off[i], off[i+1]) of (blk, shift).  False when the encoder is not ok
    // or the inverse fails the circulant check.
    bool core_terms(std::vector<int> &off, std::vector<int> &blk, std::vector<int> &shift) const;

private:
    SynCode c_;
    bool ok_ = false;
    int core_ = 0, bits_ = 0, words_ = 0;
    std::vector<uint64_t> inv_;  // bits_ x words_ rows of the core inverse
};

// rate_match (src/rate_matching.rs:181-260): E bits from the circular buffer
// of the full codeword (n_b*Z bits; the first 2Z are never sent) starting at
// k0(rv), skipping fillers, then the Qm-column bit interleaver.  bg 0/1.
bool rate_match(const uint8_t *codeword, size_t cw_len, uint8_t *e_out, size_t e_len, int rv, int qm, int bg, int z,
                size_t n_filler);

// AwgnChannel (src/channel_sim.rs:61-215): BPSK over AWGN, LLR = 2 r / sigma^2.
void awgn_transmit(float ebno_db, float code_rate, uint64_t seed, const uint8_t *bits, size_t n, float *llr_out);

}  // namespace syn
