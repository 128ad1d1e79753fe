/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A plain-C, scalar restatement of the reference's layered offset min-sum
 * (LOMS) QC-LDPC decode path (ThomasGl/syndrome, src/qc_ldpc.rs), its
 * fixed-point front end (src/quantize.rs), the AWGN input generator
 * (src/channel_sim.rs) and a codeword generator.  Only tests/, the smoke
 * check and the benchmark's CPU-baseline leg load it, and only as the
 * checker / the CPU timing baseline.  The product (syndrome_amd/) never
 * links or calls it.
 *
 * Every function cites the reference file:line it follows.
 */
#ifndef SYNDROME_LOMS_ORACLE_H
#define SYNDROME_LOMS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Table capacity of this restatement: the reference's from_raw_edges has no
 * limit but u8 row / column indices; these hold every 3GPP and 802.11 code
 * and the raw codes the tests build (rows past 255 stay empty). */
#define ORC_MAX_ROWS 512
#define ORC_MAX_COLS 1024
#define ORC_MAX_EDGES 4096
#define ORC_Z_MAX 384

/* QcLdpcParams (src/qc_ldpc.rs:233-253). */
typedef struct {
    int z;
    int m_b;                        /* num_row_blocks */
    int n_b;                        /* num_col_blocks */
    int n_edges;                    /* E */
    int max_deg;                    /* max_layer_degree */
    int layer_off[ORC_MAX_ROWS + 1];
    int cols[ORC_MAX_EDGES];
    int shifts[ORC_MAX_EDGES];      /* already reduced mod z */
} orc_code_t;

/* Status codes mirror capi/src/lib.rs:74-94. */
#define ORC_OK 0
#define ORC_INVALID_PARAM (-1)
#define ORC_BUFFER_TOO_SMALL (-4)

/* QcLdpcParams::new (qc_ldpc.rs:335-366): bg 0 = BG1, 1 = BG2. */
int orc_code_3gpp(orc_code_t *code, int bg, int z);
/* QcLdpcParams::from_raw_edges (qc_ldpc.rs:412-455): any m_b, n_b >= 0 up
 * to the capacity above, repeated (row, col) entries kept in order. */
int orc_code_raw(orc_code_t *code, const int *rows, const int *cols, const int *shifts, int n_entries,
                 int m_b, int n_b, int z);
/* wifi_ldpc_decoder / wifi_ldpc_matrix (wifi_ldpc_tables.rs:455-529). */
int orc_code_wifi(orc_code_t *code, int z, int rate_num, int rate_den);

size_t orc_n(const orc_code_t *c);          /* variable_node_count, qc_ldpc.rs:569-571 */
size_t orc_k(const orc_code_t *c);          /* info_bit_count_5g, qc_ldpc.rs:736-739 */
size_t orc_edge_buf(const orc_code_t *c);   /* required_edge_buffer, qc_ldpc.rs:581-583 */
size_t orc_layer_buf(const orc_code_t *c);  /* required_layer_buffer, qc_ldpc.rs:586-588 */

/* f32 decode: decode_layered_offset_min_sum_dispatch, scalar kernel
 * (qc_ldpc.rs:962-1221) with check_syndrome_f32 (1231-1252).  Returns the
 * iterations used (>= 0) or a negative status.  `trace`, if non-NULL,
 * receives a copy of the posterior after every completed iteration
 * (row-major [iteration][n]), the on_iteration hook of
 * decode_layered_offset_min_sum_traced (qc_ldpc.rs:853-874, 1208). */
long orc_decode_f32(const orc_code_t *c, float beta, float *llr, size_t llr_len, float *edge_r,
                    size_t edge_len, uint8_t *hard, size_t hard_len, size_t iterations, float *trace);
/* init_5g_llr (qc_ldpc.rs:711-732). */
int orc_init_5g_f32(const orc_code_t *c, float *llr, size_t llr_len, size_t n_filler);

/* int8 decode: decode_layered_offset_min_sum_i8_dispatch, scalar kernel
 * (qc_ldpc.rs:1530-1758) with check_syndrome_i16 (1766-1787). */
long orc_decode_i8(const orc_code_t *c, float beta, float scale, int16_t app_clamp, int16_t *app,
                   size_t app_len, int8_t *edge_r, size_t edge_len, uint8_t *hard, size_t hard_len,
                   size_t iterations);
/* decode_5g_i8 prologue (qc_ldpc.rs:1486-1523, lines 1505-1514). */
int orc_init_5g_i8(const orc_code_t *c, int16_t *app, size_t app_len, size_t n_filler, int16_t app_clamp);

/* quantize.rs:298-304 */
int8_t orc_beta_q(float beta, float scale);
/* quantize.rs:331-341 */
void orc_quantize_llr_i16(const float *llr, int16_t *out, size_t n, float scale);
/* quantize.rs:148-154 */
void orc_quantize_llr_i8(const float *llr, int8_t *out, size_t n, float scale);
/* quantize.rs:163-169 */
void orc_dequantize_llr_i8(const int8_t *q, float *out, size_t n, float scale);

/* Syndrome of hard bits (0/1 bytes) against H: 1 iff every check holds. */
int orc_syndrome_ok_bits(const orc_code_t *c, const uint8_t *bits);

/* FNV-1a over hard bits (ldpc_bench_export.rs:57-64). */
uint64_t orc_fnv1a_hard(const uint8_t *hard, size_t n);

/* AwgnChannel::new + transmit (channel_sim.rs:61-215). */
float orc_awgn_sigma(float ebno_db, float code_rate);
void orc_awgn_transmit(float ebno_db, float code_rate, uint64_t seed, const uint8_t *bits, size_t n,
                       float *llr_out);

/* Systematic encoder: codeword[0..K) = info, parity solves H c = 0.  The
 * parity is unique because the parity part of H is invertible, so this
 * equals QcLdpcEncoder::encode (qc_ldpc.rs:2592-2614) bit for bit.
 * Returns 0, or -1 if the parity part could not be solved. */
int orc_encode(const orc_code_t *c, const uint8_t *info, uint8_t *codeword);

/* rate_match (rate_matching.rs:181-260): codeword = full N = n_b*Z (68Z or
 * 52Z) bits; e_out gets E bits.  bg 0/1.  Returns 0 or -1. */
int orc_rate_match(const uint8_t *codeword, size_t cw_len, uint8_t *e_out, size_t e_len, int rv, int qm, int bg,
                   int z, size_t n_filler);
/* rate_dematch_llr (rate_matching.rs:307-369): cb_llr [Ncb] accumulates. */
int orc_rate_dematch(const float *e_llr, size_t e_len, float *cb_llr, size_t cb_len, int rv, int qm, int bg, int z,
                     size_t n_filler);
/* RateMatchCache::rate_dematch_llr_into (rate_matching.rs:531-551, 675-715),
 * as HarqBuffer::combine uses it: same selection, but a table lookup per
 * received index q, accumulated in received order, with slots past the walk's
 * 2*Ncb guard mapped to position 0. */
int orc_rate_dematch_cached(const float *e_llr, size_t e_len, float *cb_llr, size_t cb_len, int rv, int qm, int bg,
                            int z, size_t n_filler);

/* CRCs and transport-block assembly (crc.rs, segmentation.rs:361-410,
 * transport_block.rs:963-1044). */
uint32_t orc_crc(const uint8_t *bits, size_t n, int len, uint32_t poly);
int orc_segment(const uint8_t *tb_with_crc, size_t b, size_t c, size_t k_prime, uint8_t *blocks);
int orc_tb_check(const uint8_t *hard, const int64_t *cb_off, size_t c, size_t k_prime, size_t tb_size,
                 uint8_t *cb_ok, uint8_t *tb_out);

/* Knuth multiplicative-hash info bits used by the reference's exporters
 * (ldpc_convergence_export.rs:145-147, ldpc_pipeline_bench.rs:247-250). */
void orc_knuth_info_bits(uint8_t *bits, size_t k);

#ifdef __cplusplus
}
#endif
#endif
