/*
 * syndrome_ldpc.h -- C ABI of the MI355X (gfx950) layered offset min-sum
 * QC-LDPC decoder.
 *
 * Part 1 is a drop-in for the reference's C ABI (ThomasGl/syndrome,
 * capi/src/lib.rs): same symbol names, argument order, argument meaning,
 * status codes and validation order.  Buffers are host memory owned by the
 * caller, as in the reference; the decode itself runs on the GPU.
 *
 * Part 2 extends it with what the reference has in Rust but not in its C
 * ABI (raw-edge / Wi-Fi codes, the non-5G decode, the int8 path) and with
 * device-resident batch entry points (device pointers, many codewords per
 * launch) for callers that keep data in HBM.
 *
 * Part 3 is the work-queue batch decoder (uneven codes in one launch) and
 * the serving loop (continuous batching of code blocks).
 *
 * Every entry point returns an int32 status (below) and never lets a C++
 * exception cross the boundary (capi/src/lib.rs:117-119 `guard`).
 */
#ifndef SYNDROME_LDPC_H
#define SYNDROME_LDPC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SyndromeStatus (capi/src/lib.rs:74-94), plus two device-side codes. */
#define SYNDROME_OK 0
#define SYNDROME_INVALID_PARAM (-1)
#define SYNDROME_CRC_MISMATCH (-2)
#define SYNDROME_DECODER_NOT_CONVERGED (-3) /* reserved; never returned, as in the reference */
#define SYNDROME_BUFFER_TOO_SMALL (-4)
#define SYNDROME_NULL_POINTER (-5)
#define SYNDROME_PANIC (-6)
#define SYNDROME_DEVICE_ERROR (-7)          /* HIP runtime failure (no reference counterpart) */
#define SYNDROME_BUSY (-8)                  /* admission refused: HBM budget / slots exhausted */

typedef struct SyndromeLdpcDecoder SyndromeLdpcDecoder;
typedef struct SyndromeLdpcEncoder SyndromeLdpcEncoder;

/* ======================================================================
 * Part 1: the reference C ABI (capi/src/lib.rs)
 * ====================================================================== */

/* capi/src/lib.rs:148-178.  base_graph 0 = BG1, 1 = BG2; z a 3GPP lifting
 * size (Table 5.3.2-1).  On failure *out_decoder is set to NULL.  The
 * decoder binds to the HIP device current at creation. */
int32_t syndrome_ldpc_decoder_create(int32_t base_graph, size_t z, float offset_beta,
                                     SyndromeLdpcDecoder **out_decoder);
/* capi/src/lib.rs:187-193.  NULL is a no-op. */
void syndrome_ldpc_decoder_destroy(SyndromeLdpcDecoder *decoder);
/* capi/src/lib.rs:202-250 (QcLdpcDecoder size queries, qc_ldpc.rs:569-588). */
size_t syndrome_ldpc_decoder_variable_node_count(const SyndromeLdpcDecoder *decoder);
size_t syndrome_ldpc_decoder_info_bit_count(const SyndromeLdpcDecoder *decoder);
size_t syndrome_ldpc_decoder_required_edge_buffer(const SyndromeLdpcDecoder *decoder);
size_t syndrome_ldpc_decoder_required_layer_buffer(const SyndromeLdpcDecoder *decoder);

/* Host-buffer decodes (decode, decode_traced, decode_5g, decode_i8,
 * decode_5g_i8) may be called concurrently on one handle, as the reference's
 * decoder is Send + Sync (qc_ldpc.rs:480-484): each call takes one of up to 8
 * staging areas of the handle, each with its own stream, so up to 8 calls run
 * side by side and further ones wait for a free area.
 *
 * capi/src/lib.rs:290-343 -> QcLdpcDecoder::decode_5g (qc_ldpc.rs:644-655).
 * llr [N] in/out (channel LLRs in, a-posteriori out), edge_r [>= E*Z] out
 * (final check-to-variable messages), layer_scratch [>= maxdeg*Z] validated
 * only (its contents are left unspecified), hard [N] out, iterations = max
 * layered passes; *out_iterations_used (nullable) = passes run. */
int32_t syndrome_ldpc_decode_5g(SyndromeLdpcDecoder *decoder, float *llr, size_t llr_len, size_t n_filler,
                                float *edge_r, size_t edge_r_len, float *layer_scratch, size_t layer_scratch_len,
                                uint8_t *hard_output, size_t hard_output_len, size_t iterations,
                                size_t *out_iterations_used);

/* Allocation-free syndrome_ldpc_decode_5g.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, code that must not call hipMalloc).  Same arguments,
 * validation, status codes and results as syndrome_ldpc_decode_5g, plus the
 * HIP stream to run on (a hipStream_t; NULL = the default stream).  The
 * codeword is staged in one file-scope device array per GPU of
 * SYNDROME_STATIC_SCRATCH_BYTES, which the library holds until the process
 * ends; no hipMalloc, hipFree or stream creation happens in the call.  A
 * code whose staging (posterior N*4 + messages E*Z*4 + hard decisions N
 * bytes, each rounded up to 256, + 256) exceeds that size is refused with
 * INVALID_PARAM (every 3GPP and 802.11 code fits: BG1 Z=384 needs 616,192
 * bytes).  Calls on one GPU share its array: the library serialises them
 * with a lock, and each returns only after its results are copied back, so
 * nothing else may use the array meanwhile. */
#define SYNDROME_STATIC_SCRATCH_BYTES (640u * 1024u)
int32_t syndrome_ldpc_decode_5g_static(SyndromeLdpcDecoder *decoder, float *llr, size_t llr_len, size_t n_filler,
                                       float *edge_r, size_t edge_r_len, float *layer_scratch,
                                       size_t layer_scratch_len, uint8_t *hard_output, size_t hard_output_len,
                                       size_t iterations, size_t *out_iterations_used, void *stream);

/* capi/src/lib.rs:369-484 (QcLdpcEncoder, qc_ldpc.rs:2408-2614).  Host-side
 * systematic encoder: codeword = [info | parity] with H * codeword = 0. */
int32_t syndrome_ldpc_encoder_create(int32_t base_graph, size_t z, SyndromeLdpcEncoder **out_encoder);
void syndrome_ldpc_encoder_destroy(SyndromeLdpcEncoder *encoder);
size_t syndrome_ldpc_encoder_info_bit_count(const SyndromeLdpcEncoder *encoder);
size_t syndrome_ldpc_encoder_codeword_bit_count(const SyndromeLdpcEncoder *encoder);
int32_t syndrome_ldpc_encode_5g(const SyndromeLdpcEncoder *encoder, const uint8_t *info_bits, size_t info_len,
                                size_t n_filler, uint8_t *codeword, size_t codeword_len);

/* ======================================================================
 * Part 2: extensions (host and device buffers)
 * ====================================================================== */

/* QcLdpcDecoder::from_raw_edges (qc_ldpc.rs:412-455, 555-565): entries are
 * (row_block, col_block, shift) triples, shift already < z, z in 1..=384;
 * any number of entries, repeated (row, col) pairs allowed, m_b and n_b >= 0.
 * INVALID_PARAM for an entry out of range, and for sizes this library does
 * not index: m_b*z, n_b*z or n_entries*z at or above 2^31.  A code with more
 * than 46 row blocks, 68 column blocks or 316 edges, or with a column twice in
 * one layer, decodes on the runtime-structured wide kernel (one codeword per
 * workgroup; a posterior too large for LDS stays in HBM; a layer with a
 * repeated column applies its updates edge by edge, as the reference does).
 * Its batch decodes take stream-ordered device memory for their workspace
 * (hipMallocAsync on the call's stream); such a code has no encoder
 * (encode_batch / encode_batch_device return INVALID_PARAM).  decode_5g /
 * decode_5g_i8 on a code with n_b < m_b or n_b < 2 return PANIC, as the
 * reference's init_5g_llr panics there (qc_ldpc.rs:720-731); the batch
 * INIT_5G flag is then INVALID_PARAM. */
int32_t syndrome_ldpc_decoder_create_raw(const uint8_t *rows, const uint8_t *cols, const int16_t *shifts,
                                         size_t n_entries, size_t num_row_blocks, size_t num_col_blocks, size_t z,
                                         float offset_beta, SyndromeLdpcDecoder **out_decoder);
/* wifi_ldpc_decoder (wifi_ldpc_tables.rs:521-529): z in {27,54,81}, rate
 * rate_num/rate_den in {1/2, 2/3, 3/4, 5/6}. */
int32_t syndrome_ldpc_decoder_create_wifi(size_t z, size_t rate_num, size_t rate_den, float offset_beta,
                                          SyndromeLdpcDecoder **out_decoder);
size_t syndrome_ldpc_decoder_check_node_count(const SyndromeLdpcDecoder *decoder);
float syndrome_ldpc_decoder_offset_beta(const SyndromeLdpcDecoder *decoder);

/* QcLdpcDecoder::decode_layered_offset_min_sum (qc_ldpc.rs:794-810): as
 * decode_5g without the 5G prologue. */
int32_t syndrome_ldpc_decode(SyndromeLdpcDecoder *decoder, float *llr, size_t llr_len, float *edge_r,
                             size_t edge_r_len, float *layer_scratch, size_t layer_scratch_len,
                             uint8_t *hard_output, size_t hard_output_len, size_t iterations,
                             size_t *out_iterations_used);

/* QcLdpcDecoder::decode_layered_offset_min_sum_i8 (qc_ldpc.rs:1345-1363):
 * app [N] i16 in/out, edge_r [>= E*Z] i8 out; QuantParams{scale, app_clamp}
 * (quantize.rs:217-305). */
int32_t syndrome_ldpc_decode_i8(SyndromeLdpcDecoder *decoder, int16_t *app, size_t app_len, int8_t *edge_r,
                                size_t edge_r_len, int8_t *layer_scratch, size_t layer_scratch_len,
                                uint8_t *hard_output, size_t hard_output_len, size_t iterations, float scale,
                                int16_t app_clamp, size_t *out_iterations_used);
/* QcLdpcDecoder::decode_5g_i8 (qc_ldpc.rs:1486-1523). */
int32_t syndrome_ldpc_decode_5g_i8(SyndromeLdpcDecoder *decoder, int16_t *app, size_t app_len, size_t n_filler,
                                   int8_t *edge_r, size_t edge_r_len, int8_t *layer_scratch,
                                   size_t layer_scratch_len, uint8_t *hard_output, size_t hard_output_len,
                                   size_t iterations, float scale, int16_t app_clamp, size_t *out_iterations_used);

/* Batch decode flags. */
#define SYNDROME_BATCH_INIT_5G 1u      /* apply the decode_5g / decode_5g_i8 prologue */
#define SYNDROME_BATCH_WRITE_POST 2u   /* write posteriors to d_post */
#define SYNDROME_BATCH_WRITE_HARD 4u   /* write hard decisions to d_hard */
#define SYNDROME_BATCH_WRITE_EDGE 8u   /* write final R messages to d_edge */
#define SYNDROME_BATCH_QUANTIZE 16u    /* int8 only: d_in is f32 LLR, quantize_llr_i16 fused into the load */
#define SYNDROME_BATCH_PLAIN 256u      /* force the runtime-structured kernel (R streamed through HBM) */

/* Device-resident batch decode, f32: `batch` codewords stacked in HBM.
 * d_in f32 [batch][N]; d_post f32 [batch][N] (may equal d_in); d_edge f32
 * [batch][E*Z] (required with SYNDROME_BATCH_PLAIN or WRITE_EDGE, else may
 * be NULL); d_hard u8 [batch][N]; d_iters i32 [batch] (required).  Each
 * codeword gets exactly the result of QcLdpcDecoder::decode_5g (INIT_5G) or
 * decode_layered_offset_min_sum on it alone.  Asynchronous on `stream`
 * (a hipStream_t; NULL = default stream). */
int32_t syndrome_ldpc_decode_batch_f32(const SyndromeLdpcDecoder *decoder, const float *d_in, float *d_post,
                                       float *d_edge, uint8_t *d_hard, int32_t *d_iters, size_t batch,
                                       size_t iterations, size_t n_filler, uint32_t flags, void *stream);
/* Device-resident batch decode, int8 messages / i16 posterior.  d_in is i16
 * app [batch][N], or f32 LLR with SYNDROME_BATCH_QUANTIZE; d_post i16;
 * d_edge i8. */
int32_t syndrome_ldpc_decode_batch_i8(const SyndromeLdpcDecoder *decoder, const void *d_in, int16_t *d_post,
                                      int8_t *d_edge, uint8_t *d_hard, int32_t *d_iters, size_t batch,
                                      size_t iterations, size_t n_filler, float scale, int16_t app_clamp,
                                      uint32_t flags, void *stream);

/* Launch geometry the batch entry points use for this decoder:
 * codewords per workgroup, threads per workgroup, LDS bytes. */
int32_t syndrome_ldpc_launch_shape(const SyndromeLdpcDecoder *decoder, int32_t int8, int32_t plain,
                                   int32_t *out_codewords_per_group, int32_t *out_threads, int32_t *out_lds_bytes);

/* QcLdpcDecoder::decode_layered_offset_min_sum_traced (qc_ldpc.rs:853-874):
 * as syndrome_ldpc_decode, and row i of trace [iterations][N] receives the
 * posterior after layered pass i + 1, before that pass's syndrome check (the
 * reference's on_iteration(i + 1, llr), qc_ldpc.rs:1208).  Rows
 * 0 .. *out_iterations_used - 1 are written; a binding calls its callback
 * once per row.  trace_len < iterations * N -> BUFFER_TOO_SMALL. */
int32_t syndrome_ldpc_decode_traced(SyndromeLdpcDecoder *decoder, float *llr, size_t llr_len, float *edge_r,
                                    size_t edge_r_len, float *layer_scratch, size_t layer_scratch_len,
                                    uint8_t *hard_output, size_t hard_output_len, size_t iterations, float *trace,
                                    size_t trace_len, size_t *out_iterations_used);
/* quantize_llr_i16 (quantize.rs:331-341) on device: f32 [n] -> i16 [n]. */
int32_t syndrome_ldpc_quantize_i16_device(const float *d_llr, int16_t *d_out, size_t n, float scale, void *stream);
/* quantize_llr (quantize.rs:148-154) on device: f32 [n] -> i8 [n],
 * round half away from zero, clamp to +-127 (never -128), NaN -> 0. */
int32_t syndrome_ldpc_quantize_i8_device(const float *d_llr, int8_t *d_out, size_t n, float scale, void *stream);
/* dequantize_llr (quantize.rs:163-169) on device: i8 [n] -> f32 [n] as
 * q * (1 / scale). */
int32_t syndrome_ldpc_dequantize_i8_device(const int8_t *d_q, float *d_out, size_t n, float scale, void *stream);

/* ======================================================================
 * Part 3: host helpers for traffic generation
 * ====================================================================== */

/* AwgnChannel::new(ebno_db, code_rate, seed).transmit(bits) (channel_sim.rs:139-215). */
int32_t syndrome_awgn_transmit(float ebno_db, float code_rate, uint64_t seed, const uint8_t *bits, size_t n,
                               float *llr_out);
/* Systematic encode of `count` info blocks (info [count][K] -> codeword
 * [count][N]) for any decoder's code (3GPP, Wi-Fi or raw), `threads` host
 * threads (0 = all). */
int32_t syndrome_ldpc_encode_batch(const SyndromeLdpcDecoder *decoder, const uint8_t *info, uint8_t *codewords,
                                   size_t count, int32_t threads);
/* The same encode on device (QcLdpcEncoder::encode, qc_ldpc.rs:2592-2614,
 * whose 3GPP path is encode_sparse, qc_ldpc.rs:2620-2727): d_info
 * [count][K] and d_codewords [count][N] are device buffers of one byte per
 * bit; info bytes are copied through, parity bits are 0/1 computed from bit
 * 0 of each info byte.  Asynchronous on `stream` (NULL = default stream);
 * the first call on a handle builds the device tables (one allocation).
 * INVALID_PARAM for a raw code whose core inverse has more than 1024
 * circulant terms (every 3GPP and 802.11 code has at most 180). */
int32_t syndrome_ldpc_encode_batch_device(const SyndromeLdpcDecoder *decoder, const uint8_t *d_info,
                                          uint8_t *d_codewords, size_t count, void *stream);

/* ======================================================================
 * Part 4: work-queue batches of mixed codes (one persistent launch)
 * ====================================================================== */

typedef struct SyndromeLdpcPlan SyndromeLdpcPlan;

/* A plan for `count` code blocks: block i is 3GPP base graph bg[i] (0 = BG1,
 * 1 = BG2) at lifting size z[i] with n_filler[i] filler bits (NULL = all 0);
 * its N_i = (bg ? 52 : 68) * z[i] decoder-input elements start at element
 * offsets[i] of the caller's arenas (input, posterior and hard decisions use
 * the same offsets; iterations are indexed by i).  Blocks are grouped by
 * code into work items of as many codewords as fit a workgroup of `block`
 * lanes (0 = default) and handed out longest first by a device queue.
 * Decode semantics per block: decode_5g (SYNDROME_BATCH_INIT_5G) or the
 * plain decode, exactly as if it were decoded alone.  No reference
 * counterpart: the reference decodes one code block per call. */
int32_t syndrome_ldpc_plan_create(const int32_t *bg, const int32_t *z, const int32_t *n_filler,
                                  const int64_t *offsets, size_t count, float offset_beta, int32_t int8,
                                  int32_t block, SyndromeLdpcPlan **out_plan);
void syndrome_ldpc_plan_destroy(SyndromeLdpcPlan *plan);
/* The plan's persistent launches: one per group of its codes (codes of
 * Z > 256; codes of Z <= 256 in 256-lane workgroups; f32 BG2 codes of
 * Z <= 256; f32 BG1 Z=384 blocks two per workgroup when the posterior is
 * written back).  An eager call runs them on two streams at once: `stream`
 * and the plan's own stream, which waits for `stream` and which `stream`
 * then waits for, so the call is ordered on `stream` as one launch would
 * be.  A call under stream capture issues them in
 * sequence on `stream`.  int8
 * plans take i16 (or f32 with SYNDROME_BATCH_QUANTIZE) inputs and write i16
 * posteriors.  Each call takes the next of 16 queue counters, so up to 16
 * runs of one plan may be in flight at once (on different streams, or a
 * graph replay beside eager runs); a 17th issued while the first still runs
 * would share its counter.  A HIP graph captured from plan_run keeps the
 * counter of its capture: replay one graph on one stream at a time.
 * tune, set_dematch and set_tbs replace the plan's device tables but keep
 * the old ones until plan_destroy, so a graph captured before them still
 * replays, with the launch shape and tables it was captured with. */
int32_t syndrome_ldpc_plan_run(SyndromeLdpcPlan *plan, const void *d_in, void *d_post, uint8_t *d_hard,
                               int32_t *d_iters, size_t iterations, float scale, int16_t app_clamp, uint32_t flags,
                               void *stream);
/* The main launch's block, grid and LDS bytes (or the first launch that has
 * items, when the main launch has none); items of all launches. */
int32_t syndrome_ldpc_plan_shape(const SyndromeLdpcPlan *plan, int32_t *out_block, int32_t *out_grid,
                                 int32_t *out_items, int32_t *out_lds_bytes);
/* Launch-shape tuner: times the plan at each workgroup size from the
 * smallest that holds its largest Z up to 512 lanes, on the caller's data,
 * and keeps the fastest (written to *out_block). */
int32_t syndrome_ldpc_plan_tune(SyndromeLdpcPlan *plan, const void *d_in, void *d_post, uint8_t *d_hard,
                                int32_t *d_iters, size_t iterations, float scale, int16_t app_clamp, uint32_t flags,
                                int32_t reps, int32_t *out_block);
/* Rate de-matching on device, as the reference's decode chain does it
 * (RateMatchCache::rate_dematch_llr_into, rate_matching.rs:531-551 and
 * 675-715, into a fresh HARQ buffer, then the input placement of
 * prepare_cb_llr, transport_block.rs:473-496): block i's e_lengths[i]
 * received LLRs at element e_offsets[i] of d_e_llr, with modulation order
 * qm[i] and redundancy version rv[i] (NULL = 1 / 0), become its N_i-element
 * decoder input at offsets[i] of d_llr_out.  Repetitions are added in
 * received order; selections past the reference walk's 2*Ncb guard land on
 * circular-buffer position 0, as there.  set_dematch synchronises with the
 * device (it replaces the block table). */
int32_t syndrome_ldpc_plan_set_dematch(SyndromeLdpcPlan *plan, const int64_t *e_offsets, const int32_t *e_lengths,
                                       const int32_t *qm, const int32_t *rv);
int32_t syndrome_ldpc_plan_dematch(SyndromeLdpcPlan *plan, const float *d_e_llr, float *d_llr_out, void *stream);
/* Transport-block checks on device (DlSchDecoder::decode after its
 * code-block loop, transport_block.rs:963-1044; CRCs per crc.rs): transport
 * block t is the plan's code blocks [tb_first_cb[t], + tb_num_cb[t]), which
 * share K' = K - n_filler; each contributes P = K' - 24 payload bits (K'
 * when alone), the reassembled stream carries tb_size[t] bits and then their
 * CRC-24A.  set_tbs validates (equal K', num * P >= tb_size + 24) and
 * uploads the table (synchronising with the device); check_tbs reads the
 * decoded hard bits at the plan's offsets of d_hard and writes
 * d_cb_ok[cb] (CRC-24B passed, or no CRC-24B), d_tb_ok[t] (every block ok
 * and CRC-24A passed) and, only when d_tb_ok[t], the tb_size[t] payload
 * bits (one per byte) at element tb_out_off[t] of d_tb_out. */
int32_t syndrome_ldpc_plan_set_tbs(SyndromeLdpcPlan *plan, size_t n_tb, const int32_t *tb_first_cb,
                                   const int32_t *tb_num_cb, const int32_t *tb_size, const int64_t *tb_out_off);
int32_t syndrome_ldpc_plan_check_tbs(SyndromeLdpcPlan *plan, const uint8_t *d_hard, uint8_t *d_tb_out,
                                     uint8_t *d_tb_ok, uint8_t *d_cb_ok, void *stream);
/* HARQ combining on device (HarqBuffer::combine, harq.rs:142-158, feeding
 * prepare_cb_llr, transport_block.rs:473-496): as syndrome_ldpc_plan_dematch,
 * but each block's de-matched LLRs are added, in the reference's order, to
 * what its decoder-input slot of d_acc already holds ([2Z zeros | Ncb
 * accumulated LLRs]), so d_acc is at once the HARQ buffer and the decoder
 * input.  rv 0..3 overrides every block's redundancy version for this
 * transmission; rv < 0 keeps set_dematch's.  Flush = zero the slot. */
int32_t syndrome_ldpc_plan_harq_combine(SyndromeLdpcPlan *plan, const float *d_e_llr, float *d_acc, int32_t rv,
                                        void *stream);

/* ======================================================================
 * Part 5: serving loop (continuous batching of code blocks on one GPU)
 * ====================================================================== */

/* The protocol is the reference pipeline's (LdpcPipeline::acquire / submit /
 * try_recv / release, src/ldpc_pipeline.rs:499-576); the engine behind it is
 * the GPU: admission by an HBM budget (fixed slots sized for 68 x 384),
 * a dispatcher that launches everything submitted since the last batch as
 * soon as one of `streams` launch lanes is free (so arrivals join the next
 * batch while earlier ones run, and batches overlap on the device), and the
 * work-queue decoder for mixed codes.  Each code block is decoded with
 * decode_5g semantics (its n_filler) at `iterations`, exactly as alone. */
typedef struct SyndromeServer SyndromeServer;

typedef struct {
    uint32_t slot;
    uint64_t tag;
    int32_t iterations;   /* passes used; -1 if its batch failed to launch */
    const uint8_t *hard;  /* n hard decisions, valid until release */
    size_t n;
    double submit_ns;     /* steady-clock ns at submit */
    double done_ns;       /* steady-clock ns when its batch was seen complete */
} SyndromeServed;

/* hbm_budget_bytes buys one slot per 68 * 384 * 4 + 64 bytes (an f32 input
 * of the largest code); 1 <= max_batch <= 65535.  streams 1..8: batches of
 * what is pending are launched on that many streams; streams == 0: the
 * resident decoder, one grid that stays on the device while work flows and
 * pulls code blocks from a queue in pinned memory as they are submitted
 * (leaving after 2 ms without work, relaunched on the next submission). */
int32_t syndrome_server_create(float offset_beta, size_t iterations, size_t hbm_budget_bytes, int32_t max_batch,
                               int32_t streams, SyndromeServer **out_server);
void syndrome_server_destroy(SyndromeServer *server);
/* A slot for one code block of (base_graph, z): *out_llr points at its
 * N = (bg ? 52 : 68) * z decoder-input floats (pinned host memory) to fill
 * before submit.  SYNDROME_BUSY when the HBM budget is exhausted. */
int32_t syndrome_server_acquire(SyndromeServer *server, int32_t base_graph, size_t z, size_t n_filler,
                                float **out_llr, uint32_t *out_slot);
int32_t syndrome_server_submit(SyndromeServer *server, uint32_t slot, uint64_t tag);
/* Submit with the input read from `src` instead of the slot's staging:
 * device-accessible memory (syndrome_host_alloc or HBM) holding the block's
 * N floats, untouched until the block completes.  No CPU copy. */
int32_t syndrome_server_submit_from(SyndromeServer *server, uint32_t slot, const float *src, uint64_t tag);
/* One completed code block, or SYNDROME_BUSY when none is ready. */
int32_t syndrome_server_try_recv(SyndromeServer *server, SyndromeServed *out);
int32_t syndrome_server_release(SyndromeServer *server, uint32_t slot);

/* Batched forms of acquire / submit_from / try_recv / release: one server
 * lock per call instead of one per code block, for front ends that move
 * many blocks at a time (LdpcPipeline::submit / try_recv,
 * ldpc_pipeline.rs:536-589, called in a loop).
 * acquire_many: slots for blocks i < n of (base_graph[i], z[i],
 * n_filler[i] or 0), in order, until the budget runs out; *out_acquired =
 * how many (SYNDROME_BUSY when none); out_llr may be NULL.
 * submit_many: srcs NULL, or srcs[i] NULL, keeps slot i's pinned staging;
 * tags may be NULL (0).  All slots are checked first: a slot out of range,
 * not acquired or listed twice submits nothing (INVALID_PARAM).
 * try_recv_many: up to cap completed blocks; *out_n = how many
 * (SYNDROME_BUSY when none; DEVICE_ERROR if any carries iterations < 0).
 * release_many: all-or-nothing like submit_many. */
int32_t syndrome_server_acquire_many(SyndromeServer *server, const int32_t *base_graph, const int32_t *z,
                                     const int32_t *n_filler, size_t n, float **out_llr, uint32_t *out_slots,
                                     size_t *out_acquired);
int32_t syndrome_server_submit_many(SyndromeServer *server, const uint32_t *slots, const float *const *srcs,
                                    const uint64_t *tags, size_t n);
int32_t syndrome_server_try_recv_many(SyndromeServer *server, SyndromeServed *out, size_t cap, size_t *out_n);
int32_t syndrome_server_release_many(SyndromeServer *server, const uint32_t *slots, size_t n);

/* Transport blocks through the resident server (streams == 0): the end of
 * DlSchDecoder::decode (transport_block.rs:963-1044) on device.  The caller
 * acquires num_cb slots of one code and n_filler, fills their inputs (or
 * passes srcs[i], device-accessible, as submit_from does; srcs or an entry
 * may be NULL), then submits them together, in code-block order, with the
 * transport block's payload size and a device-accessible destination of
 * tb_size bytes.  Each block, once decoded, checks its CRC-24B (num_cb > 1)
 * and folds its share of the CRC-24A into the transport block's state
 * straight from its posterior in LDS, and writes its payload bits (one per
 * byte) to tb_out; the last block decides the transport block.  tb_out
 * holds the payload whatever the outcome: it is the transport block only
 * when crc_ok (the reference leaves its output untouched otherwise).
 * Block slots of a submitted transport block are not delivered by
 * try_recv, and their hard decisions are not copied to host memory;
 * try_recv_tb reports the transport block once it is decided and
 * all its blocks are complete, and release_tb frees its slots.  K' = K -
 * n_filler; each block carries P = K' - 24 payload bits (K' when alone);
 * num_cb * P >= tb_size + 24 is required.  SYNDROME_BUSY when 4096
 * transport blocks are in flight. */
typedef struct {
    uint32_t tb;              /* handle for syndrome_server_release_tb */
    uint64_t tag;
    int32_t crc_ok;           /* every CRC-24B and the CRC-24A passed */
    int32_t max_iterations;   /* over its blocks */
    size_t tb_size, num_cb;
    double submit_ns, done_ns;
} SyndromeServedTb;

int32_t syndrome_server_submit_tb(SyndromeServer *server, const uint32_t *slots, const float *const *srcs,
                                  size_t num_cb, size_t tb_size, uint8_t *tb_out, uint64_t tag);
/* A decided transport block, or SYNDROME_BUSY when none is ready.  A
 * transport block some of whose blocks the server could not decode (a
 * failed launch) is delivered with crc_ok 0 and max_iterations -1, and the
 * call returns SYNDROME_DEVICE_ERROR; release it as any other. */
int32_t syndrome_server_try_recv_tb(SyndromeServer *server, SyndromeServedTb *out);
/* DecodeReport::cb_crc (transport_block.rs:140-149) of a received, not yet
 * released transport block: cb_ok[i] = 1 when code block i passed its
 * CRC-24B (or is the transport block's only block), 0 when it failed or was
 * never decoded; *out_n = num_cb (BUFFER_TOO_SMALL when cap < num_cb). */
int32_t syndrome_server_tb_block_crc(SyndromeServer *server, uint32_t tb, uint8_t *cb_ok, size_t cap,
                                     size_t *out_n);
int32_t syndrome_server_release_tb(SyndromeServer *server, uint32_t tb);
/* Launches and blocks decoded so far, and the slot count.  Returns the
 * server's state: OK, or the sticky DEVICE_ERROR of a failed launch, after
 * which every submit is refused with that error and every block or
 * transport block not yet decoded is delivered as failed (iterations -1). */
int32_t syndrome_server_stats(const SyndromeServer *server, int64_t *out_batches, int64_t *out_decoded,
                              int32_t *out_slots);
/* Pinned, device-mapped host memory (usable as a submit_from source). */
void *syndrome_host_alloc(size_t bytes);
void syndrome_host_free(void *p);
/* Load generator: transport blocks drawn cyclically from a pool (TB t =
 * code blocks tb_first_cb[t] .. + tb_num_cb[t]; block b's decoder input at
 * pool + cb_off[b], pinned from syndrome_host_alloc) arrive as a Poisson
 * process of tb_rate TBs per second (tb_rate <= 0: a saturating stream)
 * for duration_s; every block is acquired, submitted from the pool and
 * received.  Latency = completion - arrival (admission waits included):
 * per code block into cb_lat_ns (first cb_cap), per TB (its last block)
 * into tb_lat_ns.  *info_bps = decoded info bits (cb_info_bits) per second
 * of wall time; *waits = acquire attempts refused for lack of a slot. */
int32_t syndrome_server_loadgen(SyndromeServer *server, const float *pool, const int64_t *cb_off, const int32_t *cb_bg,
                                const int32_t *cb_z, const int32_t *cb_nfill, const int32_t *cb_info_bits,
                                const int32_t *tb_first_cb, const int32_t *tb_num_cb, size_t n_tb, double tb_rate,
                                double duration_s, uint64_t seed, double *cb_lat_ns, size_t cb_cap, size_t *n_cb,
                                double *tb_lat_ns, size_t tb_cap, size_t *n_tbs, double *info_bps, int64_t *waits);
/* The same over transport blocks (resident server only): each arrival is
 * pool transport block p, its blocks acquired together and submitted with
 * syndrome_server_submit_tb (inputs from the pool, payload to tb_out +
 * tb_out_off[p], tb_size[p] bits); latency = transport-block completion,
 * CRCs decided on the device, minus scheduled arrival.  Reports completed
 * transport blocks and their code blocks, info bits/s (sum of tb_size),
 * goodput (tb_size of the transport blocks whose CRCs passed, per second),
 * CRC failures, and for each latency entry whether its transport block
 * passed (tb_lat_ok, may be NULL). */
int32_t syndrome_server_loadgen_tb(SyndromeServer *server, const float *pool, const int64_t *cb_off,
                                   const int32_t *cb_bg, const int32_t *cb_z, const int32_t *cb_nfill,
                                   const int32_t *tb_first_cb, const int32_t *tb_num_cb, const int32_t *tb_size,
                                   uint8_t *tb_out, const int64_t *tb_out_off, size_t n_tb, double tb_rate,
                                   double duration_s, uint64_t seed, double *tb_lat_ns, uint8_t *tb_lat_ok,
                                   size_t lat_cap, size_t *n_tbs, size_t *n_cbs, double *info_bps,
                                   double *goodput_bps, int64_t *crc_fails, int64_t *waits);

/* Synthetic code-block traffic, as the reference's benchmarks build it
 * (channel_sim.rs, QcLdpcEncoder, rate_match): for block i, K'_i = K_i -
 * n_filler[i] seeded info bits (written at info + info_off[i] when info is
 * not NULL), fillers zero, encoded for (bg[i], z[i]), rate matched to
 * e_len[i] bits (rv 0, Qm 1), BPSK over AWGN at ebno_db[i] for code rate
 * rate[i]; the e_len[i] LLRs are written at e_llr + e_off[i].  `threads`
 * host threads (0 = all). */
int32_t syndrome_traffic_blocks(const int32_t *bg, const int32_t *z, const int32_t *n_filler, const int32_t *e_len,
                                const float *ebno_db, const float *rate, uint64_t seed, size_t count, float *e_llr,
                                const int64_t *e_off, uint8_t *info, const int64_t *info_off, int32_t threads);
/* The same with the info bits given: block i's K'_i bits are read from
 * info + info_off[i] (one bit per byte), so transport-block payloads with
 * their CRCs and segmentation can be sent through the same chain. */
int32_t syndrome_traffic_blocks_info(const int32_t *bg, const int32_t *z, const int32_t *n_filler,
                                     const int32_t *e_len, const float *ebno_db, const float *rate, uint64_t seed,
                                     size_t count, const uint8_t *info, const int64_t *info_off, float *e_llr,
                                     const int64_t *e_off, int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
