/*
 * A C program against include/syndrome_ldpc.h and libsyndrome_amd.so,
 * compiled and linked by a C compiler as the reference's CI does for its
 * C ABI (.github/workflows/ci.yml:198-243).
 *
 *  1. Encoder round trip, as that CI check: BG1 Z=4, info bit i = (i % 3 == 0),
 *     encode_5g, LLR -5 / +5 from the codeword, decode_5g, hard == info.
 *  2. The golden vector: BG1 Z=384, llr[i] = +0.5 for even i and -0.5 for
 *     odd, offset 0.25, 10 iterations (the decode never converges), FNV-1a
 *     over the hard bits (offset 0xcbf29ce484222325, prime 0x100000001b3, one
 *     byte per bit) == 0xa2448e8cf5272825 (bench/results/ldpc_rust.checksum,
 *     ldpc_bench_export.rs:239-250).
 *
 * Prints one line per check and exits 0 when both hold.  ABI_SYMBOL_TABLE,
 * when defined, names a generated file that takes every declared entry
 * point's address, so that linking proves each one is exported.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "syndrome_ldpc.h"

#ifdef ABI_SYMBOL_TABLE
#include ABI_SYMBOL_TABLE
#endif

static int round_trip(void)
{
    SyndromeLdpcEncoder *enc = NULL;
    SyndromeLdpcDecoder *dec = NULL;
    if (syndrome_ldpc_encoder_create(0, 4, &enc) != SYNDROME_OK || !enc) return 1;
    if (syndrome_ldpc_decoder_create(0, 4, 0.5f, &dec) != SYNDROME_OK || !dec) return 1;
    size_t k = syndrome_ldpc_encoder_info_bit_count(enc), n = syndrome_ldpc_encoder_codeword_bit_count(enc);
    size_t el = syndrome_ldpc_decoder_required_edge_buffer(dec), ll = syndrome_ldpc_decoder_required_layer_buffer(dec);
    uint8_t *info = malloc(k), *cw = malloc(n), *hard = malloc(n);
    float *llr = malloc(n * sizeof(float)), *er = calloc(el, sizeof(float)), *ls = calloc(ll, sizeof(float));
    int bad = !info || !cw || !hard || !llr || !er || !ls;
    for (size_t i = 0; !bad && i < k; i++) info[i] = (uint8_t)(i % 3 == 0);
    if (!bad && syndrome_ldpc_encode_5g(enc, info, k, 0, cw, n) != SYNDROME_OK) bad = 1;
    for (size_t i = 0; !bad && i < n; i++) llr[i] = cw[i] ? -5.0f : 5.0f;
    size_t iters = 0;
    if (!bad && syndrome_ldpc_decode_5g(dec, llr, n, 0, er, el, ls, ll, hard, n, 10, &iters) != SYNDROME_OK) bad = 1;
    for (size_t i = 0; !bad && i < k; i++)
        if (hard[i] != info[i]) {
            fprintf(stderr, "mismatch at %zu\n", i);
            bad = 1;
        }
    if (!bad) printf("round trip ok: k=%zu n=%zu iterations=%zu\n", k, n, iters);
    free(info), free(cw), free(hard), free(llr), free(er), free(ls);
    syndrome_ldpc_encoder_destroy(enc);
    syndrome_ldpc_decoder_destroy(dec);
    return bad;
}

static int golden(void)
{
    SyndromeLdpcDecoder *dec = NULL;
    if (syndrome_ldpc_decoder_create(0, 384, 0.25f, &dec) != SYNDROME_OK || !dec) return 1;
    size_t n = syndrome_ldpc_decoder_variable_node_count(dec);
    size_t el = syndrome_ldpc_decoder_required_edge_buffer(dec), ll = syndrome_ldpc_decoder_required_layer_buffer(dec);
    float *llr = malloc(n * sizeof(float)), *er = calloc(el, sizeof(float)), *ls = calloc(ll, sizeof(float));
    uint8_t *hard = malloc(n);
    int bad = !llr || !er || !ls || !hard;
    for (size_t i = 0; !bad && i < n; i++) llr[i] = (i % 2 == 0) ? 0.5f : -0.5f;
    size_t iters = 0;
    /* the plain decode: decode_5g would zero the 2Z punctured positions */
    if (!bad && syndrome_ldpc_decode(dec, llr, n, er, el, ls, ll, hard, n, 10, &iters) != SYNDROME_OK) bad = 1;
    uint64_t h = 0xcbf29ce484222325ULL;
    for (size_t i = 0; !bad && i < n; i++) {
        h ^= (uint64_t)(hard[i] & 1u);
        h *= 0x100000001b3ULL;
    }
    if (!bad) {
        printf("golden checksum %016llx after %zu iterations\n", (unsigned long long)h, iters);
        bad = h != 0xa2448e8cf5272825ULL || iters != 10;
    }
    free(llr), free(er), free(ls), free(hard);
    syndrome_ldpc_decoder_destroy(dec);
    return bad;
}

int main(void)
{
    if (round_trip()) {
        fprintf(stderr, "round trip failed\n");
        return 1;
    }
    if (golden()) {
        fprintf(stderr, "golden checksum failed\n");
        return 1;
    }
    printf("C ABI ok\n");
    return 0;
}
