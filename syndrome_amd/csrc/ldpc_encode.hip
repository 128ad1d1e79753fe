// Systematic QC-LDPC encoder on device, for generating decoder traffic.
//
// Same codewords as QcLdpcEncoder::encode (src/qc_ldpc.rs:2592-2614): the
// info bytes are copied through, parity solves H c = 0.  The reference
// solves the 3GPP core with the double-diagonal recurrence of encode_sparse
// (qc_ldpc.rs:2620-2727) and other codes with a dense generator; here every
// code takes one route:
//   1. core syndrome: for core block row r, bit zi is the XOR of bit 0 of
//      info[col*Z + (zi + shift) mod Z] over the row's info edges;
//   2. core parity: block i of the first core*Z parity bits is the XOR of
//      the core syndrome blocks rotated by the circulant terms of the core
//      inverse (block-circulant, built once on the host);
//   3. every later block row closes on the identity block in its own parity
//      column, so its parity is the XOR of its other edges, all of them
//      reading info or core parity: the rows are independent.
//
// Bit layout.  Every Z-bit block that is read under rotation is kept in LDS
// packed 32 bits a word in "wrapped" form: word q, bit i holds
// v[(32q + i) mod Z] for L = 2*ceil(Z/32) + 1 words, so bits zi .. zi+31 of
// the block rotated by s are one funnel shift of two adjacent words, read
// at bit s + zi.  A lane therefore XORs 32 check bits per edge with one
// two-dword LDS read.
//
// Two kernels.  When Z is a multiple of 32 the wrapped form is the packed
// block repeated, so encode_stream packs each 32-byte run of info straight
// from its global load (copying it through to the codeword on the way) and
// expands each parity word straight into a 32-byte global store: nothing
// but the wrapped forms lives in LDS.  Other Z (encode_staged) assemble the
// codeword in LDS one byte per bit and build the wrapped forms a wave at a
// time (lane l takes bit 64c + l of the form, the wave ballot packs 64).
//
// HBM traffic is the compulsory K bytes read and N bytes written per block.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "ldpc_common.h"

namespace syn {

namespace {

constexpr int ENC_THREADS = 128;
constexpr int ENC_SMALL_N = 8192;  // codeword bits up to which a workgroup is one wave

// Sizes of one code, shared by every phase.
struct Shape {
    int z, nw, len, kb, mb, core, k, n, n_edges, n_terms;
};

__device__ __forceinline__ Shape shape_of(const SynCode *code, const SynEncCode *enc)
{
    Shape s;
    s.z = code->z;
    s.mb = code->m_b;
    s.kb = code->n_b - s.mb;
    s.core = enc->core;
    s.nw = (s.z + 31) >> 5;
    s.len = 2 * s.nw + 1;
    s.k = s.kb * s.z;
    s.n = code->n_b * s.z;
    s.n_edges = code->n_edges;
    s.n_terms = enc->n_terms;
    return s;
}

// LDS carve-up, mirrored by encode_lds_bytes: edge and term tables, row and
// term offsets, wrapped forms of the info and core parity blocks, of the
// core syndrome blocks, the core syndrome packed, then (staged kernel) the
// codeword bytes.
struct Lds {
    short *e_col, *e_sh, *t_blk, *t_sh;
    int *row_off, *term_off;
    uint32_t *wr, *wsyn, *syn;
    uint8_t *bytes;
};

__host__ __device__ inline size_t lds_words_end(int n_edges, int n_terms, int mb, int core, int kb, int len, int nw)
{
    return (size_t)2 * (n_edges + n_terms) * sizeof(short) + (size_t)(mb + 1 + core + 1) * sizeof(int) +
           ((size_t)(kb + 2 * core) * len + (size_t)core * nw) * sizeof(uint32_t);
}

template <int T>
__device__ __forceinline__ Lds carve(uint8_t *base, const Shape &s, const SynCode *code, const SynEncCode *enc)
{
    Lds l;
    l.e_col = reinterpret_cast<short *>(base);
    l.e_sh = l.e_col + s.n_edges;
    l.t_blk = l.e_sh + s.n_edges;
    l.t_sh = l.t_blk + s.n_terms;
    l.row_off = reinterpret_cast<int *>(l.t_sh + s.n_terms);  // 2*(n_edges + n_terms) shorts: 4-aligned
    l.term_off = l.row_off + s.mb + 1;
    l.wr = reinterpret_cast<uint32_t *>(l.term_off + s.core + 1);  // kb + core blocks
    l.wsyn = l.wr + (s.kb + s.core) * s.len;                       // core blocks
    l.syn = l.wsyn + s.core * s.len;                               // core * nw words
    l.bytes = base + ((lds_words_end(s.n_edges, s.n_terms, s.mb, s.core, s.kb, s.len, s.nw) + 15) & ~(size_t)15);
    for (int i = (int)threadIdx.x; i < s.n_edges; i += T) {
        l.e_col[i] = (short)code->cols[i];
        l.e_sh[i] = (short)code->shifts[i];
    }
    for (int i = (int)threadIdx.x; i < s.n_terms; i += T) {
        l.t_blk[i] = enc->term_blk[i];
        l.t_sh[i] = enc->term_shift[i];
    }
    for (int i = (int)threadIdx.x; i <= s.mb; i += T) l.row_off[i] = code->layer_off[i];
    for (int i = (int)threadIdx.x; i <= s.core; i += T) l.term_off[i] = enc->term_off[i];
    return l;
}

// Bits zi .. zi+31 (zi = p - shift) of a wrapped block read at bit p.
__device__ __forceinline__ uint32_t read32(const uint32_t *w, int p)
{
    const int q = p >> 5;
    return __builtin_amdgcn_alignbit(w[q + 1], w[q], (uint32_t)(p & 31));
}

// Bit 0 of each of the four bytes of x, as four bits.
__device__ __forceinline__ uint32_t pack4(uint32_t x) { return (((x & 0x01010101u) * 0x01020408u) >> 24) & 0xFu; }

// Four bits as four bytes 0/1.
__device__ __forceinline__ uint32_t expand4(uint32_t b) { return ((b & 0xFu) * 0x00204081u) & 0x01010101u; }

// Core syndrome word j of block row r: XOR of the row's info edges.
__device__ __forceinline__ uint32_t syndrome_word(const Lds &l, const Shape &s, int r, int j)
{
    uint32_t acc = 0;
    for (int e = l.row_off[r]; e < l.row_off[r + 1]; ++e) {
        const int col = l.e_col[e];
        if (col < s.kb) acc ^= read32(l.wr + col * s.len, l.e_sh[e] + 32 * j);
    }
    return acc;
}

// Core parity word j of block i: the core inverse's circulant terms.
__device__ __forceinline__ uint32_t core_word(const Lds &l, const Shape &s, int i, int j)
{
    uint32_t acc = 0;
    for (int t = l.term_off[i]; t < l.term_off[i + 1]; ++t) acc ^= read32(l.wsyn + l.t_blk[t] * s.len, l.t_sh[t] + 32 * j);
    return acc;
}

// Parity word j of extension row r: every edge but its own identity block.
__device__ __forceinline__ uint32_t ext_word(const Lds &l, const Shape &s, int r, int j)
{
    uint32_t acc = 0;
    for (int e = l.row_off[r]; e < l.row_off[r + 1]; ++e) {
        const int col = l.e_col[e];
        if (col != s.kb + r) acc ^= read32(l.wr + col * s.len, l.e_sh[e] + 32 * j);
    }
    return acc;
}

// Z a multiple of 32: word j of a block goes to wrapped words j and j + nw
// (and 2*nw for j = 0).
__device__ __forceinline__ void put_wrapped(uint32_t *w, const Shape &s, int j, uint32_t v)
{
    w[j] = v;
    w[j + s.nw] = v;
    if (j == 0) w[2 * s.nw] = v;
}

// 32 output bits as 32 bytes: two 16-byte global stores.
__device__ __forceinline__ void store_bits32(uint8_t *dst, uint32_t w)
{
    uint4 a, b;
    a.x = expand4(w);
    a.y = expand4(w >> 4);
    a.z = expand4(w >> 8);
    a.w = expand4(w >> 12);
    b.x = expand4(w >> 16);
    b.y = expand4(w >> 20);
    b.z = expand4(w >> 24);
    b.w = expand4(w >> 28);
    reinterpret_cast<uint4 *>(dst)[0] = a;
    reinterpret_cast<uint4 *>(dst)[1] = b;
}

// Z % 32 == 0, info and codeword buffers 16-byte aligned.  One workgroup
// per codeword (grid-stride); item it of a phase is 32-bit word it % nw of
// block row it / nw, so a wave's items are consecutive 32-byte runs of the
// codeword.
__global__ void __launch_bounds__(ENC_THREADS) encode_stream(const SynCode *__restrict__ code,
                                                             const SynEncCode *__restrict__ enc,
                                                             const uint8_t *__restrict__ info,
                                                             uint8_t *__restrict__ cw, long long count)
{
    extern __shared__ __align__(16) uint8_t lds[];
    const Shape s = shape_of(code, enc);
    const Lds l = carve<ENC_THREADS>(lds, s, code, enc);
    const int tid = (int)threadIdx.x;
    for (long long b = blockIdx.x; b < count; b += gridDim.x) {
        const uint8_t *src = info + b * s.k;
        uint8_t *dst = cw + b * s.n;
        // info: copy through, pack into the wrapped forms
        for (int it = tid; it < s.kb * s.nw; it += ENC_THREADS) {
            const uint4 a = reinterpret_cast<const uint4 *>(src)[2 * it];
            const uint4 c = reinterpret_cast<const uint4 *>(src)[2 * it + 1];
            reinterpret_cast<uint4 *>(dst)[2 * it] = a;
            reinterpret_cast<uint4 *>(dst)[2 * it + 1] = c;
            const uint32_t w = pack4(a.x) | pack4(a.y) << 4 | pack4(a.z) << 8 | pack4(a.w) << 12 | pack4(c.x) << 16 |
                               pack4(c.y) << 20 | pack4(c.z) << 24 | pack4(c.w) << 28;
            const int blk = it / s.nw, j = it - blk * s.nw;
            put_wrapped(l.wr + blk * s.len, s, j, w);
        }
        __syncthreads();
        for (int it = tid; it < s.core * s.nw; it += ENC_THREADS) {
            const int r = it / s.nw, j = it - r * s.nw;
            put_wrapped(l.wsyn + r * s.len, s, j, syndrome_word(l, s, r, j));
        }
        __syncthreads();
        for (int it = tid; it < s.core * s.nw; it += ENC_THREADS) {
            const int i = it / s.nw, j = it - i * s.nw;
            const uint32_t w = core_word(l, s, i, j);
            put_wrapped(l.wr + (s.kb + i) * s.len, s, j, w);
            store_bits32(dst + s.k + 32 * it, w);
        }
        __syncthreads();
        const int ext0 = s.core * s.nw;
        for (int it = tid; it < (s.mb - s.core) * s.nw; it += ENC_THREADS) {
            const int rr = it / s.nw, j = it - rr * s.nw;
            store_bits32(dst + s.k + 32 * (ext0 + it), ext_word(l, s, s.core + rr, j));
        }
        __syncthreads();
    }
}

// Bytes 0/1 of the low `cnt` bits of w at dst (4-byte LDS stores when Z % 4
// == 0 keeps them aligned).
__device__ __forceinline__ void put_bits(uint8_t *dst, uint32_t w, int cnt, bool words)
{
    if (words) {
        for (int q = 0; q < cnt; q += 4) *reinterpret_cast<uint32_t *>(dst + q) = expand4(w >> q);
    } else {
        for (int i = 0; i < cnt; ++i) dst[i] = (uint8_t)((w >> i) & 1u);
    }
}

// Wrapped forms of `nblk` blocks: dst + b*len, bit j of block b = src(b, j).
// A wave takes whole blocks; lane l holds j = (64c + l) mod Z for chunk c,
// advanced by 64 mod Z per chunk.  Four chunks' bits are fetched before
// their ballots.
template <int T, class Src>
__device__ __forceinline__ void wrap_blocks(uint32_t *dst, int nblk, const Shape &s, int j0, int step, Src src)
{
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    const int chunks = (s.len + 1) >> 1;  // 64-bit chunks per block
    for (int b = wave; b < nblk; b += T / 64) {
        uint32_t *d = dst + b * s.len;
        int j = j0;
        for (int c0 = 0; c0 < chunks; c0 += 4) {
            uint32_t bit[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                bit[u] = src(b, j);
                j += step;
                if (j >= s.z) j -= s.z;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const unsigned long long m = __ballot(bit[u] != 0u);
                const int q = 2 * (c0 + u) + (lane & 1);
                if (c0 + u < chunks && lane < 2 && q < s.len) d[q] = lane ? (uint32_t)(m >> 32) : (uint32_t)m;
            }
        }
    }
}

// Any Z.  One workgroup per codeword (grid-stride); the codeword is
// assembled in LDS bytes and written out at the end.
template <int T>
__global__ void __launch_bounds__(T) encode_staged(const SynCode *__restrict__ code,
                                                             const SynEncCode *__restrict__ enc,
                                                             const uint8_t *__restrict__ info,
                                                             uint8_t *__restrict__ cw, long long count, int vec)
{
    extern __shared__ __align__(16) uint8_t lds[];
    const Shape s = shape_of(code, enc);
    const Lds l = carve<T>(lds, s, code, enc);
    uint8_t *bytes = l.bytes;
    const bool words = (s.z & 3) == 0;
    const int tid = (int)threadIdx.x;
    const int j0 = ((int)threadIdx.x & 63) % s.z, step = 64 % s.z;
    for (long long b = blockIdx.x; b < count; b += gridDim.x) {
        const uint8_t *src = info + b * s.k;
        uint8_t *dst = cw + b * s.n;
        if (vec) {
            for (int i = tid; i < s.k / 16; i += T)
                reinterpret_cast<uint4 *>(bytes)[i] = reinterpret_cast<const uint4 *>(src)[i];
        } else {
            for (int i = tid; i < s.k; i += T) bytes[i] = src[i];
        }
        __syncthreads();
        wrap_blocks<T>(l.wr, s.kb, s, j0, step, [&](int blk, int j) { return bytes[blk * s.z + j] & 1u; });
        __syncthreads();
        for (int it = tid; it < s.core * s.nw; it += T) {
            const int r = it / s.nw, j = it - r * s.nw;
            l.syn[it] = syndrome_word(l, s, r, j);
        }
        __syncthreads();
        wrap_blocks<T>(l.wsyn, s.core, s, j0, step,
                    [&](int blk, int j) { return (l.syn[blk * s.nw + (j >> 5)] >> (j & 31)) & 1u; });
        __syncthreads();
        for (int it = tid; it < s.core * s.nw; it += T) {
            const int i = it / s.nw, j = it - i * s.nw;
            put_bits(bytes + (s.kb + i) * s.z + 32 * j, core_word(l, s, i, j), min(32, s.z - 32 * j), words);
        }
        __syncthreads();
        wrap_blocks<T>(l.wr + s.kb * s.len, s.core, s, j0, step,
                    [&](int blk, int j) { return (uint32_t)bytes[(s.kb + blk) * s.z + j]; });
        __syncthreads();
        for (int it = tid; it < (s.mb - s.core) * s.nw; it += T) {
            const int rr = it / s.nw, j = it - rr * s.nw, r = s.core + rr;
            put_bits(bytes + (s.kb + r) * s.z + 32 * j, ext_word(l, s, r, j), min(32, s.z - 32 * j), words);
        }
        __syncthreads();
        if (vec) {
            for (int i = tid; i < s.n / 16; i += T)
                reinterpret_cast<uint4 *>(dst)[i] = reinterpret_cast<const uint4 *>(bytes)[i];
        } else {
            for (int i = tid; i < s.n; i += T) dst[i] = bytes[i];
        }
        __syncthreads();
    }
}

bool stream_ok(int z, const void *info, const void *cw)
{
    return z % 32 == 0 && ((uintptr_t)info & 15) == 0 && ((uintptr_t)cw & 15) == 0;
}

}  // namespace

// LDS bytes of the encoder launch for one code and buffer pair.
size_t encode_lds_bytes(const SynCode &c, const SynEncCode &e, const void *info, const void *cw)
{
    const int nw = (c.z + 31) / 32, len = 2 * nw + 1, kb = c.n_b - c.m_b;
    const size_t words = lds_words_end(c.n_edges, e.n_terms, c.m_b, e.core, kb, len, nw);
    if (stream_ok(c.z, info, cw)) return words;
    return ((words + 15) & ~(size_t)15) + (size_t)c.n_b * c.z;
}

// Encodes `count` blocks of one code: info [count][K] -> cw [count][N]
// (device pointers).  `lds` from encode_lds_bytes for the same buffers.
int launch_encode(const SynCode *d_code, const SynEncCode *d_enc, int z, int k, int n, size_t lds,
                  const uint8_t *info, uint8_t *cw, long long count, hipStream_t stream)
{
    if (count <= 0) return 0;
    const unsigned grid = (unsigned)(count < 65536 ? count : 65536);
    if (stream_ok(z, info, cw)) {
        hipLaunchKernelGGL(encode_stream, dim3(grid), dim3(ENC_THREADS), lds, stream, d_code, d_enc, info, cw, count);
    } else {
        const bool vec = k % 16 == 0 && n % 16 == 0 && ((uintptr_t)info & 15) == 0 && ((uintptr_t)cw & 15) == 0;
        // a small codeword gets a one-wave workgroup: its phases are too short
        // to keep four waves busy between barriers
        if (n <= ENC_SMALL_N)
            hipLaunchKernelGGL(encode_staged<64>, dim3(grid), dim3(64), lds, stream, d_code, d_enc, info, cw, count,
                               vec ? 1 : 0);
        else
            hipLaunchKernelGGL(encode_staged<ENC_THREADS>, dim3(grid), dim3(ENC_THREADS), lds, stream, d_code, d_enc,
                               info, cw, count, vec ? 1 : 0);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace syn
