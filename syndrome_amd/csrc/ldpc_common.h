// Shared host/device definitions for the MI355X LOMS decoder.
//
// Layout conventions (one codeword):
//   posterior / channel LLR : f32 [N] or i16 [N], N = n_b * Z, index col*Z + z
//   check-to-variable R     : f32 / i8 [E*Z], index (layer_begin + edge)*Z + z
//                             (the reference's edge_r layout, src/qc_ldpc.rs:1070-1076)
//   hard decisions          : u8 [N], 0/1
// A batch stacks codewords contiguously: codeword b starts at b*N (b*E*Z for R).
#pragma once
#include <stddef.h>
#include <stdint.h>

#define SYN_MAX_ROWS 46
#define SYN_MAX_COLS 68
#define SYN_MAX_EDGES 316
#define SYN_MAX_DEG 19
#define SYN_Z_MAX 384

// Status codes: capi/src/lib.rs:74-94.
#define SYN_OK 0
#define SYN_INVALID_PARAM (-1)
#define SYN_CRC_MISMATCH (-2)
#define SYN_NOT_CONVERGED (-3)
#define SYN_BUFFER_TOO_SMALL (-4)
#define SYN_NULL_POINTER (-5)
#define SYN_PANIC (-6)
#define SYN_DEVICE_ERROR (-7)
#define SYN_BUSY (-8)

// Base-graph family a code was built from: selects the register-resident
// kernel specialisation (kind 0/1) or the runtime-structured one (raw).
#define SYN_KIND_BG1 0
#define SYN_KIND_BG2 1
#define SYN_KIND_RAW 2
#define SYN_KIND_WIDE 3    // a raw code outside the inline tables below (SynWideCode)
#define SYN_KIND_WIFI0 16  // + matrix index: the 12 IEEE 802.11 matrices (src/wifi_ldpc_tables.rs)

// Runtime description of one lifted code, QcLdpcParams (src/qc_ldpc.rs:233-253).
// Lives in device memory; every field is read with wave-uniform (scalar) loads.
struct SynCode {
    int kind;
    int z;
    int m_b;
    int n_b;
    int n_edges;
    int max_deg;
    int layer_off[SYN_MAX_ROWS + 1];
    int cols[SYN_MAX_EDGES];
    int shifts[SYN_MAX_EDGES];  // reduced mod z
};

// A raw-edge code of any size (QcLdpcParams::from_raw_edges,
// src/qc_ldpc.rs:412-455, whose domain is u8 row and column indices, any
// number of edges and repeated (row, col) pairs): the tables live in device
// memory of their own.  Only the layers holding edges are listed (an empty
// layer changes nothing and every one of its checks holds).  dup[l] marks a
// layer in which one column block appears more than once, so that two check
// rows of the layer share a variable; such a layer is decoded edge by edge
// (src/qc_ldpc.rs:1181-1204 applies the posterior deltas in edge order).
// m_b * z, n_b * z and n_edges * z are below 2^31.
struct SynWideCode {
    int z;
    int m_b;          // num_row_blocks (layers with no edges included)
    int n_b;          // num_col_blocks
    int n_layers;     // layers holding edges
    int n_edges;
    int max_deg;
    int any_dup;
    int pad;
    const int *layer_off;      // [n_layers + 1] into cols / shifts
    const int *cols;           // [n_edges]
    const int *shifts;         // [n_edges], reduced mod z
    const unsigned char *dup;  // [n_layers]
};

// Launch geometry of a wide decode: one codeword per workgroup of `block`
// lanes; per codeword `work_stride` bytes of HBM workspace, holding the
// posterior at 0 unless it sits in LDS (`lds` bytes), and from `q_off` the Q
// messages of a dup layer.
struct WideShape {
    int block = 0;
    bool p_in_lds = false;
    size_t lds = 0, q_off = 0, work_stride = 0;
};

// Decode-call flags.
#define SYN_F_INIT_5G 1u        // apply init_5g_llr / decode_5g_i8 prologue on load
#define SYN_F_WRITE_POST 2u     // write the posterior back
#define SYN_F_WRITE_HARD 4u     // write hard decisions
#define SYN_F_WRITE_EDGE 8u     // write final R messages (edge_r)
#define SYN_F_QUANTIZE_IN 16u   // i8 path: input is f32 LLR, quantize_llr_i16 on load
#define SYN_F_PUBLIC 31u        // the flags a caller may pass; the rest are set by the launchers
#define SYN_F_ALIGNED 32u       // element offsets and N are multiples of 4, in/post/hard 16-byte aligned
#define SYN_F_TRACE 64u         // f32 path: posterior after every pass to `trace`
#define SYN_F_XPAIR 128u        // resident server: BG1 Z=384 items run the two-codeword kernel

// Per-launch arguments shared by every decode kernel.
struct SynLaunch {
    const void *in;        // f32 [B][N] (f32 path or SYN_F_QUANTIZE_IN) or i16 [B][N]
    void *post;            // f32 / i16 [B][N]   (may equal `in`)
    void *edge;            // f32 / i8 [B][E*Z]  (plain kernels: required working buffer)
    uint8_t *hard;         // u8 [B][N]
    int *iters;            // i32 [B]
    int batch;
    int iterations;
    int cws;               // codewords per workgroup
    unsigned flags;
    long long n_filler;
    float beta;            // f32 offset
    float scale;           // i8 path: quantizer scale
    int beta_q;            // i8 path: quantized offset
    int app_clamp;         // i8 path: max(app_clamp, 0)
    float *trace;          // SYN_F_TRACE: f32 [B][iterations][N], row it = posterior after pass it+1
};

// Most codewords one work item of a work-queue launch holds (the size of
// the kernel's per-item tables in LDS).
#define SYN_ITEM_MAX_CWS 256

// One work item of a work-queue launch: `count` codewords (ids
// ids[first .. first+count)) of code `code` (an index into the launch's code
// table), whose base graph is `kind`.
struct SynItem {
    int code;
    int kind;  // the code's base-graph kind, | SYN_ITEM_PACKED for an int8 BG2 item decoded pair-packed
    int first;
    int count;
};
#define SYN_ITEM_PACKED 0x100

// Resident serving decoder: one queued item = up to SYN_WORK_MAX code
// blocks of one code, named by their server slots.  Items live in a ring in
// host memory; `tail` (items published) and `stop` in a control block there.
#define SYN_WORK_MAX 32
struct SynWork {
    int code;     // index into the server's code table
    int count;    // code blocks in the item
    int aligned;  // every input pointer 16-byte aligned
    int pad;
    int slot[SYN_WORK_MAX];          // output slot of each block
    int nfill[SYN_WORK_MAX];         // its filler count
    const void *src[SYN_WORK_MAX];   // its decoder input (HBM or mapped host memory)
    int tb[SYN_WORK_MAX];            // its transport block's ring index, or -1
    int tb_ci[SYN_WORK_MAX];         // its index within that transport block
};

// A transport block of the resident server (ring in host memory): num_cb
// blocks of P payload bits each carry tb_size payload bits then CRC-24A;
// each block carries its CRC-24B after its P bits when num_cb > 1.
struct SynTbWork {
    unsigned id;     // submission number; its done word becomes id + 1
    int num_cb;
    int p;
    int tb_size;
    unsigned char *out;  // payload destination, one bit per byte (device-accessible)
};

// Device-side fold of one transport block's checks (device memory).
struct SynTbState {
    unsigned acc;    // XOR of the blocks' contributions to the CRC-24A remainder
    unsigned rx;     // the received CRC-24A bits, OR-ed in by the blocks holding them
    unsigned bad;    // some block failed its CRC-24B
    unsigned count;  // blocks folded so far
};

struct SynResidentTb {
    const SynTbWork *work;  // ring (host memory)
    SynTbState *state;      // per ring index (device memory)
    unsigned *done;         // per ring index: id + 1 once decided (host memory)
    unsigned *ok;           // per ring index: 1 if every check passed (host memory)
    unsigned *cbok;         // per server slot: 1 if its CRC-24B passed or it is alone (host memory)
};

// Bytes of the resident decoder's device-side queue block (two 128-byte
// cache lines).
#define SYN_RESIDENT_QUEUE_BYTES 256

struct SynResidentCtl {
    unsigned int tail;  // items published; written by the host after the item
    unsigned int stop;  // non-zero: every workgroup leaves at its next wait
    unsigned int pad[14];
};

// One transport block for the device TB checks: code blocks
// [first_cb, first_cb + num_cb) of a plan, each with K' = k_prime bits
// (its CRC-24B included when num_cb > 1), carrying tb_size payload bits
// followed by the 24-bit CRC-24A; the payload goes to tb_out + out_off.
struct SynTb {
    long long out_off;
    int first_cb;
    int num_cb;
    int k_prime;
    int tb_size;
};

// One code block for the device rate de-matcher.
struct SynDematchCb {
    long long e_off;    // offset of its E received LLRs
    long long out_off;  // offset of its N-element decoder input
    int e;              // E
    int qm;             // modulation order (bit interleaver columns)
    int z;
    int bg;             // 0 = BG1, 1 = BG2
    int rv;             // redundancy version 0..3
    int n_filler;
    // derived once on the host (dematch_derive), read by every workgroup
    int n;              // decoder input length: the circular buffer + 2Z
    int fs, fe;         // filler range in the circular buffer
    int m;              // selectable (non-filler) positions per cycle
    int rows;           // e / qm
    int kmax;           // selections that follow the walk: min(e, 2m + 1)
    int bncb;           // non-filler positions in the whole buffer
    int qsh;            // log2(qm) when qm is a power of two, else -1
    int k0[4];          // walk start per redundancy version
    int bk0[4];         // non-filler positions before k0[rv]
};

// Parity solve of the device encoder for one code: block rows [0, core)
// are solved jointly through the core inverse, given as circulant terms
// (block row i: XOR over terms [term_off[i], term_off[i+1]) of the core
// syndrome block term_blk rotated by term_shift); rows [core, m_b) each
// close on the identity block in their own parity column.
#define SYN_ENC_MAX_TERMS 1024
struct SynEncCode {
    int core;
    int n_terms;
    int term_off[SYN_MAX_ROWS + 1];
    short term_blk[SYN_ENC_MAX_TERMS];
    short term_shift[SYN_ENC_MAX_TERMS];
};

// One unit of device rate de-matching: output elements
// [p0, p0 + SYN_DEMATCH_CHUNK) of code block `cb` (clipped to its N).
#define SYN_DEMATCH_CHUNK 4096
struct SynDematchChunk {
    int cb;
    int p0;
};
