"""ORACLE -- test infrastructure, not product code.

ctypes binding for ``liborc.so`` (the plain-C restatement of the reference's
LOMS decode path, ``loms_oracle.c``).  Only tests, the smoke check and the
benchmark's CPU-baseline leg import this module, as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libloms_ref.so")

MAX_ROWS, MAX_EDGES = 512, 4096  # loms_oracle.h ORC_MAX_ROWS / ORC_MAX_EDGES


class Code(C.Structure):
    """orc_code_t: QcLdpcParams (src/qc_ldpc.rs:233-253)."""

    _fields_ = [
        ("z", C.c_int), ("m_b", C.c_int), ("n_b", C.c_int), ("n_edges", C.c_int), ("max_deg", C.c_int),
        ("layer_off", C.c_int * (MAX_ROWS + 1)), ("cols", C.c_int * MAX_EDGES), ("shifts", C.c_int * MAX_EDGES),
    ]

    @property
    def n(self):
        return self.n_b * self.z

    @property
    def k(self):
        return (self.n_b - self.m_b) * self.z

    @property
    def edge_buf(self):
        return self.n_edges * self.z

    @property
    def layer_buf(self):
        return self.max_deg * self.z


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_code_3gpp.argtypes = [C.POINTER(Code), C.c_int, C.c_int]
        L.orc_code_wifi.argtypes = [C.POINTER(Code), C.c_int, C.c_int, C.c_int]
        L.orc_code_raw.argtypes = [C.POINTER(Code), P, P, P, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_decode_f32.argtypes = [C.POINTER(Code), C.c_float, P, C.c_size_t, P, C.c_size_t, P, C.c_size_t,
                                     C.c_size_t, P]
        L.orc_decode_f32.restype = C.c_long
        L.orc_init_5g_f32.argtypes = [C.POINTER(Code), P, C.c_size_t, C.c_size_t]
        L.orc_decode_i8.argtypes = [C.POINTER(Code), C.c_float, C.c_float, C.c_int16, P, C.c_size_t, P, C.c_size_t,
                                    P, C.c_size_t, C.c_size_t]
        L.orc_decode_i8.restype = C.c_long
        L.orc_init_5g_i8.argtypes = [C.POINTER(Code), P, C.c_size_t, C.c_size_t, C.c_int16]
        L.orc_beta_q.argtypes = [C.c_float, C.c_float]
        L.orc_beta_q.restype = C.c_int8
        L.orc_quantize_llr_i16.argtypes = [P, P, C.c_size_t, C.c_float]
        L.orc_quantize_llr_i8.argtypes = [P, P, C.c_size_t, C.c_float]
        L.orc_dequantize_llr_i8.argtypes = [P, P, C.c_size_t, C.c_float]
        L.orc_syndrome_ok_bits.argtypes = [C.POINTER(Code), P]
        L.orc_fnv1a_hard.argtypes = [P, C.c_size_t]
        L.orc_fnv1a_hard.restype = C.c_uint64
        L.orc_awgn_sigma.argtypes = [C.c_float, C.c_float]
        L.orc_awgn_sigma.restype = C.c_float
        L.orc_awgn_transmit.argtypes = [C.c_float, C.c_float, C.c_uint64, P, C.c_size_t, P]
        L.orc_encode.argtypes = [C.POINTER(Code), P, P]
        L.orc_knuth_info_bits.argtypes = [P, C.c_size_t]
        L.orc_rate_match.argtypes = [P, C.c_size_t, P, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int, C.c_size_t]
        L.orc_rate_dematch.argtypes = [P, C.c_size_t, P, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int, C.c_size_t]
        L.orc_crc.argtypes = [P, C.c_size_t, C.c_int, C.c_uint32]
        L.orc_crc.restype = C.c_uint32
        L.orc_segment.argtypes = [P, C.c_size_t, C.c_size_t, C.c_size_t, P]
        L.orc_tb_check.argtypes = [P, P, C.c_size_t, C.c_size_t, C.c_size_t, P, P]
        L.orc_rate_dematch_cached.argtypes = [P, C.c_size_t, P, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int,
                                              C.c_size_t]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def code_3gpp(bg, z):
    c = Code()
    rc = lib().orc_code_3gpp(C.byref(c), bg, z)
    if rc != 0:
        raise ValueError(f"invalid BG{bg + 1} Z={z}")
    return c


def code_wifi(z, rn, rd):
    c = Code()
    if lib().orc_code_wifi(C.byref(c), z, rn, rd) != 0:
        raise ValueError("invalid wifi code")
    return c


def code_raw(rows, cols, shifts, m_b, n_b, z):
    c = Code()
    r = np.ascontiguousarray(rows, np.int32)
    cc = np.ascontiguousarray(cols, np.int32)
    s = np.ascontiguousarray(shifts, np.int32)
    if lib().orc_code_raw(C.byref(c), _p(r), _p(cc), _p(s), len(r), m_b, n_b, z) != 0:
        raise ValueError("invalid raw code")
    return c


def decode_f32(code, beta, llr, iterations, trace=False):
    """Returns (iters, posterior, edge_r, hard[, trace])."""
    llr = np.array(llr, dtype=np.float32, copy=True)
    er = np.zeros(code.edge_buf, np.float32)
    hard = np.zeros(code.n, np.uint8)
    tr = np.zeros((max(iterations, 1), code.n), np.float32) if trace else None
    it = lib().orc_decode_f32(C.byref(code), beta, _p(llr), llr.size, _p(er), er.size, _p(hard), hard.size,
                              iterations, _p(tr) if trace else None)
    if it < 0:
        raise ValueError(f"oracle decode failed: {it}")
    return (it, llr, er, hard, tr[:it]) if trace else (it, llr, er, hard)


def init_5g_f32(code, llr, n_filler):
    llr = np.array(llr, dtype=np.float32, copy=True)
    assert lib().orc_init_5g_f32(C.byref(code), _p(llr), llr.size, n_filler) == 0
    return llr


def decode_i8(code, beta, app, iterations, scale=8.0, app_clamp=32767):
    """Returns (iters, app, edge_r, hard)."""
    app = np.array(app, dtype=np.int16, copy=True)
    er = np.zeros(code.edge_buf, np.int8)
    hard = np.zeros(code.n, np.uint8)
    it = lib().orc_decode_i8(C.byref(code), beta, scale, app_clamp, _p(app), app.size, _p(er), er.size,
                             _p(hard), hard.size, iterations)
    if it < 0:
        raise ValueError(f"oracle decode failed: {it}")
    return it, app, er, hard


def init_5g_i8(code, app, n_filler, app_clamp=32767):
    app = np.array(app, dtype=np.int16, copy=True)
    assert lib().orc_init_5g_i8(C.byref(code), _p(app), app.size, n_filler, app_clamp) == 0
    return app


def beta_q(beta, scale=8.0):
    return int(lib().orc_beta_q(beta, scale))


def quantize_i16(llr, scale=8.0):
    llr = np.ascontiguousarray(llr, np.float32)
    out = np.zeros(llr.size, np.int16)
    lib().orc_quantize_llr_i16(_p(llr), _p(out), llr.size, scale)
    return out


def quantize_i8(llr, scale=8.0):
    llr = np.ascontiguousarray(llr, np.float32)
    out = np.zeros(llr.size, np.int8)
    lib().orc_quantize_llr_i8(_p(llr), _p(out), llr.size, scale)
    return out


def dequantize_i8(q, scale=8.0):
    q = np.ascontiguousarray(q, np.int8)
    out = np.zeros(q.size, np.float32)
    lib().orc_dequantize_llr_i8(_p(q), _p(out), q.size, scale)
    return out


def syndrome_ok(code, bits):
    bits = np.ascontiguousarray(bits, np.uint8)
    return bool(lib().orc_syndrome_ok_bits(C.byref(code), _p(bits)))


def fnv1a(hard):
    hard = np.ascontiguousarray(hard, np.uint8)
    return int(lib().orc_fnv1a_hard(_p(hard), hard.size))


def awgn(ebno_db, rate, seed, bits):
    bits = np.ascontiguousarray(bits, np.uint8)
    out = np.zeros(bits.size, np.float32)
    lib().orc_awgn_transmit(ebno_db, rate, seed, _p(bits), bits.size, _p(out))
    return out


def sigma(ebno_db, rate):
    return float(lib().orc_awgn_sigma(ebno_db, rate))


def encode(code, info):
    info = np.ascontiguousarray(info, np.uint8)
    assert info.size == code.k
    cw = np.zeros(code.n, np.uint8)
    if lib().orc_encode(C.byref(code), _p(info), _p(cw)) != 0:
        raise RuntimeError("oracle encode failed")
    return cw


def knuth_info(k):
    out = np.zeros(k, np.uint8)
    lib().orc_knuth_info_bits(_p(out), k)
    return out


def xorshift_bits(seed, k):
    """Xorshift64 middle-bit info bits (tests/ldpc_int8_kernel_equivalence.rs:56-68):
    state = seed | 1; bit = (next >> 27) & 1."""
    s = (seed | 1) & 0xFFFFFFFFFFFFFFFF
    out = np.zeros(k, np.uint8)
    for i in range(k):
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        out[i] = (s >> 27) & 1
    return out


# -- the reference's own C++ port (only where it was built) --------------------

def ref_available():
    return os.path.exists(REF_LIB_PATH)


def ref_decode_bg1_z384_10it(json_text, llr):
    """Run the reference's bench/cpp/loms_decode.cpp decode (BG1 Z=384,
    beta=0.25, 10 iterations, no early exit) on `llr`; returns (llr, edge_r)."""
    L = C.CDLL(REF_LIB_PATH)
    L.ref_loms_bg1_z384_decode10.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p]
    llr = np.array(llr, dtype=np.float32, copy=True)
    er = np.zeros(316 * 384, np.float32)
    rc = L.ref_loms_bg1_z384_decode10(json_text.encode(), _p(llr), _p(er))
    assert rc == 0
    return llr, er


# -- CPU baseline: AVX2 restatement + worker pipeline (libcpubase.so) ---------

CPB_PATH = os.path.join(HERE, "libcpubase.so")
_cpb = None


def cpb():
    global _cpb
    if _cpb is None:
        if not os.path.exists(CPB_PATH):
            build()
        L = C.CDLL(CPB_PATH)
        L.cpb_decode_f32_avx2.argtypes = [C.POINTER(Code), C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]
        L.cpb_decode_f32_avx2.restype = C.c_long
        L.cpb_pipeline_run.argtypes = [C.POINTER(Code), C.c_float, C.c_void_p, C.c_long, C.c_long, C.c_int, C.c_int,
                                       C.c_long, C.c_void_p, C.c_void_p]
        L.cpb_pipeline_run.restype = C.c_longlong
        L.cpb_decode_i8_avx2.argtypes = [C.POINTER(Code), C.c_int8, C.c_int16, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_long]
        L.cpb_decode_i8_avx2.restype = C.c_long
        L.cpb_pipeline_mixed.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long,
                                         C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_int8, C.c_int16, C.c_long,
                                         C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.cpb_pipeline_mixed.restype = C.c_longlong
        _cpb = L
    return _cpb


def avx2_decode_f32(code, beta, llr, iterations):
    llr = np.array(llr, dtype=np.float32, copy=True)
    er = np.zeros(code.edge_buf, np.float32)
    hard = np.zeros(code.n, np.uint8)
    it = cpb().cpb_decode_f32_avx2(C.byref(code), beta, _p(llr), _p(er), _p(hard), iterations)
    return it, llr, er, hard


def pipeline_run(code, beta, llr_batch, iterations, workers, init5g=True, n_filler=0):
    """Returns (seconds, iters[count], hard[count, n])."""
    llr_batch = np.ascontiguousarray(llr_batch, np.float32)
    count = llr_batch.shape[0]
    iters = np.zeros(count, np.int32)
    hard = np.zeros((count, code.n), np.uint8)
    ns = cpb().cpb_pipeline_run(C.byref(code), beta, _p(llr_batch), count, iterations, workers, int(init5g),
                                n_filler, _p(iters), _p(hard))
    return ns * 1e-9, iters, hard


def avx2_decode_i8(code, beta_q_value, app, iterations, app_clamp=32767):
    """The AVX2 int8 restatement: returns (iters, app, edge_r, hard)."""
    app = np.array(app, dtype=np.int16, copy=True)
    er = np.zeros(code.edge_buf, np.int8)
    hard = np.zeros(code.n, np.uint8)
    it = cpb().cpb_decode_i8_avx2(C.byref(code), beta_q_value, app_clamp, _p(app), _p(er), _p(hard), iterations)
    if it < 0:
        raise ValueError("AVX2 int8 decode refused the code")
    return it, app, er, hard


def pipeline_mixed(codes, frame_code, in_off, llr_arena, iterations, workers, order=None, int8=False, beta=0.5,
                   scale=8.0, app_clamp=32767, init5g=True, n_filler=None):
    """A stream through the CPU pipeline over mixed codes (cpb_pipeline_mixed).
    codes: list of Code; frame f has code codes[frame_code[f]] and its N_f
    LLRs at llr_arena[in_off[f]:]; order: the frames in stream order
    (default each once).  Returns (seconds, iters[frames], hard arena)."""
    arr = (Code * len(codes))(*codes)
    fc = np.ascontiguousarray(frame_code, np.int32)
    off = np.ascontiguousarray(in_off, np.int64)
    nf = None if n_filler is None else np.ascontiguousarray(n_filler, np.int64)
    order = np.arange(fc.size, dtype=np.int32) if order is None else np.ascontiguousarray(order, np.int32)
    x = np.ascontiguousarray(llr_arena, np.float32)
    iters = np.zeros(fc.size, np.int32)
    hard = np.zeros(x.size, np.uint8)
    ns = cpb().cpb_pipeline_mixed(arr, _p(fc), _p(off), None if nf is None else _p(nf), _p(order), order.size,
                                  _p(x), int(bool(int8)), beta, scale, beta_q(beta, scale) if int8 else 0,
                                  app_clamp, iterations, workers, int(bool(init5g)), _p(iters), _p(hard))
    return ns * 1e-9, iters, hard


def rate_match(codeword, e_len, rv, qm, bg, z, n_filler):
    cw = np.ascontiguousarray(codeword, np.uint8)
    out = np.zeros(e_len, np.uint8)
    assert lib().orc_rate_match(_p(cw), cw.size, _p(out), e_len, rv, qm, bg, z, n_filler) == 0
    return out


def rate_dematch(e_llr, rv, qm, bg, z, n_filler, cb=None, cached=False):
    """rate_dematch_llr (cached=False) or RateMatchCache::rate_dematch_llr_into
    (cached=True, the HARQ path); accumulates into cb (default zeros)."""
    e = np.ascontiguousarray(e_llr, np.float32)
    ncb = (66 if bg == 0 else 50) * z
    cb = np.zeros(ncb, np.float32) if cb is None else np.array(cb, np.float32, copy=True)
    fn = lib().orc_rate_dematch_cached if cached else lib().orc_rate_dematch
    assert fn(_p(e), e.size, _p(cb), cb.size, rv, qm, bg, z, n_filler) == 0
    return cb


def decoder_input_from_cb(cb, z):
    """prepare_cb_llr placement (transport_block.rs:473-496): [2Z zeros | cb]."""
    return np.concatenate([np.zeros(2 * z, np.float32), np.asarray(cb, np.float32)])


CRC24A, CRC24B = 0x864CFB, 0x800063


def crc(bits, length=24, poly=CRC24A):
    """Crc24::compute (crc.rs:258-300) for a CRC of `length` bits."""
    b = np.ascontiguousarray(bits, np.uint8)
    return int(lib().orc_crc(_p(b), b.size, length, poly))


def attach_crc(bits, poly=CRC24A):
    """Crc24::attach (crc.rs:323-330): bits followed by their 24-bit CRC, MSB first."""
    r = crc(bits, 24, poly)
    return np.concatenate([np.asarray(bits, np.uint8), np.array([(r >> (23 - i)) & 1 for i in range(24)], np.uint8)])


def segment(tb_with_crc, c, k_prime):
    """segment (segmentation.rs:361-410): c blocks of K' bits, [c, K']."""
    t = np.ascontiguousarray(tb_with_crc, np.uint8)
    out = np.zeros((c, k_prime), np.uint8)
    assert lib().orc_segment(_p(t), t.size, c, k_prime, _p(out)) == 0
    return out


def tb_check(hard, cb_off, c, k_prime, tb_size):
    """DlSchDecoder::decode's CRC and reassembly tail: (ok, cb_ok[c], tb bits or None)."""
    h = np.ascontiguousarray(hard, np.uint8)
    offs = np.ascontiguousarray(cb_off, np.int64)
    cb_ok = np.zeros(c, np.uint8)
    out = np.zeros(tb_size, np.uint8)
    r = lib().orc_tb_check(_p(h), _p(offs), c, k_prime, tb_size, _p(cb_ok), _p(out))
    assert r >= 0
    return bool(r), cb_ok, (out if r else None)
