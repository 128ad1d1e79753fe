"""C ABI checks that need no GPU: the library loads, exports every symbol
include/syndrome_ldpc.h declares, and the host-side parts (argument
validation, encoder, channel model) behave as the reference's
(capi/src/lib.rs:494-621, qc_ldpc.rs:2547-2614, channel_sim.rs)."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import f32_bits


@pytest.fixture(scope="module")
def L():
    from syndrome_amd import _lib
    return _lib.lib()


def test_exports_every_declared_symbol(L):
    from syndrome_amd._lib import SIGNATURES, declared_symbols
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
        assert s in SIGNATURES, f"{s} has no ctypes signature"


def test_decoder_create_rejects_bad_params_before_touching_the_device(L):
    from syndrome_amd import _lib
    out = C.c_void_p(1234)
    assert L.syndrome_ldpc_decoder_create(2, 384, 0.5, C.byref(out)) == _lib.INVALID_PARAM
    assert out.value is None  # *out = NULL on failure (capi/src/lib.rs:160)
    for z in (0, 1, 383, 385, 2 ** 40):
        out = C.c_void_p(99)
        assert L.syndrome_ldpc_decoder_create(0, z, 0.5, C.byref(out)) == _lib.INVALID_PARAM
        assert out.value is None
    assert L.syndrome_ldpc_decoder_create(0, 384, 0.5, None) == _lib.NULL_POINTER
    assert L.syndrome_ldpc_decoder_create_wifi(27, 1, 3, 0.5, C.byref(out)) == _lib.INVALID_PARAM
    # raw edges (from_raw_edges, qc_ldpc.rs:418-440): z outside 1..=384, shift
    # >= z or < 0, row or column out of range are refused before any device
    # call; so are sizes whose element counts reach 2^31 (this library's
    # limit, the reference has none)
    def raw(rows, cols, sh, m_b, n_b, z):
        r = np.array(rows, np.uint8)
        c = np.array(cols, np.uint8)
        s = np.array(sh, np.int16)
        o = C.c_void_p(77)
        rc = L.syndrome_ldpc_decoder_create_raw(r.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p),
                                                s.ctypes.data_as(C.c_void_p), len(r), m_b, n_b, z, 0.5, C.byref(o))
        if rc == _lib.OK:
            L.syndrome_ldpc_decoder_destroy(o)
        else:
            assert o.value is None
        return rc

    for args in (([0], [0], [4], 2, 3, 4), ([2], [0], [0], 2, 3, 4), ([0], [3], [0], 2, 3, 4),
                 ([0], [0], [-1], 2, 3, 4), ([0], [0], [0], 2, 3, 0), ([0], [0], [0], 2, 3, 385),
                 ([0], [0], [0], 1, 2 ** 31 // 4, 4), ([0], [0], [0], 2 ** 40, 3, 4)):
        assert raw(*args) == _lib.INVALID_PARAM, args
    # the rest of the domain is accepted: a repeated (row, col), more rows,
    # columns and edges than any 3GPP graph, empty codes; without a GPU the
    # call then fails at the device, not at validation
    rng = np.random.default_rng(3)
    big = (rng.integers(0, 256, 700), rng.integers(0, 256, 700), rng.integers(0, 8, 700))
    for args in (([0, 0], [1, 1], [0, 1], 2, 3, 4), (*big, 300, 300, 8), ([], [], [], 0, 0, 5),
                 ([], [], [], 3, 0, 5), ([0], [0], [0], 1, 1, 384)):
        assert raw(*args) != _lib.INVALID_PARAM, args


def test_null_decoder_is_reported_not_dereferenced(L):
    from syndrome_amd import _lib
    buf = np.zeros(16, np.float32)
    p = buf.ctypes.data_as(C.c_void_p)
    assert L.syndrome_ldpc_decode_5g(None, p, 16, 0, p, 16, p, 16, p, 16, 5, None) == _lib.NULL_POINTER
    assert L.syndrome_ldpc_decode(None, p, 16, p, 16, p, 16, p, 16, 5, None) == _lib.NULL_POINTER
    assert L.syndrome_ldpc_encode_batch_device(None, p, p, 1, None) == _lib.NULL_POINTER
    n = C.c_size_t(0)
    assert L.syndrome_server_acquire_many(None, p, p, None, 1, None, p, C.byref(n)) == _lib.NULL_POINTER
    assert L.syndrome_server_submit_many(None, p, None, None, 1) == _lib.NULL_POINTER
    assert L.syndrome_server_try_recv_many(None, p, 1, C.byref(n)) == _lib.NULL_POINTER
    assert L.syndrome_server_release_many(None, p, 1) == _lib.NULL_POINTER
    L.syndrome_ldpc_decoder_destroy(None)  # no-op
    L.syndrome_ldpc_encoder_destroy(None)


def test_encoder_matches_oracle_and_reference_conventions(L, oracle):
    from syndrome_amd import QcLdpcEncoder, SyndromeError, _lib
    for bg, z in ((0, 384), (1, 208), (0, 2), (1, 3), (0, 240)):
        enc = QcLdpcEncoder(bg, z)
        c = oracle.code_3gpp(bg, z)
        assert (enc.info_bit_count(), enc.codeword_bit_count()) == (c.k, c.n)
        for nf in (0, 5):
            info = oracle.xorshift_bits(z + nf, c.k - nf)
            cw = np.zeros(c.n, np.uint8)
            enc.encode_5g(info, nf, cw)
            full = np.concatenate([info, np.zeros(nf, np.uint8)])
            assert np.array_equal(cw, oracle.encode(c, full))
    enc = QcLdpcEncoder(0, 4)
    cw = np.zeros(68 * 4, np.uint8)
    with pytest.raises(SyndromeError) as e:
        enc.encode_5g(np.zeros(88 + 1, np.uint8), 0, cw)  # wrong info length
    assert e.value.status == _lib.BUFFER_TOO_SMALL
    with pytest.raises(SyndromeError) as e:
        enc.encode_5g(np.zeros(0, np.uint8), 89, cw)  # n_filler > k
    assert e.value.status == _lib.INVALID_PARAM
    with pytest.raises(SyndromeError) as e:
        enc.encode_5g(np.zeros(88, np.uint8), 0, np.zeros(10, np.uint8))
    assert e.value.status == _lib.BUFFER_TOO_SMALL


def test_awgn_matches_oracle(oracle):
    from syndrome_amd import awgn_transmit
    bits = oracle.xorshift_bits(3, 5000)
    for ebno, rate, seed in ((2.0, 0.5, 1), (-1.0, 1 / 3, 7), (4.5, 8448 / 26112, 42)):
        assert np.array_equal(f32_bits(awgn_transmit(ebno, rate, seed, bits)), f32_bits(oracle.awgn(ebno, rate, seed, bits)))


def test_product_has_no_oracle_dependency():
    """The product package never imports, links or runs the oracle."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for d, _, files in os.walk(os.path.join(root, "syndrome_amd")):
        if "_build" in d:
            continue
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h", ".inc")):
                text = open(os.path.join(d, f)).read()
                assert not re.search(r"\boracle\b|liborc|loms_oracle", text), f


def test_traffic_generator_matches_oracle_pipeline(oracle):
    """syndrome_traffic_blocks = encode -> rate_match (rv 0, Qm 1) -> AWGN,
    each stage as the oracle restates the reference's."""
    from syndrome_amd import traffic_blocks
    cases = [(0, 384, 0, 25344, 1.5), (1, 208, 0, 10400, 0.7), (0, 52, 17, 3000, 2.0), (1, 13, 5, 900, 3.0),
             (1, 96, 40, 200, 1.0),
             (0, 8, 40, 1100, 1.0),   # past the selection walk's 2*Ncb guard
             (1, 4, 30, 450, 2.0)]    # the same with a filler range
    bg = [c[0] for c in cases]
    z = [c[1] for c in cases]
    nf = [c[2] for c in cases]
    el = [c[3] for c in cases]
    eb = [c[4] for c in cases]
    rate = [((22 if b == 0 else 10) * zz - f) / e for b, zz, f, e, _ in cases]
    seed = 1234
    e_llr, e_off, info, i_off = traffic_blocks(bg, z, nf, el, eb, rate, seed, threads=3)
    for i, (b, zz, f, e, ebno) in enumerate(cases):
        code = oracle.code_3gpp(b, zz)
        kp = code.k - f
        inf = np.concatenate([info[i_off[i]:i_off[i] + kp], np.zeros(f, np.uint8)])
        cw = oracle.encode(code, inf)
        bits = oracle.rate_match(cw, e, 0, 1, b, zz, f)
        s = (seed ^ ((0xD1B54A32D192ED03 * (i + 1)) & 0xFFFFFFFFFFFFFFFF)) & 0xFFFFFFFFFFFFFFFF
        ref = oracle.awgn(np.float32(ebno), np.float32(rate[i]), s, bits)
        assert np.array_equal(f32_bits(e_llr[e_off[i]:e_off[i] + e]), f32_bits(ref)), i


def test_traffic_generator_with_given_info(oracle):
    """syndrome_traffic_blocks_info: caller's info bits through the same chain."""
    from syndrome_amd import traffic_blocks_info
    cases = [(0, 64, 12, 5000, 1.0), (1, 30, 0, 700, 2.5)]
    rng = np.random.default_rng(2)
    bg = [c[0] for c in cases]
    z = [c[1] for c in cases]
    nf = [c[2] for c in cases]
    el = [c[3] for c in cases]
    eb = [c[4] for c in cases]
    kp = [(22 if b == 0 else 10) * zz - f for b, zz, f in zip(bg, z, nf)]
    info = rng.integers(0, 2, sum(kp)).astype(np.uint8)
    i_off = np.concatenate([[0], np.cumsum(kp)[:-1]]).astype(np.int64)
    rate = [k / e for k, e in zip(kp, el)]
    e_llr, e_off = traffic_blocks_info(bg, z, nf, el, eb, rate, info, i_off, 99)
    for i in range(len(cases)):
        code = oracle.code_3gpp(bg[i], z[i])
        cw = oracle.encode(code, np.concatenate([info[i_off[i]:i_off[i] + kp[i]], np.zeros(nf[i], np.uint8)]))
        bits = oracle.rate_match(cw, el[i], 0, 1, bg[i], z[i], nf[i])
        s = (99 ^ ((0xD1B54A32D192ED03 * (i + 1)) & 0xFFFFFFFFFFFFFFFF)) & 0xFFFFFFFFFFFFFFFF
        ref = oracle.awgn(np.float32(eb[i]), np.float32(rate[i]), s, bits)
        assert np.array_equal(f32_bits(e_llr[e_off[i]:e_off[i] + el[i]]), f32_bits(ref)), i
