"""Robustness and API-equivalence checks on the GPU, after the reference's
own: LLRs with NaN, +-inf, subnormals and signed zeros (tests/robustness.rs
fuzz values) decode without error and as the oracle does; the workspace path
equals the raw-buffer path (qc_ldpc.rs:2787-2816); concurrent decodes on one
handle from several threads (the decoder is Send + Sync, qc_ldpc.rs:480-484)
each give the oracle's result."""
import threading

import numpy as np
import pytest

from tests.helpers import channel, codewords, f32_bits, say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _fuzz_llrs(rng, n):
    """The fuzz distribution of robustness.rs:101-116."""
    kind = rng.integers(0, 16, n)
    raw = (rng.integers(0, 20001, n) / 100.0).astype(np.float32) * np.where(rng.integers(0, 2, n) == 1, -1, 1)
    special = np.array([np.nan, -np.nan, np.inf, -np.inf, np.float32(1.17549435e-38) / 4,
                        -np.float32(1.17549435e-38) / 4, 0.0, -0.0], np.float32)
    out = raw.astype(np.float32)
    m = kind < 8
    out[m] = special[kind[m]]
    return out


def _same_f32(a, b):
    """Bit-equal, except that NaN payloads may differ (IEEE leaves the
    propagated payload to the implementation)."""
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(f32_bits(a[~na]), f32_bits(b[~nb]))


@pytest.mark.parametrize("plain", [False, True], ids=["fast", "plain"])
def test_special_float_llrs_decode_like_the_oracle(oracle, plain):
    import syndrome_amd
    from syndrome_amd import QcLdpcDecoder, _lib as L
    syndrome_amd.lib()
    rng = np.random.default_rng(17)
    for bg, z in ((0, 16), (1, 52), (0, 384)):
        say(f"[fuzz] BG{bg + 1} Z={z} plain={plain}: NaN / inf / subnormal / signed-zero LLRs")
        code = oracle.code_3gpp(bg, z)
        dec = QcLdpcDecoder.with_lifting_size(bg, z, 0.5)
        B = 6
        llr = np.stack([_fuzz_llrs(rng, code.n) for _ in range(B)])
        d_in = torch.from_numpy(llr).cuda()
        d_post = torch.empty_like(d_in)
        d_hard = torch.empty(llr.shape, dtype=torch.uint8, device="cuda")
        d_edge = torch.empty((B, dec.required_edge_buffer()), dtype=torch.float32, device="cuda")
        d_it = torch.empty(B, dtype=torch.int32, device="cuda")
        flags = L.WRITE_POST | L.WRITE_HARD | L.WRITE_EDGE | (L.PLAIN if plain else 0)
        dec.decode_batch_f32(d_in, d_post, d_hard, d_it, 6, flags=flags, d_edge=d_edge)
        torch.cuda.synchronize()
        post, hard, edge, its = (d_post.cpu().numpy(), d_hard.cpu().numpy(), d_edge.cpu().numpy(),
                                 d_it.cpu().numpy())
        for b in range(B):
            ri, rp, re, rh = oracle.decode_f32(code, 0.5, llr[b], 6)
            assert its[b] == ri and np.array_equal(hard[b], rh), f"BG{bg + 1} Z={z} codeword {b}"
            assert _same_f32(post[b], rp) and _same_f32(edge[b], re), f"BG{bg + 1} Z={z} codeword {b}: values"
        # the int8 path quantizes them (NaN -> 0, +-inf -> +-127) and stays exact
        d_post8 = torch.empty(llr.shape, dtype=torch.int16, device="cuda")
        dec.decode_batch_i8(d_in, d_post8, d_hard, d_it, 6, flags=L.WRITE_POST | L.WRITE_HARD | L.QUANTIZE |
                            (L.PLAIN if plain else 0), d_edge=torch.empty((B, dec.required_edge_buffer()),
                                                                           dtype=torch.int8, device="cuda"))
        torch.cuda.synchronize()
        post8, hard8, its8 = d_post8.cpu().numpy(), d_hard.cpu().numpy(), d_it.cpu().numpy()
        for b in range(B):
            ri, ra, _, rh = oracle.decode_i8(code, 0.5, oracle.quantize_i16(llr[b], 8.0), 6)
            assert its8[b] == ri and np.array_equal(post8[b], ra) and np.array_equal(hard8[b], rh)


@pytest.mark.parametrize("plain", [False, True], ids=["fast", "plain"])
def test_special_offsets_decode_like_the_oracle(oracle, plain):
    """The offset floor mag = max(me - beta, 0) at zero, negative, huge,
    infinite and NaN beta.  For a NaN beta the oracle follows the scalar
    reference (qc_ldpc.rs:1195-1196, where "mag < 0" keeps the NaN); the
    AVX2 path's max(0.0) (simd_avx2.rs:276) would give 0 instead."""
    import syndrome_amd
    from syndrome_amd import QcLdpcDecoder, _lib as L
    syndrome_amd.lib()
    bg, z = 0, 384
    code = oracle.code_3gpp(bg, z)
    cws = codewords(oracle, code, 3, seed=41)
    llr = channel(oracle, code, cws, [1.0, 2.0, 3.0], seed=43)
    for beta in (0.0, -0.5, 1e30, float("inf"), float("nan")):
        say(f"[offset] BG1 Z=384 beta={beta} plain={plain}")
        dec = QcLdpcDecoder.with_lifting_size(bg, z, beta)
        B = llr.shape[0]
        d_in = torch.from_numpy(llr).cuda()
        d_post = torch.empty_like(d_in)
        d_hard = torch.empty(llr.shape, dtype=torch.uint8, device="cuda")
        d_edge = torch.empty((B, dec.required_edge_buffer()), dtype=torch.float32, device="cuda")
        d_it = torch.empty(B, dtype=torch.int32, device="cuda")
        flags = L.WRITE_POST | L.WRITE_HARD | L.WRITE_EDGE | (L.PLAIN if plain else 0)
        dec.decode_batch_f32(d_in, d_post, d_hard, d_it, 5, flags=flags, d_edge=d_edge)
        torch.cuda.synchronize()
        post, hard, edge, its = (d_post.cpu().numpy(), d_hard.cpu().numpy(), d_edge.cpu().numpy(),
                                 d_it.cpu().numpy())
        for b in range(B):
            ri, rp, re, rh = oracle.decode_f32(code, beta, llr[b], 5)
            assert its[b] == ri and np.array_equal(hard[b], rh), f"beta {beta} codeword {b}"
            assert _same_f32(post[b], rp) and _same_f32(edge[b], re), f"beta {beta} codeword {b}: values"


def test_workspace_path_equals_raw_path(oracle):
    from syndrome_amd import QcLdpcDecoder
    say("[workspace] decode_with_workspace == decode_layered_offset_min_sum (BG1 Z=96)")
    code = oracle.code_3gpp(0, 96)
    dec = QcLdpcDecoder.with_lifting_size(0, 96, 0.5)
    cw = codewords(oracle, code, 1, seed=21)[0]
    llr = channel(oracle, code, cw[None, :], [1.0], seed=4)[0]
    a = llr.copy()
    er = np.zeros(dec.required_edge_buffer(), np.float32)
    sc = np.zeros(dec.required_layer_buffer(), np.float32)
    hd = np.zeros(dec.variable_node_count(), np.uint8)
    it_raw = dec.decode_layered_offset_min_sum(a, er, sc, hd, 8)
    b = llr.copy()
    ws = dec.workspace()
    it_ws = dec.decode_with_workspace(b, ws, 8)
    assert it_raw == it_ws and np.array_equal(f32_bits(a), f32_bits(b))
    assert np.array_equal(hd, ws.hard_output()) and np.array_equal(f32_bits(er), f32_bits(ws.edge_r))


def test_concurrent_decodes_share_one_handle(oracle):
    """12 threads on one handle, more than its 8 staging areas, mixing the
    f32 and int8 host-buffer decodes: every call returns the oracle's result
    (the calls overlap on the device, each on its own staging and stream)."""
    from syndrome_amd import QcLdpcDecoder, QuantParams
    T, K = 12, 36
    say(f"[threads] {T} threads x {K // T} decode_5g / decode_5g_i8 calls on one BG2 Z=104 handle")
    code = oracle.code_3gpp(1, 104)
    dec = QcLdpcDecoder.with_lifting_size(1, 104, 0.5)
    cws = codewords(oracle, code, K, seed=33)
    llrs = channel(oracle, code, cws, list(np.linspace(0.0, 2.5, K)), seed=8)
    results = [None] * K
    errors = []

    def work(tid):
        try:
            for j in range(tid, K, T):
                hd = np.zeros(dec.variable_node_count(), np.uint8)
                if j % 3 == 2:
                    a = oracle.quantize_i16(llrs[j], 8.0)
                    er = np.zeros(dec.required_edge_buffer(), np.int8)
                    sc = np.zeros(dec.required_layer_buffer(), np.int8)
                    used = dec.decode_5g_i8(a, 0, er, sc, hd, 8, QuantParams())
                else:
                    a = llrs[j].copy()
                    er = np.zeros(dec.required_edge_buffer(), np.float32)
                    sc = np.zeros(dec.required_layer_buffer(), np.float32)
                    used = dec.decode_5g(a, 0, er, sc, hd, 8)
                results[j] = (used, a, er, hd)
        except Exception as e:  # surfaced below
            errors.append(e)

    th = [threading.Thread(target=work, args=(t,)) for t in range(T)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=60)
    assert not errors and all(r is not None for r in results)
    for j in range(K):
        used, a, er, hd = results[j]
        if j % 3 == 2:
            app = oracle.init_5g_i8(code, oracle.quantize_i16(llrs[j], 8.0), 0)
            ri, ra, re, rh = oracle.decode_i8(code, 0.5, app, 8)
            assert used == ri and np.array_equal(a, ra), f"decode {j}"
            assert np.array_equal(er, re) and np.array_equal(hd, rh), f"decode {j}"
        else:
            ri, rp, re, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, llrs[j], 0), 8)
            assert used == ri and np.array_equal(f32_bits(a), f32_bits(rp)), f"decode {j}"
            assert np.array_equal(f32_bits(er), f32_bits(re)) and np.array_equal(hd, rh), f"decode {j}"


def test_device_quantizer_known_answers_and_oracle(oracle):
    """syndrome_ldpc_quantize_i16_device against quantize.rs's known answers
    (quantize.rs:372-437: round-half-away, saturation at +-127, NaN -> 0,
    -inf -> -127) and against the oracle's quantize_llr_i16 on a large vector
    holding every special value, at two scales, on a caller stream."""
    import syndrome_amd
    from syndrome_amd.ldpc import quantize_llr_i16_device
    syndrome_amd.lib()
    cases = [([0, 1, -1, 1e9, -np.inf], 8.0, [0, 8, -8, 127, -127]),
             ([np.nan], 8.0, [0]),
             ([1.0, -0.25, 99.0], 8.0, [8, -2, 127]),
             ([0.0625, -0.0625, 0.1875], 8.0, [1, -1, 2])]
    for x, scale, want in cases:
        d_x = torch.tensor(x, dtype=torch.float32, device="cuda")
        d_q = torch.full((len(x),), 999, dtype=torch.int16, device="cuda")
        quantize_llr_i16_device(d_x, d_q, scale)
        torch.cuda.synchronize()
        assert d_q.cpu().tolist() == want, (x, scale)
    rng = np.random.default_rng(23)
    n = (1 << 20) + 37  # ragged tail past any vector width
    x = _fuzz_llrs(rng, n)
    x[: n // 2] = rng.normal(0, 20, n // 2).astype(np.float32)
    s = torch.cuda.Stream()
    for scale in (8.0, 3.5):
        say(f"[quantize] {n} LLRs, scale {scale}, caller stream")
        d_x = torch.from_numpy(x).cuda()
        d_q = torch.empty(n, dtype=torch.int16, device="cuda")
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            quantize_llr_i16_device(d_x, d_q, scale, stream=s.cuda_stream)
        s.synchronize()
        assert np.array_equal(d_q.cpu().numpy(), oracle.quantize_i16(x, scale))


def test_device_quantize_llr_i8_and_dequantize(oracle):
    """syndrome_ldpc_quantize_i8_device / dequantize_i8_device: quantize_llr
    and dequantize_llr (quantize.rs:148-169) on the doc example's known
    answers and on a large fuzz vector against the oracle, two scales."""
    import syndrome_amd
    from syndrome_amd.ldpc import dequantize_llr_i8_device, quantize_llr_i8_device
    syndrome_amd.lib()
    d_x = torch.tensor([1.5, -3.0, 0.0, 200.0, float("nan"), float("-inf")], dtype=torch.float32, device="cuda")
    d_q = torch.full((6,), 99, dtype=torch.int8, device="cuda")
    quantize_llr_i8_device(d_x, d_q, 8.0)
    d_back = torch.empty(6, dtype=torch.float32, device="cuda")
    dequantize_llr_i8_device(d_q, d_back, 8.0)
    torch.cuda.synchronize()
    assert d_q.cpu().tolist() == [12, -24, 0, 127, 0, -127]
    assert d_back.cpu().tolist() == [1.5, -3.0, 0.0, 15.875, 0.0, -15.875]
    rng = np.random.default_rng(29)
    n = (1 << 20) + 13
    x = _fuzz_llrs(rng, n)
    x[: n // 2] = rng.normal(0, 20, n // 2).astype(np.float32)
    for scale in (8.0, 3.0):
        say(f"[quantize i8] {n} LLRs, scale {scale}")
        d_x = torch.from_numpy(x).cuda()
        d_q = torch.empty(n, dtype=torch.int8, device="cuda")
        d_f = torch.empty(n, dtype=torch.float32, device="cuda")
        quantize_llr_i8_device(d_x, d_q, scale)
        dequantize_llr_i8_device(d_q, d_f, scale)
        torch.cuda.synchronize()
        q = oracle.quantize_i8(x, scale)
        assert np.array_equal(d_q.cpu().numpy(), q)
        assert np.array_equal(f32_bits(d_f.cpu().numpy()), f32_bits(oracle.dequantize_i8(q, scale)))
