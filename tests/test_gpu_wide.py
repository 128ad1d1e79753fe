"""Raw-edge codes over from_raw_edges' whole domain on the GPU
(qc_ldpc.rs:412-455: u8 row and column indices, any number of edges,
repeated (row, col) pairs).  Codes past the inline tables (more than 46 rows,
68 columns or 316 edges, or a column twice in one layer) run the wide
kernels; every result is compared bit for bit with the oracle: iteration
counts, posteriors, R messages, hard decisions, traces, on the host-buffer
entry points and the device batch ones, f32 and int8.  Inputs are AWGN-like
LLRs around the all-zero codeword (a codeword of every code), strong enough
that some codewords converge early and others run every iteration."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import f32_bits, say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    import syndrome_amd
    syndrome_amd.lib()
    return torch.device("cuda:0")


def _dup_rows(rng, m_b, n_b, z, deg, dups):
    """m_b rows of `deg` distinct columns plus `dups` repeats of columns
    already in the row (random shift, sometimes the same one), shuffled."""
    ent = []
    for r in range(m_b):
        cs = list(rng.choice(n_b, deg, replace=False))
        row = [(r, int(c), int(rng.integers(0, z))) for c in cs]
        for j in range(dups):
            c, s = row[j % len(row)][1], row[j % len(row)][2]
            row.append((r, c, s if j % 3 == 0 else int(rng.integers(0, z))))
        rng.shuffle(row)
        ent += row
    return ent


def _codes():
    rng = np.random.default_rng(2024)
    out = {}
    # the worked example of the oracle's tests: column 0 twice in one layer
    out["hand"] = ([(0, 0, 0), (0, 0, 1), (0, 1, 0)], 1, 2, 2)
    # repeated columns in most layers, a degree-1 row, Z=24
    ent = _dup_rows(rng, 6, 14, 24, 5, 2) + [(6, 3, 5)]
    out["dup"] = (ent, 7, 14, 24)
    # more rows, columns and edges than any 3GPP graph: rows 0..255 used with
    # gaps, 300 rows in all, 320 columns (u8 indices name 0..255, the rest
    # stay unused), ~880 edges, no repeats
    ent = []
    for r in range(256):
        if r % 7 == 3:
            continue
        for c in rng.choice(256, 4, replace=False):
            ent.append((r, int(c), int(rng.integers(0, 4))))
    out["tall"] = (ent, 300, 320, 4)
    # a posterior larger than one workgroup's LDS (255 * 384 variables):
    # the wide kernel keeps it in HBM; some layers repeat a column
    ent = _dup_rows(rng, 24, 255, 384, 7, 0) + _dup_rows(rng, 4, 255, 384, 5, 1)
    ent = [(r if i < 24 * 7 else r + 24, c, s) for i, (r, c, s) in enumerate(ent)]
    out["hbm"] = (ent, 28, 255, 384)
    return out


CODES = _codes()


def _mk(oracle, name):
    from syndrome_amd import QcLdpcDecoder
    ent, m_b, n_b, z = CODES[name]
    e = np.array(ent, np.int64).reshape(-1, 3)
    dec = QcLdpcDecoder.from_raw_edges(e, m_b, n_b, z, 0.5)
    code = oracle.code_raw(e[:, 0], e[:, 1], e[:, 2], m_b, n_b, z)
    assert (dec.variable_node_count(), dec.required_edge_buffer(), dec.required_layer_buffer()) == \
        (code.n, code.edge_buf, code.layer_buf)
    return dec, code


def _llr(code, count, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(2.5, 2.0, (count, code.n)).astype(np.float32)


@pytest.mark.parametrize("name", list(CODES))
def test_wide_host_f32_matches_oracle(dev, oracle, name):
    dec, code = _mk(oracle, name)
    say(f"[wide f32 host] {name}: N={code.n} E={code.n_edges} Z={code.z} shape={dec.launch_shape()}")
    for i, x in enumerate(_llr(code, 2, 11)):
        a = x.copy()
        er = np.full(dec.required_edge_buffer(), np.nan, np.float32)
        sc = np.zeros(dec.required_layer_buffer(), np.float32)
        hd = np.full(code.n, 7, np.uint8)
        used = dec.decode_layered_offset_min_sum(a, er, sc, hd, 6)
        ri, rp, re, rh = oracle.decode_f32(code, 0.5, x, 6)
        assert used == ri, (name, i)
        assert np.array_equal(f32_bits(a), f32_bits(rp)), (name, i)
        assert np.array_equal(f32_bits(er), f32_bits(re)), (name, i)
        assert np.array_equal(hd, rh), (name, i)


@pytest.mark.parametrize("name", list(CODES))
def test_wide_host_i8_matches_oracle(dev, oracle, name):
    from syndrome_amd import QuantParams
    dec, code = _mk(oracle, name)
    say(f"[wide i8 host] {name}")
    for clamp in (32767, 200):
        app0 = oracle.quantize_i16(_llr(code, 1, 12)[0], 8.0)
        app = app0.copy()
        er = np.zeros(dec.required_edge_buffer(), np.int8)
        sc = np.zeros(dec.required_layer_buffer(), np.int8)
        hd = np.zeros(code.n, np.uint8)
        used = dec.decode_layered_offset_min_sum_i8(app, er, sc, hd, 6, QuantParams(8.0, clamp))
        ri, ra, re, rh = oracle.decode_i8(code, 0.5, np.clip(app0, -clamp, clamp), 6, 8.0, clamp)
        assert used == ri, (name, clamp)
        assert np.array_equal(app, ra) and np.array_equal(er, re) and np.array_equal(hd, rh), (name, clamp)


@pytest.mark.parametrize("name", ["dup", "tall", "hbm"])
def test_wide_batch_matches_oracle(dev, oracle, name):
    from syndrome_amd import QuantParams, _lib as L
    dec, code = _mk(oracle, name)
    B = 3
    x = _llr(code, B, 13)
    say(f"[wide batch] {name}: B={B}")
    flags = L.WRITE_POST | L.WRITE_HARD | L.WRITE_EDGE
    d_in = torch.from_numpy(x).cuda()
    d_post = torch.empty_like(d_in)
    d_hard = torch.empty(x.shape, dtype=torch.uint8, device="cuda")
    d_edge = torch.empty((B, dec.required_edge_buffer()), dtype=torch.float32, device="cuda")
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    dec.decode_batch_f32(d_in, d_post, d_hard, d_it, 5, flags=flags, d_edge=d_edge)
    torch.cuda.synchronize()
    post, hard, edge, its = d_post.cpu().numpy(), d_hard.cpu().numpy(), d_edge.cpu().numpy(), d_it.cpu().numpy()
    for b in range(B):
        ri, rp, re, rh = oracle.decode_f32(code, 0.5, x[b], 5)
        assert its[b] == ri and np.array_equal(f32_bits(post[b]), f32_bits(rp)), (name, b)
        assert np.array_equal(f32_bits(edge[b]), f32_bits(re)) and np.array_equal(hard[b], rh), (name, b)
    # int8 with the quantizer fused into the load, on a caller stream
    s = torch.cuda.Stream()
    d_p8 = torch.empty(x.shape, dtype=torch.int16, device="cuda")
    d_e8 = torch.empty((B, dec.required_edge_buffer()), dtype=torch.int8, device="cuda")
    with torch.cuda.stream(s):
        dec.decode_batch_i8(d_in, d_p8, d_hard, d_it, 5, flags=flags | L.QUANTIZE, d_edge=d_e8,
                            quant=QuantParams(8.0, 300), stream=s.cuda_stream)
    s.synchronize()
    p8, h8, e8, i8 = d_p8.cpu().numpy(), d_hard.cpu().numpy(), d_e8.cpu().numpy(), d_it.cpu().numpy()
    for b in range(B):
        app = np.clip(oracle.quantize_i16(x[b], 8.0), -300, 300)
        ri, ra, re, rh = oracle.decode_i8(code, 0.5, app, 5, 8.0, 300)
        assert i8[b] == ri and np.array_equal(p8[b], ra) and np.array_equal(e8[b], re), (name, b)
        assert np.array_equal(h8[b], rh), (name, b)


@pytest.mark.parametrize("name", ["hand", "dup", "hbm"])
def test_wide_traced_matches_oracle(dev, oracle, name):
    dec, code = _mk(oracle, name)
    say(f"[wide traced] {name}")
    x = _llr(code, 1, 14)[0] * 0.5
    rows = []
    a = x.copy()
    er = np.zeros(dec.required_edge_buffer(), np.float32)
    sc = np.zeros(dec.required_layer_buffer(), np.float32)
    hd = np.zeros(code.n, np.uint8)
    used = dec.decode_layered_offset_min_sum_traced(a, er, sc, hd, 4, lambda i, p: rows.append((i, p.copy())))
    ri, rp, re, rh, rt = oracle.decode_f32(code, 0.5, x, 4, trace=True)
    assert used == ri and [i for i, _ in rows] == list(range(1, ri + 1))
    for (i, p), t in zip(rows, rt):
        assert np.array_equal(f32_bits(p), f32_bits(t)), (name, i)
    assert np.array_equal(f32_bits(a), f32_bits(rp))


def test_wide_decode_5g_and_static(dev, oracle):
    """decode_5g / decode_5g_i8 (the 5G prologue on a raw code: K = (n_b -
    m_b) Z, 2Z punctured) and the allocation-free decode_5g on a wide code
    whose staging fits its array."""
    from syndrome_amd import QuantParams
    L = __import__("syndrome_amd").lib()
    for name, nf in (("dup", 9), ("tall", 0)):
        dec, code = _mk(oracle, name)
        if code.n_b < code.m_b:
            continue
        say(f"[wide decode_5g] {name} n_filler={nf}")
        x = _llr(code, 1, 15)[0]
        a = x.copy()
        er = np.zeros(dec.required_edge_buffer(), np.float32)
        sc = np.zeros(dec.required_layer_buffer(), np.float32)
        hd = np.zeros(code.n, np.uint8)
        used = dec.decode_5g(a, nf, er, sc, hd, 6)
        ri, rp, re, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, x, nf), 6)
        assert used == ri and np.array_equal(f32_bits(a), f32_bits(rp)) and np.array_equal(f32_bits(er), f32_bits(re))
        # the static entry point gives the same bits
        b = x.copy()
        er2 = np.zeros_like(er)
        hd2 = np.zeros_like(hd)
        u2 = C.c_size_t(0)
        p = lambda v: v.ctypes.data_as(C.c_void_p)
        rc = L.syndrome_ldpc_decode_5g_static(dec._h, p(b), b.size, nf, p(er2), er2.size, p(sc), sc.size, p(hd2),
                                              hd2.size, 6, C.byref(u2), None)
        assert rc == 0 and u2.value == used
        assert np.array_equal(f32_bits(b), f32_bits(a)) and np.array_equal(f32_bits(er2), f32_bits(er))
        app0 = oracle.quantize_i16(x, 8.0)
        app = app0.copy()
        e8 = np.zeros(dec.required_edge_buffer(), np.int8)
        s8 = np.zeros(dec.required_layer_buffer(), np.int8)
        used = dec.decode_5g_i8(app, nf, e8, s8, hd, 6, QuantParams())
        ri, ra, re, rh = oracle.decode_i8(code, 0.5, oracle.init_5g_i8(code, app0, nf), 6)
        assert used == ri and np.array_equal(app, ra) and np.array_equal(e8, re)


def test_wide_degenerate_codes(dev, oracle):
    """Codes the reference accepts and cannot decode as 5G: no variables,
    no checks, more rows than columns.  Plain decodes run one pass (every
    check holds), leave the posterior as given; decode_5g returns PANIC where
    init_5g_llr panics (qc_ldpc.rs:720-731) and the batch prologue is
    refused."""
    from syndrome_amd import QcLdpcDecoder, SyndromeError, _lib as L
    # no checks: every check holds after the first pass
    dec = QcLdpcDecoder.from_raw_edges(np.zeros((0, 3), np.int64), 0, 5, 8, 0.5)
    x = np.linspace(-3, 3, 40).astype(np.float32)
    a = x.copy()
    hd = np.zeros(40, np.uint8)
    used = dec.decode_layered_offset_min_sum(a, np.zeros(0, np.float32), np.zeros(0, np.float32), hd, 5)
    assert used == 1 and np.array_equal(a, x) and np.array_equal(hd, (x < 0).astype(np.uint8))
    # no variables
    dec = QcLdpcDecoder.from_raw_edges(np.zeros((0, 3), np.int64), 3, 0, 8, 0.5)
    assert dec.variable_node_count() == 0 and dec.check_node_count() == 24
    # more rows than columns: decode_5g's prologue is undefined
    dec = QcLdpcDecoder.from_raw_edges(np.array([(0, 0, 1), (2, 1, 0), (1, 1, 3)]), 3, 2, 4, 0.5)
    x = np.linspace(-1, 1, 8).astype(np.float32)
    a = x.copy()
    with pytest.raises(SyndromeError) as e:
        dec.decode_5g(a, 0, np.zeros(12, np.float32), np.zeros(4, np.float32), np.zeros(8, np.uint8), 3)
    assert e.value.status == L.PANIC and np.array_equal(a, x)
    d_in = torch.from_numpy(x[None, :]).cuda()
    with pytest.raises(SyndromeError) as e:
        dec.decode_batch_f32(d_in, torch.empty_like(d_in), torch.empty((1, 8), dtype=torch.uint8, device="cuda"),
                             torch.empty(1, dtype=torch.int32, device="cuda"), 3,
                             flags=L.INIT_5G | L.WRITE_POST | L.WRITE_HARD,
                             d_edge=torch.empty((1, 12), dtype=torch.float32, device="cuda"))
    assert e.value.status == L.INVALID_PARAM
    # and decodes as the oracle without the prologue
    code = oracle.code_raw([0, 2, 1], [0, 1, 1], [1, 0, 3], 3, 2, 4)
    er = np.zeros(12, np.float32)
    used = dec.decode_layered_offset_min_sum(a, er, np.zeros(4, np.float32), hd[:8], 3)
    ri, rp, re, rh = oracle.decode_f32(code, 0.5, x, 3)
    assert used == ri and np.array_equal(f32_bits(a), f32_bits(rp)) and np.array_equal(f32_bits(er), f32_bits(re))


def test_wide_equals_inline_on_a_padded_3gpp_code(dev, oracle):
    """BG2 Z=52 given as raw edges with 9 empty rows and 2 unused columns
    runs the wide kernel; on the 3GPP code's variables it must give exactly
    what the 3GPP decoder gives (register-resident kernel), and leave the
    extra variables as they came in."""
    from syndrome_amd import QcLdpcDecoder, _lib as L
    base = oracle.code_3gpp(1, 52)
    ent = [(r, base.cols[e], base.shifts[e]) for r in range(base.m_b)
           for e in range(base.layer_off[r], base.layer_off[r + 1])]
    wide = QcLdpcDecoder.from_raw_edges(np.array(ent), base.m_b + 9, base.n_b + 2, 52, 0.5)
    ref = QcLdpcDecoder.with_lifting_size(1, 52, 0.5)
    B, n0, n1 = 8, base.n, base.n + 2 * 52
    rng = np.random.default_rng(31)
    x = rng.normal(1.8, 2.0, (B, n1)).astype(np.float32)
    say(f"[wide == inline] BG2 Z=52 padded, B={B}")
    outs = []
    for dec, n in ((wide, n1), (ref, n0)):
        d_in = torch.from_numpy(np.ascontiguousarray(x[:, :n])).cuda()
        d_post = torch.empty_like(d_in)
        d_hard = torch.empty((B, n), dtype=torch.uint8, device="cuda")
        d_it = torch.empty(B, dtype=torch.int32, device="cuda")
        d_edge = torch.empty((B, dec.required_edge_buffer()), dtype=torch.float32, device="cuda")
        dec.decode_batch_f32(d_in, d_post, d_hard, d_it, 8, flags=L.WRITE_POST | L.WRITE_HARD | L.WRITE_EDGE,
                             d_edge=d_edge)
        torch.cuda.synchronize()
        outs.append((d_post.cpu().numpy(), d_hard.cpu().numpy(), d_it.cpu().numpy(), d_edge.cpu().numpy()))
    (pw, hw, iw, ew), (pr, hr, ir, er) = outs
    assert np.array_equal(iw, ir)
    assert np.array_equal(f32_bits(pw[:, :n0]), f32_bits(pr)) and np.array_equal(pw[:, n0:], x[:, n0:])
    assert np.array_equal(hw[:, :n0], hr) and np.array_equal(f32_bits(ew), f32_bits(er))
