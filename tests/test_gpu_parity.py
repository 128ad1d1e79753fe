"""GPU parity: the HIP decode kernels against the oracle, through the C ABI.

The bar (bit-exact):
* f32: iterations used, hard decisions, posterior LLRs and final R messages
  identical bit for bit to the scalar restatement of qc_ldpc.rs:962-1252.
  (The stated tolerance for posteriors is 1e-6 relative; these kernels
  perform the reference's own IEEE operations in its order, so they meet
  it with zero difference, and the tests assert exact equality.)
* int8: iterations, app, edge_r and hard identical (qc_ldpc.rs:1530-1787).
"""
import numpy as np
import pytest

from tests.helpers import channel, codewords, f32_bits, say

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    import syndrome_amd
    syndrome_amd.lib()
    props = torch.cuda.get_device_properties(0)
    say(f"device: {props.name} CUs={props.multi_processor_count} "
        f"lds/block={getattr(props, 'shared_memory_per_block', '?')}")
    return torch.device("cuda:0")


def _oracle_f32(orc, code, llr, beta, iters, init5g, n_filler=0):
    out = []
    for i in range(llr.shape[0]):
        x = orc.init_5g_f32(code, llr[i], n_filler) if init5g else llr[i]
        out.append(orc.decode_f32(code, beta, x, iters))
    return out


def _run_batch_f32(dec, llr, iters, init5g, plain, n_filler=0):
    import syndrome_amd._lib as L
    B, N = llr.shape
    d_in = torch.from_numpy(llr).cuda()
    d_post = torch.empty_like(d_in)
    d_hard = torch.empty((B, N), dtype=torch.uint8, device="cuda")
    d_edge = torch.empty((B, dec.required_edge_buffer()), dtype=torch.float32, device="cuda")
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    flags = L.WRITE_POST | L.WRITE_HARD | L.WRITE_EDGE | (L.INIT_5G if init5g else 0) | (L.PLAIN if plain else 0)
    dec.decode_batch_f32(d_in, d_post, d_hard, d_it, iters, n_filler=n_filler, flags=flags, d_edge=d_edge)
    torch.cuda.synchronize()
    return d_post.cpu().numpy(), d_hard.cpu().numpy(), d_edge.cpu().numpy(), d_it.cpu().numpy()


F32_CASES = [(0, 384), (0, 2), (0, 7), (0, 44), (0, 96), (0, 208), (0, 352), (1, 208), (1, 13), (1, 52),
             (1, 128), (1, 384), (1, 3)]


@pytest.mark.parametrize("plain", [False, True], ids=["fast", "plain"])
@pytest.mark.parametrize("bg,z", F32_CASES)
def test_f32_batch_matches_oracle(dev, oracle, bg, z, plain):
    from syndrome_amd import QcLdpcDecoder
    say(f"[f32 batch] BG{bg + 1} Z={z} plain={plain}")
    code = oracle.code_3gpp(bg, z)
    dec = QcLdpcDecoder.with_lifting_size(bg, z, 0.25)
    B = 6
    cws = codewords(oracle, code, B, seed=1000 * z + bg)
    llr = channel(oracle, code, cws, [0.0, 1.0, 1.5, 2.5, 4.0, 0.5], seed=17 * z)
    iters = 8
    post, hard, edge, it = _run_batch_f32(dec, llr, iters, init5g=True, plain=plain)
    ref = _oracle_f32(oracle, code, llr, 0.25, iters, init5g=True)
    for b in range(B):
        ri, rp, re, rh = ref[b]
        assert it[b] == ri, f"codeword {b}: iterations {it[b]} != {ri}"
        assert np.array_equal(hard[b], rh), f"codeword {b}: hard decisions differ"
        assert np.array_equal(f32_bits(post[b]), f32_bits(rp)), f"codeword {b}: posterior differs"
        assert np.array_equal(f32_bits(edge[b][: code.edge_buf]), f32_bits(re)), f"codeword {b}: edge_r differs"


def test_checksum_golden_on_gpu(dev, oracle):
    """The reference's cross-language gate: FNV-1a of hard bits after 10
    iterations of the +-0.5 input, BG1 Z=384, beta 0.25 = a2448e8cf5272825
    (bench/results/ldpc_rust.checksum)."""
    from syndrome_amd import QcLdpcDecoder
    say("[checksum] BG1 Z=384 +-0.5, 10 iterations, fast and plain")
    dec = QcLdpcDecoder.with_lifting_size(0, 384, 0.25)
    n = dec.variable_node_count()
    llr = np.where(np.arange(n) % 2 == 1, -0.5, 0.5).astype(np.float32)[None, :]
    for plain in (False, True):
        _, hard, _, it = _run_batch_f32(dec, llr, 10, init5g=False, plain=plain)
        assert it[0] == 10
        assert oracle.fnv1a(hard[0]) == 0xA2448E8CF5272825


@pytest.mark.parametrize("plain", [False, True], ids=["fast", "plain"])
def test_wifi_convergence_fixture_on_gpu(dev, oracle, plain):
    """ldpc_convergence.json: Wi-Fi Z=27 R=1/2, beta 0.25, seed 139, 13 iterations."""
    import json
    import os
    from syndrome_amd import QcLdpcDecoder
    say("[wifi fixture] Z=27 R=1/2")
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ldpc_convergence.json")))
    code = oracle.code_wifi(27, 1, 2)
    cw = np.array(d["transmitted_codeword"], np.uint8)
    llr = oracle.awgn(d["channel"]["eb_n0_db"], 0.5, d["channel"]["seed"], cw)
    dec = QcLdpcDecoder.wifi(27, 1, 2, 0.25)
    for it_budget in (d["max_iterations"],):
        post, hard, edge, it = _run_batch_f32(dec, llr[None, :], it_budget, init5g=False, plain=plain)
        assert it[0] == d["iterations_used"]
        assert np.array_equal(hard[0], np.array(d["frames"][-1]["hard_bits"], np.uint8))
    # every intermediate frame: decode with a budget of k iterations
    for k in range(1, d["iterations_used"] + 1):
        post, hard, edge, it = _run_batch_f32(dec, llr[None, :], k, init5g=False, plain=plain)
        assert np.array_equal((post[0] < 0).astype(np.uint8), np.array(d["frames"][k]["hard_bits"], np.uint8)), k


WIFI = [(z, rn, rd) for z in (27, 54, 81) for rn, rd in ((1, 2), (2, 3), (3, 4), (5, 6))]


@pytest.mark.parametrize("plain", [False, True], ids=["fast", "plain"])
@pytest.mark.parametrize("z,rn,rd", WIFI)
def test_wifi_matrices_match_oracle(dev, oracle, z, rn, rd, plain):
    """All 12 IEEE 802.11 matrices (wifi_ldpc_tables.rs): f32 and int8,
    bit-exact in iterations, posteriors / app, R and hard decisions."""
    import syndrome_amd._lib as L
    from syndrome_amd import QcLdpcDecoder
    say(f"[wifi] Z={z} R={rn}/{rd} plain={plain}")
    code = oracle.code_wifi(z, rn, rd)
    dec = QcLdpcDecoder.wifi(z, rn, rd, 0.5)
    B = 5
    cws = codewords(oracle, code, B, seed=31 * z + rn)
    llr = channel(oracle, code, cws, [0.5, 1.5, 2.5, 3.5, 5.0], seed=7 * z + rd)
    post, hard, edge, it = _run_batch_f32(dec, llr, 10, init5g=False, plain=plain)
    for b in range(B):
        ri, rp, re, rh = oracle.decode_f32(code, 0.5, llr[b], 10)
        assert it[b] == ri and np.array_equal(hard[b], rh), f"f32 codeword {b}"
        assert np.array_equal(f32_bits(post[b]), f32_bits(rp)), f"f32 codeword {b}: posterior"
        assert np.array_equal(f32_bits(edge[b][: code.edge_buf]), f32_bits(re)), f"f32 codeword {b}: edge_r"
    app = np.stack([oracle.quantize_i16(x, 8.0) for x in llr])
    d_in = torch.from_numpy(app).cuda()
    d_post = torch.empty_like(d_in)
    d_hard = torch.empty(app.shape, dtype=torch.uint8, device="cuda")
    d_edge = torch.empty((B, dec.required_edge_buffer()), dtype=torch.int8, device="cuda")
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    dec.decode_batch_i8(d_in, d_post, d_hard, d_it, 10, flags=L.WRITE_POST | L.WRITE_HARD | L.WRITE_EDGE |
                        (L.PLAIN if plain else 0), d_edge=d_edge)
    torch.cuda.synchronize()
    pa, ha, ea, ia = d_post.cpu().numpy(), d_hard.cpu().numpy(), d_edge.cpu().numpy(), d_it.cpu().numpy()
    for b in range(B):
        ri, ra, re, rh = oracle.decode_i8(code, 0.5, app[b], 10)
        assert ia[b] == ri and np.array_equal(pa[b], ra) and np.array_equal(ha[b], rh), f"i8 codeword {b}"
        assert np.array_equal(ea[b][: code.edge_buf], re), f"i8 codeword {b}: edge_r"


def test_pipeline_codeword_five_iterations(dev, oracle):
    """ldpc_pipeline_bench: BG1 Z=384, Knuth info, AWGN 2.0 dB seed 42 -> exactly 5 iterations."""
    from syndrome_amd import QcLdpcDecoder
    say("[pipeline codeword] BG1 Z=384 2.0 dB seed 42")
    code = oracle.code_3gpp(0, 384)
    cw = oracle.encode(code, oracle.knuth_info(code.k))
    llr = oracle.awgn(2.0, code.k / code.n, 42, cw)
    dec = QcLdpcDecoder.with_lifting_size(0, 384, 0.25)
    post, hard, edge, it = _run_batch_f32(dec, llr[None, :], 10, init5g=False, plain=False)
    assert it[0] == 5
    assert np.array_equal(hard[0], cw)


I8_CASES = [(0, 384), (0, 7), (0, 44), (0, 96), (1, 208), (1, 13), (1, 52), (1, 128), (1, 2)]


@pytest.mark.parametrize("plain", [False, True], ids=["fast", "plain"])
@pytest.mark.parametrize("bg,z", I8_CASES)
def test_i8_batch_matches_oracle(dev, oracle, bg, z, plain):
    import syndrome_amd._lib as L
    from syndrome_amd import QcLdpcDecoder
    say(f"[i8 batch] BG{bg + 1} Z={z} plain={plain}")
    code = oracle.code_3gpp(bg, z)
    dec = QcLdpcDecoder.with_lifting_size(bg, z, 0.5)
    B = 6
    cws = codewords(oracle, code, B, seed=300 * z + bg)
    llr = channel(oracle, code, cws, [0.0, 1.5, 3.0, 0.7, 2.2, 4.0], seed=31 * z)
    iters = 8
    for clamp in (32767, 127):
        # quantize fused into the load, 5G prologue on
        B_, N = llr.shape
        d_in = torch.from_numpy(llr).cuda()
        d_post = torch.empty((B_, N), dtype=torch.int16, device="cuda")
        d_hard = torch.empty((B_, N), dtype=torch.uint8, device="cuda")
        d_edge = torch.empty((B_, dec.required_edge_buffer()), dtype=torch.int8, device="cuda")
        d_it = torch.empty(B_, dtype=torch.int32, device="cuda")
        from syndrome_amd import QuantParams
        flags = L.WRITE_POST | L.WRITE_HARD | L.WRITE_EDGE | L.INIT_5G | L.QUANTIZE | (L.PLAIN if plain else 0)
        dec.decode_batch_i8(d_in, d_post, d_hard, d_it, iters, flags=flags, d_edge=d_edge,
                            quant=QuantParams(8.0, clamp))
        torch.cuda.synchronize()
        post, hard, edge, it = (d_post.cpu().numpy(), d_hard.cpu().numpy(), d_edge.cpu().numpy(),
                                d_it.cpu().numpy())
        for b in range(B):
            app = oracle.init_5g_i8(code, oracle.quantize_i16(llr[b], 8.0), 0, clamp)
            ri, ra, re, rh = oracle.decode_i8(code, 0.5, app, iters, 8.0, clamp)
            assert it[b] == ri, f"clamp {clamp} codeword {b}: iterations {it[b]} != {ri}"
            assert np.array_equal(post[b], ra), f"clamp {clamp} codeword {b}: app differs"
            assert np.array_equal(edge[b], re), f"clamp {clamp} codeword {b}: edge_r differs"
            assert np.array_equal(hard[b], rh), f"clamp {clamp} codeword {b}: hard differs"


@pytest.mark.parametrize("plain", [False, True], ids=["fast", "plain"])
@pytest.mark.parametrize("bg,z", [(0, 384), (1, 208), (0, 20)])
def test_i8_saturated_rows_match_oracle(dev, oracle, bg, z, plain):
    """Inputs far beyond the message range: every |q| of most rows exceeds
    127, so min1 = min2 = 127 there and the clamp of q (qc_ldpc.rs Q-build)
    is what the kernels' unclamped form must reproduce; a few flipped signs
    keep the decode going for several iterations."""
    import syndrome_amd._lib as L
    from syndrome_amd import QcLdpcDecoder, QuantParams
    say(f"[i8 saturated] BG{bg + 1} Z={z} plain={plain}")
    code = oracle.code_3gpp(bg, z)
    dec = QcLdpcDecoder.with_lifting_size(bg, z, 0.5)
    B = 4
    cws = codewords(oracle, code, B, seed=77 * z + bg)
    rng = np.random.default_rng(5 * z + bg)
    mag = rng.integers(160, 480, size=cws.shape)  # int16 app, beyond the message range
    app_in = np.where(cws == 0, mag, -mag).astype(np.int16)
    for b in range(B):
        flip = rng.choice(code.n, size=(b + 1) * 3, replace=False)
        app_in[b, flip] = -(app_in[b, flip] // (2 if b % 2 else 8))  # strong or weak wrong signs
    iters = 6
    for clamp in (32767, 400):
        d_in = torch.from_numpy(app_in).cuda()
        d_post = torch.empty((B, code.n), dtype=torch.int16, device="cuda")
        d_hard = torch.empty((B, code.n), dtype=torch.uint8, device="cuda")
        d_edge = torch.empty((B, dec.required_edge_buffer()), dtype=torch.int8, device="cuda")
        d_it = torch.empty(B, dtype=torch.int32, device="cuda")
        flags = L.WRITE_POST | L.WRITE_HARD | L.WRITE_EDGE | (L.PLAIN if plain else 0)
        dec.decode_batch_i8(d_in, d_post, d_hard, d_it, iters, flags=flags, d_edge=d_edge,
                            quant=QuantParams(8.0, clamp))
        torch.cuda.synchronize()
        post, hard, edge, it = (d_post.cpu().numpy(), d_hard.cpu().numpy(), d_edge.cpu().numpy(),
                                d_it.cpu().numpy())
        for b in range(B):
            app = app_in[b]
            ri, ra, re, rh = oracle.decode_i8(code, 0.5, app, iters, 8.0, clamp)
            assert it[b] == ri, f"clamp {clamp} codeword {b}: iterations {it[b]} != {ri}"
            assert np.array_equal(post[b], ra), f"clamp {clamp} codeword {b}: app differs"
            assert np.array_equal(edge[b], re), f"clamp {clamp} codeword {b}: edge_r differs"
            assert np.array_equal(hard[b], rh), f"clamp {clamp} codeword {b}: hard differs"


def test_capi_decode_5g_host_buffers(dev, oracle):
    """The reference-shaped entry point (capi/src/lib.rs:290-343) on host buffers."""
    from syndrome_amd import QcLdpcDecoder
    say("[capi decode_5g] BG1 Z=384 / BG2 Z=52 with filler")
    for bg, z, nf in ((0, 384, 0), (1, 52, 37)):
        code = oracle.code_3gpp(bg, z)
        dec = QcLdpcDecoder.with_lifting_size(bg, z, 0.5)
        cw = codewords(oracle, code, 1, seed=5)[0]
        cw[code.k - nf: code.k] = 0  # filler bits are known zeros...
        cw = oracle.encode(code, cw[: code.k])  # ...re-encode so the codeword is valid
        llr = channel(oracle, code, cw[None, :], [1.2], seed=99)[0]
        a = llr.copy()
        er = np.zeros(dec.required_edge_buffer(), np.float32)
        sc = np.zeros(dec.required_layer_buffer(), np.float32)
        hd = np.zeros(dec.variable_node_count(), np.uint8)
        used = dec.decode_5g(a, nf, er, sc, hd, 10)
        ri, rp, re, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, llr, nf), 10)
        assert used == ri
        assert np.array_equal(f32_bits(a), f32_bits(rp))
        assert np.array_equal(f32_bits(er), f32_bits(re))
        assert np.array_equal(hd, rh)


def test_capi_i8_host_buffers(dev, oracle):
    from syndrome_amd import QcLdpcDecoder, QuantParams
    say("[capi decode_i8 / decode_5g_i8] BG2 Z=208")
    code = oracle.code_3gpp(1, 208)
    dec = QcLdpcDecoder.with_lifting_size(1, 208, 0.5)
    cw = codewords(oracle, code, 1, seed=9)[0]
    llr = channel(oracle, code, cw[None, :], [1.0], seed=3)[0]
    app0 = oracle.quantize_i16(llr, 8.0)
    for five_g in (False, True):
        app = app0.copy()
        er = np.zeros(dec.required_edge_buffer(), np.int8)
        sc = np.zeros(dec.required_layer_buffer(), np.int8)
        hd = np.zeros(dec.variable_node_count(), np.uint8)
        if five_g:
            used = dec.decode_5g_i8(app, 0, er, sc, hd, 8, QuantParams())
            ri, ra, re, rh = oracle.decode_i8(code, 0.5, oracle.init_5g_i8(code, app0, 0), 8)
        else:
            used = dec.decode_layered_offset_min_sum_i8(app, er, sc, hd, 8, QuantParams())
            ri, ra, re, rh = oracle.decode_i8(code, 0.5, app0, 8)
        assert used == ri
        assert np.array_equal(app, ra) and np.array_equal(er, re) and np.array_equal(hd, rh)


@pytest.mark.parametrize("bg,z", [(1, 208), (0, 384), (1, 52), (0, 26)])
def test_i8_degree1_bypass_and_its_fallback(dev, oracle, bg, z):
    """The int8 kernels' degree-1 bypass (a degree-1 variable's posterior
    kept as channel value + R) holds only while every degree-1 |c| of a
    workgroup is at most C - 127; a workgroup with one value past it runs
    the full update.  One batch mixes both kinds of workgroup, at an app
    clamp of 300 (bound 173), with values up to the clamp on degree-1
    columns of every other codeword; every codeword equals the oracle."""
    import syndrome_amd._lib as L
    from syndrome_amd import QcLdpcDecoder, QuantParams
    code = oracle.code_3gpp(bg, z)
    dec = QcLdpcDecoder.with_lifting_size(bg, z, 0.5)
    cws_per_group, _, _ = dec.launch_shape(int8=True)
    B = 4 * cws_per_group
    say(f"[i8 degree-1] BG{bg + 1} Z={z}: {B} codewords, {cws_per_group} per workgroup")
    cw = codewords(oracle, code, B, seed=91 * z + bg)
    rng = np.random.default_rng(7 * z + bg)
    mag = rng.integers(1, 120, size=cw.shape)
    app_in = np.where(cw == 0, mag, -mag).astype(np.int16)
    first_deg1 = (26 if bg == 0 else 14) * z  # degree-1 columns: BG1 26..67, BG2 14..51
    for g in range(1, 4, 2):  # workgroups 1 and 3: one degree-1 value past the bound each codeword
        for b in range(g * cws_per_group, (g + 1) * cws_per_group):
            v = first_deg1 + int(rng.integers(0, code.n - first_deg1))
            app_in[b, v] = np.int16(250 if cw[b, v] == 0 else -250)
    flip = rng.choice(code.n, size=(B, 6))
    for b in range(B):
        app_in[b, flip[b]] = -app_in[b, flip[b]] // 2
    for clamp in (300, 32767):
        d_in = torch.from_numpy(app_in).cuda()
        d_post = torch.empty((B, code.n), dtype=torch.int16, device="cuda")
        d_hard = torch.empty((B, code.n), dtype=torch.uint8, device="cuda")
        d_edge = torch.empty((B, dec.required_edge_buffer()), dtype=torch.int8, device="cuda")
        d_it = torch.empty(B, dtype=torch.int32, device="cuda")
        flags = L.WRITE_POST | L.WRITE_HARD | L.WRITE_EDGE
        dec.decode_batch_i8(d_in, d_post, d_hard, d_it, 8, flags=flags, d_edge=d_edge, quant=QuantParams(8.0, clamp))
        torch.cuda.synchronize()
        post, hard, edge, it = (d_post.cpu().numpy(), d_hard.cpu().numpy(), d_edge.cpu().numpy(),
                                d_it.cpu().numpy())
        for b in range(B):
            ri, ra, re, rh = oracle.decode_i8(code, 0.5, app_in[b], 8, 8.0, clamp)
            assert it[b] == ri, f"clamp {clamp} codeword {b}: iterations {it[b]} != {ri}"
            assert np.array_equal(post[b], ra), f"clamp {clamp} codeword {b}: app differs"
            assert np.array_equal(edge[b], re), f"clamp {clamp} codeword {b}: edge_r differs"
            assert np.array_equal(hard[b], rh), f"clamp {clamp} codeword {b}: hard differs"
