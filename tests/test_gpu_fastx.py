"""The BG1 f32 kernel that keeps the 42 extension columns' posteriors in the
posterior buffer (HBM) instead of LDS, two codewords per workgroup
(syndrome_amd/csrc/ldpc_fastx.h).  It runs whenever a BG1 f32 batch of
Z = 384, 352, 320 or 288 writes its posterior back; a batch that writes only
hard decisions runs the all-LDS kernel, so the two can be set against each
other on the same input.  At 352 and 288 the last wave of each codeword
has idle lanes (the lifting size is not a whole number of waves).

Checked against the oracle (qc_ldpc.rs:962-1252 restated) bit for bit:
posteriors, final messages (edge_r), hard decisions, iteration counts, over
* odd batches (the last workgroup's second codeword does not exist),
* codewords stopping at different iterations inside one workgroup,
* in-place decodes (posterior buffer = input buffer) and inputs that are
  not 16-byte aligned (the element-wise load and store path),
* -0.0, NaN and infinities in the extension columns, whose hard decision
  and syndrome bit are `llr < 0` (so -0.0 and NaN read as 0),
* one iteration, and the cap reached without convergence."""
import numpy as np
import pytest

from tests.helpers import f32_bits, say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

Z = 384


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    import syndrome_amd
    syndrome_amd.lib()
    return torch.device("cuda:0")


def _decoder(beta, z=Z):
    import syndrome_amd
    return syndrome_amd.QcLdpcDecoder.with_lifting_size(syndrome_amd.BG1, z, beta)


def _channel(dec, ebnos, seed, z=Z):
    """5G-style inputs: encoded random info bits, BPSK/AWGN, first 2Z zero."""
    from syndrome_amd import awgn_transmit
    Z = z
    n, k = dec.variable_node_count(), dec.info_bit_count_5g()
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, size=(len(ebnos), k), dtype=np.uint8)
    cws = dec.encode_batch(info, threads=4)
    llr = np.zeros((len(ebnos), n), np.float32)
    for i, e in enumerate(ebnos):
        llr[i, 2 * Z:] = awgn_transmit(e, k / (n - 2 * Z), seed * 1000 + i, cws[i, 2 * Z:])
    return llr


def _run(dec, d_in, d_post, iters, flags, edge=True):
    from syndrome_amd import _lib as L
    B, n = d_in.shape
    d_hard = torch.empty((B, n), dtype=torch.uint8, device="cuda")
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    d_edge = torch.empty((B, dec.required_edge_buffer()), dtype=torch.float32, device="cuda") if edge else None
    dec.decode_batch_f32(d_in, d_post, d_hard, d_it, iters, flags=flags | (L.WRITE_EDGE if edge else 0),
                         d_edge=d_edge)
    torch.cuda.synchronize()
    return (d_post.cpu().numpy() if d_post is not None else None, d_hard.cpu().numpy(), d_it.cpu().numpy(),
            d_edge.cpu().numpy() if edge else None)


def _check_oracle(oracle, beta, llr_in, post, hard, its, edge, iters, init5g=True, label="", z=Z):
    code = oracle.code_3gpp(0, z)
    for b in range(llr_in.shape[0]):
        x = oracle.init_5g_f32(code, llr_in[b], 0) if init5g else llr_in[b]
        ri, rp, re, rh = oracle.decode_f32(code, beta, x, iters)
        assert its[b] == ri, f"{label} codeword {b}: iterations {its[b]} != {ri}"
        assert np.array_equal(hard[b], rh), f"{label} codeword {b}: hard decisions differ"
        if post is not None:
            assert np.array_equal(f32_bits(post[b]), f32_bits(rp)), f"{label} codeword {b}: posterior differs"
        if edge is not None:
            assert np.array_equal(f32_bits(edge[b][:re.size]), f32_bits(re)), f"{label} codeword {b}: edge_r differs"


def test_fastx_odd_batch_mixed_convergence_matches_oracle(dev, oracle):
    from syndrome_amd import _lib as L
    dec = _decoder(0.5)
    ebnos = [1.5, 0.4, 3.0, 1.0, 1.5, 2.0, 0.8]  # 7 codewords: four workgroups, the last half empty
    llr = _channel(dec, ebnos, 11)
    say("[fastx] B=7, iterations stop at different passes inside a workgroup")
    d_in = torch.from_numpy(llr).cuda()
    post, hard, its, edge = _run(dec, d_in, torch.empty_like(d_in), 8, L.INIT_5G | L.WRITE_POST | L.WRITE_HARD)
    assert len(set(its.tolist())) > 1, "the batch should stop at different iterations"
    _check_oracle(oracle, 0.5, llr, post, hard, its, edge, 8, label="odd batch")


def test_fastx_equals_all_lds_kernel(dev):
    from syndrome_amd import _lib as L
    dec = _decoder(0.5)
    llr = _channel(dec, [1.2, 1.5, 1.8, 0.9, 2.5, 1.5], 12)
    d_in = torch.from_numpy(llr).cuda()
    say("[fastx] same batch with and without the posterior written back")
    _, h1, i1, e1 = _run(dec, d_in, torch.empty_like(d_in), 8, L.INIT_5G | L.WRITE_POST | L.WRITE_HARD)
    _, h2, i2, e2 = _run(dec, d_in, None, 8, L.INIT_5G | L.WRITE_HARD)
    assert np.array_equal(i1, i2)
    assert np.array_equal(h1, h2)
    assert np.array_equal(f32_bits(e1), f32_bits(e2))


def test_fastx_in_place_and_unaligned(dev, oracle):
    from syndrome_amd import _lib as L
    dec = _decoder(0.5)
    llr = _channel(dec, [1.5, 1.1, 2.2], 13)
    B, n = llr.shape
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD
    say("[fastx] in-place decode")
    d_io = torch.from_numpy(llr.copy()).cuda()
    post, hard, its, edge = _run(dec, d_io, d_io, 8, flags)
    _check_oracle(oracle, 0.5, llr, post, hard, its, edge, 8, label="in place")
    say("[fastx] input and posterior 4 bytes off 16-byte alignment")
    buf_in = torch.zeros(B * n + 1, dtype=torch.float32, device="cuda")
    buf_in[1:] = torch.from_numpy(llr.reshape(-1)).cuda()
    buf_post = torch.zeros(B * n + 1, dtype=torch.float32, device="cuda")
    d_in, d_post = buf_in[1:].view(B, n), buf_post[1:].view(B, n)
    assert d_in.data_ptr() % 16 != 0
    post, hard, its, edge = _run(dec, d_in, d_post, 8, flags)
    _check_oracle(oracle, 0.5, llr, post, hard, its, edge, 8, label="unaligned")


def test_fastx_special_values_in_extension_columns(dev, oracle):
    from syndrome_amd import _lib as L
    dec = _decoder(0.25)
    n = dec.variable_node_count()
    rng = np.random.default_rng(14)
    llr = rng.normal(1.0, 2.0, size=(4, n)).astype(np.float32)
    ext = np.arange(26 * Z, n)
    special = np.array([-0.0, np.nan, np.inf, -np.inf, 0.0], np.float32)
    for b in range(4):
        pick = rng.choice(ext, 40, replace=False)
        llr[b, pick] = special[rng.integers(0, special.size, 40)]
    llr[1, ext] = -0.0  # every extension variable -0.0 in one codeword
    say("[fastx] -0.0, NaN and infinities in the extension columns (no 5G prologue)")
    d_in = torch.from_numpy(llr).cuda()
    post, hard, its, edge = _run(dec, d_in, torch.empty_like(d_in), 6, L.WRITE_POST | L.WRITE_HARD)
    _check_oracle(oracle, 0.25, llr, post, hard, its, edge, 6, init5g=False, label="special values")


def test_fastx_one_iteration_and_cap(dev, oracle):
    from syndrome_amd import _lib as L
    dec = _decoder(0.25)
    n = dec.variable_node_count()
    alt = np.where(np.arange(n) % 2 == 0, 0.5, -0.5).astype(np.float32)  # never converges
    llr = np.stack([alt, _channel(dec, [1.5], 15)[0], alt])
    d_in = torch.from_numpy(llr).cuda()
    for iters in (1, 10):
        say(f"[fastx] {iters} iteration(s)")
        post, hard, its, edge = _run(dec, d_in, torch.empty_like(d_in), iters, L.WRITE_POST | L.WRITE_HARD)
        _check_oracle(oracle, 0.25, llr, post, hard, its, edge, iters, init5g=False, label=f"{iters} it")
    assert oracle.fnv1a(hard[0]) == 0xA2448E8CF5272825  # the reference's golden checksum


@pytest.mark.parametrize("codes", [
    [(0, 384), (1, 208), (0, 384), (0, 52), (0, 384), (1, 384)],
    [(0, 352), (0, 320), (1, 208), (0, 288), (0, 384), (0, 352), (0, 288), (1, 384), (0, 320), (0, 96)],
], ids=["z384", "mixed_sizes"])
def test_plan_extension_launch_with_and_without_posterior(dev, oracle, codes):
    """A work-queue plan decodes its f32 BG1 blocks of Z > 256 in a launch
    of the two-codeword kernel when the run writes the posterior back (the
    Z=384-only kernel, or the kernel of all four sizes when the plan has
    others), and in the main launch, one codeword per workgroup, when it
    does not: the same bits both ways, and each block as the oracle decodes
    it alone (fillers included).  Odd counts per size leave items that hold
    one codeword."""
    import syndrome_amd._lib as L
    from syndrome_amd import DecodePlan
    rng = np.random.default_rng(21)
    bg, z, nf, blocks = [], [], [], []
    for i, (b, zz) in enumerate(codes):
        code = oracle.code_3gpp(b, zz)
        f = int(rng.integers(0, 40))
        info = oracle.xorshift_bits(3100 + i, code.k)
        info[code.k - f:] = 0
        blocks.append(oracle.awgn(float(rng.uniform(0.5, 2.5)), code.k / code.n, 900 + i, oracle.encode(code, info)))
        bg.append(b)
        z.append(zz)
        nf.append(f)
    offs = np.concatenate([[0], np.cumsum([x.size for x in blocks])[:-1]]).astype(np.int64)
    arena = np.concatenate(blocks).astype(np.float32)
    plan = DecodePlan(np.array(bg), np.array(z), np.array(nf), offs, 0.5)
    say("[fastx plan] shape", plan.shape())
    d_in = torch.from_numpy(arena).cuda()
    out = []
    for flags in (L.INIT_5G | L.WRITE_POST | L.WRITE_HARD, L.INIT_5G | L.WRITE_HARD):
        d_post = torch.zeros_like(d_in)
        d_hard = torch.empty(arena.size, dtype=torch.uint8, device="cuda")
        d_it = torch.empty(len(bg), dtype=torch.int32, device="cuda")
        plan.run(d_in, d_post, d_hard, d_it, 8, flags)
        torch.cuda.synchronize()
        out.append((d_post.cpu().numpy(), d_hard.cpu().numpy(), d_it.cpu().numpy()))
    (post, hard, its), (_, hard2, its2) = out
    assert np.array_equal(its, its2) and np.array_equal(hard, hard2)
    for i in range(len(bg)):
        code = oracle.code_3gpp(bg[i], z[i])
        ri, rp, _, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, blocks[i], nf[i]), 8)
        o, n = offs[i], code.n
        assert its[i] == ri, f"block {i}: iterations {its[i]} != {ri}"
        assert np.array_equal(hard[o:o + n], rh), f"block {i}: hard differs"
        assert np.array_equal(f32_bits(post[o:o + n]), f32_bits(rp)), f"block {i}: posterior differs"


@pytest.mark.parametrize("z", [352, 320, 288])
def test_fastx_other_sizes_match_oracle_and_all_lds_kernel(dev, oracle, z):
    from syndrome_amd import _lib as L
    dec = _decoder(0.5, z)
    llr = _channel(dec, [1.5, 0.4, 3.0, 1.0, 1.8], 30 + z, z)  # odd: the last workgroup half empty
    say(f"[fastx] BG1 Z={z}, B=5")
    d_in = torch.from_numpy(llr).cuda()
    post, hard, its, edge = _run(dec, d_in, torch.empty_like(d_in), 8, L.INIT_5G | L.WRITE_POST | L.WRITE_HARD)
    _check_oracle(oracle, 0.5, llr, post, hard, its, edge, 8, label=f"Z={z}", z=z)
    _, h2, i2, e2 = _run(dec, d_in, None, 8, L.INIT_5G | L.WRITE_HARD)
    assert np.array_equal(its, i2) and np.array_equal(hard, h2)
    assert np.array_equal(f32_bits(edge), f32_bits(e2))
