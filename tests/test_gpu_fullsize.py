"""Parity at the benchmark sizes: one launch over the whole batch, as
bench.py times it, checked three ways.
* 64 sampled codewords (the first, the last and 62 seeded ones) against the
  oracle bit for bit: iteration counts, posteriors, hard decisions.
* Every codeword: the work-queue (persistent) launch of a DecodePlan over
  the same batch returns the same bits as the per-code batch launch, two
  different kernels; every codeword that stopped before the iteration cap
  has a zero syndrome (the early exit is the reference's, qc_ldpc.rs:1211 /
  1748); hard decisions are the signs of the posteriors.
* The decoded information bits of the converged codewords are the
  transmitted ones.
Workloads: BG1 Z=384 rate 1/3, 1.5 dB, f32, max 8 iterations, 4,096
codewords; BG2 Z=208 rate 1/5, Eb/N0 uniform in [0, 3] dB, int8 with the
quantizer fused into the load, max 8 iterations, 16,384 codewords."""
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


def _sample(B, seed):
    rng = np.random.default_rng(seed)
    return np.unique(np.concatenate([[0, B - 1], rng.choice(np.arange(1, B - 1), 62, replace=False)]))


def _plan_run(bg, z, B, n, int8, d_in, flags, quant=None):
    from syndrome_amd import DecodePlan
    plan = DecodePlan(np.full(B, bg), np.full(B, z), np.zeros(B), np.arange(B, dtype=np.int64) * n, 0.5, int8=int8)
    d_post = torch.empty(B * n, dtype=torch.int16 if int8 else torch.float32, device="cuda")
    d_hard = torch.empty(B * n, dtype=torch.uint8, device="cuda")
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    plan.run(d_in.view(-1), d_post, d_hard, d_it, 8, flags, quant=quant)
    torch.cuda.synchronize()
    return d_post.view(B, n), d_hard.view(B, n), d_it


def test_bg1_z384_batch_4096_f32(dev, oracle):
    import syndrome_amd
    from syndrome_amd import QcLdpcDecoder, awgn_transmit, _lib as L
    B, z = 4096, 384
    dec = QcLdpcDecoder.with_lifting_size(syndrome_amd.BG1, z, 0.5)
    n, k = dec.variable_node_count(), dec.info_bit_count_5g()
    say(f"[full size f32] BG1 Z=384, B={B}: encoding")
    rng = np.random.default_rng(41)
    info = rng.integers(0, 2, size=(B, k), dtype=np.uint8)
    cws = dec.encode_batch(info, threads=16)
    llr = np.zeros((B, n), np.float32)
    for i in range(B):
        llr[i, 2 * z:] = awgn_transmit(1.5, k / (n - 2 * z), 7000 + i, cws[i, 2 * z:])
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD
    d_in = torch.from_numpy(llr).cuda()
    d_post = torch.empty_like(d_in)
    d_hard = torch.empty((B, n), dtype=torch.uint8, device="cuda")
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    say("[full size f32] one launch over the batch")
    dec.decode_batch_f32(d_in, d_post, d_hard, d_it, 8, flags=flags)
    torch.cuda.synchronize()
    post, hard, its = d_post.cpu().numpy(), d_hard.cpu().numpy(), d_it.cpu().numpy()
    code = oracle.code_3gpp(0, z)
    for b in _sample(B, 5):
        ri, rp, _, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, llr[b], 0), 8)
        assert its[b] == ri, f"codeword {b}: iterations {its[b]} != {ri}"
        assert np.array_equal(f32_bits(post[b]), f32_bits(rp)), f"codeword {b}: posterior differs"
        assert np.array_equal(hard[b], rh), f"codeword {b}: hard decisions differ"
    assert ((its >= 1) & (its <= 8)).all()
    assert np.array_equal(hard, (post < 0).astype(np.uint8))
    early = np.nonzero(its < 8)[0]
    for b in early:
        assert oracle.syndrome_ok(code, hard[b]), f"codeword {b} stopped at {its[b]} with a nonzero syndrome"
    ok = (hard[:, :k] == info).all(axis=1)
    assert ok[early].all()
    say(f"[full size f32] {len(early)} of {B} stopped early, {int(ok.sum())} decoded; persistent launch")
    pp, ph, pi = _plan_run(0, z, B, n, False, d_in, flags)
    assert torch.equal(pi.cpu(), torch.from_numpy(its))
    assert torch.equal(pp.view(torch.int32), d_post.view(torch.int32)) and torch.equal(ph, d_hard)


def test_bg2_z208_batch_16384_int8(dev, oracle):
    import syndrome_amd
    from syndrome_amd import QcLdpcDecoder, QuantParams, traffic_blocks, _lib as L
    B, z = 16384, 208
    dec = QcLdpcDecoder.with_lifting_size(syndrome_amd.BG2, z, 0.5)
    n, k = dec.variable_node_count(), dec.info_bit_count_5g()
    say(f"[full size int8] BG2 Z=208, B={B}: traffic")
    rng = np.random.default_rng(43)
    ebno = rng.uniform(0, 3, B)
    e_llr, _, info, _ = traffic_blocks(np.ones(B), np.full(B, z), np.zeros(B), np.full(B, 50 * z), ebno,
                                       np.full(B, k / (50 * z)), 4343)
    llr = np.zeros((B, n), np.float32)
    llr[:, 2 * z:] = e_llr.reshape(B, 50 * z)  # E = Ncb at rv 0, no filler: identity placement
    info = info.reshape(B, k)
    q = QuantParams(8.0, 32767)
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD | L.QUANTIZE
    d_in = torch.from_numpy(llr).cuda()
    d_post = torch.empty((B, n), dtype=torch.int16, device="cuda")
    d_hard = torch.empty((B, n), dtype=torch.uint8, device="cuda")
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    say("[full size int8] one launch over the batch")
    dec.decode_batch_i8(d_in, d_post, d_hard, d_it, 8, flags=flags, quant=q)
    torch.cuda.synchronize()
    post, hard, its = d_post.cpu().numpy(), d_hard.cpu().numpy(), d_it.cpu().numpy()
    code = oracle.code_3gpp(1, z)
    for b in _sample(B, 6):
        app = oracle.init_5g_i8(code, oracle.quantize_i16(llr[b], 8.0), 0, 32767)
        ri, ra, _, rh = oracle.decode_i8(code, 0.5, app, 8)
        assert its[b] == ri, f"codeword {b}: iterations {its[b]} != {ri}"
        assert np.array_equal(post[b], ra), f"codeword {b}: app differs"
        assert np.array_equal(hard[b], rh), f"codeword {b}: hard decisions differ"
    assert ((its >= 1) & (its <= 8)).all()
    assert np.array_equal(hard, (post < 0).astype(np.uint8))
    early = np.nonzero(its < 8)[0]
    for b in early[:: max(1, len(early) // 2000)]:
        assert oracle.syndrome_ok(code, hard[b]), f"codeword {b} stopped at {its[b]} with a nonzero syndrome"
    ok = (hard[:, :k] == info).all(axis=1)
    assert ok[early].mean() > 0.999
    say(f"[full size int8] {len(early)} of {B} stopped early, {int(ok.sum())} decoded; persistent launch")
    pp, ph, pi = _plan_run(1, z, B, n, True, d_in, flags, quant=q)
    assert torch.equal(pi.cpu(), torch.from_numpy(its))
    assert torch.equal(pp, d_post) and torch.equal(ph, d_hard)
