"""HIP-graph capture of the batched decode: the rate de-matcher, the
work-queue reset and the persistent mixed-code decode of a DecodePlan, and
the per-config batch decode, recorded once in a graph and replayed.  Every
replay, with eager work and synchronisations between replays, must equal
the eager result bit for bit (iterations, posterior, hard decisions), and
the eager result must equal the oracle's decode of each block alone.
"""
import numpy as np
import pytest

from tests.helpers import f32_bits, say
from tests.test_gpu_batch import MIX

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _traffic(count, seed):
    from syndrome_amd import traffic_blocks
    rng = np.random.default_rng(seed)
    bg = np.array([MIX[i % len(MIX)][0] for i in range(count)])
    z = np.array([MIX[i % len(MIX)][1] for i in range(count)])
    k = np.where(bg == 0, 22, 10) * z
    nf = np.array([int(rng.integers(0, min(40, kk // 4) + 1)) for kk in k])
    rate = rng.choice([1 / 3, 1 / 2, 2 / 3, 5 / 6], count)
    e_len = np.maximum(np.round((k - nf) / rate).astype(np.int64), 8)
    e_llr, e_off, _, _ = traffic_blocks(bg, z, nf, e_len, rng.uniform(0.5, 3.0, count), (k - nf) / e_len, seed,
                                        want_info=False)
    return bg, z, nf, e_len, e_llr, e_off


def test_plan_dematch_and_decode_replay_in_a_graph(oracle):
    import syndrome_amd
    from syndrome_amd import DecodePlan, _lib as L
    assert torch.cuda.is_available()
    syndrome_amd.lib()
    say("[graph] de-match + persistent decode of 28 mixed blocks captured, 4 replays")
    bg, z, nf, e_len, e_llr, e_off = _traffic(28, 21)
    n = np.where(bg == 0, 68, 52) * z
    offs = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    plan = DecodePlan(bg, z, nf, offs, 0.5)
    plan.set_dematch(e_off, e_len)
    d_e = torch.from_numpy(e_llr).cuda()
    d_llr = torch.empty(int(n.sum()), dtype=torch.float32, device="cuda")
    d_post = torch.empty_like(d_llr)
    d_hard = torch.empty(int(n.sum()), dtype=torch.uint8, device="cuda")
    d_it = torch.empty(len(bg), dtype=torch.int32, device="cuda")
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.dematch(d_e, d_llr, stream=s.cuda_stream)
        plan.run(d_llr, d_post, d_hard, d_it, 8, flags, stream=s.cuda_stream)
    torch.cuda.synchronize()
    ref = (d_llr.cpu().numpy(), d_post.cpu().numpy(), d_hard.cpu().numpy(), d_it.cpu().numpy())
    # the eager result is the oracle's, block by block
    for i in range(len(bg)):
        code = oracle.code_3gpp(int(bg[i]), int(z[i]))
        cb = oracle.rate_dematch(e_llr[e_off[i]:e_off[i] + e_len[i]], 0, 1, int(bg[i]), int(z[i]), int(nf[i]),
                                 cached=True)
        inp = oracle.decoder_input_from_cb(cb, int(z[i]))
        assert np.array_equal(f32_bits(ref[0][offs[i]:offs[i] + n[i]]), f32_bits(inp)), f"block {i}: de-match"
        ri, rp, _, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, inp, int(nf[i])), 8)
        assert ref[3][i] == ri and np.array_equal(ref[2][offs[i]:offs[i] + n[i]], rh), f"block {i}"
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.dematch(d_e, d_llr, stream=s.cuda_stream)
        plan.run(d_llr, d_post, d_hard, d_it, 8, flags, stream=s.cuda_stream)
    for rep in range(4):
        for t in (d_llr, d_post):
            t.fill_(float("nan"))
        d_hard.fill_(7)
        d_it.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = (d_llr.cpu().numpy(), d_post.cpu().numpy(), d_hard.cpu().numpy(), d_it.cpu().numpy())
        assert np.array_equal(f32_bits(got[0]), f32_bits(ref[0])), f"replay {rep}: de-matched input"
        assert np.array_equal(f32_bits(got[1]), f32_bits(ref[1])), f"replay {rep}: posterior"
        assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3]), f"replay {rep}"
        # eager work on the default stream between replays
        plan.run(d_llr, d_post, d_hard, d_it, 8, flags)
        torch.cuda.synchronize()
        assert np.array_equal(d_it.cpu().numpy(), ref[3])


def test_batch_decode_replay_in_a_graph():
    import syndrome_amd
    from syndrome_amd import QcLdpcDecoder, _lib as L
    syndrome_amd.lib()
    say("[graph] fast_f32 BG1 Z=384 batch of 64 captured, 3 replays")
    rng = np.random.default_rng(4)
    dec = QcLdpcDecoder.with_lifting_size(0, 384, 0.5)
    n = dec.variable_node_count()
    x = torch.from_numpy(rng.normal(0.4, 1.0, (64, n)).astype(np.float32)).cuda()
    post, hard = torch.empty_like(x), torch.empty(x.shape, dtype=torch.uint8, device="cuda")
    it = torch.empty(64, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD
    with torch.cuda.stream(s):
        dec.decode_batch_f32(x, post, hard, it, 8, flags=flags, stream=s.cuda_stream)
    torch.cuda.synchronize()
    ref = (post.cpu().numpy(), hard.cpu().numpy(), it.cpu().numpy())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        dec.decode_batch_f32(x, post, hard, it, 8, flags=flags, stream=s.cuda_stream)
    for rep in range(3):
        post.zero_()
        hard.zero_()
        it.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(f32_bits(post.cpu().numpy()), f32_bits(ref[0])), f"replay {rep}"
        assert np.array_equal(hard.cpu().numpy(), ref[1]) and np.array_equal(it.cpu().numpy(), ref[2])
