"""GPU parity of the work-queue (persistent) decoder over mixed codes, its
launch-shape tuner, and the device rate de-matcher.

Every code block in a mixed batch must come out exactly as the oracle
decodes it alone (decode_5g with its own filler count): iterations,
posterior and hard decisions bit for bit; and exactly as the per-config
batch path decodes it alone.
"""
import numpy as np
import pytest

from tests.helpers import f32_bits, say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MIX = [(0, 384), (1, 208), (0, 2), (1, 3), (0, 44), (1, 52), (0, 104), (1, 384), (0, 15), (1, 120), (0, 256),
       (1, 7), (0, 352), (1, 26)]


def _mixed_inputs(oracle, count, seed, with_filler=True):
    rng = np.random.default_rng(seed)
    bg, z, nf, blocks = [], [], [], []
    for i in range(count):
        b, zz = MIX[i % len(MIX)]
        code = oracle.code_3gpp(b, zz)
        f = int(rng.integers(0, min(64, code.k // 4) + 1)) if with_filler else 0
        info = oracle.xorshift_bits(seed * 1000 + i, code.k)
        info[code.k - f:] = 0  # fillers are known zeros
        cw = oracle.encode(code, info)
        ebno = float(rng.uniform(-0.5, 3.0))
        llr = oracle.awgn(ebno, code.k / code.n, seed + 77 * i, cw)
        bg.append(b)
        z.append(zz)
        nf.append(f)
        blocks.append(llr)
    offs = np.concatenate([[0], np.cumsum([x.size for x in blocks])[:-1]]).astype(np.int64)
    return np.array(bg), np.array(z), np.array(nf), offs, np.concatenate(blocks).astype(np.float32), blocks


def test_plan_f32_mixed_matches_oracle_and_alone(dev_ok, oracle):
    import syndrome_amd._lib as L
    from syndrome_amd import DecodePlan, QcLdpcDecoder
    say("[plan f32] 28 mixed code blocks")
    bg, z, nf, offs, arena, blocks = _mixed_inputs(oracle, 28, 5)
    plan = DecodePlan(bg, z, nf, offs, 0.5)
    say("  shape", plan.shape())
    d_in = torch.from_numpy(arena).cuda()
    d_post = torch.empty_like(d_in)
    d_hard = torch.empty(arena.size, dtype=torch.uint8, device="cuda")
    d_it = torch.empty(len(bg), dtype=torch.int32, device="cuda")
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD
    plan.run(d_in, d_post, d_hard, d_it, 8, flags)
    torch.cuda.synchronize()
    post, hard, its = d_post.cpu().numpy(), d_hard.cpu().numpy(), d_it.cpu().numpy()
    for i in range(len(bg)):
        code = oracle.code_3gpp(int(bg[i]), int(z[i]))
        ri, rp, _, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, blocks[i], int(nf[i])), 8)
        o, n = offs[i], code.n
        assert its[i] == ri, f"block {i} (BG{bg[i] + 1} Z={z[i]}): iterations {its[i]} != {ri}"
        assert np.array_equal(hard[o:o + n], rh), f"block {i}: hard differs"
        assert np.array_equal(f32_bits(post[o:o + n]), f32_bits(rp)), f"block {i}: posterior differs"
    # alone, through the per-config batch path: identical
    for i in (0, 1, 5, 9):
        dec = QcLdpcDecoder.with_lifting_size(int(bg[i]), int(z[i]), 0.5)
        x = torch.from_numpy(blocks[i]).cuda()[None, :]
        p1, h1, i1 = torch.empty_like(x), torch.empty(x.shape, dtype=torch.uint8, device="cuda"), \
            torch.empty(1, dtype=torch.int32, device="cuda")
        dec.decode_batch_f32(x, p1, h1, i1, 8, n_filler=int(nf[i]), flags=flags)
        torch.cuda.synchronize()
        o, n = offs[i], blocks[i].size
        assert int(i1.item()) == its[i]
        assert np.array_equal(f32_bits(p1.cpu().numpy()[0]), f32_bits(post[o:o + n]))
    # the tuner may change the launch shape, never the results
    best = plan.tune(d_in, d_post, d_hard, d_it, 8, flags)
    say("  tuned block", best, plan.shape())
    d_post2, d_hard2, d_it2 = torch.empty_like(d_post), torch.empty_like(d_hard), torch.empty_like(d_it)
    plan.run(d_in, d_post2, d_hard2, d_it2, 8, flags)
    torch.cuda.synchronize()
    assert torch.equal(d_it2, d_it) and torch.equal(d_hard2, d_hard)
    assert torch.equal(d_post2.view(torch.int32), d_post.view(torch.int32))


def test_plan_i8_mixed_matches_oracle(dev_ok, oracle):
    import syndrome_amd._lib as L
    from syndrome_amd import DecodePlan, QuantParams
    say("[plan i8] 28 mixed code blocks, fused quantizer")
    bg, z, nf, offs, arena, blocks = _mixed_inputs(oracle, 28, 9)
    plan = DecodePlan(bg, z, nf, offs, 0.5, int8=True)
    d_in = torch.from_numpy(arena).cuda()
    d_post = torch.empty(arena.size, dtype=torch.int16, device="cuda")
    d_hard = torch.empty(arena.size, dtype=torch.uint8, device="cuda")
    d_it = torch.empty(len(bg), dtype=torch.int32, device="cuda")
    plan.run(d_in, d_post, d_hard, d_it, 8, L.INIT_5G | L.WRITE_POST | L.WRITE_HARD | L.QUANTIZE,
             quant=QuantParams(8.0, 32767))
    torch.cuda.synchronize()
    post, hard, its = d_post.cpu().numpy(), d_hard.cpu().numpy(), d_it.cpu().numpy()
    for i in range(len(bg)):
        code = oracle.code_3gpp(int(bg[i]), int(z[i]))
        app = oracle.init_5g_i8(code, oracle.quantize_i16(blocks[i], 8.0), int(nf[i]), 32767)
        ri, ra, _, rh = oracle.decode_i8(code, 0.5, app, 8)
        o, n = offs[i], code.n
        assert its[i] == ri, f"block {i}: iterations"
        assert np.array_equal(post[o:o + n], ra), f"block {i}: app differs"
        assert np.array_equal(hard[o:o + n], rh), f"block {i}: hard differs"


def test_plan_i8_packed_and_one_per_lane_items(dev_ok, oracle):
    """int8 BG2 items run pair-packed, except where one pair would overfill
    the small launch's share of a CU (BG2 Z=256 beside a Z > 256 code):
    those run one codeword per lane.  Odd counts leave a pair with one live
    codeword.  Every block equals the oracle."""
    import syndrome_amd._lib as L
    from syndrome_amd import DecodePlan, QuantParams
    codes = [(1, 256)] * 3 + [(0, 384)] + [(1, 13)] * 5 + [(1, 104)] * 3 + [(1, 384)] * 3
    rng = np.random.default_rng(17)
    bg, z, nf, blocks = [], [], [], []
    for i, (b, zz) in enumerate(codes):
        code = oracle.code_3gpp(b, zz)
        f = int(rng.integers(0, min(32, code.k // 4) + 1))
        info = oracle.xorshift_bits(4242 + i, code.k)
        info[code.k - f:] = 0
        llr = oracle.awgn(float(rng.uniform(0.0, 3.0)), code.k / code.n, 99 + 13 * i, oracle.encode(code, info))
        bg.append(b)
        z.append(zz)
        nf.append(f)
        blocks.append(llr)
    offs = np.concatenate([[0], np.cumsum([x.size for x in blocks])[:-1]]).astype(np.int64)
    arena = np.concatenate(blocks).astype(np.float32)
    bg, z, nf = np.array(bg), np.array(z), np.array(nf)
    plan = DecodePlan(bg, z, nf, offs, 0.5, int8=True)
    say(f"[plan i8 packed] {len(bg)} blocks, shape {plan.shape()}")
    d_in = torch.from_numpy(arena).cuda()
    d_post = torch.empty(arena.size, dtype=torch.int16, device="cuda")
    d_hard = torch.empty(arena.size, dtype=torch.uint8, device="cuda")
    d_it = torch.empty(len(bg), dtype=torch.int32, device="cuda")
    for clamp in (32767, 300):
        plan.run(d_in, d_post, d_hard, d_it, 8, L.INIT_5G | L.WRITE_POST | L.WRITE_HARD | L.QUANTIZE,
                 quant=QuantParams(8.0, clamp))
        torch.cuda.synchronize()
        post, hard, its = d_post.cpu().numpy(), d_hard.cpu().numpy(), d_it.cpu().numpy()
        for i in range(len(bg)):
            code = oracle.code_3gpp(int(bg[i]), int(z[i]))
            app = oracle.init_5g_i8(code, oracle.quantize_i16(blocks[i], 8.0), int(nf[i]), clamp)
            ri, ra, _, rh = oracle.decode_i8(code, 0.5, app, 8, 8.0, clamp)
            o, n = offs[i], code.n
            tag = f"clamp {clamp} block {i} (BG{bg[i] + 1} Z={z[i]})"
            assert its[i] == ri, f"{tag}: iterations {its[i]} != {ri}"
            assert np.array_equal(post[o:o + n], ra), f"{tag}: app differs"
            assert np.array_equal(hard[o:o + n], rh), f"{tag}: hard differs"


def test_device_dematch_matches_oracle(dev_ok, oracle):
    from syndrome_amd import DecodePlan
    say("[dematch] 20 code blocks, Qm in {1,2,4,6}, rv 0..3, repetition and puncturing")
    rng = np.random.default_rng(3)
    bg, z, nf, e_len, qm, rv = [], [], [], [], [], []
    for i in range(20):
        b, zz = MIX[i % len(MIX)]
        k = (22 if b == 0 else 10) * zz
        q = [1, 2, 4, 6][i % 4]
        f = int(rng.integers(0, min(40, k // 4) + 1))
        ncb = (66 if b == 0 else 50) * zz
        hi = 3 * ncb if i % 6 == 5 else 2 * ncb
        e = int(rng.integers(max(q, ncb // 3), hi)) // q * q  # punctured, repeated, past the walk's guard
        bg.append(b)
        z.append(zz)
        nf.append(f)
        e_len.append(e)
        qm.append(q)
        rv.append(i % 4)
    n = [(68 if b == 0 else 52) * zz for b, zz in zip(bg, z)]
    offs = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    e_offs = np.concatenate([[0], np.cumsum(e_len)[:-1]]).astype(np.int64)
    e_llr = rng.normal(0, 3, int(sum(e_len))).astype(np.float32)
    plan = DecodePlan(bg, z, nf, offs, 0.5)
    plan.set_dematch(e_offs, e_len, qm, rv)
    d_e = torch.from_numpy(e_llr).cuda()
    d_out = torch.full((int(sum(n)),), 7.0, dtype=torch.float32, device="cuda")
    plan.dematch(d_e, d_out)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    for i in range(20):
        cb = oracle.rate_dematch(e_llr[e_offs[i]:e_offs[i] + e_len[i]], rv[i], qm[i], bg[i], z[i], nf[i], cached=True)
        ref = oracle.decoder_input_from_cb(cb, z[i])
        assert np.array_equal(f32_bits(out[offs[i]:offs[i] + n[i]]), f32_bits(ref)), f"block {i}"


def test_device_dematch_no_repetition_matches_oracle(dev_ok, oracle):
    """The de-match kernel's direct path (E <= m, no combining): Qm 1, 2, 4,
    6 and 8, every rv, fillers, signed zeros in the received LLRs (the
    reference's += turns -0 into +0), and blocks at unaligned output offsets,
    which take the general path beside it."""
    from syndrome_amd import DecodePlan
    say("[dematch] 24 code blocks, Qm 1..8, E <= m, rv 0..3, -0.0 inputs")
    rng = np.random.default_rng(11)
    bg, z, nf, e_len, rv, qm = [], [], [], [], [], []
    for i in range(24):
        b, zz = MIX[i % len(MIX)]
        k = (22 if b == 0 else 10) * zz
        f = int(rng.integers(0, min(40, k // 4) + 1)) if i % 3 else 0
        ncb = (66 if b == 0 else 50) * zz
        m = ncb - f  # non-filler positions of the circular buffer
        q = [1, 2, 4, 6, 8, 1][i % 6]
        e = (int(rng.integers(max(q, ncb // 4), m + 1)) if i % 5 else m) // q * q
        qm.append(q)
        bg.append(b)
        z.append(zz)
        nf.append(f)
        e_len.append(e)
        rv.append(i % 4)
    n = [(68 if b == 0 else 52) * zz for b, zz in zip(bg, z)]
    pad = [(i % 4 == 3) * (1 + i % 3) for i in range(24)]  # some offsets not a multiple of 4
    offs = np.concatenate([[0], np.cumsum(np.array(n) + np.array(pad))[:-1]]).astype(np.int64) + np.array(pad)
    e_offs = np.concatenate([[0], np.cumsum(e_len)[:-1]]).astype(np.int64)
    e_llr = rng.normal(0, 3, int(sum(e_len))).astype(np.float32)
    e_llr[rng.choice(e_llr.size, e_llr.size // 50, replace=False)] = np.float32(-0.0)
    plan = DecodePlan(bg, z, nf, offs, 0.5)
    plan.set_dematch(e_offs, e_len, qm, rv)
    d_e = torch.from_numpy(e_llr).cuda()
    d_out = torch.full((int(offs[-1] + n[-1]),), 7.0, dtype=torch.float32, device="cuda")
    plan.dematch(d_e, d_out)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    for i in range(24):
        cb = oracle.rate_dematch(e_llr[e_offs[i]:e_offs[i] + e_len[i]], rv[i], qm[i], bg[i], z[i], nf[i], cached=True)
        ref = oracle.decoder_input_from_cb(cb, z[i])
        assert np.array_equal(f32_bits(out[offs[i]:offs[i] + n[i]]), f32_bits(ref)), f"block {i}"


@pytest.fixture(scope="module")
def dev_ok():
    assert torch.cuda.is_available()
    import syndrome_amd
    syndrome_amd.lib()
    return True


def test_device_harq_combine_matches_oracle(dev_ok, oracle):
    """Four transmissions (rv 0, 2, 3, 1) of mixed code blocks accumulated on
    the device equal the oracle's HarqBuffer restatement after every one."""
    from syndrome_amd import DecodePlan
    say("[harq] 14 code blocks x 4 transmissions, rv 0/2/3/1, Qm 1/2/4/6, E up to 3 cycles")
    rng = np.random.default_rng(9)
    bg = np.array([b for b, _ in MIX])
    z = np.array([zz for _, zz in MIX])
    k = np.where(bg == 0, 22, 10) * z
    nf = np.array([int(rng.integers(0, min(40, kk // 4) + 1)) for kk in k])
    ncb = np.where(bg == 0, 66, 50) * z
    n = ncb + 2 * z
    qm = np.array([[1, 2, 4, 6][i % 4] for i in range(len(bg))])
    # lengths from punctured to beyond the selection walk's guard (> 2 cycles)
    e_len = np.array([int(rng.integers(max(1, c // 4), c * (3 if i % 5 == 0 else 2) + 7)) // q * q
                      for i, (c, q) in enumerate(zip(ncb, qm))])
    offs = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    e_off = np.concatenate([[0], np.cumsum(e_len)[:-1]]).astype(np.int64)
    plan = DecodePlan(bg, z, nf, offs, 0.5)
    plan.set_dematch(e_off, e_len, qm)
    d_acc = torch.zeros(int(n.sum()), dtype=torch.float32, device="cuda")
    acc = [np.zeros(c, np.float32) for c in ncb]
    for rv in (0, 2, 3, 1):
        e_llr = rng.normal(0, 2.5, int(e_len.sum())).astype(np.float32)
        plan.harq_combine(torch.from_numpy(e_llr).cuda(), d_acc, rv=rv)
        torch.cuda.synchronize()
        got = d_acc.cpu().numpy()
        for i in range(len(bg)):
            acc[i] = oracle.rate_dematch(e_llr[e_off[i]:e_off[i] + e_len[i]], rv, int(qm[i]), int(bg[i]), int(z[i]),
                                         int(nf[i]), cb=acc[i], cached=True)
            ref = oracle.decoder_input_from_cb(acc[i], int(z[i]))
            assert np.array_equal(f32_bits(got[offs[i]:offs[i] + n[i]]), f32_bits(ref)), f"rv {rv} block {i}"


def test_harq_recovers_blocks_that_fail_alone(dev_ok, oracle):
    """As transport_block.rs:1255-1292: BG2 Z=56 (400-bit TB at rate 1/2, so
    K' = 424 and 136 fillers), E = 700, -1.5 dB, 20 iterations, beta 0.25:
    the rv-0 shot alone fails for each block, rv 0 + rv 3 combined on the
    device decodes each exactly (the oracle agrees bit for bit)."""
    from syndrome_amd import DecodePlan, _lib as L
    say("[harq] 4 blocks that need two transmissions")
    bgv, zv, nfv, E = 1, 56, 136, 700
    code = oracle.code_3gpp(bgv, zv)
    kp = code.k - nfv
    trials = [1, 2, 17, 19]  # seeds where shot 1 fails and the combination succeeds
    infos, e0, e3 = [], [], []
    for t in trials:
        info = oracle.xorshift_bits(100 + t, code.k)
        info[kp:] = 0
        cw = oracle.encode(code, info)
        infos.append(info)
        e0.append(oracle.awgn(np.float32(-1.5), np.float32(0.5), 1 + t, oracle.rate_match(cw, E, 0, 1, bgv, zv, nfv)))
        e3.append(oracle.awgn(np.float32(-1.5), np.float32(0.5), 1001 + t, oracle.rate_match(cw, E, 3, 1, bgv, zv, nfv)))
    cnt, n = len(trials), code.n
    offs = (np.arange(cnt) * n).astype(np.int64)
    plan = DecodePlan(np.full(cnt, bgv), np.full(cnt, zv), np.full(cnt, nfv), offs, 0.25)
    plan.set_dematch((np.arange(cnt) * E).astype(np.int64), np.full(cnt, E))
    d_acc = torch.zeros(cnt * n, dtype=torch.float32, device="cuda")
    d_post = torch.empty_like(d_acc)
    d_hard = torch.empty(cnt * n, dtype=torch.uint8, device="cuda")
    d_it = torch.empty(cnt, dtype=torch.int32, device="cuda")
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD
    acc = [np.zeros(n - 2 * zv, np.float32) for _ in trials]
    for shot, (rv, es) in enumerate(((0, e0), (3, e3))):
        plan.harq_combine(torch.from_numpy(np.concatenate(es)).cuda(), d_acc, rv=rv)
        plan.run(d_acc, d_post, d_hard, d_it, 20, flags)
        torch.cuda.synchronize()
        hard, its = d_hard.cpu().numpy(), d_it.cpu().numpy()
        for j in range(cnt):
            acc[j] = oracle.rate_dematch(es[j], rv, 1, bgv, zv, nfv, cb=acc[j], cached=True)
            inp = oracle.init_5g_f32(code, oracle.decoder_input_from_cb(acc[j], zv), nfv)
            ri, _, _, rh = oracle.decode_f32(code, 0.25, inp, 20)
            h = hard[j * n:(j + 1) * n]
            assert its[j] == ri and np.array_equal(h, rh), f"shot {shot} block {j}: differs from the oracle"
            ok = np.array_equal(h[:kp], infos[j][:kp])
            assert ok == (shot == 1), f"shot {shot} block {j}: decoded={ok}"


def test_plan_runs_in_flight_together_match_one_run(dev_ok, oracle):
    """Runs of one plan issued back to back on three streams, with no wait
    between them, overlap on the device; each takes its own queue counter,
    so each output set equals what one run alone gives."""
    import syndrome_amd._lib as L
    from syndrome_amd import DecodePlan
    say("[plan] three runs in flight on three streams")
    bg, z, nf, offs, arena, blocks = _mixed_inputs(oracle, 42, 9)
    plan = DecodePlan(bg, z, nf, offs, 0.5)
    d_in = torch.from_numpy(arena).cuda()
    flags = L.INIT_5G | L.WRITE_POST | L.WRITE_HARD
    ref = (torch.empty_like(d_in), torch.empty(arena.size, dtype=torch.uint8, device="cuda"),
           torch.empty(len(bg), dtype=torch.int32, device="cuda"))
    plan.run(d_in, ref[0], ref[1], ref[2], 8, flags)
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream() for _ in range(3)]
    outs = [(torch.full_like(d_in, float("nan")), torch.full((arena.size,), 9, dtype=torch.uint8, device="cuda"),
             torch.full((len(bg),), -1, dtype=torch.int32, device="cuda")) for _ in streams]
    torch.cuda.synchronize()
    for rep in range(3):
        for s, (p, h, it) in zip(streams, outs):
            plan.run(d_in, p, h, it, 8, flags, stream=s.cuda_stream)
    torch.cuda.synchronize()
    for p, h, it in outs:
        assert torch.equal(it, ref[2]) and torch.equal(h, ref[1])
        assert torch.equal(p.view(torch.int32), ref[0].view(torch.int32))
