"""GPU parity of the transport-block tail on device: CRC-24B per code block,
reassembly and CRC-24A (transport_block.rs:963-1044), after the device
de-matcher and the persistent decoder, against the oracle's restatement on
the same decoded bits; clean blocks deliver the sent transport blocks, a
planted bit error or an undecodable block is caught."""
import numpy as np
import pytest

from tests.helpers import say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TBS = [(100, 0.5, 3.0), (1000, 1 / 3, 2.0), (8425, 0.5, 2.5), (30000, 2 / 3, 3.5), (60000, 5 / 6, 4.5),
       (3000, 0.5, -2.0)]  # the last one is sent far below its threshold


def _build(oracle):
    from syndrome_amd.transport import compute_segmentation
    cbs, tbs = [], []
    for t, (a, rate, ebno) in enumerate(TBS):
        rng = np.random.default_rng(t + 7)
        tb = rng.integers(0, 2, a).astype(np.uint8)
        p = compute_segmentation(a, rate)
        blocks = oracle.segment(oracle.attach_crc(tb, oracle.CRC24A), p.c, p.k_prime)
        code = oracle.code_3gpp(p.bg, p.z)
        e = max(int(round((a + 24) / rate / p.c)), 8)
        first = len(cbs)
        for i in range(p.c):
            info = np.concatenate([blocks[i], np.zeros(p.n_filler, np.uint8)])
            cw = oracle.encode(code, info)
            llr = oracle.awgn(np.float32(ebno), np.float32(rate), 1000 * t + i,
                              oracle.rate_match(cw, e, 0, 1, p.bg, p.z, p.n_filler))
            cbs.append((p.bg, p.z, p.n_filler, llr))
        tbs.append((first, p.c, a, p.k_prime, tb))
    return cbs, tbs


def test_tb_checks_match_oracle(oracle):
    import syndrome_amd
    from syndrome_amd import DecodePlan, _lib as L
    assert torch.cuda.is_available()
    syndrome_amd.lib()
    cbs, tbs = _build(oracle)
    say(f"[tb] {len(tbs)} transport blocks, {len(cbs)} code blocks: de-match, decode, CRC checks")
    bg = np.array([c[0] for c in cbs])
    z = np.array([c[1] for c in cbs])
    nf = np.array([c[2] for c in cbs])
    e_len = np.array([c[3].size for c in cbs])
    n = np.where(bg == 0, 68, 52) * z
    offs = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    e_off = np.concatenate([[0], np.cumsum(e_len)[:-1]]).astype(np.int64)
    plan = DecodePlan(bg, z, nf, offs, 0.5)
    plan.set_dematch(e_off, e_len)
    d_llr = torch.empty(int(n.sum()), dtype=torch.float32, device="cuda")
    plan.dematch(torch.from_numpy(np.concatenate([c[3] for c in cbs])).cuda(), d_llr)
    d_hard = torch.empty(int(n.sum()), dtype=torch.uint8, device="cuda")
    d_it = torch.empty(len(cbs), dtype=torch.int32, device="cuda")
    plan.run(d_llr, None, d_hard, d_it, 8, L.INIT_5G | L.WRITE_HARD)
    out_off = np.concatenate([[0], np.cumsum([t[2] for t in tbs])[:-1]]).astype(np.int64)
    plan.set_tbs([t[0] for t in tbs], [t[1] for t in tbs], [t[2] for t in tbs], out_off)
    d_out = torch.full((int(sum(t[2] for t in tbs)),), 9, dtype=torch.uint8, device="cuda")
    d_tb_ok = torch.full((len(tbs),), 7, dtype=torch.uint8, device="cuda")
    d_cb_ok = torch.full((len(cbs),), 7, dtype=torch.uint8, device="cuda")

    def check_against_oracle(hard, what):
        plan.check_tbs(torch.from_numpy(hard).cuda(), d_out, d_tb_ok, d_cb_ok)
        torch.cuda.synchronize()
        out, tb_ok, cb_ok = d_out.cpu().numpy(), d_tb_ok.cpu().numpy(), d_cb_ok.cpu().numpy()
        for ti, (first, c, a, kp, tb) in enumerate(tbs):
            ok, rcb, rout = oracle.tb_check(hard, offs[first:first + c], c, kp, a)
            assert tb_ok[ti] == ok, f"{what}: TB {ti} ok {tb_ok[ti]} != {ok}"
            assert np.array_equal(cb_ok[first:first + c], rcb), f"{what}: TB {ti} block flags"
            if ok:
                assert np.array_equal(out[out_off[ti]:out_off[ti] + a], rout), f"{what}: TB {ti} payload"
                assert np.array_equal(rout, tb), f"{what}: TB {ti} not the sent bits"
        return tb_ok

    hard = d_hard.cpu().numpy()
    tb_ok = check_against_oracle(hard, "decoded")
    assert tb_ok[:-1].all(), "clean transport blocks must pass"
    assert not tb_ok[-1], "the block sent far below threshold must fail"
    # a planted error in one payload bit of the 3rd TB's 2nd block
    first, c, a, kp, _ = tbs[2]
    bad = hard.copy()
    bad[offs[first + 1] + 5] ^= 1
    d_out.fill_(9)
    tb_ok = check_against_oracle(bad, "planted")
    assert not tb_ok[2] and tb_ok[0] and tb_ok[1]
