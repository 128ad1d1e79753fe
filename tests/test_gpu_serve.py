"""GPU tests of the serving loop: continuous batching over mixed code blocks.

* Every code block served in the mix (batches of varying composition,
  overlapping on several launch lanes) equals the oracle's decode of it
  alone and the per-config batch decode of it alone: hard decisions and
  iterations.
* Admission by HBM budget: with the slots exhausted, acquire refuses
  (Busy) until a slot is released.
"""
import time

import numpy as np
import pytest

from tests.helpers import say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MIX = [(0, 384), (1, 208), (0, 2), (1, 3), (0, 44), (1, 52), (0, 104), (1, 384), (0, 15), (1, 120), (0, 256),
       (1, 7), (0, 352), (1, 26), (0, 96), (1, 160)]


def _blocks(oracle, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        b, z = MIX[int(rng.integers(0, len(MIX)))]
        code = oracle.code_3gpp(b, z)
        nf = int(rng.integers(0, min(48, code.k // 4) + 1))
        info = oracle.xorshift_bits(seed * 7919 + i, code.k)
        info[code.k - nf:] = 0
        cw = oracle.encode(code, info)
        llr = oracle.awgn(float(rng.uniform(0.0, 3.0)), code.k / code.n, seed + 31 * i, cw)
        out.append((b, z, nf, llr))
    return out


@pytest.mark.parametrize("streams", [3, 0], ids=["lanes", "resident"])
def test_server_concurrent_clients(oracle, streams):
    """Four client threads share one server, each acquiring, submitting,
    receiving (whichever blocks are ready, its own or not) and releasing at
    once: the server's lock and its lock-free queue counts under contention.
    Every block must come back exactly once, equal to the oracle's decode."""
    import threading
    import syndrome_amd
    from syndrome_amd import Server
    syndrome_amd.lib()
    say(f"[serve] 4 concurrent clients x 30 blocks, streams={streams}")
    blocks = _blocks(oracle, 120, 23)
    srv = Server(0.5, 8, hbm_budget_bytes=64 << 20, max_batch=32, streams=streams)
    results, lock = {}, threading.Lock()
    received = [0]
    errors = []

    def client(k):
        try:
            mine = list(range(k, len(blocks), 4))
            t_end = time.time() + 60
            i = 0
            while time.time() < t_end:
                if i < len(mine):
                    b, z, nf, llr = blocks[mine[i]]
                    got = srv.acquire(b, z, nf)
                    if got is not None:
                        slot, buf = got
                        buf[:] = llr
                        srv.submit(slot, mine[i])
                        i += 1
                r = srv.try_recv()
                if r is not None:
                    slot, tag, its, hard, _, _ = r
                    with lock:
                        assert tag not in results, f"block {tag} delivered twice"
                        results[tag] = (its, hard)
                        received[0] += 1
                    srv.release(slot)
                with lock:
                    if received[0] == len(blocks):
                        return
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(90)
    srv.close()
    assert not errors, errors
    assert len(results) == len(blocks), f"only {len(results)} of {len(blocks)} completed"
    for tag, (b, z, nf, llr) in enumerate(blocks):
        code = oracle.code_3gpp(b, z)
        ri, _, _, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, llr, nf), 8)
        its, hard = results[tag]
        assert its == ri, f"block {tag} (BG{b + 1} Z={z}): iterations {its} != {ri}"
        assert np.array_equal(hard, rh), f"block {tag}: hard decisions differ"


@pytest.mark.parametrize("streams", [3, 0], ids=["lanes", "resident"])
def test_server_results_equal_alone_and_in_the_mix(oracle, streams):
    import syndrome_amd
    from syndrome_amd import QcLdpcDecoder, Server, _lib as L
    assert torch.cuda.is_available()
    syndrome_amd.lib()
    say(f"[serve] 160 mixed code blocks, streams={streams}, bursts of varying size")
    blocks = _blocks(oracle, 160, 11)
    srv = Server(0.5, 8, hbm_budget_bytes=96 << 20, max_batch=48, streams=streams)
    rng = np.random.default_rng(2)
    results = {}
    i = 0
    t_end = time.time() + 60
    while len(results) < len(blocks) and time.time() < t_end:
        burst = int(rng.integers(1, 12))
        for _ in range(burst):
            if i >= len(blocks):
                break
            b, z, nf, llr = blocks[i]
            got = srv.acquire(b, z, nf)
            if got is None:
                break
            slot, buf = got
            buf[:] = llr
            srv.submit(slot, i)
            i += 1
        time.sleep(float(rng.uniform(0, 0.002)))
        while True:
            r = srv.try_recv()
            if r is None:
                break
            slot, tag, its, hard, t0, t1 = r
            results[tag] = (its, hard)
            srv.release(slot)
    assert len(results) == len(blocks), f"only {len(results)} of {len(blocks)} completed"
    st = srv.stats()
    say("  stats", st)
    if streams:
        assert st["batches"] > 3  # really batched in several, varying groups
    else:
        assert st["batches"] >= 1  # grid launches
    for tag, (b, z, nf, llr) in enumerate(blocks):
        code = oracle.code_3gpp(b, z)
        ri, _, _, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, llr, nf), 8)
        its, hard = results[tag]
        assert its == ri, f"block {tag} (BG{b + 1} Z={z}): iterations {its} != {ri}"
        assert np.array_equal(hard, rh), f"block {tag}: hard decisions differ"
    # and alone on the GPU
    for tag in (0, 1, 2, 3):
        b, z, nf, llr = blocks[tag]
        dec = QcLdpcDecoder.with_lifting_size(b, z, 0.5)
        x = torch.from_numpy(llr).cuda()[None, :]
        h = torch.empty(x.shape, dtype=torch.uint8, device="cuda")
        it = torch.empty(1, dtype=torch.int32, device="cuda")
        dec.decode_batch_f32(x, None, h, it, 8, n_filler=nf, flags=L.INIT_5G | L.WRITE_HARD)
        torch.cuda.synchronize()
        assert int(it.item()) == results[tag][0]
        assert np.array_equal(h.cpu().numpy()[0], results[tag][1])
    srv.close()


def test_server_admission_by_budget():
    import syndrome_amd
    from syndrome_amd import Server
    syndrome_amd.lib()
    say("[serve] admission: 3-slot budget")
    slot_bytes = 68 * 384 * 4 + 64
    srv = Server(0.5, 4, hbm_budget_bytes=3 * slot_bytes, max_batch=8, streams=1)
    assert srv.stats()["slots"] == 3
    held = [srv.acquire(1, 16) for _ in range(3)]
    assert all(h is not None for h in held)
    assert srv.acquire(1, 16) is None  # Busy: budget exhausted
    srv.release(held[0][0])
    again = srv.acquire(1, 16)
    assert again is not None
    # a strong all-zero block decodes to zeros in one iteration
    slot, buf = again
    buf[:] = 8.0
    srv.submit(slot, 99)
    t_end = time.time() + 20
    r = None
    while r is None and time.time() < t_end:
        r = srv.try_recv()
    assert r is not None and r[1] == 99 and r[2] == 1 and not r[3].any()
    srv.release(slot)
    for h in held[1:]:
        srv.release(h[0])
    srv.close()


@pytest.mark.parametrize("streams", [2, 0], ids=["lanes", "resident"])
def test_loadgen_accounts_every_block(streams):
    """The C++ load generator drives acquire/submit_from/try_recv/release:
    saturating and Poisson modes both finish every admitted transport block,
    and a TB's latency is that of its last block (same arrival instant)."""
    import syndrome_amd
    from syndrome_amd import DecodePlan, HostPool, Server, serve_loadgen, traffic_blocks
    from syndrome_amd.transport import compute_segmentation
    syndrome_amd.lib()
    say(f"[serve] loadgen streams={streams}: 6 TBs, saturating 0.3 s then Poisson 200 TB/s for 0.3 s")
    bg, z, nf, e, kp, first, num, info = [], [], [], [], [], [], [], []
    for a, r in ((1000, 0.5), (20000, 0.75), (3000, 1 / 3), (60000, 5 / 6), (300, 0.5), (9000, 2 / 3)):
        sg = compute_segmentation(a, r)
        first.append(len(bg))
        num.append(sg.c)
        for _ in range(sg.c):
            bg.append(sg.bg)
            z.append(sg.z)
            nf.append(sg.n_filler)
            e.append(int(round((a + 24) / r / sg.c)))
            kp.append(sg.k_prime)
            info.append(a // sg.c)
    bg, z, nf, e, kp = map(np.array, (bg, z, nf, e, kp))
    e_llr, e_off, _, _ = traffic_blocks(bg, z, nf, e, np.full(len(bg), 3.0), kp / e, 5, want_info=False)
    n = np.where(bg == 0, 68, 52) * z
    offs = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    plan = DecodePlan(bg, z, nf, offs, 0.5)
    plan.set_dematch(e_off, e)
    d = torch.empty(int(n.sum()), dtype=torch.float32, device="cuda")
    plan.dematch(torch.from_numpy(e_llr).cuda(), d)
    pool = HostPool(int(n.sum()))
    pool.array[:] = d.cpu().numpy()
    srv = Server(0.5, 8, hbm_budget_bytes=64 << 20, max_batch=64, streams=streams)
    cbl, tbl, bps, _ = serve_loadgen(srv, pool, offs, bg, z, nf, info, first, num, 0, 0.3)
    say(f"  saturating: {len(cbl)} blocks, {len(tbl)} TBs, {bps / 1e9:.3f} Gbit/s")
    assert len(cbl) > 0 and len(tbl) > 0 and bps > 0 and (cbl > 0).all() and (tbl > 0).all()
    cbl, tbl, bps, _ = serve_loadgen(srv, pool, offs, bg, z, nf, info, first, num, 200.0, 0.3, seed=3)
    say(f"  poisson: {len(cbl)} blocks, {len(tbl)} TBs, p50 {np.median(cbl) / 1e6:.3f} ms")
    assert len(tbl) > 10  # ~60 arrivals expected
    assert len(cbl) == sum(num[i % 6] for i in range(len(tbl)))  # every admitted TB completed
    assert (cbl > 0).all() and tbl.max() == cbl.max() and tbl.min() >= cbl.min()
    st = srv.stats()
    assert st["decoded"] >= len(cbl)
    srv.close()


def test_server_transport_blocks_match_oracle(oracle):
    """Transport blocks through the resident server: each block's CRC-24B,
    the CRC-24A and the reassembled payload are decided on the device from
    the decoded posteriors; they equal the oracle's decode and TB check of
    the same blocks, and plain code blocks submitted in between are
    delivered as before."""
    import syndrome_amd
    from syndrome_amd import HostPool, Server
    from tests.test_gpu_tb import TBS, _build
    syndrome_amd.lib()
    cbs, tbs = _build(oracle)
    say(f"[serve tb] {len(tbs)} transport blocks ({len(cbs)} code blocks) + loose blocks, resident mode")
    inputs = []
    for bg, z, nf, e_llr in cbs:
        cb = oracle.rate_dematch(e_llr, 0, 1, bg, z, nf, cached=True)
        inputs.append(oracle.decoder_input_from_cb(cb, z))
    total = int(sum(t[2] for t in tbs))
    out = HostPool(total, np.uint8)
    out.array[:] = 9
    out_off = np.concatenate([[0], np.cumsum([t[2] for t in tbs])[:-1]]).astype(np.int64)
    loose = _blocks(oracle, 12, 5)
    srv = Server(0.5, 8, hbm_budget_bytes=256 << 20, max_batch=64, streams=0)
    for ti, (first, c, a, kp, tb) in enumerate(tbs):
        slots = []
        for ci in range(c):
            bg, z, nf, _ = cbs[first + ci]
            got = srv.acquire(bg, z, nf)
            assert got is not None
            slots.append(got[0])
            got[1][:] = inputs[first + ci]
        assert srv.submit_tb(slots, a, out.ptr + int(out_off[ti]), 1000 + ti)
        if ti < len(loose):  # interleave loose code blocks
            b, z, nf, llr = loose[ti]
            slot, buf = srv.acquire(b, z, nf)
            buf[:] = llr
            srv.submit(slot, ti)
    got_tb, got_cb = {}, {}
    t_end = time.time() + 60
    while (len(got_tb) < len(tbs) or len(got_cb) < min(len(loose), len(tbs))) and time.time() < t_end:
        r = srv.try_recv_tb()
        if r is not None:
            ti = r[1] - 1000
            got_tb[ti] = r + (srv.tb_block_crc(r[0], tbs[ti][1]),)
            srv.release_tb(r[0])
        r = srv.try_recv()
        if r is not None:
            got_cb[r[1]] = r
            srv.release(r[0])
    assert len(got_tb) == len(tbs), f"{len(got_tb)} of {len(tbs)} transport blocks"
    for ti, (first, c, a, kp, tb) in enumerate(tbs):
        hards, its = [], []
        for ci in range(c):
            bg, z, nf, _ = cbs[first + ci]
            code = oracle.code_3gpp(bg, z)
            ri, _, _, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, inputs[first + ci], nf), 8)
            hards.append(rh)
            its.append(ri)
        n = hards[0].size
        ok, cb_ok, _ = oracle.tb_check(np.concatenate(hards), np.arange(c) * n, c, kp, a)
        _, tag, crc_ok, max_it, _, _, cb_crc = got_tb[ti]
        assert crc_ok == ok and max_it == max(its), f"TB {ti}: crc_ok {crc_ok} (oracle {ok}), iterations"
        assert np.array_equal(cb_crc, cb_ok.astype(bool)), f"TB {ti}: per-block CRC-24B {cb_crc} (oracle {cb_ok})"
        p = kp - 24 if c > 1 else kp
        stream = np.concatenate([h[:p] for h in hards])[:a]
        assert np.array_equal(out.array[out_off[ti]:out_off[ti] + a], stream), f"TB {ti}: payload bits"
        if ok:
            assert np.array_equal(stream, tb)
    for t, r in got_cb.items():
        b, z, nf, llr = loose[t]
        code = oracle.code_3gpp(b, z)
        ri, _, _, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, llr, nf), 8)
        assert r[2] == ri and np.array_equal(r[3], rh), f"loose block {t}"
    srv.close()


def test_loadgen_transport_blocks():
    """The transport-block load generator: every admitted transport block
    completes, at 5 dB (above every rate's threshold) every CRC passes, and
    the payloads written on the device are the sent transport blocks."""
    import syndrome_amd
    from syndrome_amd import DecodePlan, Server, serve_loadgen_tb, traffic_blocks_info
    from syndrome_amd.transport import attach_crc24, compute_segmentation, segment
    syndrome_amd.lib()
    say("[serve tb loadgen] 5 TBs, saturating 0.3 s then Poisson 150 TB/s for 0.3 s")
    rng = np.random.default_rng(12)
    bg, z, nf, e, kp, first, num, sizes, infos, payloads = [], [], [], [], [], [], [], [], [], []
    for a, r in ((1000, 0.5), (20000, 0.75), (3000, 1 / 3), (45000, 5 / 6), (300, 0.5)):
        sg = compute_segmentation(a, r)
        tb = rng.integers(0, 2, a).astype(np.uint8)
        blocks = segment(attach_crc24(tb), sg)
        first.append(len(bg))
        num.append(sg.c)
        sizes.append(a)
        payloads.append(tb)
        for i in range(sg.c):
            bg.append(sg.bg)
            z.append(sg.z)
            nf.append(sg.n_filler)
            e.append(int(round((a + 24) / r / sg.c)))
            kp.append(sg.k_prime)
            infos.append(blocks[i])
    bg, z, nf, e, kp = map(np.array, (bg, z, nf, e, kp))
    i_off = np.concatenate([[0], np.cumsum(kp)[:-1]]).astype(np.int64)
    e_llr, e_off = traffic_blocks_info(bg, z, nf, e, np.full(len(bg), 5.0), kp / e, np.concatenate(infos), i_off, 4)
    n = np.where(bg == 0, 68, 52) * z
    offs = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    plan = DecodePlan(bg, z, nf, offs, 0.5)
    plan.set_dematch(e_off, e)
    d = torch.empty(int(n.sum()), dtype=torch.float32, device="cuda")
    plan.dematch(torch.from_numpy(e_llr).cuda(), d)
    out_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    d_out = torch.zeros(int(sum(sizes)), dtype=torch.uint8, device="cuda")
    srv = Server(0.5, 8, hbm_budget_bytes=64 << 20, max_batch=64, streams=0)
    lat, ncb, bps, fails, _, good, okf = serve_loadgen_tb(srv, d, offs, bg, z, nf, first, num, sizes, d_out, out_off, 0,
                                                         0.3)
    say(f"  saturating: {len(lat)} TBs, {ncb} blocks, {bps / 1e9:.3f} Gbit/s, {fails} CRC failures")
    assert len(lat) > 0 and fails == 0 and (lat > 0).all()
    assert okf.all() and good == pytest.approx(bps)  # every transport block passed: goodput = throughput
    torch.cuda.synchronize()
    got = d_out.cpu().numpy()
    for t, tb in enumerate(payloads):
        assert np.array_equal(got[out_off[t]:out_off[t] + sizes[t]], tb), f"TB {t} payload"
    lat, ncb, bps, fails, _, good, okf = serve_loadgen_tb(srv, d, offs, bg, z, nf, first, num, sizes, d_out, out_off,
                                                         150.0, 0.3, seed=2)
    say(f"  poisson: {len(lat)} TBs, p50 {np.median(lat) / 1e6:.3f} ms")
    assert len(lat) > 10 and fails == 0
    assert ncb == sum(num[i % 5] for i in range(len(lat)))
    srv.close()


@pytest.mark.parametrize("streams", [3, 0], ids=["lanes", "resident"])
def test_server_batched_calls(oracle, streams):
    """acquire_many / submit_many / try_recv_many / release_many: partial
    acquisition at the budget, inputs from pinned staging and from HBM, the
    same results as the oracle alone, all-or-nothing validation."""
    import syndrome_amd
    from syndrome_amd import Server, SyndromeError
    syndrome_amd.lib()
    say(f"[serve] batched calls, streams={streams}")
    blocks = _blocks(oracle, 96, 23)
    slot_bytes = 68 * 384 * 4 + 64
    srv = Server(0.5, 8, hbm_budget_bytes=40 * slot_bytes, max_batch=32, streams=streams)
    # validation first: a repeated slot submits nothing, an unacquired slot releases nothing
    sl, views = srv.acquire_many([0, 1], [8, 8])
    assert len(sl) == 2 and views[0].size == 68 * 8 and views[1].size == 52 * 8
    with pytest.raises(SyndromeError):
        srv.submit_many([sl[0], sl[0]])
    with pytest.raises(SyndromeError):
        srv.release_many([sl[0], next(x for x in range(40) if x not in sl)])
    srv.release_many(sl)  # both still only acquired (the failed calls rolled back): released now
    d_pool = torch.from_numpy(np.concatenate([llr for _, _, _, llr in blocks])).cuda()
    offs = np.concatenate([[0], np.cumsum([llr.size for _, _, _, llr in blocks])[:-1]])
    results, i, partial = {}, 0, False
    t_end = time.time() + 60
    while len(results) < len(blocks) and time.time() < t_end:
        want = blocks[i:i + 24]
        if want:
            got, views = srv.acquire_many([b for b, _, _, _ in want], [z for _, z, _, _ in want],
                                          [nf for _, _, nf, _ in want])
            partial = partial or 0 < len(got) < len(want)
            srcs = []
            for j, slot in enumerate(got):
                k = i + j
                if k % 2:
                    views[j][:] = blocks[k][3]
                    srcs.append(0)  # pinned staging
                else:
                    srcs.append(d_pool.data_ptr() + 4 * int(offs[k]))
            if len(got):
                srv.submit_many(got, srcs, [i + j for j in range(len(got))])
            i += len(got)
        done = srv.try_recv_many(16)
        for slot, tag, its, hard, t0, t1 in done:
            results[tag] = (its, hard)
        if done:
            srv.release_many([d[0] for d in done])
    assert len(results) == len(blocks), f"only {len(results)} of {len(blocks)} completed"
    assert partial, "the 40-slot budget never cut an acquire_many short"
    for tag, (b, z, nf, llr) in enumerate(blocks):
        code = oracle.code_3gpp(b, z)
        ri, _, _, rh = oracle.decode_f32(code, 0.5, oracle.init_5g_f32(code, llr, nf), 8)
        its, hard = results[tag]
        assert its == ri, f"block {tag} (BG{b + 1} Z={z}): iterations {its} != {ri}"
        assert np.array_equal(hard, rh), f"block {tag}: hard decisions differ"
    srv.close()
