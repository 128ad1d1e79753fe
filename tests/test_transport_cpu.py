"""Transport-block segmentation used to build serving traffic, against the
reference's own segmentation tests (src/segmentation.rs:466-640) and
TS 38.212 5.2.2 values worked by hand."""
import pytest

from syndrome_amd.transport import MAX_TB_SIZE_BITS, compute_segmentation, select_base_graph

AWKWARD = [1, 100, 292, 293, 3800, 3824, 3825, 4000, 8424, 8425, 8426, 8447, 8448, 8449, 10000, 10001, 12345,
           16896, 16897, 20000, 50000, 100_000, 500_000, 1_000_000, MAX_TB_SIZE_BITS]


def test_base_graph_rules():
    assert select_base_graph(292, 0.9) == 1 and select_base_graph(293, 0.9) == 0
    assert select_base_graph(3824, 0.67) == 1 and select_base_graph(3825, 0.67) == 0
    assert select_base_graph(4000, 0.5) == 0
    assert select_base_graph(9999, 0.25) == 1
    assert select_base_graph(3824, 0.6701) == 0


def test_worked_examples():
    p = compute_segmentation(100, 0.5)  # B = 124 <= 3840, BG2, one block, K_b = 6
    assert (p.bg, p.c, p.has_cb_crc, p.k_prime, p.z, p.k, p.n_filler) == (1, 1, False, 124, 22, 220, 96)
    p = compute_segmentation(10000, 0.5)  # C = ceil(10024/8424) = 2, K' = (10024+48)/2
    assert (p.bg, p.b, p.c, p.k_prime, p.z, p.n_filler, p.n) == (0, 10024, 2, 5036, 240, 5280 - 5036, 66 * 240)
    p = compute_segmentation(8425, 0.5)  # B' = 8497 not divisible by C = 2
    assert (p.bg, p.c, p.k_prime) == (0, 2, 4249)
    p = compute_segmentation(500000, 0.5)
    assert (p.bg, p.z, p.c, p.k_prime) == (0, 384, 60, 8358)


@pytest.mark.parametrize("rate", [0.25, 0.5, 0.67, 0.9])
def test_awkward_sizes_self_consistent(rate):
    for a in AWKWARD:
        p = compute_segmentation(a, rate)
        l = 24 if p.has_cb_crc else 0
        b_prime = p.b + p.c * l
        assert p.k_prime * p.c >= b_prime and p.k_prime * p.c - b_prime < p.c, a
        assert p.k >= p.k_prime and p.n_filler == p.k - p.k_prime
        assert p.n == (66 if p.bg == 0 else 50) * p.z
        assert p.has_cb_crc == (p.c > 1)


def test_rejects_out_of_range():
    for a in (0, MAX_TB_SIZE_BITS + 1):
        with pytest.raises(ValueError):
            compute_segmentation(a, 0.5)
    compute_segmentation(MAX_TB_SIZE_BITS, 0.5)


# --- CRCs and transport-block assembly in the oracle (crc.rs, segmentation.rs:361-410,
# transport_block.rs:963-1044) ---

def _ascii_bits(s):
    return [(b >> (7 - i)) & 1 for b in s for i in range(8)]


def test_crc_catalogue_check_values(oracle):
    # the published check values (CRC of ASCII "123456789") of the CRC-24/LTE-A,
    # CRC-24/LTE-B and CRC-16/XMODEM parameter sets: poly as crc.rs:84-92,
    # init 0, no reflection, no final XOR
    bits = _ascii_bits(b"123456789")
    assert oracle.crc(bits, 24, 0x864CFB) == 0xCDE703
    assert oracle.crc(bits, 24, 0x800063) == 0x23EF52
    assert oracle.crc(bits, 16, 0x11021) == 0x31C3


def test_crc_algebra(oracle):
    import numpy as np
    rng = np.random.default_rng(5)
    for poly in (oracle.CRC24A, oracle.CRC24B):
        assert oracle.crc(np.zeros(100, np.uint8), 24, poly) == 0            # all_zeros_produces_zero_remainder
        g = [1] + [(poly >> (23 - i)) & 1 for i in range(24)]                 # g(x) itself: x^24 g mod g = 0
        assert oracle.crc(g, 24, poly) == 0
        a, b = rng.integers(0, 2, 500).astype(np.uint8), rng.integers(0, 2, 500).astype(np.uint8)
        assert oracle.crc(a ^ b, 24, poly) == oracle.crc(a, 24, poly) ^ oracle.crc(b, 24, poly)  # linearity
        m = oracle.attach_crc(a, poly)
        assert oracle.crc(m[:-24], 24, poly) == int("".join(map(str, m[-24:])), 2)  # attach_then_check_roundtrip
        for i in (0, 17, 499):                                                       # single_bit_flip_detected
            m2 = m.copy()
            m2[i] ^= 1
            assert oracle.crc(m2[:-24], 24, poly) != int("".join(map(str, m2[-24:])), 2)
        assert oracle.crc([0] * 7 + list(a), 24, poly) == oracle.crc(a, 24, poly)   # leading zeros change nothing


def test_segment_round_trip_8425(oracle):
    # segment_round_trip_awkward_tb_size_8425 (segmentation.rs:606-645)
    import numpy as np
    a = 8425
    tb = np.array([1 if (i * 7 + 3) % 5 < 2 else 0 for i in range(a)], np.uint8)
    tbc = oracle.attach_crc(tb, oracle.CRC24A)
    p = compute_segmentation(a, 0.5)
    assert len(tbc) == p.b
    blocks = oracle.segment(tbc, p.c, p.k_prime)
    assert blocks.shape == (p.c, p.k_prime)
    pay = p.k_prime - 24
    rec = blocks[:, :pay].reshape(-1)
    assert np.array_equal(rec[:p.b], tbc) and not rec[p.b:].any()
    for blk in blocks:  # each block carries its CRC-24B
        assert oracle.crc(blk[:pay], 24, oracle.CRC24B) == int("".join(map(str, blk[pay:])), 2)


@pytest.mark.parametrize("a,rate", [(100, 0.5), (8425, 0.5), (30000, 2 / 3)])
def test_tb_check_passes_and_catches(oracle, a, rate):
    import numpy as np
    rng = np.random.default_rng(a)
    tb = rng.integers(0, 2, a).astype(np.uint8)
    p = compute_segmentation(a, rate)
    blocks = oracle.segment(oracle.attach_crc(tb, oracle.CRC24A), p.c, p.k_prime)
    stride = p.k_prime + 11  # blocks need not be packed
    hard = np.zeros(p.c * stride, np.uint8)
    offs = np.arange(p.c) * stride
    for i in range(p.c):
        hard[offs[i]:offs[i] + p.k_prime] = blocks[i]
    ok, cb_ok, out = oracle.tb_check(hard, offs, p.c, p.k_prime, a)
    assert ok and cb_ok.all() and np.array_equal(out, tb)
    hard[offs[-1] + 3] ^= 1  # one payload bit of the last block
    ok, cb_ok, out = oracle.tb_check(hard, offs, p.c, p.k_prime, a)
    assert not ok and out is None
    assert (cb_ok[-1] == 0) == (p.c > 1) and cb_ok[:-1].all()


def test_host_crc_and_segment_match_oracle(oracle):
    """The product's host CRC / segmentation (traffic building) against the
    oracle's restatements, which are pinned above."""
    import numpy as np
    from syndrome_amd.transport import CRC24A, CRC24B, attach_crc24, crc24, segment
    rng = np.random.default_rng(3)
    for n in (0, 1, 7, 8, 9, 100, 1001):
        bits = rng.integers(0, 2, n).astype(np.uint8)
        assert crc24(bits, CRC24A) == oracle.crc(bits, 24, CRC24A)
        assert crc24(bits, CRC24B) == oracle.crc(bits, 24, CRC24B)
    for a, r in ((100, 0.5), (8425, 0.5), (30000, 2 / 3), (200000, 5 / 6)):
        p = compute_segmentation(a, r)
        tb = rng.integers(0, 2, a).astype(np.uint8)
        tbc = attach_crc24(tb)
        assert np.array_equal(tbc, oracle.attach_crc(tb, oracle.CRC24A))
        assert np.array_equal(segment(tbc, p), oracle.segment(tbc, p.c, p.k_prime))
