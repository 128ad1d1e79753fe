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
