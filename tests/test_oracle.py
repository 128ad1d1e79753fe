"""The oracle, pinned against what the reference itself holds (CPU only).

Pins, by part of the oracle:
* f32 decode (qc_ldpc.rs:962-1252): the cross-language golden checksum
  a2448e8cf5272825 (bench/results/ldpc_{rust,cpp}.checksum); the full
  14-frame hard-decision trajectory and the 13-iteration early exit of
  bench/results/ldpc_convergence.json; the exactly-5-iteration decode of the
  pipeline benchmark codeword (ldpc_pipeline_rust.json); and, where the
  reference tree is mounted, value-for-value equality of posteriors and R
  messages with the reference's own C++ port (bench/cpp/loms_decode.cpp).
* channel model (channel_sim.rs): frame 0 of the convergence fixture is the
  raw channel hard decision, and the transmitted codeword is stored there.
* encoder: the fixture's transmitted codeword; H c = 0 for every code.
* quantizer (quantize.rs): the known answers of quantize.rs:372-437.
* int8 decode (qc_ldpc.rs:1530-1787): no reference vector exists for it.
  It is pinned by the exact identity between the two number formats: on
  integer-valued inputs whose messages never reach the i8 clamp, with an
  integer offset, every f32 operation of the f32 decoder is exact, so the
  int8 decoder must return the same posteriors, messages, hard decisions
  and iteration counts as the (reference-pinned) f32 decoder.  A dropped
  term, a wrong sign or a misplaced clamp breaks that identity.  The
  clamp paths are pinned by a brute-force one-row pass with every message
  and the posterior in their clamps (it tells MSG_MAX 127 from 128), by
  saturation invariants and by the reference's own documented behaviour
  (strong channel -> all-zero, filler pinned).
* raw-edge codes (from_raw_edges, qc_ldpc.rs:412-455): a hand-worked layer
  with a repeated column (Q from the pre-layer posterior, deltas applied in
  edge order), f32 and int8; empty rows and unused columns are inert (a 3GPP
  code padded with them decodes bit for bit as the code itself).
"""
import json
import os

import numpy as np
import pytest

from tests.helpers import channel, codewords, f32_bits

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _golden_checksum():
    a = open(os.path.join(GOLDEN, "ldpc_rust.checksum")).read().strip()
    b = open(os.path.join(GOLDEN, "ldpc_cpp.checksum")).read().strip()
    assert a == b
    return int(a, 16)


def test_tables_dimensions(oracle):
    # qc_ldpc.rs:2830-2845 and the JSON's own counts
    c1 = oracle.code_3gpp(0, 384)
    assert (c1.m_b, c1.n_b, c1.n_edges, c1.max_deg) == (46, 68, 316, 19)
    assert (c1.n, c1.k, c1.edge_buf, c1.layer_buf) == (26112, 8448, 121344, 7296)
    c2 = oracle.code_3gpp(1, 208)
    assert (c2.m_b, c2.n_b, c2.n_edges, c2.max_deg) == (42, 52, 197, 10)
    assert (c2.n, c2.k, c2.edge_buf) == (10816, 2080, 40976)
    w = oracle.code_wifi(27, 1, 2)
    assert (w.m_b, w.n_b, w.n_edges, w.n, w.k) == (12, 24, 88, 648, 324)
    with pytest.raises(ValueError):
        oracle.code_3gpp(0, 383)
    with pytest.raises(ValueError):
        oracle.code_3gpp(0, 0)


def test_golden_checksum(oracle):
    c = oracle.code_3gpp(0, 384)
    llr = np.where(np.arange(c.n) % 2 == 1, -0.5, 0.5).astype(np.float32)
    it, post, er, hard = oracle.decode_f32(c, 0.25, llr, 10)
    assert it == 10
    assert oracle.fnv1a(hard) == _golden_checksum()


def test_golden_checksum_catches_a_planted_mistake(oracle):
    """The checksum is sensitive: the same input at beta 0.5 (a plausible
    mistake: the default offset instead of the benchmark's) misses it."""
    c = oracle.code_3gpp(0, 384)
    llr = np.where(np.arange(c.n) % 2 == 1, -0.5, 0.5).astype(np.float32)
    _, _, _, hard = oracle.decode_f32(c, 0.5, llr, 10)
    assert oracle.fnv1a(hard) != _golden_checksum()


def test_convergence_fixture(oracle):
    d = json.load(open(os.path.join(GOLDEN, "ldpc_convergence.json")))
    assert d["code"]["z"] == 27 and d["code"]["rate_num"] == 1 and d["code"]["rate_den"] == 2
    w = oracle.code_wifi(27, 1, 2)
    info = oracle.knuth_info(w.k)
    cw = oracle.encode(w, info)
    assert np.array_equal(cw, np.array(d["transmitted_codeword"], np.uint8))
    llr = oracle.awgn(d["channel"]["eb_n0_db"], d["channel"]["code_rate"], d["channel"]["seed"], cw)
    assert np.array_equal((llr < 0).astype(np.uint8), np.array(d["frames"][0]["hard_bits"], np.uint8))
    it, post, er, hard, tr = oracle.decode_f32(w, d["code"]["offset_beta"], llr, d["max_iterations"], trace=True)
    assert it == d["iterations_used"] == 13
    for k in range(it):
        f = d["frames"][k + 1]
        assert f["iteration"] == k + 1
        assert np.array_equal((tr[k] < 0).astype(np.uint8), np.array(f["hard_bits"], np.uint8)), k
    assert d["frames"][-1]["bit_errors"] == 0 and np.array_equal(hard, cw)


def test_pipeline_codeword_exactly_five_iterations(oracle):
    c = oracle.code_3gpp(0, 384)
    cw = oracle.encode(c, oracle.knuth_info(c.k))
    llr = oracle.awgn(2.0, c.k / c.n, 42, cw)
    it, post, er, hard = oracle.decode_f32(c, 0.25, llr, 10)
    assert it == 5
    assert np.array_equal(hard, cw)


@pytest.mark.skipif(not os.path.exists("/root/reference/bench/cpp/loms_decode.cpp"),
                    reason="reference tree not mounted")
def test_matches_reference_cpp_port_value_for_value(oracle):
    import subprocess
    if not oracle.ref_available():
        subprocess.run(["make", "-s", "-C", os.path.dirname(oracle.LIB_PATH), "ref"], check=True)
    js = open("/root/reference/data/bg_tables.json").read()
    c = oracle.code_3gpp(0, 384)
    rng = np.random.default_rng(7)
    inputs = [np.where(np.arange(c.n) % 2 == 1, -0.5, 0.5).astype(np.float32),
              rng.uniform(-3, 3, c.n).astype(np.float32)]
    cw = oracle.encode(c, oracle.knuth_info(c.k))
    inputs.append(oracle.awgn(1.0, c.k / c.n, 5, cw))
    inputs.append(oracle.awgn(0.3, c.k / c.n, 6, cw))
    for x in inputs:
        rl, rer = oracle.ref_decode_bg1_z384_10it(js, x)
        # the port runs exactly 10 iterations with no early exit; run the
        # oracle without its exit by checking it did not converge earlier,
        # or compare only when it ran the full budget
        it, post, er, hard = oracle.decode_f32(c, 0.25, x, 10)
        if it == 10:
            assert np.array_equal(f32_bits(post), f32_bits(rl))
            assert np.array_equal(f32_bits(er), f32_bits(rer))
        else:
            # converged early: the port keeps iterating on a valid codeword;
            # hard decisions must still agree
            assert np.array_equal(hard, (rl < 0).astype(np.uint8))


def test_encoder_satisfies_parity(oracle):
    for bg, z in ((0, 2), (0, 384), (1, 3), (1, 208), (0, 15)):
        c = oracle.code_3gpp(bg, z)
        cw = codewords(oracle, c, 2, seed=z)
        for x in cw:
            assert oracle.syndrome_ok(c, x)
    for z, rn, rd in ((27, 1, 2), (54, 2, 3), (81, 3, 4), (27, 5, 6), (81, 1, 2)):
        w = oracle.code_wifi(z, rn, rd)
        x = oracle.encode(w, oracle.xorshift_bits(z * rn, w.k))
        assert oracle.syndrome_ok(w, x)


def test_quantizer_known_answers(oracle):
    # quantize.rs:325-329, 372-437
    assert list(oracle.quantize_i16(np.array([0, 1, -1, 1e9, -np.inf], np.float32), 8.0)) == [0, 8, -8, 127, -127]
    assert list(oracle.quantize_i16(np.array([np.nan], np.float32), 8.0)) == [0]
    assert list(oracle.quantize_i8(np.array([np.nan, 1e9, -1e9, -np.inf], np.float32), 8.0)) == [0, 127, -127, -127]
    assert list(oracle.quantize_i16(np.array([1.0, -0.25, 99.0], np.float32), 8.0)) == [8, -2, 127]
    # ties away from zero (f32::round), not to even
    assert list(oracle.quantize_i16(np.array([0.0625, -0.0625, 0.1875], np.float32), 8.0)) == [1, -1, 2]
    # quantize_llr / dequantize_llr doc example (quantize.rs:123-133): 200 saturates at 127
    q = oracle.quantize_i8(np.array([1.5, -3.0, 0.0, 200.0], np.float32), 8.0)
    assert list(q) == [12, -24, 0, 127]
    assert list(oracle.dequantize_i8(q, 8.0)) == [1.5, -3.0, 0.0, 15.875]
    # round trip on the grid: every multiple of 1/8 in [-15.875, 15.875] comes back
    grid = (np.arange(-127, 128) / 8.0).astype(np.float32)
    assert np.array_equal(oracle.dequantize_i8(oracle.quantize_i8(grid, 8.0), 8.0), grid)
    # 1/scale is one f32 division, then one multiply: scale 3 gives q * 0.33333334f
    assert oracle.dequantize_i8(np.array([3], np.int8), 3.0)[0] == np.float32(3) * (np.float32(1) / np.float32(3))
    assert oracle.beta_q(0.5) == 4 and oracle.beta_q(0.125) == 1 and oracle.beta_q(0.0) == 0
    assert oracle.beta_q(-1.0) == 0 and oracle.beta_q(1e6) == 127
    assert oracle.beta_q(float("inf")) == 127 and oracle.beta_q(float("nan")) == 127


def _int_inputs(oracle, code, seed, lo=-12, hi=12):
    rng = np.random.default_rng(seed)
    cw = codewords(oracle, code, 1, seed)[0]
    mag = rng.integers(lo, hi + 1, code.n)
    return np.where(cw == 1, -1, 1) * np.abs(mag) * np.where(rng.random(code.n) < 0.15, -1, 1)


@pytest.mark.parametrize("bg,z,beta", [(0, 8, 1.0), (1, 13, 0.0), (1, 52, 2.0), (0, 44, 1.0)])
def test_i8_equals_f32_on_exact_integer_inputs(oracle, bg, z, beta):
    """int8 decode == f32 decode when nothing saturates (see module doc)."""
    c = oracle.code_3gpp(bg, z)
    checked = 0
    for seed in range(6):
        x = _int_inputs(oracle, c, seed, -6, 6)
        for iters in (1, 2, 3):
            it_f, post_f, er_f, hard_f = oracle.decode_f32(c, beta, x.astype(np.float32), iters)
            # the identity needs |q| <= 127 everywhere, which holds when the
            # posterior never exceeds 127 - max|R| ... check the outcome bound
            if np.abs(post_f).max() > 100 or np.abs(er_f).max() > 100:
                continue
            it_i, app, er_i, hard_i = oracle.decode_i8(c, beta, x.astype(np.int16), iters, scale=1.0)
            assert it_i == it_f
            assert np.array_equal(app.astype(np.float32), post_f)
            assert np.array_equal(er_i.astype(np.float32), er_f)
            assert np.array_equal(hard_i, hard_f)
            checked += 1
    assert checked >= 6


def test_i8_identity_catches_planted_mistakes(oracle):
    """The identity above is not vacuous: a wrong offset or a flipped input
    breaks it."""
    c = oracle.code_3gpp(1, 13)
    x = _int_inputs(oracle, c, 3, -6, 6)
    _, post_f, _, _ = oracle.decode_f32(c, 1.0, x.astype(np.float32), 2)
    _, app, _, _ = oracle.decode_i8(c, 2.0, x.astype(np.int16), 2, scale=1.0)
    assert not np.array_equal(app.astype(np.float32), post_f)
    y = x.copy()
    y[5] = -y[5] if y[5] else 3
    _, app2, _, _ = oracle.decode_i8(c, 1.0, y.astype(np.int16), 2, scale=1.0)
    assert not np.array_equal(app2.astype(np.float32), post_f)


def test_i8_saturation_invariants(oracle):
    c = oracle.code_3gpp(0, 16)
    cw = codewords(oracle, c, 1, 11)[0]
    llr = channel(oracle, c, cw[None, :], [0.0], 4)[0]
    for clamp in (127, 32767, 300, 0):
        app0 = oracle.quantize_i16(llr * 30.0, 8.0)  # saturating channel
        it, app, er, hard = oracle.decode_i8(c, 0.5, app0, 10, 8.0, clamp)
        assert np.abs(app).max() <= clamp
        assert np.abs(er.astype(np.int32)).max() <= 127
        assert np.array_equal(hard, (app < 0).astype(np.uint8))


def test_i8_reference_documented_behaviour(oracle):
    # qc_ldpc.rs:1322-1340 doc example: strong all-zero channel -> all zero
    c = oracle.code_3gpp(0, 16)
    app = oracle.quantize_i16(np.full(c.n, 5.0, np.float32), 8.0)
    it, a, er, hard = oracle.decode_i8(c, 0.5, app, 10)
    assert not hard.any()
    # decode_5g_i8 pins filler at app_clamp, punctured at 0 (qc_ldpc.rs:1505-1514)
    c2 = oracle.code_3gpp(1, 16)
    app = oracle.init_5g_i8(c2, np.full(c2.n, 8, np.int16), 7, 32767)
    assert (app[: 2 * 16] == 0).all() and (app[c2.k - 7: c2.k] == 32767).all()


def test_f32_closed_forms(oracle):
    # all-zero codeword, strong LLRs: BG2 Z=2, 50 iterations -> <= 5 its, all zero (qc_ldpc.rs:2863-2885)
    c = oracle.code_3gpp(1, 2)
    it, post, er, hard = oracle.decode_f32(c, 0.25, np.full(c.n, 10.0, np.float32), 50)
    assert it <= 5 and not hard.any()
    # Wi-Fi all-zero codeword converges in exactly one iteration (wifi_ldpc_integration.rs:132-155)
    for z, rn, rd in ((27, 1, 2), (54, 2, 3), (81, 3, 4), (27, 5, 6)):
        w = oracle.code_wifi(z, rn, rd)
        it, _, _, hard = oracle.decode_f32(w, 0.25, np.full(w.n, 5.0, np.float32), 10)
        assert it == 1 and not hard.any()
    # +-5 round trip and one flipped LLR corrected (tests/ldpc_integration.rs:35-187)
    for bg in (0, 1):
        c = oracle.code_3gpp(bg, 16)
        cw = codewords(oracle, c, 1, 21)[0]
        llr = np.where(cw == 1, -5.0, 5.0).astype(np.float32)
        llr[40] = -llr[40]
        it, _, _, hard = oracle.decode_f32(c, 0.25, llr, 20)
        assert np.array_equal(hard, cw)
    # iterations = 0: hard = sign of the input, R zeroed
    c = oracle.code_3gpp(0, 4)
    x = np.linspace(-3, 3, c.n).astype(np.float32)
    it, post, er, hard = oracle.decode_f32(c, 0.25, x, 0)
    assert it == 0 and np.array_equal(post, x) and not er.any()


def test_f32_min_sum_by_brute_force(oracle):
    """A single layered pass on a one-row toy code equals the textbook
    check update computed by brute force (excluded-edge minimum and sign
    product, offset, floor at zero)."""
    z = 3
    rows, cols, shifts = [0, 0, 0, 0], [0, 1, 2, 3], [0, 1, 2, 0]
    c = oracle.code_raw(rows, cols, shifts, 1, 4, z)
    rng = np.random.default_rng(1)
    x = rng.normal(0, 2, c.n).astype(np.float32)
    beta = np.float32(0.3)
    it, post, er, hard = oracle.decode_f32(c, float(beta), x, 1)
    for zi in range(z):
        vs = [cols[e] * z + (zi + shifts[e]) % z for e in range(4)]
        q = [x[v] for v in vs]
        for e in range(4):
            others = [q[j] for j in range(4) if j != e]
            mag = max(np.float32(min(abs(o) for o in others)) - beta, np.float32(0))
            sign = -1 if sum(o < 0 for o in others) % 2 else 1
            assert er[e * z + zi] == np.float32(sign * mag)
            assert post[vs[e]] == np.float32(x[vs[e]] + np.float32(sign * mag))


def test_i8_min_sum_by_brute_force_at_saturation(oracle):
    """One layered int8 pass on a one-row toy code, every message and the
    posterior driven into their clamps, equals the check update computed by
    brute force (qc_ldpc.rs:1619-1744 read as the textbook rule): Q clamped
    to +-127 (MSG_MAX), excluded-edge minimum of |Q| and sign product, offset
    beta_q, floor at zero, then the posterior clamped to +-app_clamp."""
    z = 4
    rows, cols, shifts = [0, 0, 0, 0, 0], [0, 1, 2, 3, 4], [0, 1, 2, 3, 1]
    c = oracle.code_raw(rows, cols, shifts, 1, 5, z)
    rng = np.random.default_rng(8)
    beta_q = 4  # round(0.5 * 8)
    for clamp in (32767, 200, 127):
        # mostly saturating inputs: |app| >= 127 on most edges, some small
        x = rng.choice([-1000, -300, -128, -127, -126, 5, 126, 127, 128, 190, 640], c.n).astype(np.int16)
        x = np.clip(x, -clamp, clamp).astype(np.int16)
        it, post, er, hard = oracle.decode_i8(c, 0.5, x, 1, 8.0, clamp)
        assert it == 1
        for zi in range(z):
            vs = [cols[e] * z + (zi + shifts[e]) % z for e in range(5)]
            q = [max(-127, min(127, int(x[v]))) for v in vs]
            for e in range(5):
                others = [q[j] for j in range(5) if j != e]
                mag = max(min(abs(o) for o in others) - beta_q, 0)
                r = -mag if sum(o < 0 for o in others) % 2 else mag
                assert er[e * z + zi] == r, (clamp, zi, e)
                assert post[vs[e]] == max(-clamp, min(clamp, int(x[vs[e]]) + r)), (clamp, zi, e)


def test_i8_saturation_brute_force_catches_a_wrong_message_clamp(oracle):
    """The brute force above separates MSG_MAX = 127 from 128: with every
    input past the clamp the excluded minimum is 127, so R = 127 - 4."""
    z = 1
    c = oracle.code_raw([0, 0, 0], [0, 1, 2], [0, 0, 0], 1, 3, z)
    _, post, er, _ = oracle.decode_i8(c, 0.5, np.array([1000, 1000, -1000], np.int16), 1, 8.0, 32767)
    assert list(er) == [-123, -123, 123]
    assert list(post) == [877, 877, -877]


def test_raw_repeated_column_worked_example(oracle):
    """from_raw_edges keeps a repeated (row, col) (qc_ldpc.rs:412-455 has no
    check against it).  Worked by hand from qc_ldpc.rs:1072-1204 on a Z=2
    layer holding column 0 twice (shifts 0 and 1) and column 1 once: row 0
    reads variables (0, 1, 2), row 1 reads (1, 0, 3).  Every Q comes from the
    posterior before the layer, and the deltas land edge by edge, so
    variable 0 gets 1 + (-1.5) + 0.25 and variable 1 gets 2 + 0.25 - 0.5.
    Reading the posterior as it is being updated would flip row 1's Q on
    edge 1 (to -0.5), and letting the last writer win would leave variable 1
    at 1.5: both fail here."""
    c = oracle.code_raw([0, 0, 0], [0, 0, 1], [0, 1, 0], 1, 2, 2)
    x = np.array([1.0, 2.0, -3.0, 0.75], np.float32)
    it, post, er, hard = oracle.decode_f32(c, 0.5, x, 1)
    # row 0: Q = (1, 2, -3), one negative, min1 1, min2 2; row 1: Q = (2, 1, 0.75), min1 0.75, min2 1
    assert it == 1
    assert list(er) == [-1.5, 0.25, -0.5, 0.25, 0.5, 0.5]
    assert list(post) == [-0.25, 1.75, -2.5, 1.25]
    assert list(hard) == [1, 0, 1, 0]
    # the int8 twin, app = 8 * x, beta_q = 4
    it, app, er8, hard8 = oracle.decode_i8(c, 0.5, np.array([8, 16, -24, 6], np.int16), 1)
    assert it == 1
    assert list(er8) == [-12, 2, -4, 2, 4, 4]
    assert list(app) == [-2, 14, -20, 10]


def test_raw_empty_rows_and_columns_change_nothing(oracle):
    """Rows past the last used one and columns no edge names are inert
    (an empty layer has no Q, no update, and its checks hold): BG2 Z=4 given
    as raw edges with 7 extra rows and 3 extra columns decodes exactly as
    the 3GPP code on its variables, and leaves the extra ones untouched."""
    base = oracle.code_3gpp(1, 4)
    rows, cols, shifts = [], [], []
    for r in range(base.m_b):
        for e in range(base.layer_off[r], base.layer_off[r + 1]):
            rows.append(r)
            cols.append(base.cols[e])
            shifts.append(base.shifts[e])
    wide = oracle.code_raw(rows, cols, shifts, base.m_b + 7, base.n_b + 3, 4)
    rng = np.random.default_rng(5)
    for it_max in (1, 6):
        x = rng.normal(1.0, 2.0, wide.n).astype(np.float32)
        i0, p0, e0, h0 = oracle.decode_f32(base, 0.5, x[: base.n], it_max)
        i1, p1, e1, h1 = oracle.decode_f32(wide, 0.5, x, it_max)
        assert i0 == i1
        assert np.array_equal(f32_bits(p1[: base.n]), f32_bits(p0)) and np.array_equal(p1[base.n:], x[base.n:])
        assert np.array_equal(f32_bits(e1), f32_bits(e0))
        a = oracle.quantize_i16(x, 8.0)
        j0, q0, r0, _ = oracle.decode_i8(base, 0.5, a[: base.n], it_max)
        j1, q1, r1, _ = oracle.decode_i8(wide, 0.5, a, it_max)
        assert j0 == j1 and np.array_equal(q1[: base.n], q0) and np.array_equal(q1[base.n:], a[base.n:])
        assert np.array_equal(r1, r0)
