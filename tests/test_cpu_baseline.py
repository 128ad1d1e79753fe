"""The CPU baseline (oracle/cpu_baseline.cpp) against the scalar oracle, CPU only.

The baseline restates the reference's AVX2 kernels (simd_avx2.rs:78-286,
464-507 f32; 546-847 int8) and its multi-worker pipeline
(ldpc_pipeline.rs:98-589).  bench.py times it as `cpu_baseline`, so it must
decode exactly as the oracle does: iteration counts, posteriors, R messages
and hard decisions, bit for bit, one frame at a time and through the
pipeline at one and several workers, over one code and over mixed codes.
The golden checksum a2448e8cf5272825 (bench/results/ldpc_{rust,cpp}.checksum)
pins the f32 kernels directly.
"""
import os

import numpy as np
import pytest

from tests.helpers import channel, codewords, f32_bits

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _code(oracle, spec):
    return oracle.code_wifi(*spec[1:]) if spec[0] == "wifi" else oracle.code_3gpp(*spec)


def test_avx2_f32_golden_checksum(oracle):
    """ldpc_bench_export.rs:239-250: +-0.5 alternating input, beta 0.25, 10
    iterations (never converges), FNV-1a of the hard bits."""
    want = int(open(os.path.join(GOLDEN, "ldpc_rust.checksum")).read().strip(), 16)
    c = oracle.code_3gpp(0, 384)
    llr = np.where(np.arange(c.n) % 2 == 0, 0.5, -0.5).astype(np.float32)
    it, post, er, hard = oracle.avx2_decode_f32(c, 0.25, llr, 10)
    assert it == 10 and oracle.fnv1a(hard) == want
    ri, rp, re, rh = oracle.decode_f32(c, 0.25, llr, 10)
    assert np.array_equal(f32_bits(post), f32_bits(rp)) and np.array_equal(f32_bits(er), f32_bits(re))


@pytest.mark.parametrize("spec", [(0, 384), (1, 208), (0, 15), (1, 13), (0, 52), ("wifi", 27, 1, 2),
                                  ("wifi", 81, 5, 6)], ids=str)
def test_avx2_f32_equals_oracle(oracle, spec):
    c = _code(oracle, spec)
    cws = codewords(oracle, c, 4, 31)
    llr = channel(oracle, c, cws, [-0.5, 1.0, 2.0, 4.0], 5)
    for b in range(4):
        x = oracle.init_5g_f32(c, llr[b], 0) if spec[0] != "wifi" else llr[b]
        for beta in (0.5, 0.25):
            it, post, er, hard = oracle.avx2_decode_f32(c, beta, x, 8)
            ri, rp, re, rh = oracle.decode_f32(c, beta, x, 8)
            assert it == ri, (spec, b, beta)
            assert np.array_equal(hard, rh)
            assert np.array_equal(f32_bits(post), f32_bits(rp)) and np.array_equal(f32_bits(er), f32_bits(re))


@pytest.mark.parametrize("spec", [(1, 208), (0, 384), (1, 13), (0, 52), ("wifi", 27, 1, 2)], ids=str)
@pytest.mark.parametrize("clamp", [32767, 400, 127])
def test_avx2_i8_equals_oracle(oracle, spec, clamp):
    """int8 kernels, including saturating posteriors (clamps 400 and 127)
    and lifting sizes that leave 16- and 32-lane tails."""
    c = _code(oracle, spec)
    cws = codewords(oracle, c, 3, 77)
    llr = channel(oracle, c, cws, [0.0, 1.5, 3.0], 9)
    bq = oracle.beta_q(0.5, 8.0)
    for b in range(3):
        app = oracle.quantize_i16(llr[b], 8.0)
        if spec[0] != "wifi":
            app = oracle.init_5g_i8(c, app, 0, clamp)
        it, ra, re, rh = oracle.avx2_decode_i8(c, bq, app, 8, clamp)
        oi, oa, oe, oh = oracle.decode_i8(c, 0.5, app, 8, 8.0, clamp)
        assert it == oi, (spec, clamp, b)
        assert np.array_equal(ra, oa) and np.array_equal(re, oe) and np.array_equal(rh, oh)


@pytest.mark.parametrize("workers", [1, 4])
def test_pipeline_equals_oracle(oracle, workers):
    """The f32 pipeline over one code (the headline CPU baseline's form):
    every frame's iterations and hard decisions equal the oracle's."""
    c = oracle.code_3gpp(0, 384)
    cws = codewords(oracle, c, 12, 3)
    llr = channel(oracle, c, cws, [1.0, 1.5, 2.0], 11)
    llr[:, :2 * c.z] = 0.0
    sec, its, hard = oracle.pipeline_run(c, 0.5, llr, 8, workers)
    assert sec > 0
    for b in range(12):
        ri, _, _, rh = oracle.decode_f32(c, 0.5, oracle.init_5g_f32(c, llr[b], 0), 8)
        assert its[b] == ri and np.array_equal(hard[b], rh), b


@pytest.mark.parametrize("int8", [False, True], ids=["f32", "int8"])
@pytest.mark.parametrize("workers", [1, 3])
def test_pipeline_mixed_codes_equals_oracle(oracle, int8, workers):
    """Mixed base graphs, lifting sizes and filler counts in one stream, each
    frame decoded as alone, with frames repeated in the stream order."""
    specs = [(0, 384, 0), (1, 208, 16), (0, 13, 0), (1, 56, 40), (0, 120, 8)]
    codes = [oracle.code_3gpp(bg, z) for bg, z, _ in specs]
    frame_code, offs, nfill, xs = [], [], [], []
    off = 0
    for f in range(10):
        ci = f % len(specs)
        c = codes[ci]
        cw = codewords(oracle, c, 1, 100 + f)[0]
        x = oracle.awgn(1.0 + 0.25 * f, c.k / c.n, 500 + f, cw)
        frame_code.append(ci)
        offs.append(off)
        nfill.append(specs[ci][2])
        xs.append(x)
        off += c.n
    arena = np.concatenate(xs)
    order = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 3, 7, 0], np.int32)
    sec, its, hard = oracle.pipeline_mixed(codes, frame_code, offs, arena, 8, workers, order=order, int8=int8,
                                           n_filler=nfill)
    for f in range(10):
        c = codes[frame_code[f]]
        x = arena[offs[f]:offs[f] + c.n]
        if int8:
            ri, _, _, rh = oracle.decode_i8(c, 0.5, oracle.init_5g_i8(c, oracle.quantize_i16(x, 8.0), nfill[f]), 8)
        else:
            ri, _, _, rh = oracle.decode_f32(c, 0.5, oracle.init_5g_f32(c, x, nfill[f]), 8)
        assert its[f] == ri, f
        assert np.array_equal(hard[offs[f]:offs[f] + c.n], rh), f
