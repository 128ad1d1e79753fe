"""GPU parity of the pair-packed int8 kernel (BG2: two codewords per lane,
one per 16-bit half), through the C ABI, against the oracle's restatement
of qc_ldpc.rs:1530-1787: iterations, app, edge_r and hard decisions equal.

Cases the packing adds to the int8 tests of test_gpu_parity.py:
* odd batch sizes, so a workgroup's last pair has one live codeword;
* the two codewords of a pair stopping at different iterations (a stopped
  half must keep its R and posterior while the other half goes on);
* posteriors at the edge of the 16-bit range with app_clamp 32767, where
  A - R_old and A + delta saturate in 16 bits (the reference's Q-build
  clamp and saturating add);
* misaligned device buffers (the element-wise load and store path);
* int16 input without the fused quantizer, at several lifting sizes.
"""
import numpy as np
import pytest

from tests.helpers import channel, codewords, say

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    import syndrome_amd
    syndrome_amd.lib()
    return torch.device("cuda:0")


def _decode_i16(dec, app_in, iters, clamp, misalign=0):
    """int16 input decode with every output; misalign shifts each buffer
    by that many elements off its 16-byte alignment."""
    import syndrome_amd._lib as L
    from syndrome_amd import QuantParams
    B, N = app_in.shape
    EZ = dec.required_edge_buffer()

    def buf(dtype, n, src=None):
        raw = torch.zeros(B * n + misalign, dtype=dtype, device="cuda")
        v = raw[misalign:].view(B, n)
        if src is not None:
            v.copy_(torch.from_numpy(src))
        return v

    d_in = buf(torch.int16, N, app_in)
    d_post = buf(torch.int16, N)
    d_hard = buf(torch.uint8, N)
    d_edge = buf(torch.int8, EZ)
    d_it = torch.empty(B, dtype=torch.int32, device="cuda")
    flags = L.WRITE_POST | L.WRITE_HARD | L.WRITE_EDGE
    dec.decode_batch_i8(d_in, d_post, d_hard, d_it, iters, flags=flags, d_edge=d_edge, quant=QuantParams(8.0, clamp))
    torch.cuda.synchronize()
    return d_post.cpu().numpy(), d_hard.cpu().numpy(), d_edge.cpu().numpy(), d_it.cpu().numpy()


def _check(oracle, code, app_in, out, iters, clamp, tag):
    post, hard, edge, it = out
    for b in range(app_in.shape[0]):
        ri, ra, re, rh = oracle.decode_i8(code, 0.5, app_in[b], iters, 8.0, clamp)
        assert it[b] == ri, f"{tag} codeword {b}: iterations {it[b]} != {ri}"
        assert np.array_equal(post[b], ra), f"{tag} codeword {b}: app differs"
        assert np.array_equal(edge[b], re), f"{tag} codeword {b}: edge_r differs"
        assert np.array_equal(hard[b], rh), f"{tag} codeword {b}: hard differs"
    return it


@pytest.mark.parametrize("z", [208, 13, 52, 2, 384])
def test_pack_odd_batch_and_split_pairs(dev, oracle, z):
    """An odd batch whose pairs mix a strong and a weak codeword: the two
    halves of a pair stop at different iterations."""
    from syndrome_amd import QcLdpcDecoder
    code = oracle.code_3gpp(1, z)
    dec = QcLdpcDecoder.with_lifting_size(1, z, 0.5)
    cws_per_group, _, _ = dec.launch_shape(int8=True)
    assert cws_per_group % 2 == 0
    B = cws_per_group + 3 if cws_per_group > 2 else 7
    say(f"[pack odd batch] BG2 Z={z}: {B} codewords, {cws_per_group} per workgroup")
    cw = codewords(oracle, code, B, seed=17 * z)
    ebno = [4.0 if b % 2 == 0 else 0.3 for b in range(B)]
    llr = channel(oracle, code, cw, ebno, seed=23 * z)
    app_in = np.stack([oracle.init_5g_i8(code, oracle.quantize_i16(llr[b], 8.0), 0, 32767) for b in range(B)])
    for clamp in (32767, 127):
        app = np.clip(app_in, -clamp, clamp).astype(np.int16)
        it = _check(oracle, code, app, _decode_i16(dec, app, 8, clamp), 8, clamp, f"Z={z} clamp {clamp}")
        say(f"[pack odd batch] clamp {clamp}: iterations {it.tolist()}")
    assert len(set(it[0::2].tolist()) | set(it[1::2].tolist())) > 1


@pytest.mark.parametrize("z", [208, 26])
def test_pack_sixteen_bit_edges(dev, oracle, z):
    """Posteriors near +-32767 with app_clamp 32767: A - R_old leaves the
    16-bit range and A + delta saturates; small values elsewhere keep the
    checks failing so the decode runs every iteration."""
    from syndrome_amd import QcLdpcDecoder
    code = oracle.code_3gpp(1, z)
    dec = QcLdpcDecoder.with_lifting_size(1, z, 0.5)
    B = 6
    cw = codewords(oracle, code, B, seed=5 * z + 1)
    rng = np.random.default_rng(z)
    mag = rng.integers(1, 40, size=cw.shape)
    app = np.where(cw == 0, mag, -mag).astype(np.int64)
    big = rng.random(cw.shape) < 0.15
    app[big] = np.where(cw[big] == 0, 1, -1) * rng.integers(32600, 32768, size=int(big.sum()))
    wrong = rng.random(cw.shape) < 0.08
    app[wrong] = -app[wrong]  # some saturated values with the wrong sign
    app_in = np.clip(app, -32767, 32767).astype(np.int16)
    it = _check(oracle, code, app_in, _decode_i16(dec, app_in, 6, 32767), 6, 32767, f"Z={z} 16-bit edges")
    say(f"[pack 16-bit edges] Z={z}: iterations {it.tolist()}")


@pytest.mark.parametrize("misalign", [1, 3])
def test_pack_misaligned_buffers(dev, oracle, misalign):
    """Buffers off their 16-byte alignment take the element-wise load and
    store paths of the packed kernel."""
    from syndrome_amd import QcLdpcDecoder
    z = 208
    code = oracle.code_3gpp(1, z)
    dec = QcLdpcDecoder.with_lifting_size(1, z, 0.5)
    B = 5
    cw = codewords(oracle, code, B, seed=99 + misalign)
    llr = channel(oracle, code, cw, [0.5, 1.0, 2.5, 0.0, 3.0], seed=7 + misalign)
    app_in = np.stack([oracle.init_5g_i8(code, oracle.quantize_i16(llr[b], 8.0), 0, 32767) for b in range(B)])
    _check(oracle, code, app_in, _decode_i16(dec, app_in, 8, 32767, misalign), 8, 32767, f"misalign {misalign}")
