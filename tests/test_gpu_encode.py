"""GPU parity of the device encoder (QcLdpcEncoder::encode,
src/qc_ldpc.rs:2592-2614; encode_sparse 2620-2727).

The device codewords must equal the oracle's encoder and the library's host
encoder byte for byte, and satisfy H c = 0, for every 3GPP lifting size of
both base graphs, every 802.11 matrix and a raw code with a non-trivial core
inverse; non-binary info bytes, unaligned buffers and a grid-stride launch
(more blocks than workgroups) take the other load/store paths.
"""
import numpy as np
import pytest

from tests.helpers import say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ZS = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40, 44, 48, 52, 56,
      60, 64, 72, 80, 88, 96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384]
WIFI = [(z, rn, rd) for z in (27, 54, 81) for (rn, rd) in ((1, 2), (2, 3), (3, 4), (5, 6))]
# m_b = 3, n_b = 5, Z = 5: core [[P1, I], [I, 0]] (inverse [[0, I], [I, P1]]), then one extension row
RAW = [(0, 0, 2), (0, 1, 4), (0, 2, 1), (0, 3, 0), (1, 0, 3), (1, 1, 1), (1, 2, 0), (2, 0, 1), (2, 3, 2), (2, 4, 0)]


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    import syndrome_amd
    syndrome_amd.lib()
    return torch.device("cuda:0")


def _info(rng, count, k):
    return rng.integers(0, 2, (count, k), dtype=np.uint8)


def _device_encode(dec, info):
    d_info = torch.from_numpy(info).cuda()
    d_cw = torch.full((info.shape[0], dec.variable_node_count()), 0xEE, dtype=torch.uint8, device="cuda")
    dec.encode_batch_device(d_info, d_cw)
    torch.cuda.synchronize()
    return d_cw.cpu().numpy()


def _check(oracle, code, dec, info, n_oracle):
    cw = _device_encode(dec, info)
    host = dec.encode_batch(info)
    assert np.array_equal(cw, host), "device codewords differ from the host encoder"
    for i in range(min(n_oracle, info.shape[0])):
        ref = oracle.encode(code, info[i])
        assert np.array_equal(cw[i], ref), f"block {i} differs from the oracle"
        assert oracle.syndrome_ok(code, cw[i])
    return cw


@pytest.mark.parametrize("bg", [0, 1])
def test_encode_every_lifting_size(dev, oracle, bg):
    from syndrome_amd import QcLdpcDecoder
    rng = np.random.default_rng(11 + bg)
    for z in ZS:
        say(f"[encode] BG{bg + 1} Z={z}")
        code = oracle.code_3gpp(bg, z)
        dec = QcLdpcDecoder.with_lifting_size(bg, z, 0.5)
        _check(oracle, code, dec, _info(rng, 9, code.k), 3)
        dec.close()


def test_encode_headline_batch(dev, oracle):
    """BG1 Z=384 and BG2 Z=208 at 256 blocks: every block against the oracle."""
    from syndrome_amd import QcLdpcDecoder
    rng = np.random.default_rng(5)
    for bg, z in ((0, 384), (1, 208)):
        say(f"[encode] BG{bg + 1} Z={z} x256")
        code = oracle.code_3gpp(bg, z)
        dec = QcLdpcDecoder.with_lifting_size(bg, z, 0.5)
        _check(oracle, code, dec, _info(rng, 256, code.k), 256)


@pytest.mark.parametrize("z,rn,rd", WIFI)
def test_encode_wifi(dev, oracle, z, rn, rd):
    from syndrome_amd import QcLdpcDecoder
    say(f"[encode] Wi-Fi Z={z} R={rn}/{rd}")
    code = oracle.code_wifi(z, rn, rd)
    dec = QcLdpcDecoder.wifi(z, rn, rd, 0.25)
    _check(oracle, code, dec, _info(np.random.default_rng(z * 10 + rn), 33, code.k), 33)


def test_encode_wifi_fixture_codeword(dev, oracle):
    """ldpc_convergence.json holds the transmitted Wi-Fi Z=27 R=1/2 codeword."""
    import json
    import os
    from syndrome_amd import QcLdpcDecoder
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ldpc_convergence.json")))
    cw = np.array(d["transmitted_codeword"], np.uint8)
    dec = QcLdpcDecoder.wifi(27, 1, 2, 0.25)
    k = dec.info_bit_count_5g()
    out = _device_encode(dec, np.ascontiguousarray(cw[None, :k]))
    assert np.array_equal(out[0], cw)


def test_encode_raw_code(dev, oracle):
    from syndrome_amd import QcLdpcDecoder
    say("[encode] raw code, 2-block core + extension row")
    e = np.array(RAW)
    code = oracle.code_raw(e[:, 0], e[:, 1], e[:, 2], 3, 5, 5)
    dec = QcLdpcDecoder.from_raw_edges(RAW, 3, 5, 5, 0.5)
    info = np.array([[int(b) for b in f"{v:010b}"] for v in range(1024)], np.uint8)
    cw = _check(oracle, code, dec, info, 1024)
    assert len({c.tobytes() for c in cw}) == 1024  # systematic: distinct info, distinct codewords


def test_encode_nonbinary_bytes_and_unaligned(dev, oracle):
    """Info bytes are copied through; parity uses bit 0.  Buffers offset by
    one byte take the byte-wise load/store path."""
    from syndrome_amd import QcLdpcDecoder
    rng = np.random.default_rng(3)
    dec = QcLdpcDecoder.with_lifting_size(0, 384, 0.5)
    k, n = dec.info_bit_count_5g(), dec.variable_node_count()
    info = rng.integers(0, 256, (17, k), dtype=np.uint8)
    ref = dec.encode_batch(info)
    assert np.array_equal(ref[:, :k], info)
    assert np.array_equal(_device_encode(dec, info), ref)
    d_in = torch.zeros(17 * k + 1, dtype=torch.uint8, device="cuda")
    d_out = torch.zeros(17 * n + 1, dtype=torch.uint8, device="cuda")
    d_in[1:] = torch.from_numpy(info.reshape(-1)).cuda()
    dec.encode_batch_device(d_in[1:], d_out[1:])
    torch.cuda.synchronize()
    assert np.array_equal(d_out[1:].cpu().numpy().reshape(17, n), ref)
    assert int(d_out[0]) == 0


def test_encode_grid_stride(dev, oracle):
    """More blocks than the launch has workgroups (65536)."""
    from syndrome_amd import QcLdpcDecoder
    say("[encode] BG2 Z=2 x70001")
    dec = QcLdpcDecoder.with_lifting_size(1, 2, 0.5)
    info = _info(np.random.default_rng(9), 70001, dec.info_bit_count_5g())
    assert np.array_equal(_device_encode(dec, info), dec.encode_batch(info))


def test_encode_zero_count_and_nulls(dev):
    import ctypes as C
    import syndrome_amd._lib as L
    from syndrome_amd import QcLdpcDecoder
    dec = QcLdpcDecoder.with_lifting_size(0, 384, 0.5)
    assert L.lib().syndrome_ldpc_encode_batch_device(dec.handle, None, None, 0, None) == L.OK
    buf = torch.zeros(64, dtype=torch.uint8, device="cuda")
    assert L.lib().syndrome_ldpc_encode_batch_device(dec.handle, None, C.c_void_p(buf.data_ptr()), 1,
                                                     None) == L.NULL_POINTER
    assert L.lib().syndrome_ldpc_encode_batch_device(None, None, None, 0, None) == L.NULL_POINTER
