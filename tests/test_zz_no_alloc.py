"""The allocation-free decode (syndrome_ldpc_decode_5g_static) against the
decode it varies (syndrome_ldpc_decode_5g), through the C ABI, on the GPU.

Two calls in a row on one device, the second with a smaller code (BG1 Z=384,
then BG2 Z=52, then BG1 Z=384 again), each bit-exact with decode_5g on the
same input: posterior, R messages, hard decisions and iteration count.  The
array is shared by every call on a GPU, so a result that leaked from the
previous, larger call would show here.
"""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import channel, codewords, say

pytestmark = pytest.mark.gpu


def _decode(L, fn, dec, llr, n_filler, iters, stream=None):
    n = L.syndrome_ldpc_decoder_variable_node_count(dec)
    ez = L.syndrome_ldpc_decoder_required_edge_buffer(dec)
    lb = L.syndrome_ldpc_decoder_required_layer_buffer(dec)
    x = np.array(llr, np.float32, copy=True)
    er = np.full(ez, np.nan, np.float32)
    sc = np.zeros(lb, np.float32)
    hard = np.full(n, 7, np.uint8)
    used = C.c_size_t(99)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    args = [dec, p(x), x.size, n_filler, p(er), er.size, p(sc), sc.size, p(hard), hard.size, iters, C.byref(used)]
    if fn == "syndrome_ldpc_decode_5g_static":
        args.append(stream)
    rc = getattr(L, fn)(*args)
    return rc, x, er, hard, used.value


def test_static_decode_equals_decode_5g(oracle):
    import torch
    import syndrome_amd
    L = syndrome_amd.lib()
    torch.cuda.init()
    stream = torch.cuda.Stream()
    cases = [(0, 384, 0, 1.5), (1, 52, 24, 1.0), (0, 384, 0, 0.5)]
    for i, (bg, z, nf, ebno) in enumerate(cases):
        say(f"[static] call {i}: BG{bg + 1} Z={z} n_filler={nf} {ebno} dB")
        code = oracle.code_3gpp(bg, z)
        llr = channel(oracle, code, codewords(oracle, code, 1, 40 + i), [ebno], 90 + i)[0]
        dec = C.c_void_p()
        assert L.syndrome_ldpc_decoder_create(bg, z, 0.5, C.byref(dec)) == 0
        try:
            # on the default stream, then on a caller's stream
            for s in (None, stream.cuda_stream):
                a = _decode(L, "syndrome_ldpc_decode_5g_static", dec, llr, nf, 8, s)
                b = _decode(L, "syndrome_ldpc_decode_5g", dec, llr, nf, 8)
                assert a[0] == 0 and b[0] == 0, (a[0], b[0])
                assert a[4] == b[4], (i, a[4], b[4])
                assert np.array_equal(a[3], b[3])
                assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
                assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
        finally:
            L.syndrome_ldpc_decoder_destroy(dec)


def test_static_decode_validates_like_decode_5g():
    import syndrome_amd
    L = syndrome_amd.lib()
    dec = C.c_void_p()
    assert L.syndrome_ldpc_decoder_create(1, 52, 0.5, C.byref(dec)) == 0
    try:
        n = L.syndrome_ldpc_decoder_variable_node_count(dec)
        x = np.zeros(n, np.float32)
        buf = np.zeros(n * 8, np.float32)
        h = np.zeros(n, np.uint8)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        f = L.syndrome_ldpc_decode_5g_static
        assert f(None, p(x), n, 0, p(buf), buf.size, p(buf), buf.size, p(h), n, 8, None, None) == -5
        assert f(dec, p(x), n - 1, 0, p(buf), buf.size, p(buf), buf.size, p(h), n, 8, None, None) == -1
        assert f(dec, p(x), n, 0, p(buf), buf.size, p(buf), buf.size, p(h), n - 1, 8, None, None) == -1
        assert f(dec, p(x), n, 0, p(buf), 10, p(buf), buf.size, p(h), n, 8, None, None) == -4
    finally:
        L.syndrome_ldpc_decoder_destroy(dec)
