"""GPU parity of the traced decode (decode_layered_offset_min_sum_traced,
src/qc_ldpc.rs:853-874; on_iteration at :1208).

The callback must see, once per completed pass and before that pass's
syndrome check, exactly the posterior the reference passes: bit-exact with
the oracle's trace, and on the reference's Wi-Fi convergence fixture
(bench/results/ldpc_convergence.json) the per-iteration hard decisions of
every recorded frame.  A traced decode runs on the plain kernel for every
code: the Wi-Fi fixture, BG1 Z=384 and a raw-edge code.
"""
import json
import os

import numpy as np
import pytest

from tests.helpers import say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _traced(dec, llr, iters):
    n = dec.variable_node_count()
    x = np.array(llr, np.float32, copy=True)
    er = np.zeros(dec.required_edge_buffer(), np.float32)
    sc = np.zeros(dec.required_layer_buffer(), np.float32)
    hard = np.zeros(n, np.uint8)
    seen = []
    used = dec.decode_layered_offset_min_sum_traced(x, er, sc, hard, iters, lambda i, p: seen.append((i, p.copy())))
    return used, x, er, hard, seen


def _check_vs_oracle(oracle, code, dec, beta, llr, iters):
    used, post, er, hard, seen = _traced(dec, llr, iters)
    it, rp, rer, rh, rtr = oracle.decode_f32(code, beta, llr, iters, trace=True)
    assert used == it and len(seen) == it
    assert [i for i, _ in seen] == list(range(1, it + 1))
    for k, (_, p) in enumerate(seen):
        assert np.array_equal(p.view(np.uint32), rtr[k].view(np.uint32)), f"pass {k + 1} posterior differs"
    assert np.array_equal(post.view(np.uint32), rp.view(np.uint32))
    assert np.array_equal(er.view(np.uint32), rer.view(np.uint32))
    assert np.array_equal(hard, rh)
    return seen


def test_traced_wifi_convergence_fixture(oracle):
    from syndrome_amd import QcLdpcDecoder
    say("[traced] Wi-Fi Z=27 R=1/2 fixture, 14 frames")
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ldpc_convergence.json")))
    code = oracle.code_wifi(27, 1, 2)
    cw = np.array(d["transmitted_codeword"], np.uint8)
    llr = oracle.awgn(d["channel"]["eb_n0_db"], 0.5, d["channel"]["seed"], cw)
    dec = QcLdpcDecoder.wifi(27, 1, 2, d["code"]["offset_beta"])
    seen = _check_vs_oracle(oracle, code, dec, d["code"]["offset_beta"], llr, d["max_iterations"])
    assert len(seen) == d["iterations_used"]
    for i, p in seen:
        f = d["frames"][i]
        assert f["iteration"] == i
        hb = (p < 0).astype(np.uint8)
        assert np.array_equal(hb, np.array(f["hard_bits"], np.uint8)), f"frame {i}"
        assert int((hb != cw).sum()) == f["bit_errors"]


def test_traced_bg1_z384(oracle):
    from syndrome_amd import QcLdpcDecoder
    say("[traced] BG1 Z=384, 1.0 dB")
    code = oracle.code_3gpp(0, 384)
    cw = oracle.encode(code, oracle.xorshift_bits(77, code.k))
    llr = oracle.awgn(1.0, code.k / code.n, 5, cw)
    dec = QcLdpcDecoder.with_lifting_size(0, 384, 0.5)
    seen = _check_vs_oracle(oracle, code, dec, 0.5, llr, 12)
    assert len(seen) >= 2


def test_traced_raw_code_plain_kernel(oracle):
    """A raw-edge copy of the Wi-Fi Z=54 R=2/3 matrix."""
    from syndrome_amd import QcLdpcDecoder
    say("[traced] raw-edge code (plain kernel)")
    w = oracle.code_wifi(54, 2, 3)
    rows, cols, shifts = [], [], []
    for r in range(w.m_b):
        for e in range(w.layer_off[r], w.layer_off[r + 1]):
            rows.append(r)
            cols.append(w.cols[e])
            shifts.append(w.shifts[e])
    dec = QcLdpcDecoder.from_raw_edges(list(zip(rows, cols, shifts)), w.m_b, w.n_b, w.z, 0.5)
    code = oracle.code_raw(rows, cols, shifts, w.m_b, w.n_b, w.z)
    cw = oracle.encode(code, oracle.xorshift_bits(3, code.k))
    llr = oracle.awgn(1.5, code.k / code.n, 9, cw)
    _check_vs_oracle(oracle, code, dec, 0.5, llr, 10)


def test_traced_validation():
    from syndrome_amd import QcLdpcDecoder, SyndromeError, _lib as L
    import ctypes as C
    dec = QcLdpcDecoder.with_lifting_size(1, 16, 0.5)
    n = dec.variable_node_count()
    x = np.zeros(n, np.float32)
    er = np.zeros(dec.required_edge_buffer(), np.float32)
    sc = np.zeros(dec.required_layer_buffer(), np.float32)
    hard = np.zeros(n, np.uint8)
    tr = np.zeros(3 * n - 1, np.float32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = L.lib().syndrome_ldpc_decode_traced(dec.handle, p(x), n, p(er), er.size, p(sc), sc.size, p(hard), n, 3,
                                             p(tr), tr.size, None)
    assert rc == L.BUFFER_TOO_SMALL
    # zero iterations: no pass, no callback, hard decisions of the input
    calls = []
    x[:] = -1.0
    assert dec.decode_layered_offset_min_sum_traced(x, er, sc, hard, 0, lambda i, q: calls.append(i)) == 0
    assert calls == [] and hard.all()
    with pytest.raises(SyndromeError):
        dec.decode_layered_offset_min_sum_traced(x[:-1], er, sc, hard, 2, lambda i, q: None)
