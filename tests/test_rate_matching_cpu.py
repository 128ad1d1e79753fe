"""Oracle pins for rate matching and de-matching (rate_matching.rs:72-715,
harq.rs:142-158): the reference's own unit-test facts plus closed forms of
TS 38.212 5.4.2 bit selection and interleaving, so that a wrong k0, a
transposed interleaver, a filler that is not skipped or a repetition in the
wrong place fails here."""
import numpy as np
import pytest

K0 = {0: (0, 17, 33, 56), 1: (0, 13, 25, 43)}  # rv_k0 numerators (rate_matching.rs:72-94)


def _selection(bg, z, rv, nf, e):
    """Closed form of the bit-selection walk: the circular-buffer positions
    of selections 0..e-1 (no guard), skipping the filler range."""
    ncb = (66 if bg == 0 else 50) * z
    k = (22 if bg == 0 else 10) * z
    kp = max(k - nf, 0)
    fs, fe = max(kp - 2 * z, 0), max(k - 2 * z, 0)
    walk = [(K0[bg][rv] * z + j) % ncb for j in range(3 * ncb)]
    return [p for p in walk if not (fs <= p < fe)][:e], ncb


@pytest.mark.parametrize("bg,z,rv,qm,nf,e", [(0, 384, 1, 1, 0, 500), (0, 4, 0, 1, 0, 64), (0, 8, 1, 2, 6, 120),
                                             (1, 6, 2, 4, 4, 96), (1, 16, 3, 6, 20, 180), (1, 56, 3, 1, 136, 700)])
def test_rate_match_is_selection_then_interleave(oracle, bg, z, rv, qm, nf, e):
    rng = np.random.default_rng(bg * 1000 + z)
    n = (68 if bg == 0 else 52) * z
    k = (22 if bg == 0 else 10) * z
    cw = rng.integers(0, 2, n).astype(np.uint8)
    cw[k - nf:k] = 0
    sel, _ = _selection(bg, z, rv, nf, e)
    picked = cw[2 * z + np.array(sel)]           # selection reads codeword bit 2Z + position
    rows = e // qm
    expect = picked.reshape(rows, qm).T.reshape(-1)  # written row by row, read column by column
    assert np.array_equal(oracle.rate_match(cw, e, rv, qm, bg, z, nf), expect)


def test_k0_known_values(oracle):
    # rv_k0(BG1, 1, 384) = 17*384 (rate_matching.rs:733-737): the first selected bit
    z = 384
    cw = np.zeros(68 * z, np.uint8)
    cw[2 * z + 17 * z] = 1
    e = oracle.rate_match(cw, 8, 1, 1, 0, z, 0)
    assert e[0] == 1 and not e[1:].any()
    cw[:] = 0
    cw[2 * z] = 1  # rv 0 starts at 0 for both graphs
    assert oracle.rate_match(cw, 4, 0, 1, 0, z, 0)[0] == 1


@pytest.mark.parametrize("cached", [False, True])
def test_dematch_inverts_selection_without_repetition(oracle, cached):
    bg, z, rv, qm, nf = 1, 16, 3, 6, 20
    ncb = 50 * z
    e_len = 6 * 100
    sel, _ = _selection(bg, z, rv, nf, e_len)
    vals = np.arange(1, e_len + 1, dtype=np.float32)  # exact integers
    deint = np.empty(e_len, np.float32)
    rows = e_len // qm
    interleaved = vals.reshape(rows, qm).T.reshape(-1)  # what the channel delivers for selection order vals
    cb = oracle.rate_dematch(interleaved, rv, qm, bg, z, nf, cached=cached)
    expect = np.zeros(ncb, np.float32)
    expect[np.array(sel)] = vals
    assert np.array_equal(cb, expect)
    del deint


@pytest.mark.parametrize("cached", [False, True])
def test_repetition_accumulates(oracle, cached):
    # m < E <= 2m: positions of rank r < E - m get two LLRs, the rest one
    bg, z, nf = 0, 4, 10
    ncb = 66 * z
    sel, _ = _selection(bg, z, 0, nf, 2 * ncb)
    m = len(set(sel))
    e_len = m + 37
    vals = np.arange(1, e_len + 1, dtype=np.float32)
    cb = oracle.rate_dematch(vals, 0, 1, bg, z, nf, cached=cached)
    expect = np.zeros(ncb, np.float32)
    for i, p in enumerate(sel[:e_len]):
        expect[p] += vals[i]
    assert np.array_equal(cb, expect)


def test_cached_equals_uncached_on_the_reference_configs(oracle):
    # cached_rate_dematch_matches_uncached_random (rate_matching.rs): HARQ
    # sequences of transmissions accumulate identically in both forms
    rng = np.random.default_rng(0xDECA11ED)
    for bg, z, qm, nf, e, rvs in ((0, 4, 1, 0, 64, (0, 0, 1)), (1, 10, 4, 8, 160, (0, 2, 3))):
        a = b = None
        for rv in rvs:
            llr = (rng.integers(0, 2000, e) / 100.0 - 10.0).astype(np.float32)
            a = oracle.rate_dematch(llr, rv, qm, bg, z, nf, cb=a, cached=True)
            b = oracle.rate_dematch(llr, rv, qm, bg, z, nf, cb=b, cached=False)
        assert np.array_equal(a, b)


def test_cached_guard_sends_the_excess_to_position_zero(oracle):
    # the selection table has 2m+1 real entries (the walk's 2*Ncb guard);
    # later slots keep their initial 0, so their LLRs all land on position 0
    bg, z, nf = 1, 2, 0
    ncb = 50 * z
    sel, _ = _selection(bg, z, 0, nf, 3 * ncb)
    m = ncb
    e_len = 2 * m + 1 + 9
    vals = np.ones(e_len, np.float32)
    cb = oracle.rate_dematch(vals, 0, 1, bg, z, nf, cached=True)
    expect = np.zeros(ncb, np.float32)
    for i in range(2 * m + 1):
        expect[sel[i]] += 1.0
    expect[0] += 9.0
    assert np.array_equal(cb, expect)
    # the uncached function stops at the guard instead
    cb2 = oracle.rate_dematch(vals, 0, 1, bg, z, nf, cached=False)
    expect[0] -= 9.0
    assert np.array_equal(cb2, expect)


def test_cached_adds_in_received_order(oracle):
    # with a non-zero accumulator and two repetitions in different
    # interleaver columns, the order of the two adds is visible in f32:
    # the cached form adds in received (interleaved) order
    bg, z, qm, nf = 1, 2, 2, 3
    sel, _ = _selection(bg, z, 0, nf, 400)
    m = len(set(sel))
    assert m % qm == 1  # repetitions of one position fall in different columns
    e_len = (m + 5) // qm * qm + qm
    rows = e_len // qm
    rng = np.random.default_rng(1)
    llr = rng.normal(0, 1e4, e_len).astype(np.float32)
    acc0 = rng.normal(0, 1.0, 50 * z).astype(np.float32)
    got = oracle.rate_dematch(llr, 0, qm, bg, z, nf, cb=acc0, cached=True)
    expect = acc0.copy()
    for q in range(e_len):
        ks = (q % rows) * qm + q // rows
        expect[sel[ks]] = np.float32(expect[sel[ks]] + llr[q])
    assert np.array_equal(got.view(np.uint32), expect.view(np.uint32))
    # ... and selection order would give a different result here
    korder = acc0.copy()
    for ks in range(e_len):
        q = (ks % qm) * rows + ks // qm
        korder[sel[ks]] = np.float32(korder[sel[ks]] + llr[q])
    assert not np.array_equal(korder.view(np.uint32), expect.view(np.uint32))
