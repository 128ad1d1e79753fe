"""Input generation shared by the tests (seeded, reproducible)."""
import numpy as np


def say(*a):
    print(*a, flush=True)


def codewords(orc, code, count, seed):
    """`count` codewords of `code` from xorshift middle-bit info
    (tests/ldpc_int8_kernel_equivalence.rs:56-68 style) encoded by the oracle."""
    out = np.zeros((count, code.n), np.uint8)
    for i in range(count):
        out[i] = orc.encode(code, orc.xorshift_bits((seed + 7919 * i) ^ 0xA5A5A5A5, code.k))
    return out


def channel(orc, code, cws, ebno_list, seed):
    """AWGN LLRs for each codeword (channel_sim.rs semantics), Eb/N0 cycling over ebno_list."""
    rate = code.k / code.n
    llr = np.zeros(cws.shape, np.float32)
    for i in range(cws.shape[0]):
        llr[i] = orc.awgn(ebno_list[i % len(ebno_list)], rate, seed + i, cws[i])
    return llr


def f32_bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)
