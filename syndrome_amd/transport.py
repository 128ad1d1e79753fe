"""Transport-block segmentation and CRCs on the host, for building serving
traffic (the device checks the CRCs; ldpc_crc.h).

* compute_segmentation (src/segmentation.rs:265-331, TS 38.212 5.2.2 /
  7.2.2): base-graph selection by TB size and code rate, the number of code
  blocks C, K' per block (with a 24-bit CB CRC when C > 1), the smallest
  lifting size with k_b * Z >= K', and the filler count K - K'.
* crc24 / attach_crc24 (src/crc.rs:258-330): MSB-first, initial 0, no
  reflection, no final XOR; a byte-at-a-time table over packed bits.
* segment (src/segmentation.rs:361-410): the TB with its CRC-24A cut into
  C blocks of K' - 24 payload bits (K' when C == 1), zero-padded, each
  followed by its CRC-24B when C > 1.
"""
from dataclasses import dataclass

import numpy as np

# TS 38.212 Table 5.3.2-1, ascending (segmentation.rs:75-79)
VALID_Z = (2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40, 44, 48,
           52, 56, 60, 64, 72, 80, 88, 96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240, 256, 288, 320,
           352, 384)
MAX_TB_SIZE_BITS = 1_277_992  # segmentation.rs:90


@dataclass
class Segmentation:
    bg: int          # 0 = BG1, 1 = BG2
    z: int
    c: int           # code blocks
    b: int           # TBS + 24 (TB CRC)
    k_prime: int     # bits per code block incl. CB CRC
    k: int           # k_b * Z (encoder systematic length)
    n: int           # n_b * Z circular-buffer length (n_b = 66 / 50)
    n_filler: int
    has_cb_crc: bool


def select_base_graph(a, r):
    """segmentation.rs:125-131; the rate is compared in f32 as there."""
    r = np.float32(r)
    return 1 if (a <= 292 or (a <= 3824 and r <= np.float32(0.67)) or r <= np.float32(0.25)) else 0


def _k_b(bg, b):
    if bg == 0:
        return 22
    if b > 640:
        return 10
    if b > 560:
        return 9
    if b > 192:
        return 8
    return 6


def compute_segmentation(a, target_rate):
    if a <= 0 or a > MAX_TB_SIZE_BITS:
        raise ValueError("transport block size out of range")
    bg = select_base_graph(a, target_rate)
    k_cb = 8448 if bg == 0 else 3840
    n_b = 66 if bg == 0 else 50
    b = a + 24
    if b <= k_cb:
        c, l = 1, 0
    else:
        c = -(-b // (k_cb - 24))
        l = 24
    b_prime = b + c * l
    k_prime = -(-b_prime // c)
    kb = _k_b(bg, b)
    z = next((zz for zz in VALID_Z if kb * zz >= k_prime), None)
    if z is None:
        raise ValueError("K' too large")
    k = (22 if bg == 0 else 10) * z
    return Segmentation(bg, z, c, b, k_prime, k, n_b * z, k - k_prime, l > 0)


CRC24A = 0x864CFB
CRC24B = 0x800063


def _crc_table(poly):
    t = np.zeros(256, np.uint32)
    for i in range(256):
        reg = i << 16
        for _ in range(8):
            reg = ((reg << 1) ^ poly) if reg & 0x800000 else (reg << 1)
            reg &= 0xFFFFFF
        t[i] = reg
    return t


_TABLES = {}


def crc24(bits, poly=CRC24A):
    """Remainder of the bit sequence (one bit per element, MSB first) times
    x^24 modulo g."""
    b = np.asarray(bits, np.uint8) & 1
    table = _TABLES.setdefault(poly, _crc_table(poly))
    full = b.size // 8 * 8
    reg = 0
    for byte in np.packbits(b[:full]).tolist():
        reg = ((reg << 8) & 0xFFFFFF) ^ int(table[((reg >> 16) ^ byte) & 0xFF])
    for bit in b[full:].tolist():
        fb = ((reg >> 23) ^ bit) & 1
        reg = (reg << 1) & 0xFFFFFF
        if fb:
            reg ^= poly
    return reg


def attach_crc24(bits, poly=CRC24A):
    r = crc24(bits, poly)
    return np.concatenate([np.asarray(bits, np.uint8), ((r >> np.arange(23, -1, -1)) & 1).astype(np.uint8)])


def segment(tb_with_crc, seg):
    """[C, K'] blocks of a transport block with its CRC-24A attached."""
    t = np.asarray(tb_with_crc, np.uint8)
    c, kp = seg.c, seg.k_prime
    p = kp - 24 if c > 1 else kp
    out = np.zeros((c, kp), np.uint8)
    padded = np.zeros(c * p, np.uint8)
    padded[:min(t.size, seg.b)] = t[:seg.b]
    out[:, :p] = padded.reshape(c, p)
    if c > 1:
        for i in range(c):
            r = crc24(out[i, :p], CRC24B)
            out[i, p:] = (r >> np.arange(23, -1, -1)) & 1
    return out
