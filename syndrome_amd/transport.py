"""Transport-block segmentation for building serving traffic.

compute_segmentation (src/segmentation.rs:265-331, TS 38.212 5.2.2 / 7.2.2):
base-graph selection by TB size and code rate, the number of code blocks C,
K' per block (with a 24-bit CB CRC when C > 1), the smallest lifting size
with k_b * Z >= K', and the filler count K - K'.
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
