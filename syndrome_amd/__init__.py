"""MI355X-native layered offset min-sum QC-LDPC decoder.

The decode runs in hand-written HIP kernels for gfx950 (``csrc/``), behind
a C ABI (``include/syndrome_ldpc.h``) that is a drop-in for the
reference's ``syndrome_ldpc_*`` symbols.  ``QcLdpcDecoder`` mirrors the
reference's Rust decoder API over that ABI.
"""
from ._lib import SyndromeError, lib  # noqa: F401
from .ldpc import (APP_CLAMP_I8, APP_CLAMP_WIDE, BG1, BG2, DEFAULT_SCALE, MSG_MAX, DecodePlan, LdpcWorkspace,  # noqa: F401,E501
                   HostPool, QcLdpcDecoder, QcLdpcEncoder, QuantParams, Server, awgn_transmit,
                   quantize_llr_i16_device, serve_loadgen, serve_loadgen_tb, traffic_blocks,
                   traffic_blocks_info)
