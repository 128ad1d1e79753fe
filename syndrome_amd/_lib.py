"""ctypes loader for ``libsyndrome_amd.so``.

The library is built in-tree by ``syndrome_amd.build``.  Nothing here falls
back to a CPU implementation: if the library is missing the import of
:func:`lib` raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsyndrome_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "syndrome_ldpc.h")

# status codes (capi/src/lib.rs:74-94 plus device codes)
OK = 0
INVALID_PARAM = -1
CRC_MISMATCH = -2
DECODER_NOT_CONVERGED = -3
BUFFER_TOO_SMALL = -4
NULL_POINTER = -5
PANIC = -6
DEVICE_ERROR = -7
BUSY = -8

STATUS_NAMES = {0: "Ok", -1: "InvalidParam", -2: "CrcMismatch", -3: "DecoderNotConverged",
                -4: "BufferTooSmall", -5: "NullPointer", -6: "Panic", -7: "DeviceError", -8: "Busy"}

# batch flags
INIT_5G = 1
WRITE_POST = 2
WRITE_HARD = 4
WRITE_EDGE = 8
QUANTIZE = 16
PLAIN = 256

P = C.c_void_p
SZ = C.c_size_t
I32 = C.c_int32

# name -> (restype, argtypes); every symbol declared in include/syndrome_ldpc.h
SIGNATURES = {
    "syndrome_ldpc_decoder_create": (I32, [I32, SZ, C.c_float, C.POINTER(P)]),
    "syndrome_ldpc_decoder_destroy": (None, [P]),
    "syndrome_ldpc_decoder_variable_node_count": (SZ, [P]),
    "syndrome_ldpc_decoder_info_bit_count": (SZ, [P]),
    "syndrome_ldpc_decoder_required_edge_buffer": (SZ, [P]),
    "syndrome_ldpc_decoder_required_layer_buffer": (SZ, [P]),
    "syndrome_ldpc_decode_5g": (I32, [P, P, SZ, SZ, P, SZ, P, SZ, P, SZ, SZ, C.POINTER(SZ)]),
    "syndrome_ldpc_decode_5g_static": (I32, [P, P, SZ, SZ, P, SZ, P, SZ, P, SZ, SZ, C.POINTER(SZ), P]),
    "syndrome_ldpc_encoder_create": (I32, [I32, SZ, C.POINTER(P)]),
    "syndrome_ldpc_encoder_destroy": (None, [P]),
    "syndrome_ldpc_encoder_info_bit_count": (SZ, [P]),
    "syndrome_ldpc_encoder_codeword_bit_count": (SZ, [P]),
    "syndrome_ldpc_encode_5g": (I32, [P, P, SZ, SZ, P, SZ]),
    "syndrome_ldpc_decoder_create_raw": (I32, [P, P, P, SZ, SZ, SZ, SZ, C.c_float, C.POINTER(P)]),
    "syndrome_ldpc_decoder_create_wifi": (I32, [SZ, SZ, SZ, C.c_float, C.POINTER(P)]),
    "syndrome_ldpc_decoder_check_node_count": (SZ, [P]),
    "syndrome_ldpc_decoder_offset_beta": (C.c_float, [P]),
    "syndrome_ldpc_decode": (I32, [P, P, SZ, P, SZ, P, SZ, P, SZ, SZ, C.POINTER(SZ)]),
    "syndrome_ldpc_decode_traced": (I32, [P, P, SZ, P, SZ, P, SZ, P, SZ, SZ, P, SZ, C.POINTER(SZ)]),
    "syndrome_ldpc_decode_i8": (I32, [P, P, SZ, P, SZ, P, SZ, P, SZ, SZ, C.c_float, C.c_int16, C.POINTER(SZ)]),
    "syndrome_ldpc_decode_5g_i8": (I32, [P, P, SZ, SZ, P, SZ, P, SZ, P, SZ, SZ, C.c_float, C.c_int16,
                                         C.POINTER(SZ)]),
    "syndrome_ldpc_decode_batch_f32": (I32, [P, P, P, P, P, P, SZ, SZ, SZ, C.c_uint32, P]),
    "syndrome_ldpc_decode_batch_i8": (I32, [P, P, P, P, P, P, SZ, SZ, SZ, C.c_float, C.c_int16, C.c_uint32, P]),
    "syndrome_ldpc_launch_shape": (I32, [P, I32, I32, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]),
    "syndrome_ldpc_quantize_i16_device": (I32, [P, P, SZ, C.c_float, P]),
    "syndrome_ldpc_quantize_i8_device": (I32, [P, P, SZ, C.c_float, P]),
    "syndrome_ldpc_dequantize_i8_device": (I32, [P, P, SZ, C.c_float, P]),
    "syndrome_awgn_transmit": (I32, [C.c_float, C.c_float, C.c_uint64, P, SZ, P]),
    "syndrome_ldpc_encode_batch": (I32, [P, P, P, SZ, I32]),
    "syndrome_ldpc_encode_batch_device": (I32, [P, P, P, SZ, P]),
    "syndrome_ldpc_plan_create": (I32, [P, P, P, P, SZ, C.c_float, I32, I32, C.POINTER(P)]),
    "syndrome_ldpc_plan_destroy": (None, [P]),
    "syndrome_ldpc_plan_run": (I32, [P, P, P, P, P, SZ, C.c_float, C.c_int16, C.c_uint32, P]),
    "syndrome_ldpc_plan_shape": (I32, [P, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]),
    "syndrome_ldpc_plan_tune": (I32, [P, P, P, P, P, SZ, C.c_float, C.c_int16, C.c_uint32, I32, C.POINTER(I32)]),
    "syndrome_ldpc_plan_set_dematch": (I32, [P, P, P, P, P]),
    "syndrome_ldpc_plan_dematch": (I32, [P, P, P, P]),
    "syndrome_ldpc_plan_harq_combine": (I32, [P, P, P, I32, P]),
    "syndrome_ldpc_plan_set_tbs": (I32, [P, SZ, P, P, P, P]),
    "syndrome_ldpc_plan_check_tbs": (I32, [P, P, P, P, P, P]),
    "syndrome_traffic_blocks": (I32, [P, P, P, P, P, P, C.c_uint64, SZ, P, P, P, P, I32]),
    "syndrome_traffic_blocks_info": (I32, [P, P, P, P, P, P, C.c_uint64, SZ, P, P, P, P, I32]),
    "syndrome_server_create": (I32, [C.c_float, SZ, SZ, I32, I32, C.POINTER(P)]),
    "syndrome_server_destroy": (None, [P]),
    "syndrome_server_acquire": (I32, [P, I32, SZ, SZ, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32)]),
    "syndrome_server_submit": (I32, [P, C.c_uint32, C.c_uint64]),
    "syndrome_server_try_recv": (I32, [P, P]),
    "syndrome_server_release": (I32, [P, C.c_uint32]),
    "syndrome_server_acquire_many": (I32, [P, P, P, P, SZ, P, P, C.POINTER(SZ)]),
    "syndrome_server_submit_many": (I32, [P, P, P, P, SZ]),
    "syndrome_server_try_recv_many": (I32, [P, P, SZ, C.POINTER(SZ)]),
    "syndrome_server_release_many": (I32, [P, P, SZ]),
    "syndrome_server_submit_tb": (I32, [P, P, P, SZ, SZ, P, C.c_uint64]),
    "syndrome_server_try_recv_tb": (I32, [P, P]),
    "syndrome_server_release_tb": (I32, [P, C.c_uint32]),
    "syndrome_server_stats": (I32, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(I32)]),
    "syndrome_server_submit_from": (I32, [P, C.c_uint32, P, C.c_uint64]),
    "syndrome_host_alloc": (P, [SZ]),
    "syndrome_host_free": (None, [P]),
    "syndrome_server_loadgen": (I32, [P, P, P, P, P, P, P, P, P, SZ, C.c_double, C.c_double, C.c_uint64, P, SZ,
                                      C.POINTER(SZ), P, SZ, C.POINTER(SZ), C.POINTER(C.c_double),
                                      C.POINTER(C.c_int64)]),
    "syndrome_server_loadgen_tb": (I32, [P, P, P, P, P, P, P, P, P, P, P, SZ, C.c_double, C.c_double, C.c_uint64,
                                         P, P, SZ, P, P, P, P, P, P]),
    "syndrome_server_tb_block_crc": (I32, [P, C.c_uint32, P, SZ, C.POINTER(SZ)]),
}


class Served(C.Structure):
    _fields_ = [("slot", C.c_uint32), ("tag", C.c_uint64), ("iterations", C.c_int32),
                ("hard", C.POINTER(C.c_uint8)), ("n", C.c_size_t), ("submit_ns", C.c_double),
                ("done_ns", C.c_double)]


class ServedTb(C.Structure):
    _fields_ = [("tb", C.c_uint32), ("tag", C.c_uint64), ("crc_ok", C.c_int32), ("max_iterations", C.c_int32),
                ("tb_size", C.c_size_t), ("num_cb", C.c_size_t), ("submit_ns", C.c_double),
                ("done_ns", C.c_double)]

_lib = None


class SyndromeError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)} ({status})")


def check(status, what="syndrome call"):
    if status != OK:
        raise SyndromeError(status, what)
    return status


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m syndrome_amd.build` "
                              "(the decoder has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Every function name declared in include/syndrome_ldpc.h."""
    import re
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(syndrome_\w+)\s*\(", text)))
