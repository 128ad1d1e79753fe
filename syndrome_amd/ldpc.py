"""Python mirror of the reference's decoder API over the HIP C ABI.

``QcLdpcDecoder`` follows ``syndrome::qc_ldpc::QcLdpcDecoder``
(src/qc_ldpc.rs:481-1963): the same constructors, size queries, decode
entry points, argument meaning and error behaviour (errors surface as
:class:`SyndromeError` carrying the reference's status code).  Host arrays
are numpy; the decode runs on the GPU.  ``decode_batch_*`` take torch
tensors that already live in HBM.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import SyndromeError, check, lib

BG1, BG2 = 0, 1
DEFAULT_SCALE = 8.0          # quantize.rs:182
MSG_MAX = 127                # quantize.rs:190
APP_CLAMP_I8 = 127           # quantize.rs:194
APP_CLAMP_WIDE = 32767       # quantize.rs:205


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _need(t, name, what, dtypes, numel):
    """A device buffer handed to a kernel: a contiguous torch tensor of one
    of `dtypes` with at least `numel` elements (residency is checked after
    every buffer's shape, by _resident).  Anything else raises ValueError
    here, before a launch could read or write past it."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what}: {name} must be a torch tensor")
    if t.dtype not in dtypes:
        names = " or ".join(str(d).replace("torch.", "") for d in dtypes)
        raise ValueError(f"{what}: {name} must be {names}, got {str(t.dtype).replace('torch.', '')}")
    if not t.is_contiguous():
        raise ValueError(f"{what}: {name} must be contiguous")
    if t.numel() < numel:
        raise ValueError(f"{what}: {name} holds {t.numel()} elements, {numel} needed")


def _resident(what, **tensors):
    for name, t in tensors.items():
        if t is not None and not t.is_cuda:
            raise ValueError(f"{what}: {name} must be in device memory (a CUDA tensor)")


def check_batch_tensors(what, n, ez, int8, flags, d_in, d_post, d_hard, d_iters, d_edge):
    """Validate the tensors of a device batch decode of d_iters.numel()
    codewords of N = n variables and E*Z = ez messages (f32 path, or the
    int8 path with i16 / f32-QUANTIZE input).  Returns the batch size."""
    import torch
    _need(d_iters, "d_iters", what, (torch.int32,), 0)
    batch = d_iters.numel()
    if int8:
        in_types = (torch.float32,) if flags & _lib.QUANTIZE else (torch.int16,)
        post_t, edge_t = (torch.int16,), (torch.int8,)
    else:
        if flags & _lib.QUANTIZE:
            raise ValueError(f"{what}: QUANTIZE is an int8 flag")
        in_types, post_t, edge_t = (torch.float32,), (torch.float32,), (torch.float32,)
    _need(d_in, "d_in", what, in_types, batch * n)
    if flags & _lib.WRITE_POST:
        _need(d_post, "d_post", what, post_t, batch * n)
    if flags & _lib.WRITE_HARD:
        _need(d_hard, "d_hard", what, (torch.uint8,), batch * n)
    if d_edge is not None or flags & (_lib.WRITE_EDGE | _lib.PLAIN):
        _need(d_edge, "d_edge", what, edge_t, batch * ez)
    _resident(what, d_in=d_in, d_post=d_post, d_hard=d_hard, d_iters=d_iters, d_edge=d_edge)
    return batch


def check_plan_tensors(what, arena, count, int8, flags, d_in, d_post, d_hard, d_iters):
    """Validate a plan run's arenas: `arena` = max over blocks of offset + N_i
    elements; iterations for `count` blocks."""
    import torch
    if int8:
        in_types = (torch.float32,) if flags & _lib.QUANTIZE else (torch.int16,)
        post_t = (torch.int16,)
    else:
        if flags & _lib.QUANTIZE:
            raise ValueError(f"{what}: QUANTIZE is an int8 flag")
        in_types, post_t = (torch.float32,), (torch.float32,)
    _need(d_in, "d_in", what, in_types, arena)
    if flags & _lib.WRITE_POST:
        _need(d_post, "d_post", what, post_t, arena)
    if flags & _lib.WRITE_HARD:
        _need(d_hard, "d_hard", what, (torch.uint8,), arena)
    _need(d_iters, "d_iters", what, (torch.int32,), count)
    _resident(what, d_in=d_in, d_post=d_post, d_hard=d_hard, d_iters=d_iters)


class QuantParams:
    """QuantParams (src/quantize.rs:217-305)."""

    def __init__(self, scale=DEFAULT_SCALE, app_clamp=APP_CLAMP_WIDE):
        self.scale = float(scale)
        self.app_clamp = int(app_clamp)


class QcLdpcDecoder:
    """GPU layered offset min-sum decoder behind the reference's API."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    # -- constructors (qc_ldpc.rs:506-565, wifi_ldpc_tables.rs:521-529) --
    @classmethod
    def with_lifting_size(cls, bg, z, offset_beta):
        out = C.c_void_p()
        check(lib().syndrome_ldpc_decoder_create(bg, z, offset_beta, C.byref(out)), "decoder_create")
        return cls(out.value)

    @classmethod
    def new(cls, bg, offset_beta):
        return cls.with_lifting_size(bg, 384 if bg == BG1 else 128, offset_beta)

    @classmethod
    def from_raw_edges(cls, entries, num_row_blocks, num_col_blocks, z, offset_beta):
        e = np.asarray(entries, dtype=np.int64).reshape(-1, 3)
        if (e < 0).any() or (e[:, :2] > 255).any() or (e[:, 2] > 32767).any():
            raise SyndromeError(_lib.INVALID_PARAM, "from_raw_edges")
        rows = np.ascontiguousarray(e[:, 0], np.uint8)
        cols = np.ascontiguousarray(e[:, 1], np.uint8)
        sh = np.ascontiguousarray(e[:, 2], np.int16)
        out = C.c_void_p()
        check(lib().syndrome_ldpc_decoder_create_raw(_ptr(rows), _ptr(cols), _ptr(sh), len(e), num_row_blocks,
                                                     num_col_blocks, z, offset_beta, C.byref(out)),
              "decoder_create_raw")
        return cls(out.value)

    @classmethod
    def wifi(cls, z, rate_num, rate_den, offset_beta):
        out = C.c_void_p()
        check(lib().syndrome_ldpc_decoder_create_wifi(z, rate_num, rate_den, offset_beta, C.byref(out)),
              "decoder_create_wifi")
        return cls(out.value)

    def close(self):
        if self._h and self._h.value:
            lib().syndrome_ldpc_decoder_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- size queries (qc_ldpc.rs:569-588, 736-739) --
    def variable_node_count(self):
        return lib().syndrome_ldpc_decoder_variable_node_count(self._h)

    def check_node_count(self):
        return lib().syndrome_ldpc_decoder_check_node_count(self._h)

    def info_bit_count_5g(self):
        return lib().syndrome_ldpc_decoder_info_bit_count(self._h)

    def required_edge_buffer(self):
        return lib().syndrome_ldpc_decoder_required_edge_buffer(self._h)

    def required_layer_buffer(self):
        return lib().syndrome_ldpc_decoder_required_layer_buffer(self._h)

    @property
    def offset_beta(self):
        return lib().syndrome_ldpc_decoder_offset_beta(self._h)

    # -- host-buffer decodes (in place, like the reference) --
    def decode_5g(self, llr, n_filler, edge_r, layer_scratch, hard_output, iterations):
        """QcLdpcDecoder::decode_5g (qc_ldpc.rs:644-655).  Returns iterations used."""
        used = C.c_size_t(0)
        check(lib().syndrome_ldpc_decode_5g(self._h, _ptr(llr), llr.size, n_filler, _ptr(edge_r), edge_r.size,
                                            _ptr(layer_scratch), layer_scratch.size, _ptr(hard_output),
                                            hard_output.size, iterations, C.byref(used)), "decode_5g")
        return used.value

    def decode_layered_offset_min_sum(self, llr, edge_r, layer_scratch, hard_output, iterations):
        """QcLdpcDecoder::decode_layered_offset_min_sum (qc_ldpc.rs:794-810)."""
        used = C.c_size_t(0)
        check(lib().syndrome_ldpc_decode(self._h, _ptr(llr), llr.size, _ptr(edge_r), edge_r.size,
                                         _ptr(layer_scratch), layer_scratch.size, _ptr(hard_output),
                                         hard_output.size, iterations, C.byref(used)), "decode")
        return used.value

    def decode_layered_offset_min_sum_traced(self, llr, edge_r, layer_scratch, hard_output, iterations,
                                             on_iteration):
        """QcLdpcDecoder::decode_layered_offset_min_sum_traced (qc_ldpc.rs:853-874):
        on_iteration(i, posterior) once per completed pass i = 1, 2, ..., with
        the posterior before that pass's syndrome check.  Returns iterations used."""
        n = self.variable_node_count()
        trace = np.zeros((max(iterations, 1), n), np.float32)
        used = C.c_size_t(0)
        check(lib().syndrome_ldpc_decode_traced(self._h, _ptr(llr), llr.size, _ptr(edge_r), edge_r.size,
                                                _ptr(layer_scratch), layer_scratch.size, _ptr(hard_output),
                                                hard_output.size, iterations, _ptr(trace), trace.size,
                                                C.byref(used)), "decode_traced")
        for i in range(used.value):
            on_iteration(i + 1, trace[i])
        return used.value

    def decode_layered_offset_min_sum_i8(self, app, edge_r, layer_scratch, hard_output, iterations, quant=None):
        """QcLdpcDecoder::decode_layered_offset_min_sum_i8 (qc_ldpc.rs:1345-1363)."""
        q = quant or QuantParams()
        used = C.c_size_t(0)
        check(lib().syndrome_ldpc_decode_i8(self._h, _ptr(app), app.size, _ptr(edge_r), edge_r.size,
                                            _ptr(layer_scratch), layer_scratch.size, _ptr(hard_output),
                                            hard_output.size, iterations, q.scale, q.app_clamp, C.byref(used)),
              "decode_i8")
        return used.value

    def decode_5g_i8(self, app, n_filler, edge_r, layer_scratch, hard_output, iterations, quant=None):
        """QcLdpcDecoder::decode_5g_i8 (qc_ldpc.rs:1486-1523)."""
        q = quant or QuantParams()
        used = C.c_size_t(0)
        check(lib().syndrome_ldpc_decode_5g_i8(self._h, _ptr(app), app.size, n_filler, _ptr(edge_r), edge_r.size,
                                               _ptr(layer_scratch), layer_scratch.size, _ptr(hard_output),
                                               hard_output.size, iterations, q.scale, q.app_clamp, C.byref(used)),
              "decode_5g_i8")
        return used.value

    # -- workspace convenience (qc_ldpc.rs:1817-1963) --
    def workspace(self):
        return LdpcWorkspace(self)

    def decode_with_workspace(self, llr, ws, iterations):
        return self.decode_layered_offset_min_sum(llr, ws.edge_r, ws.layer_scratch, ws.hard, iterations)

    def decode_5g_with_workspace(self, llr, n_filler, ws, iterations):
        return self.decode_5g(llr, n_filler, ws.edge_r, ws.layer_scratch, ws.hard, iterations)

    # -- device batch (torch tensors in HBM) --
    def launch_shape(self, int8=False, plain=False):
        cws, thr, lds = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().syndrome_ldpc_launch_shape(self._h, int(int8), int(plain), C.byref(cws), C.byref(thr),
                                               C.byref(lds)), "launch_shape")
        return cws.value, thr.value, lds.value

    def decode_batch_f32(self, d_in, d_post, d_hard, d_iters, iterations, n_filler=0, flags=None, d_edge=None,
                         stream=None):
        import torch
        if flags is None:
            flags = _lib.WRITE_POST | _lib.WRITE_HARD
        batch = check_batch_tensors("decode_batch_f32", self.variable_node_count(), self.required_edge_buffer(),
                                    False, flags, d_in, d_post, d_hard, d_iters, d_edge)
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().syndrome_ldpc_decode_batch_f32(
            self._h, d_in.data_ptr(), d_post.data_ptr() if d_post is not None else None,
            d_edge.data_ptr() if d_edge is not None else None, d_hard.data_ptr() if d_hard is not None else None,
            d_iters.data_ptr(), batch, iterations, n_filler, flags, s), "decode_batch_f32")

    def decode_batch_i8(self, d_in, d_post, d_hard, d_iters, iterations, n_filler=0, flags=None, d_edge=None,
                        quant=None, stream=None):
        import torch
        q = quant or QuantParams()
        if flags is None:
            flags = _lib.WRITE_POST | _lib.WRITE_HARD
        batch = check_batch_tensors("decode_batch_i8", self.variable_node_count(), self.required_edge_buffer(),
                                    True, flags, d_in, d_post, d_hard, d_iters, d_edge)
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().syndrome_ldpc_decode_batch_i8(
            self._h, d_in.data_ptr(), d_post.data_ptr() if d_post is not None else None,
            d_edge.data_ptr() if d_edge is not None else None, d_hard.data_ptr() if d_hard is not None else None,
            d_iters.data_ptr(), batch, iterations, n_filler, q.scale, q.app_clamp, flags, s), "decode_batch_i8")

    # -- traffic generation --
    def encode_batch(self, info, threads=0):
        """Encode info [count, K] (u8) to codewords [count, N]."""
        info = np.ascontiguousarray(info, np.uint8)
        count = info.shape[0] if info.ndim == 2 else 1
        out = np.zeros((count, self.variable_node_count()), np.uint8)
        check(lib().syndrome_ldpc_encode_batch(self._h, _ptr(info), _ptr(out), count, threads), "encode_batch")
        return out

    def encode_batch_device(self, d_info, d_cw, stream=None):
        """Encode on device: d_info u8 [count, K] -> d_cw u8 [count, N] (torch tensors in HBM)."""
        import torch
        k, n = self.info_bit_count_5g(), self.variable_node_count()
        if d_info.dtype != torch.uint8 or d_cw.dtype != torch.uint8:
            raise ValueError("encode_batch_device: u8 tensors expected")
        count = d_info.numel() // k if k else 0
        if d_info.numel() != count * k or d_cw.numel() != count * n:
            raise ValueError("encode_batch_device: shapes must be [count, K] and [count, N]")
        if not (d_info.is_contiguous() and d_cw.is_contiguous()):
            raise ValueError("encode_batch_device: contiguous tensors expected")
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().syndrome_ldpc_encode_batch_device(self._h, d_info.data_ptr(), d_cw.data_ptr(), count, s),
              "encode_batch_device")


class LdpcWorkspace:
    """LdpcWorkspace (src/qc_ldpc.rs:1919-1963)."""

    def __init__(self, dec):
        n = dec.variable_node_count()
        self.codeword_llr = np.zeros(n, np.float32)
        self.edge_r = np.zeros(dec.required_edge_buffer(), np.float32)
        self.layer_scratch = np.zeros(dec.required_layer_buffer(), np.float32)
        self.hard = np.zeros(n, np.uint8)

    def hard_output(self):
        return self.hard


class QcLdpcEncoder:
    """QcLdpcEncoder (src/qc_ldpc.rs:2408-2614) through the C ABI."""

    def __init__(self, bg, z):
        out = C.c_void_p()
        check(lib().syndrome_ldpc_encoder_create(bg, z, C.byref(out)), "encoder_create")
        self._h = out

    def __del__(self):
        try:
            if self._h and self._h.value:
                lib().syndrome_ldpc_encoder_destroy(self._h)
        except Exception:
            pass

    def info_bit_count(self):
        return lib().syndrome_ldpc_encoder_info_bit_count(self._h)

    def codeword_bit_count(self):
        return lib().syndrome_ldpc_encoder_codeword_bit_count(self._h)

    def encode_5g(self, info_bits, n_filler, codeword):
        check(lib().syndrome_ldpc_encode_5g(self._h, _ptr(info_bits), info_bits.size, n_filler, _ptr(codeword),
                                            codeword.size), "encode_5g")


def quantize_llr_i16_device(d_llr, d_out, scale=DEFAULT_SCALE, stream=None):
    import torch
    _need(d_llr, "d_llr", "quantize_llr_i16_device", (torch.float32,), 0)
    _need(d_out, "d_out", "quantize_llr_i16_device", (torch.int16,), d_llr.numel())
    _resident("quantize_llr_i16_device", d_llr=d_llr, d_out=d_out)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().syndrome_ldpc_quantize_i16_device(d_llr.data_ptr(), d_out.data_ptr(), d_llr.numel(), scale, s),
          "quantize")


def quantize_llr_i8_device(d_llr, d_out, scale=DEFAULT_SCALE, stream=None):
    """quantize_llr (quantize.rs:148-154) on device: f32 tensor -> i8 tensor."""
    import torch
    _need(d_llr, "d_llr", "quantize_llr_i8_device", (torch.float32,), 0)
    _need(d_out, "d_out", "quantize_llr_i8_device", (torch.int8,), d_llr.numel())
    _resident("quantize_llr_i8_device", d_llr=d_llr, d_out=d_out)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().syndrome_ldpc_quantize_i8_device(d_llr.data_ptr(), d_out.data_ptr(), d_llr.numel(), scale, s),
          "quantize_i8")


def dequantize_llr_i8_device(d_q, d_out, scale=DEFAULT_SCALE, stream=None):
    """dequantize_llr (quantize.rs:163-169) on device: i8 tensor -> f32 tensor."""
    import torch
    _need(d_q, "d_q", "dequantize_llr_i8_device", (torch.int8,), 0)
    _need(d_out, "d_out", "dequantize_llr_i8_device", (torch.float32,), d_q.numel())
    _resident("dequantize_llr_i8_device", d_q=d_q, d_out=d_out)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().syndrome_ldpc_dequantize_i8_device(d_q.data_ptr(), d_out.data_ptr(), d_q.numel(), scale, s),
          "dequantize_i8")


def awgn_transmit(ebno_db, code_rate, seed, bits):
    """AwgnChannel::new(ebno_db, code_rate, seed).transmit(bits) (channel_sim.rs:139-215)."""
    bits = np.ascontiguousarray(bits, np.uint8)
    out = np.zeros(bits.size, np.float32)
    check(lib().syndrome_awgn_transmit(ebno_db, code_rate, seed, _ptr(bits), bits.size, _ptr(out)), "awgn")
    return out


class DecodePlan:
    """Work-queue batch of mixed 3GPP code blocks decoded in one persistent
    launch (include/syndrome_ldpc.h Part 4).  Block i: base graph bg[i]
    (0/1), lifting size z[i], n_filler[i]; its N_i inputs start at element
    offsets[i] of the arenas."""

    def __init__(self, bg, z, n_filler, offsets, offset_beta, int8=False, block=0):
        self.bg = np.ascontiguousarray(bg, np.int32)
        self.z = np.ascontiguousarray(z, np.int32)
        self.n_filler = np.ascontiguousarray(n_filler if n_filler is not None else np.zeros(len(self.bg)), np.int32)
        self.offsets = np.ascontiguousarray(offsets, np.int64)
        if not (self.bg.ndim == self.z.ndim == self.n_filler.ndim == self.offsets.ndim == 1 and
                len(self.bg) == len(self.z) == len(self.n_filler) == len(self.offsets)):
            raise ValueError("DecodePlan: bg, z, n_filler and offsets must be 1-D arrays of one length")
        self.int8 = bool(int8)
        # elements every arena (input, posterior, hard decisions) must hold
        self.arena = int((self.offsets + np.where(self.bg == 0, 68, 52).astype(np.int64) * self.z).max()) \
            if len(self.bg) else 0
        out = C.c_void_p()
        check(lib().syndrome_ldpc_plan_create(_ptr(self.bg), _ptr(self.z), _ptr(self.n_filler), _ptr(self.offsets),
                                              len(self.bg), offset_beta, int(self.int8), block, C.byref(out)),
              "plan_create")
        self._h = out

    def __del__(self):
        try:
            if self._h and self._h.value:
                lib().syndrome_ldpc_plan_destroy(self._h)
        except Exception:
            pass

    @property
    def count(self):
        return len(self.bg)

    def n_of(self):
        return np.where(self.bg == 0, 68, 52) * self.z

    def shape(self):
        b, g, it, lds = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().syndrome_ldpc_plan_shape(self._h, C.byref(b), C.byref(g), C.byref(it), C.byref(lds)), "plan_shape")
        return {"block": b.value, "grid": g.value, "items": it.value, "lds": lds.value}

    def run(self, d_in, d_post, d_hard, d_iters, iterations, flags, quant=None, stream=None):
        import torch
        q = quant or QuantParams()
        check_plan_tensors("DecodePlan.run", self.arena, self.count, self.int8, flags, d_in, d_post, d_hard, d_iters)
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().syndrome_ldpc_plan_run(self._h, d_in.data_ptr(), d_post.data_ptr() if d_post is not None else None,
                                           d_hard.data_ptr() if d_hard is not None else None, d_iters.data_ptr(),
                                           iterations, q.scale, q.app_clamp, flags, s), "plan_run")

    def tune(self, d_in, d_post, d_hard, d_iters, iterations, flags, quant=None, reps=2):
        q = quant or QuantParams()
        check_plan_tensors("DecodePlan.tune", self.arena, self.count, self.int8, flags, d_in, d_post, d_hard,
                           d_iters)
        best = C.c_int32()
        check(lib().syndrome_ldpc_plan_tune(self._h, d_in.data_ptr(), d_post.data_ptr() if d_post is not None else None,
                                            d_hard.data_ptr() if d_hard is not None else None, d_iters.data_ptr(),
                                            iterations, q.scale, q.app_clamp, flags, reps, C.byref(best)),
              "plan_tune")
        return best.value

    def set_dematch(self, e_offsets, e_lengths, qm=None, rv=None):
        self._eo = np.ascontiguousarray(e_offsets, np.int64)
        self._el = np.ascontiguousarray(e_lengths, np.int32)
        for name, a in (("e_offsets", self._eo), ("e_lengths", self._el), ("qm", qm), ("rv", rv)):
            if a is not None and np.size(a) != self.count:
                raise ValueError(f"DecodePlan.set_dematch: {name} must have one entry per block")
        self._e_arena = int((self._eo + self._el).max()) if self.count else 0
        self._qm = None if qm is None else np.ascontiguousarray(qm, np.int32)
        self._rv = None if rv is None else np.ascontiguousarray(rv, np.int32)
        check(lib().syndrome_ldpc_plan_set_dematch(self._h, _ptr(self._eo), _ptr(self._el),
                                                   None if self._qm is None else _ptr(self._qm),
                                                   None if self._rv is None else _ptr(self._rv)), "plan_set_dematch")

    def dematch(self, d_e_llr, d_llr_out, stream=None):
        import torch
        if not hasattr(self, "_e_arena"):
            raise ValueError("DecodePlan.dematch: set_dematch first")
        _need(d_e_llr, "d_e_llr", "DecodePlan.dematch", (torch.float32,), self._e_arena)
        _need(d_llr_out, "d_llr_out", "DecodePlan.dematch", (torch.float32,), self.arena)
        _resident("DecodePlan.dematch", d_e_llr=d_e_llr, d_llr_out=d_llr_out)
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().syndrome_ldpc_plan_dematch(self._h, d_e_llr.data_ptr(), d_llr_out.data_ptr(), s), "plan_dematch")

    def set_tbs(self, tb_first_cb, tb_num_cb, tb_size, tb_out_off):
        """Transport blocks over the plan's code blocks (syndrome_ldpc_plan_set_tbs)."""
        self._tf = np.ascontiguousarray(tb_first_cb, np.int32)
        self._tn = np.ascontiguousarray(tb_num_cb, np.int32)
        self._ts = np.ascontiguousarray(tb_size, np.int32)
        self._to = np.ascontiguousarray(tb_out_off, np.int64)
        if not (len(self._tf) == len(self._tn) == len(self._ts) == len(self._to)):
            raise ValueError("DecodePlan.set_tbs: one entry per transport block in every array")
        self._tb_arena = int((self._to + self._ts).max()) if len(self._tf) else 0
        check(lib().syndrome_ldpc_plan_set_tbs(self._h, len(self._tf), _ptr(self._tf), _ptr(self._tn), _ptr(self._ts),
                                               _ptr(self._to)), "plan_set_tbs")

    def check_tbs(self, d_hard, d_tb_out, d_tb_ok, d_cb_ok, stream=None):
        """CRC-24B per block, reassembly and CRC-24A per transport block."""
        import torch
        if not hasattr(self, "_tb_arena"):
            raise ValueError("DecodePlan.check_tbs: set_tbs first")
        w = "DecodePlan.check_tbs"
        _need(d_hard, "d_hard", w, (torch.uint8,), self.arena)
        _need(d_tb_out, "d_tb_out", w, (torch.uint8,), self._tb_arena)
        _need(d_tb_ok, "d_tb_ok", w, (torch.uint8,), len(self._tf))
        _need(d_cb_ok, "d_cb_ok", w, (torch.uint8,), self.count)
        _resident(w, d_hard=d_hard, d_tb_out=d_tb_out, d_tb_ok=d_tb_ok, d_cb_ok=d_cb_ok)
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().syndrome_ldpc_plan_check_tbs(self._h, d_hard.data_ptr(), d_tb_out.data_ptr(), d_tb_ok.data_ptr(),
                                                 d_cb_ok.data_ptr(), s), "plan_check_tbs")

    def harq_combine(self, d_e_llr, d_acc, rv=-1, stream=None):
        """HarqBuffer::combine on device (src/harq.rs:142-158): add this
        transmission's de-matched LLRs to the decoder-input slots of d_acc.
        rv 0..3 overrides every block's redundancy version."""
        import torch
        if not hasattr(self, "_e_arena"):
            raise ValueError("DecodePlan.harq_combine: set_dematch first")
        _need(d_e_llr, "d_e_llr", "DecodePlan.harq_combine", (torch.float32,), self._e_arena)
        _need(d_acc, "d_acc", "DecodePlan.harq_combine", (torch.float32,), self.arena)
        _resident("DecodePlan.harq_combine", d_e_llr=d_e_llr, d_acc=d_acc)
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().syndrome_ldpc_plan_harq_combine(self._h, d_e_llr.data_ptr(), d_acc.data_ptr(), int(rv), s),
              "plan_harq_combine")


def traffic_blocks(bg, z, n_filler, e_len, ebno_db, rate, seed, threads=0, want_info=True):
    """Synthetic code-block traffic (include/syndrome_ldpc.h
    syndrome_traffic_blocks).  Returns (e_llr, e_offsets, info, info_offsets)."""
    bg = np.ascontiguousarray(bg, np.int32)
    z = np.ascontiguousarray(z, np.int32)
    nf = np.ascontiguousarray(n_filler, np.int32)
    el = np.ascontiguousarray(e_len, np.int32)
    eb = np.ascontiguousarray(ebno_db, np.float32)
    rt = np.ascontiguousarray(rate, np.float32)
    e_off = np.concatenate([[0], np.cumsum(el.astype(np.int64))[:-1]]).astype(np.int64)
    e_llr = np.zeros(int(el.astype(np.int64).sum()), np.float32)
    kp = np.where(bg == 0, 22, 10).astype(np.int64) * z - nf
    i_off = np.concatenate([[0], np.cumsum(kp)[:-1]]).astype(np.int64)
    info = np.zeros(int(kp.sum()), np.uint8) if want_info else None
    check(lib().syndrome_traffic_blocks(_ptr(bg), _ptr(z), _ptr(nf), _ptr(el), _ptr(eb), _ptr(rt), seed, len(bg),
                                        _ptr(e_llr), _ptr(e_off), _ptr(info) if want_info else None,
                                        _ptr(i_off) if want_info else None, threads), "traffic_blocks")
    return e_llr, e_off, info, i_off


def traffic_blocks_info(bg, z, n_filler, e_len, ebno_db, rate, info, info_off, seed, threads=0):
    """syndrome_traffic_blocks_info: the given info bits (K'_i at info_off[i])
    encoded, rate matched and sent over AWGN.  Returns (e_llr, e_offsets)."""
    bg = np.ascontiguousarray(bg, np.int32)
    z = np.ascontiguousarray(z, np.int32)
    nf = np.ascontiguousarray(n_filler, np.int32)
    el = np.ascontiguousarray(e_len, np.int32)
    eb = np.ascontiguousarray(ebno_db, np.float32)
    rt = np.ascontiguousarray(rate, np.float32)
    inf = np.ascontiguousarray(info, np.uint8)
    ioff = np.ascontiguousarray(info_off, np.int64)
    e_off = np.concatenate([[0], np.cumsum(el.astype(np.int64))[:-1]]).astype(np.int64)
    e_llr = np.zeros(int(el.astype(np.int64).sum()), np.float32)
    check(lib().syndrome_traffic_blocks_info(_ptr(bg), _ptr(z), _ptr(nf), _ptr(el), _ptr(eb), _ptr(rt), seed, len(bg),
                                             _ptr(inf), _ptr(ioff), _ptr(e_llr), _ptr(e_off), threads),
          "traffic_blocks_info")
    return e_llr, e_off


class Server:
    """Serving loop (include/syndrome_ldpc.h Part 5): acquire / submit /
    try_recv / release, as LdpcPipeline (src/ldpc_pipeline.rs:499-576).
    streams >= 1: batches launched on that many streams; streams == 0: the
    resident decoder (one grid stays on the device and pulls code blocks
    from a queue as they are submitted)."""

    def __init__(self, offset_beta, iterations, hbm_budget_bytes=256 << 20, max_batch=256, streams=3):
        out = C.c_void_p()
        check(lib().syndrome_server_create(offset_beta, iterations, hbm_budget_bytes, max_batch, streams,
                                           C.byref(out)), "server_create")
        self._h = out

    def close(self):
        if self._h and self._h.value:
            lib().syndrome_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def acquire(self, bg, z, n_filler=0):
        """Returns (slot, numpy view of the slot's N decoder-input floats) or None if busy."""
        p = C.POINTER(C.c_float)()
        slot = C.c_uint32()
        rc = lib().syndrome_server_acquire(self._h, bg, z, n_filler, C.byref(p), C.byref(slot))
        if rc == _lib.BUSY:
            return None
        check(rc, "server_acquire")
        n = (68 if bg == 0 else 52) * z
        return slot.value, np.ctypeslib.as_array(p, shape=(n,))

    def submit(self, slot, tag):
        check(lib().syndrome_server_submit(self._h, slot, tag), "server_submit")

    def try_recv(self):
        """Returns (slot, tag, iterations, hard copy, submit_ns, done_ns) or None."""
        r = _lib.Served()
        rc = lib().syndrome_server_try_recv(self._h, C.byref(r))
        if rc == _lib.BUSY:
            return None
        check(rc, "server_try_recv")
        hard = np.ctypeslib.as_array(r.hard, shape=(r.n,)).copy()
        return r.slot, r.tag, r.iterations, hard, r.submit_ns, r.done_ns

    def release(self, slot):
        check(lib().syndrome_server_release(self._h, slot), "server_release")

    # -- batched forms (one server lock per call) --
    def acquire_many(self, bg, z, n_filler=None):
        """Slots for blocks (bg[i], z[i], n_filler[i]) in order until the
        budget runs out: (uint32 slots, numpy views of their pinned input
        staging), possibly shorter than bg (empty when busy)."""
        bg = np.ascontiguousarray(bg, np.int32)
        z = np.ascontiguousarray(z, np.int32)
        nf = None if n_filler is None else np.ascontiguousarray(n_filler, np.int32)
        slots = np.zeros(bg.size, np.uint32)
        ptrs = (C.POINTER(C.c_float) * max(bg.size, 1))()
        k = C.c_size_t(0)
        rc = lib().syndrome_server_acquire_many(self._h, _ptr(bg), _ptr(z), None if nf is None else _ptr(nf),
                                                bg.size, ptrs, _ptr(slots), C.byref(k))
        if rc != _lib.BUSY:
            check(rc, "server_acquire_many")
        views = [np.ctypeslib.as_array(ptrs[i], shape=((68 if bg[i] == 0 else 52) * int(z[i]),))
                 for i in range(k.value)]
        return slots[:k.value], views

    def submit_many(self, slots, srcs=None, tags=None):
        """srcs: None (each slot's pinned staging) or device-accessible addresses (ints)."""
        sl = np.ascontiguousarray(slots, np.uint32)
        sp = None if srcs is None else np.ascontiguousarray(srcs, np.uint64)
        tg = None if tags is None else np.ascontiguousarray(tags, np.uint64)
        check(lib().syndrome_server_submit_many(self._h, _ptr(sl), None if sp is None else _ptr(sp),
                                                None if tg is None else _ptr(tg), sl.size), "server_submit_many")

    def try_recv_many(self, cap=256):
        """Up to cap completed blocks as (slot, tag, iterations, hard copy, submit_ns, done_ns) tuples."""
        out = (_lib.Served * cap)()
        k = C.c_size_t(0)
        rc = lib().syndrome_server_try_recv_many(self._h, out, cap, C.byref(k))
        if rc == _lib.BUSY:
            return []
        check(rc, "server_try_recv_many")
        res = []
        for r in out[:k.value]:
            hard = np.ctypeslib.as_array(r.hard, shape=(r.n,)).copy()
            res.append((r.slot, r.tag, r.iterations, hard, r.submit_ns, r.done_ns))
        return res

    def release_many(self, slots):
        sl = np.ascontiguousarray(slots, np.uint32)
        check(lib().syndrome_server_release_many(self._h, _ptr(sl), sl.size), "server_release_many")

    def submit_tb(self, slots, tb_size, tb_out, tag):
        """Submit acquired, filled slots of one code as one transport block
        (syndrome_server_submit_tb).  tb_out: a HostPool of uint8, a device
        uint8 tensor, or a device-accessible address, of at least tb_size
        bytes.  Returns False when busy."""
        sl = np.ascontiguousarray(slots, np.uint32)
        if isinstance(tb_out, int):
            out = tb_out
        else:
            out = tb_out.ptr if isinstance(tb_out, HostPool) else tb_out.data_ptr()
        rc = lib().syndrome_server_submit_tb(self._h, _ptr(sl), None, sl.size, tb_size, out, tag)
        if rc == _lib.BUSY:
            return False
        check(rc, "server_submit_tb")
        return True

    def try_recv_tb(self):
        """Returns (tb, tag, crc_ok, max_iterations, submit_ns, done_ns) or None."""
        r = _lib.ServedTb()
        rc = lib().syndrome_server_try_recv_tb(self._h, C.byref(r))
        if rc == _lib.BUSY:
            return None
        check(rc, "server_try_recv_tb")
        return r.tb, r.tag, bool(r.crc_ok), r.max_iterations, r.submit_ns, r.done_ns

    def tb_block_crc(self, tb, num_cb):
        """Per-code-block CRC-24B outcome of a received transport block (DecodeReport::cb_crc)."""
        out = np.zeros(max(num_cb, 1), np.uint8)
        n = C.c_size_t(0)
        check(lib().syndrome_server_tb_block_crc(self._h, tb, _ptr(out), out.size, C.byref(n)), "server_tb_block_crc")
        return out[:n.value].astype(bool)

    def release_tb(self, tb):
        check(lib().syndrome_server_release_tb(self._h, tb), "server_release_tb")

    def stats(self):
        b, d, s = C.c_int64(), C.c_int64(), C.c_int32()
        check(lib().syndrome_server_stats(self._h, C.byref(b), C.byref(d), C.byref(s)), "server_stats")
        return {"batches": b.value, "decoded": d.value, "slots": s.value}


class HostPool:
    """Pinned, device-mapped host memory (syndrome_host_alloc) as a numpy array."""

    def __init__(self, count, dtype=np.float32):
        nbytes = int(count) * np.dtype(dtype).itemsize
        self._p = lib().syndrome_host_alloc(nbytes)
        if not self._p:
            raise SyndromeError(_lib.DEVICE_ERROR, "host_alloc")
        buf = (C.c_char * nbytes).from_address(self._p)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(count))

    @property
    def ptr(self):
        return self._p

    def __del__(self):
        try:
            if self._p:
                lib().syndrome_host_free(self._p)
                self._p = None
        except Exception:
            pass


def serve_loadgen(server, pool, cb_off, cb_bg, cb_z, cb_nfill, cb_info_bits, tb_first, tb_num, tb_rate, duration_s,
                  seed=1):
    """Run the C++ load generator (syndrome_server_loadgen) over a pool of
    decoder inputs: a HostPool (pinned host memory, read over PCIe) or a
    device tensor (inputs already resident in HBM).  Returns
    (cb_latency_ns, tb_latency_ns, info_bits_per_s, refused_acquires)."""
    pool_ptr = pool.ptr if isinstance(pool, HostPool) else pool.data_ptr()
    arrs = [np.ascontiguousarray(x, t) for x, t in ((cb_off, np.int64), (cb_bg, np.int32), (cb_z, np.int32),
                                                     (cb_nfill, np.int32), (cb_info_bits, np.int32),
                                                     (tb_first, np.int32), (tb_num, np.int32))]
    cap = 4_000_000
    cb_lat = np.zeros(cap, np.float64)
    tb_lat = np.zeros(cap // 4, np.float64)
    n_cb, n_tb = C.c_size_t(), C.c_size_t()
    bps = C.c_double()
    waits = C.c_int64()
    check(lib().syndrome_server_loadgen(server._h, pool_ptr, *[_ptr(a) for a in arrs], len(arrs[5]), tb_rate,
                                        duration_s, seed, _ptr(cb_lat), cap, C.byref(n_cb), _ptr(tb_lat), cap // 4,
                                        C.byref(n_tb), C.byref(bps), C.byref(waits)), "server_loadgen")
    return cb_lat[:min(n_cb.value, cap)], tb_lat[:min(n_tb.value, cap // 4)], bps.value, waits.value


def serve_loadgen_tb(server, pool, cb_off, cb_bg, cb_z, cb_nfill, tb_first, tb_num, tb_size, tb_out, tb_out_off,
                     tb_rate, duration_s, seed=1):
    """Transport-block load generator (syndrome_server_loadgen_tb) over a
    resident Server.  tb_out: device uint8 tensor or HostPool holding every
    pool transport block's payload at tb_out_off.  Returns (tb_latency_ns,
    code_blocks, info_bits_per_s, crc_failures, refused_acquires,
    goodput_bits_per_s, tb_crc_ok per latency entry)."""
    pool_ptr = pool.ptr if isinstance(pool, HostPool) else pool.data_ptr()
    out_ptr = tb_out.ptr if isinstance(tb_out, HostPool) else tb_out.data_ptr()
    arrs = [np.ascontiguousarray(x, t) for x, t in ((cb_off, np.int64), (cb_bg, np.int32), (cb_z, np.int32),
                                                     (cb_nfill, np.int32), (tb_first, np.int32), (tb_num, np.int32),
                                                     (tb_size, np.int32), (tb_out_off, np.int64))]
    cap = 2_000_000
    lat = np.zeros(cap, np.float64)
    ok = np.zeros(cap, np.uint8)
    n_tb, n_cb = C.c_size_t(), C.c_size_t()
    bps, good = C.c_double(), C.c_double()
    fails, waits = C.c_int64(), C.c_int64()
    check(lib().syndrome_server_loadgen_tb(server._h, pool_ptr, *[_ptr(a) for a in arrs[:7]], out_ptr, _ptr(arrs[7]),
                                           len(arrs[4]), tb_rate, duration_s, seed, _ptr(lat), _ptr(ok), cap,
                                           C.byref(n_tb), C.byref(n_cb), C.byref(bps), C.byref(good), C.byref(fails),
                                           C.byref(waits)),
          "server_loadgen_tb")
    m = min(n_tb.value, cap)
    return lat[:m], n_cb.value, bps.value, fails.value, waits.value, good.value, ok[:m].astype(bool)
