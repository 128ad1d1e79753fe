"""Host-side validation of device buffers (CPU only, no GPU needed).

Every Python entry point that hands tensors to a kernel checks them first:
dtype, contiguity, element count for the batch or plan, and device
residency, in that order, so that a wrong tensor raises ValueError instead
of becoming out-of-bounds device reads and writes.  These tests run the
checks on CPU tensors: each mistake raises its own error, and a correct set
fails only on residency.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from syndrome_amd import _lib as L  # noqa: E402
from syndrome_amd.ldpc import DecodePlan, check_batch_tensors, check_plan_tensors  # noqa: E402

N, EZ, B = 26112, 121344, 3


def _ok_f32():
    return dict(d_in=torch.zeros(B * N), d_post=torch.zeros(B * N), d_hard=torch.zeros(B * N, dtype=torch.uint8),
                d_iters=torch.zeros(B, dtype=torch.int32), d_edge=None)


def _expect(match, **kw):
    args = _ok_f32()
    args.update(kw)
    flags = args.pop("flags", L.WRITE_POST | L.WRITE_HARD)
    int8 = args.pop("int8", False)
    with pytest.raises(ValueError, match=match):
        check_batch_tensors("decode_batch", N, EZ, int8, flags, **args)


def test_correct_cpu_tensors_fail_only_on_residency():
    _expect("device memory")


def test_wrong_dtype_rejected():
    _expect("d_in must be float32", d_in=torch.zeros(B * N, dtype=torch.float64))
    _expect("d_hard must be uint8", d_hard=torch.zeros(B * N, dtype=torch.int8))
    _expect("d_iters must be int32", d_iters=torch.zeros(B, dtype=torch.int64))
    _expect("d_in must be int16", int8=True)  # int8 path without QUANTIZE takes i16
    _expect("QUANTIZE is an int8 flag", flags=L.WRITE_POST | L.QUANTIZE)


def test_short_buffers_rejected():
    _expect("d_in holds", d_in=torch.zeros(B * N - 1))
    _expect("d_post holds", d_post=torch.zeros(N))
    _expect("d_hard holds", d_hard=torch.zeros(B * N - 4, dtype=torch.uint8))
    _expect("d_edge holds", d_edge=torch.zeros(B * EZ - 1), flags=L.WRITE_POST | L.PLAIN)
    _expect("d_edge must be a torch tensor", flags=L.WRITE_POST | L.WRITE_EDGE)  # required, missing


def test_non_contiguous_rejected():
    _expect("d_in must be contiguous", d_in=torch.zeros(B * N * 2)[::2])
    _expect("d_post must be a torch tensor", d_post=np.zeros(B * N, np.float32))


def test_unused_outputs_may_be_absent():
    # no WRITE_POST / WRITE_HARD: only the input and the iteration counts are read
    _expect("device memory", d_post=None, d_hard=None, flags=L.INIT_5G)


def test_plan_arena_checks():
    arena, count = 10_000, 4
    ok = dict(d_in=torch.zeros(arena), d_post=torch.zeros(arena), d_hard=torch.zeros(arena, dtype=torch.uint8),
              d_iters=torch.zeros(count, dtype=torch.int32))
    flags = L.WRITE_POST | L.WRITE_HARD
    with pytest.raises(ValueError, match="device memory"):
        check_plan_tensors("run", arena, count, False, flags, **ok)
    with pytest.raises(ValueError, match="d_in holds"):
        check_plan_tensors("run", arena, count, False, flags, **dict(ok, d_in=torch.zeros(arena - 1)))
    with pytest.raises(ValueError, match="d_iters holds"):
        check_plan_tensors("run", arena, count, False, flags, **dict(ok, d_iters=torch.zeros(3, dtype=torch.int32)))
    with pytest.raises(ValueError, match="d_post must be int16"):
        check_plan_tensors("run", arena, count, True, flags | L.QUANTIZE, **ok)


def test_plan_arrays_of_unequal_length_rejected():
    with pytest.raises(ValueError, match="one length"):
        DecodePlan([0, 1], [384], [0, 0], [0, 26112], 0.5)
    with pytest.raises(ValueError, match="one length"):
        DecodePlan([0, 1], [384, 208], [0], [0, 26112], 0.5)
