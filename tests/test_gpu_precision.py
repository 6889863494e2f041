"""Precision forms against the FP64 sweep on the GPU: FP32 storage and
arithmetic stays within 1e-5 (max relative error per point, floor = RMS of
the FP64 field) after 1, 10 and 100 sweeps; BF16 storage with FP32
accumulation is bounded by its 8-bit significand; the FP64 comparisons
themselves are exact on identical fields."""
import pytest

from bricklib_amd import BK_BF16, BK_F32, precision

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("stencil", ["7pt", "125pt", "25pt"])
def test_fp32_within_1e5_of_fp64(gpu, stencil):
    r = precision.study(64, stencil, dtypes=(BK_F32,), sweeps=(1, 10, 100))
    for k, v in r.items():
        print(stencil, k, v)
        if v["underflow"]:
            # the FP64 field left FP32's range: the FP32 form must be ~0 too
            assert v["l2"] <= 1e-30, (k, v)
            continue
        assert v["max_rel"] <= 1e-5, (k, v)
        assert v["rms_rel"] <= 1e-6, (k, v)


def test_bf16_storage_error_is_bounded(gpu):
    r = precision.study(64, "7pt", dtypes=(BK_BF16,), sweeps=(1, 10))
    for k, v in r.items():
        print(k, v)
        # each sweep rounds its output to BF16 (relative 2^-9); the averaging
        # stencil keeps the accumulated error within a few roundings
        assert v["max_rel"] <= 2 ** -5, (k, v)
        assert v["rms_rel"] <= 2 ** -7, (k, v)


def test_fp64_self_comparison_is_zero(gpu):
    g = gpu.PeriodicGrid((4, 4, 4))
    st = gpu.Stencil.make("7pt")
    a, keep = precision.run_sweeps(g, st, 0, 3, 5)
    b, keep2 = precision.run_sweeps(g, st, 0, 3, 5)
    c = precision.compare(g, a, b)
    assert c["max_abs"] == 0.0 and c["max_rel"] == 0.0 and c["l2"] == c["l2_ref"]
