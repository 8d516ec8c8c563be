"""The written-out correctly rounded division and square root the kernels use (pbf_math.h) equal
the compiler's '/' and sqrtf bit for bit on the device, over 2^28 random inputs per check."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def test_written_out_division_and_sqrt_are_correctly_rounded():
    import torch
    if torch.cuda.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    from akuaengine_amd import _lib
    out = (ctypes.c_uint64 * 4)()
    _lib.check(_lib.load().akua_selftest_arith(1 << 28, 12345, out))
    assert list(out) == [0, 0, 0, 0], list(out)
