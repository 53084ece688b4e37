"""The FP64 kernels divide several numerators by one denominator with the
denominator's part of the correctly rounded division (scaling, reciprocal,
two Newton steps) formed once.  Every quotient must keep the bits of plain
division: checked on random operands over exponents [-126, 126] with signed
zeros among the numerators, on the device."""
import pytest

import bcm_cgli_amd as B

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 20260503])
def test_shared_denominator_division_is_bitwise_plain_division(seed):
    assert B.selftest_division(1 << 24, seed=seed) == 0
