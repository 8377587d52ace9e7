"""The kernels' fp32 division by cs2, 2 cs2 and rho_eff cs2 is one fp64
product with the divisor's fp64 reciprocal, rounded to fp32
(pytorch-lbm_amd/csrc/lbm_kernels.hpp, divk).  The argument that this equals
IEEE division for every fp32 x whenever d is not an even integer is checked
here on the host against numpy's IEEE float32 division: every special value
(zeros, subnormals, the largest finite values, infinities) and 2^22 random bit
patterns per divisor.  The GPU test test_gpu_parity.py::
test_division_by_constants_is_ieee_exact covers all 2^32 inputs on the
device."""
import numpy as np
import pytest

F32 = np.float32
CS2 = F32(1.0 / 3.0)
DIVISORS = [CS2, F32(2.0) * CS2, F32(0.5) * CS2, F32(0.7312) * CS2, F32(0.9999) * CS2,
            F32(1.0 / 828.0) * CS2, F32(1.0), F32(3.0), F32(5.0)]


def product_form(x, d):
    y = 1.0 / np.float64(d)
    with np.errstate(over="ignore", invalid="ignore"):
        return (x.astype(np.float64) * y).astype(np.float32)


def special_values():
    b = [0x00000000, 0x80000000, 0x00000001, 0x80000001, 0x007FFFFF, 0x807FFFFF,
         0x00800000, 0x80800000, 0x7F7FFFFF, 0xFF7FFFFF, 0x7F800000, 0xFF800000,
         0x3F800000, 0xBF800000, 0x0DA24260, 0x8DA24260, 0x71B4CE6C]
    # every subnormal step near the ends, and the binades around 2^-100 / 2^100
    b += list(range(0x00000000, 0x00001000)) + list(range(0x007FF000, 0x00801000))
    b += list(range(0x0D7FF000, 0x0D801000)) + list(range(0x717FF000, 0x71801000))
    b += list(range(0x7F7FF000, 0x7F800000))
    u = np.array(b, dtype=np.uint32)
    return np.concatenate([u, u | np.uint32(0x80000000)]).view(np.float32)


@pytest.mark.parametrize("d", DIVISORS, ids=lambda d: f"{float(d):.6g}")
def test_product_form_equals_ieee_division(d):
    rng = np.random.default_rng(int(d.view(np.uint32)))
    x = np.concatenate([special_values(),
                        rng.integers(0, 1 << 32, 1 << 22, dtype=np.uint64).astype(np.uint32)
                        .view(np.float32)])
    x = x[~np.isnan(x)]
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        ref = x / d
    got = product_form(x, d)
    bad = np.flatnonzero(ref.view(np.uint32) != got.view(np.uint32))
    assert bad.size == 0, [(hex(int(v)), float(r), float(g)) for v, r, g in
                           zip(x[bad[:5]].view(np.uint32), ref[bad[:5]], got[bad[:5]])]


def test_even_integer_divisor_tie_exists_in_principle():
    """The one case the argument leaves out: for d an even integer that is not
    a power of two, x / d can equal a midpoint of the subnormal grid exactly
    (x = 3 * 2^-149, d = 6: x / d = 2^-150, the midpoint between 0 and the
    smallest subnormal).  Such divisors use the Markstein form instead."""
    x = np.array([3], dtype=np.uint32).view(np.float32)
    q = np.float64(x[0]) / 6.0
    assert q == 2.0 ** -150                       # exactly on the boundary
    assert (x / F32(6.0))[0] == F32(0.0)          # ties to even: 0
