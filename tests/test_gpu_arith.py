"""Device arithmetic the substep kernels rely on: the force pass's reciprocal
(hardware reciprocal + one Newton-Raphson step with a fused residual) equals
the IEEE division 1.0f / x for every float in [2^-100, 2^100], the range that
the smoothing-radius bounds keep the force terms' divisors in."""
import ctypes as C
import struct
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

LIB = Path(__file__).resolve().parent / "arith" / "libarith_check.so"


def _bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def _check(lo: float, hi: float, mode: int):
    if not LIB.exists():
        raise RuntimeError(f"{LIB} is not built (tests/arith/Makefile)")
    lib = C.CDLL(str(LIB))
    lib.rcp_check.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint32)]
    bad, first = C.c_ulonglong(0), C.c_uint32(0)
    rc = lib.rcp_check(_bits(lo), _bits(hi) + 1, mode, C.byref(bad), C.byref(first))
    assert rc == 0, f"HIP error {rc}"
    return bad.value, struct.unpack("<f", struct.pack("<I", first.value))[0]


def test_reciprocal_is_the_ieee_division_over_its_range():
    lo, hi = 2.0 ** -100, 2.0 ** 100
    print(f"case rcp_rn: every float in [2^-100, 2^100] ({_bits(hi) - _bits(lo) + 1} values)", flush=True)
    bad, first = _check(lo, hi, 0)
    assert bad == 0, f"{bad} floats differ from 1.0f / x, first {first!r}"


def test_bare_hardware_reciprocal_is_caught():
    # the same comparison over [1, 2] flags the unrefined reciprocal
    bad, first = _check(1.0, 2.0, 1)
    assert bad > 1000 and 1.0 <= first <= 2.0

