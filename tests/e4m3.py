"""Test-side OCP e4m3fn quantiser (numpy, float64 in and out).

e4m3fn: 1 sign bit, 4 exponent bits (bias 7), 3 mantissa bits, normals from
2^-6 up to 448, subnormals in steps of 2^-9, no infinities.  Rounding is to
nearest with ties to even.  tests/test_oracle.py pins this function against
torch's float8_e4m3fn conversion (ties and subnormals included); the FP8
parity tests then use it as the reference quantiser.
"""
import numpy as np

E4M3_MAX = 448.0


def e4m3_rne(y):
    """round float64 values to e4m3fn, nearest-even; |y| <= 448 assumed"""
    a = np.abs(y)
    E = np.floor(np.log2(np.maximum(a, 2.0 ** -6)))
    q = 2.0 ** (E - 3)
    return np.sign(y) * np.rint(a / q) * q
