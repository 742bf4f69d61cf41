"""The correction step behind the bit-exact FP64 forms' RK4 division by 6
(gray_amd/csrc/fpdiv.h): for normal x, q = RN(x RN(1/6)) followed by
RN(q + RN(x - 6 q) RN(1/6)) with fused multiply-adds must equal the
correctly rounded x / 6.  The FMA is evaluated exactly with fractions and
rounded once (CPython's int / int true division is correctly rounded), and
compared with IEEE division on random doubles across the exponent range the
device path accepts, mantissas with long runs of ones or zeros, and small
multiples of 1/2."""
import random
import struct
from fractions import Fraction

Y = 1.0 / 6.0


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def div6(x):
    q = x * Y
    return fma(fma(-6.0, q, x), Y, q)


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def from_bits(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def test_div6_random_doubles():
    rng = random.Random(0x6)
    m52 = (1 << 52) - 1
    for i in range(200000):
        m = rng.getrandbits(52)
        if i % 3 == 1:
            m |= m52 & ~((1 << rng.randrange(24)) - 1)
        elif i % 3 == 2:
            m &= (1 << rng.randrange(24)) - 1
        e = 1023 + rng.randrange(-960, 1000)
        x = from_bits(m | (e << 52) | (rng.getrandbits(1) << 63))
        assert bits(div6(x)) == bits(x / 6.0), x.hex()


def test_div6_half_integers():
    for k in range(1, 20001):
        x = 0.5 * k
        assert div6(x) == x / 6.0 and div6(-x) == -x / 6.0, k


def test_naive_multiply_is_not_enough():
    # without the correction the product x RN(1/6) misses the correctly
    # rounded quotient for some x, which is why the device needs the FMA step
    rng = random.Random(7)
    misses = sum(1 for _ in range(20000)
                 if (lambda x: x * Y != x / 6.0)(rng.uniform(1.0, 2.0)))
    assert misses > 0
