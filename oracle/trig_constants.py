"""Constants of the trigonometric functions the oracle and the GPU library
both implement (orc_sin/orc_cos/orc_atan2, edmd_sin/edmd_cos/edmd_atan2):
pi/2 split into a 33-bit head, a 33-bit middle and a 53-bit tail (so that
k * head and k * middle are exact for |k| <= 2), pi and pi/2 as head + tail,
and atan(0.4), atan(0.8) of the doubles nearest 0.4 and 0.8, each computed
with 60 decimal digits and printed as C hex literals.  Run to reprint."""
from decimal import Decimal, getcontext
from fractions import Fraction

getcontext().prec = 80


def pi():
    """Machin's formula, pi = 16 atan(1/5) - 4 atan(1/239)."""
    def atan_inv(n):
        x = Decimal(1) / n
        x2 = x * x
        s, t, k = Decimal(0), x, 1
        while True:
            term = t / k
            if abs(term) < Decimal(10) ** -75:
                break
            s += term if (k // 2) % 2 == 0 else -term
            t *= x2
            k += 2
        return s
    return 16 * atan_inv(5) - 4 * atan_inv(239)


def atan(x):
    """atan by the series after halving the argument twice:
    atan(x) = 2 atan(x / (1 + sqrt(1 + x^2)))."""
    x = Decimal(x)
    mult = 1
    for _ in range(3):
        x = x / (1 + (1 + x * x).sqrt())
        mult *= 2
    x2 = x * x
    s, t, k = Decimal(0), x, 1
    while True:
        term = t / k
        if abs(term) < Decimal(10) ** -75:
            break
        s += term if (k // 2) % 2 == 0 else -term
        t *= x2
        k += 2
    return mult * s


def to_double(d, bits=53):
    """d rounded to nearest with `bits` significant bits (as a float)."""
    f = Fraction(d)
    if f == 0:
        return 0.0
    e = 0
    a = abs(f)
    while a >= 2:
        a /= 2
        e += 1
    while a < 1:
        a *= 2
        e -= 1
    m = round(a * 2 ** (bits - 1))
    return float(Fraction(m * (1 if f > 0 else -1), 2 ** (bits - 1)) * Fraction(2) ** e)


def trunc_bits(d, bits):
    """d truncated to `bits` significant bits."""
    f = Fraction(d)
    a, e = abs(f), 0
    while a >= 2:
        a /= 2
        e += 1
    while a < 1:
        a *= 2
        e -= 1
    m = int(a * 2 ** (bits - 1))
    return float(Fraction(m, 2 ** (bits - 1)) * Fraction(2) ** e)


if __name__ == "__main__":
    p = pi()
    h = p / 2
    p1 = trunc_bits(h, 33)
    p2 = trunc_bits(h - Decimal(p1), 33)
    p3 = to_double(h - Decimal(p1) - Decimal(p2))
    print("PIO2_1 %s\nPIO2_2 %s\nPIO2_3 %s" % (p1.hex(), p2.hex(), p3.hex()))
    ph = to_double(h)
    print("PIO2_HI %s\nPIO2_LO %s" % (ph.hex(), to_double(h - Decimal(ph)).hex()))
    pp = to_double(p)
    print("PI_HI %s\nPI_LO %s" % (pp.hex(), to_double(p - Decimal(pp)).hex()))
    print("TWO_OVER_PI %s" % to_double(2 / p).hex())
    for c in (0.4, 0.8):
        print("ATAN_%s %s" % (str(c).replace(".", ""), to_double(atan(Decimal(c))).hex()))
