"""The shading kernels evaluate math.Pow's integer-power loop without its
frexp scaling when x lies in a range the host computes (every intermediate
stays within 2^-1000 .. 2^1000).  This checks the claim that makes that
exact -- in the normal range, scaling by a power of two commutes with
rounding, so both forms give the same double -- on the CPU with Python
floats (IEEE binary64, as on the device), for the powers shaders use and
random ones, including fractional and negative exponents."""
import math
import random


def go_pow_loop(x, y):
    """math.Pow from the Modf step on (x > 0, finite, y not a special case)."""
    yi, yf = math.modf(abs(y))[1], math.modf(abs(y))[0]
    a1, ae = 1.0, 0
    if yf != 0:
        if yf > 0.5:
            yf -= 1
            yi += 1
        a1 = math.exp(yf * math.log(x))
    x1, xe = math.frexp(x)
    i = int(yi)
    while i != 0:
        if xe < -(1 << 12) or (1 << 12) < xe:
            ae += xe
            break
        if i & 1:
            a1 *= x1
            ae += xe
        x1 *= x1
        xe <<= 1
        if x1 < 0.5:
            x1 += x1
            xe -= 1
        i >>= 1
    if y < 0:
        a1 = 1 / a1
        ae = -ae
    return math.ldexp(a1, ae)


def bounds(y):
    yi = math.modf(abs(y))[1]
    if math.modf(abs(y))[0] > 0.5:
        yi += 1
    L = min(1000.0 / (yi + 0.5), 1000.0)
    return 2.0 ** -L, 2.0 ** L, int(yi)


def fast_pow(x, y):
    yf = math.modf(abs(y))[0]
    _, _, yi = bounds(y)
    if yf > 0.5:
        yf -= 1
    a1 = math.exp(yf * math.log(x)) if yf != 0 else 1.0
    x1 = x
    i = yi
    while i != 0:
        if i & 1:
            a1 *= x1
        i >>= 1
        if i != 0:
            x1 *= x1
    return 1 / a1 if y < 0 else a1


def test_fast_form_is_exact_in_range():
    rng = random.Random(7)
    powers = [32.0, 2.0, 3.0, 7.0, 10.0, 64.0, 100.0, 255.0, 1000.0, 12345.0, 2.5, 0.75, 33.3, -4.0, -32.5, 0.25]
    powers += [rng.uniform(-300, 300) for _ in range(40)]
    checked = 0
    for y in powers:
        lo, hi, _ = bounds(y)
        xs = [lo, hi, 1.0, math.nextafter(1.0, 2.0), math.nextafter(1.0, 0.0)]
        xs += [math.exp(rng.uniform(math.log(lo), math.log(hi))) for _ in range(400)]
        xs += [rng.uniform(0.0, 1.0) for _ in range(200)]
        for x in xs:
            if not (lo <= x <= hi) or x == 1.0:
                continue
            assert fast_pow(x, y) == go_pow_loop(x, y), (x, y)
            checked += 1
    assert checked > 20000


def test_outside_range_the_forms_can_differ():
    """With y = 1.75 (yi = 2, yf = -0.25) the loop squares x, and x**2 can
    be subnormal while x**1.75 is normal: the frexp form keeps the square's
    53 bits, the unscaled one does not.  The range check excludes such x."""
    rng = random.Random(3)
    lo, _, _ = bounds(1.75)
    diffs = 0
    for _ in range(200):
        x = 2.0 ** -584 * rng.uniform(1.0, 1.5)
        assert x < lo and x * x < 2.0 ** -1022 and go_pow_loop(x, 1.75) >= 2.0 ** -1022
        diffs += fast_pow(x, 1.75) != go_pow_loop(x, 1.75)
    assert diffs > 100
