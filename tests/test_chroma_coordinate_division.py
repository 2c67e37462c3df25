"""The YUV->RGB kernel's chroma coordinate floor(n / d) (color.hip div_floor: a binary64 estimate
n * (1.0 / d), truncated, corrected by one from the remainder) equals integer division for every
numerator below 2^32 and divisor up to 65535: random pairs and the pairs around every multiple
boundary where an estimate is most likely off, for the largest divisors and numerators."""
import numpy as np


def div_floor(n, d, inv):
    q = int(float(n) * inv)
    r = n - q * d
    if r < 0:
        q -= 1
    elif r >= d:
        q += 1
    return q


def test_div_floor_matches_integer_division():
    rng = np.random.default_rng(5)
    ds = np.concatenate([rng.integers(1, 65536, 3000), np.arange(1, 200), np.arange(65335, 65536)])
    for d in ds:
        d = int(d)
        inv = 1.0 / d
        ns = list(rng.integers(0, 2 ** 32, 60))
        top = (2 ** 32 - 1) // d
        for k in (1, 2, 3, top - 1, top, int(rng.integers(1, top + 1))):
            ns += [k * d - 1, k * d, k * d + 1]
        for n in ns:
            n = int(n)
            if 0 <= n < 2 ** 32:
                assert div_floor(n, d, inv) == n // d, (n, d)
