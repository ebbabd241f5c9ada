"""The sample positions of the device's sample sort (k_keys, k_splitters,
k_splitters_wide) are floor(a * b / c) with a * b < 2^53, computed as the
floor of one correctly rounded FP64 division (muldiv_floor) instead of a
64-bit integer division. The claim: both operands are exact in FP64, a
quotient that is not an integer lies at least 1/c from the integers around
it, and the division's rounding moves it by at most q 2^-53 < 1/c, so the
floor is exact. Checked here against Python's integer floor division, on
random operands over the whole range and on adversarial ones: products one
below, at and one above a multiple of c, quotients next to powers of two, and
products at the 2^53 bound. Python floats divide with IEEE round-to-nearest,
as the device's FP64 division does."""
import math
import random

import pytest

BOUND = 1 << 53


def muldiv_floor(a, b, c):
    return int(math.floor(float(a * b) / float(c)))


def check(a, b, c):
    assert a * b < BOUND and c > 0
    assert muldiv_floor(a, b, c) == (a * b) // c, (a, b, c)


def test_random_operands_in_range():
    rng = random.Random(20261203)
    for _ in range(200_000):
        # the device's shapes: an index below 2^21 (a sample, quantile or
        # bucket position) times a body count below 2^32, over a count of
        # samples or buckets
        b = rng.randrange(1, 1 << 32)
        a = rng.randrange(0, min(1 << 21, (BOUND - 1) // b + 1))
        c = rng.randrange(1, 1 << 21)
        check(a, b, c)


def test_products_next_to_multiples():
    rng = random.Random(7)
    for _ in range(50_000):
        c = rng.randrange(2, 1 << 21)
        k = rng.randrange(1, (BOUND - 2) // c)
        for p in (k * c - 1, k * c, k * c + 1):
            if 0 < p < BOUND:
                check(p, 1, c)


@pytest.mark.parametrize("c", [1, 2, 3, 7, 255, 256, 257, 1023, 1024, 1025, 8191, 8192,
                               (1 << 21) - 1, 1 << 21])
def test_products_at_the_bound(c):
    for p in (BOUND - 1, BOUND - 2, BOUND - c, BOUND - c - 1, (BOUND // c) * c,
              (BOUND // c) * c - 1):
        if 0 < p < BOUND:
            check(p, 1, c)


def test_quotients_next_to_powers_of_two():
    for e in range(1, 52):
        for c in (3, 5, 7, 1023, 8191, (1 << 21) - 1):
            for p in ((1 << e) * c - 1, (1 << e) * c, (1 << e) * c + 1):
                if 0 < p < BOUND:
                    check(p, 1, c)


def test_the_claim_needs_its_bound():
    # beyond 2^53 the product is no longer exact in FP64 and the shortcut
    # fails: the range check in the kernels' static_assert is what keeps
    # every position below it
    c = 3
    p = (1 << 60) - 1
    assert int(math.floor(float(p) / c)) != p // c
