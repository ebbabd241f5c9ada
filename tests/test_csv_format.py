"""csvsink's number format (utilities/src/csvsink.rs:73-79 prints each
coordinate with Rust's `{}`): the library's pg_format_f64 against an
independent shortest round-trip formatter (Python's float repr, David Gay's
algorithm) written out positionally -- Rust's Display uses no exponent and
no trailing ".0"; NaN, inf, -inf and -0 as Rust prints them. Host code only:
no GPU is touched."""
from decimal import Decimal

import numpy as np

import physim_amd as P


def rust_display(x):
    """Rust's {} for f64 restated from its definition: the shortest digits
    that round-trip (Python repr), positional, minimal."""
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "inf" if x > 0 else "-inf"
    return format(Decimal(repr(x)).normalize(), "f")


def test_known_values():
    cases = [(0.0, "0"), (-0.0, "-0"), (1.0, "1"), (-2.5, "-2.5"), (0.1, "0.1"),
             (1e-7, "0.0000001"), (100.0, "100"), (1e21, "1000000000000000000000"), (0.3, "0.3"),
             (1.0 / 3.0, "0.3333333333333333"), (5e-324, "0." + "0" * 323 + "5"),
             (1.7976931348623157e308, "17976931348623157" + "0" * 292),
             (float("nan"), "NaN"), (float("inf"), "inf"), (float("-inf"), "-inf"),
             (0.2 + 0.1, "0.30000000000000004"), (123456.789, "123456.789")]
    for x, want in cases:
        assert P.format_f64(x) == want, (x, P.format_f64(x), want)


def test_random_doubles_match_shortest_positional():
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2**63, 20000, dtype=np.uint64) | (
        rng.integers(0, 2, 20000, dtype=np.uint64) << np.uint64(63))
    vals = bits.view(np.float64)
    vals = np.concatenate([vals[np.isfinite(vals)], rng.uniform(-2, 2, 5000),
                           rng.uniform(-1, 1, 5000) * 10.0 ** rng.integers(-30, 30, 5000)])
    for x in vals:
        x = float(x)
        s = P.format_f64(x)
        assert s == rust_display(x), (x, s, rust_display(x))
        assert float(s) == x
