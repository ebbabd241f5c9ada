"""Pins of the oracle's ChaCha block function and of the RNG layer built on it.

The block function is pinned by RFC 8439 section 2.3.2 / A.1 and by keystreams
that OpenSSL produced (tests/golden/make_chacha20_fixture.sh). The rand 0.9 /
rand_core 0.9 layer above it (seed_from_u64 via PCG32, u64 assembly, float
conversion) has no reference output in this environment. What the
mathematics fixes is checked: which lattice each call site's draws lie on
(2^-52 steps for random_range's [1, 2) construction, 2^-53 for StandardUniform)
with every mantissa bit in use. seed_from_u64's PCG32 expansion and the order
in which two keystream words form a u64 remain "parity unpinned": no output
of the reference's rand 0.9 stack exists here to compare against.
"""
import os

import numpy as np

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "chacha20_openssl.txt")


def words_le(hexstr):
    b = bytes.fromhex(hexstr)
    return [int.from_bytes(b[i:i + 4], "little") for i in range(0, len(b), 4)]


def test_rfc8439_block_function_vector():
    # RFC 8439 section 2.3.2: key 00..1f, counter 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00.
    # The oracle's layout puts a 64-bit counter in words 12-13 and a 64-bit stream id in
    # words 14-15, so word 13 (nonce bytes 0-3) is the counter's high half.
    key = words_le(bytes(range(32)).hex())
    counter = 1 | (0x09000000 << 32)
    stream = 0x4a000000
    out = O.chacha_block(key, counter, stream, rounds=20)
    expect = [0xe4e7f110, 0x15593bd1, 0x1fdd0f50, 0xc47120a3, 0xc7f4d1c7, 0x0368c033,
              0x9aaa2204, 0x4e6cd4c3, 0x466482d2, 0x09aa9f07, 0x05d7c214, 0xa2028bd9,
              0xd19c12b5, 0xb94e16de, 0xe883d0cb, 0x4e3c50a2]
    assert list(out) == expect


def test_openssl_keystreams():
    lines = open(GOLDEN).read().split("\n")
    assert len([l for l in lines if l.strip()]) == 4
    for line in lines:
        if not line.strip():
            continue
        key_hex, ctr_hex, ks_hex = line.split()
        key = words_le(key_hex)
        ctr = int.from_bytes(bytes.fromhex(ctr_hex), "little")
        ks = words_le(ks_hex)
        for blk in range(len(ks) // 16):
            got = O.chacha_block(key, ctr + blk, 0, rounds=20)
            assert list(got) == ks[16 * blk:16 * blk + 16], (key_hex, blk)


def test_chacha8_differs_only_by_rounds():
    key = list(range(8))
    b8 = O.chacha_block(key, 0, 0, rounds=8)
    b20 = O.chacha_block(key, 0, 0, rounds=20)
    assert not np.array_equal(b8, b20)
    # zero rounds leaves out = 2*input for every word (state + state)
    b0 = O.chacha_block(key, 7, 0, rounds=0)
    assert list(b0[4:12]) == [2 * k for k in key]
    assert b0[12] == 14


def _lattice(values, lo, step):
    """(v - lo) / step for every value, and whether all are integers."""
    k = (np.asarray(values, dtype=np.float64) - lo) / step
    return k, bool(np.all(k == np.floor(k)))


def test_range_draws_lie_on_the_2_52_lattice():
    """rand's random_range(lo..hi) for f64 builds a value in [1, 2) from the
    top 52 bits of a u64 (its mantissa), subtracts 1 and scales: every draw
    of random_range(-1.0..1.0) is -1 + k * 2^-51 and every draw of
    random_range(0.0..1.0) is k * 2^-52 for an integer k, with all 52 bits
    in use (the lowest lattice bit set in about half of the draws). A draw
    built as StandardUniform's (u >> 11) * 2^-53 (a different call), or from
    fewer bits, fails one of these."""
    r = O.Rng(9)
    xs = [r.random_range(-1.0, 1.0) for _ in range(4000)]
    zs = [r.random_range(0.0, 1.0) for _ in range(4000)]
    for vals, lo, step in ((xs, -1.0, 2.0**-51), (zs, 0.0, 2.0**-52)):
        k, integral = _lattice(vals, lo, step)
        assert integral
        odd = np.mean(np.mod(k, 2) == 1)
        assert 0.45 < odd < 0.55, odd
        assert min(vals) >= lo and max(vals) < lo + 2.0**52 * step


def test_standard_draws_lie_on_the_2_53_lattice():
    """StandardUniform f64 (what Plummer's random_iter draws, initialisers.rs:
    296-299) is (u >> 11) * 2^-53: every draw is k * 2^-53 in [0, 1), with
    the lowest bit set in about half of them; the [1, 2) construction of
    random_range (2^-52 steps) would never set it."""
    r = O.Rng(5)
    vals = [r.random_f64() for _ in range(4000)]
    k, integral = _lattice(vals, 0.0, 2.0**-53)
    assert integral
    assert 0.45 < np.mean(np.mod(k, 2) == 1) < 0.55
    assert min(vals) >= 0.0 and max(vals) < 1.0


def test_cube_positions_come_from_range_draws():
    """The cube generator's coordinates (size 1, centre 0) are the three
    random_range draws of Entity::random (physim-core/src/lib.rs:115-118):
    x, y on the -1 + k 2^-51 lattice, z on the k 2^-52 lattice, each with
    its lowest bit in use; so is a shifted, scaled cube's (size 0.5, centre
    (0.25, -0.5, 2)) once the exact power-of-two scaling is undone."""
    s = O.gen_cube(3000, seed=1, spin=1000.0)
    for ax, lo, step in (("x", -1.0, 2.0**-51), ("y", -1.0, 2.0**-51), ("z", 0.0, 2.0**-52)):
        k, integral = _lattice(s[ax], lo, step)
        assert integral, ax
        assert 0.4 < np.mean(np.mod(k, 2) == 1) < 0.6, ax
    t = O.gen_cube(3000, seed=1, spin=1000.0, size=0.5, centre=(0.25, -0.5, 2.0))
    # x * 0.5 + 0.25 and y * 0.5 - 0.5 are exact on these lattices; z * 0.5 + 2.0 rounds
    np.testing.assert_array_equal((t["x"] - 0.25) * 2.0, s["x"])
    np.testing.assert_array_equal((t["y"] + 0.5) * 2.0, s["y"])


def test_cube_generator_layout():
    """RandomCube: 3 draws per body, x,y in [-1,1), z in [0,1), v = (y, -x) * spin,
    mass = total / n (astro/src/initialisers.rs:81-107)."""
    s = O.gen_cube(1000, seed=1, spin=1000.0)
    r = O.Rng(1)
    for i in range(50):
        x = r.random_range(-1.0, 1.0)
        y = r.random_range(-1.0, 1.0)
        z = r.random_range(0.0, 1.0)
        assert (s["x"][i], s["y"][i], s["z"][i]) == (x, y, z)
        assert s["vx"][i] == y * 1000.0 and s["vy"][i] == -x * 1000.0 and s["vz"][i] == 0.0
    assert np.all(s["mass"] == 1.0 / 1000)
    assert np.all(s["radius"] == 0.02)
    assert s["x"].min() >= -1 and s["x"].max() < 1 and s["z"].min() >= 0 and s["z"].max() < 1
