"""Pins of the oracle's ChaCha block function and of the RNG layer built on it.

The block function is pinned by RFC 8439 section 2.3.2 / A.1 and by keystreams
that OpenSSL produced (tests/golden/make_chacha20_fixture.sh). The rand 0.9 /
rand_core 0.9 layer above it (seed_from_u64 via PCG32, u64 assembly, float
conversion) has no reference output in this environment: it is "parity
unpinned", and only its documented structure is checked here.
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


def test_rng_stream_is_the_chacha8_keystream():
    """next_u64 joins two keystream words (low first) of ChaCha8 blocks at
    consecutive counters under the PCG32-expanded key."""
    r = O.Rng(1)
    u = [r.next_u64() for _ in range(40)]
    # recompute the key the documented way: PCG32 outputs as key words
    state = 1
    mul, inc = 6364136223846793005, 11634580027462260723
    key = []
    for _ in range(8):
        state = (state * mul + inc) % 2**64
        xs = (((state >> 18) ^ state) >> 27) & 0xFFFFFFFF
        rot = state >> 59
        key.append(((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF)
    words = np.concatenate([O.chacha_block(key, c, 0, rounds=8) for c in range(5)])
    expect = [int(words[2 * i]) | (int(words[2 * i + 1]) << 32) for i in range(40)]
    assert u == expect


def test_float_draws_structure():
    r1, r2 = O.Rng(5), O.Rng(5)
    for _ in range(100):
        u = r2.next_u64()
        f = r1.random_f64()
        assert f == (u >> 11) * 2.0**-53
    r1, r2 = O.Rng(9), O.Rng(9)
    for _ in range(100):
        u = r2.next_u64()
        f = r1.random_range(-1.0, 1.0)
        m = (u >> 12) * 2.0**-52
        assert f == m * 2.0 - 1.0
        assert -1.0 <= f < 1.0


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
