"""The shortcut behind the device's Morton keys (k_keys): a coordinate p that
lies farther than 2^(lmax - 41) from every cell boundary, in units of the
level-lmax cell (w = (p / E + 1) 2^(lmax - 1)), has the octant digits of the
reference's FP64 descent (octree.rs:160-215 / quadtree.rs:161-194) in the
bits of floor(w). Checked here against that descent, restated level by level
over many coordinates at once: random ones, and adversarial ones on cell
boundaries and one ulp either side of them (which the shortcut must refuse,
leaving them to the descent). IEEE float64 in numpy rounds as the device does
(the device code is built without contraction)."""
import numpy as np
import pytest


def descent_digits(p, ext, lmax):
    """Per-level digits (p > c) of the FP64 descent along one axis, as the
    integer whose bit (lmax - 1 - l) is the level-l digit."""
    c = np.zeros_like(p)
    h = np.full_like(p, ext)
    q = np.zeros(p.shape, dtype=np.uint64)
    for _ in range(lmax):
        d = p > c
        h = h * 0.5
        c = np.where(d, c + h, c - h)
        q = (q << np.uint64(1)) | d.astype(np.uint64)
    return q


def shortcut(p, ext, lmax):
    inv = 1.0 / ext
    w = (p * inv + 1.0) * np.ldexp(1.0, lmax - 1)
    f = w - np.floor(w)
    margin = np.ldexp(1.0, lmax - 41)
    ok = (f > margin) & (f < 1.0 - margin)
    q = np.where(ok, w, 0.0).astype(np.uint64)
    return ok, q


def adversarial(ext, lmax, rng, n):
    """Coordinates on boundaries of every level, and their neighbours."""
    lev = rng.integers(1, lmax + 1, n)
    k = rng.integers(0, 2 ** 20, n).astype(np.float64)
    frac = np.mod(k, np.ldexp(1.0, lev)) / np.ldexp(1.0, lev)  # in [0, 1)
    p = ext * (2.0 * frac - 1.0)
    return np.concatenate([p, np.nextafter(p, np.inf), np.nextafter(p, -np.inf),
                           np.nextafter(np.nextafter(p, np.inf), np.inf)])


@pytest.mark.parametrize("lmax", [21, 32, 10])
@pytest.mark.parametrize("ext", [1.0, 0.7312345678901234, 3.1e5, 2.2e-7])
def test_shortcut_agrees_with_descent(lmax, ext):
    rng = np.random.default_rng(lmax * 7 + int(ext * 1000) % 97)
    p = np.concatenate([rng.uniform(-ext, ext, 200000), adversarial(ext, lmax, rng, 20000),
                        [ext, -ext, 0.0, np.nextafter(ext, 0.0), np.nextafter(-ext, 0.0)]])
    ok, q = shortcut(p, ext, lmax)
    ref = descent_digits(p, ext, lmax)
    assert np.array_equal(q[ok], ref[ok])
    # on a boundary the shortcut always refuses
    on = adversarial(ext, lmax, rng, 2000)[:2000]
    assert not shortcut(on, ext, lmax)[0].any()
    # and random coordinates almost never need the descent
    r_ok, _ = shortcut(rng.uniform(-ext, ext, 100000), ext, lmax)
    assert r_ok.mean() > 1.0 - 1e-2


def test_descent_restatement_pinned():
    """The vectorised descent matches a scalar one written as the reference
    states it (a child's centre is the parent's +/- extent / 2.0)."""
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1.0, 1.0, (200, 3))
    ext = float(np.abs(pts).max())
    qs = [descent_digits(pts[:, a].copy(), ext, 21) for a in range(3)]
    for i in range(len(pts)):
        # independent scalar descent of all three axes together
        c = [0.0, 0.0, 0.0]
        h = ext
        want = [0, 0, 0]
        for _ in range(21):
            o = [pts[i, a] > c[a] for a in range(3)]
            h = h / 2.0
            for a in range(3):
                c[a] = c[a] + h if o[a] else c[a] - h
                want[a] = (want[a] << 1) | int(o[a])
        assert [int(qs[a][i]) for a in range(3)] == want
