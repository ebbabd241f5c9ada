"""The allocation-free transform entry point (pg_transform_aos_noalloc)
against pg_transform_aos, through the C-ABI: bit-identical accelerations,
accumulated into the caller's array, on consecutive calls of falling size
(the static device array keeps the previous call's contents), for every
kind and precision; the size limit is a clear error, not a fault."""
import numpy as np
import pytest

import physim_amd as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAX_BODIES = 65536


def prefilled(n, seed):
    """An Acceleration AoS holding arbitrary values, so that += is checked."""
    a = np.zeros(n, P.ACCEL)
    v = np.random.default_rng(seed).uniform(-1.0, 1.0, (n, 3))
    a["x"], a["y"], a["z"] = v[:, 0], v[:, 1], v[:, 2]
    return a


@pytest.mark.parametrize("kind", ["astro2", "astro", "simple_astro"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_equals_transform(engine, kind, precision):
    kw = dict(kind=kind, theta=0.9, e=0.25, precision=precision)
    for seed, n in ((3, 40000), (4, 5000)):   # the second call is smaller
        s = P.cube_and_stars(n, seed=seed)
        ref = engine.transform(s, acc=prefilled(len(s), seed), **kw)
        got = engine.transform_noalloc(s, acc=prefilled(len(s), seed), **kw)
        assert got.tobytes() == ref.tobytes(), (kind, precision, n)


def test_noalloc_largest_input_and_limit(engine):
    s = P.gen_plummer_virial(MAX_BODIES, seed=9, mass=1.0, a=0.1)
    kw = dict(kind="astro2", theta=0.7, e=0.01, precision="mixed")
    got = engine.transform_noalloc(s, **kw)
    ref = engine.transform(s, **kw)
    assert got.tobytes() == ref.tobytes()
    over = P.gen_plummer_virial(MAX_BODIES + 1, seed=9, mass=1.0, a=0.1)
    with pytest.raises(P.PgError) as ei:
        engine.transform_noalloc(over, **kw)
    assert ei.value.code == 2  # PG_ERR_ARG
    # the entry point still works after the rejection, and matches the oracle
    s = P.cube_and_stars(3000, seed=5)
    a = engine.transform_noalloc(s, kind="astro2", theta=1.5, e=0.5, precision="fp64")
    oref, _, fault = O.bh_transform(s, 1.5, 0.5)
    assert fault == 0
    a = a.view(np.float64).reshape(-1, 3)
    r = np.linalg.norm(a - oref, axis=1) / np.linalg.norm(oref, axis=1)
    assert np.quantile(r, 0.99) <= 1e-11 and r.max() <= 1e-8


def test_noalloc_reports_nonfinite_state(engine):
    s = P.cube_and_stars(2000, seed=6)
    s["z"][11] = np.inf
    with pytest.raises(P.PgError) as ei:
        engine.transform_noalloc(s, kind="astro2", theta=1.0, e=0.5)
    assert ei.value.code == 7  # PG_ERR_NONFINITE, as pg_transform_aos
    good = P.cube_and_stars(2000, seed=6)
    ref = engine.transform(good, kind="astro2", theta=1.0, e=0.5)
    assert engine.transform_noalloc(good, kind="astro2", theta=1.0, e=0.5).tobytes() == ref.tobytes()
