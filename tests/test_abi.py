"""CPU checks of the C-ABI library: it loads, exports every symbol the
headers declare, speaks the physim plugin protocol (discover.rs /
transform.rs) without touching a GPU, and its initial-condition generators
equal the oracle's bit for bit."""
import os
import re

import numpy as np
import pytest

import physim_amd as P
from physim_amd.plugin import PluginLibrary, TransformElementHandler
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("physim_abi.h", "physim_gpu.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"^\s*#.*$", "", src, flags=re.M)
        src = re.sub(r"typedef\s+(struct|enum)[^{;]*\{.*?\}[^;]*;", "", src, flags=re.S)
        src = re.sub(r"typedef[^;]*;", "", src, flags=re.S)
        src = re.sub(r"\b_?[sS]tatic_assert\s*\(.*?\)\s*;", "", src, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{]*)\)\s*;", src):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import ctypes

    lib = ctypes.CDLL(P.library_path())
    names = declared_functions()
    assert len(names) >= 45
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_plugin_protocol():
    pl = PluginLibrary()
    assert pl.abi_info() == "C"
    assert pl.element_names() == ["astro", "astro2", "simple_astro"]
    for name in pl.element_names():
        m = pl.meta(name)
        assert m["kind"] == 1  # ElementKind::Transform
        assert m["name"] == name
        assert m["plugin"] == "physim_gpu"


def test_init_parses_properties_like_the_reference():
    pl = PluginLibrary()
    # discovery instantiates every transform with {} (discover.rs:383-386)
    for name in pl.element_names():
        el = pl.load(name, {})
        props = el.get_property_descriptions()
        assert "e" in props
        assert ("theta" in props) == (name != "simple_astro")
        el.close()
    el = pl.load("astro2", {"theta": 1.5, "e": -0.5, "other": [1, {"x": None}]})
    el.close()
    # invalid JSON and non-objects give NULL (physim-attribute/src/lib.rs:67-70)
    for raw in (b"", b"{", b"[1,2]", b"3", b'{"theta": 1.5,}', b'{"theta": 01}', b"{'a':1}"):
        assert TransformElementHandler.from_raw_json(pl, "astro2", raw) is None, raw
    assert TransformElementHandler.from_raw_json(pl, "astro2", b' {"theta": "x"} ') is not None


def same_entities(a, b):
    """Field-wise bit equality (padding bytes are unspecified)."""
    return len(a) == len(b) and all(
        a[f].tobytes() == b[f].tobytes() for f in P.ENTITY.names)


def test_generators_equal_oracle():
    a = P.gen_cube(5000, seed=1, spin=1000.0)
    b = O.gen_cube(5000, seed=1, spin=1000.0)
    assert same_entities(a, b)
    a = P.gen_cube(777, seed=9, spin=-3.0, size=2.5, centre=(1.0, -2.0, 0.25), mass=7.0, id=4)
    b = O.gen_cube(777, seed=9, spin=-3.0, size=2.5, centre=(1.0, -2.0, 0.25), mass=7.0, id=4)
    assert same_entities(a, b)
    a = P.gen_plummer_ref(3000, seed=3, mass=2.0, a=0.3, centre=(0, 1, 0), v=(0.1, 0, 0), spin=0.5)
    b = O.gen_plummer_ref(3000, seed=3, mass=2.0, a=0.3, centre=(0, 1, 0), v=(0.1, 0, 0), spin=0.5)
    assert same_entities(a, b)
    assert same_entities(P.cube_and_stars(100), O.cube_and_stars(100))


def test_virial_plummer_moments():
    """Plummer (1911) sphere: half-mass radius 1.3048 a; virial ratio 2K/|W| = 1
    (W from the Newtonian pair sum); centre-of-mass frame."""
    n, a = 20000, 0.1
    s = P.gen_plummer_virial(n, seed=0, mass=1.0, a=a)
    r = np.sqrt(s["x"] ** 2 + s["y"] ** 2 + s["z"] ** 2)
    assert np.median(r) == pytest.approx(1.3048 * a, rel=0.03)
    assert abs(s["x"].mean()) < 1e-12 and abs(s["vx"].mean()) < 1e-12
    sub = s[:4000]
    K = 0.5 * np.sum(sub["mass"] * (sub["vx"] ** 2 + sub["vy"] ** 2 + sub["vz"] ** 2)) * (n / 4000)
    p = np.stack([sub["x"], sub["y"], sub["z"]], 1)
    d = np.sqrt(((p[:, None, :] - p[None, :, :]) ** 2).sum(-1))
    iu = np.triu_indices(len(sub), 1)
    W = -np.sum(1.0 / d[iu]) * (1.0 / n) ** 2 * (n * (n - 1) / 2) / len(iu[0])
    assert 2 * K / abs(W) == pytest.approx(1.0, abs=0.06)
    # the radial profile: fraction inside r = a is 1/(2 sqrt 2)
    assert np.mean(r < a) == pytest.approx(2 ** -1.5, abs=0.01)


def test_memory_model():
    assert P.lib().pg_sim_bytes(1_000_000, 3) > P.lib().pg_sim_bytes(1_000_000, 0)
    assert P.lib().pg_sim_bytes(2_000_000, 3) > 1.9 * P.lib().pg_sim_bytes(1_000_000, 3)
