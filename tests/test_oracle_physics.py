"""Pins of the oracle's force law, transforms, integrator and energy.

Closed forms: the two-body force (astro/src/lib.rs:84-113), antisymmetry,
a symmetric shell around a probe, the Stormer integrator under a constant
acceleration (integrators/src/verlet.rs:24-82), the force-consistent
two-body potential. Reductions: theta <= 0 Barnes-Hut equals the direct sum
(transformers.rs:123-160 vs :219-244). Reference edge cases: fixed bodies
get no acceleration (transformers.rs:139-141), zero-mass bodies get NaN.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O


def ent(x, y, z, m, vx=0.0, vy=0.0, vz=0.0, fixed=False):
    e = np.zeros(1, O.ENTITY)
    e["x"], e["y"], e["z"], e["mass"] = x, y, z, m
    e["vx"], e["vy"], e["vz"] = vx, vy, vz
    e["fixed"] = fixed
    return e


def test_two_body_force_closed_form():
    a, b = ent(0, 0, 0, 2.0), ent(3.0, 4.0, 0.0, 5.0)
    f = O.gravity(a, b, 0.5)
    mag = 2.0 * 5.0 / (25.0 + 0.5)
    np.testing.assert_allclose(f, [0.6 * mag, 0.8 * mag, 0.0], rtol=1e-15)
    # the direction is not softened: |F| is m_a m_b / (r^2 + e) for any e
    for e in (0.0, 1e-3, 1.0, 10.0):
        f = O.gravity(a, b, e)
        assert np.linalg.norm(f) == pytest.approx(10.0 / (25.0 + e), rel=1e-15)


def test_antisymmetry():
    """F_ab = -F_ba up to the order of the two mass products (a few ulp); with
    equal masses the arithmetic is mirrored exactly."""
    r = np.random.default_rng(0)
    for _ in range(50):
        p, q = r.normal(size=3), r.normal(size=3)
        a, b = ent(*p, r.random() + 0.1), ent(*q, r.random() + 0.1)
        np.testing.assert_allclose(O.gravity(a, b, 0.3), -O.gravity(b, a, 0.3), rtol=1e-15)
        a, b = ent(*p, 0.7), ent(*q, 0.7)
        assert np.array_equal(O.gravity(a, b, 0.3), -O.gravity(b, a, 0.3))


@pytest.mark.parametrize("kind", ["simple_astro", "astro2"])
def test_symmetric_shell_gives_zero(kind):
    pts = [(sx, sy, sz) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]
    s = O.concat([ent(*p, 1.0) for p in pts] + [ent(0.0, 0.0, 0.0, 1.0)])
    if kind == "simple_astro":
        acc = O.simple_transform(s, 0.5)
    else:
        acc, _, _ = O.bh_transform(s, 0.0, 0.5)
    assert np.abs(acc[8]).max() < 1e-15


@pytest.mark.parametrize("dim", [3, 2])
def test_theta_nonpositive_equals_direct_sum(dim):
    s = O.cube_and_stars(2000, seed=3)
    d = O.simple_transform(s, 0.5)
    for theta in (0.0, -1.0):
        a, counts, fault = O.bh_transform(s, theta, 0.5, dim=dim, counts=True)
        assert fault == 0
        assert np.all(counts == len(s))
        rel = np.abs(a - d).max() / np.abs(d).max()
        assert rel < 1e-13


def test_fixed_and_massless_bodies():
    s = O.concat([ent(0, 0, 0, 1.0, fixed=True), ent(1, 0, 0, 1.0), ent(0, 1, 0, 0.0)])
    for acc in (O.simple_transform(s, 0.1), O.bh_transform(s, 0.5, 0.1)[0]):
        assert np.all(acc[0] == 0.0)
        assert np.all(np.isfinite(acc[1]))
        assert np.all(np.isnan(acc[2]))


def test_stormer_constant_acceleration():
    """x(t) = x0 + v0 t + a t^2 / 2 is reproduced (Taylor first step, then the
    second difference of a quadratic is exactly a dt^2)."""
    n = 4
    s = np.zeros(n, O.ENTITY)
    s["x"] = [0.0, 1.0, -2.0, 0.5]
    s["vx"] = [1.0, -0.5, 0.25, 0.0]
    a = np.zeros((n, 3))
    a[:, 0] = [2.0, -1.0, 0.5, 4.0]
    dt = 1e-3
    v = O.Verlet()
    state = s.copy()
    new = np.zeros(n, O.ENTITY)
    for k in range(1, 201):
        v.integrate(state, new, a, dt)
        state = new.copy()
        t = k * dt
        expect = s["x"] + s["vx"] * t + 0.5 * a[:, 0] * t * t
        np.testing.assert_allclose(state["x"], expect, rtol=1e-10, atol=1e-13)
    # first step is exactly the Taylor step
    v2 = O.Verlet()
    out = np.zeros(n, O.ENTITY)
    v2.integrate(s, out, a, dt)
    assert np.array_equal(out["x"], s["x"] + s["vx"] * dt + 0.5 * a[:, 0] * (dt * dt))
    assert np.array_equal(out["vx"], s["vx"] + a[:, 0] * dt)


def test_stormer_velocity_is_backward_difference():
    s = O.cube_and_stars(50, seed=2)
    v = O.Verlet()
    st = s.copy()
    new = np.zeros(len(s), O.ENTITY)
    acc = np.random.default_rng(0).normal(size=(len(s), 3))
    v.integrate(st, new, acc, 1e-4)
    st2 = new.copy()
    v.integrate(st2, new, acc, 1e-4)
    assert np.array_equal(new["vx"], (new["x"] - st2["x"]) / 1e-4)
    # non-position fields are carried over unchanged
    for f in ("radius", "mass", "id", "fixed"):
        assert np.array_equal(new[f], s[f])


def test_two_body_energy_closed_form():
    e = 0.5
    s = O.concat([ent(0, 0, 0, 2.0, vx=1.0), ent(3, 4, 0, 5.0, vy=-2.0)])
    K, V, E, px, py, pz = O.energy(s, e)
    assert K == pytest.approx(0.5 * 2 * 1 + 0.5 * 5 * 4, rel=1e-15)
    assert V == pytest.approx(-(10.0 / math.sqrt(e)) * math.atan(math.sqrt(e) / 5.0), rel=1e-14)
    assert (px, py, pz) == pytest.approx((2.0, -10.0, 0.0))
    # dU/dr equals the force magnitude
    h = 1e-6
    s2 = s.copy()
    s2["x"][1] += 0.6 * h
    s2["y"][1] += 0.8 * h
    dV = (O.energy(s2, e)[1] - V) / h
    assert dV == pytest.approx(10.0 / (25.0 + e), rel=1e-5)


def test_pipeline_matches_manual_steps():
    s = O.cube_and_stars(300, seed=4)
    a = s.copy()
    O.pipeline_steps(a, "astro2", 1.5, 0.5, 1e-5, 5)
    b = s.copy()
    v = O.Verlet()
    new = np.zeros(len(s), O.ENTITY)
    for _ in range(5):
        acc, _, _ = O.bh_transform(b, 1.5, 0.5)
        v.integrate(b, new, acc, 1e-5)
        b = new.copy()
    assert np.array_equal(a, b)


# ------------------------------------------------------------ euler / rk4
def _two_body_orbit():
    """Two unit masses 1 apart, unsoftened (e = 0): each accelerates by 1
    toward the other, so a speed of sqrt(0.5) about their centre keeps the
    circular orbit of period pi / sqrt(0.5)."""
    v = np.sqrt(0.5)
    return O.concat([O.gen_star(x=0.5, vy=v, mass=1.0), O.gen_star(x=-0.5, vy=-v, mass=1.0)])


def _orbit_error(integrator, steps):
    s = _two_body_orbit()
    s0 = s.copy()
    dt = np.pi / np.sqrt(0.5) / steps
    assert O.integrate_steps(s, integrator, "simple_astro", 1.0, 0.0, dt, steps) == 0
    return max(abs(s["x"] - s0["x"]).max(), abs(s["y"] - s0["y"]).max())


def test_rk4_is_fourth_order_and_euler_first_order():
    """Global error after one period on the closed orbit: halving dt divides
    rk4's by ~16 and euler's by ~2. A wrong stage weight or a stage taken at
    the wrong state lowers rk4's order and fails the ratio."""
    r_rk4 = _orbit_error("rk4", 200) / _orbit_error("rk4", 400)
    r_eul = _orbit_error("euler", 4000) / _orbit_error("euler", 8000)
    assert 14.0 < r_rk4 < 18.0, r_rk4
    assert 1.8 < r_eul < 2.2, r_eul
    assert _orbit_error("rk4", 400) < 1e-6


def test_euler_step_is_verlets_taylor_step():
    """euler.rs:26-45 and verlet.rs:24-48 are the same formulas: one step of
    each from the same state agrees bit for bit (x and v)."""
    s = O.cube_and_stars(500, seed=3)
    a, b = s.copy(), s.copy()
    O.integrate_steps(a, "euler", "astro2", 1.0, 0.5, 1e-4, 1)
    O.pipeline_steps(b, "astro2", 1.0, 0.5, 1e-4, 1)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(a[f], b[f])


def test_rk4_fixed_bodies_stay_with_zero_velocity():
    """rk4.rs:168-174: a fixed body keeps its position and gets v = 0, while
    its stage positions (which move with its velocity) still pull the
    others; euler has no such case (the fixed body drifts with v)."""
    s = O.cube_and_stars(300, seed=5)
    s["fixed"][-1] = 1
    s["vx"][-1] = 3.0
    r = s.copy()
    O.integrate_steps(r, "rk4", "astro2", 1.0, 0.5, 1e-3, 2)
    assert r["x"][-1] == s["x"][-1] and r["y"][-1] == s["y"][-1] and r["vx"][-1] == 0.0
    e = s.copy()
    O.integrate_steps(e, "euler", "astro2", 1.0, 0.5, 1e-3, 1)
    assert e["x"][-1] == s["x"][-1] + 3.0 * 1e-3 and e["vx"][-1] == 3.0


OPT = "/opt/rocm/llvm/bin/opt"

# What rustc emits for `s.powf(0.5)` (astro/src/lib.rs:92-93): a call of the
# memory-free llvm.pow.f64 intrinsic with a constant exponent.
POW_HALF_IR = """
declare double @llvm.pow.f64(double, double)
define double @r_norm(double %s) {
  %p = call double @llvm.pow.f64(double %s, double 5.000000e-01)
  ret double %p
}
"""


@pytest.mark.skipif(not os.path.exists(OPT), reason="LLVM opt not installed")
def test_powf_half_is_lowered_to_sqrt(tmp_path):
    """The reference's r_norm is s.powf(0.5) for a sum of squares s, built at
    opt-level 3 (astro/Cargo.toml). The oracle takes IEEE sqrt. That is what
    LLVM's optimizer makes of the intrinsic (fabs(sqrt(s)), +inf for -inf;
    both equal sqrt(s) for s >= 0), and the choice matters: this libm's pow
    is off by one ulp on some inputs, and the oracle's force follows sqrt."""
    src = tmp_path / "pow.ll"
    src.write_text(POW_HALF_IR)
    r = subprocess.run([OPT, "-O3", "-S", str(src), "-o", "-"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    body = r.stdout.split("define double @r_norm", 1)[1].split("}", 1)[0]
    assert "@llvm.sqrt.f64(double %s)" in body and "@llvm.pow" not in body, body

    # a separation whose s has pow(s, 0.5) != sqrt(s) in this libm
    rng = np.random.default_rng(11)
    for _ in range(100000):
        d = rng.uniform(0.1, 2.0, 3)
        s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
        if math.pow(s, 0.5) != math.sqrt(s):
            break
    else:
        pytest.fail("no input where pow and sqrt differ")
    a, b = ent(0.0, 0.0, 0.0, 2.0), ent(d[0], d[1], d[2], 3.0)
    f = O.gravity(a, b, 0.5)

    def force(r_norm):
        re = s + 0.5
        return [(d[k] / r_norm) * 1.0 * 2.0 * 3.0 / re for k in range(3)]

    assert list(f) == force(math.sqrt(s))
    assert list(f) != force(math.pow(s, 0.5))
