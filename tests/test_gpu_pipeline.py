"""The verlet element binding and the device-resident pipeline runner,
through the C-ABI.

* tests/harness/pipeline_harness (a C program) runs the pipeline loop of
  physim-core/src/pipeline.rs:143-182 over pg_verlet_* with a GPU astro2
  element -- synths, integrate, transmutes, state = new_state -- including a
  transmute that edits a body (same length: the Stormer history carries on,
  verlet.rs:104-107) and a synth that appends one (a Taylor restart). Its
  final state equals the oracle's run of the same loop: bit for bit where the
  accelerations are (three bodies, theta 0.01), to FP64 summation order
  otherwise.
* pg_runner_* writes csvsink frames (first state, then every print_n-th):
  the text equals the oracle's states printed in Rust's f64 Display, and
  parses back to the device state.
"""
import os
import subprocess
from decimal import Decimal

import numpy as np
import pytest

import physim_amd as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HARNESS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "harness", "pipeline_harness")
FIELDS = ("x", "y", "z", "vx", "vy", "vz", "radius", "mass", "id", "fixed")


def three_bodies():
    return O.concat([O.gen_star(x=0.31, y=0.12, z=-0.07, vx=0.3, vy=-1.0, vz=2.5, mass=1.7),
                     O.gen_star(x=-0.45, y=0.25, z=0.5, vx=-0.7, vy=0.1, vz=0.0, mass=2.3),
                     O.gen_star(x=0.05, y=-0.61, z=0.2, vx=0.2, vy=0.4, vz=-0.3, mass=0.9)])


def run_harness(tmp_path, n_cube, steps, theta, e, dt, precision, edit_at, append_at):
    out = tmp_path / "final.bin"
    args = [HARNESS, str(n_cube), str(steps), repr(theta), repr(e), repr(dt), precision,
            str(edit_at), str(append_at), str(out)]
    print(" ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return np.fromfile(out, dtype=O.ENTITY)


def oracle_loop(state, steps, theta, e, dt, edit_at, append_at):
    """The same loop on the oracle: one Verlet element across iterations."""
    v = O.Verlet()
    for it in range(1, steps + 1):
        if it == append_at:
            state = O.concat([state, O.gen_star(x=0.05, y=-0.3, z=0.25, vx=1.0, mass=3.0)])
        _, fault = O.pipeline_steps(state, "astro2", theta, e, dt, 1, verlet=v)
        assert fault == 0
        if it == edit_at:
            state["x"][7 if len(state) > 7 else 0] += 1e-3
    return state


def test_harness_three_bodies_bit_exact(tmp_path):
    got = run_harness(tmp_path, 0, 7, 0.01, 0.1, 1e-3, "fp64", 3, 0)
    ref = oracle_loop(three_bodies(), 7, 0.01, 0.1, 1e-3, 3, 0)
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f)


def test_harness_cube_with_edit_and_synth(tmp_path):
    got = run_harness(tmp_path, 2000, 10, 1.0, 0.5, 1e-5, "fp64", 4, 7)
    ref = oracle_loop(O.cube_and_stars(2000, seed=1, spin=1000.0), 10, 1.0, 0.5, 1e-5, 4, 7)
    assert len(got) == len(ref) == 2003
    p = np.stack([got[f] for f in ("x", "y", "z")], 1)
    r = np.stack([ref[f] for f in ("x", "y", "z")], 1)
    err = np.linalg.norm(p - r, axis=1) / np.linalg.norm(r, axis=1)
    print(f"harness cube: position rel err max {err.max():.2e} (cube bodies "
          f"{err[:2000].max():.2e})")
    # the two heavy stars accept their own ancestor cells, whose centres of
    # mass (rounded differently in FP64) sit ~1e-10 from them: their
    # accelerations carry up to 1e-8 relative, the cube's 1e-12
    assert err[:2000].max() <= 1e-12 and err.max() <= 1e-10
    for f in ("radius", "mass", "id", "fixed"):
        np.testing.assert_array_equal(got[f], ref[f])


# ------------------------------------------------------------- runner
def rust_display(x):
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "inf" if x > 0 else "-inf"
    return format(Decimal(repr(float(x))).normalize(), "f")


def csv_line(state):
    return "".join(f"{rust_display(a)},{rust_display(b)},{rust_display(c)},"
                   for a, b, c in zip(state["x"], state["y"], state["z"])) + "\n"


@pytest.mark.parametrize("integrator", ["verlet", "rk4"])
def test_runner_csvsink_text_equals_oracle(engine, tmp_path, integrator):
    s = three_bodies()
    kw = dict(kind="astro2", theta=0.01, e=0.1, dt=1e-3, precision="fp64")
    r = P.Runner(engine, s, integrator=integrator, **kw)
    f4, f1 = tmp_path / "every4.csv", tmp_path / "every1.csv"
    r.add_csvsink(f4, print_n=4)
    r.add_csvsink(f1, print_n=1)
    r.run(5)
    r.run(7)   # a second run continues the iteration count
    r.flush()
    r.close()
    ref, lines4, lines1 = s.copy(), [csv_line(s)], [csv_line(s)]
    v = O.Verlet()
    for it in range(1, 13):
        if integrator == "verlet":
            O.pipeline_steps(ref, "astro2", 0.01, 0.1, 1e-3, 1, verlet=v)
        else:
            O.integrate_steps(ref, "rk4", "astro2", 0.01, 0.1, 1e-3, 1)
        lines1.append(csv_line(ref))
        if it % 4 == 0:
            lines4.append(csv_line(ref))
    assert f4.read_text() == "".join(lines4)
    assert f1.read_text() == "".join(lines1)


def test_runner_frames_parse_back_to_device_state(engine, tmp_path):
    s = P.cube_and_stars(1000, seed=3)
    r = P.Runner(engine, s, kind="astro2", theta=1.5, e=0.5, dt=1e-5, precision="mixed")
    path = tmp_path / "frames.csv"
    r.add_csvsink(path, print_n=10)
    r.run(20)
    r.flush()
    last = r.state()
    r.close()
    lines = path.read_text().splitlines()
    assert len(lines) == 3  # the initial state, iterations 10 and 20
    for line, st in ((lines[0], s), (lines[-1], last)):
        vals = np.array([float(t) for t in line.rstrip(",").split(",")]).reshape(-1, 3)
        np.testing.assert_array_equal(vals, np.stack([st["x"], st["y"], st["z"]], 1))
    # the runner's device steps equal a simulation stepped alone
    sim = P.Sim(engine, s, kind="astro2", theta=1.5, e=0.5, dt=1e-5, precision="mixed")
    sim.step(20)
    assert sim.download().tobytes() == last.tobytes()
    sim.close()
