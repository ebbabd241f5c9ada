"""GPU parity of the astro2 / astro / simple_astro transforms and the Stormer
step against the oracle, through libphysim_gpu.so's C-ABI.

Bars (stated per test):
  * tree topology, node geometric centres and half-widths, interaction-list
    lengths: bit-exact;
  * FP64 mode accelerations: median <= 1e-12, p99 <= 1e-11 relative, max
    <= 1e-8. The only
    differences are summation order and the rounding order of each cell's
    centre of mass (the reference folds bodies in insertion order, the device
    sums a cell's bodies in Morton order; the reference's fold over thousands
    of bodies itself accumulates ~1e-13 relative rounding). A heavy body whose own ancestor cell is accepted
    (the reference's opening rule allows it) sees that cell's centre of mass
    ~1e-10 away, so the pair direction is a cancellation and inherits the COM
    rounding difference; with theta = 0 (no pseudo-particles) the bar is 1e-12
    for every body;
  * mixed mode (FP32 pair forces, compensated sums): p99.9 relative error
    <= 1e-4, max <= 1e-3;
  * Stormer update: bit-exact given identical accelerations.
"""
import numpy as np
import pytest

import physim_amd as P
from physim_amd.plugin import PluginLibrary, zero_acc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CFG1 = dict(theta=1.5, e=0.5)


def rel_err(a, ref):
    na = np.linalg.norm(ref, axis=1)
    return np.linalg.norm(a - ref, axis=1) / np.where(na > 0, na, 1.0)


def fp64_ok(r):
    print(f"fp64 rel err median {np.median(r):.2e} p99 {np.quantile(r, 0.99):.2e} "
          f"max {r.max():.2e}")
    return np.median(r) <= 1e-12 and np.quantile(r, 0.99) <= 1e-11 and r.max() <= 1e-8


def small_cube(n, seed=1):
    return P.cube_and_stars(n, seed=seed)


# ------------------------------------------------------------- transforms
@pytest.mark.parametrize("kind,dim", [("astro2", 3), ("astro", 2)])
def test_bh_fp64_matches_oracle(engine, kind, dim):
    s = small_cube(10000)
    ref, counts, fault = O.bh_transform(s, CFG1["theta"], CFG1["e"], dim=dim, counts=True)
    assert fault == 0
    sim = P.Sim(engine, s, kind=kind, precision="fp64", **CFG1)
    sim.set_counting(True)
    a = sim.accel()
    np.testing.assert_array_equal(sim.interaction_counts(), counts)
    r = rel_err(a, ref)
    print(f"{kind} fp64 rel err p99.9 {np.quantile(r, 0.999):.3e} max {r.max():.3e}")
    assert fp64_ok(r)


@pytest.mark.parametrize("kind,dim", [("astro2", 3), ("astro", 2)])
@pytest.mark.parametrize("theta", [0.3, 0.7, 1.0, 1.5])
def test_bh_mixed_matches_oracle(engine, kind, dim, theta):
    s = small_cube(6000, seed=2)
    ref, counts, fault = O.bh_transform(s, theta, 0.5, dim=dim, counts=True)
    sim = P.Sim(engine, s, kind=kind, precision="mixed", theta=theta, e=0.5)
    sim.set_counting(True)
    a = sim.accel()
    np.testing.assert_array_equal(sim.interaction_counts(), counts)
    r = rel_err(a, ref)
    p999 = np.quantile(r, 0.999)
    print(f"{kind} theta {theta} mixed rel err p99.9 {p999:.3e} max {r.max():.3e}")
    assert p999 <= 1e-4 and r.max() <= 1e-3


@pytest.mark.parametrize("kind,dim", [("astro2", 3), ("astro", 2)])
@pytest.mark.parametrize("near", [2.0, 5.0])
def test_near_field_matches_oracle(engine, kind, dim, near):
    """The near field (cells within `near` widths of the body's own leaf
    opened, so close bodies add pair forces): interaction counts bit-exact
    against the oracle's walk with the same rule, accelerations to the FP64
    and mixed bars; stepping with it works alike in a batch."""
    s = small_cube(6000, seed=4)
    ref, counts, fault = O.bh_transform_near(s, 0.5, 0.5, near, dim=dim, counts=True)
    assert fault == 0
    for prec in ("fp64", "mixed"):
        sim = P.Sim(engine, s, kind=kind, precision=prec, theta=0.5, e=0.5)
        sim.set_near_field(near)
        sim.set_counting(True)
        a = sim.accel()
        np.testing.assert_array_equal(sim.interaction_counts(), counts)
        r = rel_err(a, ref)
        if prec == "fp64":
            assert fp64_ok(r)
        else:
            assert np.quantile(r, 0.999) <= 1e-4 and r.max() <= 1e-3
        sim.close()


@pytest.mark.parametrize("kind,dim", [("astro2", 3), ("astro", 2)])
def test_tree_topology_bit_exact(engine, kind, dim):
    for seed, n in ((1, 10000), (5, 3000), (9, 257)):
        s = small_cube(n, seed=seed)
        sim = P.Sim(engine, s, kind=kind, precision="fp64", **CFG1)
        sim.accel()
        rec, fl = sim.tree()
        t = O.Tree.from_state(s, dim)
        orec, ofl = t.serialize()
        assert t.merges == 0 and t.fault == 0
        assert rec.shape == orec.shape
        np.testing.assert_array_equal(rec[:, :3], orec[:, :3])      # level, octant, leaf body
        np.testing.assert_array_equal(fl[:, 4:8], ofl[:, 4:8])      # centres, half-widths: exact
        np.testing.assert_allclose(fl[:, 0], ofl[:, 0], rtol=1e-13)  # mass
        np.testing.assert_allclose(fl[:, 1:4], ofl[:, 1:4], rtol=1e-13, atol=1e-15)
        leaf = rec[:, 2] >= 0
        np.testing.assert_array_equal(fl[leaf, 1:4], ofl[leaf, 1:4])  # leaves hold exact bodies


def close_pairs_state(n=3000, seed=4):
    """A cube with close pairs and a close triple inserted: separations from
    3e-9 to 2e-7 in a unit-size tree, so their keys tie over all 21 levels
    a 63-bit octree key holds (the reference's tree keeps splitting them,
    octree.rs:92-125)."""
    s = small_cube(n, seed=seed)
    extra = []
    offsets = [(3e-9, 0.0, 0.0), (0.0, 2e-8, -1e-8), (1e-7, 1e-7, 1e-7), (-2e-7, 5e-8, 0.0)]
    for k, (dx, dy, dz) in enumerate(offsets):
        b = s[[17 * k + 3]].copy()
        b["x"] += dx
        b["y"] += dy
        b["z"] += dz
        extra.append(b)
    t = s[[101]].copy()  # a triple: two more bodies around body 101
    t2 = s[[101]].copy()
    t["x"] += 4e-9
    t2["y"] -= 6e-9
    return P.concat([s] + extra + [t, t2])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pairs_deeper_than_one_key_word(engine, precision):
    """Bodies that share every level of the first key word: the tree goes on
    to the levels the reference reaches, topology and centres bit-exact,
    accelerations as for any other state."""
    s = close_pairs_state()
    sim = P.Sim(engine, s, kind="astro2", precision=precision, **CFG1)
    a = sim.accel()
    rec, fl = sim.tree()
    st = sim.stats()
    t = O.Tree.from_state(s, 3)
    orec, ofl = t.serialize()
    assert t.merges == 0 and t.fault == 0
    assert st["max_level"] > 21 or int(rec[:, 0].max()) > 21
    np.testing.assert_array_equal(rec[:, :3], orec[:, :3])
    np.testing.assert_array_equal(fl[:, 4:8], ofl[:, 4:8])
    ref, _, _ = O.bh_transform(s, CFG1["theta"], CFG1["e"])
    r = rel_err(a, ref)
    if precision == "fp64":
        assert fp64_ok(r)
    else:
        assert np.quantile(r, 0.999) <= 1e-4
    # and it steps like any other state
    sim.step(3)
    assert np.all(np.isfinite(sim.download()["x"]))


def deep_pairs_state(extent, n=1500, seed=9):
    """A small cube with close pairs (2e-9 to 4e-8 apart along x) and a light
    body at (extent, 0, 0) that sets the root extent: a pair d apart shares
    about log2(2 extent / d) levels of the reference's descent."""
    s = small_cube(n, seed=seed)
    extra = []
    for k, dx in enumerate((2e-9, 7e-9, 4e-8)):
        b = s[[31 * k + 5]].copy()
        b["x"] += dx
        extra.append(b)
    far = s[[1]].copy()
    far["x"], far["y"], far["z"], far["mass"] = extent, 0.0, 0.0, 1e-6
    return P.concat([s] + extra + [far])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pairs_deeper_than_two_key_words(engine, precision):
    """Pairs that share about 47 levels (a root extent of 2e5): past what a
    first key word and one more word of the same length hold, within the 64
    levels the reference reaches (octree.rs:60-62). Topology and centres
    bit-exact against the oracle's insertion tree, accelerations as for any
    other state."""
    s = deep_pairs_state(2e5)
    t = O.Tree.from_state(s, 3)
    assert t.merges == 0 and t.fault == 0
    orec, ofl = t.serialize()
    assert int(orec[:, 0].max()) > 42
    sim = P.Sim(engine, s, kind="astro2", precision=precision, **CFG1)
    a = sim.accel()
    rec, fl = sim.tree()
    np.testing.assert_array_equal(rec[:, :3], orec[:, :3])
    np.testing.assert_array_equal(fl[:, 4:8], ofl[:, 4:8])
    ref, _, fault = O.bh_transform(s, CFG1["theta"], CFG1["e"])
    assert fault == 0
    r = rel_err(a, ref)
    if precision == "fp64":
        assert fp64_ok(r)
    else:
        assert np.quantile(r, 0.999) <= 1e-4


def test_depth_does_not_depend_on_the_batch(engine):
    """Pairs about 41 levels deep (a root extent of 2e3) in a batch of 512
    simulations, whose keys give the segment 9 bits and the tree 18 levels
    per word: the batch steps them as the simulation does alone, bit for bit."""
    big = deep_pairs_state(2e3, n=1000, seed=11)
    assert int(O.Tree.from_state(big, 3).serialize()[0][:, 0].max()) > 36
    kw = dict(theta=1.0, e=0.5, dt=1e-5, precision="fp64")
    sim = P.Sim(engine, big, **kw)
    sim.step(3)
    alone = sim.download()
    sim.close()
    smalls = [small_cube(4, seed=100 + k) for k in range(511)]
    sims = [P.Sim(engine, big, **kw)] + [P.Sim(engine, st, **kw) for st in smalls]
    engine.batch_step(sims, 3)
    assert sims[0].download().tobytes() == alone.tobytes()
    for x in sims:
        x.close()


def test_pairs_past_depth_64_are_an_error(engine):
    """Two bodies 2e-9 apart under a root extent of 1e12 share about 69
    levels: the reference panics past depth 64 (octree.rs:60-62) and the
    oracle reports the fault; the device reports PG_ERR_DEPTH."""
    s = P.concat([P.gen_star(x=0.3, y=0.2, z=0.1, mass=1.0),
                  P.gen_star(x=0.3 + 2e-9, y=0.2, z=0.1, mass=1.0),
                  P.gen_star(x=1e12, y=0.0, z=0.0, mass=1e-6)])
    assert O.Tree.from_state(s, 3).fault != 0
    with pytest.raises(P.PgError) as ei:
        P.Sim(engine, s, theta=1.0, e=0.5, precision="fp64").accel()
    assert ei.value.code == 5


def test_clustered_tile_tree(engine):
    """Close pairs throughout a tile (each pair ~2e-7 apart, ~24 levels deep):
    the tile has far more nodes than bodies, so it is built body by body, and
    its upper cells hold hundreds of bodies (their sums taken by a wave).
    Topology and centres bit-exact, masses and centres of mass to rounding,
    accelerations as for any other state."""
    rng = np.random.default_rng(21)
    s = small_cube(700, seed=21)
    twin = s.copy()
    for ax in ("x", "y", "z"):
        twin[ax] += rng.uniform(-1e-7, 1e-7, len(s)) + np.where(rng.random(len(s)) < 0.5, -2e-8, 2e-8)
    s = P.concat([s, twin])
    sim = P.Sim(engine, s, kind="astro2", precision="fp64", **CFG1)
    a = sim.accel()
    rec, fl = sim.tree()
    t = O.Tree.from_state(s, 3)
    orec, ofl = t.serialize()
    assert t.merges == 0 and t.fault == 0
    assert len(rec) > 4 * len(s)  # the first tile alone exceeds the node-parallel table
    np.testing.assert_array_equal(rec[:, :3], orec[:, :3])
    np.testing.assert_array_equal(fl[:, 4:8], ofl[:, 4:8])
    np.testing.assert_allclose(fl[:, 0], ofl[:, 0], rtol=1e-13)
    np.testing.assert_allclose(fl[:, 1:4], ofl[:, 1:4], rtol=1e-13, atol=1e-15)
    leaf = rec[:, 2] >= 0
    np.testing.assert_array_equal(fl[leaf, 1:4], ofl[leaf, 1:4])
    ref, _, _ = O.bh_transform(s, CFG1["theta"], CFG1["e"])
    assert fp64_ok(rel_err(a, ref))


def short_tile_state(npairs=200, sep=5e-4, seed=31):
    """1024 bodies in the all-negative corner (the first node tile, sorted),
    then `npairs` pairs `sep` apart in the all-positive corner: a second tile
    of 2 npairs bodies whose pairs hang on chains of single-child cells, so
    it holds far more internal nodes than a full tile's share."""
    rng = np.random.default_rng(seed)
    a = np.zeros(1024, P.ENTITY)
    for ax in ("x", "y", "z"):
        a[ax] = rng.uniform(-1.0, -0.5, 1024)
    a["mass"] = rng.uniform(0.5, 1.5, 1024)
    c = np.zeros(npairs, P.ENTITY)
    for ax in ("x", "y", "z"):
        c[ax] = rng.uniform(0.5, 1.0, npairs)
    c["mass"] = rng.uniform(0.5, 1.5, npairs)
    t = c.copy()
    for ax in ("x", "y", "z"):
        t[ax] += sep * rng.uniform(0.5, 1.0, npairs) * np.where(rng.random(npairs) < 0.5, -1, 1)
    s = P.concat([a, c, t])
    s["id"] = np.arange(len(s))
    return s


@pytest.mark.parametrize("kind,dim", [("astro2", 3), ("astro", 2)])
def test_short_tile_many_internal_nodes(engine, kind, dim):
    """A node tile cut short by the segment's end (400 bodies) that holds
    more internal nodes than a full tile's node-parallel list (768) while
    its node count stays under a full tile's limit (1792): every node is
    built, topology and centres bit-exact, masses and centres of mass to
    rounding."""
    s = short_tile_state()
    t = O.Tree.from_state(s, dim)
    orec, ofl = t.serialize()
    assert t.merges == 0 and t.fault == 0
    leaves = np.nonzero(orec[:, 2] >= 0)[0]
    nloc = len(orec) - (leaves[1023] + 1)  # nodes owned by the second tile's bodies
    assert nloc - (len(s) - 1024) > 768 and nloc <= 1792
    sim = P.Sim(engine, s, kind=kind, precision="fp64", **CFG1)
    a = sim.accel()
    rec, fl = sim.tree()
    assert rec.shape == orec.shape
    np.testing.assert_array_equal(rec[:, :3], orec[:, :3])
    np.testing.assert_array_equal(fl[:, 4:8], ofl[:, 4:8])
    np.testing.assert_allclose(fl[:, 0], ofl[:, 0], rtol=1e-13)
    np.testing.assert_allclose(fl[:, 1:4], ofl[:, 1:4], rtol=1e-13, atol=1e-15)
    ref, _, _ = O.bh_transform(s, CFG1["theta"], CFG1["e"], dim=dim)
    assert fp64_ok(rel_err(a, ref))


def octant_path_key(p, extent, dim, levels):
    c = [0.0, 0.0, 0.0]
    h = extent
    key = 0
    for _ in range(levels):
        o = int(p[0] > c[0]) | (int(p[1] > c[1]) << 1)
        if dim == 3:
            o |= int(p[2] > c[2]) << 2
        h2 = h / 2.0
        c[0] = c[0] + h2 if o & 1 else c[0] - h2
        c[1] = c[1] + h2 if o & 2 else c[1] - h2
        if dim == 3:
            c[2] = c[2] + h2 if o & 4 else c[2] - h2
        h = h2
        key = (key << dim) | o
    return key


@pytest.mark.parametrize("kind,dim,levels", [("astro2", 3, 21), ("astro", 2, 32)])
def test_morton_order_and_permutation(engine, kind, dim, levels):
    s = small_cube(3000, seed=3)
    sim = P.Sim(engine, s, kind=kind, **CFG1)
    sim.accel()
    keys, perm = sim.morton()
    ext = O.extent(s)
    ref = np.array([octant_path_key((b["x"], b["y"], b["z"]), ext, dim, levels) for b in s],
                   dtype=np.uint64)
    order = np.lexsort((np.arange(len(s)), ref))  # stable: ties by upload index
    np.testing.assert_array_equal(keys, ref[order])
    np.testing.assert_array_equal(perm, order.astype(np.uint32))


@pytest.mark.parametrize("kind,dim,levels", [("astro2", 3, 21), ("astro", 2, 32)])
def test_morton_keys_on_cell_boundaries(engine, kind, dim, levels):
    """Coordinates on cell boundaries of every level, and one ulp either side
    (the device keys bodies clear of every boundary from their scaled
    coordinates, these by the FP64 descent itself): keys equal the descent's."""
    rng = np.random.default_rng(17)
    s = small_cube(3000, seed=17)
    s["x"][0] = s["y"][0] = s["z"][0] = 1.0  # the extent: 1
    for i in range(1, len(s)):
        ax = ("x", "y", "z")[i % dim]
        lev = int(rng.integers(1, levels + 1))
        frac = float(int(rng.integers(0, 2 ** 20)) % (1 << lev)) / float(1 << lev)
        p = 2.0 * frac - 1.0
        p = (p, np.nextafter(p, 2.0), np.nextafter(p, -2.0))[i % 3]
        s[ax][i] = max(p, -1.0)
    assert O.extent(s) == 1.0
    sim = P.Sim(engine, s, kind=kind, **CFG1)
    sim.accel()
    keys, perm = sim.morton()
    ref = np.array([octant_path_key((b["x"], b["y"], b["z"]), 1.0, dim, levels) for b in s],
                   dtype=np.uint64)
    order = np.lexsort((np.arange(len(s)), ref))
    np.testing.assert_array_equal(keys, ref[order])
    np.testing.assert_array_equal(perm, order.astype(np.uint32))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-4)])
def test_simple_astro_matches_oracle(engine, precision, tol):
    s = small_cube(3000, seed=4)
    ref = O.simple_transform(s, 0.5)
    sim = P.Sim(engine, s, kind="simple_astro", precision=precision, e=0.5)
    a = sim.accel()
    r = rel_err(a, ref)
    print(f"simple_astro {precision} p99.9 {np.quantile(r, 0.999):.3e} max {r.max():.3e}")
    assert np.quantile(r, 0.999) <= tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-4)])
def test_simple_astro_close_pairs(engine, precision, tol):
    """Pairs 3e-9 to 2e-7 apart: their separations carry the softened pair
    force's direction, so they must keep their FP64 precision until the
    difference is taken."""
    s = close_pairs_state()
    ref = O.simple_transform(s, 0.5)
    a = P.Sim(engine, s, kind="simple_astro", precision=precision, e=0.5).accel()
    r = rel_err(a, ref)
    print(f"simple_astro close pairs {precision} p99.9 {np.quantile(r, 0.999):.3e} max {r.max():.3e}")
    assert np.quantile(r, 0.999) <= tol and r.max() <= 10 * tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-4)])
def test_simple_astro_extreme_ranges(engine, precision, tol):
    """Direct sum over separations near 1e7 with masses near 1e30, and close
    pairs 1e-7 apart without softening (e = 0): the mixed-precision weight
    1/(r (r^2 + e)) must stay inside FP32's range at both ends (the product
    r^2 (r^2 + e)^2 does not), matching the FP64 oracle."""
    rng = np.random.default_rng(17)
    far = P.gen_cube(200, seed=21)
    for ax in ("x", "y", "z"):
        far[ax] = far[ax] * 1.0e7
    far["mass"] = rng.uniform(0.5e30, 2e30, len(far))
    near = P.gen_cube(200, seed=22)
    twin = near.copy()
    twin["x"] = twin["x"] + 1e-7 * rng.uniform(0.5, 1.0, len(twin))
    twin["y"] = twin["y"] - 1e-7 * rng.uniform(0.5, 1.0, len(twin))
    for s, e in ((far, 0.5), (P.concat([near, twin]), 0.0)):
        ref = O.simple_transform(s, e)
        a = P.Sim(engine, s, kind="simple_astro", precision=precision, e=e).accel()
        assert np.all(np.isfinite(a))
        r = rel_err(a, ref)
        print(f"simple_astro extreme ranges e={e} {precision} p99.9 {np.quantile(r, 0.999):.3e} "
              f"max {r.max():.3e}")
        assert np.quantile(r, 0.999) <= tol and r.max() <= 10 * tol


def test_theta_zero_is_direct_sum(engine):
    s = small_cube(2000, seed=6)
    ref = O.simple_transform(s, 0.5)
    for kind in ("astro2", "astro"):
        a = P.Sim(engine, s, kind=kind, precision="fp64", theta=0.0, e=0.5).accel()
        assert rel_err(a, ref).max() <= 1e-12


def test_cube_bodies_fp64_exact_to_rounding(engine):
    """Without the two heavy stars the self-ancestor cancellation is mild:
    p99 <= 1e-12 and every body within 1e-10 in FP64 mode."""
    s = P.gen_cube(8000, seed=11, spin=1000.0)
    ref, _, _ = O.bh_transform(s, 1.5, 0.5)
    a = P.Sim(engine, s, kind="astro2", precision="fp64", theta=1.5, e=0.5).accel()
    r = rel_err(a, ref)
    assert fp64_ok(r) and r.max() <= 1e-10


# ------------------------------------------------------------ integration
def test_stormer_steps_fp64_match_oracle(engine):
    """cfg1 pipeline (cube 10000 + 2 stars, astro2 theta 1.5 e 0.5, verlet
    dt 1e-5) for 100 steps. FP64 device steps track the oracle to summation
    order; positions agree to 1e-9 relative after 100 steps."""
    s = small_cube(10000)
    ref = s.copy()
    O.pipeline_steps(ref, "astro2", 1.5, 0.5, 1e-5, 100)
    sim = P.Sim(engine, s, kind="astro2", precision="fp64", theta=1.5, e=0.5, dt=1e-5)
    sim.step(100)
    out = sim.download()
    for f in ("radius", "mass", "id", "fixed"):
        np.testing.assert_array_equal(out[f], s[f])
    pos = np.stack([out["x"], out["y"], out["z"]], 1)
    rpos = np.stack([ref["x"], ref["y"], ref["z"]], 1)
    err = np.linalg.norm(pos - rpos, axis=1) / np.linalg.norm(rpos, axis=1)
    print(f"fp64 100-step position rel err max {err.max():.3e}")
    assert err.max() <= 1e-9
    assert sim.stats()["steps_done"] == 100


def test_stormer_steps_mixed_within_tolerance(engine):
    """Same pipeline in mixed precision: positions after 100 steps within
    1e-6 relative (p99.9) of the FP64 oracle."""
    s = small_cube(10000)
    ref = s.copy()
    O.pipeline_steps(ref, "astro2", 1.5, 0.5, 1e-5, 100)
    sim = P.Sim(engine, s, kind="astro2", precision="mixed", theta=1.5, e=0.5, dt=1e-5)
    sim.step(100)
    out = sim.download()
    pos = np.stack([out["x"], out["y"], out["z"]], 1)
    rpos = np.stack([ref["x"], ref["y"], ref["z"]], 1)
    err = np.linalg.norm(pos - rpos, axis=1) / np.linalg.norm(rpos, axis=1)
    print(f"mixed 100-step position rel err p99.9 {np.quantile(err, 0.999):.3e} max {err.max():.3e}")
    assert np.quantile(err, 0.999) <= 1e-6


@pytest.mark.parametrize("kind", ["astro2", "astro", "simple_astro"])
def test_stormer_update_bit_exact(engine, kind):
    """Three free bodies with non-zero accelerations, FP64 mode: each body's
    force is a sum of two pair terms (equal in either order), and theta 0.01
    opens every cell, so the device's accelerations equal the reference
    arithmetic's bit for bit. The Taylor first step and the Stormer steps
    after it (verlet.rs:24-82: 0.5*a*dt^2, a*dt^2, 2x - x_prev, the backward
    difference for v) must then give bit-identical x and v after 1, 2 and 7
    steps; an FMA contraction or a reordered term changes the last bits."""
    s = O.concat([O.gen_star(x=0.31, y=0.12, z=-0.07, vx=0.3, vy=-1.0, vz=2.5, mass=1.7),
                  O.gen_star(x=-0.45, y=0.25, z=0.5, vx=-0.7, vy=0.1, vz=0.0, mass=2.3),
                  O.gen_star(x=0.05, y=-0.61, z=0.2, vx=0.2, vy=0.4, vz=-0.3, mass=0.9)])
    for steps in (1, 2, 7):
        ref = s.copy()
        O.pipeline_steps(ref, kind, 0.01, 0.1, 1e-3, steps)
        sim = P.Sim(engine, s, kind=kind, theta=0.01, e=0.1, dt=1e-3, precision="fp64")
        sim.step(steps)
        out = sim.download()
        sim.close()
        moved = max(abs(out["vx"] - s["vx"]).max(), abs(out["vy"] - s["vy"]).max())
        assert moved > 1e-6  # the accelerations are not zero
        for f in ("x", "y", "z", "vx", "vy", "vz"):
            np.testing.assert_array_equal(out[f], ref[f], err_msg=f"{kind} {steps} steps {f}")


@pytest.mark.parametrize("kind", ["astro2", "astro", "simple_astro"])
def test_graph_and_eager_steps_identical(engine, kind):
    s = small_cube(4000, seed=8)
    a = P.Sim(engine, s, kind=kind, theta=1.0, e=0.5, dt=1e-5)
    a.step(6)
    engine.set_graphs(False)
    try:
        b = P.Sim(engine, s, kind=kind, theta=1.0, e=0.5, dt=1e-5)
        for _ in range(6):
            b.step(1)
    finally:
        engine.set_graphs(True)
    assert a.download().tobytes() == b.download().tobytes()


# ------------------------------------------------------------ key sorting
def clustered_state(n=40000, seed=13):
    """A cube with tight clusters of 20 to 200 bodies (2e-5 across): a
    bucket's keys crowd into a sliver of its key range, so slots of the
    bucket sort overflow (sorted by a wave, or the bucket by the bitonic
    sort)."""
    rng = np.random.default_rng(seed)
    s = small_cube(n, seed=seed)
    extra = []
    for size in (20, 30, 45, 60, 100, 200):
        b = np.repeat(s[[int(rng.integers(n))]], size)
        for ax in ("x", "y", "z"):
            b[ax] = b[ax] + rng.uniform(-1e-5, 1e-5, size)
        extra.append(b)
    return P.concat([s] + extra)


@pytest.mark.parametrize("make", [
    lambda: P.gen_plummer_virial(60000, seed=12, mass=1.0, a=0.1),  # random order: uneven buckets
    lambda: small_cube(100003, seed=3),
    lambda: close_pairs_state(3000, seed=6),                         # equal keys (ties)
    clustered_state,
], ids=["plummer60k", "cube100k", "ties", "clusters"])
def test_sample_sort_equals_radix_sort(engine, make):
    """The sample sort (buckets in LDS, or in HBM past a small bucket cap;
    splitters in every tile's LDS, or, with more buckets than that holds, in
    HBM behind a coarse table) and the LSD radix sort give the same Morton
    order, tree and steps. Small buckets and a small LDS limit put these
    states on the many-bucket path: up to 391 buckets, which a tile of
    random-order keys spans (its two-pass rank by bucket)."""
    s = make()
    out = {}
    default = engine.tuning()
    small = dict(default, bucket_keys=256, lds_buckets=16)
    try:
        for name, tuning, limits in (("radix", default, (0, 4096)),
                                     ("sample", default, (8192, 4096)),
                                     ("sample_hbm", default, (8192, 64)),
                                     ("all_hbm", default, (8192, 0)),
                                     ("wide", small, (8192, 4096)),
                                     ("wide_hbm", small, (8192, 64))):
            engine.set_tuning(**tuning)
            engine.set_sort(*limits)
            sim = P.Sim(engine, s, kind="astro2", theta=0.9, e=0.05, dt=1e-5)
            sim.step(3)
            keys, perm = sim.morton()
            rec, fl = sim.tree()
            out[name] = (sim.download().tobytes(), keys.tobytes(), perm.tobytes(),
                         rec.tobytes(), fl.tobytes())
            sim.close()
    finally:
        engine.set_tuning(**default)
        engine.set_sort(8192, 4096)
    for name in ("sample", "sample_hbm", "all_hbm", "wide", "wide_hbm"):
        for a, b, what in zip(out["radix"], out[name], ("state", "keys", "perm", "tree", "nodes")):
            assert a == b, (name, what)


def test_many_bucket_sort_over_steps(engine):
    """A 300,000-body cube of 1,172 buckets (more than a tile's LDS holds
    splitters for) over 40 steps, graphs on: the first step samples
    splitters from random-order keys, the later ones take the previous
    step's quantiles, and every step's order is the radix sort's."""
    s = small_cube(300000, seed=21)
    kw = dict(kind="astro2", theta=1.0, e=0.05, dt=1e-4)
    default = engine.tuning()
    out = {}
    try:
        for name, tuning in (("radix", dict(default, max_buckets=0)),
                             ("wide", dict(default, bucket_keys=256))):
            engine.set_tuning(**tuning)
            sim = P.Sim(engine, s, **kw)
            sim.step(40)
            keys, perm = sim.morton()
            out[name] = (sim.download().tobytes(), keys.tobytes(), perm.tobytes())
            sim.close()
    finally:
        engine.set_tuning(**default)
    for a, b, what in zip(out["radix"], out["wide"], ("state", "keys", "perm")):
        assert a == b, what


# ---------------------------------------------------------------- batching
def test_batch_equals_alone(engine):
    states = [small_cube(n, seed=k) for k, n in enumerate((5000, 1, 2, 777, 12000))]
    thetas = (1.5, 0.7, 1.0, 0.5, 1.2)
    alone = []
    for st, th in zip(states, thetas):
        sim = P.Sim(engine, st, theta=th, e=0.5, dt=1e-5)
        sim.step(5)
        alone.append(sim.download())
    sims = [P.Sim(engine, st, theta=th, e=0.5, dt=1e-5) for st, th in zip(states, thetas)]
    engine.batch_step(sims, 5)
    for sim, ref in zip(sims, alone):
        assert sim.download().tobytes() == ref.tobytes()


# ------------------------------------------------------------ plugin path
def test_plugin_transform_accumulates(engine):
    pl = PluginLibrary()
    s = small_cube(5000, seed=7)
    el = pl.load("astro2", {"theta": 1.5, "e": 0.5, "precision": "fp64"})
    acc = zero_acc(len(s))
    acc["x"] = 1.0
    el.transform(s, acc)
    ref, _, _ = O.bh_transform(s, 1.5, 0.5)
    got = acc.view(np.float64).reshape(-1, 3).copy()
    got[:, 0] -= 1.0
    assert fp64_ok(rel_err(got, ref))
    el.close()
    for name, kind in (("astro", 2), ("simple_astro", 0)):
        el = pl.load(name, {"theta": 0.9, "e": 0.25, "precision": "fp64"})
        acc = zero_acc(len(s))
        el.transform(s, acc)
        if kind:
            ref, _, _ = O.bh_transform(s, 0.9, 0.25, dim=2)
        else:
            ref = O.simple_transform(s, 0.25)
        assert fp64_ok(rel_err(acc.view(np.float64).reshape(-1, 3), ref))
        el.close()


# -------------------------------------------------------------- edge cases
def test_small_and_special_states(engine):
    one = P.gen_star(x=0.3, y=0.1, z=0.2, mass=2.0)
    a = P.Sim(engine, one, theta=1.0, e=0.5).accel()
    assert np.all(a == 0.0)
    two = P.concat([P.gen_star(x=0.0, mass=2.0), P.gen_star(x=3.0, y=4.0, mass=5.0)])
    ref, _, _ = O.bh_transform(two, 1.0, 0.5)
    for prec in ("fp64", "mixed"):
        a = P.Sim(engine, two, theta=1.0, e=0.5, precision=prec).accel()
        assert rel_err(a, ref).max() <= (1e-15 if prec == "fp64" else 1e-6)
    # fixed bodies get no acceleration; zero-mass bodies get NaN (as the reference)
    s = small_cube(500, seed=2)
    s["fixed"][:10] = 1
    s["mass"][10] = 0.0
    ref, _, _ = O.bh_transform(s, 1.0, 0.5)
    for prec in ("fp64", "mixed"):
        a = P.Sim(engine, s, theta=1.0, e=0.5, precision=prec).accel()
        assert np.all(a[:10] == 0.0)
        assert np.all(np.isnan(a[10])) and np.all(np.isnan(ref[10]))
        ok = np.arange(len(s)) > 10
        r = rel_err(a[ok], ref[ok])
        assert fp64_ok(r) if prec == "fp64" else r.max() <= 1e-3


def test_errors_are_reported(engine):
    s = small_cube(300, seed=3)
    bad = s.copy()
    bad["x"][5] = np.nan
    with pytest.raises(P.PgError) as ei:
        P.Sim(engine, bad, theta=1.0, e=0.5).accel()
    assert ei.value.code == 7
    dup = s.copy()
    dup["x"][7], dup["y"][7], dup["z"][7] = dup["x"][8], dup["y"][8], dup["z"][8]
    with pytest.raises(P.PgError) as ei:
        P.Sim(engine, dup, theta=1.0, e=0.5).accel()
    assert ei.value.code in (5, 6)
    # the engine keeps working after a failed simulation
    a = P.Sim(engine, s, theta=1.0, e=0.5, precision="fp64").accel()
    ref, _, _ = O.bh_transform(s, 1.0, 0.5)
    assert fp64_ok(rel_err(a, ref))
