"""Pins of the oracle's insertion tree and walk.

1. The reference's own 30 tree tests (astro/src/octree.rs:227-518 and
   astro/src/quadtree.rs:206-491), restated one for one on the oracle tree,
   for dim=3 (octree) and dim=2 (quadtree).
2. Brute force on small inputs: the tree topology equals an independent
   construction from each body's octant path (a cell exists iff its parent
   holds >= 2 bodies), and the interaction list equals the recursive
   definition of the opening rule evaluated over the serialised tree.
3. Invariants: insertion-order independence, mass conservation, centres of
   mass as mass-weighted means.
"""
import numpy as np
import pytest

from oracle import oracle as O

DIMS = [3, 2]


def tree(dim, extent, centre=(0.0, 0.0, 0.0)):
    return O.Tree(dim=dim, centre=centre, extent=extent)


def push_xyz(t, x, y, z, m=1.0, body=-1):
    t.push(O.entity_new(x, y, z, m), body)


def leaves(t, loc, theta):
    out, _ = t.get_leaves_with_resolution(loc, theta)
    return out


# ------------------------------------------------ the reference's tests
@pytest.mark.parametrize("dim", DIMS)
def test_empty_tree(dim):
    t = tree(dim, 1.0)
    assert len(leaves(t, [0, 0, 0], 0.5)) == 0


@pytest.mark.parametrize("dim", DIMS)
def test_single_entity(dim):
    t = tree(dim, 1.0)
    push_xyz(t, 0, 0, 0)
    assert len(leaves(t, [0, 0, 0], 0.5)) == 1


@pytest.mark.parametrize("dim", DIMS)
def test_entities_at_origin(dim):
    t = tree(dim, 2.0)
    for _ in range(10):
        push_xyz(t, 0, 0, 0)
    assert len(leaves(t, [0, 0, 0], 0.5)) == 1
    assert t.merges == 9


@pytest.mark.parametrize("dim", DIMS)
def test_entities_in_different_octants(dim):
    t = tree(dim, 2.0)
    pos = [[0.5, 0.5, 0.5], [-0.5, 0.5, 0.5], [0.5, -0.5, 0.5], [-0.5, -0.5, 0.5]]
    if dim == 3:
        pos += [[0.5, 0.5, -0.5], [-0.5, 0.5, -0.5], [0.5, -0.5, -0.5], [-0.5, -0.5, -0.5]]
    for p in pos:
        push_xyz(t, *p)
    assert len(leaves(t, [0, 0, 0], 0.5)) == (8 if dim == 3 else 4)


@pytest.mark.parametrize("dim", DIMS)
def test_bh_factor_filtering(dim):
    t = tree(dim, 10.0)
    for x in (0.1, 5.0, 8.0):
        push_xyz(t, x, 0, 0)
    assert len(leaves(t, [0, 0, 0], -0.1)) == 3
    assert len(leaves(t, [0, 0, 0], 1.0)) > 0


@pytest.mark.parametrize("dim", DIMS)
def test_boundary_conditions(dim):
    ext = 1.0
    t = tree(dim, ext)
    pts = [(ext, 0, 0), (-ext, 0, 0), (0, ext, 0), (0, -ext, 0)]
    if dim == 3:
        pts += [(0, 0, ext), (0, 0, -ext)]
    for p in pts:
        push_xyz(t, *p)
    assert len(leaves(t, [0, 0, 0], -0.1)) == len(pts)


@pytest.mark.parametrize("dim", DIMS)
def test_query_from_different_locations(dim):
    t = tree(dim, 10.0)
    for i in range(20):
        push_xyz(t, i * 0.1, 0, 0)
    assert len(leaves(t, [0, 0, 0], 0.5)) > 0
    assert len(leaves(t, [100.0, 0, 0], 0.5)) > 0


@pytest.mark.parametrize("dim", DIMS)
def test_dense_cluster(dim):
    t = tree(dim, 1.0)
    r = O.Rng(42)
    for _ in range(1000):
        x = (r.random_f64() - 0.5) * 0.2
        y = (r.random_f64() - 0.5) * 0.2
        z = (r.random_f64() - 0.5) * 0.2
        push_xyz(t, x, y, z)
    assert len(leaves(t, [0, 0, 0], -0.1)) == 1000


@pytest.mark.parametrize("dim", DIMS)
def test_sparse_distribution(dim):
    t = tree(dim, 100.0)
    r = O.Rng(123)
    for _ in range(100):
        x = (r.random_f64() - 0.5) * 200.0
        y = (r.random_f64() - 0.5) * 200.0
        z = (r.random_f64() - 0.5) * 200.0
        push_xyz(t, x, y, z)
    assert len(leaves(t, [0, 0, 0], -0.1)) == 100


@pytest.mark.parametrize("dim", DIMS)
def test_resolution_threshold(dim):
    t = tree(dim, 10.0)
    for x in (1.0, 5.0, 9.0):
        push_xyz(t, x, 0, 0)
    loose = len(leaves(t, [0, 0, 0], 0.1))
    medium = len(leaves(t, [0, 0, 0], 0.5))
    strict = len(leaves(t, [0, 0, 0], 2.0))
    assert loose >= medium >= strict


@pytest.mark.parametrize("dim", DIMS)
def test_lots(dim):
    n = 100_000
    r = O.Rng(0)
    t = O.Tree(dim=dim, extent=1.0, reserve=2 * n)
    for _ in range(n):
        t.push(r.entity_random())
    assert len(leaves(t, [0, 0, 0], -0.1)) == n


@pytest.mark.parametrize("dim", DIMS)
def test_extreme_coordinates(dim):
    t = tree(dim, 1000.0)
    push_xyz(t, 999.0, 999.0, 999.0)
    push_xyz(t, -999.0, -999.0, -999.0)
    assert len(leaves(t, [0, 0, 0], -0.1)) == 2


@pytest.mark.parametrize("dim", DIMS)
def test_zero_extent(dim):
    t = tree(dim, 0.001)
    push_xyz(t, 0, 0, 0)
    assert len(leaves(t, [0, 0, 0], -0.1)) == 1


@pytest.mark.parametrize("dim", DIMS)
def test_query_outside_bounds(dim):
    t = tree(dim, 1.0)
    push_xyz(t, 0, 0, 0)
    assert len(leaves(t, [1000.0, 1000.0, 1000.0], 0.5)) > 0


@pytest.mark.parametrize("dim", DIMS)
def test_reproducibility(dim):
    r1, r2 = O.Rng(999), O.Rng(999)
    t1, t2 = tree(dim, 1.0), tree(dim, 1.0)
    for _ in range(1000):
        t1.push(r1.entity_random())
        t2.push(r2.entity_random())
    assert len(leaves(t1, [0, 0, 0], -0.1)) == len(leaves(t2, [0, 0, 0], -0.1))


# ----------------------------------------------- brute force: topology
def octant_path(p, extent, dim, depth):
    """The reference's descent (octree.rs:160-215), in plain Python floats."""
    c = [0.0, 0.0, 0.0]
    h = extent
    path = []
    for _ in range(depth):
        o = int(p[0] > c[0]) | (int(p[1] > c[1]) << 1)
        if dim == 3:
            o |= int(p[2] > c[2]) << 2
        h2 = h / 2.0
        c[0] = c[0] + h2 if o & 1 else c[0] - h2
        c[1] = c[1] + h2 if o & 2 else c[1] - h2
        if dim == 3:
            c[2] = c[2] + h2 if o & 4 else c[2] - h2
        h = h2
        path.append(o)
    return tuple(path)


def brute_topology(state, dim, depth=40):
    """Preorder (level, octant, leaf body or -1) of the tree whose cells are the
    octant-path prefixes of parents holding >= 2 bodies."""
    ext = O.extent(state)
    paths = [octant_path((b["x"], b["y"], b["z"]), ext, dim, depth) for b in state]
    out = []

    def visit(prefix, bodies, level, oct_):
        if len(bodies) == 1:
            out.append((level, oct_, bodies[0]))
            return
        out.append((level, oct_, -1))
        for o in range(8 if dim == 3 else 4):
            sub = [b for b in bodies if paths[b][level] == o]
            if sub:
                visit(prefix + (o,), sub, level + 1, o)

    visit((), list(range(len(state))), 0, -1)
    return out


def random_state(n, seed, scale=1.0, offset=(0.0, 0.0, 0.5)):
    r = np.random.default_rng(seed)
    s = np.zeros(n, O.ENTITY)
    s["x"] = (r.random(n) - 0.5) * 2 * scale + offset[0]
    s["y"] = (r.random(n) - 0.5) * 2 * scale + offset[1]
    s["z"] = (r.random(n) - 0.5) * scale + offset[2]
    s["mass"] = r.random(n) + 0.1
    return s


@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("seed", range(12))
def test_topology_equals_prefix_construction(dim, seed):
    n = [2, 3, 5, 9, 17, 33, 64, 100, 128, 200, 250, 300][seed]
    s = random_state(n, seed)
    if seed % 3 == 0:  # points exactly on cell centres go to the low octant
        s["x"][: n // 2] = 0.0
        s["y"][n // 2: n // 2 + n // 3] = 0.5 * O.extent(s)
    t = O.Tree.from_state(s, dim)
    assert t.fault == 0 and t.merges == 0
    rec, _ = t.serialize()
    got = [tuple(int(v) for v in r[:3]) for r in rec]
    assert got == brute_topology(s, dim)


# -------------------------------------------------- brute force: walk
def recursive_list(rec, fl, loc, theta, dim):
    """Interaction list by the recursive form of the opening rule: a node
    contributes itself if accepted or a leaf, else its children's lists."""
    # rebuild child lists from the preorder serialisation
    n = len(rec)
    children = [[] for _ in range(n)]
    stack = []
    for k in range(n):
        level = rec[k][0]
        while stack and rec[stack[-1]][0] >= level:
            stack.pop()
        if stack:
            children[stack[-1]].append(k)
        stack.append(k)

    def go(k):
        c = fl[k][4:7]
        r = np.sqrt((loc[0] - c[0]) ** 2 + (loc[1] - c[1]) ** 2 + (loc[2] - c[2]) ** 2)
        with np.errstate(divide="ignore"):
            accepted = fl[k][7] / r < theta
        if accepted or not children[k]:
            return [int(rec[k][3])]
        out = []
        for ch in children[k]:
            out += go(ch)
        return out

    return go(0) if n else []


@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("theta", [-1.0, 0.0, 0.3, 0.7, 1.0, 1.5, 3.0])
def test_walk_equals_recursive_rule(dim, theta):
    s = random_state(150, 7)
    t = O.Tree.from_state(s, dim)
    rec, fl = t.serialize()
    for i in range(0, 150, 7):
        loc = (s["x"][i], s["y"][i], s["z"][i])
        _, ids = t.get_leaves_with_resolution(loc, theta)
        assert sorted(ids.tolist()) == sorted(recursive_list(rec, fl, loc, theta, dim))
    # theta <= 0: every leaf, i.e. the direct sum's source set
    if theta <= 0:
        _, ids = t.get_leaves_with_resolution((0.1, 0.2, 0.3), theta)
        assert len(ids) == 150


# ---------------------------------------------------------- invariants
@pytest.mark.parametrize("dim", DIMS)
def test_insertion_order_independence(dim):
    s = random_state(400, 3)
    rec1, fl1 = O.Tree.from_state(s, dim).serialize()
    perm = np.random.default_rng(1).permutation(len(s))
    s2 = s[perm]
    rec2, fl2 = O.Tree.from_state(s2, dim).serialize()
    # same shape; leaf bodies map through the permutation
    assert np.array_equal(rec1[:, :2], rec2[:, :2])
    leaf1 = rec1[:, 2]
    leaf2 = np.where(rec2[:, 2] >= 0, perm[np.maximum(rec2[:, 2], 0)], rec2[:, 2])
    assert np.array_equal(leaf1, leaf2)
    np.testing.assert_allclose(fl1, fl2, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dim", DIMS)
def test_mass_and_com_conservation(dim):
    s = random_state(300, 11)
    rec, fl = O.Tree.from_state(s, dim).serialize()
    n = len(rec)
    stack = []
    kids = [[] for _ in range(n)]
    for k in range(n):
        while stack and rec[stack[-1]][0] >= rec[k][0]:
            stack.pop()
        if stack:
            kids[stack[-1]].append(k)
        stack.append(k)
    assert sum(1 for r in rec if r[2] >= 0) == len(s)
    for k in range(n):
        if kids[k]:
            assert len(kids[k]) >= 1
            m = sum(fl[c][0] for c in kids[k])
            assert fl[k][0] == pytest.approx(m, rel=1e-13)
            com = sum(fl[c][0] * fl[c][1:4] for c in kids[k]) / m
            np.testing.assert_allclose(fl[k][1:4], com, rtol=1e-12, atol=1e-14)
        else:
            b = rec[k][2]
            assert fl[k][0] == s["mass"][b]
            assert tuple(fl[k][1:4]) == (s["x"][b], s["y"][b], s["z"][b])


def _same_tree(a, b):
    ra, fa = a.serialize()
    rb, fb = b.serialize()
    return (np.array_equal(ra[:, :3], rb[:, :3]) and fa.tobytes() == fb.tobytes()
            and a.merges == b.merges and a.fault == b.fault)


@pytest.mark.parametrize("dim", [3, 2])
def test_parallel_build_equals_insertion_build(dim):
    """The multi-threaded build (used by the all-cores CPU baseline) gives
    the insertion tree bit for bit: topology, centres, masses and centres of
    mass (the same insertion-order folds), merges included; where a node
    above its split depth starts with a coincident pair it falls back to the
    serial build."""
    s = O.cube_and_stars(30000, seed=4)
    assert _same_tree(O.Tree.from_state(s, dim), O.Tree.from_state_mt(s, dim, threads=4))
    # coincident bodies deep in the tree (merged by the reference's rule)
    t = s.copy()
    t["x"][500], t["y"][500], t["z"][500] = t["x"][900], t["y"][900], t["z"][900] + 1e-12
    a, b = O.Tree.from_state(t, dim), O.Tree.from_state_mt(t, dim, threads=4)
    assert a.merges == 1 and _same_tree(a, b)
    # the first two bodies coincide: the root would be a merged leaf
    u = s.copy()
    u["x"][1], u["y"][1], u["z"][1] = u["x"][0], u["y"][0], u["z"][0]
    assert _same_tree(O.Tree.from_state(u, dim), O.Tree.from_state_mt(u, dim, threads=4))
    # a Plummer sphere (deep, uneven cells)
    p = O.gen_plummer_ref(20000, seed=3, mass=1.0, a=0.2)
    assert _same_tree(O.Tree.from_state(p, dim), O.Tree.from_state_mt(p, dim, threads=4))


def test_parallel_baseline_steps_equal_serial():
    s = O.cube_and_stars(20000, seed=2)
    a, b = s.copy(), s.copy()
    O.pipeline_steps_timed(a, 1.5, 0.5, 1e-5, 3)
    O.pipeline_steps_timed_mt(b, 1.5, 0.5, 1e-5, 3, 4)
    for f in ("x", "y", "z", "vx", "vy", "vz", "mass", "id"):
        np.testing.assert_array_equal(a[f], b[f])


def _brute_near_counts(s, theta, near, dim=3):
    """Interaction-list lengths by a direct recursive application of the
    opening rule plus the near field over the serialized tree: a node is
    taken when half_width / r < theta and not (its half-width is the body's
    leaf half-width or a half or a quarter of it, and the body lies within
    half_width + near * leaf width of its centre on every split axis); a
    leaf is always taken."""
    rec, fl = O.Tree.from_state(s, dim).serialize()
    kids = [[] for _ in range(len(rec))]
    stack = []
    for k in range(len(rec)):
        while stack and rec[stack[-1], 0] >= rec[k, 0]:
            stack.pop()
        if stack:
            kids[stack[-1]].append(k)
        stack.append(k)
    leaf_of = {int(rec[k, 2]): k for k in range(len(rec)) if rec[k, 2] >= 0}
    out = []
    for i in range(len(s)):
        loc = np.array([s["x"][i], s["y"][i], s["z"][i]])
        lx = fl[leaf_of[i], 7]
        margin = near * (2.0 * lx)
        cnt, todo = 0, [0]
        while todo:
            k = todo.pop()
            c, hw = fl[k, 4:7], fl[k, 7]
            d = loc - c
            r = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
            take = hw / r < theta if r > 0 else False
            if take and 0.25 * lx <= hw <= lx:
                reach = hw + margin
                axes = 3 if dim == 3 else 2
                take = not all(abs(d[a]) < reach for a in range(axes))
            if take or not kids[k]:
                cnt += 1
            else:
                todo.extend(kids[k])
        out.append(cnt)
    return np.array(out)


def test_near_field_rule():
    """The build's near field in the oracle's walk: 0 leaf widths is the
    reference's walk exactly; widening it only opens cells (lists grow); the
    counts equal a brute-force application of the rule over the serialized
    tree (the own leaf's scale only, so a body alone in a large leaf does not
    take the finely split rest of the tree directly)."""
    s = O.cube_and_stars(3000, seed=2)
    a0, c0, _ = O.bh_transform(s, 0.5, 0.5, counts=True)
    a1, c1, _ = O.bh_transform_near(s, 0.5, 0.5, 0.0, counts=True)
    np.testing.assert_array_equal(a0, a1)
    np.testing.assert_array_equal(c0, c1)
    _, c2, _ = O.bh_transform_near(s, 0.5, 0.5, 2.0, counts=True)
    _, c3, _ = O.bh_transform_near(s, 0.5, 0.5, 6.0, counts=True)
    assert np.all(c2 >= c0) and np.all(c3 >= c2) and c2.sum() > c0.sum()
    small = O.concat([O.gen_plummer_ref(400, seed=4, mass=1.0, a=0.2),
                      O.gen_star(x=0.9, y=0.8, z=-0.7, mass=1.0)])  # one body far out
    for dim in (3, 2):
        _, cs, _ = O.bh_transform_near(small, 0.5, 0.5, 2.0, dim=dim, counts=True)
        np.testing.assert_array_equal(cs, _brute_near_counts(small, 0.5, 2.0, dim))
        assert cs[-1] < len(small) // 4  # the far body's list stays short
    # rows agree with the full pass
    rows = np.array([0, 17, 2999, 3001])
    ar, cr, _ = O.bh_transform_near(s, 0.5, 0.5, 2.0, rows=rows, counts=True)
    np.testing.assert_array_equal(cr, c2[rows])
