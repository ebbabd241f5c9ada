"""ctypes front end of the CPU oracle (oracle/physim_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg. The product (physim_amd/)
never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# physim-core/src/lib.rs:16-29: repr(C) Entity, 80 bytes.
ENTITY = np.dtype(
    {
        "names": ["x", "y", "z", "vx", "vy", "vz", "radius", "mass", "id", "fixed"],
        "formats": ["<f8"] * 8 + ["<u8", "u1"],
        "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 64, 72],
        "itemsize": 80,
    }
)
# physim-core/src/lib.rs:31-37: Acceleration, 24 bytes.
ACCEL = np.dtype({"names": ["x", "y", "z"], "formats": ["<f8"] * 3, "itemsize": 24})

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        D = ctypes.c_double
        U64 = ctypes.c_uint64
        SZ = ctypes.c_size_t
        I = ctypes.c_int
        I64 = ctypes.c_int64
        sig = {
            "or_chacha_block": (None, [P, U64, U64, I, P]),
            "or_rng_sizeof": (SZ, []),
            "or_rng_seed_from_u64": (None, [P, U64]),
            "or_rng_next_u64": (U64, [P]),
            "or_rng_f64_standard": (D, [P]),
            "or_rng_f64_range": (D, [P, D, D]),
            "or_entity_new": (None, [D, D, D, D, P]),
            "or_entity_random": (None, [P, P]),
            "or_gen_cube": (None, [U64, U64, D, D, P, D, U64, P]),
            "or_gen_star": (None, [D, D, D, D, D, D, D, D, U64, I, P]),
            "or_gen_plummer_ref": (None, [U64, U64, D, D, P, P, D, U64, P]),
            "or_gravity": (None, [P, P, D, P]),
            "or_tree_new": (P, [I, P, D, SZ]),
            "or_tree_free": (None, [P]),
            "or_tree_push": (None, [P, P, I64]),
            "or_tree_merges": (U64, [P]),
            "or_tree_fault": (I, [P]),
            "or_tree_node_count": (SZ, [P]),
            "or_tree_walk": (SZ, [P, P, D, P, P, SZ]),
            "or_tree_serialize": (SZ, [P, P, P, SZ]),
            "or_extent": (D, [P, SZ]),
            "or_build_tree": (P, [I, P, SZ]),
            "or_build_tree_mt": (P, [I, P, SZ, I]),
            "or_bh_transform": (I, [I, P, SZ, D, D, P, P]),
            "or_bh_transform_timed": (I, [I, P, SZ, D, D, P]),
            "or_simple_transform": (None, [P, SZ, D, P]),
            "or_simple_transform_rows": (None, [P, SZ, D, P, SZ, P]),
            "or_verlet_new": (P, []),
            "or_verlet_free": (None, [P]),
            "or_verlet_integrate": (None, [P, P, P, SZ, P, D]),
            "or_pipeline_steps": (I, [P, P, SZ, I, D, D, D, I]),
            "or_pipeline_steps_timed": (I, [P, P, SZ, D, D, D, I]),
            "or_energy": (None, [P, SZ, D, P]),
            "or_bh_transform_rows": (I, [I, P, SZ, D, D, P, SZ, P, P]),
            "or_bh_transform_near": (I, [I, P, SZ, D, D, D, P, SZ, P, P]),
            "or_pipeline_steps_timed_mt": (I, [P, P, SZ, D, D, D, I, I]),
            "or_max_threads": (I, []),
            "or_compiler": (ctypes.c_char_p, []),
            "or_euler_step": (I, [P, SZ, I, D, D, D]),
            "or_rk4_step": (I, [P, SZ, I, D, D, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ---------------------------------------------------------------- RNG layer
class Rng:
    """ChaCha8Rng::seed_from_u64 restatement (rand_chacha 0.9.0)."""

    def __init__(self, seed):
        self._buf = ctypes.create_string_buffer(lib().or_rng_sizeof())
        lib().or_rng_seed_from_u64(self._buf, seed)

    def next_u64(self):
        return lib().or_rng_next_u64(self._buf)

    def random_f64(self):
        return lib().or_rng_f64_standard(self._buf)

    def random_range(self, lo, hi):
        return lib().or_rng_f64_range(self._buf, lo, hi)

    def entity_random(self):
        e = np.zeros(1, ENTITY)
        lib().or_entity_random(self._buf, _p(e))
        return e[0]


def chacha_block(key_words, counter, stream=0, rounds=20):
    key = np.asarray(key_words, dtype=np.uint32)
    out = np.zeros(16, np.uint32)
    lib().or_chacha_block(_p(key), counter, stream, rounds, _p(out))
    return out


# ------------------------------------------------------------ initialisers
def entity_new(x, y, z, mass):
    e = np.zeros(1, ENTITY)
    lib().or_entity_new(x, y, z, mass, _p(e))
    return e[0]


def gen_cube(n, seed=0, spin=0.0, size=1.0, centre=(0.0, 0.0, 0.0), mass=1.0, id=0):
    out = np.zeros(n, ENTITY)
    c = np.asarray(centre, dtype=np.float64)
    lib().or_gen_cube(n, seed, spin, size, _p(c), mass, id, _p(out))
    return out


def gen_star(x=0.0, y=0.0, z=0.0, vx=0.0, vy=0.0, vz=0.0, radius=0.1, mass=0.0, id=0,
             fixed=False):
    out = np.zeros(1, ENTITY)
    lib().or_gen_star(x, y, z, vx, vy, vz, radius, mass, id, int(fixed), _p(out))
    return out


def gen_plummer_ref(n, seed=0, mass=1.0, a=1.0, centre=(0.0, 0.0, 0.0), v=(0.0, 0.0, 0.0),
                    spin=0.0, id=0):
    out = np.zeros(n, ENTITY)
    c = np.asarray(centre, dtype=np.float64)
    vv = np.asarray(v, dtype=np.float64)
    lib().or_gen_plummer_ref(n, seed, mass, a, _p(c), _p(vv), spin, id, _p(out))
    return out


def concat(parts):
    """Concatenates Entity arrays keeping the 80-byte repr(C) layout
    (np.concatenate would pack the structured dtype and drop the padding)."""
    out = np.zeros(sum(len(p) for p in parts), ENTITY)
    k = 0
    for p in parts:
        for f in ENTITY.names:
            out[f][k:k + len(p)] = p[f]
        k += len(p)
    return out


def cube_and_stars(n, seed=1, spin=1000.0):
    """The benchmark pipeline's initial state, in CLI add order
    (physim-core/src/help.txt:3): cube n, then the two 1e5-mass stars."""
    return concat(
        [
            gen_cube(n, seed=seed, spin=spin),
            gen_star(x=0.2, y=0.2, z=0.5, radius=0.1, mass=1e5),
            gen_star(x=-0.2, y=-0.2, z=0.5, radius=0.1, mass=1e5),
        ]
    )


# -------------------------------------------------------------------- trees
class Tree:
    """Insertion octree (dim=3) or quadtree (dim=2), astro/src/octree.rs."""

    def __init__(self, dim=3, centre=(0.0, 0.0, 0.0), extent=1.0, reserve=0, handle=None):
        self.dim = dim
        if handle is not None:
            self.h = handle
        else:
            c = np.asarray(centre, dtype=np.float64)
            self.h = lib().or_tree_new(dim, _p(c), extent, reserve)

    @classmethod
    def from_state(cls, state, dim=3):
        state = np.ascontiguousarray(state, ENTITY)
        return cls(dim=dim, handle=lib().or_build_tree(dim, _p(state), len(state)))

    @classmethod
    def from_state_mt(cls, state, dim=3, threads=4):
        """The same tree built on `threads` host threads (subtrees of the
        depth-3 cells in parallel)."""
        state = np.ascontiguousarray(state, ENTITY)
        return cls(dim=dim, handle=lib().or_build_tree_mt(dim, _p(state), len(state), threads))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_tree_free(self.h)
            self.h = None

    def push(self, entity, body=-1):
        e = np.asarray([entity], ENTITY) if not isinstance(entity, np.ndarray) else entity
        e = np.ascontiguousarray(e.reshape(1), ENTITY)
        lib().or_tree_push(self.h, _p(e), body)

    @property
    def merges(self):
        return lib().or_tree_merges(self.h)

    @property
    def fault(self):
        return lib().or_tree_fault(self.h)

    def node_count(self):
        return lib().or_tree_node_count(self.h)

    def get_leaves_with_resolution(self, location, theta):
        """Returns (positions+mass [k,4], arena node ids [k])."""
        loc = np.asarray(location, dtype=np.float64)
        n = lib().or_tree_walk(self.h, _p(loc), theta, None, None, 0)
        out = np.zeros((n, 4))
        ids = np.zeros(n, np.int64)
        lib().or_tree_walk(self.h, _p(loc), theta, _p(out), _p(ids), n)
        return out, ids

    def serialize(self):
        n = lib().or_tree_serialize(self.h, None, None, 0)
        rec = np.zeros((n, 4), np.int64)
        fl = np.zeros((n, 8))
        lib().or_tree_serialize(self.h, _p(rec), _p(fl), n)
        return rec, fl


# --------------------------------------------------------------- transforms
def extent(state):
    state = np.ascontiguousarray(state, ENTITY)
    return lib().or_extent(_p(state), len(state))


def bh_transform(state, theta, e, dim=3, acc=None, counts=False):
    """astro2 (dim=3) / astro (dim=2) transform. Returns (acc [n,3], counts or
    None, fault)."""
    state = np.ascontiguousarray(state, ENTITY)
    n = len(state)
    a = np.zeros(n, ACCEL) if acc is None else acc
    c = np.zeros(n, np.int64) if counts else None
    fault = lib().or_bh_transform(dim, _p(state), n, theta, e, _p(a), _p(c))
    return a.view(np.float64).reshape(n, 3), c, fault


def simple_transform(state, e):
    state = np.ascontiguousarray(state, ENTITY)
    n = len(state)
    a = np.zeros(n, ACCEL)
    lib().or_simple_transform(_p(state), n, e, _p(a))
    return a.view(np.float64).reshape(n, 3)


def simple_transform_rows(state, e, rows):
    state = np.ascontiguousarray(state, ENTITY)
    rows = np.ascontiguousarray(rows, np.int64)
    a = np.zeros(len(rows), ACCEL)
    lib().or_simple_transform_rows(_p(state), len(state), e, _p(rows), len(rows), _p(a))
    return a.view(np.float64).reshape(len(rows), 3)


def bh_transform_rows(state, theta, e, rows, dim=3, counts=False):
    """astro2 / astro accelerations of the bodies `rows` only (the whole tree
    is built). Returns (acc [len(rows), 3], counts or None, fault)."""
    state = np.ascontiguousarray(state, ENTITY)
    rows = np.ascontiguousarray(rows, np.int64)
    a = np.zeros(len(rows), ACCEL)
    c = np.zeros(len(rows), np.int64) if counts else None
    fault = lib().or_bh_transform_rows(dim, _p(state), len(state), theta, e, _p(rows), len(rows),
                                       _p(a), _p(c))
    return a.view(np.float64).reshape(len(rows), 3), c, fault


def bh_transform_near(state, theta, e, near, rows=None, dim=3, counts=False):
    """astro2 / astro with the build's near field (`near` widths of each
    body's own leaf cell); rows None: every body. Returns (acc, counts or
    None, fault)."""
    state = np.ascontiguousarray(state, ENTITY)
    m = len(state) if rows is None else len(rows)
    r = None if rows is None else np.ascontiguousarray(rows, np.int64)
    a = np.zeros(m, ACCEL)
    c = np.zeros(m, np.int64) if counts else None
    fault = lib().or_bh_transform_near(dim, _p(state), len(state), theta, e, near, _p(r), m,
                                       _p(a), _p(c))
    return a.view(np.float64).reshape(m, 3), c, fault


def gravity(a, b, e):
    ea = np.ascontiguousarray(np.asarray(a, ENTITY).reshape(1))
    eb = np.ascontiguousarray(np.asarray(b, ENTITY).reshape(1))
    f = np.zeros(3)
    lib().or_gravity(_p(ea), _p(eb), e, _p(f))
    return f


class Verlet:
    """integrators/src/verlet.rs: Stormer verlet with stored previous state."""

    def __init__(self):
        self.h = lib().or_verlet_new()

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_verlet_free(self.h)
            self.h = None

    def integrate(self, entities, new_state, acc, dt):
        entities = np.ascontiguousarray(entities, ENTITY)
        assert new_state.dtype == ENTITY and new_state.flags["C_CONTIGUOUS"]
        acc = np.ascontiguousarray(np.asarray(acc, np.float64).reshape(-1, 3))
        lib().or_verlet_integrate(self.h, _p(entities), _p(new_state), len(entities), _p(acc), dt)


KIND = {"simple_astro": 0, "astro": 2, "astro2": 3}


def pipeline_steps(state, kind, theta, e, dt, steps, verlet=None):
    """Runs `steps` pipeline iterations in place on `state` (reference order);
    returns (verlet, fault)."""
    assert state.dtype == ENTITY and state.flags["C_CONTIGUOUS"]
    v = verlet or Verlet()
    fault = lib().or_pipeline_steps(v.h, _p(state), len(state), KIND[kind], theta, e, dt, steps)
    return v, fault


def pipeline_steps_timed(state, theta, e, dt, steps, verlet=None):
    assert state.dtype == ENTITY and state.flags["C_CONTIGUOUS"]
    v = verlet or Verlet()
    fault = lib().or_pipeline_steps_timed(v.h, _p(state), len(state), theta, e, dt, steps)
    return v, fault


def pipeline_steps_timed_mt(state, theta, e, dt, steps, threads, verlet=None):
    """The timed astro2 + verlet baseline with the walks and the update split
    over `threads` host threads (serial insertion build)."""
    assert state.dtype == ENTITY and state.flags["C_CONTIGUOUS"]
    v = verlet or Verlet()
    fault = lib().or_pipeline_steps_timed_mt(v.h, _p(state), len(state), theta, e, dt, steps,
                                             threads)
    return v, fault


def max_threads():
    return lib().or_max_threads()


def compiler():
    return lib().or_compiler().decode()


def integrate_steps(state, integrator, kind, theta, e, dt, steps):
    """`steps` pipeline iterations in place with the euler or rk4 integrator
    (integrators/src/euler.rs, rk4.rs) and one gravity transform."""
    assert state.dtype == ENTITY and state.flags["C_CONTIGUOUS"]
    fn = {"euler": lib().or_euler_step, "rk4": lib().or_rk4_step}[integrator]
    fault = 0
    for _ in range(steps):
        fault |= fn(_p(state), len(state), KIND[kind], theta, e, dt)
    return fault


def energy(state, e):
    state = np.ascontiguousarray(state, ENTITY)
    out = np.zeros(6)
    lib().or_energy(_p(state), len(state), e, _p(out))
    return out
