"""ctypes bindings of libphysim_gpu.so (include/physim_gpu.h)."""
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libphysim_gpu.so"

# physim-core/src/lib.rs:16-29: repr(C) Entity, 80 bytes.
ENTITY = np.dtype(
    {
        "names": ["x", "y", "z", "vx", "vy", "vz", "radius", "mass", "id", "fixed"],
        "formats": ["<f8"] * 8 + ["<u8", "u1"],
        "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 64, 72],
        "itemsize": 80,
    }
)
# physim-core/src/lib.rs:31-37
ACCEL = np.dtype({"names": ["x", "y", "z"], "formats": ["<f8"] * 3, "itemsize": 24})

KIND = {"simple_astro": 0, "astro": 2, "astro2": 3}
PRECISION = {"mixed": 0, "fp64": 1}

STATUS = {
    0: "PG_OK",
    1: "PG_ERR_HIP",
    2: "PG_ERR_ARG",
    3: "PG_ERR_NOMEM",
    4: "PG_ERR_BUDGET",
    5: "PG_ERR_DEPTH",
    6: "PG_ERR_COINCIDENT",
    7: "PG_ERR_NONFINITE",
    8: "PG_ERR_STATE",
    9: "PG_ERR_INTERNAL",
}


class PgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("theta", ctypes.c_double),
        ("e", ctypes.c_double),
        ("dt", ctypes.c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("steps_done", ctypes.c_uint64),
        ("nodes", ctypes.c_uint64),
        ("interactions", ctypes.c_uint64),
        ("max_level", ctypes.c_uint32),
        ("key_levels", ctypes.c_uint32),
        ("error", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("max_buckets", ctypes.c_uint32), ("bucket_keys", ctypes.c_uint32),
                ("walk_run_log2", ctypes.c_uint32), ("lds_buckets", ctypes.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("n", ctypes.c_uint64),
        ("steps", ctypes.c_uint32),
        ("status", ctypes.c_uint32),
        ("submit_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("complete_s", ctypes.c_double),
    ]


_lib = None


def library_path():
    return os.path.join(HERE, _LIB_NAME)


def lib():
    """Loads the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m physim_amd.build` "
            "(there is no CPU fallback)"
        )
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    PP = ctypes.POINTER(ctypes.c_void_p)
    I = ctypes.c_int
    U32 = ctypes.c_uint32
    U64 = ctypes.c_uint64
    D = ctypes.c_double
    SZ = ctypes.c_size_t
    sig = {
        "pg_last_error": (ctypes.c_char_p, []),
        "pg_version": (ctypes.c_char_p, []),
        "pg_engine_create": (I, [I, U64, PP]),
        "pg_engine_destroy": (None, [P]),
        "pg_engine_memory": (I, [P, ctypes.POINTER(U64), ctypes.POINTER(U64)]),
        "pg_sim_bytes": (U64, [U64, ctypes.c_int32]),
        "pg_engine_set_graphs": (I, [P, I]),
        "pg_engine_set_profiling": (I, [P, I]),
        "pg_engine_set_sort": (I, [P, U32, U32]),
        "pg_engine_profile_read": (I, [P, ctypes.c_char_p, SZ, P, P, U32, ctypes.POINTER(U32)]),
        "pg_engine_profile_reset": (I, [P]),
        "pg_engine_synchronize": (I, [P]),
        "pg_engine_set_tuning": (I, [P, ctypes.POINTER(Tuning)]),
        "pg_engine_get_tuning": (I, [P, ctypes.POINTER(Tuning)]),
        "pg_engine_autotune": (I, [P, P, U32, U32, ctypes.POINTER(Tuning), P, P, U32,
                                   ctypes.POINTER(U32)]),
        "pg_engine_copy_probe": (I, [P, U64, U32, ctypes.POINTER(D)]),
        "pg_sim_create": (I, [P, U64, ctypes.POINTER(Params), PP]),
        "pg_sim_destroy": (None, [P]),
        "pg_sim_set_params": (I, [P, ctypes.POINTER(Params)]),
        "pg_sim_upload": (I, [P, P, U64]),
        "pg_sim_download": (I, [P, P, U64]),
        "pg_sim_step": (I, [P, U32]),
        "pg_sim_upload_state": (I, [P, P, P, U64]),
        "pg_sim_checkpoint": (I, [P, P, P, U64, ctypes.POINTER(I)]),
        "pg_verlet_create": (I, [PP]),
        "pg_verlet_destroy": (None, [P]),
        "pg_verlet_bind": (I, [P, P]),
        "pg_verlet_integrate": (I, [P, P, P, U64, D]),
        "pg_runner_create": (I, [P, ctypes.POINTER(Params), ctypes.c_int32, P, U64, PP]),
        "pg_runner_add_csvsink": (I, [P, ctypes.c_char_p, U32]),
        "pg_runner_run": (I, [P, U32]),
        "pg_runner_flush": (I, [P]),
        "pg_runner_state": (I, [P, P, U64]),
        "pg_runner_destroy": (None, [P]),
        "pg_format_f64": (SZ, [D, ctypes.c_char_p, SZ]),
        "pg_sim_set_integrator": (I, [P, ctypes.c_int32]),
        "pg_sim_set_near_field": (I, [P, D]),
        "pg_sim_accel": (I, [P, P]),
        "pg_sim_energy": (I, [P, P]),
        "pg_sim_stats": (I, [P, ctypes.POINTER(Stats)]),
        "pg_sim_set_counting": (I, [P, I]),
        "pg_sim_interaction_counts": (I, [P, P]),
        "pg_sim_tree": (I, [P, P, P, U64, ctypes.POINTER(U64)]),
        "pg_sim_morton": (I, [P, P, P, U64]),
        "pg_batch_step": (I, [P, P, U32, U32]),
        "pg_transform_aos": (I, [P, ctypes.POINTER(Params), P, U64, P]),
        "pg_transform_aos_noalloc": (I, [P, ctypes.POINTER(Params), P, U64, P]),
        "pg_server_create": (I, [P, U32, PP]),
        "pg_server_destroy": (None, [P]),
        "pg_server_submit": (I, [P, P, U64, ctypes.POINTER(Params), U32, ctypes.POINTER(U64)]),
        "pg_server_tick": (I, [P, U32, ctypes.POINTER(U32), ctypes.POINTER(U32)]),
        "pg_server_take": (I, [P, U64, P, U64, ctypes.POINTER(RequestInfo)]),
        "pg_server_info": (I, [P, U64, ctypes.POINTER(RequestInfo)]),
        "pg_server_clock": (D, [P]),
        "pg_gen_cube": (I, [U64, U64, D, D, P, D, U64, P]),
        "pg_gen_star": (I, [D, D, D, D, D, D, D, D, U64, I, P]),
        "pg_gen_plummer_ref": (I, [U64, U64, D, D, P, P, D, U64, P]),
        "pg_gen_plummer_virial": (I, [U64, U64, D, D, P, P, U64, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def check(rc):
    if rc != 0:
        raise PgError(rc, lib().pg_last_error().decode())


def params(kind="astro2", theta=1.0, e=1.0, dt=1e-6, precision="mixed"):
    return Params(
        KIND[kind] if isinstance(kind, str) else kind,
        PRECISION[precision] if isinstance(precision, str) else precision,
        theta,
        abs(e),
        dt,
    )


# ------------------------------------------------------------ generators
def gen_cube(n, seed=0, spin=0.0, size=1.0, centre=(0.0, 0.0, 0.0), mass=1.0, id=0):
    out = np.zeros(n, ENTITY)
    c = np.asarray(centre, np.float64)
    check(lib().pg_gen_cube(n, seed, spin, size, _p(c), mass, id, _p(out)))
    return out


def gen_star(x=0.0, y=0.0, z=0.0, vx=0.0, vy=0.0, vz=0.0, radius=0.1, mass=0.0, id=0,
             fixed=False):
    out = np.zeros(1, ENTITY)
    check(lib().pg_gen_star(x, y, z, vx, vy, vz, radius, mass, id, int(fixed), _p(out)))
    return out


def gen_plummer_ref(n, seed=0, mass=1.0, a=1.0, centre=(0.0, 0.0, 0.0), v=(0.0, 0.0, 0.0),
                    spin=0.0, id=0):
    out = np.zeros(n, ENTITY)
    c = np.asarray(centre, np.float64)
    vv = np.asarray(v, np.float64)
    check(lib().pg_gen_plummer_ref(n, seed, mass, a, _p(c), _p(vv), spin, id, _p(out)))
    return out


def gen_plummer_virial(n, seed=0, mass=1.0, a=1.0, centre=(0.0, 0.0, 0.0), v=(0.0, 0.0, 0.0),
                       id=0):
    out = np.zeros(n, ENTITY)
    c = np.asarray(centre, np.float64)
    vv = np.asarray(v, np.float64)
    check(lib().pg_gen_plummer_virial(n, seed, mass, a, _p(c), _p(vv), id, _p(out)))
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


# --------------------------------------------------------------- objects
class Engine:
    def __init__(self, device=0, budget=0):
        h = ctypes.c_void_p()
        check(lib().pg_engine_create(device, budget, ctypes.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()

    def close(self):
        """Destroys the engine; servers and simulations still open on it are
        closed first (they must not outlive it)."""
        if getattr(self, "h", None):
            for child in sorted(self._children, key=lambda c: not isinstance(c, Server)):
                child.close()
            lib().pg_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def memory(self):
        used, budget = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().pg_engine_memory(self.h, ctypes.byref(used), ctypes.byref(budget)))
        return used.value, budget.value

    def set_graphs(self, on):
        check(lib().pg_engine_set_graphs(self.h, int(on)))

    def set_profiling(self, on):
        check(lib().pg_engine_set_profiling(self.h, int(on)))

    def set_sort(self, max_buckets=8192, bucket_cap=4096):
        """Key sort of the tree build: sample sort into at most max_buckets
        buckets (0: the LSD radix sort), buckets above bucket_cap keys
        sorted in HBM instead of LDS. Both orders are the same."""
        check(lib().pg_engine_set_sort(self.h, int(max_buckets), int(bucket_cap)))

    def profile_reset(self):
        check(lib().pg_engine_profile_reset(self.h))

    def profile(self):
        """{kernel: (total_ms, launches)} accumulated while profiling."""
        cap = 4096
        names = ctypes.create_string_buffer(cap)
        ms = np.zeros(64)
        n = np.zeros(64, np.uint64)
        count = ctypes.c_uint32()
        check(lib().pg_engine_profile_read(self.h, names, cap, _p(ms), _p(n), 64,
                                           ctypes.byref(count)))
        keys = names.value.decode().split("\n")[: count.value]
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(keys)}

    def synchronize(self):
        check(lib().pg_engine_synchronize(self.h))

    def transform(self, state, kind="astro2", theta=1.0, e=1.0, precision="mixed", acc=None):
        """acc += accelerations of `state` (host AoS in and out)."""
        state = np.ascontiguousarray(state, ENTITY)
        a = np.zeros(len(state), ACCEL) if acc is None else acc
        p = params(kind, theta, e, 0.0, precision)
        check(lib().pg_transform_aos(self.h, ctypes.byref(p), _p(state), len(state), _p(a)))
        return a

    def transform_noalloc(self, state, kind="astro2", theta=1.0, e=1.0, precision="mixed",
                          acc=None):
        """transform() through the allocation-free entry point (at most
        65,536 bodies; its device buffers are the library's static array)."""
        state = np.ascontiguousarray(state, ENTITY)
        a = np.zeros(len(state), ACCEL) if acc is None else acc
        p = params(kind, theta, e, 0.0, precision)
        check(lib().pg_transform_aos_noalloc(self.h, ctypes.byref(p), _p(state), len(state),
                                             _p(a)))
        return a

    def tuning(self):
        t = Tuning()
        check(lib().pg_engine_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, max_buckets=8192, bucket_keys=1664, walk_run_log2=4, lds_buckets=1024):
        t = Tuning(max_buckets, bucket_keys, walk_run_log2, lds_buckets)
        check(lib().pg_engine_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, sims, trial_steps=4):
        """Times each candidate tuning on the batch (state saved and restored),
        keeps the fastest; returns (best, [(tuning, ms_per_step), ...])."""
        arr = (ctypes.c_void_p * len(sims))(*[s.h for s in sims])
        best, cap = Tuning(), 64
        tried = (Tuning * cap)()
        ms = np.zeros(cap)
        k = ctypes.c_uint32()
        check(lib().pg_engine_autotune(self.h, arr, len(sims), trial_steps, ctypes.byref(best),
                                       ctypes.cast(tried, ctypes.c_void_p), _p(ms), cap,
                                       ctypes.byref(k)))
        return best.as_dict(), [(tried[i].as_dict(), float(ms[i])) for i in range(k.value)]

    def batch_step(self, sims, steps):
        arr = (ctypes.c_void_p * len(sims))(*[s.h for s in sims])
        check(lib().pg_batch_step(self.h, arr, len(sims), steps))


INTEGRATOR = {"verlet": 0, "euler": 1, "rk4": 2}


class Sim:
    def __init__(self, engine, state=None, n=None, kind="astro2", theta=1.0, e=1.0, dt=1e-6,
                 precision="mixed", integrator="verlet"):
        self.engine = engine
        self.n = len(state) if state is not None else int(n)
        self.params = params(kind, theta, e, dt, precision)
        h = ctypes.c_void_p()
        check(lib().pg_sim_create(engine.h, self.n, ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h
        engine._children.add(self)
        if integrator != "verlet":
            self.set_integrator(integrator)
        if state is not None:
            self.upload(state)

    def checkpoint(self):
        """(state, previous or None): the state and its Stormer history."""
        st, prev, h = np.zeros(self.n, ENTITY), np.zeros(self.n, ENTITY), ctypes.c_int()
        check(lib().pg_sim_checkpoint(self.h, _p(st), _p(prev), self.n, ctypes.byref(h)))
        return st, (prev if h.value else None)

    def resume(self, state, previous=None):
        """Uploads a checkpoint: the next step continues its history."""
        state = np.ascontiguousarray(state, ENTITY)
        prev = None if previous is None else np.ascontiguousarray(previous, ENTITY)
        check(lib().pg_sim_upload_state(self.h, _p(state), _p(prev), len(state)))

    def set_near_field(self, leaf_widths):
        """Direct-sum near field: cells within `leaf_widths` widths of each
        body's own leaf cell are opened (0: the reference's walk)."""
        check(lib().pg_sim_set_near_field(self.h, float(leaf_widths)))

    def set_integrator(self, name):
        """verlet (default), euler or rk4: the reference's integrator elements."""
        check(lib().pg_sim_set_integrator(self.h, INTEGRATOR[name]))

    def close(self):
        if getattr(self, "h", None):
            lib().pg_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_params(self, **kw):
        cur = {
            "kind": self.params.kind,
            "theta": self.params.theta,
            "e": self.params.e,
            "dt": self.params.dt,
            "precision": self.params.precision,
        }
        cur.update(kw)
        self.params = params(**cur)
        check(lib().pg_sim_set_params(self.h, ctypes.byref(self.params)))

    def upload(self, state):
        state = np.ascontiguousarray(state, ENTITY)
        assert len(state) == self.n
        check(lib().pg_sim_upload(self.h, _p(state), self.n))

    def download(self):
        out = np.zeros(self.n, ENTITY)
        check(lib().pg_sim_download(self.h, _p(out), self.n))
        return out

    def step(self, steps=1):
        check(lib().pg_sim_step(self.h, steps))

    def accel(self):
        a = np.zeros((self.n, 3))
        check(lib().pg_sim_accel(self.h, _p(a)))
        return a

    def energy(self):
        out = np.zeros(6)
        check(lib().pg_sim_energy(self.h, _p(out)))
        return out

    def stats(self):
        s = Stats()
        check(lib().pg_sim_stats(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def set_counting(self, on):
        check(lib().pg_sim_set_counting(self.h, int(on)))

    def interaction_counts(self):
        c = np.zeros(self.n, np.int64)
        check(lib().pg_sim_interaction_counts(self.h, _p(c)))
        return c

    def tree(self):
        count = ctypes.c_uint64()
        check(lib().pg_sim_tree(self.h, None, None, 0, ctypes.byref(count)))
        n = count.value
        rec = np.zeros((n, 4), np.int64)
        fl = np.zeros((n, 8))
        check(lib().pg_sim_tree(self.h, _p(rec), _p(fl), n, ctypes.byref(count)))
        return rec, fl

    def morton(self):
        keys = np.zeros(self.n, np.uint64)
        perm = np.zeros(self.n, np.uint32)
        check(lib().pg_sim_morton(self.h, _p(keys), _p(perm), self.n))
        return keys, perm


class Server:
    def __init__(self, engine, max_active=0):
        self.engine = engine
        h = ctypes.c_void_p()
        check(lib().pg_server_create(engine.h, max_active, ctypes.byref(h)))
        self.h = h
        engine._children.add(self)

    def close(self):
        if getattr(self, "h", None):
            lib().pg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, state, steps, kind="astro2", theta=1.0, e=1.0, dt=1e-6, precision="mixed"):
        state = np.ascontiguousarray(state, ENTITY)
        p = params(kind, theta, e, dt, precision)
        rid = ctypes.c_uint64()
        check(lib().pg_server_submit(self.h, _p(state), len(state), ctypes.byref(p), steps,
                                     ctypes.byref(rid)))
        return rid.value

    def tick(self, max_steps=1):
        a, q = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().pg_server_tick(self.h, max_steps, ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def info(self, rid):
        info = RequestInfo()
        rc = lib().pg_server_info(self.h, rid, ctypes.byref(info))
        return info, rc == 0

    def take(self, rid, n):
        out = np.zeros(n, ENTITY)
        info = RequestInfo()
        check(lib().pg_server_take(self.h, rid, _p(out), n, ctypes.byref(info)))
        return out, info

    def clock(self):
        return lib().pg_server_clock(self.h)


def copy_probe(engine, nbytes, reps=20):
    """Achieved (read + write) GB/s of a plain streaming copy of nbytes."""
    g = ctypes.c_double()
    check(lib().pg_engine_copy_probe(engine.h, nbytes, reps, ctypes.byref(g)))
    return g.value


def format_f64(v):
    """Rust's `{}` of an f64 (csvsink's number format), from the library."""
    buf = ctypes.create_string_buffer(400)
    lib().pg_format_f64(float(v), buf, 400)
    return buf.value.decode()


class Runner:
    """A device-resident pipeline (one GPU gravity transform + integrator)
    with csvsink renderers."""

    def __init__(self, engine, state, kind="astro2", theta=1.0, e=1.0, dt=1e-6,
                 precision="mixed", integrator="verlet"):
        state = np.ascontiguousarray(state, ENTITY)
        self.n = len(state)
        p = params(kind, theta, e, dt, precision)
        h = ctypes.c_void_p()
        check(lib().pg_runner_create(engine.h, ctypes.byref(p), INTEGRATOR[integrator],
                                     _p(state), self.n, ctypes.byref(h)))
        self.h = h

    def add_csvsink(self, file, print_n=1):
        check(lib().pg_runner_add_csvsink(self.h, str(file).encode(), print_n))

    def run(self, iterations):
        check(lib().pg_runner_run(self.h, iterations))

    def flush(self):
        check(lib().pg_runner_flush(self.h))

    def state(self):
        out = np.zeros(self.n, ENTITY)
        check(lib().pg_runner_state(self.h, _p(out), self.n))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().pg_runner_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
