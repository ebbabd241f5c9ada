"""physim_amd: MI355X gravity and integration engine for physim's astro2 /
astro / simple_astro transforms and the Stormer `verlet` step.

The product is the in-tree shared library libphysim_gpu.so (HIP kernels for
gfx950 + C-ABI). This package is its Python front end: ctypes bindings of
include/physim_gpu.h (`lib()`), a loader that drives the library's physim
plugin symbols the way the reference host does (`plugin`), and thin
object wrappers (`Engine`, `Sim`, `Server`).

There is no CPU fallback: every entry point raises if the library is missing.
"""
from ._lib import (  # noqa: F401
    ACCEL,
    ENTITY,
    KIND,
    PRECISION,
    PgError,
    Engine,
    Runner,
    format_f64,
    Server,
    Sim,
    gen_cube,
    gen_plummer_ref,
    gen_plummer_virial,
    gen_star,
    concat,
    cube_and_stars,
    lib,
    library_path,
)
