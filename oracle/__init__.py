"""ctypes wrapper of the CPU oracle (test infrastructure only).

The oracle (cpu_ref.hpp) restates the reference's CPU algorithm for the
patch-AMR advection step.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module; the product path under
gpu-amr_amd/ never does.  load_ref() opens the reference's own physics
functions (_ref/libref_physics.so, compiled from the reference headers where
they are present), against which the tests check the restatement.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"

# Reference device metadata record (include/cuda/halo_exchange.hpp:11-26).
HALO_META_DTYPE = np.dtype(
    [
        ("neighbor", "<i4"),
        ("finer_ids", "<i4", (4,)),
        ("quadrant", "<i4", (3,)),
        ("relation", "i1"),
        ("pad", "i1", (3,)),
    ]
)
assert HALO_META_DTYPE.itemsize == 36

_lib = None


def load(path=None):
    """Load liboracle.so (or the build at `path`, e.g. build_native()'s; the
    first call decides which build every OracleTree / OracleEuler uses)."""
    global _lib
    if _lib is not None:
        return _lib
    path = pathlib.Path(path) if path else _LIB_PATH
    if not path.exists():
        raise RuntimeError(f"{path} is missing: run `make -C oracle` (or __graft_entry__.build())")
    lib = ctypes.CDLL(str(path))
    vp, sz, dp, i8p, u64p = (
        ctypes.c_void_p,
        ctypes.c_size_t,
        ctypes.POINTER(ctypes.c_double),
        ctypes.POINTER(ctypes.c_int8),
        ctypes.POINTER(ctypes.c_uint64),
    )
    lib.oracle_last_error.restype = ctypes.c_char_p
    lib.oracle_create.restype = vp
    lib.oracle_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, sz, dp]
    lib.oracle_destroy.argtypes = [vp]
    lib.oracle_size.restype = sz
    lib.oracle_size.argtypes = [vp]
    lib.oracle_keys.argtypes = [vp, u64p]
    lib.oracle_metadata.argtypes = [vp, vp]
    lib.oracle_reconstruct.argtypes = [vp, i8p, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_last_sets.argtypes = [vp, u64p, ctypes.POINTER(sz), u64p, ctypes.POINTER(sz)]
    for name in ("oracle_set_interior", "oracle_get_interior", "oracle_set_padded", "oracle_get_padded", "oracle_cell_centers"):
        getattr(lib, name).argtypes = [vp, dp]
    lib.oracle_halo_exchange.argtypes = [vp]
    lib.oracle_set_params.argtypes = [vp, dp, ctypes.c_double, ctypes.c_int]
    lib.oracle_compute_dt.restype = ctypes.c_double
    lib.oracle_compute_dt.argtypes = [vp]
    lib.oracle_step.argtypes = [vp, ctypes.c_double]
    lib.oracle_step_kernel_only.argtypes = [vp, ctypes.c_double]
    lib.oracle_swap_buffers.argtypes = [vp]
    lib.oracle_advance_batch.argtypes = [vp, sz, ctypes.c_double, dp, ctypes.POINTER(sz)]
    lib.oracle_criterion.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, i8p]
    # Euler system
    lib.oracle_euler_create.restype = vp
    lib.oracle_euler_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, sz, dp]
    lib.oracle_euler_destroy.argtypes = [vp]
    lib.oracle_euler_size.restype = sz
    lib.oracle_euler_size.argtypes = [vp]
    lib.oracle_euler_fields.argtypes = [vp]
    lib.oracle_euler_keys.argtypes = [vp, u64p]
    lib.oracle_euler_metadata.argtypes = [vp, vp]
    lib.oracle_euler_reconstruct.argtypes = [vp, i8p, ctypes.POINTER(ctypes.c_int)]
    for name in ("oracle_euler_set_interior", "oracle_euler_get_interior", "oracle_euler_get_padded", "oracle_euler_cell_centers"):
        getattr(lib, name).argtypes = [vp, dp]
    lib.oracle_euler_halo_exchange.argtypes = [vp]
    lib.oracle_euler_set_params.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.oracle_euler_compute_dt.restype = ctypes.c_double
    lib.oracle_euler_compute_dt.argtypes = [vp]
    lib.oracle_euler_step.argtypes = [vp, ctypes.c_double]
    lib.oracle_euler_advance_batch.argtypes = [vp, sz, ctypes.c_double, dp, ctypes.POINTER(sz)]
    lib.oracle_euler_criterion.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, i8p]
    lib.oracle_euler_prim_to_cons.argtypes = [ctypes.c_int, dp, dp, sz, ctypes.c_double]
    lib.oracle_euler_flux.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_int, ctypes.c_double]
    lib.oracle_euler_max_speed.restype = ctypes.c_double
    lib.oracle_euler_max_speed.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_double]
    _lib = lib
    return lib


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class _Shaped:
    """Patch shape of an oracle instance: n is the extent of every dim (int)
    or the extents of layout dims 0..rank-1 (static_shape<N, M[, O]>, the
    last is physical x)."""

    def _shape(self, n):
        shape = (int(n),) * self.rank if isinstance(n, (int, np.integer)) else tuple(int(x) for x in n)
        assert len(shape) == self.rank
        self.shape = shape
        self.n = shape[0] if len(set(shape)) == 1 else shape
        self.ni = int(np.prod(shape))
        self.pad_shape = tuple(x + 2 * self.halo for x in shape)
        self.pad = self.pad_shape[0] if len(set(self.pad_shape)) == 1 else self.pad_shape
        self.flat = int(np.prod(self.pad_shape))
        self._ext = (ctypes.c_int * 3)(*(list(shape) + [1] * (3 - len(shape))))


class OracleTree(_Shaped):
    """The reference's ndtree + amr_solver (advection) restated on the CPU."""

    CRIT_MAX_VALUE = 0
    CRIT_GRADIENT = 1

    def __init__(self, rank, n, halo, depth, capacity, lengths=(1.0, 1.0, 1.0)):
        self.lib = load()
        self.rank, self.halo, self.depth = rank, halo, depth
        self._shape(n)
        L = np.ascontiguousarray(np.array(list(lengths) + [1.0] * (3 - len(lengths)), dtype=np.float64)[:3])
        self.h = self.lib.oracle_create(rank, self._ext, halo, depth, capacity, _ptr(L, ctypes.c_double))
        if not self.h:
            raise RuntimeError(self.lib.oracle_last_error().decode())

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.oracle_destroy(h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.oracle_last_error().decode())

    def size(self):
        return int(self.lib.oracle_size(self.h))

    def keys(self):
        out = np.zeros(self.size(), dtype=np.uint64)
        self._check(self.lib.oracle_keys(self.h, _ptr(out, ctypes.c_uint64)))
        return out

    def metadata(self):
        out = np.zeros(self.size() * 2 * self.rank, dtype=HALO_META_DTYPE)
        self._check(self.lib.oracle_metadata(self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def reconstruct(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int8)
        assert f.shape == (self.size(),)
        changed = ctypes.c_int(0)
        self._check(self.lib.oracle_reconstruct(self.h, _ptr(f, ctypes.c_int8), ctypes.byref(changed)))
        return bool(changed.value)

    def refine_all(self):
        return self.reconstruct(np.ones(self.size(), dtype=np.int8))

    def last_sets(self):
        nr, nc = ctypes.c_size_t(0), ctypes.c_size_t(0)
        self._check(self.lib.oracle_last_sets(self.h, None, ctypes.byref(nr), None, ctypes.byref(nc)))
        r = np.zeros(nr.value, dtype=np.uint64)
        c = np.zeros(nc.value, dtype=np.uint64)
        self._check(self.lib.oracle_last_sets(self.h, _ptr(r, ctypes.c_uint64), ctypes.byref(nr), _ptr(c, ctypes.c_uint64), ctypes.byref(nc)))
        return r, c

    def set_interior(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        assert v.size == self.size() * self.ni
        self._check(self.lib.oracle_set_interior(self.h, _ptr(v, ctypes.c_double)))

    def get_interior(self):
        out = np.zeros(self.size() * self.ni, dtype=np.float64)
        self._check(self.lib.oracle_get_interior(self.h, _ptr(out, ctypes.c_double)))
        return out.reshape(self.size(), *self.shape)

    def set_padded(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        assert v.size == self.size() * self.flat
        self._check(self.lib.oracle_set_padded(self.h, _ptr(v, ctypes.c_double)))

    def get_padded(self):
        out = np.zeros(self.size() * self.flat, dtype=np.float64)
        self._check(self.lib.oracle_get_padded(self.h, _ptr(out, ctypes.c_double)))
        return out.reshape(self.size(), *self.pad_shape)

    def cell_centers(self):
        out = np.zeros(self.size() * self.ni * self.rank, dtype=np.float64)
        self._check(self.lib.oracle_cell_centers(self.h, _ptr(out, ctypes.c_double)))
        return out.reshape(self.size(), self.ni, self.rank)

    def halo_exchange(self):
        self._check(self.lib.oracle_halo_exchange(self.h))

    def set_params(self, velocity, cfl, threads=1):
        v = np.zeros(3, dtype=np.float64)
        v[: len(velocity)] = velocity
        self._check(self.lib.oracle_set_params(self.h, _ptr(v, ctypes.c_double), float(cfl), int(threads)))

    def compute_dt(self):
        return float(self.lib.oracle_compute_dt(self.h))

    def step(self, dt):
        self._check(self.lib.oracle_step(self.h, float(dt)))

    def step_kernel_only(self, dt):
        self._check(self.lib.oracle_step_kernel_only(self.h, float(dt)))

    def swap_buffers(self):
        """ndtree::swap_buffers: the current and next state exchange roles."""
        self._check(self.lib.oracle_swap_buffers(self.h))

    def advance_batch(self, steps, remaining=np.finfo(np.float64).max):
        s = ctypes.c_double(0.0)
        ex = ctypes.c_size_t(0)
        self._check(self.lib.oracle_advance_batch(self.h, steps, float(remaining), ctypes.byref(s), ctypes.byref(ex)))
        return s.value, ex.value

    def advance(self):
        return self.advance_batch(1)[0]

    def criterion(self, kind, refine_thr, coarsen_thr, min_level, max_level):
        out = np.zeros(self.size(), dtype=np.int8)
        self._check(self.lib.oracle_criterion(self.h, int(kind), float(refine_thr), float(coarsen_thr), int(min_level), int(max_level), _ptr(out, ctypes.c_int8)))
        return out


class OracleEuler(_Shaped):
    """The reference's ndtree + amr_solver<..., EulerPhysics<R>, R> restated on
    the CPU (EulerPhysics.hpp, amr_solver.hpp:105-413).  Host arrays are
    field-major: all leaves of rho, then rho*u, rho*v, (rho*w), E."""

    def __init__(self, rank, n, halo, depth, capacity, lengths=(1.0, 1.0, 1.0)):
        self.lib = load()
        self.rank, self.halo, self.depth = rank, halo, depth
        self.nv = rank + 2
        self._shape(n)
        L = np.ascontiguousarray(np.array(list(lengths) + [1.0] * (3 - len(lengths)), dtype=np.float64)[:3])
        self.h = self.lib.oracle_euler_create(rank, self._ext, halo, depth, capacity, _ptr(L, ctypes.c_double))
        if not self.h:
            raise RuntimeError(self.lib.oracle_last_error().decode())

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.oracle_euler_destroy(h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.oracle_last_error().decode())

    def size(self):
        return int(self.lib.oracle_euler_size(self.h))

    def keys(self):
        out = np.zeros(self.size(), dtype=np.uint64)
        self._check(self.lib.oracle_euler_keys(self.h, _ptr(out, ctypes.c_uint64)))
        return out

    def metadata(self):
        out = np.zeros(self.size() * 2 * self.rank, dtype=HALO_META_DTYPE)
        self._check(self.lib.oracle_euler_metadata(self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def reconstruct(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int8)
        assert f.shape == (self.size(),)
        changed = ctypes.c_int(0)
        self._check(self.lib.oracle_euler_reconstruct(self.h, _ptr(f, ctypes.c_int8), ctypes.byref(changed)))
        return bool(changed.value)

    def refine_all(self):
        return self.reconstruct(np.ones(self.size(), dtype=np.int8))

    def set_interior(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        assert v.size == self.nv * self.size() * self.ni
        self._check(self.lib.oracle_euler_set_interior(self.h, _ptr(v, ctypes.c_double)))

    def get_interior(self):
        out = np.zeros(self.nv * self.size() * self.ni, dtype=np.float64)
        self._check(self.lib.oracle_euler_get_interior(self.h, _ptr(out, ctypes.c_double)))
        return out.reshape(self.nv, self.size(), *self.shape)

    def get_padded(self):
        out = np.zeros(self.nv * self.size() * self.flat, dtype=np.float64)
        self._check(self.lib.oracle_euler_get_padded(self.h, _ptr(out, ctypes.c_double)))
        return out.reshape(self.nv, self.size(), *self.pad_shape)

    def cell_centers(self):
        out = np.zeros(self.size() * self.ni * self.rank, dtype=np.float64)
        self._check(self.lib.oracle_euler_cell_centers(self.h, _ptr(out, ctypes.c_double)))
        return out.reshape(self.size(), self.ni, self.rank)

    def halo_exchange(self):
        self._check(self.lib.oracle_euler_halo_exchange(self.h))

    def set_params(self, gamma, cfl, threads=1):
        self._check(self.lib.oracle_euler_set_params(self.h, float(gamma), float(cfl), int(threads)))

    def compute_dt(self):
        return float(self.lib.oracle_euler_compute_dt(self.h))

    def step(self, dt):
        self._check(self.lib.oracle_euler_step(self.h, float(dt)))

    def advance_batch(self, steps, remaining=np.finfo(np.float64).max):
        s = ctypes.c_double(0.0)
        ex = ctypes.c_size_t(0)
        self._check(self.lib.oracle_euler_advance_batch(self.h, steps, float(remaining), ctypes.byref(s), ctypes.byref(ex)))
        return s.value, ex.value

    def criterion(self, kind, refine_thr, coarsen_thr, min_level, max_level):
        out = np.zeros(self.size(), dtype=np.int8)
        self._check(self.lib.oracle_euler_criterion(self.h, int(kind), float(refine_thr), float(coarsen_thr), int(min_level), int(max_level), _ptr(out, ctypes.c_int8)))
        return out


def euler_prim_to_cons(rank, prim, gamma):
    """primitiveToConservative per cell; prim[..., rank + 2] -> cons[..., rank + 2]."""
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    cons = np.zeros_like(prim)
    lib = load()
    cells = prim.size // (rank + 2)
    rc = lib.oracle_euler_prim_to_cons(rank, _ptr(prim, ctypes.c_double), _ptr(cons, ctypes.c_double), cells, float(gamma))
    if rc != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return cons


def euler_flux(rank, UL, UR, direction, gamma):
    """rusanovFlux(UL, UR, flux, direction, gamma)."""
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    out = np.zeros(rank + 2, dtype=np.float64)
    lib = load()
    rc = lib.oracle_euler_flux(rank, _ptr(UL, ctypes.c_double), _ptr(UR, ctypes.c_double), _ptr(out, ctypes.c_double), int(direction), float(gamma))
    if rc != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return out


def euler_max_speed(rank, U, direction, gamma):
    """getMaxSpeed of one conservative state."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    return float(load().oracle_euler_max_speed(rank, _ptr(U, ctypes.c_double), int(direction), float(gamma)))


def advection_flux(ul, ur, v):
    """rusanovFlux of AdvectionPhysics for velocity component v."""
    lib = load()
    lib.oracle_advection_flux.restype = ctypes.c_double
    lib.oracle_advection_flux.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
    return float(lib.oracle_advection_flux(float(ul), float(ur), float(v)))


# ---------------------------------------------------------------------------
# The reference's own physics functions (oracle/_ref/libref_physics.so, built
# from the reference headers by `make -C oracle ref` where the reference
# sources are present): the checker of the restatement above.
# ---------------------------------------------------------------------------
_REF_PATH = _HERE / "_ref" / "libref_physics.so"


def load_ref():
    """The reference-physics library, or None where it is not built."""
    if not _REF_PATH.exists():
        return None
    lib = ctypes.CDLL(str(_REF_PATH))
    d = ctypes.POINTER(ctypes.c_double)
    lib.ref_euler_prim_to_cons.argtypes = [ctypes.c_int, d, d, ctypes.c_double]
    lib.ref_euler_flux.argtypes = [ctypes.c_int, d, d, d, ctypes.c_int, ctypes.c_double]
    lib.ref_euler_max_speed.argtypes = [ctypes.c_int, d, ctypes.c_int, ctypes.c_double]
    lib.ref_euler_max_speed.restype = ctypes.c_double
    lib.ref_advection_flux.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.ref_advection_flux.restype = ctypes.c_double
    return lib


def load_ref_morton():
    """ref_morton_key(rank, x, y, z, level) from the reference's libmorton,
    or None where it is not built."""
    path = _HERE / "_ref" / "libref_morton.so"
    if not path.exists():
        return None
    lib = ctypes.CDLL(str(path))
    lib.ref_morton_key.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    lib.ref_morton_key.restype = ctypes.c_uint64
    return lib


def build_ref(reference="/root/reference", verbose=False):
    """Compile oracle/_ref/ (the reference's physics and Morton encoding) from
    the reference headers (only where the reference sources are present)."""
    import subprocess

    r = subprocess.run(["make", "-s", "-C", str(_HERE), "ref", f"REFERENCE={reference}"], capture_output=not verbose,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference physics build failed:\n{r.stdout}\n{r.stderr}")
    return str(_REF_PATH)


def build_native(out_dir):
    """Start compiling the oracle for this host's CPU (-march=native, same
    -ffp-contract=off arithmetic) into out_dir/liboracle_native.so; returns
    (Popen, path).  The CPU baseline uses it where the compile succeeds."""
    import subprocess

    out = pathlib.Path(out_dir) / "liboracle_native.so"
    cmd = ["g++", "-std=c++20", "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
           "-o", str(out), str(_HERE / "cpu_ref_capi.cpp"), "-lpthread"]
    return subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE), out


def build(verbose=False):
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    r = subprocess.run(["make", "-s", "-C", str(_HERE)], capture_output=not verbose, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return str(_LIB_PATH)
