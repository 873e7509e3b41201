"""gpu-amr_amd: MI355X-native patch-AMR advection step (Python face).

A thin ctypes binding of libamr_hip.so (C ABI in include/amr_hip.h) that
mirrors the reference's ndtree / amr_solver interface
(include/ndtree/ndtree.hpp, include/solver/amr_solver.hpp).  There is no CPU
fallback: every call runs the HIP kernels, and importing the package fails
loudly when the shared library has not been built.
"""
from __future__ import annotations

import ctypes
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libamr_hip.so"

AMR_OK, AMR_ERR_HIP, AMR_ERR_CAPACITY, AMR_ERR_ARGUMENT, AMR_ERR_UNSUPPORTED, AMR_ERR_TOPOLOGY = range(6)
CRITERION_MAX_VALUE, CRITERION_GRADIENT = 0, 1
STABLE, REFINE, COARSEN = 0, 1, 2

# amr_halo_direction_metadata (include/cuda/halo_exchange.hpp:11-26 layout)
HALO_META_DTYPE = np.dtype(
    [
        ("neighbor", "<i4"),
        ("finer_ids", "<i4", (4,)),
        ("quadrant", "<i4", (3,)),
        ("relation", "i1"),
        ("pad", "i1", (3,)),
    ]
)


class AmrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class TreeConfig(ctypes.Structure):
    _fields_ = [
        ("rank", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("patch_size", ctypes.c_int32),
        ("halo_width", ctypes.c_int32),
        ("capacity", ctypes.c_uint64),
        ("domain_length", ctypes.c_double * 3),
        ("patch_extents", ctypes.c_int32 * 3),
        ("reserved", ctypes.c_int32),
    ]


def patch_shape(rank, patch_size):
    """Interior extents per layout dim (the reference's static_shape<N, M[, O]>,
    last = physical x) from an int (square / cubic) or a sequence."""
    if isinstance(patch_size, (int, np.integer)):
        return (int(patch_size),) * rank
    shape = tuple(int(x) for x in patch_size)
    if len(shape) != rank:
        raise ValueError(f"{rank}D patch needs {rank} extents, got {shape}")
    return shape


class Criterion(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("min_level", ctypes.c_int32),
        ("max_level", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("refine_threshold", ctypes.c_double),
        ("coarsen_threshold", ctypes.c_double),
    ]


_lib = None

# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    "amr_tree_create": [ctypes.POINTER(TreeConfig), ctypes.POINTER(ctypes.c_void_p)],
    "amr_tree_destroy": [ctypes.c_void_p],
    "amr_tree_size": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)],
    "amr_tree_get_keys": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
    "amr_tree_get_neighbor_metadata": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
    "amr_tree_reconstruct_flags": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)],
    "amr_tree_reconstruct_criterion": [ctypes.c_void_p, ctypes.POINTER(Criterion), ctypes.POINTER(ctypes.c_int)],
    "amr_tree_compute_flags": [ctypes.c_void_p, ctypes.POINTER(Criterion), ctypes.c_void_p, ctypes.c_uint64],
    "amr_tree_compute_flags_noalloc": [ctypes.c_void_p, ctypes.POINTER(Criterion), ctypes.c_void_p, ctypes.c_uint64],
    "amr_tree_set_interior": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
    "amr_tree_get_interior": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
    "amr_tree_get_patches": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
    "amr_tree_halo_exchange_update": [ctypes.c_void_p],
    "amr_tree_swap_buffers": [ctypes.c_void_p],
    "amr_tree_device_interior": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)],
    "amr_tree_stream": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)],
    "amr_solver_configure": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_double],
    "amr_solver_compute_dt": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)],
    "amr_solver_advance_batch_async": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double],
    "amr_solver_finish_advance_batch": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)],
    "amr_solver_advance": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)],
    "amr_solver_step_variants": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)],
    "amr_solver_step_variant_name": [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p)],
    "amr_solver_set_step_variant": [ctypes.c_void_p, ctypes.c_int32],
    "amr_solver_get_step_variant": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)],
    "amr_solver_autotune": [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float)],
    "amr_tree_fields": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)],
    "amr_tree_write_vtk": [ctypes.c_void_p, ctypes.c_char_p],
    "amr_euler_create": [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)],
    "amr_euler_configure": [ctypes.c_void_p, ctypes.c_double, ctypes.c_double],
    "amr_euler_set_primitive": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
    "amr_solver_set_step_fusion": [ctypes.c_void_p, ctypes.c_int32],
    "amr_solver_get_step_fusion": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)],
    "amr_ctx_synchronize": [ctypes.c_void_p],
    "amr_batch_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)],
    "amr_batch_destroy": [ctypes.c_void_p],
    "amr_batch_size": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)],
    "amr_batch_advance": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int],
    "amr_batch_synchronize": [ctypes.c_void_p],
    "amr_batch_scatter": [ctypes.c_void_p],
    "amr_batch_stream": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)],
    "amr_batch_set_step_variant": [ctypes.c_void_p, ctypes.c_int32],
    "amr_batch_autotune": [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float)],
    "amr_tape_record": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                        ctypes.POINTER(ctypes.c_void_p)],
    "amr_tape_destroy": [ctypes.c_void_p],
    "amr_tape_get_info": [ctypes.c_void_p, ctypes.c_void_p],
    "amr_tape_get_dts": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
    "amr_tape_get_regrid": [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                            ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                            ctypes.c_uint64],
    "amr_tape_get_keys": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                          ctypes.POINTER(ctypes.c_uint64)],
    "amr_tape_forward": [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p],
    "amr_tape_adjoint": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)],
}


class TapeInfo(ctypes.Structure):
    _fields_ = [
        (name, ctypes.c_uint64)
        for name in ("steps", "checkpoint_every", "checkpoints", "regrids", "topologies", "initial_leaves",
                     "final_leaves", "max_leaves", "cells_per_leaf")
    ]


def exported_symbols():
    """The entry points include/amr_hip.h declares."""
    return ["amr_last_error", *_SIGS]


def load():
    """Load libamr_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or gpu-amr_amd/build.py")
    # PyTorch ships its own HIP runtime under the same soname as the system's
    # (libamdhip64.so.7).  Whichever is loaded first serves the whole
    # process, and torch cannot initialise the device on a runtime loaded
    # before it, so torch (when installed) is loaded first and the library
    # shares its runtime; tensors and this library's buffers then live in one
    # runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(str(LIB_PATH))
    lib.amr_last_error.restype = ctypes.c_char_p
    lib.amr_last_error.argtypes = []
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != AMR_OK:
        raise AmrError(rc, load().amr_last_error().decode())


def _vp(a):
    return ctypes.c_void_p(a.ctypes.data)


def decode_keys(keys, rank):
    """Morton keys -> (coords[n, rank] at 2**depth resolution, level[n])."""
    keys = np.asarray(keys, dtype=np.uint64)
    m = keys >> np.uint64(6)
    lvl = (keys & np.uint64(63)).astype(np.int64)
    c = np.zeros((keys.size, rank), dtype=np.int64)
    for b in range(21):
        for i in range(rank):
            c[:, i] |= ((m >> np.uint64(rank * b + i)) & np.uint64(1)).astype(np.int64) << b
    return c, lvl


class NdTree:
    """Periodic quad-/octree of leaf patches on one GPU (ndtree.hpp:94-2157)."""

    def __init__(self, rank, max_depth, patch_size, halo_width, capacity, domain_lengths=(1.0, 1.0, 1.0), _euler=None):
        """patch_size: interior cells per dim (int), or the extents of layout
        dims 0..rank-1 (static_shape<N, M[, O]>, the last is physical x)."""
        self.lib = load()
        shape = patch_shape(rank, patch_size)
        cfg = TreeConfig()
        cfg.rank, cfg.max_depth, cfg.halo_width = rank, max_depth, halo_width
        cfg.patch_size = 0
        for d in range(rank):
            cfg.patch_extents[d] = shape[d]
        cfg.capacity = capacity
        L = list(domain_lengths) + [1.0] * (3 - len(domain_lengths))
        for i in range(3):
            cfg.domain_length[i] = L[i]
        h = ctypes.c_void_p()
        if _euler is None:
            _check(self.lib.amr_tree_create(ctypes.byref(cfg), ctypes.byref(h)))
        else:
            gamma, cfl = _euler
            _check(self.lib.amr_euler_create(ctypes.byref(cfg), float(gamma), float(cfl), ctypes.byref(h)))
        self.h = h
        nv = ctypes.c_int32()
        _check(self.lib.amr_tree_fields(self.h, ctypes.byref(nv)))
        self.fields = int(nv.value)
        self.rank, self.depth, self.halo = rank, max_depth, halo_width
        self.shape = shape
        self.n = shape[0] if len(set(shape)) == 1 else shape  # square / cubic: the extent
        self.lengths = tuple(L[:rank])
        self.ni = int(np.prod(shape))
        self.pad_shape = tuple(x + 2 * halo_width for x in shape)
        self.flat = int(np.prod(self.pad_shape))

    def close(self):
        if getattr(self, "h", None):
            self.lib.amr_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- topology ---------------------------------------------------------
    def size(self):
        n = ctypes.c_uint64()
        _check(self.lib.amr_tree_size(self.h, ctypes.byref(n)))
        return int(n.value)

    def keys(self):
        out = np.zeros(self.size(), dtype=np.uint64)
        _check(self.lib.amr_tree_get_keys(self.h, _vp(out), out.size))
        return out

    def get_node_index_at(self, i):
        return int(self.keys()[i])

    def neighbor_metadata(self):
        out = np.zeros(self.size() * 2 * self.rank, dtype=HALO_META_DTYPE)
        _check(self.lib.amr_tree_get_neighbor_metadata(self.h, _vp(out), out.size))
        return out

    def reconstruct_tree(self, fn):
        """reconstruct_tree(fn) (ndtree.hpp:1249-1271).

        fn is either a Criterion (evaluated on the device), an array of
        per-leaf flags, or a callable key -> flag evaluated per leaf.
        Returns True when the topology changed."""
        changed = ctypes.c_int(0)
        if isinstance(fn, Criterion):
            _check(self.lib.amr_tree_reconstruct_criterion(self.h, ctypes.byref(fn), ctypes.byref(changed)))
            return bool(changed.value)
        if callable(fn):
            flags = np.array([int(fn(int(k))) for k in self.keys()], dtype=np.int8)
        else:
            flags = np.ascontiguousarray(fn, dtype=np.int8)
        _check(self.lib.amr_tree_reconstruct_flags(self.h, _vp(flags), flags.size, ctypes.byref(changed)))
        return bool(changed.value)

    def refine_all(self):
        return self.reconstruct_tree(np.full(self.size(), REFINE, dtype=np.int8))

    def compute_flags(self, crit: Criterion):
        out = np.zeros(self.size(), dtype=np.int8)
        _check(self.lib.amr_tree_compute_flags(self.h, ctypes.byref(crit), _vp(out), out.size))
        return out

    def compute_flags_noalloc(self, crit: Criterion):
        """compute_flags staged in the library's fixed device array (no
        device allocation; trees of up to 2**20 leaves)."""
        out = np.zeros(self.size(), dtype=np.int8)
        _check(self.lib.amr_tree_compute_flags_noalloc(self.h, ctypes.byref(crit), _vp(out), out.size))
        return out

    # -- fields -----------------------------------------------------------
    def _field_shape(self, per_leaf):
        lead = (self.fields,) if self.fields > 1 else ()
        return lead + (self.size(),) + tuple(per_leaf)

    def set_interior(self, values):
        """Interiors of all leaves; with several fields, field-major
        [fields, leaves, n, ...]."""
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        _check(self.lib.amr_tree_set_interior(self.h, _vp(v), v.size))

    def get_interior(self):
        out = np.zeros(self.fields * self.size() * self.ni, dtype=np.float64)
        _check(self.lib.amr_tree_get_interior(self.h, _vp(out), out.size))
        return out.reshape(self._field_shape(self.shape))

    def get_patches(self):
        """Padded patches after halo_exchange_update (corners 0)."""
        out = np.zeros(self.fields * self.size() * self.flat, dtype=np.float64)
        _check(self.lib.amr_tree_get_patches(self.h, _vp(out), out.size))
        return out.reshape(self._field_shape(self.pad_shape))

    def write_vtk(self, path):
        """vtk_print::print (vtk_print.hpp:53-254): legacy-VTK binary
        unstructured grid of the current leaves and fields."""
        _check(self.lib.amr_tree_write_vtk(self.h, str(path).encode()))

    def get_patch(self, i):
        return self.get_patches()[i]

    def halo_exchange_update(self):
        _check(self.lib.amr_tree_halo_exchange_update(self.h))

    def swap_buffers(self):
        _check(self.lib.amr_tree_swap_buffers(self.h))

    def device_interior_ptr(self):
        p = ctypes.c_void_p()
        _check(self.lib.amr_tree_device_interior(self.h, ctypes.byref(p)))
        return p.value

    def stream(self):
        p = ctypes.c_void_p()
        _check(self.lib.amr_tree_stream(self.h, ctypes.byref(p)))
        return p.value

    def synchronize(self):
        _check(self.lib.amr_ctx_synchronize(self.h))

    # -- geometry (physics_system.hpp:150-231) -----------------------------
    def cell_sizes(self, levels):
        lv = np.asarray(levels, dtype=np.int64)
        out = np.zeros((lv.size, self.rank))
        maxc = float(1 << self.depth)
        for i in range(self.rank):
            pm = np.array([float(1 << (self.depth - int(x))) for x in lv])
            out[:, i] = (self.lengths[i] * pm / maxc) / float(self.shape[self.rank - 1 - i])
        return out

    def cell_centers(self):
        """Cell centres of every interior cell: [size, n^rank, rank] (x first)."""
        coords, lv = decode_keys(self.keys(), self.rank)
        maxc = float(1 << self.depth)
        cs = self.cell_sizes(lv)
        idx = np.indices(self.shape).reshape(self.rank, -1)  # layout order
        out = np.zeros((coords.shape[0], self.ni, self.rank))
        for i in range(self.rank):
            origin = self.lengths[i] * coords[:, i].astype(np.float64) / maxc
            data_idx = idx[self.rank - 1 - i].astype(np.float64)
            cell_origin = origin[:, None] + data_idx[None, :] * cs[:, i][:, None]
            out[:, :, i] = cell_origin + 0.5 * cs[:, i][:, None]
        return out


class AmrSolver:
    """amr_solver<Tree, Geometry, Advection, DIM> (amr_solver.hpp:28-544)."""

    def __init__(self, tree: NdTree, velocity=(1.0, 0.5, 0.0), cfl=0.4):
        self.tree = tree
        self.configure(velocity, cfl)

    def configure(self, velocity, cfl):
        v = (ctypes.c_double * 3)(*(list(velocity) + [0.0] * (3 - len(velocity))))
        _check(self.tree.lib.amr_solver_configure(self.tree.h, v, float(cfl)))
        self.velocity, self.cfl = tuple(velocity), cfl

    def get_tree(self):
        return self.tree

    def initialize(self, ic):
        """initialize(ic) (amr_solver.hpp:105-145): ic(coords[..., rank]) -> values."""
        cen = self.tree.cell_centers()
        self.tree.set_interior(np.asarray(ic(cen), dtype=np.float64))

    def compute_dt(self):
        dt = ctypes.c_double()
        _check(self.tree.lib.amr_solver_compute_dt(self.tree.h, ctypes.byref(dt)))
        return dt.value

    def advance_batch_async(self, step_count, remaining_time=np.finfo(np.float64).max):
        _check(self.tree.lib.amr_solver_advance_batch_async(self.tree.h, int(step_count), float(remaining_time)))

    def finish_advance_batch(self):
        s, n = ctypes.c_double(), ctypes.c_uint64()
        _check(self.tree.lib.amr_solver_finish_advance_batch(self.tree.h, ctypes.byref(s), ctypes.byref(n)))
        return s.value, int(n.value)

    def advance(self):
        dt = ctypes.c_double()
        _check(self.tree.lib.amr_solver_advance(self.tree.h, ctypes.byref(dt)))
        return dt.value

    # -- step-kernel autotuner ----------------------------------------------
    def step_variants(self):
        n = ctypes.c_int32()
        _check(self.tree.lib.amr_solver_step_variants(self.tree.h, ctypes.byref(n)))
        names = []
        for v in range(n.value):
            s = ctypes.c_char_p()
            _check(self.tree.lib.amr_solver_step_variant_name(self.tree.h, v, ctypes.byref(s)))
            names.append(s.value.decode())
        return names

    def set_step_variant(self, v):
        _check(self.tree.lib.amr_solver_set_step_variant(self.tree.h, int(v)))

    def step_variant(self):
        v = ctypes.c_int32()
        _check(self.tree.lib.amr_solver_get_step_variant(self.tree.h, ctypes.byref(v)))
        return v.value

    def autotune(self, trial_steps=10):
        """Times every step-kernel variant on the current tree, keeps the
        fastest; returns (variant index, microseconds per step)."""
        v, us = ctypes.c_int32(), ctypes.c_float()
        _check(self.tree.lib.amr_solver_autotune(self.tree.h, int(trial_steps), ctypes.byref(v), ctypes.byref(us)))
        return v.value, us.value

    def set_step_fusion(self, enable):
        """Two steps per launch on single-level trees (default on; results
        are bit-identical either way)."""
        _check(self.tree.lib.amr_solver_set_step_fusion(self.tree.h, 1 if enable else 0))

    def step_fusion(self):
        """(enabled, applies to the current tree)."""
        e, a = ctypes.c_int32(), ctypes.c_int32()
        _check(self.tree.lib.amr_solver_get_step_fusion(self.tree.h, ctypes.byref(e), ctypes.byref(a)))
        return bool(e.value), bool(a.value)


class EulerSolver:
    """amr_solver<tree, physics, EulerPhysics<rank>, rank>(capacity, gamma,
    cfl) (amr_solver.hpp:28-97, EulerPhysics.hpp) with its tree: the cells
    carry [rho, rho u, rho v, (rho w), E]; `tree` is the NdTree view of the
    same context (regrids, keys, fields [fields, leaves, n, ...])."""

    def __init__(self, rank, max_depth, patch_size, halo_width, capacity, domain_lengths=(1.0, 1.0, 1.0),
                 gamma=1.4, cfl=0.3):
        self.tree = NdTree(rank, max_depth, patch_size, halo_width, capacity, domain_lengths, _euler=(gamma, cfl))
        self.gamma, self.cfl = gamma, cfl

    def get_tree(self):
        return self.tree

    def configure(self, gamma, cfl):
        _check(self.tree.lib.amr_euler_configure(self.tree.h, float(gamma), float(cfl)))
        self.gamma, self.cfl = gamma, cfl

    def initialize(self, ic):
        """initialize(ic) (amr_solver.hpp:105-145): ic(coords[..., rank]) ->
        primitive values [..., rank + 2] = [rho, u, v, (w), p], converted to
        conservative variables by primitiveToConservative."""
        cen = self.tree.cell_centers()
        prim = np.asarray(ic(cen), dtype=np.float64)  # [leaves, n^rank, fields]
        v = np.ascontiguousarray(np.moveaxis(prim, -1, 0)).reshape(-1)
        _check(self.tree.lib.amr_euler_set_primitive(self.tree.h, _vp(v), v.size))

    def compute_dt(self):
        dt = ctypes.c_double()
        _check(self.tree.lib.amr_solver_compute_dt(self.tree.h, ctypes.byref(dt)))
        return dt.value

    def advance_batch_async(self, step_count, remaining_time=np.finfo(np.float64).max):
        _check(self.tree.lib.amr_solver_advance_batch_async(self.tree.h, int(step_count), float(remaining_time)))

    def finish_advance_batch(self):
        """(summed dt, executed steps)."""
        s, n = ctypes.c_double(), ctypes.c_uint64()
        _check(self.tree.lib.amr_solver_finish_advance_batch(self.tree.h, ctypes.byref(s), ctypes.byref(n)))
        return s.value, int(n.value)

    def advance(self):
        dt = ctypes.c_double()
        _check(self.tree.lib.amr_solver_advance(self.tree.h, ctypes.byref(dt)))
        return dt.value


class Batch:
    """Many independent trees stepped as one launch per step (optionally one
    hipGraph launch for many steps); each keeps its own velocity and dt."""

    def __init__(self, trees):
        self.lib = load()
        self.trees = list(trees)
        arr = (ctypes.c_void_p * len(self.trees))(*[t.h.value for t in self.trees])
        h = ctypes.c_void_p()
        _check(self.lib.amr_batch_create(arr, len(self.trees), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.amr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def size(self):
        n = ctypes.c_uint64()
        _check(self.lib.amr_batch_size(self.h, ctypes.byref(n)))
        return int(n.value)

    def advance(self, steps, use_graph=True):
        _check(self.lib.amr_batch_advance(self.h, int(steps), 1 if use_graph else 0))

    def synchronize(self):
        _check(self.lib.amr_batch_synchronize(self.h))

    def scatter(self):
        _check(self.lib.amr_batch_scatter(self.h))

    def stream(self):
        p = ctypes.c_void_p()
        _check(self.lib.amr_batch_stream(self.h, ctypes.byref(p)))
        return p.value

    def set_step_variant(self, v):
        _check(self.lib.amr_batch_set_step_variant(self.h, int(v)))

    def autotune(self, trial_steps=10):
        v, us = ctypes.c_int32(), ctypes.c_float()
        _check(self.lib.amr_batch_autotune(self.h, int(trial_steps), ctypes.byref(v), ctypes.byref(us)))
        return v.value, us.value


class Tape:
    """Recorded forward solve for the discrete adjoint (amr_tape_* in
    include/amr_hip.h).  Recording advances the solver's tree `steps` steps,
    regridding with `criterion` before every step k > 0 with
    k % regrid_every == 0, and keeps the state before every
    checkpoint_every-th step.  forward()/adjoint() take device pointers
    (ints) -- torch tensors on the GPU pass tensor.data_ptr()."""

    def __init__(self, solver: AmrSolver, steps, criterion: Criterion | None = None, regrid_every=0,
                 checkpoint_every=20):
        self.lib = load()
        self.tree = solver.tree
        h = ctypes.c_void_p()
        crit = ctypes.byref(criterion) if criterion is not None else None
        _check(self.lib.amr_tape_record(self.tree.h, crit, int(steps), int(regrid_every), int(checkpoint_every),
                                        ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.amr_tape_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def info(self):
        i = TapeInfo()
        _check(self.lib.amr_tape_get_info(self.h, ctypes.byref(i)))
        return {name: int(getattr(i, name)) for name, _ in TapeInfo._fields_}

    def dts(self):
        n = self.info()["steps"]
        out = np.zeros(n)
        _check(self.lib.amr_tape_get_dts(self.h, _vp(out), n))
        return out

    def regrid(self, index):
        """(step, leaves before, changed, flags) of recorded regrid `index`."""
        step, nb, ch = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        _check(self.lib.amr_tape_get_regrid(self.h, int(index), ctypes.byref(step), ctypes.byref(nb), ctypes.byref(ch),
                                            None, 0))
        flags = np.zeros(nb.value, dtype=np.int8)
        _check(self.lib.amr_tape_get_regrid(self.h, int(index), None, None, None, _vp(flags), flags.size))
        return int(step.value), int(nb.value), bool(ch.value), flags

    def keys(self, topology):
        n = ctypes.c_uint64()
        _check(self.lib.amr_tape_get_keys(self.h, int(topology), None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.uint64)
        _check(self.lib.amr_tape_get_keys(self.h, int(topology), _vp(out), out.size, ctypes.byref(n)))
        return out

    def forward(self, u0_ptr, velocity, uT_ptr):
        v = (ctypes.c_double * 3)(*(list(velocity) + [0.0] * (3 - len(velocity))))
        _check(self.lib.amr_tape_forward(self.h, ctypes.c_void_p(u0_ptr), v, ctypes.c_void_p(uT_ptr)))

    def adjoint(self, lamT_ptr, grad_u0_ptr):
        g = (ctypes.c_double * 3)()
        _check(self.lib.amr_tape_adjoint(self.h, ctypes.c_void_p(lamT_ptr), ctypes.c_void_p(grad_u0_ptr), g))
        return np.array(g[:])


def make_criterion(kind, refine_threshold, coarsen_threshold, min_level, max_level):
    c = Criterion()
    c.kind, c.min_level, c.max_level = kind, min_level, max_level
    c.refine_threshold, c.coarsen_threshold = refine_threshold, coarsen_threshold
    return c
