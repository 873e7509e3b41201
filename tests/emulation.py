"""ctypes face of tests/cpp/libkernel_emulation.so: the HIP kernels' item
bodies executed on the CPU (test infrastructure)."""
import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
SRC = HERE / "cpp" / "kernel_emulation.hip"
LIB = HERE / "cpp" / "libkernel_emulation.so"

FACE_LINK = np.dtype([("ids", "<i4", (4,))])
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE / "cpp"), str(LIB)], check=True, capture_output=True)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists() or LIB.stat().st_mtime < max(p.stat().st_mtime for p in (ROOT / "gpu-amr_amd" / "csrc").glob("*.hpp")):
            build()
        _lib = ctypes.CDLL(str(LIB))
        for name in ("emu_build_table", "emu_step", "emu_padded", "emu_flags", "emu_regrid", "emu_adj_step",
                     "emu_adj_regrid", "emu_runs", "emu_uniform_neighbors", "emu_step_box", "emu_padded_box",
                     "emu_flags_box", "emu_transfer_box", "emu_adj_step_box", "emu_adj_regrid_box"):
            getattr(_lib, name).restype = ctypes.c_int
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def build_table(rank, depth, keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    n = keys.size
    links = np.zeros(n * 2 * rank, dtype=FACE_LINK)
    metas = np.zeros(n * 2 * rank, dtype=np.uint32)
    err = load().emu_build_table(rank, depth, _p(keys), ctypes.c_uint32(n), _p(links), _p(metas))
    return links, metas, err


def to_reference_metadata(rank, links, metas):
    from oracle import HALO_META_DTYPE

    out = np.zeros(metas.size, dtype=HALO_META_DTYPE)
    out["neighbor"] = -1
    out["finer_ids"] = -1
    rel = (metas & 3).astype(np.int8)
    out["relation"] = rel
    same_or_coarse = (rel == 1) | (rel == 3)
    out["neighbor"][same_or_coarse] = links["ids"][same_or_coarse, 0]
    fin = rel == 2
    k = 1 << (rank - 1)
    out["finer_ids"][fin, :k] = links["ids"][fin, :k]
    co = rel == 3
    for d in range(rank):
        out["quadrant"][co, d] = (metas[co] >> (2 + d)) & 1
    return out


def step_params(rank, n, depth, max_level, vel, dt, lengths=(1.0, 1.0, 1.0)):
    """n: cells per dim (int) or the extents of layout dims 0..rank-1."""
    ext = (n,) * rank if isinstance(n, int) else tuple(n)
    dtdx = np.zeros((17, 3))
    for lv in range(depth + 1):
        for d in range(rank):
            cs = (lengths[d] * float(1 << (depth - lv)) / float(1 << depth)) / float(ext[rank - 1 - d])
            dtdx[lv, d] = dt / cs
    v = np.zeros(3)
    v[:rank] = vel
    hs = 0.5 * np.abs(v)
    return dtdx, v, hs


def step(rank, n, keys, links, metas, u, dtdx, vel, hs):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    un = np.zeros_like(u)
    rc = load().emu_step(rank, n, _p(keys), ctypes.c_uint32(keys.size), _p(links), _p(metas), _p(u), _p(un),
                         _p(np.ascontiguousarray(dtdx)), _p(vel), _p(hs))
    assert rc == 0
    return un


def padded(rank, n, h, links, metas, u, nleaves):
    u = np.ascontiguousarray(u, dtype=np.float64)
    pad = np.zeros(nleaves * (n + 2 * h) ** rank)
    rc = load().emu_padded(rank, n, h, ctypes.c_uint32(nleaves), _p(links), _p(metas), _p(u), _p(pad))
    assert rc == 0
    return pad.reshape((nleaves,) + (n + 2 * h,) * rank)


def flags(rank, n, h, kind, keys, links, metas, u, rthr, cthr, mn, mx):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(keys.size, dtype=np.int8)
    rc = load().emu_flags(rank, n, h, kind, _p(keys), ctypes.c_uint32(keys.size), _p(links), _p(metas), _p(u),
                          ctypes.c_double(rthr), ctypes.c_double(cthr), mn, mx, _p(out))
    assert rc == 0
    return out


def regrid(rank, n, depth, keys, fl, links, metas, u, max_level, cap, with_maps=False):
    """The regrid kernels in sequence; with_maps also returns the (4, n_old)
    uint32 rows mark, head, counts, pos of the old->new leaf map."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    fl = np.ascontiguousarray(fl, dtype=np.int8)
    keys_out = np.zeros(cap, dtype=np.uint64)
    u_out = np.zeros(cap * n**rank)
    summary = np.zeros(4, dtype=np.uint32)
    maps = np.zeros((4, keys.size), dtype=np.uint32)
    changed = load().emu_regrid(rank, n, depth, _p(keys), ctypes.c_uint32(keys.size), _p(fl), _p(links), _p(metas),
                                _p(u), max_level, _p(keys_out), _p(u_out), ctypes.c_uint32(cap), _p(summary), _p(maps))
    total = int(summary[0])
    out = (bool(changed), keys_out[:total], u_out[: total * n**rank], summary)
    return out + (maps,) if with_maps else out


def adj_step(rank, n, keys, links, metas, lam_next, u, dtdx, vel):
    """Transposed step (lam) and per-leaf velocity partials (n_leaves, 3)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    lam_next = np.ascontiguousarray(lam_next, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.zeros(3)
    v[: len(vel)] = vel
    lam = np.zeros_like(lam_next)
    vpart = np.zeros((keys.size, 3))
    rc = load().emu_adj_step(rank, n, _p(keys), ctypes.c_uint32(keys.size), _p(links), _p(metas), _p(lam_next), _p(u),
                             _p(np.ascontiguousarray(dtdx)), _p(v), _p(lam), _p(vpart))
    assert rc == 0
    return lam, vpart


def adj_regrid(rank, n, lam_new, maps, n_old):
    lam_new = np.ascontiguousarray(lam_new, dtype=np.float64)
    maps = np.ascontiguousarray(maps, dtype=np.uint32)
    lam_old = np.zeros(n_old * n**rank)
    rc = load().emu_adj_regrid(rank, n, _p(lam_new), _p(maps), ctypes.c_uint32(n_old), _p(lam_old))
    assert rc == 0
    return lam_old


def runs(rank, depth, keys, links, metas, K):
    """(head, len) rows of the +x run list of a table."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.zeros((keys.size, 2), dtype=np.uint32)
    count = load().emu_runs(rank, depth, _p(keys), ctypes.c_uint32(keys.size), _p(links), _p(metas),
                            ctypes.c_uint32(K), _p(out))
    return out[:count]



def uniform_neighbors(level, slot):
    """The 27 neighbour slots (offsets (dx, dy, dz) in {-1,0,1}^3, dx fastest)
    of a slot in a single-level 3D tree, by the fused kernel's arithmetic."""
    out = np.zeros(27, dtype=np.uint32)
    load().emu_uniform_neighbors(ctypes.c_uint32(level), ctypes.c_uint32(slot), _p(out))
    return out


# ---- the any-shape (box) kernels; ext = extents of layout dims 0..rank-1 ----
def _ext(ext):
    e = (ctypes.c_int * 3)(*(list(ext) + [1] * (3 - len(ext))))
    return e


def step_box(rank, ext, keys, links, metas, u, dtdx, vel, hs):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    un = np.zeros_like(u)
    load().emu_step_box(rank, _ext(ext), _p(keys), ctypes.c_uint32(keys.size), _p(links), _p(metas), _p(u), _p(un),
                        _p(np.ascontiguousarray(dtdx)), _p(vel), _p(hs))
    return un


def padded_box(rank, ext, h, links, metas, u, nleaves):
    u = np.ascontiguousarray(u, dtype=np.float64)
    pshape = tuple(x + 2 * h for x in ext)
    pad = np.zeros(nleaves * int(np.prod(pshape)))
    load().emu_padded_box(rank, _ext(ext), h, ctypes.c_uint32(nleaves), _p(links), _p(metas), _p(u), _p(pad))
    return pad.reshape((nleaves,) + pshape)


def flags_box(rank, ext, h, kind, keys, links, metas, u, rthr, cthr, mn, mx):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(keys.size, dtype=np.int8)
    load().emu_flags_box(rank, _ext(ext), h, kind, _p(keys), ctypes.c_uint32(keys.size), _p(links), _p(metas), _p(u),
                         ctypes.c_double(rthr), ctypes.c_double(cthr), mn, mx, _p(out))
    return out


def transfer_box(rank, ext, depth, keys, maps, u):
    """The regrid transfer (emit) of the box kernels for a regrid plan."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    maps = np.ascontiguousarray(maps, dtype=np.uint32)
    u = np.ascontiguousarray(u, dtype=np.float64)
    total = int(maps[2].sum())
    ni = int(np.prod(ext))
    keys_out = np.zeros(total, dtype=np.uint64)
    u_out = np.zeros(total * ni)
    load().emu_transfer_box(rank, _ext(ext), depth, _p(keys), ctypes.c_uint32(keys.size), _p(maps), _p(u),
                            _p(keys_out), _p(u_out))
    return keys_out, u_out


def adj_step_box(rank, ext, keys, links, metas, lam_next, u, dtdx, vel):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    lam_next = np.ascontiguousarray(lam_next, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.zeros(3)
    v[: len(vel)] = vel
    lam = np.zeros_like(lam_next)
    vpart = np.zeros((keys.size, 3))
    load().emu_adj_step_box(rank, _ext(ext), _p(keys), ctypes.c_uint32(keys.size), _p(links), _p(metas),
                            _p(lam_next), _p(u), _p(np.ascontiguousarray(dtdx)), _p(v), _p(lam), _p(vpart))
    return lam, vpart


def adj_regrid_box(rank, ext, lam_new, maps, n_old):
    lam_new = np.ascontiguousarray(lam_new, dtype=np.float64)
    maps = np.ascontiguousarray(maps, dtype=np.uint32)
    lam_old = np.zeros(n_old * int(np.prod(ext)))
    load().emu_adj_regrid_box(rank, _ext(ext), _p(lam_new), _p(maps), ctypes.c_uint32(n_old), _p(lam_old))
    return lam_old
