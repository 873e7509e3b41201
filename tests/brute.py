"""Brute-force geometric checkers, independent of the oracle's index formulas.

Everything here is derived from geometry (which leaf contains which point of
the periodic unit square / cube), not from the reference's incremental
neighbour bookkeeping, so agreement with the oracle pins the oracle's
restatement of neighbor.hpp / patch_utils.hpp / ndtree.hpp.

Conventions (the reference's): key = (interleave(x, y[, z]) << 6) | level with
x in bit 0 of each group (lib/libmorton/morton2D.h:113-116, morton3D.h:119-121);
coordinates are integers at resolution 2**depth; patch data is row-major with
the last layout index fastest and physical axis a stored in layout dim
rank-1-a (include/solver/physics_system.hpp:126-135).
"""
from __future__ import annotations

import itertools

import numpy as np

from oracle import HALO_META_DTYPE


def extents(rank, n):
    """Interior cells per layout dim: n (every dim) or a sequence of rank
    extents (static_shape<N, M[, O]>, the last is physical x)."""
    if isinstance(n, (int, np.integer)):
        return (int(n),) * rank
    ext = tuple(int(x) for x in n)
    assert len(ext) == rank
    return ext


def interleave(coords):
    m = 0
    r = len(coords)
    for b in range(21):
        for i, c in enumerate(coords):
            m |= ((c >> b) & 1) << (r * b + i)
    return m


def deinterleave(m, rank):
    c = [0] * rank
    for b in range(21):
        for i in range(rank):
            c[i] |= ((m >> (rank * b + i)) & 1) << b
    return tuple(c)


def make_key(coords, level):
    return (interleave(coords) << 6) | (level & 0x3F)


def decode(key, rank):
    key = int(key)
    return deinterleave(key >> 6, rank), key & 0x3F


def uniform_keys(rank, depth, level):
    """Sorted keys of the uniform tree at `level` (closed form)."""
    s = 1 << (depth - level)
    n = 1 << level
    keys = [make_key(tuple(c * s for c in reversed(idx)), level) for idx in itertools.product(range(n), repeat=rank)]
    return np.array(sorted(keys), dtype=np.uint64)


class Leaves:
    """Set of leaves with point location by geometric containment."""

    def __init__(self, keys, rank, depth):
        self.rank, self.depth = rank, depth
        self.M = 1 << depth
        self.keys = [int(k) for k in keys]
        self.dec = [decode(k, rank) for k in self.keys]
        self.index = {(lv, c): i for i, (c, lv) in enumerate(self.dec)}

    def size(self, i):
        return 1 << (self.depth - self.dec[i][1])

    def find(self, point):
        point = tuple(p % self.M for p in point)
        for lv in range(self.depth, -1, -1):
            s = 1 << (self.depth - lv)
            c = tuple((p // s) * s for p in point)
            i = self.index.get((lv, c))
            if i is not None:
                return i
        raise AssertionError(f"no leaf contains {point}: leaves do not tile the domain")


def neighbor_entry(L: Leaves, i, k):
    """Geometric face neighbour of leaf i in direction k.

    Returns (relation, payload): ('same', j), ('coarser', (j, quadrant)),
    ('finer', [j0..]) with the finer list ordered little-endian over the
    non-normal layout dims.
    """
    rank, M = L.rank, L.M
    c, lv = L.dec[i]
    s = L.size(i)
    dim, pos = k // 2, (k % 2) == 1
    a = rank - 1 - dim
    p = list(c)
    p[a] = (c[a] + s) % M if pos else (c[a] - 1) % M
    j = L.find(p)
    lq = L.dec[j][1]
    if lq == lv:
        return "same", j
    if lq == lv - 1:
        cj = L.dec[j][0]
        quad = [0] * rank
        for jd in range(rank):
            b = rank - 1 - jd
            if jd == dim:
                quad[jd] = 0 if pos else 1
            else:
                quad[jd] = (c[b] - cj[b]) // s
        return "coarser", (j, quad)
    if lq == lv + 1:
        tang = [jd for jd in range(rank) if jd != dim]
        ids = []
        for q in range(1 << (rank - 1)):
            p2 = list(c)
            p2[a] = (c[a] + s) % M if pos else (c[a] - s // 2) % M
            for bit, jd in enumerate(tang):
                b = rank - 1 - jd
                p2[b] = c[b] + ((q >> bit) & 1) * (s // 2)
            jj = L.find(p2)
            assert L.dec[jj][1] == lv + 1, "tree is not 2:1 face balanced"
            ids.append(jj)
        return "finer", ids
    raise AssertionError(f"tree is not 2:1 face balanced (levels {lv} vs {lq})")


def expected_metadata(keys, rank, depth):
    """Neighbour table in the reference's linear-index metadata format."""
    L = Leaves(keys, rank, depth)
    n = len(L.keys)
    out = np.zeros(n * 2 * rank, dtype=HALO_META_DTYPE)
    out["neighbor"] = -1
    out["finer_ids"] = -1
    for i in range(n):
        for k in range(2 * rank):
            e = out[i * 2 * rank + k]
            kind, payload = neighbor_entry(L, i, k)
            if kind == "same":
                e["relation"], e["neighbor"] = 1, payload
            elif kind == "coarser":
                e["relation"], e["neighbor"] = 3, payload[0]
                e["quadrant"][:rank] = payload[1]
            else:
                e["relation"] = 2
                e["finer_ids"][: len(payload)] = payload
    return out


def check_balanced(keys, rank, depth):
    """Raises if the leaves do not tile the domain or violate 2:1 face balance."""
    L = Leaves(keys, rank, depth)
    vol = sum(L.size(i) ** rank for i in range(len(L.keys)))
    assert vol == L.M**rank, "leaves do not cover the domain exactly once"
    for i in range(len(L.keys)):
        for k in range(2 * rank):
            neighbor_entry(L, i, k)


def expected_halo(keys, interior, rank, n, h, depth):
    """Padded patches with face halos filled by geometric containment.

    same -> copy, coarser -> injection of the containing coarse cell, finer ->
    mean of the 2**rank covered fine cells summed in row-major order from 0.0.
    Corner cells (two or more halo coordinates) are left NaN.
    """
    L = Leaves(keys, rank, depth)
    M, D = L.M, depth
    nl = len(L.keys)
    ext = extents(rank, n)
    na = [ext[rank - 1 - a] for a in range(rank)]  # cells per patch along physical axis a
    out = np.full((nl,) + tuple(x + 2 * h for x in ext), np.nan)
    interior = np.asarray(interior).reshape((nl,) + ext)
    for i in range(nl):
        out[(i,) + tuple(slice(h, h + x) for x in ext)] = interior[i]
        c, lv = L.dec[i]
        cs = 1 << (D - lv)
        for idx in itertools.product(*[range(x + 2 * h) for x in ext]):
            halo_dims = [d for d in range(rank) if idx[d] < h or idx[d] >= h + ext[d]]
            if len(halo_dims) != 1:
                continue
            cg = [x - h for x in idx]  # compact layout coordinates
            O = [0] * rank  # ghost origin per physical axis, level-`depth` cell units
            for a in range(rank):
                O[a] = (c[a] * na[a] + cg[rank - 1 - a] * cs) % (M * na[a])
            j = L.find([O[a] // na[a] for a in range(rank)])
            cj, lq = L.dec[j]
            fs = 1 << (D - lq)
            assert abs(lq - lv) <= 1
            loc = [(O[a] - cj[a] * na[a]) // fs for a in range(rank)]
            lay = [loc[rank - 1 - jd] for jd in range(rank)]
            if lq <= lv:
                out[(i,) + tuple(idx)] = interior[j][tuple(lay)]
            else:
                tot = 0.0
                for off in itertools.product(range(2), repeat=rank):
                    tot += interior[j][tuple(lay[d] + off[d] for d in range(rank))]
                out[(i,) + tuple(idx)] = tot / float(1 << rank)
    return out


def expected_regrid(keys, flags, rank, depth):
    """Refine closure + coarsen veto of the reference, restated geometrically.

    Returns (new_sorted_keys, refine_set, coarsen_parents) as Python ints.
    """
    L = Leaves(keys, rank, depth)
    nl = len(L.keys)
    F = 1 << rank
    lev = [L.dec[i][1] for i in range(nl)]
    R = {i for i in range(nl) if flags[i] == 1 and lev[i] < depth}
    changed = True
    while changed:
        changed = False
        for i in list(R):
            for k in range(2 * rank):
                kind, payload = neighbor_entry(L, i, k)
                if kind == "coarser" and payload[0] not in R:
                    R.add(payload[0])
                    changed = True

    def child(c, lv, o):
        s = 1 << (depth - lv - 1)
        return tuple(c[a] + ((o >> a) & 1) * s for a in range(rank)), lv + 1

    parents = {}
    for i in range(nl):
        c, lv = L.dec[i]
        if lv == 0:
            continue
        s = 1 << (depth - lv + 1)
        parents[(tuple((x // s) * s for x in c), lv - 1)] = True
    coarsen = []
    for (pc, pl) in sorted(parents, key=lambda t: make_key(t[0], t[1])):
        kids = [L.index.get((child(pc, pl, o)[1], child(pc, pl, o)[0])) for o in range(F)]
        if any(k is None for k in kids) or any(flags[k] != 2 for k in kids):
            continue
        ps = 1 << (depth - pl)
        veto = False
        for kid in kids:
            kc, kl = L.dec[kid]
            for k in range(2 * rank):
                dim, pos = k // 2, (k % 2) == 1
                a = rank - 1 - dim
                inside = (kc[a] + (1 << (depth - kl)) < pc[a] + ps) if pos else (kc[a] > pc[a])
                if inside:
                    continue
                kind, payload = neighbor_entry(L, kid, k)
                if kind == "finer" or (kind == "same" and payload in R):
                    veto = True
        if not veto:
            coarsen.append((pc, pl, kids))
    if not R and not coarsen:
        return [int(k) for k in L.keys], set(), []
    gone = set(R)
    new = []
    for pc, pl, kids in coarsen:
        gone.update(kids)
        new.append(make_key(pc, pl))
    for i in R:
        c, lv = L.dec[i]
        for o in range(F):
            cc, cl = child(c, lv, o)
            new.append(make_key(cc, cl))
    new.extend(L.keys[i] for i in range(nl) if i not in gone)
    return sorted(new), {L.keys[i] for i in R}, [make_key(pc, pl) for pc, pl, _ in coarsen]


def expected_transfer(old_keys, old_interior, new_keys, rank, n, depth):
    """Interiors after a regrid: copy / injection / row-major mean."""
    old = Leaves(old_keys, rank, depth)
    new = Leaves(new_keys, rank, depth)
    ext = extents(rank, n)
    old_int = np.asarray(old_interior).reshape((len(old.keys),) + ext)
    out = np.zeros((len(new.keys),) + ext)
    h2 = [x // 2 for x in ext]
    for i, key in enumerate(new.keys):
        c, lv = new.dec[i]
        j = old.index.get((lv, c))
        if j is not None:
            out[i] = old_int[j]
            continue
        s = 1 << (depth - lv)
        pc = tuple((x // (2 * s)) * (2 * s) for x in c)
        j = old.index.get((lv - 1, pc)) if lv > 0 else None
        if j is not None:  # injected child
            pos = [(c[a] - pc[a]) // s for a in range(rank)]
            for t in itertools.product(*[range(x) for x in ext]):
                src = tuple(pos[rank - 1 - jd] * h2[jd] + t[jd] // 2 for jd in range(rank))
                out[(i,) + t] = old_int[j][src]
            continue
        kids = {}
        for o in range(1 << rank):
            cc = tuple(c[a] + ((o >> a) & 1) * (s // 2) for a in range(rank))
            kids[tuple((o >> (rank - 1 - jd)) & 1 for jd in range(rank))] = old.index[(lv + 1, cc)]
        for t in itertools.product(*[range(x) for x in ext]):
            kid = kids[tuple(t[jd] // h2[jd] for jd in range(rank))]
            base = [(t[jd] % h2[jd]) * 2 for jd in range(rank)]
            tot = 0.0
            for off in itertools.product(range(2), repeat=rank):
                tot += old_int[kid][tuple(base[d] + off[d] for d in range(rank))]
            out[(i,) + t] = tot / float(1 << rank)
    return out


def dense_from_uniform(keys, interior, rank, n, depth):
    """Assemble a uniform tree's interiors into one dense periodic array
    (layout order: last axis = physical x)."""
    L = Leaves(keys, rank, depth)
    lv = L.dec[0][1]
    assert all(d[1] == lv for d in L.dec)
    s = 1 << (depth - lv)
    nb = 1 << lv
    ext = extents(rank, n)
    out = np.zeros(tuple(nb * x for x in ext))
    interior = np.asarray(interior).reshape((len(L.keys),) + ext)
    for i, (c, _) in enumerate(L.dec):
        blk = [c[rank - 1 - jd] // s for jd in range(rank)]
        out[tuple(slice(b * x, (b + 1) * x) for b, x in zip(blk, ext))] = interior[i]
    return out


def split_uniform(keys, dense, rank, n, depth):
    L = Leaves(keys, rank, depth)
    s = 1 << (depth - L.dec[0][1])
    ext = extents(rank, n)
    out = np.zeros((len(L.keys),) + ext)
    for i, (c, _) in enumerate(L.dec):
        blk = [c[rank - 1 - jd] // s for jd in range(rank)]
        out[i] = dense[tuple(slice(b * x, (b + 1) * x) for b, x in zip(blk, ext))]
    return out


def dense_upwind_step(u, vel, dt, dx):
    """Textbook periodic first-order Rusanov/upwind sweep on a dense array,
    written in the same operation order as AdvectionPhysics::rusanovFlux and
    amr_solver::time_step_cpu.  Physical axis d is array axis rank-1-d."""
    rank = u.ndim
    total = np.zeros_like(u)
    for d in range(rank):
        ax = rank - 1 - d
        ul = np.roll(u, 1, axis=ax)
        ur = np.roll(u, -1, axis=ax)
        v = vel[d]
        s = abs(v)
        fL = 0.5 * (ul * v + u * v) - 0.5 * s * (u - ul)
        fR = 0.5 * (u * v + ur * v) - 0.5 * s * (ur - u)
        dtdx = dt / dx[d]
        total -= dtdx * (fR - fL)
    return u + total


# ---------------------------------------------------------------------------
# Refinement criteria, restated from their definitions on the padded patches
# that expected_halo produces (face halos filled geometrically, corners NaN).
# ---------------------------------------------------------------------------
def criterion_values(padded, rank, n, h, kind, start=0.0):
    """Per-leaf criterion value.

    kind 0 (maximum value): max over the interior and face-halo cells, and
    `start` -- the reference's CPU criteria start from 0
    (examples/fvm_solver_advection.e.cpp:79, tests/advection_equation_amr.t.cpp:159);
    corner cells are never written by a halo exchange and are not read.
    kind 1 (undivided gradient): max over interior cells c and layout dims d
    of max(|u[c+e_d] - u[c]|, |u[c] - u[c-e_d]|), and `start`.
    """
    pad = np.asarray(padded, dtype=np.float64)
    nl = pad.shape[0]
    ext = extents(rank, n)
    out = np.full(nl, start, dtype=np.float64)
    inner = [slice(h, h + x) for x in ext]
    if kind == 0:
        for i in range(nl):
            p = pad[i]
            vals = [p[tuple(inner)].ravel()]
            for d in range(rank):
                for band in (slice(0, h), slice(h + ext[d], ext[d] + 2 * h)):
                    sl = tuple(band if e == d else inner[e] for e in range(rank))
                    vals.append(p[sl].ravel())
            out[i] = max(start, float(np.max(np.concatenate(vals))))
        return out
    for i in range(nl):
        p = pad[i]
        u = p[tuple(inner)]
        g = start
        for d in range(rank):
            up = p[tuple(slice(h + 1, h + ext[d] + 1) if e == d else inner[e] for e in range(rank))]
            um = p[tuple(slice(h - 1, h + ext[d] - 1) if e == d else inner[e] for e in range(rank))]
            g = max(g, float(np.max(np.abs(up - u))), float(np.max(np.abs(u - um))))
        out[i] = g
    return out


def decide(values, keys, refine_thr, coarsen_thr, min_level, max_level):
    """Refine if value > refine_thr and level < max_level, else Coarsen if
    value < coarsen_thr and level > min_level, else Stable
    (examples/fvm_solver_advection.e.cpp:84-89, src/cuda/fvm_refinement_criterion.cu:53-66)."""
    levels = np.asarray(keys, dtype=np.uint64) & np.uint64(63)
    flags = np.zeros(len(values), dtype=np.int8)
    for i, (v, lv) in enumerate(zip(values, levels)):
        if v > refine_thr and int(lv) < max_level:
            flags[i] = 1
        elif v < coarsen_thr and int(lv) > min_level:
            flags[i] = 2
    return flags


def expected_flags(keys, padded, rank, n, h, kind, refine_thr, coarsen_thr, min_level, max_level):
    return decide(criterion_values(padded, rank, n, h, kind), keys, refine_thr, coarsen_thr, min_level, max_level)
