"""The HIP kernels' per-item bodies, executed on the CPU, against the oracle.

These run without a GPU: they pin the device-side logic (geometric neighbour
table, ghost gather, upwind update, criteria, regrid expand/contract) to the
oracle bit for bit; the GPU tests then add the launch configuration and
host sequencing on real hardware.
"""
import numpy as np
import pytest

import brute
import emulation as emu
from oracle import OracleTree
from test_oracle_tree import random_flags


def emulated_sequence(rank, n, h, depth, start, rounds, seed, vel):
    rng = np.random.default_rng(seed)
    cap = (1 << (rank * depth)) + 64
    o = OracleTree(rank, n, h, depth, 2 * cap + 64)
    o.set_params(vel, 0.4)
    for _ in range(start):
        o.refine_all()
    keys = o.keys()
    u = rng.standard_normal(o.size() * n**rank)
    o.set_interior(u)
    o.halo_exchange()
    for rd in range(rounds):
        links, metas, err = emu.build_table(rank, depth, keys)
        assert err == 0
        assert emu.to_reference_metadata(rank, links, metas).tobytes() == o.metadata().tobytes(), f"round {rd}: table"
        # padded view vs the oracle's halo exchange (corners excluded)
        pad = emu.padded(rank, n, h, links, metas, u, keys.size)
        eh = brute.expected_halo(keys, u, rank, n, h, depth)
        m = ~np.isnan(eh)
        assert np.array_equal(pad[m].view(np.uint64), o.get_padded()[m].view(np.uint64)), f"round {rd}: halo"
        # criteria
        for kind, thr in ((0, (0.5, -1.0)), (1, (1.0, 0.2))):
            fe = emu.flags(rank, n, h, kind, keys, links, metas, u, thr[0], thr[1], 1, depth)
            fo = o.criterion(kind, thr[0], thr[1], 1, depth)
            np.testing.assert_array_equal(fe, fo, err_msg=f"round {rd}: criterion {kind}")
        # regrid
        flags = random_flags(rng, keys, rank, depth)
        max_level = int(max(int(k) & 63 for k in keys))
        changed, k2, u2, summ = emu.regrid(rank, n, depth, keys, flags, links, metas, u, max_level, cap)
        assert summ[3] == 0
        co = o.reconstruct(flags)
        assert changed == co, f"round {rd}: changed"
        if changed:
            keys, u = k2, u2
        np.testing.assert_array_equal(keys, o.keys(), err_msg=f"round {rd}: keys")
        assert np.array_equal(u.view(np.uint64), o.get_interior().reshape(-1).view(np.uint64)), f"round {rd}: transfer"
        o.halo_exchange()
        # two steps
        links, metas, err = emu.build_table(rank, depth, keys)
        max_level = int(max(int(k) & 63 for k in keys))
        for _ in range(2):
            dt = o.compute_dt()
            dtdx, v, hs = emu.step_params(rank, n, depth, max_level, vel, dt)
            u = emu.step(rank, n, keys, links, metas, u, dtdx, v, hs)
            o.step(dt)
        assert np.array_equal(u.view(np.uint64), o.get_interior().reshape(-1).view(np.uint64)), f"round {rd}: step"


@pytest.mark.parametrize(
    "rank,n,h,depth,start,rounds,seed",
    [(2, 4, 1, 5, 2, 8, 1), (2, 4, 2, 5, 2, 6, 2), (2, 16, 1, 5, 2, 5, 3), (3, 4, 1, 4, 1, 5, 4), (3, 8, 1, 3, 1, 3, 5)],
)
def test_emulated_kernels_match_oracle(rank, n, h, depth, start, rounds, seed):
    emulated_sequence(rank, n, h, depth, start, rounds, seed, (1.0, 0.5, 0.25)[:rank])


def test_emulated_table_detects_unbalanced_tree():
    # a level-1 leaf next to level-3 leaves violates 2:1
    keys = [brute.make_key((0, 0), 1)]
    s = 1 << 4  # depth 5, level 1 -> size 16; level 3 -> size 4
    for x in range(0, 32, 16):
        for y in range(0, 32, 16):
            if (x, y) == (0, 0):
                continue
            if (x, y) == (16, 0):
                for xx in range(16, 32, 4):
                    for yy in range(0, 16, 4):
                        keys.append(brute.make_key((xx, yy), 3))
            else:
                keys.append(brute.make_key((x, y), 1))
    keys = np.array(sorted(keys), dtype=np.uint64)
    _, _, err = emu.build_table(2, 5, keys)
    assert err == 1


@pytest.mark.parametrize("rank,n,depth,start,rounds,seed", [(2, 16, 5, 2, 5, 21), (3, 8, 4, 1, 4, 22), (2, 4, 6, 3, 6, 23)])
def test_run_lists_partition_the_leaves(rank, n, depth, start, rounds, seed):
    """The +x run lists used by the run-streaming step kernels: every leaf in
    exactly one run, consecutive leaves of a run are same-level +x
    neighbours, runs hold at most K leaves and start where the x index at the
    leaf's level is a multiple of K or the -x neighbour is not same-level."""
    rng = np.random.default_rng(seed)
    o = OracleTree(rank, n, 1, depth, 2 * ((1 << (rank * depth)) + 64))
    for _ in range(start):
        o.refine_all()
    for rd in range(rounds + 1):
        keys = o.keys()
        links, metas, err = emu.build_table(rank, depth, keys)
        assert err == 0
        mx, px = 2 * (rank - 1), 2 * (rank - 1) + 1
        for K in (1, 4, 8):
            seen = np.zeros(keys.size, dtype=np.int64)
            for head, ln in emu.runs(rank, depth, keys, links, metas, K):
                assert 1 <= ln <= K
                j = int(head)
                key = int(keys[j])
                level = key & 63
                x = brute.deinterleave(key >> 6, rank)[0] >> (depth - level)
                assert x % K == 0 or (metas[j * 2 * rank + mx] & 3) != 1
                for t in range(ln):
                    seen[j] += 1
                    if t + 1 < ln:
                        e = j * 2 * rank + px
                        assert (metas[e] & 3) == 1, "run continues through a same-level +x neighbour"
                        j = int(links["ids"][e, 0])
            assert np.all(seen == 1), f"round {rd} K {K}: leaves not covered exactly once"
        if rd < rounds:
            o.reconstruct(random_flags(rng, keys, rank, depth))



@pytest.mark.parametrize("level", [0, 1, 2, 3, 5])
def test_uniform_neighbour_arithmetic(level):
    """The fused two-step kernel finds neighbours without the table: on a
    single-level 3D tree, slot i holds the leaf of Morton index i, and the
    slot at offset (dx, dy, dz) is a dilated-integer step of i.  Checked
    against the keys' decoded coordinates (periodic wrap) and against the
    device neighbour table's face entries."""
    depth = 5
    o = OracleTree(3, 8, 1, depth, 2 * (1 << (3 * level)) + 64)
    for _ in range(level):
        o.refine_all()
    keys = o.keys()
    assert keys.size == 1 << (3 * level)
    coords = np.array([brute.deinterleave(int(k) >> 6, 3) for k in keys]) >> (depth - level)
    slot_of = {tuple(c): i for i, c in enumerate(coords)}
    for i, c in enumerate(coords):
        assert brute.interleave([int(v) for v in c]) == i, "slot == Morton index at the level"
    links, metas, err = emu.build_table(3, depth, keys)
    assert err == 0
    rng = np.random.default_rng(level)
    sample = range(keys.size) if keys.size <= 512 else rng.choice(keys.size, 300, replace=False)
    ext = 1 << level
    for i in sample:
        got = emu.uniform_neighbors(level, int(i))
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    want = slot_of[tuple((coords[i] + (dx, dy, dz)) % ext)]
                    assert got[9 * (dz + 1) + 3 * (dy + 1) + (dx + 1)] == want, (i, dx, dy, dz)
        if level >= 1:
            # directions: layout dim 0 = z, 1 = y, 2 = x; negative if even
            for dirn, (dz, dy, dx) in enumerate([(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]):
                e = int(i) * 6 + dirn
                assert metas[e] & 3 == 1
                assert links["ids"][e, 0] == got[9 * (dz + 1) + 3 * (dy + 1) + (dx + 1)]


# ---- the any-shape (box) kernels, for non-square and untabled shapes ------
BOX_CASES = [
    (2, (16, 8), 1, 5, 2, 4, 31),
    (2, (8, 16), 1, 5, 2, 4, 32),
    (2, (12, 8), 2, 5, 2, 4, 33),
    (3, (4, 6, 8), 1, 4, 1, 3, 34),
    (2, (6, 6), 1, 5, 2, 3, 35),
    (2, (4, 4), 2, 5, 2, 3, 36),
    (3, (4, 4, 4), 1, 4, 1, 3, 37),
]


@pytest.mark.parametrize("rank,ext,h,depth,start,rounds,seed", BOX_CASES)
def test_emulated_box_kernels_match_oracle(rank, ext, h, depth, start, rounds, seed):
    """Padded view, both criteria, regrid transfer and the upwind step of the
    any-shape kernels against the oracle (and the halo against brute force),
    bit for bit, over regrid rounds on mixed-level trees."""
    rng = np.random.default_rng(seed)
    vel = (1.0, -0.5, 0.25)[:rank]
    cap = (1 << (rank * depth)) + 64
    o = OracleTree(rank, ext, h, depth, 2 * cap + 64)
    o.set_params(vel, 0.4)
    for _ in range(start):
        o.refine_all()
    keys = o.keys()
    u = rng.standard_normal(o.size() * o.ni)
    o.set_interior(u)
    o.halo_exchange()
    for rd in range(rounds):
        links, metas, err = emu.build_table(rank, depth, keys)
        assert err == 0
        pad = emu.padded_box(rank, ext, h, links, metas, u, keys.size)
        eh = brute.expected_halo(keys, u, rank, ext, h, depth)
        m = ~np.isnan(eh)
        assert np.array_equal(pad[m].view(np.uint64), eh[m].view(np.uint64)), f"round {rd}: halo vs brute"
        assert np.array_equal(pad[m].view(np.uint64), o.get_padded()[m].view(np.uint64)), f"round {rd}: halo"
        assert np.all(pad[~m] == 0.0), "corners are 0"
        for kind, thr in ((0, (0.5, -1.0)), (1, (1.0, 0.2))):
            fe = emu.flags_box(rank, ext, h, kind, keys, links, metas, u, thr[0], thr[1], 1, depth)
            fo = o.criterion(kind, thr[0], thr[1], 1, depth)
            np.testing.assert_array_equal(fe, fo, err_msg=f"round {rd}: criterion {kind}")
        flags = random_flags(rng, keys, rank, depth)
        max_level = int(max(int(k) & 63 for k in keys))
        # the regrid plan does not depend on the patch shape
        changed, _, _, summ, maps = emu.regrid(rank, 4, depth, keys, flags, links, metas,
                                               np.zeros(keys.size * 4**rank), max_level, cap, with_maps=True)
        assert summ[3] == 0
        co = o.reconstruct(flags)
        assert changed == co
        if changed:
            keys, u = emu.transfer_box(rank, ext, depth, keys, maps, u)
        np.testing.assert_array_equal(keys, o.keys(), err_msg=f"round {rd}: keys")
        assert np.array_equal(u.view(np.uint64), o.get_interior().reshape(-1).view(np.uint64)), f"round {rd}: transfer"
        o.halo_exchange()
        links, metas, err = emu.build_table(rank, depth, keys)
        max_level = int(max(int(k) & 63 for k in keys))
        for _ in range(2):
            dt = o.compute_dt()
            dtdx, v, hs = emu.step_params(rank, ext, depth, max_level, vel, dt)
            u = emu.step_box(rank, ext, keys, links, metas, u, dtdx, v, hs)
            o.step(dt)
        assert np.array_equal(u.view(np.uint64), o.get_interior().reshape(-1).view(np.uint64)), f"round {rd}: step"


@pytest.mark.parametrize("rank,n,depth,seed", [(2, 4, 5, 41), (2, 16, 5, 42), (3, 4, 4, 43)])
def test_emulated_box_adjoint_equals_fixed_shape(rank, n, depth, seed):
    """On a cubic shape the any-shape adjoint step and regrid transpose give
    the fixed-shape kernels' bits (lambda and the velocity partials)."""
    rng = np.random.default_rng(seed)
    o = OracleTree(rank, n, 1, depth, 4 * (1 << (rank * depth)) + 64)
    o.refine_all()
    o.refine_all()
    for _ in range(2):
        o.reconstruct(random_flags(rng, o.keys(), rank, depth))
    keys = o.keys()
    links, metas, _ = emu.build_table(rank, depth, keys)
    u = rng.standard_normal(keys.size * n**rank)
    lam_next = rng.standard_normal(keys.size * n**rank)
    vel = (0.7, -0.4, 0.3)[:rank]
    max_level = int(max(int(k) & 63 for k in keys))
    dtdx, _, _ = emu.step_params(rank, n, depth, max_level, vel, 1e-3)
    a, va = emu.adj_step(rank, n, keys, links, metas, lam_next, u, dtdx, vel)
    b, vb = emu.adj_step_box(rank, (n,) * rank, keys, links, metas, lam_next, u, dtdx, vel)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert np.array_equal(va.view(np.uint64), vb.view(np.uint64))
    flags = random_flags(rng, keys, rank, depth)
    changed, k2, _, _, maps = emu.regrid(rank, n, depth, keys, flags, links, metas, u, max_level,
                                         4 * (1 << (rank * depth)), with_maps=True)
    assert changed
    lam_new = rng.standard_normal(k2.size * n**rank)
    a = emu.adj_regrid(rank, n, lam_new, maps, keys.size)
    b = emu.adj_regrid_box(rank, (n,) * rank, lam_new, maps, keys.size)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("rank,ext,depth,seed", [(2, (16, 8), 5, 51), (2, (8, 12), 5, 52), (3, (4, 6, 8), 4, 53)])
def test_emulated_box_adjoint_transposes(rank, ext, depth, seed):
    """Non-square patches: the any-shape adjoint step is the transpose of the
    step (with its ghost rules) and the adjoint regrid the transpose of the
    transfer: <lam, A u> = <A^T lam, u> to rounding, and the velocity
    partials sum to the derivative of <lam, A u> along v (the step is linear
    in v for fixed upwind signs): central difference of the emulated step."""
    rng = np.random.default_rng(seed)
    o = OracleTree(rank, 4, 1, depth, 4 * (1 << (rank * depth)) + 64)
    o.refine_all()
    o.refine_all()
    for _ in range(2):
        o.reconstruct(random_flags(rng, o.keys(), rank, depth))
    keys = o.keys()
    ni = int(np.prod(ext))
    links, metas, _ = emu.build_table(rank, depth, keys)
    u = rng.standard_normal(keys.size * ni)
    lam = rng.standard_normal(keys.size * ni)
    vel = np.array((0.7, -0.4, 0.3)[:rank])
    max_level = int(max(int(k) & 63 for k in keys))

    def fwd(v, x):
        dtdx, vv, hs = emu.step_params(rank, ext, depth, max_level, v, 1e-3)
        return emu.step_box(rank, ext, keys, links, metas, x, dtdx, vv, hs)

    dtdx, _, _ = emu.step_params(rank, ext, depth, max_level, vel, 1e-3)
    at, vpart = emu.adj_step_box(rank, ext, keys, links, metas, lam, u, dtdx, vel)
    lhs, rhs = float(lam @ fwd(vel, u)), float(at @ u)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    for d in range(rank):
        e = np.zeros(rank)
        e[d] = 1e-4
        fd = (float(lam @ fwd(vel + e, u)) - float(lam @ fwd(vel - e, u))) / 2e-4
        g = float(vpart[:, d].sum())
        assert abs(fd - g) <= 1e-8 * max(abs(g), 1.0), (d, fd, g)
    flags = random_flags(rng, keys, rank, depth)
    changed, _, _, _, maps = emu.regrid(rank, 4, depth, keys, flags, links, metas, np.zeros(keys.size * 4**rank),
                                        max_level, 4 * (1 << (rank * depth)), with_maps=True)
    assert changed
    k2, u2 = emu.transfer_box(rank, ext, depth, keys, maps, u)
    lam_new = rng.standard_normal(k2.size * ni)
    back = emu.adj_regrid_box(rank, ext, lam_new, maps, keys.size)
    lhs, rhs = float(lam_new @ u2), float(back @ u)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
