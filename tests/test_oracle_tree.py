"""Oracle regrid, neighbour tables, transfer and halo exchange against the
brute-force geometric checkers (tests/brute.py)."""
import numpy as np
import pytest

import brute
from oracle import OracleTree


def leaf_centres(keys, rank, depth):
    out = []
    for k in keys:
        c, lv = brute.decode(k, rank)
        s = 1 << (depth - lv)
        out.append([(x + 0.5 * s) / (1 << depth) for x in c])
    return np.array(out)


def random_flags(rng, keys, rank, depth):
    """Spatially correlated flags so that whole sibling families coarsen."""
    n = len(keys)
    if rng.random() < 0.25:
        return no_root_coarsen(rng.choice(np.array([0, 1, 2], dtype=np.int8), size=n, p=[0.5, 0.2, 0.3]), keys)
    cen = leaf_centres(keys, rank, depth)
    c0 = rng.random(rank)
    r = 0.1 + 0.2 * rng.random()
    d = np.abs(cen - c0)
    d = np.minimum(d, 1.0 - d)
    dist = np.sqrt((d**2).sum(axis=1))
    flags = np.zeros(n, dtype=np.int8)
    flags[(dist < r) & (rng.random(n) < 0.7)] = 1
    flags[(dist > 1.5 * r) & (rng.random(n) < 0.9)] = 2
    return no_root_coarsen(flags, keys)


def no_root_coarsen(flags, keys):
    # the reference aborts when the root's children recombine; criteria keep
    # min_level >= 1 (tests/advection_equation_amr.t.cpp:167)
    flags = flags.copy()
    flags[((np.asarray(keys, dtype=np.uint64) & np.uint64(63)) <= 1) & (flags == 2)] = 0
    return flags


def run_sequence(rank, n, h, depth, start_level, rounds, seed, check_halo=True):
    rng = np.random.default_rng(seed)
    cap = 4 * (1 << (rank * depth)) + 64
    t = OracleTree(rank, n, h, depth, cap)
    for _ in range(start_level):
        t.refine_all()
    changed_rounds = 0
    coarsened = 0
    for rd in range(rounds):
        keys = t.keys()
        vals = rng.standard_normal(t.size() * t.ni)
        t.set_interior(vals)
        flags = random_flags(rng, keys, rank, depth)
        exp_keys, exp_r, exp_c = brute.expected_regrid(keys, flags, rank, depth)
        changed = t.reconstruct(flags)
        new_keys = t.keys()
        assert [int(k) for k in new_keys] == exp_keys, f"round {rd}: keys"
        r_set, c_set = t.last_sets()
        assert {int(x) for x in r_set} == exp_r, f"round {rd}: refine closure"
        assert [int(x) for x in c_set] == exp_c, f"round {rd}: coarsen set"
        brute.check_balanced(new_keys, rank, depth)
        meta = t.metadata()
        assert meta.tobytes() == brute.expected_metadata(new_keys, rank, depth).tobytes(), f"round {rd}: table"
        got = t.get_interior()
        exp = brute.expected_transfer(keys, vals, new_keys, rank, n, depth)
        assert np.array_equal(got.view(np.uint64), exp.view(np.uint64)), f"round {rd}: transfer"
        if check_halo:
            t.halo_exchange()
            pad = t.get_padded()
            eh = brute.expected_halo(new_keys, got, rank, n, h, depth)
            m = ~np.isnan(eh)
            assert np.array_equal(pad[m].view(np.uint64), eh[m].view(np.uint64)), f"round {rd}: halo"
        changed_rounds += int(changed)
        coarsened += len(exp_c)
    return changed_rounds, coarsened


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_regrid_sequences_2d(seed):
    ch, co = run_sequence(2, 4, 1, 5, 2, 12, seed)
    assert ch >= 6 and co >= 1


def test_regrid_sequences_2d_halo2():
    run_sequence(2, 4, 2, 5, 2, 8, 11)


@pytest.mark.parametrize("seed", [4, 5])
def test_regrid_sequences_3d(seed):
    ch, co = run_sequence(3, 4, 1, 4, 1, 6, seed, check_halo=(seed == 4))
    assert ch >= 3


# Non-square patches (static_shape<N, M[, O]>): extents per layout dim, the
# last is physical x; both orientations, a halo-2 shape and two 3D boxes
# (halo 1 and 2).
@pytest.mark.parametrize(
    "rank,ext,h,depth,start,seed",
    [(2, (16, 8), 1, 5, 2, 21), (2, (8, 16), 1, 5, 2, 22), (2, (12, 8), 2, 5, 2, 23), (3, (4, 6, 8), 1, 4, 1, 24),
     (3, (6, 4, 4), 2, 4, 2, 25)],
)
def test_regrid_sequences_non_square(rank, ext, h, depth, start, seed):
    ch, _ = run_sequence(rank, ext, h, depth, start, 6 if rank == 2 else 4, seed)
    assert ch >= 2


@pytest.mark.parametrize("ext,h", [((16, 8), 1), ((8, 16), 1), ((12, 8), 2)])
def test_dense_uniform_step_non_square_bitwise(ext, h):
    """Uniform tree of non-square patches = one dense periodic array of
    (2^level n_0) x (2^level n_1) cells with dx per axis: the textbook sweep,
    bit for bit."""
    rank, depth, lv = 2, 7, 3
    t = OracleTree(rank, ext, h, depth, 2 * 4**lv + 8)
    for _ in range(lv):
        t.refine_all()
    rng = np.random.default_rng(sum(ext))
    t.set_interior(rng.random(t.size() * t.ni))
    t.halo_exchange()
    t.set_params((1.0, -0.5), 0.4)
    dx = (1.0 / (8 * ext[1]), 1.0 / (8 * ext[0]))  # physical x is layout dim 1
    dt = t.compute_dt()
    assert dt == 0.4 * min(dx[0] / 1.0, dx[1] / 0.5)
    keys = t.keys()
    dense = brute.dense_from_uniform(keys, t.get_interior(), rank, ext, depth)
    assert dense.shape == (8 * ext[0], 8 * ext[1])
    for _ in range(4):
        dense = brute.dense_upwind_step(dense, (1.0, -0.5), dt, dx)
        t.step(dt)
    exp = brute.split_uniform(keys, dense, rank, ext, depth)
    assert np.array_equal(t.get_interior().view(np.uint64), exp.view(np.uint64))


def test_finer_ghost_rule_for_non_square_patches():
    """The finer-neighbour ghost of a non-square patch: the fine cells lie at
    (2 v) mod n(d) along every dim d (the reference's device rule,
    src/cuda/halo_exchange.cu:174-176), which is where geometry puts them
    (expected_halo, checked in the regrid sequences above).  The reference's
    CPU rule reduces every dim mod the face normal's extent
    (include/ndtree/patch_utils.hpp:358): for an 8 x 16 patch (x = layout
    dim 1) its -x/+x ghosts along y would start at padded row up to
    (2*7) % 16 + h = 15, past the 8 + 2h = 10 rows of the fine patch, so
    that rule has no result to reproduce there; for square patches the two
    rules coincide."""
    h = 1
    for ext in ((8, 16), (16, 8), (12, 8)):
        for dim in range(2):  # face normal
            tang = 1 - dim
            device = [((v * 2) % ext[tang]) + h for v in range(ext[tang])]
            cpu = [((v * 2) % ext[dim]) + h for v in range(ext[tang])]
            # the device rule stays inside the fine patch's interior
            assert all(h <= b and b + 1 < h + ext[tang] for b in device)
            if ext[dim] > ext[tang]:
                assert cpu != device
                if ext[dim] >= 2 * ext[tang]:
                    # past the fine patch's padded rows
                    assert max(cpu) + 1 >= ext[tang] + 2 * h


def test_dense_uniform_step_2d_bitwise():
    # C1 shape: uniform depth-4 quadtree of 16x16 patches, velocity (1, 0.5).
    rank, n, depth, lv = 2, 16, 7, 4
    t = OracleTree(rank, n, 1, depth, 2 * 4**lv + 8)
    for _ in range(lv):
        t.refine_all()
    cen = t.cell_centers()
    u0 = np.exp(-((cen[..., 0] - 0.3) ** 2 + (cen[..., 1] - 0.3) ** 2) / (2 * 0.1**2))
    t.set_interior(u0)
    t.halo_exchange()
    t.set_params((1.0, 0.5), 0.4)
    dt = t.compute_dt()
    assert dt == 0.4 * (1.0 / 256.0)
    keys = t.keys()
    dense = brute.dense_from_uniform(keys, t.get_interior(), rank, n, depth)
    dx = (1.0 / 256.0, 1.0 / 256.0)
    for _ in range(5):
        dense = brute.dense_upwind_step(dense, (1.0, 0.5), dt, dx)
        t.step(dt)
    got = t.get_interior()
    exp = brute.split_uniform(keys, dense, rank, n, depth)
    assert np.array_equal(got.view(np.uint64), exp.view(np.uint64))


def test_dense_uniform_step_3d_bitwise():
    rank, n, depth, lv = 3, 8, 5, 2
    t = OracleTree(rank, n, 1, depth, 2 * 8**lv + 8)
    for _ in range(lv):
        t.refine_all()
    rng = np.random.default_rng(3)
    t.set_interior(rng.random(t.size() * n**3))
    t.halo_exchange()
    t.set_params((1.0, 0.5, 0.25), 0.3)
    dt = t.compute_dt()
    assert dt == 0.3 * (1.0 / 32.0)
    keys = t.keys()
    dense = brute.dense_from_uniform(keys, t.get_interior(), rank, n, depth)
    for _ in range(3):
        dense = brute.dense_upwind_step(dense, (1.0, 0.5, 0.25), dt, (1 / 32, 1 / 32, 1 / 32))
        t.step(dt)
    exp = brute.split_uniform(keys, dense, rank, n, depth)
    assert np.array_equal(t.get_interior().view(np.uint64), exp.view(np.uint64))


def test_mass_conserved_uniform():
    t = OracleTree(2, 16, 1, 7, 200)
    for _ in range(3):
        t.refine_all()
    rng = np.random.default_rng(9)
    t.set_interior(rng.random(t.size() * 256))
    t.halo_exchange()
    t.set_params((1.0, 0.5), 0.4)
    m0 = t.get_interior().sum()
    t.advance_batch(50)
    assert abs(t.get_interior().sum() - m0) <= 1e-12 * abs(m0)


def test_advance_batch_clamps_to_remaining():
    t = OracleTree(2, 8, 1, 7, 40)
    t.refine_all()
    t.set_interior(np.ones(t.size() * 64))
    t.halo_exchange()
    t.set_params((1.0, 0.5), 0.4)
    dt = t.compute_dt()
    total, ex = t.advance_batch(10, 2.5 * dt)
    assert ex == 3 and total == dt + dt + (2.5 * dt - dt - dt)


def gaussian_ic(cen, x0=(0.2, 0.2), w=0.005):
    r2 = np.zeros(cen.shape[:-1])
    for d in range(cen.shape[-1]):
        r2 += (cen[..., d] - x0[d]) ** 2
    return np.exp(-r2 / w)


@pytest.mark.parametrize("n", [10, 50])
def test_reference_pulse_center_movement(n):
    """The reference's PulseCenterMovement (tests/advection_equation_amr.t.cpp:50-130)
    on a single periodic root patch with halo width 2."""
    t = OracleTree(2, n, 2, 7, 2000)
    t.set_params((1.0, 0.5), 0.4)
    t.set_interior(gaussian_ic(t.cell_centers()))
    time, target = 0.0, 0.5
    while time < target:
        dt = t.advance()
        if time + dt > target:
            dt = target - time
        t.halo_exchange()
        time += dt
    u = t.get_interior()[0]
    cen = t.cell_centers()[0].reshape(n, n, 2)
    iy, ix = np.unravel_index(np.argmax(u), u.shape)
    pos = cen[iy, ix]
    for d, v in enumerate((1.0, 0.5)):
        assert abs(pos[d] - (0.2 + v * target)) <= 1.0 / n
