"""HIP path vs the CPU oracle, through the C ABI.

Bar: keys, neighbour tables and flags bit-exact; FP64 interiors and halos
bit-exact (both sides compute without contraction); the stated tolerance of
1e-12 relative is asserted as well so a failure report shows the size of any
deviation.
"""
import sys

import numpy as np
import pytest

import brute
from amrlib import load_package
from oracle import OracleTree
from test_oracle_tree import gaussian_ic, random_flags

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def amr():
    return load_package()


def pair(amr, rank, n, h, depth, cap, vel, cfl):
    g = amr.NdTree(rank, depth, n, h, cap)
    s = amr.AmrSolver(g, vel, cfl)
    # the oracle appends children before compacting (ndtree.hpp:727-778):
    # twice the final capacity covers the transient
    o = OracleTree(rank, n, h, depth, 2 * cap + 64)
    o.set_params(vel, cfl, threads=8)
    return g, s, o


def assert_fields(got, exp, what):
    got = np.asarray(got)
    exp = np.asarray(exp)
    assert got.shape == exp.shape, what
    scale = max(1.0, float(np.max(np.abs(exp)))) if exp.size else 1.0
    err = float(np.max(np.abs(got - exp))) if exp.size else 0.0
    assert err <= REL_TOL * scale, f"{what}: max abs err {err}"
    assert np.array_equal(got.view(np.uint64), exp.view(np.uint64)), f"{what}: not bit-exact (max err {err})"


def assert_topology(g, o, what):
    np.testing.assert_array_equal(g.keys(), o.keys(), err_msg=f"{what}: keys")
    gm, om = g.neighbor_metadata(), o.metadata()
    assert gm.tobytes() == om.tobytes(), f"{what}: neighbour table"


def test_c1_uniform_2d_100_steps(amr):
    say("case: C1 2D uniform depth 4, 16^2, 100 steps")
    g, s, o = pair(amr, 2, 16, 1, 7, 256, (1.0, 0.5), 0.4)
    for _ in range(4):
        g.refine_all()
        o.refine_all()
    assert_topology(g, o, "C1")
    cen = o.cell_centers()
    assert np.array_equal(g.cell_centers().view(np.uint64), cen.view(np.uint64))
    u0 = np.exp(-((cen[..., 0] - 0.3) ** 2 + (cen[..., 1] - 0.3) ** 2) / (2 * 0.1**2))
    g.set_interior(u0)
    o.set_interior(u0)
    o.halo_exchange()
    assert s.compute_dt() == o.compute_dt() == 0.4 / 256.0
    s.advance_batch_async(100)
    tot, ex = s.finish_advance_batch()
    otot, oex = o.advance_batch(100)
    assert ex == oex == 100 and tot == otot
    assert_fields(g.get_interior(), o.get_interior(), "C1 after 100 steps")
    m0 = u0.sum()
    assert abs(g.get_interior().sum() - m0) <= 1e-12 * m0


def test_padded_view_matches_halo_exchange(amr):
    say("case: padded view 2D 16^2 depth 3")
    g, s, o = pair(amr, 2, 16, 1, 7, 256, (1.0, 0.5), 0.4)
    for _ in range(3):
        g.refine_all()
        o.refine_all()
    rng = np.random.default_rng(5)
    v = rng.standard_normal(g.size() * 256)
    g.set_interior(v)
    o.set_interior(v)
    o.halo_exchange()
    gp, op = g.get_patches(), o.get_padded()
    mask = np.ones(gp.shape[1:], dtype=bool)
    mask[0, 0] = mask[0, -1] = mask[-1, 0] = mask[-1, -1] = False  # corners are never written
    assert_fields(gp[:, mask], op[:, mask], "padded view")


def test_3d_uniform_steps(amr):
    say("case: 3D uniform depth 2, 8^3, v=(1,.5,.25), 20 steps")
    g, s, o = pair(amr, 3, 8, 1, 5, 64, (1.0, 0.5, 0.25), 0.3)
    for _ in range(2):
        g.refine_all()
        o.refine_all()
    assert_topology(g, o, "3D uniform")
    rng = np.random.default_rng(7)
    v = rng.random(g.size() * 512)
    g.set_interior(v)
    o.set_interior(v)
    o.halo_exchange()
    s.advance_batch_async(20)
    s.finish_advance_batch()
    o.advance_batch(20)
    assert_fields(g.get_interior(), o.get_interior(), "3D after 20 steps")


@pytest.mark.parametrize(
    "rank,n,h,depth,start,rounds,seed",
    [(2, 4, 1, 5, 2, 10, 1), (2, 4, 2, 5, 2, 6, 2), (2, 16, 1, 6, 2, 8, 3), (3, 4, 1, 4, 1, 6, 4), (3, 8, 1, 4, 2, 4, 5)],
)
def test_regrid_sequences(amr, rank, n, h, depth, start, rounds, seed):
    say(f"case: regrid sequence rank={rank} n={n} h={h} depth={depth} seed={seed}")
    vel = (1.0, 0.5, 0.25)[:rank]
    cap = (1 << (rank * depth)) + 64
    g, s, o = pair(amr, rank, n, h, depth, cap, vel, 0.4)
    for _ in range(start):
        g.refine_all()
        o.refine_all()
    rng = np.random.default_rng(seed)
    for rd in range(rounds):
        keys = o.keys()
        v = rng.standard_normal(o.size() * n**rank)
        g.set_interior(v)
        o.set_interior(v)
        flags = random_flags(rng, keys, rank, depth)
        cg = g.reconstruct_tree(flags)
        co = o.reconstruct(flags)
        assert cg == co, f"round {rd}: changed flag"
        assert_topology(g, o, f"round {rd}")
        assert_fields(g.get_interior(), o.get_interior(), f"round {rd}: transfer")
        o.halo_exchange()
        gp, op = g.get_patches(), o.get_padded()
        eh = brute.expected_halo(o.keys(), o.get_interior(), rank, n, h, depth)
        m = ~np.isnan(eh)
        assert_fields(gp[m], op[m], f"round {rd}: halos")
        s.advance_batch_async(3)
        s.finish_advance_batch()
        o.advance_batch(3)
        assert_fields(g.get_interior(), o.get_interior(), f"round {rd}: 3 steps")


def test_regrid_above_one_workgroup_planning(amr):
    """Regrids of trees above 4096 leaves plan with one launch per pass (the
    smaller ones in one workgroup): keys, transfer and steps against the
    oracle on a 2D depth-7 tree of 4^2 patches that grows past the limit."""
    say("case: regrids above the one-workgroup planning limit")
    rank, n, depth = 2, 4, 7
    cap = (1 << (rank * depth)) + 64
    g, s, o = pair(amr, rank, n, 1, depth, cap, (1.0, 0.5), 0.4)
    for _ in range(6):
        g.refine_all()
        o.refine_all()
    rng = np.random.default_rng(11)
    large = 0
    for rd in range(6):
        keys = o.keys()
        v = rng.standard_normal(o.size() * n**rank)
        g.set_interior(v)
        o.set_interior(v)
        if rd == 0:
            flags = (rng.random(len(keys)) < 0.3).astype(np.int8)  # grow past 4096 leaves
        else:
            flags = random_flags(rng, keys, rank, depth)
        large += g.size() > 4096
        assert g.reconstruct_tree(flags) == o.reconstruct(flags), f"round {rd}: changed flag"
        assert_topology(g, o, f"round {rd}")
        assert_fields(g.get_interior(), o.get_interior(), f"round {rd}: transfer")
        o.halo_exchange()
        s.advance_batch_async(2)
        s.finish_advance_batch()
        o.advance_batch(2)
        assert_fields(g.get_interior(), o.get_interior(), f"round {rd}: 2 steps")
    say("regrids above 4096 leaves:", large, "final leaves:", g.size())
    assert large >= 2


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("rank,n,h", [(2, 16, 1), (2, 10, 2), (3, 8, 1)])
def test_criterion_flags(amr, kind, rank, n, h):
    say(f"case: criterion kind={kind} rank={rank} n={n} h={h}")
    depth = 7 if rank == 2 else 5
    g, s, o = pair(amr, rank, n, h, depth, 4200, (1.0, 0.5, 0.25)[:rank], 0.4)
    for _ in range(3 if rank == 2 else 2):
        g.refine_all()
        o.refine_all()
    rng = np.random.default_rng(11)
    for rd in range(4):
        cen = o.cell_centers()
        c0 = rng.random(rank)
        r2 = sum((cen[..., d] - c0[d]) ** 2 for d in range(rank))
        u = np.exp(-r2 / (2 * 0.05**2))
        g.set_interior(u)
        o.set_interior(u)
        o.halo_exchange()
        thr = (0.5, 0.05) if kind == 0 else (0.05, 0.01)
        crit = amr.make_criterion(kind, thr[0], thr[1], 2 if rank == 2 else 1, 6 if rank == 2 else 4)
        fg = g.compute_flags(crit)
        fo = o.criterion(kind, thr[0], thr[1], 2 if rank == 2 else 1, 6 if rank == 2 else 4)
        np.testing.assert_array_equal(fg, fo)
        assert g.reconstruct_tree(crit) == o.reconstruct(fo)
        assert_topology(g, o, f"criterion round {rd}")


@pytest.mark.parametrize("n", [10, 50])
def test_reference_pulse_center_movement(amr, n):
    """tests/advection_equation_amr.t.cpp:50-130 through the C ABI."""
    say(f"case: PulseCenterMovement n={n}")
    g, s, o = pair(amr, 2, n, 2, 7, 2000, (1.0, 0.5), 0.4)
    s.initialize(gaussian_ic)
    o.set_interior(gaussian_ic(o.cell_centers()))
    o.halo_exchange()
    t, target = 0.0, 0.5
    while t < target:
        dt = s.advance()
        o.advance()
        if t + dt > target:
            dt = target - t
        g.halo_exchange_update()
        t += dt
    u = g.get_interior()
    assert_fields(u, o.get_interior(), "pulse")
    cen = g.cell_centers()[0].reshape(n, n, 2)
    iy, ix = np.unravel_index(np.argmax(u[0]), u[0].shape)
    for d, v in enumerate((1.0, 0.5)):
        assert abs(cen[iy, ix][d] - (0.2 + v * target)) <= 1.0 / n


def test_reference_pulse_center_movement_amr(amr):
    """tests/advection_equation_amr.t.cpp:132-231 (10x10 patches, halo 2,
    max-value criterion refine > 0.1 below level 5, coarsen < 0.05 above 1,
    regrid after every step)."""
    say("case: PulseCenterMovementAMR n=10")
    n = 10
    g, s, o = pair(amr, 2, n, 2, 7, 2000, (1.0, 0.5), 0.4)
    s.initialize(gaussian_ic)
    o.set_interior(gaussian_ic(o.cell_centers()))
    o.halo_exchange()
    crit = amr.make_criterion(0, 0.1, 0.05, 1, 5)
    t, target, steps = 0.0, 0.5, 0
    while t < target:
        dt = s.advance()
        assert dt == o.advance()
        fo = o.criterion(0, 0.1, 0.05, 1, 5)
        np.testing.assert_array_equal(g.compute_flags(crit), fo, err_msg=f"step {steps}")
        g.reconstruct_tree(crit)
        o.reconstruct(fo)
        o.halo_exchange()
        t += dt
        steps += 1
        if steps % 10 == 0:
            assert_topology(g, o, f"step {steps}")
            assert_fields(g.get_interior(), o.get_interior(), f"step {steps}")
    assert_topology(g, o, "final")
    u = g.get_interior()
    assert_fields(u, o.get_interior(), "final field")
    keys = g.keys()
    cen = g.cell_centers()
    i = int(np.argmax(u.reshape(-1)))
    pos = cen.reshape(-1, 2)[i]
    _, lv = amr.decode_keys(keys[:1], 2)
    cs = g.cell_sizes(lv)[0]
    for d, v in enumerate((1.0, 0.5)):
        assert abs(pos[d] - (0.2 + v * target)) <= cs[d]


def test_active_amr_gradient_short(amr):
    """C2-style run (seeded 8-bump mixture, gradient criterion every 10 steps,
    levels 3..7) for 60 steps, compared with the oracle at every regrid."""
    say("case: active AMR gradient, 60 steps")
    g, s, o = pair(amr, 2, 16, 1, 7, 20000, (1.0, 0.5), 0.4)
    for _ in range(3):
        g.refine_all()
        o.refine_all()
    rng = np.random.default_rng(2026)
    amp, sig, cx = rng.uniform(0.5, 1, 8), rng.uniform(0.03, 0.1, 8), rng.uniform(0, 1, (8, 2))

    def ic(cen):
        u = np.zeros(cen.shape[:-1])
        for a, sg, c in zip(amp, sig, cx):
            u += a * np.exp(-((cen[..., 0] - c[0]) ** 2 + (cen[..., 1] - c[1]) ** 2) / (2 * sg * sg))
        return u

    s.initialize(ic)
    o.set_interior(ic(o.cell_centers()))
    o.halo_exchange()
    crit = amr.make_criterion(1, 0.05, 0.01, 3, 7)
    for it in range(6):
        fo = o.criterion(1, 0.05, 0.01, 3, 7)
        np.testing.assert_array_equal(g.compute_flags(crit), fo)
        g.reconstruct_tree(crit)
        o.reconstruct(fo)
        o.halo_exchange()
        assert_topology(g, o, f"regrid {it}")
        s.advance_batch_async(10)
        s.finish_advance_batch()
        o.advance_batch(10)
        assert_fields(g.get_interior(), o.get_interior(), f"after batch {it}")
    say("leaves at end:", g.size())


def test_c3_full_size_periodic_tiling(amr):
    """Full C3 shape (32768 leaves of 8^3, 16.8M cells): a field with period
    1/4 evolves exactly like the oracle's 1/4-size periodic box, tiled."""
    say("case: C3 full size, tiling property, 10 steps")
    small = OracleTree(3, 8, 1, 5, 1200, lengths=(0.25, 0.25, 0.25))
    small.set_params((1.0, 0.5, 0.25), 0.3, threads=8)
    for _ in range(3):
        small.refine_all()
    rng = np.random.default_rng(42)
    small.set_interior(rng.random(small.size() * 512))
    small.halo_exchange()
    dense_small = brute.dense_from_uniform(small.keys(), small.get_interior(), 3, 8, 5)
    g = amr.NdTree(3, 5, 8, 1, 32768)
    s = amr.AmrSolver(g, (1.0, 0.5, 0.25), 0.3)
    for _ in range(5):
        g.refine_all()
    assert g.size() == 32768
    keys = g.keys()
    big = np.tile(dense_small, (4, 4, 4))
    g.set_interior(brute.split_uniform(keys, big, 3, 8, 5))
    assert s.compute_dt() == small.compute_dt() == 0.3 / 256.0
    s.advance_batch_async(10)
    s.finish_advance_batch()
    small.advance_batch(10)
    got = brute.dense_from_uniform(keys, g.get_interior(), 3, 8, 5)
    exp = np.tile(brute.dense_from_uniform(small.keys(), small.get_interior(), 3, 8, 5), (4, 4, 4))
    assert_fields(got, exp, "C3 tiling")
    assert abs(got.sum() - big.sum()) <= 1e-12 * big.sum()


def test_capacity_and_root_errors(amr):
    say("case: error paths")
    g = amr.NdTree(2, 7, 16, 1, 10)
    g.refine_all()
    with pytest.raises(amr.AmrError) as e:
        g.refine_all()  # 16 > capacity 10
    assert e.value.code == amr.AMR_ERR_CAPACITY
    assert g.size() == 4
    with pytest.raises(amr.AmrError) as e:
        g.reconstruct_tree(np.full(4, 2, dtype=np.int8))
    assert e.value.code == amr.AMR_ERR_UNSUPPORTED
    # both refusals left the tree as it was: it still steps, like the oracle
    o = OracleTree(2, 16, 1, 7, 64)
    o.refine_all()
    np.testing.assert_array_equal(g.keys(), o.keys())
    s = amr.AmrSolver(g, (1.0, 0.5), 0.4)
    o.set_params((1.0, 0.5), 0.4, threads=1)
    v = np.random.default_rng(3).standard_normal(o.size() * o.ni)
    g.set_interior(v)
    o.set_interior(v)
    o.halo_exchange()  # the oracle's step reads the halos of the last exchange
    s.advance_batch_async(3)
    assert s.finish_advance_batch() == o.advance_batch(3)
    assert_fields(g.get_interior(), o.get_interior(), "steps after the refusals")
    g.close()
    # a regrid that fills the capacity exactly is accepted
    g2 = amr.NdTree(2, 7, 16, 1, 16)
    g2.refine_all()
    g2.refine_all()
    assert g2.size() == 16
    s2 = amr.AmrSolver(g2, (1.0, 0.5), 0.4)
    s2.advance_batch_async(2)
    assert s2.finish_advance_batch()[1] == 2
    # empty batches: no steps, no time
    before = g2.get_interior().copy()
    s2.advance_batch_async(0)
    assert s2.finish_advance_batch() == (0.0, 0)
    s2.advance_batch_async(5, 0.0)
    assert s2.finish_advance_batch() == (0.0, 0)
    assert np.array_equal(g2.get_interior().view(np.uint64), before.view(np.uint64))
    g2.close()
    # any even patch extent runs (12^2 has no specialised kernels); an odd
    # one cannot be split between children
    amr.NdTree(2, 7, 12, 1, 10).close()
    with pytest.raises(amr.AmrError) as e:
        amr.NdTree(2, 7, 13, 1, 10)
    assert e.value.code == amr.AMR_ERR_ARGUMENT


def test_reference_tests_through_cpp_api():
    """tests/advection_equation_amr.t.cpp ported to the amr::hip C++ API."""
    import subprocess

    from amrlib import ROOT

    say("case: C++ host API port of the reference's advection tests")
    exe = ROOT / "tests" / "cpp" / "reference_advection_tests.bin"
    assert exe.exists(), "build the native tests first (__graft_entry__.build())"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    say(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


@pytest.mark.parametrize("rank,n,levels,mixed", [(3, 8, 3, False), (3, 8, 3, True), (2, 16, 4, True)])
def test_step_variants_bit_identical(amr, rank, n, levels, mixed):
    """Every step-kernel variant (and the autotuner's pick) reproduces the
    oracle bit for bit, on uniform and on mixed-level trees."""
    say(f"case: step variants rank={rank} n={n} levels={levels} mixed={mixed}")
    depth = 5 if rank == 3 else 7
    vel = (1.0, 0.5, 0.25)[:rank]
    cap = 8**levels + 4096 if rank == 3 else 4**levels + 4096
    g, s, o = pair(amr, rank, n, 1, depth, cap, vel, 0.3)
    for _ in range(levels):
        g.refine_all()
        o.refine_all()
    if mixed:
        flags = np.zeros(g.size(), dtype=np.int8)
        flags[::7] = 1
        g.reconstruct_tree(flags)
        o.reconstruct(flags)
        assert_topology(g, o, "mixed tree")
    rng = np.random.default_rng(17)
    u0 = rng.random(g.size() * n**rank)
    o.set_interior(u0)
    o.halo_exchange()
    o.advance_batch(3)
    exp = o.get_interior()
    for v, name in enumerate(s.step_variants()):
        s.set_step_variant(v)
        g.set_interior(u0)
        s.advance_batch_async(3)
        s.finish_advance_batch()
        assert_fields(g.get_interior(), exp, f"variant {name}")
    best, us = s.autotune(5)
    assert s.step_variant() == best and us > 0
    g.set_interior(u0)
    s.advance_batch_async(3)
    s.finish_advance_batch()
    assert_fields(g.get_interior(), exp, "autotuned variant")


@pytest.mark.parametrize("rank,n,start", [(2, 16, 3), (3, 8, 2)])
def test_run_variant_across_regrids_and_calls(amr, rank, n, start):
    """The run-streaming variant keeps its run lists current across single
    steps, batches of steps, regrids, set_interior and variant switches:
    every state matches the oracle bit for bit."""
    say(f"case: run8 across regrids rank={rank}")
    depth = 7 if rank == 2 else 4
    vel = (0.8, -0.45, 0.3)[:rank]
    g, s, o = pair(amr, rank, n, 1, depth, 16384, vel, 0.35)
    for _ in range(start):
        g.refine_all()
        o.refine_all()
    names = s.step_variants()
    rv = names.index("run8")
    s.set_step_variant(rv)
    rng = np.random.default_rng(31 + rank)
    u0 = rng.random(g.size() * n**rank)
    g.set_interior(u0)
    o.set_interior(u0)
    o.halo_exchange()
    for rd in range(4):
        flags = random_flags(rng, g.keys(), rank, depth)
        assert g.reconstruct_tree(flags) == o.reconstruct(flags)
        o.halo_exchange()
        assert_topology(g, o, f"round {rd}")
        # single steps, then a batch of steps, then a switch away and back
        for _ in range(2):
            s.advance_batch_async(1)
            s.finish_advance_batch()
            o.advance_batch(1)
        s.advance_batch_async(3)
        s.finish_advance_batch()
        o.advance_batch(3)
        s.set_step_variant(0)
        s.advance_batch_async(1)
        s.finish_advance_batch()
        o.advance_batch(1)
        s.set_step_variant(rv)
        s.advance_batch_async(2)
        s.finish_advance_batch()
        o.advance_batch(2)
        assert_fields(g.get_interior(), o.get_interior(), f"round {rd}")
    # an interior written from the host
    u1 = rng.random(g.size() * n**rank)
    g.set_interior(u1)
    o.set_interior(u1)
    o.halo_exchange()
    s.advance_batch_async(2)
    s.finish_advance_batch()
    o.advance_batch(2)
    assert_fields(g.get_interior(), o.get_interior(), "after set_interior")


@pytest.mark.parametrize("level,steps,seed", [(0, 5, 1), (1, 6, 2), (2, 9, 3), (3, 8, 4)])
def test_fused_pairs_bit_identical(amr, level, steps, seed):
    """Two steps per launch on single-level 3D trees: bit-identical to single
    steps and to the oracle, including a pair whose second step is clamped
    to the remaining time (even step counts) and an odd step count (the last
    step runs alone)."""
    say(f"case: fused steps, 3D 8^3 single level {level}, {steps} steps, seed {seed}")
    vel = (1.0, -0.5, 0.25)
    cap = (1 << (3 * level)) + 8
    fused, sf, o = pair(amr, 3, 8, 1, 5, cap, vel, 0.3)
    plain = amr.NdTree(3, 5, 8, 1, cap)
    sp = amr.AmrSolver(plain, vel, 0.3)
    sp.set_step_fusion(False)
    for _ in range(level):
        fused.refine_all()
        plain.refine_all()
        o.refine_all()
    assert sf.step_fusion() == (True, True)
    assert sp.step_fusion() == (False, True)
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(fused.size() * 512)
    for t in (fused, plain, o):
        t.set_interior(v)
    o.halo_exchange()
    dt = sf.compute_dt()
    remaining = dt * (steps - 0.5)
    sf.advance_batch_async(steps, remaining)
    got_dt = sf.finish_advance_batch()
    sp.advance_batch_async(steps, remaining)
    assert got_dt == sp.finish_advance_batch() == o.advance_batch(steps, remaining)
    got = fused.get_interior()
    assert np.array_equal(got.view(np.uint64), plain.get_interior().view(np.uint64)), "fused vs single steps"
    assert_fields(got, o.get_interior(), "fused vs oracle")


def test_fusion_follows_the_topology(amr):
    """Fusion applies exactly while every leaf sits on one level."""
    say("case: fusion applicability across regrids")
    g = amr.NdTree(3, 5, 8, 1, 600)
    s = amr.AmrSolver(g, (1.0, 0.5, 0.25), 0.3)
    g.refine_all()
    assert s.step_fusion() == (True, True)
    flags = np.zeros(g.size(), dtype=np.int8)
    flags[3] = 1
    assert g.reconstruct_tree(flags)
    assert s.step_fusion() == (True, False)
    g.refine_all()
    assert s.step_fusion() == (True, False)
    g2 = amr.NdTree(2, 7, 16, 1, 100)
    s2 = amr.AmrSolver(g2, (1.0, 0.5), 0.4)
    g2.refine_all()
    assert s2.step_fusion() == (True, False), "no fused kernel for 2D patches"


def test_swap_buffers_semantics(amr):
    """amr_tree_swap_buffers is ndtree::swap_buffers (ndtree.hpp:1558-1579):
    the current and next state exchange roles.  After one step the next
    buffer holds the state the step read, so a swap returns to it (bit for
    bit, as the oracle's swap does), a second swap returns to the stepped
    state, and stepping on continues from there.  A fused pair of steps
    (single-level trees) writes u^(n+2) over no intermediate state: the next
    buffer then holds u^n."""
    say("case: swap_buffers after single and fused steps")
    g, s, o = pair(amr, 2, 16, 1, 7, 300, (1.0, 0.5), 0.4)
    for _ in range(2):
        g.refine_all()
        o.refine_all()
    g.reconstruct_tree(np.array([1] + [0] * 15, dtype=np.int8))  # two levels: no fusion
    o.reconstruct(np.array([1] + [0] * 15, dtype=np.int8))
    rng = np.random.default_rng(13)
    u0 = rng.standard_normal(g.size() * 256)
    g.set_interior(u0)
    o.set_interior(u0)
    o.halo_exchange()
    assert s.step_fusion() == (True, False)
    dt = s.advance()
    o.step(dt)
    u1 = o.get_interior().copy()
    assert_fields(g.get_interior(), u1, "after one step")
    g.swap_buffers()
    o.swap_buffers()
    assert_fields(g.get_interior(), o.get_interior(), "swapped")
    assert np.array_equal(g.get_interior().reshape(-1).view(np.uint64), u0.view(np.uint64)), "swap returns to u^n"
    g.swap_buffers()
    o.swap_buffers()
    assert_fields(g.get_interior(), u1, "swapped back")
    o.halo_exchange()
    s.advance_batch_async(3)
    s.finish_advance_batch()
    o.advance_batch(3)
    assert_fields(g.get_interior(), o.get_interior(), "3 more steps")
    g.close()
    # single-level 8^3 tree: a fused pair
    g = amr.NdTree(3, 5, 8, 1, 128)
    s = amr.AmrSolver(g, (1.0, 0.5, 0.25), 0.3)
    for _ in range(2):
        g.refine_all()
    w0 = rng.standard_normal(g.size() * 512)
    g.set_interior(w0)
    assert s.step_fusion() == (True, True)
    s.advance_batch_async(2)
    s.finish_advance_batch()
    u2 = g.get_interior().copy()
    g.swap_buffers()
    assert np.array_equal(g.get_interior().reshape(-1).view(np.uint64), w0.view(np.uint64))
    g.swap_buffers()
    assert np.array_equal(g.get_interior().view(np.uint64), u2.view(np.uint64))
    g.close()


@pytest.mark.parametrize("rank,n,depth", [(2, 16, 16), (3, 8, 16), (3, 4, 16)])
def test_deep_graded_tree(amr, rank, n, depth):
    """A tree refined towards one point down to the deepest level the key
    format holds (16): keys and tables after every regrid, then halos
    (against the oracle and the geometric brute force) and steps with
    16 levels of cell sizes in one tree."""
    say(f"case: graded tree to level {depth} rank={rank} n={n}")
    cap = 16384
    vel = (1.0, -0.5, 0.25)[:rank]
    g, s, o = pair(amr, rank, n, 1, depth, cap, vel, 0.4)
    g.refine_all()
    o.refine_all()
    point = np.array([0.3141592653, 0.2718281828, 0.1618033988][:rank])
    at = np.floor(point * (1 << depth)).astype(np.int64)
    for lv in range(2, depth + 1):
        keys = o.keys()
        flags = np.zeros(len(keys), dtype=np.int8)
        for i, k in enumerate(keys):
            c, level = brute.decode(k, rank)
            size = 1 << (depth - level)
            if all(c[d] <= at[d] < c[d] + size for d in range(rank)):
                flags[i] = 1
        assert flags.sum() == 1
        assert g.reconstruct_tree(flags) == o.reconstruct(flags), f"level {lv}: changed flag"
        assert_topology(g, o, f"level {lv}")
    levels = sorted({int(k) & 63 for k in o.keys()})
    say(f"leaves {g.size()}, levels {levels[0]}..{levels[-1]}")
    assert levels[-1] == depth
    rng = np.random.default_rng(depth + rank)
    v = rng.standard_normal(o.size() * o.ni)
    g.set_interior(v)
    o.set_interior(v)
    o.halo_exchange()
    gp, op = g.get_patches(), o.get_padded()
    eh = brute.expected_halo(o.keys(), o.get_interior(), rank, n, 1, depth)
    m = ~np.isnan(eh)
    assert_fields(gp[m], op[m], "halos")
    assert_fields(gp[m], eh[m], "halos vs geometry")
    assert s.compute_dt() == o.compute_dt()
    s.advance_batch_async(4)
    tot, ex = s.finish_advance_batch()
    otot, oex = o.advance_batch(4)
    assert ex == oex == 4 and tot == otot
    assert_fields(g.get_interior(), o.get_interior(), "4 steps")
    g.close()
