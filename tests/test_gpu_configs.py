"""The benchmark configurations at their full sizes, HIP path against the
oracle bit for bit (which also meets the 1e-12 relative bound):

C2: 2D active AMR, uniform depth 3, seeded 8-bump mixture (seed 2026),
    gradient criterion every 10 steps (refine > 0.05, coarsen < 0.01,
    levels 3..7), v = (1, 0.5), CFL 0.4, 500 steps;
C3: 3D uniform octree of depth 5 (32768 leaves of 8^3, 16.8M cells),
    seeded Gaussian bump, v = (1, 0.5, 0.25), CFL 0.3, 200 steps (mass
    conserved to round-off);
C4: 64 independent 2D instances stepped as one graph launch, 200 steps.
"""
import sys

import numpy as np
import pytest

from amrlib import load_package
from oracle import OracleTree
from test_gpu_parity import assert_fields, assert_topology

pytestmark = pytest.mark.gpu
THREADS = 16


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def amr():
    mod = load_package()
    import gpu_amr_amd.workloads  # noqa: F401

    return mod


def test_c3_config_200_steps(amr):
    from gpu_amr_amd import workloads

    say("case: C3, 200 steps, GPU vs oracle")
    g, s = workloads.build_c3(amr)
    o = OracleTree(3, 8, 1, 5, 2 * 32768 + 64)
    o.set_params((1.0, 0.5, 0.25), 0.3, threads=THREADS)
    for _ in range(5):
        o.refine_all()
    np.testing.assert_array_equal(o.keys(), g.keys())
    u0 = g.get_interior()
    o.set_interior(u0)
    o.halo_exchange()
    assert s.compute_dt() == o.compute_dt() == 0.3 / 256.0
    s.advance_batch_async(200)
    dt_sum, steps = s.finish_advance_batch()
    o.advance_batch(200)
    assert steps == 200
    got = g.get_interior()
    assert_fields(got, o.get_interior(), "C3 after 200 steps")
    # uniform periodic tree: mass conserved to round-off
    assert abs(got.sum() - u0.sum()) <= 1e-12 * abs(u0.sum())
    say(f"C3: 200 steps bit-identical, mass drift {abs(got.sum() - u0.sum()) / u0.sum():.2e}")


def test_c2_config_500_steps(amr):
    from gpu_amr_amd import workloads

    say("case: C2, 500 steps with a regrid every 10, GPU vs oracle")
    g = amr.NdTree(2, 7, 16, 1, 16384 + 64)
    s = amr.AmrSolver(g, (1.0, 0.5), 0.4)
    o = OracleTree(2, 16, 1, 7, 2 * 16384 + 128)
    o.set_params((1.0, 0.5), 0.4, threads=THREADS)
    for _ in range(3):
        g.refine_all()
        o.refine_all()
    ic = workloads.c2_mixture()
    s.initialize(ic)
    o.set_interior(ic(o.cell_centers()))
    o.halo_exchange()
    assert_fields(g.get_interior(), o.get_interior(), "C2 initial field")
    crit = amr.make_criterion(amr.CRITERION_GRADIENT, 0.05, 0.01, 3, 7)
    most = 0
    for it in range(50):
        fo = o.criterion(1, 0.05, 0.01, 3, 7)
        np.testing.assert_array_equal(g.compute_flags(crit), fo, err_msg=f"regrid {it}: flags")
        assert g.reconstruct_tree(crit) == o.reconstruct(fo)
        o.halo_exchange()
        assert_topology(g, o, f"regrid {it}")
        most = max(most, g.size())
        assert s.compute_dt() == o.compute_dt()
        s.advance_batch_async(10)
        s.finish_advance_batch()
        o.advance_batch(10)
        if it % 10 == 9:
            assert_fields(g.get_interior(), o.get_interior(), f"after {10 * (it + 1)} steps")
    levels = np.bincount((g.keys() & np.uint64(63)).astype(int))
    say(f"C2: 500 steps bit-identical; up to {most} leaves, final levels {levels.tolist()}")


def test_c4_config_64_instances_200_steps(amr):
    from gpu_amr_amd import workloads

    say("case: C4, 64 instances, 200 steps in one graph launch, GPU vs oracle")
    built = [workloads.c4_instance(amr, i) for i in range(64)]
    trees = [b[0] for b in built]
    init = [t.get_interior().copy() for t in trees]
    batch = amr.Batch(trees)
    batch.advance(200, use_graph=True)
    batch.synchronize()
    batch.scatter()
    for i, (tree, solver, history, vel, _) in enumerate(built):
        o = OracleTree(2, 16, 1, 7, 2 * 16384 + 128)
        o.set_params(vel, 0.4, threads=THREADS)
        for _ in range(3):
            o.refine_all()
        for f in history:
            o.reconstruct(f)
        np.testing.assert_array_equal(o.keys(), tree.keys(), err_msg=f"instance {i}: keys")
        o.set_interior(init[i])
        o.halo_exchange()
        o.advance_batch(200)
        assert_fields(tree.get_interior(), o.get_interior(), f"instance {i}")
    say("C4: 64 instances, 200 steps, bit-identical")
