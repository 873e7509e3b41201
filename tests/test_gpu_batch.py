"""Batched form: many independent trees stepped as one launch (and as one
hipGraph launch) must reproduce the per-tree plain solves and the oracle bit
for bit, with each instance's own velocity and dt."""
import sys

import numpy as np
import pytest

from amrlib import load_package
from oracle import OracleTree

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def amr():
    mod = load_package()
    import gpu_amr_amd.workloads  # noqa: F401

    return mod


def build(amr, count, rounds=3, max_depth=6):
    from gpu_amr_amd import workloads

    out = [workloads.c4_instance(amr, i, max_depth=max_depth, rounds=rounds, capacity=4096 + 64) for i in range(count)]
    return out


def test_batch_matches_sequential_and_oracle(amr):
    say("case: batch of 6 C4-style instances, 20 steps, eager launches")
    built = build(amr, 6)
    trees = [b[0] for b in built]
    solvers = [b[1] for b in built]
    init = [t.get_interior().copy() for t in trees]
    batch = amr.Batch(trees)
    assert batch.size() == sum(t.size() for t in trees)
    batch.advance(20, use_graph=False)
    batch.synchronize()
    for s in solvers:
        s.advance_batch_async(20)
        s.finish_advance_batch()
    seq = [t.get_interior().copy() for t in trees]
    batch.scatter()
    for i, (t, u) in enumerate(zip(trees, seq)):
        assert np.array_equal(t.get_interior().view(np.uint64), u.view(np.uint64)), f"instance {i}"
    # the oracle on two instances (replaying the same regrid flags)
    for i in (0, 3):
        _, _, history, vel, _ = built[i]
        o = OracleTree(2, 16, 1, 7, 8192 + 64)
        o.set_params(vel, 0.4, threads=8)
        for _ in range(3):
            o.refine_all()
        for f in history:
            o.reconstruct(f)
        np.testing.assert_array_equal(o.keys(), trees[i].keys())
        o.set_interior(init[i])
        o.halo_exchange()
        o.advance_batch(20)
        assert np.array_equal(o.get_interior().view(np.uint64), seq[i].view(np.uint64)), f"oracle instance {i} v={vel}"


def test_batch_graph_replays(amr):
    say("case: batch hipGraph replays vs eager")
    built = build(amr, 4)
    trees = [b[0] for b in built]
    g_batch = amr.Batch(trees)
    e_batch = amr.Batch(trees)
    for rep in range(4):
        g_batch.advance(5, use_graph=True)
        g_batch.synchronize()
        e_batch.advance(5, use_graph=False)
        e_batch.synchronize()
        _ = trees[0].get_interior()  # unrelated eager work between replays
        e_batch.scatter()
        e_res = [t.get_interior().copy() for t in trees]
        g_batch.scatter()
        for i, t in enumerate(trees):
            assert np.array_equal(t.get_interior().view(np.uint64), e_res[i].view(np.uint64)), f"replay {rep} instance {i}"


def test_batch_variants_and_autotune(amr):
    say("case: batch variants")
    built = build(amr, 3)
    trees = [b[0] for b in built]
    ref = amr.Batch(trees)
    ref.advance(4, use_graph=False)
    ref.scatter()
    exp = [t.get_interior().copy() for t in trees]
    ref.close()
    nvar = len(built[0][1].step_variants())
    for v in range(nvar):
        # rebuild from the same initial state each time
        fresh = build(amr, 3)
        tr = [b[0] for b in fresh]
        bt = amr.Batch(tr)
        bt.set_step_variant(v)
        bt.advance(4, use_graph=(v % 2 == 0))
        bt.scatter()
        for i, t in enumerate(tr):
            assert np.array_equal(t.get_interior().view(np.uint64), exp[i].view(np.uint64)), f"variant {v} instance {i}"
    best, us = bt.autotune(3)
    assert 0 <= best < nvar and us > 0


def test_batch_run_variant_with_graphs_and_switches(amr):
    say("case: batch run8 graph replays with variant switches")
    built = build(amr, 4)
    trees = [b[0] for b in built]
    names = built[0][1].step_variants()
    rv = names.index("run8")
    ref = amr.Batch(trees)
    bt = amr.Batch(trees)
    plan = [(rv, True, 4), (0, True, 3), (rv, True, 4), (rv, False, 1), (1, False, 2), (rv, True, 4), (rv, True, 3)]
    for i, (v, graph, steps) in enumerate(plan):
        bt.set_step_variant(v)
        bt.advance(steps, use_graph=graph)
        bt.synchronize()
        ref.advance(steps, use_graph=False)
        ref.synchronize()
        ref.scatter()
        exp = [t.get_interior().copy() for t in trees]
        bt.scatter()
        for k, t in enumerate(trees):
            assert np.array_equal(t.get_interior().view(np.uint64), exp[k].view(np.uint64)), f"plan step {i} instance {k}"
