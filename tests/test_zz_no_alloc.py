"""amr_tree_compute_flags_noalloc (flags staged in the library's fixed device
array, no device allocation) against amr_tree_compute_flags, through the C
ABI: two calls in a row, the second on a smaller tree, both criteria, and the
size limit (2**20 leaves accepted, one more rejected).  Integer flags:
bit-exact."""
import sys

import numpy as np
import pytest

from amrlib import load_package

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def amr():
    return load_package()


def mixed_tree(amr, rank, n, depth, start, seed):
    """A tree refined by the gradient criterion from uniform level `start`,
    with a seeded bump field."""
    g = amr.NdTree(rank, depth, n, 1, 20000)
    s = amr.AmrSolver(g, (1.0, 0.5, 0.25)[:rank], 0.4)
    for _ in range(start):
        g.refine_all()
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.2, 0.8, size=(3, rank))

    def ic(x):
        return sum(np.exp(-np.sum((x - c) ** 2, axis=-1) / (2 * 0.05**2)) for c in centres)

    s.initialize(ic)
    g.reconstruct_tree(amr.make_criterion(amr.CRITERION_GRADIENT, 0.05, 0.01, start - 1, start + 2))
    s.initialize(ic)
    return g


def criteria(amr, start):
    return [
        amr.make_criterion(amr.CRITERION_GRADIENT, 0.05, 0.01, start - 1, start + 2),
        amr.make_criterion(amr.CRITERION_MAX_VALUE, 0.5, 0.1, start - 1, start + 2),
    ]


def test_noalloc_flags_match_two_calls_second_smaller(amr):
    say("case: 2D 16^2 mixed tree (start level 4), then a smaller 2D tree (start level 3), then 3D 8^3")
    # the larger of the two 2D trees goes first
    (big, bstart), (small, sstart) = sorted(
        ((mixed_tree(amr, 2, 16, 7, 4, seed=5), 4), (mixed_tree(amr, 2, 16, 7, 3, seed=6), 3)),
        key=lambda t: -t[0].size())
    cube = mixed_tree(amr, 3, 8, 5, 2, seed=7)
    assert small.size() < big.size()
    say(f"leaves: {big.size()}, then {small.size()}, then {cube.size()}")
    for tree, start in ((big, bstart), (small, sstart), (cube, 2)):
        for crit in criteria(amr, start):
            ref = tree.compute_flags(crit)
            got = tree.compute_flags_noalloc(crit)
            np.testing.assert_array_equal(got, ref)
            assert got.size == tree.size()
        # the flags are not trivial
        seen = set()
        for crit in criteria(amr, start):
            seen.update(tree.compute_flags_noalloc(crit).tolist())
        assert len(seen) >= 2
    # the first tree again after the smaller ones: same flags as before
    for crit in criteria(amr, bstart):
        np.testing.assert_array_equal(big.compute_flags_noalloc(crit), big.compute_flags(crit))
    for t in (big, small, cube):
        t.close()


def test_noalloc_size_limit(amr):
    say("case: 2D 4^2 uniform depth 10 (2**20 leaves), then 3 leaves refined (2**20 + 9 leaves)")
    g = amr.NdTree(2, 11, 4, 1, (1 << 20) + 64)
    for _ in range(10):
        g.refine_all()
    assert g.size() == 1 << 20
    rng = np.random.default_rng(3)
    g.set_interior(rng.standard_normal((g.size(), 4, 4)))
    crit = amr.make_criterion(amr.CRITERION_MAX_VALUE, 2.0, 0.5, 9, 11)
    ref = g.compute_flags(crit)
    np.testing.assert_array_equal(g.compute_flags_noalloc(crit), ref)
    assert len(set(ref.tolist())) == 3
    flags = np.zeros(g.size(), dtype=np.int8)
    flags[:3] = amr.REFINE
    g.reconstruct_tree(flags)
    assert g.size() == (1 << 20) + 9
    with pytest.raises(amr.AmrError) as e:
        g.compute_flags_noalloc(crit)
    assert e.value.code == amr.AMR_ERR_CAPACITY and "1048576" in str(e.value)
    # the context is unaffected: the allocating entry point still works
    assert g.compute_flags(crit).size == g.size()
    g.close()
