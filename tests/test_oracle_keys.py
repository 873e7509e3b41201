"""Oracle Morton keys and uniform trees against closed forms."""
import numpy as np
import pytest

import brute
from oracle import OracleTree


@pytest.mark.parametrize("rank,depth,levels", [(2, 7, 4), (3, 5, 3)])
def test_uniform_tree_keys_closed_form(rank, depth, levels):
    t = OracleTree(rank, 4, 1, depth, 2 * 8 ** levels + 16)
    assert list(t.keys()) == [0]
    for lv in range(1, levels + 1):
        assert t.refine_all()
        np.testing.assert_array_equal(t.keys(), brute.uniform_keys(rank, depth, lv))


def test_known_small_keys():
    # level-2 quadtree at depth 7: first leaf (0,0) -> key 2; the leaf at
    # x = 32, y = 0 (interleave -> bit 10) -> (1024 << 6) | 2.
    t = OracleTree(2, 4, 1, 7, 64)
    t.refine_all()
    t.refine_all()
    k = [int(x) for x in t.keys()]
    assert k[0] == 2
    assert (1024 << 6) | 2 in k
    # level-1 children of the root: offsets x=1 -> morton 1<<(2*6), y=1 -> 1<<(2*6+1)
    t1 = OracleTree(2, 4, 1, 7, 8)
    t1.refine_all()
    assert [int(x) for x in t1.keys()] == [1, (1 << 12 << 6) | 1, (2 << 12 << 6) | 1, (3 << 12 << 6) | 1]


def test_interleave_roundtrip():
    rng = np.random.default_rng(0)
    for rank, bits in ((2, 21), (3, 21)):
        for _ in range(200):
            c = tuple(int(x) for x in rng.integers(0, 1 << bits, size=rank))
            assert brute.deinterleave(brute.interleave(c), rank) == c


@pytest.mark.parametrize("rank,depth,levels", [(2, 7, 3), (3, 5, 2)])
def test_uniform_neighbor_tables(rank, depth, levels):
    t = OracleTree(rank, 4, 1, depth, 2 * 8 ** levels + 16)
    for lv in range(levels + 1):
        got = t.metadata()
        exp = brute.expected_metadata(t.keys(), rank, depth)
        assert got.tobytes() == exp.tobytes(), f"level {lv}"
        if lv < levels:
            t.refine_all()
