"""The oracle's two refinement criteria against their definitions.

brute.criterion_values / brute.decide restate the criteria in numpy from
what they compute (the maximum over the interior and face halos, starting
from 0, as the reference's CPU criteria do, examples/fvm_solver_advection.e.cpp:66-101,
tests/advection_equation_amr.t.cpp:155-170; the undivided gradient, which the
build defines) on padded patches filled by brute.expected_halo's geometric
rules, not by the oracle's halo exchange.  The oracle's flags must equal them
on mixed-level trees in 2D (halo 1 and 2) and 3D with:
  - thresholds equal to leaves' own criterion values (strict comparisons),
  - min_level / max_level at the levels present (the level guards),
  - all-negative fields with negative thresholds (the start value 0),
  - corner cells of the padded buffer set to 1e6 (never read).
test_mutants_are_caught then shows that the same data tell the definition
apart from plausible mistakes: >= for >, <= for < on the level guards, a start
of -inf instead of 0 (the reference's CUDA kernel starts from -DBL_MAX,
src/cuda/fvm_refinement_criterion.cu:46), corner cells included, a one-sided
difference, and the compact instead of the padded stride.
"""
import itertools

import numpy as np
import pytest

import brute
from oracle import OracleTree
from test_oracle_tree import random_flags

# (rank, patch extents, halo, depth); non-square patches as static_shape<N, M[, O]>
CASES = [(2, 4, 1, 5), (2, 4, 2, 5), (2, 8, 1, 5), (2, 10, 2, 5), (3, 4, 1, 4), (2, (16, 8), 1, 5), (2, (12, 8), 2, 5),
         (3, (4, 6, 8), 1, 4)]
CORNER = 1.0e6


def corner_mask(rank, n, h):
    ext = brute.extents(rank, n)
    m = np.zeros(tuple(x + 2 * h for x in ext), dtype=bool)
    for idx in itertools.product(*[range(x + 2 * h) for x in ext]):
        m[idx] = sum(1 for d in range(rank) if idx[d] < h or idx[d] >= h + ext[d]) >= 2
    return m


def make_case(rank, n, h, depth, seed, negative):
    rng = np.random.default_rng(seed)
    t = OracleTree(rank, n, h, depth, 4 * (1 << (rank * depth)) + 64)
    t.refine_all()
    t.refine_all()
    for _ in range(3):
        t.reconstruct(random_flags(rng, t.keys(), rank, depth))
    keys = t.keys()
    vals = rng.standard_normal(t.size() * t.ni)
    if negative:
        vals = -np.abs(vals) - 1e-3
    t.set_interior(vals)
    pad = t.get_padded()
    pad[:, corner_mask(rank, n, h)] = CORNER
    t.set_padded(pad)
    t.halo_exchange()
    exp_pad = brute.expected_halo(keys, vals, rank, n, h, depth)
    return t, keys, exp_pad, rng


def threshold_sets(values, keys, rng):
    """Thresholds equal to leaves' own values, level guards at the levels
    present, plus negative thresholds."""
    levels = (np.asarray(keys, dtype=np.uint64) & np.uint64(63)).astype(int)
    lo, hi = int(levels.min()), int(levels.max())
    srt = np.sort(values)
    out = []
    for q_r, q_c in ((0.75, 0.25), (0.5, 0.5), (0.9, 0.1)):
        r = float(srt[int(q_r * (len(srt) - 1))])
        c = float(srt[int(q_c * (len(srt) - 1))])
        for mn, mx in ((lo, hi), (lo - 1, hi + 1), (lo + 1, hi - 1)):
            out.append((r, c, mn, mx))
    out.append((-0.5, -1e-6, lo - 1, hi + 1))
    out.append((float(rng.standard_normal()), float(rng.standard_normal()) - 1.0, lo, hi))
    return out


def collect():
    """(case, kind, thresholds, oracle flags, expected padded, keys)."""
    rows = []
    for ci, (rank, n, h, depth) in enumerate(CASES):
        for negative in (False, True):
            t, keys, exp_pad, rng = make_case(rank, n, h, depth, 100 + ci + 50 * negative, negative)
            for kind in (0, 1):
                values = brute.criterion_values(exp_pad, rank, n, h, kind)
                for thr in threshold_sets(values, keys, rng):
                    got = t.criterion(kind, *thr)
                    rows.append(((rank, n, h, depth, negative), kind, thr, got, exp_pad, keys))
    return rows


@pytest.fixture(scope="module")
def rows():
    return collect()


def test_oracle_criteria_match_definition(rows):
    seen = {"refine": 0, "coarsen": 0, "level_guard": 0, "at_threshold": 0}
    for case, kind, thr, got, exp_pad, keys in rows:
        rank, n, h, _, _ = case
        values = brute.criterion_values(exp_pad, rank, n, h, kind)
        exp = brute.decide(values, keys, *thr)
        assert np.array_equal(got, exp), f"{case} kind {kind} thresholds {thr}"
        r, c, mn, mx = thr
        lv = (np.asarray(keys, dtype=np.uint64) & np.uint64(63)).astype(int)
        seen["refine"] += int((exp == 1).sum())
        seen["coarsen"] += int((exp == 2).sum())
        seen["level_guard"] += int((((values > r) & (lv >= mx)) | ((values < c) & (lv <= mn) & ~(values > r))).sum())
        seen["at_threshold"] += int(((values == r) | (values == c)).sum())
    assert all(v > 0 for v in seen.values()), seen


def mutant_values(exp_pad, rank, n, h, kind, which):
    pad = np.array(exp_pad)
    if which == "start_lowest":
        return brute.criterion_values(pad, rank, n, h, kind, start=-np.inf)
    if which == "corners" and kind == 0:
        pad[:, corner_mask(rank, n, h)] = CORNER
        return np.array([max(0.0, float(p.max())) for p in pad])
    if which in ("one_sided", "compact_stride") and not isinstance(n, int):
        return None  # written for square patches; the square cases catch them
    if which == "one_sided" and kind == 1:
        inner = slice(h, h + n)
        out = []
        for p in pad:
            u = p[(inner,) * rank]
            g = 0.0
            for d in range(rank):
                up = p[tuple(slice(h + 1, h + n + 1) if e == d else inner for e in range(rank))]
                g = max(g, float(np.max(np.abs(up - u))))
            out.append(g)
        return np.array(out)
    if which == "compact_stride" and kind == 1:
        P = n + 2 * h
        out = []
        for p in pad:
            flat = np.nan_to_num(p.ravel(), nan=0.0)
            g = 0.0
            for idx in itertools.product(range(h, h + n), repeat=rank):
                l = 0
                for d in range(rank):
                    l = l * P + idx[d]
                for d in range(rank):
                    s = n ** (rank - 1 - d)
                    g = max(g, abs(flat[l + s] - flat[l]), abs(flat[l] - flat[l - s]))
            out.append(g)
        return np.array(out)
    return None


def mutant_decide(values, keys, thr, which):
    r, c, mn, mx = thr
    lv = (np.asarray(keys, dtype=np.uint64) & np.uint64(63)).astype(int)
    flags = np.zeros(len(values), dtype=np.int8)
    for i, (v, l) in enumerate(zip(values, lv)):
        if which == "ge":
            ref, coa = v >= r and l < mx, v <= c and l > mn
        else:  # "level_le"
            ref, coa = v > r and l <= mx, v < c and l >= mn
        flags[i] = 1 if ref else (2 if coa else 0)
    return flags


@pytest.mark.parametrize("which", ["ge", "level_le", "start_lowest", "corners", "one_sided", "compact_stride"])
def test_mutants_are_caught(rows, which):
    caught = False
    for case, kind, thr, got, exp_pad, keys in rows:
        rank, n, h, _, _ = case
        if which in ("ge", "level_le"):
            values = brute.criterion_values(exp_pad, rank, n, h, kind)
            mflags = mutant_decide(values, keys, thr, which)
        else:
            mv = mutant_values(exp_pad, rank, n, h, kind, which)
            if mv is None:
                continue
            mflags = brute.decide(mv, keys, *thr)
        if not np.array_equal(mflags, got):
            caught = True
            break
    assert caught, f"the criterion tests cannot tell the mutant '{which}' from the oracle"
