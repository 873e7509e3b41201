"""The adjoint kernels' item bodies, executed on the CPU, against the oracle's
forward solver (no GPU needed).

For fixed topology, velocity and dt the step and the regrid transfer are
linear maps, so the transposes are pinned by the dot-product identity
<lam, F(x)> = <F^T(lam), x> on random vectors, with F the ORACLE's own step
(halo exchange + upwind update) and regrid.  The velocity partials are pinned
by central differences of <lam, F(u; v)> in v (exact up to rounding: with a
fixed sign of v_d one step is affine in v_d at fixed dt).
"""
import math

import numpy as np
import pytest

import emulation as emu
from oracle import OracleTree
from test_oracle_tree import random_flags


def fdot(a, b):
    return math.fsum(np.asarray(a, dtype=np.float64).ravel() * np.asarray(b, dtype=np.float64).ravel())


def amr_oracle(rank, n, h, depth, start, rounds, seed, vel):
    rng = np.random.default_rng(seed)
    cap = (1 << (rank * depth)) + 64
    o = OracleTree(rank, n, h, depth, 2 * cap + 64)
    o.set_params(vel, 0.4)
    for _ in range(start):
        o.refine_all()
    for _ in range(rounds):
        o.reconstruct(random_flags(rng, o.keys(), rank, depth))
    keys = o.keys()
    if len(set((keys & np.uint64(63)).tolist())) == 1:  # make sure the tree is adaptive
        flags = np.zeros(keys.size, dtype=np.int8)
        flags[0] = 1
        o.reconstruct(flags)
    return o, rng, cap


def oracle_step(o, x, dt):
    """One oracle step applied to interiors x (the map is linear in x)."""
    o.set_interior(x)
    o.halo_exchange()
    o.step(dt)
    return o.get_interior().reshape(-1).copy()


CASES = [(2, 4, 1, 5, 2, 4, 11), (2, 16, 1, 4, 2, 3, 12), (2, 8, 1, 5, 1, 5, 13), (3, 4, 1, 4, 1, 3, 14)]


@pytest.mark.parametrize("rank,n,h,depth,start,rounds,seed", CASES)
def test_adjoint_step_dot_product_against_oracle(rank, n, h, depth, start, rounds, seed):
    vel = (0.8, -0.6, 0.35)[:rank]
    o, rng, _ = amr_oracle(rank, n, h, depth, start, rounds, seed, vel)
    keys = o.keys()
    assert len(set((keys & np.uint64(63)).tolist())) > 1, "tree should be adaptive"
    links, metas, err = emu.build_table(rank, depth, keys)
    assert err == 0
    dt = o.compute_dt()
    max_level = int((keys & np.uint64(63)).max())
    dtdx, v, _ = emu.step_params(rank, n, depth, max_level, vel, dt)
    size = keys.size * n**rank
    u = rng.standard_normal(size)
    for trial in range(3):
        x = rng.standard_normal(size)
        lam_next = rng.standard_normal(size)
        fx = oracle_step(o, x, dt)
        lam, _ = emu.adj_step(rank, n, keys, links, metas, lam_next, u, dtdx, v)
        lhs, rhs = fdot(lam_next, fx), fdot(lam, x)
        scale = math.sqrt(fdot(lam_next, lam_next) * fdot(x, x))
        assert abs(lhs - rhs) <= 1e-13 * scale, f"trial {trial}: {lhs} vs {rhs}"


@pytest.mark.parametrize("rank,n,h,depth,start,rounds,seed", CASES)
def test_adjoint_velocity_partials_against_oracle_differences(rank, n, h, depth, start, rounds, seed):
    vel = np.array((0.8, -0.6, 0.35)[:rank])
    o, rng, _ = amr_oracle(rank, n, h, depth, start, rounds, seed + 100, tuple(vel))
    keys = o.keys()
    links, metas, _ = emu.build_table(rank, depth, keys)
    dt = o.compute_dt()  # held fixed under velocity perturbation
    max_level = int((keys & np.uint64(63)).max())
    dtdx, v, _ = emu.step_params(rank, n, depth, max_level, vel, dt)
    size = keys.size * n**rank
    u = rng.standard_normal(size)
    lam_next = rng.standard_normal(size)
    _, vpart = emu.adj_step(rank, n, keys, links, metas, lam_next, u, dtdx, v)
    hstep = 1e-3
    for d in range(rank):
        vals = []
        for sgn in (1.0, -1.0):
            vv = vel.copy()
            vv[d] += sgn * hstep
            o.set_params(tuple(vv), 0.4)
            vals.append(fdot(lam_next, oracle_step(o, u, dt)))
        fd = (vals[0] - vals[1]) / (2 * hstep)
        g = math.fsum(vpart[:, d])
        assert abs(fd - g) <= 1e-9 * max(1.0, abs(g)), f"dim {d}: fd {fd} adjoint {g}"
    # a sign flip of the velocity partials (a plausible slip) is not within tolerance
    assert any(abs(math.fsum(vpart[:, d])) > 1e-6 for d in range(rank))


@pytest.mark.parametrize("rank,n,h,depth,start,rounds,seed", CASES)
def test_adjoint_regrid_dot_product_against_oracle(rank, n, h, depth, start, rounds, seed):
    o, rng, cap = amr_oracle(rank, n, h, depth, start, rounds, seed + 200, (1.0, 0.5, 0.25)[:rank])
    done = 0
    for attempt in range(12):
        keys = o.keys()
        links, metas, _ = emu.build_table(rank, depth, keys)
        flags = random_flags(rng, keys, rank, depth)
        x = rng.standard_normal(keys.size * n**rank)
        max_level = int((keys & np.uint64(63)).max())
        changed, k2, y, summ, maps = emu.regrid(rank, n, depth, keys, flags, links, metas, x, max_level, 2 * cap,
                                                with_maps=True)
        o.set_interior(x)
        assert o.reconstruct(flags) == changed
        if not changed:
            continue
        # the emulated transfer is the oracle's (bit for bit) ...
        assert np.array_equal(y.view(np.uint64), o.get_interior().reshape(-1).view(np.uint64))
        np.testing.assert_array_equal(k2, o.keys())
        # ... and its transpose satisfies the dot-product identity
        lam_new = rng.standard_normal(y.size)
        lam_old = emu.adj_regrid(rank, n, lam_new, maps, keys.size)
        lhs, rhs = fdot(lam_new, y), fdot(lam_old, x)
        assert abs(lhs - rhs) <= 1e-13 * math.sqrt(fdot(lam_new, lam_new) * fdot(x, x)), f"attempt {attempt}"
        assert summ[1] + summ[2] > 0
        done += 1
    assert done >= 3


def test_adjoint_chain_through_regrid_against_oracle():
    """Two steps, a regrid, two more steps: the composed transposes of the
    emulated kernels against the oracle's composed forward map."""
    rank, n, h, depth = 2, 8, 1, 5
    vel = (0.9, 0.4)
    o, rng, cap = amr_oracle(rank, n, h, depth, 2, 3, 77, vel)
    keys0 = o.keys()
    flags = random_flags(rng, keys0, rank, depth)
    # record the topology sequence and dts with the oracle
    links0, metas0, _ = emu.build_table(rank, depth, keys0)
    dt0 = o.compute_dt()
    _, keys1, _, _, maps = emu.regrid(rank, n, depth, keys0, flags, links0, metas0, np.zeros(keys0.size * n**rank),
                                      int((keys0 & np.uint64(63)).max()), 2 * cap, with_maps=True)
    assert keys1.size > 0
    links1, metas1, _ = emu.build_table(rank, depth, keys1)

    def forward(x):
        o2, _, _ = amr_oracle(rank, n, h, depth, 2, 3, 77, vel)
        o2.set_interior(x)
        o2.halo_exchange()
        states = [x.copy()]
        for _ in range(2):
            o2.step(dt0)
            states.append(o2.get_interior().reshape(-1).copy())
        assert o2.reconstruct(flags)
        o2.halo_exchange()
        dt1 = o2.compute_dt()
        states.append(o2.get_interior().reshape(-1).copy())
        for _ in range(2):
            o2.step(dt1)
            states.append(o2.get_interior().reshape(-1).copy())
        return states, dt1

    x = rng.standard_normal(keys0.size * n**rank)
    states, dt1 = forward(x)
    lam_T = rng.standard_normal(states[-1].size)
    # backward
    ml1 = int((keys1 & np.uint64(63)).max())
    ml0 = int((keys0 & np.uint64(63)).max())
    d1, v1, _ = emu.step_params(rank, n, depth, ml1, vel, dt1)
    d0, v0, _ = emu.step_params(rank, n, depth, ml0, vel, dt0)
    lam = lam_T
    lam, _ = emu.adj_step(rank, n, keys1, links1, metas1, lam, states[4], d1, v1)
    lam, _ = emu.adj_step(rank, n, keys1, links1, metas1, lam, states[3], d1, v1)
    lam = emu.adj_regrid(rank, n, lam, maps, keys0.size)
    lam, _ = emu.adj_step(rank, n, keys0, links0, metas0, lam, states[1], d0, v0)
    lam, _ = emu.adj_step(rank, n, keys0, links0, metas0, lam, states[0], d0, v0)
    lhs, rhs = fdot(lam_T, states[-1]), fdot(lam, x)
    assert abs(lhs - rhs) <= 1e-13 * math.sqrt(fdot(lam_T, lam_T) * fdot(x, x)), f"{lhs} vs {rhs}"
