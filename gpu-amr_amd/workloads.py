"""Synthetic workloads of the benchmark configurations, built through the
public API (used by bench.py and the tests).

C3: 3D uniform octree of depth 5, 8^3 FP64 patches, seeded Gaussian bump,
    v = (1, 0.5, 0.25), CFL 0.3.
C2: 2D active AMR: uniform depth 3, seeded (2026) 8-bump mixture, v = (1, 0.5),
    CFL 0.4, gradient criterion (refine > 0.05, coarsen < 0.01) every 10
    steps, levels 3..7.
C5: C2 with levels 3..6 and 300 steps; the tree is first adapted to the
    initial field (criterion applied until it no longer changes the tree,
    the field re-sampled after each regrid), then the solve is recorded with
    a checkpoint every 20 steps.
C4: 64 independent 2D instances; instance i (seed 1000 + i) starts uniform at
    depth 3 and is refined four times with each leaf flagged Refine with
    probability 0.35 (levels capped at 7, 2:1 enforced by the regrid), carries
    a seeded mixture of 8 Gaussian bumps (amplitudes U[0.5, 1], sigmas
    U[0.03, 0.1], centres U[0, 1]^2) and a velocity drawn from U[-1, 1]^2
    (redrawn while min |v_d| < 1e-3); CFL 0.4.
"""
from __future__ import annotations

import numpy as np


def c3_dense_ic(ncell=256, seed=42):
    """Seeded 3D Gaussian bump (amplitude 1, sigma 0.1, centre U[.25,.75]^3)
    at the cell centres of the uniform ncell^3 grid; array order z, y, x."""
    c = np.random.default_rng(seed).uniform(0.25, 0.75, 3)
    h = 1.0 / ncell
    x = (np.arange(ncell) * h) + 0.5 * h
    gx = np.exp(-((x - c[0]) ** 2) / (2 * 0.1**2))
    gy = np.exp(-((x - c[1]) ** 2) / (2 * 0.1**2))
    gz = np.exp(-((x - c[2]) ** 2) / (2 * 0.1**2))
    return gz[:, None, None] * gy[None, :, None] * gx[None, None, :]


def dense_to_leaves(amr, keys, dense, n, depth, rank=3):
    """Cut a dense uniform-level array (layout order) into leaf interiors."""
    coords, lv = amr.decode_keys(keys, rank)
    s = 1 << (depth - int(lv[0]))
    nb = dense.shape[0] // n
    if rank == 3:
        blocks = dense.reshape(nb, n, nb, n, nb, n).transpose(0, 2, 4, 1, 3, 5)
        return np.ascontiguousarray(blocks[coords[:, 2] // s, coords[:, 1] // s, coords[:, 0] // s])
    blocks = dense.reshape(nb, n, nb, n).transpose(0, 2, 1, 3)
    return np.ascontiguousarray(blocks[coords[:, 1] // s, coords[:, 0] // s])


def build_c3(amr, depth=5, n=8):
    tree = amr.NdTree(3, depth, n, 1, 8 ** depth)
    solver = amr.AmrSolver(tree, (1.0, 0.5, 0.25), 0.3)
    for _ in range(depth):
        tree.refine_all()
    tree.set_interior(dense_to_leaves(amr, tree.keys(), c3_dense_ic(n * (1 << depth)), n, depth))
    return tree, solver


def mixture(rng, bumps=8):
    amp = rng.uniform(0.5, 1.0, bumps)
    sig = rng.uniform(0.03, 0.1, bumps)
    cen = rng.uniform(0.0, 1.0, (bumps, 2))

    def ic(c):
        u = np.zeros(c.shape[:-1])
        for a, s, p in zip(amp, sig, cen):
            u += a * np.exp(-((c[..., 0] - p[0]) ** 2 + (c[..., 1] - p[1]) ** 2) / (2 * s * s))
        return u

    return ic


def c4_velocity(rng):
    while True:
        v = rng.uniform(-1.0, 1.0, 2)
        if np.min(np.abs(v)) >= 1e-3:
            return (float(v[0]), float(v[1]))


def c4_instance(amr, i, max_depth=7, rounds=4, p_refine=0.35, start_level=3, capacity=16384 + 64):
    """Instance i of the batched configuration.  Returns (tree, solver,
    flag history, velocity, ic) so that the same tree can be rebuilt
    elsewhere (the oracle) by replaying the flags."""
    rng = np.random.default_rng(1000 + i)
    tree = amr.NdTree(2, 7, 16, 1, capacity)
    for _ in range(start_level):
        tree.refine_all()
    history = []
    for _ in range(rounds):
        keys = tree.keys()
        lv = (keys & np.uint64(63)).astype(np.int64)
        flags = np.where((rng.random(keys.size) < p_refine) & (lv < max_depth), 1, 0).astype(np.int8)
        history.append(flags)
        tree.reconstruct_tree(flags)
    vel = c4_velocity(rng)
    ic = mixture(rng)
    solver = amr.AmrSolver(tree, vel, 0.4)
    solver.initialize(ic)
    return tree, solver, history, vel, ic


def c2_mixture(seed=2026):
    """The C2 initial field: 8 Gaussian bumps drawn in the order amplitudes,
    sigmas, centres."""
    rng = np.random.default_rng(seed)
    amp, sig, cen = rng.uniform(0.5, 1.0, 8), rng.uniform(0.03, 0.1, 8), rng.uniform(0.0, 1.0, (8, 2))

    def ic(c):
        u = np.zeros(c.shape[:-1])
        for a, s, p in zip(amp, sig, cen):
            u += a * np.exp(-((c[..., 0] - p[0]) ** 2 + (c[..., 1] - p[1]) ** 2) / (2 * s * s))
        return u

    return ic


def c5_setup(amr, max_level=6, capacity=4096 + 64, max_adapt=8):
    """Tree, solver, criterion and the flag history of the initial adaptation
    (so that the oracle can rebuild the same tree)."""
    tree = amr.NdTree(2, 7, 16, 1, capacity)
    for _ in range(3):
        tree.refine_all()
    solver = amr.AmrSolver(tree, (1.0, 0.5), 0.4)
    ic = c2_mixture()
    solver.initialize(ic)
    crit = amr.make_criterion(amr.CRITERION_GRADIENT, 0.05, 0.01, 3, max_level)
    history = []
    for _ in range(max_adapt):
        flags = tree.compute_flags(crit)
        if not tree.reconstruct_tree(flags):
            break
        history.append(flags)
        solver.initialize(ic)
    return tree, solver, crit, history, ic


# ---------------------------------------------------------------------------
# Euler: the reference's three benchmarks.  Domain 1000^R, gamma 1.4, cfl 0.3,
# regrid by the maximum density over the padded patch (refine above / coarsen
# below the thresholds, within [min_level, max_level]).
#  - "3d": benchmark/bench_fvm_solver_integration3D.b.cpp: 8^3 patches, depth
#    5, two refine-all, acoustic pulse, regrid every 5 steps, 0.53 / 0.49,
#    levels 1..5, tmax 200;
#  - "2d": benchmark/bench_fvm_solver_integration.b.cpp: 64^2 patches, depth
#    7, three refine-all, acoustic pulse, regrid every 10 steps, 0.53 / 0.49,
#    levels 1..6, tmax 50;
#  - "2d_active": benchmark/bench_fvm_solver_integration_active_amr.b.cpp:
#    64^2 patches, depth 7, two refine-all, two pulses, regrid every 2 steps,
#    0.512 / 0.506, levels 1..7, tmax 30.
# ---------------------------------------------------------------------------
EULER_CONFIGS = {
    "3d": dict(rank=3, patch=8, depth=5, initial=2, every=5, refine=0.53, coarsen=0.49, min_level=1,
               max_level=5, tmax=200.0, capacity=50000, ic="pulse",
               source="benchmark/bench_fvm_solver_integration3D.b.cpp"),
    "2d": dict(rank=2, patch=64, depth=7, initial=3, every=10, refine=0.53, coarsen=0.49, min_level=1,
               max_level=6, tmax=50.0, capacity=10000, ic="pulse",
               source="benchmark/bench_fvm_solver_integration.b.cpp"),
    "2d_active": dict(rank=2, patch=64, depth=7, initial=2, every=2, refine=0.512, coarsen=0.506, min_level=1,
                      max_level=7, tmax=30.0, capacity=10000, ic="two_pulses",
                      source="benchmark/bench_fvm_solver_integration_active_amr.b.cpp"),
}


def euler_pulse(rank, length=1000.0):
    """Primitive acoustic pulse: rho = 0.5 + 0.2 a, u = 0, p = 1 + a with
    a = 10 exp(-r^2 / (0.01 L^2)) around the domain centre."""
    def ic(c):
        r2 = np.zeros(c.shape[:-1])
        for d in range(rank):
            dd = c[..., d] - 0.5 * length
            r2 = r2 + dd * dd
        pert = 10.0 * np.exp(-r2 / (0.01 * length * length))
        prim = np.zeros(c.shape[:-1] + (rank + 2,))
        prim[..., 0] = 0.5 + pert * 0.2
        prim[..., rank + 1] = 1.0 + pert
        return prim

    return ic


def euler_two_pulses(length=1000.0):
    """The active-AMR benchmark's 2D initial state
    (bench_fvm_solver_integration_active_amr.b.cpp:145-180): two pulses,
    a = 12 exp(-|x - (0.35, 0.45) L|^2 / w) + 9 exp(-|x - (0.68, 0.58) L|^2 / w),
    w = 0.004 L^2; rho = 0.5 + 0.2 a, u = v = 0, p = 1 + a."""
    w = 0.004 * length * length

    def ic(c):
        dxa, dya = c[..., 0] - 0.35 * length, c[..., 1] - 0.45 * length
        dxb, dyb = c[..., 0] - 0.68 * length, c[..., 1] - 0.58 * length
        pert = 12.0 * np.exp(-(dxa * dxa + dya * dya) / w) + 9.0 * np.exp(-(dxb * dxb + dyb * dyb) / w)
        prim = np.zeros(c.shape[:-1] + (4,))
        prim[..., 0] = 0.5 + pert * 0.2
        prim[..., 3] = 1.0 + pert
        return prim

    return ic


def euler_ic(cfg):
    """Primitive initial state of an Euler benchmark configuration."""
    if cfg["ic"] == "two_pulses":
        return euler_two_pulses()
    return euler_pulse(cfg["rank"])


def euler_benchmark(amr, name):
    """Solver (tree refined and initialised), criterion and configuration of
    one of the reference's Euler benchmarks ("3d", "2d", "2d_active"; a rank
    selects "3d" / "2d")."""
    if not isinstance(name, str):
        name = f"{int(name)}d"
    cfg = EULER_CONFIGS[name]
    rank = cfg["rank"]
    L = (1000.0,) * rank
    solver = amr.EulerSolver(rank, cfg["depth"], cfg["patch"], 1, cfg["capacity"], L, 1.4, 0.3)
    for _ in range(cfg["initial"]):
        solver.tree.refine_all()
    solver.initialize(euler_ic(cfg))
    crit = amr.make_criterion(amr.CRITERION_MAX_VALUE, cfg["refine"], cfg["coarsen"], cfg["min_level"],
                              cfg["max_level"])
    return solver, crit, cfg
