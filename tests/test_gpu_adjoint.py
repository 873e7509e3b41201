"""Discrete adjoint on the GPU: the recorded solve replays bit-exactly (and
matches the oracle's solve on the same topology and dt sequence), the adjoint
satisfies the dot-product identity, does not depend on the checkpoint
interval, and matches central finite differences (h = 1e-6) of the misfit
J = 0.5 ||u(T) - u_target||^2 on the C5 configuration (2D active AMR, levels
3..6, 300 steps, checkpoints every 20) to 1e-6 relative (with one Richardson
step where the velocity enters, see test_adjoint_c5_finite_differences)."""
import math
import sys
import time

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
    import gpu_amr_amd.adjoint  # noqa: F401
    import gpu_amr_amd.workloads  # noqa: F401

    return mod


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def fdot(a, b):
    return math.fsum((np.asarray(a, dtype=np.float64) * np.asarray(b, dtype=np.float64)).ravel())


def record_c5(amr, steps, checkpoint_every, max_level=6):
    from gpu_amr_amd import workloads

    tree, solver, crit, history, ic = workloads.c5_setup(amr, max_level=max_level)
    u0 = tree.get_interior().reshape(-1).copy()
    tape = amr.Tape(solver, steps, crit, regrid_every=10, checkpoint_every=checkpoint_every)
    return tree, solver, tape, u0, history


def oracle_replay(tape, history, u0, vel, threads=16):
    """The oracle's solve over the tape's recorded topology and dt sequence."""
    info = tape.info()
    o = OracleTree(2, 16, 1, 7, 2 * 4096 + 128)
    o.set_params(tuple(vel), 0.4, threads=threads)
    for _ in range(3):
        o.refine_all()
    for f in history:
        assert o.reconstruct(f)
    np.testing.assert_array_equal(o.keys(), tape.keys(0))
    o.set_interior(u0)
    o.halo_exchange()
    regrids = {}
    for i in range(info["regrids"]):
        step, _, changed, flags = tape.regrid(i)
        regrids[step] = (changed, flags)
    dts = tape.dts()
    for k in range(info["steps"]):
        if k in regrids:
            changed, flags = regrids[k]
            assert o.reconstruct(flags) == changed
            o.halo_exchange()
        o.step(float(dts[k]))
    return o.get_interior().reshape(-1).copy()


def replay(torch, tape, u0, vel):
    from gpu_amr_amd.adjoint import advection_solve

    with torch.no_grad():
        return advection_solve(u0, torch.tensor(vel, dtype=torch.float64), tape)


def test_tape_replay_matches_recorded_solve_and_oracle(amr, torch):
    say("case: tape record 40 steps (regrid every 10, checkpoint 7) + replay + oracle")
    tree, solver, tape, u0, history = record_c5(amr, 40, 7)
    info = tape.info()
    say("tape:", info)
    assert info["regrids"] == 3 and info["checkpoints"] == 6
    final = tree.get_interior().reshape(-1)
    assert final.size == info["final_leaves"] * 256
    u0_t = torch.from_numpy(u0).cuda()
    uT = replay(torch, tape, u0_t, (1.0, 0.5)).cpu().numpy()
    assert np.array_equal(uT.view(np.uint64), final.view(np.uint64))
    uo = oracle_replay(tape, history, u0, (1.0, 0.5))
    assert np.array_equal(uT.view(np.uint64), uo.view(np.uint64))
    # keys of the last topology are the tree's
    np.testing.assert_array_equal(tape.keys(info["topologies"] - 1), tree.keys())


def test_adjoint_dot_product_and_checkpoint_invariance(amr, torch):
    say("case: dot-product test, 40 steps, checkpoints 1 / 7 / 40")
    grads = []
    rng = np.random.default_rng(5)
    x = lam = None
    for ck in (1, 7, 40):
        tree, solver, tape, u0, _ = record_c5(amr, 40, ck)
        info = tape.info()
        if x is None:
            x = rng.standard_normal(u0.size)
            lam = rng.standard_normal(info["final_leaves"] * 256)
        x_t = torch.from_numpy(x).cuda()
        fx = replay(torch, tape, x_t, (1.0, 0.5)).cpu().numpy()
        g = torch.empty_like(x_t)
        lam_t = torch.from_numpy(lam).cuda()
        gv = tape.adjoint(lam_t.data_ptr(), g.data_ptr())
        gu = g.cpu().numpy()
        lhs, rhs = fdot(lam, fx), fdot(gu, x)
        rel = abs(lhs - rhs) / math.sqrt(fdot(lam, lam) * fdot(x, x))
        say(f"checkpoint {ck}: <lam, F x> = {lhs:.15e}, <F^T lam, x> = {rhs:.15e}, rel {rel:.2e}, grad_v {gv}")
        assert rel <= 1e-13
        grads.append((gu, gv))
    for gu, gv in grads[1:]:
        assert np.array_equal(gu.view(np.uint64), grads[0][0].view(np.uint64))
        # the v partials of every step are summed in one fixed order
        assert np.array_equal(np.asarray(gv).view(np.uint64), np.asarray(grads[0][1]).view(np.uint64))


def test_pass_graphs_replay_bit_exact(amr, torch):
    """The forward replay and the adjoint pass run as hipGraphs captured on
    the first call; a new velocity only rewrites the per-step constants the
    kernels read.  Every replay, with other passes (at the other velocity)
    and the program's own eager work between the replays, gives the bits of
    the first run at that velocity."""
    say("case: graph replays, 40 steps (regrid every 10, checkpoint 7), velocities (1, 0.5) / (0.7, -0.2)")
    tree, solver, tape, u0, _ = record_c5(amr, 40, 7)
    info = tape.info()
    rng = np.random.default_rng(9)
    x_t = torch.from_numpy(u0).cuda()
    lam_t = torch.from_numpy(rng.standard_normal(info["final_leaves"] * 256)).cuda()

    def both(vel):
        uT = replay(torch, tape, x_t, vel).cpu().numpy()
        g = torch.empty_like(x_t)
        gv = tape.adjoint(lam_t.data_ptr(), g.data_ptr())
        return uT, g.cpu().numpy(), np.asarray(gv)

    first = {v: both(v) for v in ((1.0, 0.5), (0.7, -0.2))}
    for r in range(3):
        for v in ((1.0, 0.5), (0.7, -0.2), (1.0, 0.5)):
            junk = torch.randn(1 << 20, dtype=torch.float64, device="cuda")
            (junk * 2.0).sum().item()
            got = both(v)
            for a, b in zip(got, first[v]):
                assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (r, v)
    # the two velocities give different results (the constants reach the replays)
    assert not np.array_equal(first[(1.0, 0.5)][0], first[(0.7, -0.2)][0])
    tape.close()
    tree.close()


def test_autograd_rejects_stale_tape(amr, torch):
    from gpu_amr_amd.adjoint import advection_solve

    _, _, tape, u0, _ = record_c5(amr, 12, 5)
    u0_t = torch.from_numpy(u0).cuda().requires_grad_(True)
    v = torch.tensor((1.0, 0.5), dtype=torch.float64, requires_grad=True)
    out = advection_solve(u0_t, v, tape)
    replay(torch, tape, u0_t.detach(), (1.0, 0.5))  # replays the tape in between
    with pytest.raises(RuntimeError, match="replayed"):
        out.sum().backward()


def test_adjoint_c5_finite_differences(amr, torch):
    """C5: gradient of the misfit w.r.t. u0 and v against central differences
    (h = 1e-6) on 32 random joint directions (seed 11); the GPU replay that
    the differences use is the oracle's solve bit for bit (checked at both
    perturbed points of two directions, and the oracle's own differences are
    then the same numbers)."""
    from gpu_amr_amd.adjoint import misfit

    say("case: C5 record 300 steps, checkpoints every 20")
    t0 = time.perf_counter()
    tree, solver, tape, u0, history = record_c5(amr, 300, 20)
    info = tape.info()
    say(f"tape {info} recorded in {time.perf_counter() - t0:.1f}s")
    vel = (1.0, 0.5)
    u0_t = torch.from_numpy(u0).cuda()
    # target: forward solve from u0 + 1e-3 N(0,1) (seed 7)
    pert = 1e-3 * np.random.default_rng(7).standard_normal(u0.size)
    u_target = replay(torch, tape, torch.from_numpy(u0 + pert).cuda(), vel)
    # gradient through the autograd harness
    u0_v = u0_t.clone().requires_grad_(True)
    v_v = torch.tensor(vel, dtype=torch.float64, requires_grad=True)
    J = misfit(u0_v, v_v, tape, u_target)
    t0 = time.perf_counter()
    J.backward()
    say(f"J = {J.item():.6e}; adjoint pass {1e3 * (time.perf_counter() - t0):.1f} ms")
    gu = u0_v.grad.cpu().numpy()
    gv = v_v.grad.cpu().numpy()
    say("dJ/dv =", gv)

    def J_at(du, dv):
        with torch.no_grad():
            return misfit(u0_t + du, torch.tensor(np.array(vel) + dv, dtype=torch.float64), tape, u_target).item()

    def central(du_t, dv, h):
        return (J_at(h * du_t, h * dv) - J_at(-h * du_t, -h * dv)) / (2 * h)

    # J is exactly quadratic in u0, so central differences along u0 have no
    # truncation error; along v they carry h^2 J3 / 6 with J3 the third
    # derivative, about 3 <u_v, u_vv> ~ 1e6 here (the misfit is small while
    # the state moves fast with v): a few 1e-7 absolute at h = 1e-6.  One
    # Richardson step over the central differences at h and h/2 removes it.
    rng = np.random.default_rng(11)
    h = 1e-6
    worst = worst_plain = 0.0
    dirs = []
    for i in range(32):
        du = rng.standard_normal(u0.size)
        dv = rng.standard_normal(2)
        dirs.append((du, dv))
        du_t = torch.from_numpy(du).cuda()
        f1 = central(du_t, dv, h)
        f2 = central(du_t, dv, h / 2)
        fd = (4 * f2 - f1) / 3
        ad = fdot(gu, du) + fdot(gv, dv)
        rel = abs(fd - ad) / max(abs(fd), abs(ad))
        worst = max(worst, rel)
        worst_plain = max(worst_plain, abs(f1 - ad) / max(abs(f1), abs(ad)))
        assert rel <= 1e-6, f"direction {i}: fd {fd:.12e} (h: {f1:.12e}) adjoint {ad:.12e} rel {rel:.2e}"
    say(f"32 joint directions: worst relative difference {worst:.2e} (plain central h=1e-6: {worst_plain:.2e})")
    # u0-only directions: plain central differences at h = 1e-6
    for i in range(8):
        du = rng.standard_normal(u0.size)
        fd = central(torch.from_numpy(du).cuda(), np.zeros(2), h)
        ad = fdot(gu, du)
        rel = abs(fd - ad) / max(abs(fd), abs(ad))
        assert rel <= 1e-6, f"u0 direction {i}: fd {fd:.12e} adjoint {ad:.12e} rel {rel:.2e}"
    # velocity components
    for d in range(2):
        dv = np.zeros(2)
        dv[d] = 1.0
        zero = 0.0 * u0_t
        fd = (4 * central(zero, dv, h / 2) - central(zero, dv, h)) / 3
        rel = abs(fd - gv[d]) / max(abs(fd), abs(gv[d]))
        say(f"dJ/dv_{d}: fd {fd:.12e} adjoint {gv[d]:.12e} rel {rel:.2e}")
        assert rel <= 1e-6
    # the replays the differences used are the oracle's solves
    for i in range(2):
        du, dv = dirs[i]
        for s in (1.0, -1.0):
            ug = replay(torch, tape, u0_t + torch.from_numpy(s * h * du).cuda(), tuple(np.array(vel) + s * h * dv))
            uo = oracle_replay(tape, history, u0 + s * h * du, tuple(np.array(vel) + s * h * dv))
            assert np.array_equal(ug.cpu().numpy().view(np.uint64), uo.view(np.uint64)), f"direction {i} sign {s}"
    say("oracle replays bit-identical at 4 perturbed points")


def test_adjoint_3d(amr, torch):
    """3D 4^3 patches, levels 2..4, gradient criterion every 4 steps: dot
    product in u0 and central differences in v."""
    say("case: 3D adjoint, 16 steps")
    tree = amr.NdTree(3, 4, 4, 1, 4096 + 64)
    for _ in range(2):
        tree.refine_all()
    solver = amr.AmrSolver(tree, (0.7, -0.4, 0.3), 0.3)

    def ic(c):
        return np.exp(-((c[..., 0] - 0.4) ** 2 + (c[..., 1] - 0.6) ** 2 + (c[..., 2] - 0.5) ** 2) / (2 * 0.08**2))

    solver.initialize(ic)
    crit = amr.make_criterion(amr.CRITERION_GRADIENT, 0.05, 0.01, 2, 4)
    tree.reconstruct_tree(crit)
    solver.initialize(ic)
    u0 = tree.get_interior().reshape(-1).copy()
    tape = amr.Tape(solver, 16, crit, regrid_every=4, checkpoint_every=5)
    info = tape.info()
    say("tape:", info)
    assert info["topologies"] >= 2
    rng = np.random.default_rng(9)
    x = rng.standard_normal(u0.size)
    lam = rng.standard_normal(info["final_leaves"] * 64)
    vel = (0.7, -0.4, 0.3)
    fx = replay(torch, tape, torch.from_numpy(x).cuda(), vel).cpu().numpy()
    # linearise around u0 for the velocity gradient
    uT = replay(torch, tape, torch.from_numpy(u0).cuda(), vel)
    g = torch.empty(u0.size, dtype=torch.float64, device="cuda")
    lam_t = torch.from_numpy(lam).cuda()
    gv = tape.adjoint(lam_t.data_ptr(), g.data_ptr())
    lhs, rhs = fdot(lam, fx), fdot(g.cpu().numpy(), x)
    assert abs(lhs - rhs) <= 1e-13 * math.sqrt(fdot(lam, lam) * fdot(x, x))
    def central(d, h):
        vp, vm = np.array(vel), np.array(vel)
        vp[d] += h
        vm[d] -= h
        fp = fdot(lam, replay(torch, tape, torch.from_numpy(u0).cuda(), tuple(vp)).cpu().numpy())
        fm = fdot(lam, replay(torch, tape, torch.from_numpy(u0).cuda(), tuple(vm)).cpu().numpy())
        return (fp - fm) / (2 * h)

    h = 1e-6
    for d in range(3):
        fd = (4 * central(d, h / 2) - central(d, h)) / 3  # one Richardson step
        rel = abs(fd - gv[d]) / max(abs(fd), abs(gv[d]))
        say(f"dim {d}: fd {fd:.12e} adjoint {gv[d]:.12e} rel {rel:.2e}")
        assert rel <= 1e-6
    del uT
