#!/usr/bin/env python3
"""Headline benchmark: FP64 cell-updates/s of the patch-AMR advection step.

Workload (one GPU): 3D linear advection, first-order upwind FVM on a static
uniform periodic octree of depth 5 -- 32768 leaf patches of 8^3 FP64 cells
(16.8M cells), one ghost layer, seeded Gaussian bump, velocity (1, 0.5, 0.25),
CFL 0.3.  A step is one launch of the fused halo-gather + upwind kernel over
all leaves with the state resident in HBM.

Prints one JSON line with the throughput, the roofline of the step kernel
(16 algorithmic bytes per cell update: one FP64 read of u^n, one FP64 write of
u^(n+1)) measured with HIP events on the context stream, the CPU oracle timed
on this host as a baseline, the batched form (64 independent 2D instances,
one graph launch) against 64 sequential plain solves, and the discrete
adjoint of the 2D active-AMR solve against its forward pass.
"""
from __future__ import annotations

import argparse
import ctypes
import importlib.util
import json
import os
import pathlib
import re
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
BYTES_PER_CELL_UPDATE = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def package():
    if "gpu_amr_amd" in sys.modules:
        return sys.modules["gpu_amr_amd"]
    pkg = ROOT / "gpu-amr_amd"
    spec = importlib.util.spec_from_file_location("gpu_amr_amd", pkg / "__init__.py", submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gpu_amr_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class HipEvents:
    """hipEvent timing on an arbitrary stream (the context's own)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        for fn in ("hipEventCreate", "hipEventRecord", "hipEventSynchronize", "hipEventElapsedTime", "hipEventDestroy"):
            getattr(self.hip, fn).restype = ctypes.c_int

    def create(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return float(ms.value)


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` from its committed PMC summary."""
    p = ROOT / "profiles" / f"{kernel}_c3_traffic.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def trace_from_profiles(kernel):
    """(average launch us, file) of `kernel` from the committed kernel trace
    of this benchmark's default command (rocprofv3 --kernel-trace --stats)."""
    p = ROOT / "profiles" / f"{kernel}_c3_bench_kernel_trace.md"
    if not p.exists():
        return None, None
    m = re.search(r"avg_launch_us:\s*([0-9.eE+]+)", p.read_text())
    return (float(m.group(1)) if m else None), f"profiles/{p.name}"


def cpu_baseline(steps, threads):
    """The CPU oracle (the reference algorithm restated, padded patches,
    per-step halo exchange) on the same C3 workload, `steps` steps."""
    if str(ROOT) not in sys.path:
        sys.path.insert(0, str(ROOT))
    from oracle import OracleTree

    from gpu_amr_amd import workloads

    o = OracleTree(3, 8, 1, 5, 36864 + 64)
    o.set_params((1.0, 0.5, 0.25), 0.3, threads=threads)
    t0 = time.perf_counter()
    for _ in range(5):
        o.refine_all()
    o.set_interior(workloads.dense_to_leaves(package(), o.keys(), workloads.c3_dense_ic(), 8, 5))
    o.halo_exchange()
    log(f"cpu baseline: tree built in {time.perf_counter() - t0:.1f}s, timing {steps} steps on {threads} threads")
    t0 = time.perf_counter()
    o.advance_batch(steps)
    dt = time.perf_counter() - t0
    return steps * o.size() * 512 / dt, dt


def host_cpu():
    """This host's CPU model and logical CPU count, and the threads the CPU
    baselines use: the job's CPU allotment (OMP_NUM_THREADS, else 16), at
    most the logical CPUs present."""
    model = "unknown"
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    allot = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return {"model": model, "nproc": nproc, "threads": max(1, min(allot, nproc)),
            "threads_rule": "min(OMP_NUM_THREADS or 16, nproc): the job's CPU allotment"}


def calib_lib():
    lib = ctypes.CDLL(str(ROOT / "gpu-amr_amd" / "libamr_calib.so"))
    lib.amr_calib_copy.restype = ctypes.c_int
    lib.amr_calib_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    return lib


def copy_kernel(n_doubles, reps, ev=None):
    """A plain copy of n_doubles FP64 values (8 bytes read and 8 written per
    value) with the step kernels' access pattern, `reps` times on its own
    stream; returns the per-copy time in ms when `ev` (HipEvents) is given.
    Its FETCH_SIZE / WRITE_SIZE per dispatch calibrate the HBM counters."""
    hip = ctypes.CDLL("libamdhip64.so")
    lib = calib_lib()
    nbytes = n_doubles * 8
    a, b, st = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert lib.amr_calib_copy(a, b, n_doubles, st) == 0  # warm
    ms = None
    if ev is not None:
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, st.value)
    for _ in range(reps):
        assert lib.amr_calib_copy(a, b, n_doubles, st) == 0
    if ev is not None:
        ev.record(e1, st.value)
        ms = ev.elapsed_ms(e0, e1) / reps
    assert hip.hipStreamSynchronize(st) == 0
    hip.hipStreamDestroy(st)
    hip.hipFree(a)
    hip.hipFree(b)
    return ms


def copy_calibration(mib, reps):
    n = (mib << 20) // 8
    copy_kernel(n, reps)
    print(json.dumps({"copy_calibration_bytes_read_per_dispatch": n * 8, "bytes_written_per_dispatch": n * 8,
                      "reps": reps}), flush=True)


def batched_section(amr, instances, steps):
    """C4: `instances` independent 2D trees (depth 3..7, own velocity and dt)
    stepped as one graph launch vs the same solves one after another with the
    plain per-tree kernel; results must agree bit for bit."""
    from gpu_amr_amd import workloads

    t0 = time.perf_counter()
    built = [workloads.c4_instance(amr, i) for i in range(instances)]
    trees = [b[0] for b in built]
    solvers = [b[1] for b in built]
    leaves = [t.size() for t in trees]
    log(f"batched: {instances} instances, {sum(leaves)} leaves (min {min(leaves)}, max {max(leaves)}), built in {time.perf_counter() - t0:.1f}s")
    batch = amr.Batch(trees)  # copies the initial state
    variant, _ = batch.autotune(10)
    batch.advance(steps, use_graph=True)  # graph capture + first run
    batch.synchronize()
    # sequential plain solves of the same instances (one launch per step each)
    t0 = time.perf_counter()
    for s in solvers:
        s.advance_batch_async(steps)
        s.finish_advance_batch()
    seq = time.perf_counter() - t0
    seq_out = [t.get_interior().copy() for t in trees]
    batch.scatter()
    same = all(np.array_equal(t.get_interior().view(np.uint64), u.view(np.uint64)) for t, u in zip(trees, seq_out))
    # timed: replay of the captured graph (`steps` is even, so the buffer
    # parity matches the capture)
    t0 = time.perf_counter()
    batch.advance(steps, use_graph=True)
    batch.synchronize()
    bat = time.perf_counter() - t0
    batch.close()
    cells = sum(leaves) * 256 * steps
    log(f"batched: sequential {seq * 1e3:.1f} ms, batched {bat * 1e3:.1f} ms, speedup {seq / bat:.2f}x, bit-identical {same}")
    return {
        "workload": f"{instances} independent 2D instances, 16^2 FP64 patches, depth 3..7, own velocity, CFL 0.4",
        "instances": instances,
        "steps": steps,
        "leaves": int(sum(leaves)),
        "sequential_ms": seq * 1e3,
        "batched_ms": bat * 1e3,
        "speedup": seq / bat,
        "batched_cell_updates_per_s": cells / bat,
        "bit_identical": bool(same),
        "step_variant": variant,
    }


def adjoint_section(amr, steps=300, checkpoint_every=20, reps=3):
    """C5: the recorded 2D active-AMR solve (levels 3..6, gradient criterion
    every 10 steps); time of one full forward replay vs one full adjoint pass
    (with the recompute between checkpoints), best of `reps`."""
    import torch

    from gpu_amr_amd import workloads
    from gpu_amr_amd.adjoint import advection_solve

    tree, solver, crit, _, _ = workloads.c5_setup(amr)
    u0 = torch.from_numpy(tree.get_interior().reshape(-1).copy()).cuda()
    t0 = time.perf_counter()
    tape = amr.Tape(solver, steps, crit, regrid_every=10, checkpoint_every=checkpoint_every)
    rec = time.perf_counter() - t0
    info = tape.info()
    vel = torch.tensor((1.0, 0.5), dtype=torch.float64)
    fwd, adj = [], []
    lam = torch.randn(info["final_leaves"] * info["cells_per_leaf"], dtype=torch.float64, device="cuda")
    g = torch.empty_like(u0)
    for _ in range(reps):
        t0 = time.perf_counter()
        with torch.no_grad():
            advection_solve(u0, vel, tape)
        fwd.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        tape.adjoint(lam.data_ptr(), g.data_ptr())
        adj.append(time.perf_counter() - t0)
    f, a = min(fwd), min(adj)
    # the same passes with a new velocity on every call (an optimisation loop
    # or a finite-difference sweep over v): the captured graphs are replayed
    # with the new per-step constants
    fwd_v, adj_v = [], []
    for r in range(reps):
        vr = torch.tensor((1.0 + 1e-3 * (r + 1), 0.5 - 1e-3 * (r + 1)), dtype=torch.float64)
        t0 = time.perf_counter()
        with torch.no_grad():
            advection_solve(u0, vr, tape)
        fwd_v.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        tape.adjoint(lam.data_ptr(), g.data_ptr())
        adj_v.append(time.perf_counter() - t0)
    topo_leaves = [len(tape.keys(k)) for k in range(info["topologies"])]
    log(f"adjoint: {info['steps']} steps, {info['topologies']} topologies, leaves initial {info['initial_leaves']}, "
        f"min {min(topo_leaves)}, max {max(topo_leaves)}, final {info['final_leaves']}; "
        f"forward {f * 1e3:.1f} ms, adjoint {a * 1e3:.1f} ms, ratio {a / f:.2f}; with a new velocity per call "
        f"forward {min(fwd_v) * 1e3:.1f} ms, adjoint {min(adj_v) * 1e3:.1f} ms")
    tape.close()
    tree.close()
    return {
        "workload": "C5: 2D active AMR, 16^2 FP64 patches, levels 3..6, gradient criterion every 10 steps, 300 steps",
        "steps": info["steps"],
        "checkpoint_every": info["checkpoint_every"],
        "topologies": info["topologies"],
        "regrids": info["regrids"],
        "leaves_initial": info["initial_leaves"],
        "leaves_min": min(topo_leaves),
        "leaves_max": max(topo_leaves),
        "leaves_final": info["final_leaves"],
        "record_ms": rec * 1e3,
        "forward_ms": f * 1e3,
        "adjoint_ms": a * 1e3,
        "adjoint_over_forward": a / f,
        "forward_ms_new_velocity": min(fwd_v) * 1e3,
        "adjoint_ms_new_velocity": min(adj_v) * 1e3,
    }


def c2_section(amr, cycles=50, cpu_cycles=50, threads=16):
    """C2, 2D linear advection with active AMR: 16^2 FP64 patches from
    uniform depth 3, the 8-bump field (seed 2026), v = (1, 0.5), CFL 0.4, a
    regrid by the undivided gradient (refine > 0.05, coarsen < 0.01, levels
    3..7, 2:1) then 10 steps, 50 times (500 steps).  Cell-updates/s over the
    loop's wall-clock (regrids included) and over the step launches alone
    (HIP events around each batch of 10 steps on the context's stream); the
    oracle on the first `cpu_cycles` cycles as the CPU baseline."""
    from gpu_amr_amd import workloads

    g = amr.NdTree(2, 7, 16, 1, 16384 + 64)
    s = amr.AmrSolver(g, (1.0, 0.5), 0.4)
    for _ in range(3):
        g.refine_all()
    ic = workloads.c2_mixture()
    s.initialize(ic)
    crit = amr.make_criterion(amr.CRITERION_GRADIENT, 0.05, 0.01, 3, 7)
    ev = HipEvents()
    marks = [(ev.create(), ev.create()) for _ in range(cycles)]
    st = g.stream()
    g.synchronize()
    upd, lo, hi = 0, g.size(), g.size()
    t0 = time.perf_counter()
    for it in range(cycles):
        g.reconstruct_tree(crit)
        n = g.size()
        ev.record(marks[it][0], st)
        s.advance_batch_async(10)
        ev.record(marks[it][1], st)
        _, ex = s.finish_advance_batch()
        upd += ex * n * 256
        lo, hi = min(lo, n), max(hi, n)
    wall = time.perf_counter() - t0
    step_s = sum(ev.elapsed_ms(a, b) for a, b in marks) * 1e-3
    log(f"c2: {cycles * 10} steps / {cycles} regrids, leaves {lo}..{hi}, {wall * 1e3:.1f} ms, "
        f"{upd / wall:.3e} cell-updates/s incl. regrid, {upd / step_s:.3e} in the steps ({step_s * 1e3:.1f} ms)")
    g.close()

    if str(ROOT) not in sys.path:
        sys.path.insert(0, str(ROOT))
    from oracle import OracleTree

    o = OracleTree(2, 16, 1, 7, 2 * 16384 + 128)
    o.set_params((1.0, 0.5), 0.4, threads=threads)
    for _ in range(3):
        o.refine_all()
    o.set_interior(ic(o.cell_centers()))
    o.halo_exchange()
    oupd = 0
    t0 = time.perf_counter()
    for _ in range(cpu_cycles):
        o.reconstruct(o.criterion(1, 0.05, 0.01, 3, 7))
        o.halo_exchange()
        n = o.size()
        o.advance_batch(10)
        oupd += 10 * n * 256
    cwall = time.perf_counter() - t0
    log(f"c2 cpu: {cpu_cycles * 10} steps in {cwall:.2f} s, {oupd / cwall:.3e} cell-updates/s on {threads} threads")
    return {
        "workload": "C2: 2D linear advection with active AMR, 16^2 FP64 patches, depth 3..7, 8 Gaussian bumps "
                    "(seed 2026), v = (1, 0.5), CFL 0.4, regrid every 10 steps by the undivided gradient "
                    "(0.05 / 0.01), 500 steps",
        "metric": "cell-updates/s",
        "value": upd / wall,
        "value_steps_only": upd / step_s,
        "wall_ms": wall * 1e3,
        "steps_ms": step_s * 1e3,
        "leaves_min": lo,
        "leaves_max": hi,
        "cpu_baseline": {"value": oupd / cwall, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                         "sample": f"first {cpu_cycles} regrid cycles ({cpu_cycles * 10} steps) of the same loop, "
                                   f"{cwall:.2f} s"},
    }


def euler_section(amr, name="3d", cpu_cycles=6, threads=16):
    """One of the reference's Euler benchmarks (`name`: "3d", "2d",
    "2d_active"; bench_fvm_solver_integration{3D,,_active_amr}.b.cpp): the
    pulse(s) on the adaptive tree, advance_batch_async(every, tmax - t),
    regrid by density, finish_advance_batch, until tmax; its own metric,
    cell-updates/s including the regrids (executed steps x leaves before the
    regrid x n^R, over the loop's wall-clock).  The CPU oracle runs the first
    `cpu_cycles` cycles of the same loop as the baseline."""
    from gpu_amr_amd import workloads

    solver, crit, cfg = workloads.euler_benchmark(amr, name)
    rank = cfg["rank"]
    tree = solver.tree
    ni = cfg["patch"] ** rank

    def loop(s, t_max, max_cycles):
        t, upd, steps, cycles, changed, lo, hi = 0.0, 0, 0, 0, 0, s.tree.size(), s.tree.size()
        while t < t_max and cycles < max_cycles:
            before = s.tree.size()
            s.advance_batch_async(cfg["every"], t_max - t)
            changed += bool(s.tree.reconstruct_tree(crit))
            dt, ex = s.finish_advance_batch()
            t += dt
            upd += ex * before * ni
            steps += ex
            cycles += 1
            lo, hi = min(lo, s.tree.size()), max(hi, s.tree.size())
        return t, upd, steps, cycles, changed, lo, hi

    tree.synchronize()
    t0 = time.perf_counter()
    t, upd, steps, cycles, changed, lo, hi = loop(solver, cfg["tmax"], 100000)
    wall = time.perf_counter() - t0
    rate = upd / wall
    log(f"euler {name}: t={t:.3f} in {steps} steps / {cycles} regrids ({changed} changed the tree), "
        f"leaves {lo}..{hi}, {wall * 1e3:.1f} ms, {rate:.3e} cell-updates/s (incl. regrid)")
    tree.close()

    # CPU baseline: the oracle on the first cpu_cycles cycles
    if str(ROOT) not in sys.path:
        sys.path.insert(0, str(ROOT))
    from oracle import OracleEuler, euler_prim_to_cons

    o = OracleEuler(rank, cfg["patch"], 1, cfg["depth"], 2 * cfg["capacity"] + 64, (1000.0,) * rank)
    o.set_params(1.4, 0.3, threads=threads)
    for _ in range(cfg["initial"]):
        o.refine_all()
    prim = workloads.euler_ic(cfg)(o.cell_centers())
    cons = euler_prim_to_cons(rank, prim.reshape(-1, rank + 2), 1.4).reshape(prim.shape)
    o.set_interior(np.moveaxis(cons, -1, 0))
    o.halo_exchange()
    ot, oupd, osteps = 0.0, 0, 0
    t0 = time.perf_counter()
    for _ in range(cpu_cycles):
        before = o.size()
        dt, ex = o.advance_batch(cfg["every"], cfg["tmax"] - ot)
        o.reconstruct(o.criterion(0, cfg["refine"], cfg["coarsen"], cfg["min_level"], cfg["max_level"]))
        o.halo_exchange()
        ot += dt
        oupd += ex * before * ni
        osteps += ex
    cwall = time.perf_counter() - t0
    log(f"euler {name} cpu: {osteps} steps in {cwall:.1f} s, {oupd / cwall:.3e} cell-updates/s on {threads} threads")
    ic = "two pulses" if cfg["ic"] == "two_pulses" else "acoustic pulse"
    return {
        "workload": f"reference Euler benchmark {cfg['source']}: {ic}, {cfg['patch']}^{rank} FP64 patches, "
                    f"domain 1000^{rank}, gamma 1.4, cfl 0.3, {cfg['initial']} refine-all, regrid every "
                    f"{cfg['every']} steps by density ({cfg['refine']:g} / {cfg['coarsen']:g}, levels "
                    f"{cfg['min_level']}..{cfg['max_level']}), tmax {cfg['tmax']:g}",
        "metric": "cell-updates/s including regrids",
        "value": rate,
        "t_end": t,
        "steps": steps,
        "regrids": cycles,
        "regrids_changing_topology": changed,
        "leaves_min": lo,
        "leaves_max": hi,
        "wall_ms": wall * 1e3,
        "cpu_baseline": {"value": oupd / cwall, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                         "sample": f"first {cpu_cycles} regrid cycles ({osteps} steps) of the same loop, {cwall:.1f} s"},
    }


def euler_uniform_section(amr, steps=20, rank=3):
    """The Euler step on a uniform tree: 3D, a depth-5 octree (32768 leaves
    of 8^3, domain 1000^3); 2D, a depth-5 quadtree (1024 leaves of 64^2,
    domain 1000^2); the benchmarks' acoustic pulse.  HIP-event time per step
    of advance_batch_async(steps) on the context stream (the step kernel, the
    per-step dt logic, and the dt reduction once per batch).  16 B x (R + 2)
    algorithmic bytes per cell-update (every field read and written)."""
    from gpu_amr_amd import workloads

    cfg = dict(workloads.EULER_CONFIGS["3d" if rank == 3 else "2d"])
    patch, leaves = (8, 32768) if rank == 3 else (64, 1024)
    s = amr.EulerSolver(rank, 5, patch, 1, leaves + 64, (1000.0,) * rank, 1.4, 0.3)
    for _ in range(5):
        s.tree.refine_all()
    s.initialize(workloads.euler_ic(cfg))
    cells = s.tree.size() * patch**rank
    s.advance_batch_async(3)
    s.finish_advance_batch()
    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    st = s.tree.stream()
    ev.record(e0, st)
    s.advance_batch_async(steps)
    ev.record(e1, st)
    _, ex = s.finish_advance_batch()
    us = ev.elapsed_ms(e0, e1) / ex * 1e3
    gbs = 16 * (rank + 2) * cells / (us * 1e-6) / 1e9
    log(f"euler uniform {rank}d: {s.tree.size()} leaves, {ex} steps, {us:.1f} us per step, {gbs:.0f} GB/s algorithmic")
    shape = f"{patch}^{rank}"
    tree = "octree" if rank == 3 else "quadtree"
    return {"workload": f"{rank}D Euler, uniform depth-5 {tree}, {s.tree.size()} leaves of {shape} FP64, acoustic pulse, "
                        f"domain 1000^{rank}",
            "steps": ex, "us_per_step": us, "cell_updates_per_s": cells / (us * 1e-6),
            "algorithmic_GBps": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--batched-instances", type=int, default=64)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB")
    ap.add_argument("--variant", default="", help="step-kernel variant to run instead of autotuning")
    ap.add_argument("--no-fusion", action="store_true", help="one step per launch (no two-step fusion)")
    ap.add_argument("--no-euler", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--euler-only", nargs="?", const="all", choices=["all", "3d", "2d", "2d_active"],
                    help="only the Euler benchmarks (profiling): all, or one of them")
    ap.add_argument("--c2-only", action="store_true", help="only the C2 active-AMR loop (profiling)")
    ap.add_argument("--euler-uniform", type=int, default=0, metavar="STEPS",
                    help="only the 3D Euler step on the uniform 32768-leaf tree (profiling)")
    ap.add_argument("--euler-uniform-2d", type=int, default=0, metavar="STEPS",
                    help="only the 2D Euler step on the uniform 1024-leaf 64^2 tree (profiling)")
    ap.add_argument("--no-native-oracle", action="store_true",
                    help="CPU baselines with the shipped oracle build instead of one compiled for this CPU")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the leaves would shard by contiguous "
              "Morton ranges with one ghost-face exchange per step).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, 5)
        return

    # the CPU baselines' oracle, compiled for this host's CPU while the GPU
    # sections run (started before anything touches the GPU)
    cpu_info = host_cpu()
    native = None
    want_cpu = not (args.no_cpu and args.no_c2 and args.no_euler)
    if want_cpu and not args.euler_only and not args.c2_only and not args.no_native_oracle and shutil.which("g++"):
        if str(ROOT) not in sys.path:
            sys.path.insert(0, str(ROOT))
        import oracle

        native = oracle.build_native(tempfile.mkdtemp(prefix="oracle_native_"))

    import torch

    amr = package()
    from gpu_amr_amd import workloads

    if args.euler_uniform or args.euler_uniform_2d:
        if args.euler_uniform:
            print(json.dumps({"euler_uniform_3d": euler_uniform_section(amr, args.euler_uniform)}), flush=True)
        if args.euler_uniform_2d:
            print(json.dumps({"euler_uniform_2d": euler_uniform_section(amr, args.euler_uniform_2d, 2)}), flush=True)
        return

    if args.c2_only:
        print(json.dumps({"c2": c2_section(amr, cpu_cycles=1)}), flush=True)
        return

    if args.euler_only:
        names = ["3d", "2d", "2d_active"] if args.euler_only == "all" else [args.euler_only]
        print(json.dumps({f"euler_{n}": euler_section(amr, n, cpu_cycles=1) for n in names}), flush=True)
        return

    g, s = workloads.build_c3(amr)
    leaves = g.size()
    cells = leaves * 512
    s.set_step_fusion(not args.no_fusion)
    fused = s.step_fusion() == (True, True)
    log(f"tree: {leaves} leaves, {cells} cells; warmup {args.warmup} steps")

    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    stream = g.stream()
    names = s.step_variants()
    per_variant = {}
    if args.variant:
        # one named variant only (profiling runs)
        best = names.index(args.variant)
        s.set_step_variant(best)
        how = "fixed"
    else:
        # per-variant timing of single-step launches (diagnostic), then the
        # autotuner's choice (used when a step runs alone)
        s.set_step_fusion(False)
        for v, name in enumerate(names):
            s.set_step_variant(v)
            # untimed steps first: right after a switch of kernel the first
            # ten or so launches run up to 30 % slower (a transient, not the
            # variant's steady rate)
            s.advance_batch_async(20)
            s.finish_advance_batch()
            g.synchronize()
            ev.record(e0, stream)
            s.advance_batch_async(20)
            ev.record(e1, stream)
            s.finish_advance_batch()
            per_variant[name] = round(ev.elapsed_ms(e0, e1) / 20 * 1e3, 2)
        best, best_us = s.autotune(20)
        s.set_step_fusion(not args.no_fusion)
        how = "autotuned"
        log(f"single-step variants (us/step): {per_variant}; autotuner -> {names[best]} ({best_us:.1f} us)")

    # the read+write streaming rate of a plain copy of the same state size
    copy_ms = copy_kernel(cells, 20, ev)
    copy_gbs = 2 * 8 * cells / (copy_ms * 1e-3) / 1e9
    log(f"plain copy of {8 * cells / 1e6:.0f} MB: {copy_ms * 1e3:.1f} us, {copy_gbs:.0f} GB/s")

    # the untimed warm-up steps, right before the timed ones
    s.advance_batch_async(args.warmup)
    s.finish_advance_batch()

    # timed: exactly args.steps steps.  Fused, the solver runs them as pairs
    # (one launch per two steps); an odd count ends with one single-step
    # launch.  The HIP events bracket the pairs only, so the launch time is
    # that of the fused kernel; the wall-clock covers every step.
    per_launch = 2 if fused and args.steps >= 2 else 1
    head = args.steps - (args.steps % per_launch)
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev.record(e0, stream)
    s.advance_batch_async(head)
    ev.record(e1, stream)
    _, executed = s.finish_advance_batch()
    if head < args.steps:
        s.advance_batch_async(args.steps - head)
        executed += s.finish_advance_batch()[1]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert executed == args.steps
    launch_ms = ev.elapsed_ms(e0, e1) / (head // per_launch)
    value = args.steps * cells / wall
    achieved = BYTES_PER_CELL_UPDATE * cells * per_launch / (launch_ms * 1e-3) / 1e9
    log(f"gpu: {wall * 1e3 / args.steps:.3f} ms/step wall, {launch_ms * 1e3:.1f} us/launch "
        f"({per_launch} step(s) per launch, {head // per_launch} launches timed), {achieved:.0f} GB/s algorithmic")
    g.close()
    fused = per_launch == 2

    if native is not None:
        # every CPU baseline below uses the oracle built for this CPU when
        # its compile succeeded
        proc, path = native
        rc = proc.wait(timeout=600)
        import oracle

        if rc == 0:
            oracle.load(path)
            cpu_info["oracle_build"] = "g++ -O3 -march=native -ffp-contract=off (compiled on this host)"
        else:
            log(f"native oracle build failed ({proc.stderr.read().decode()[-300:]}); shipped build used")
    cpu_info.setdefault("oracle_build", "g++ -O3 -march=x86-64-v3 -ffp-contract=off (shipped build)")
    threads = cpu_info["threads"]

    batched = None
    if not args.no_batched:
        batched = batched_section(amr, args.batched_instances, 200)

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_section(amr)

    c2 = None
    if not args.no_c2:
        c2 = c2_section(amr, threads=threads)

    euler = None
    if not args.no_euler:
        euler = {"3d": euler_section(amr, "3d", threads=threads),
                 "2d": euler_section(amr, "2d", cpu_cycles=2, threads=threads),
                 "2d_active": euler_section(amr, "2d_active", cpu_cycles=4, threads=threads)}

    cpu = None
    if not args.no_cpu:
        rate, secs = cpu_baseline(args.cpu_steps, threads)
        cpu = {
            "value": rate,
            "unit": "cell-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{args.cpu_steps} steps of the same 32768-leaf 8^3 workload (step + halo exchange, padded patches), {secs:.1f} s",
            "host": cpu_info,
        }
        log(f"cpu: {rate:.3e} cell-updates/s on {threads} threads of {cpu_info['nproc']} "
            f"({cpu_info['model']}; {cpu_info['oracle_build']})")

    if fused:
        kernel = "k_step2_3d8_uniform"
        kernel_desc = ("k_step2_3d8_uniform: two steps per launch (u^n read once, u^(n+2) written once), "
                       "one workgroup per 2x2x1 block of leaves")
    else:
        kernel = "k_step_3d8"
        kernel_desc = f"single-step kernel variant '{names[best]}' ({how})"
    traffic = traffic_from_profiles(kernel)
    prof_us, prof_file = trace_from_profiles(kernel)
    algo_bytes = BYTES_PER_CELL_UPDATE * cells * per_launch
    out = {
        "metric": "FP64 cell-updates/s (3D uniform octree, % of HBM roofline)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "3D uniform octree depth 5: 32768 leaves of 8^3 FP64 cells (16.8M), halo 1, v=(1,0.5,0.25), CFL 0.3, periodic",
            "leaves": leaves,
            "cells": cells,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": kernel_desc,
            "launch_us": launch_ms * 1e3,
            "steps_per_launch": per_launch,
            "algorithmic_bytes_per_launch": algo_bytes,
            # HBM bytes the counters see per launch over this run's launch time
            "dram_GBps": traffic / (launch_ms * 1e-3) / 1e9 if traffic else None,
            "dram_frac": traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
            # the same kernel in the committed trace of this command
            "profile": {"file": prof_file, "launch_us": prof_us,
                        "frac": algo_bytes / (prof_us * 1e-6) / 1e9 / HBM_PEAK_GBS if prof_us else None},
            "single_step_variants_us": per_variant,
            "copy_GBps": copy_gbs,
            "frac_of_copy": achieved / copy_gbs,
        },
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adjoint,
        "c2": c2,
        "euler": euler,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
