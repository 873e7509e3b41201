#!/usr/bin/env python3
"""Headline benchmark: FP64 cell-updates/s of the patch-AMR advection step.

Workload (one GPU): 3D linear advection, first-order upwind FVM on a static
uniform periodic octree of depth 5 -- 32768 leaf patches of 8^3 FP64 cells
(16.8M cells), one ghost layer, seeded Gaussian bump, velocity (1, 0.5, 0.25),
CFL 0.3.  A step is one launch of the fused halo-gather + upwind kernel over
all leaves with the state resident in HBM.

Prints one JSON line with the throughput, the roofline of the step kernel
(16 algorithmic bytes per cell update: one FP64 read of u^n, one FP64 write of
u^(n+1)) measured with HIP events on the context stream, and the CPU oracle
timed on this host as a baseline.
"""
from __future__ import annotations

import argparse
import ctypes
import importlib.util
import json
import os
import pathlib
import re
import sys
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


def c3_dense_ic(ncell=256, seed=42):
    """Seeded 3D Gaussian bump (amplitude 1, sigma 0.1, centre U[.25,.75]^3)
    evaluated at cell centres of the uniform 256^3 grid; array order z, y, x."""
    c = np.random.default_rng(seed).uniform(0.25, 0.75, 3)
    h = 1.0 / ncell
    x = (np.arange(ncell) * h) + 0.5 * h
    gx = np.exp(-((x - c[0]) ** 2) / (2 * 0.1**2))
    gy = np.exp(-((x - c[1]) ** 2) / (2 * 0.1**2))
    gz = np.exp(-((x - c[2]) ** 2) / (2 * 0.1**2))
    return gz[:, None, None] * gy[None, :, None] * gx[None, None, :]


def dense_to_leaves(amr, keys, dense, n, depth, rank=3):
    coords, lv = amr.decode_keys(keys, rank)
    s = 1 << (depth - int(lv[0]))
    nb = dense.shape[0] // n
    blocks = dense.reshape(nb, n, nb, n, nb, n).transpose(0, 2, 4, 1, 3, 5)
    bz, by, bx = coords[:, 2] // s, coords[:, 1] // s, coords[:, 0] // s
    return np.ascontiguousarray(blocks[bz, by, bx])


def traffic_from_profiles():
    """HBM bytes per step-kernel launch from the committed PMC summary."""
    p = ROOT / "profiles" / "k_step_3d_c3_traffic.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(steps, threads):
    """The CPU oracle (the reference algorithm restated, padded patches,
    per-step halo exchange) on the same C3 workload, `steps` steps."""
    if str(ROOT) not in sys.path:
        sys.path.insert(0, str(ROOT))
    from oracle import OracleTree

    o = OracleTree(3, 8, 1, 5, 36864 + 64)
    o.set_params((1.0, 0.5, 0.25), 0.3, threads=threads)
    t0 = time.perf_counter()
    for _ in range(5):
        o.refine_all()
    amr = package()
    o.set_interior(dense_to_leaves(amr, o.keys(), c3_dense_ic(), 8, 5))
    o.halo_exchange()
    log(f"cpu baseline: tree built in {time.perf_counter() - t0:.1f}s, timing {steps} steps on {threads} threads")
    t0 = time.perf_counter()
    o.advance_batch(steps)
    dt = time.perf_counter() - t0
    return steps * o.size() * 512 / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, sharding).", file=sys.stderr)
        sys.exit(2)

    amr = package()
    depth, n = 5, 8
    g = amr.NdTree(3, depth, n, 1, 32768)
    s = amr.AmrSolver(g, (1.0, 0.5, 0.25), 0.3)
    for _ in range(depth):
        g.refine_all()
    leaves = g.size()
    cells = leaves * n**3
    g.set_interior(dense_to_leaves(amr, g.keys(), c3_dense_ic(), n, depth))
    log(f"tree: {leaves} leaves, {cells} cells; warmup {args.warmup} steps")
    s.advance_batch_async(args.warmup)
    s.finish_advance_batch()

    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    stream = g.stream()
    g.synchronize()
    t0 = time.perf_counter()
    ev.record(e0, stream)
    s.advance_batch_async(args.steps)
    ev.record(e1, stream)
    _, executed = s.finish_advance_batch()
    wall = time.perf_counter() - t0
    kernel_ms = ev.elapsed_ms(e0, e1) / args.steps
    assert executed == args.steps
    value = args.steps * cells / wall
    achieved = BYTES_PER_CELL_UPDATE * cells / (kernel_ms * 1e-3) / 1e9
    log(f"gpu: {wall * 1e3 / args.steps:.2f} ms/step wall, {kernel_ms * 1e3:.1f} us/step kernel, {achieved:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        rate, secs = cpu_baseline(args.cpu_steps, threads)
        cpu = {
            "value": rate,
            "unit": "cell-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{args.cpu_steps} steps of the same 32768-leaf 8^3 workload (step + halo exchange, padded patches), {secs:.1f} s",
        }
        log(f"cpu: {rate:.3e} cell-updates/s on {threads} threads")

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
            "traffic": traffic_from_profiles(),
            "kernel": "k_step<3,8,256>",
            "kernel_us": kernel_ms * 1e3,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
