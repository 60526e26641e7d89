#!/usr/bin/env python3
"""Frame-step benchmark on one MI355X.

Workload: a single scene of 32 768 dynamic bodies (2/3 boxes with half extents
U[0.1, 0.5] m, 1/3 spheres with radius U[0.1, 0.4] m, uniform positions in a
30 x 40 x 30 m volume, uniform random orientations, density 1000 kg/m^3, seed
20260002) over a static ground plate, FP32, 60 Hz, 1 substep, 10 solver
iterations.  A step is one frame: integrate -> LBVH broadphase -> pairs ->
narrowphase -> XPBD solve, through the C ABI exactly as the reference's
stress driver sequences it (examples/stress_test_blocks/main.cpp:264-273).

value = body-substeps/s = bodies (ground included) x frames / seconds over the
timed frames, with all inputs resident in HBM (no host<->device state copies
in the timed region).

Prints one JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "body-substeps/s on 1 MI355X: plain vs batched; fwd+adjoint ms/frame; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bodies", type=int, default=32768)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="budget of the CPU-oracle baseline sample (0 disables it)")
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs the frame step on one GPU only; --gpus {args.gpus} "
              f"is not supported (see DESIGN.md for how the path would shard).", file=sys.stderr)
        sys.exit(2)

    from dyphur_amd import capi, scenes

    capi.load()
    if capi.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    sc = scenes.random_mixed(args.bodies, seed=20260002, frames=args.steps)
    s = capi.Stream()
    w = capi.World(sc, s)
    init = w.bodies.host_arrays(sc)

    def reset():
        w.bodies.upload_state(init, s)
        s.wait()

    # Warm-up frames (untimed), then restart from the initial state.
    for _ in range(args.warmup):
        w.step()
    reset()

    e0, e1 = capi.Event(), capi.Event()
    s.wait()
    t0 = time.perf_counter()
    e0.record(s)
    pairs, contacts, colors = [], [], []
    for _ in range(args.steps):
        st = w.step()
        pairs.append(st["n_pairs"]); contacts.append(st["n_contacts"]); colors.append(st["n_colors"])
    e1.record(s)
    s.wait()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_ms(e1)
    final_hash = w.fnv()
    n = sc.n
    value = n * args.steps * sc.n_substeps / wall
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * wall / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"single scene, {args.bodies} mixed boxes/spheres + static ground, "
                        f"frames 0..{args.steps - 1} at 60 Hz, 1 substep, 10 iterations",
            "bodies": n, "frames": args.steps, "substeps": sc.n_substeps, "iterations": sc.n_iters,
            "seed": 20260002, "max_pairs": sc.max_pairs, "max_contacts": sc.max_contacts,
        },
        "stream_ms_per_step": round(gpu_ms / args.steps, 4),
        "load": {"pairs_mean": float(np.mean(pairs)), "contacts_mean": float(np.mean(contacts)),
                 "colors_max": int(np.max(colors)), "pairs_max": int(np.max(pairs)),
                 "contacts_max": int(np.max(contacts))},
        "final_state_fnv1a64": f"{final_hash:016x}",
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
