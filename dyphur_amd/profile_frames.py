"""Frame stepping for profilers: the single-scene frame (integrate ->
broadphase -> narrowphase -> solve) on the mixed-body workload with nothing
else in the process (no timing passes, no replays, no other workloads), so
that counter-collecting profilers see exactly the frame's kernels.

    python -m dyphur_amd.profile_frames --bodies 32768 --frames 600
    python -m dyphur_amd.profile_frames --copy-mib 256   # copy calibration

Frames are enqueued without host synchronisation; the last line is JSON with
the frame count, the final state hash and the last frame's pair and contact
counts.  With --copy-mib the program instead copies a buffer of that size
ten times (a kernel of known traffic, to scale the profiler's byte counters)."""
from __future__ import annotations

import argparse
import json

import numpy as np

from dyphur_amd import capi, scenes


def copy_calibration(mib: int) -> dict:
    s = capi.Stream()
    n = mib << 20
    a = capi.DeviceArray.from_numpy(np.ones(n // 4, np.float32), s)
    b = capi.DeviceArray.from_numpy(np.zeros(n // 4, np.float32), s)
    for _ in range(10):
        capi.check(capi.lib().dy_memcpy_d2d_async(capi.vp(b.ptr), capi.vp(a.ptr), n, s.h), "dy_memcpy_d2d_async")
    s.wait()
    return {"copy_calibration_bytes": n, "copies": 10}


def run_frames(bodies: int, frames: int, seed: int) -> dict:
    sc = scenes.random_mixed(bodies, seed=seed)
    w = capi.World(sc)
    for _ in range(frames):
        w.step_async()
    last = w.check()
    return {"frames": frames, "bodies": sc.n, "seed": seed, "state_fnv1a64": f"{w.fnv():016x}",
            "last_frame_pairs": int(last["n_pairs"]), "last_frame_contacts": int(last["n_contacts"])}


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--bodies", type=int, default=32768)
    ap.add_argument("--frames", type=int, default=600)
    ap.add_argument("--seed", type=int, default=20260002)
    ap.add_argument("--copy-mib", type=int, default=0)
    args = ap.parse_args()
    capi.load()
    out = copy_calibration(args.copy_mib) if args.copy_mib else run_frames(args.bodies, args.frames, args.seed)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
