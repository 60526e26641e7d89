#!/usr/bin/env python3
"""Frame-step benchmark on one MI355X.

Workload: a single scene of 32 768 dynamic bodies (2/3 boxes with half extents
U[0.1, 0.5] m, 1/3 spheres with radius U[0.1, 0.4] m, uniform positions in a
30 x 40 x 30 m volume, uniform random orientations, density 1000 kg/m^3, seed
20260002) over a static ground plate, FP32, 60 Hz, 1 substep, 10 solver
iterations, 600 frames.  A step is one frame: integrate -> LBVH broadphase ->
pairs -> narrowphase -> XPBD solve, through the C ABI in the reference's
stress-driver order (examples/stress_test_blocks/main.cpp:264-273), using the
async-capable Narrowphase::run overload so that no frame waits on the host.
The timed frames are the last --steps frames of the 600-frame run (the
frames before them run untimed), so a short run times the settled pile.  An
untimed pass times every kernel over the same frames (HIP events; the
roofline names the slowest), and an untimed replay of the whole run collects
the load statistics and must end in the same state.

value = body-substeps/s = bodies (ground included) x frames / seconds over the
timed frames, with all inputs resident in HBM (no host<->device state copies
in the timed region).

The line also carries "stress_c1": the reference's own stress_test_blocks
scene (1 024 boxes, 300 frames) on the GPU with and without the reference's
trajectory I/O, the CPU oracle at all host threads and at 4 (the reference
publishes 205.5 fps on 4 threads), the largest penetration after each solve
and the settled kinetic energy against the oracle's, plus the 0.85 m
variant; "batched": the 4096-scene pile workload (seed 20260003) stepped 300
frames in one launch, its body-substeps/s and the plain per-scene loop of the
drop-in API on a sample of the same scenes; "differentiable": the 256-scene
thrown-stack workload (seed 20260004), its forward and forward + adjoint
times per frame; "differentiable_at_scale": the same for 1 024 mixed scenes
of up to 1 024 bodies at 4 substeps per frame (seed 20260005); and
"articulated": the joint-chain workload.

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


C2_FRAMES = 600  # frames of the 32 768-body workload
C1_FRAMES = 300  # frames of the stress stack (examples/stress_test_blocks/main.cpp:52)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600,
                    help="timed frames: the last STEPS frames of the 600-frame 32 768-body run")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bodies", type=int, default=32768)
    ap.add_argument("--cpu-window", type=int, default=20,
                    help="frames per CPU-baseline window (four windows spread over the run)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched-frames", type=int, default=300,
                    help="frames of the 4096-scene batched workload (0 skips it)")
    ap.add_argument("--diff-frames", type=int, default=120,
                    help="frames of the differentiable 256-scene workload (0 skips it)")
    ap.add_argument("--diff-scale-frames", type=int, default=240,
                    help="frames of the 1 024-scene differentiable leg (0 skips it)")
    ap.add_argument("--c1-frames", type=int, default=C1_FRAMES,
                    help="frames of the stress_test_blocks leg (0 skips it)")
    ap.add_argument("--joint-frames", type=int, default=120,
                    help="frames of the articulated workload (0 skips it)")
    ap.add_argument("--plain-sample", type=int, default=64,
                    help="scenes of the batched workload stepped through the plain per-scene loop")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 10 device-to-device copies of MIB MiB (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    return ap.parse_args()


def copy_calibration(mib: int) -> None:
    from dyphur_amd import capi
    capi.load()
    s = capi.Stream()
    n = mib << 20
    a = capi.DeviceArray.from_numpy(np.ones(n // 4, np.float32), s)
    b = capi.DeviceArray.from_numpy(np.zeros(n // 4, np.float32), s)
    for _ in range(10):
        capi.check(capi.lib().dy_memcpy_d2d_async(capi.vp(b.ptr), capi.vp(a.ptr), n, s.h), "dy_memcpy_d2d_async")
    s.wait()
    print(json.dumps({"copy_calibration_bytes": n, "copies": 10}), flush=True)


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of `kernel` measured with rocprofv3 PMC counters
    (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a copy of known
    size), from profiles/traffic_<kernel>.json when that measurement exists."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    return d.get("hbm_bytes_per_launch")


def plan_frames(steps: int, config_frames: int, cpu_window: int, n_windows: int = 4) -> dict:
    """Which frames the single-scene leg times and where the CPU windows sit.

    The run is the workload's config_frames frames (more if steps asks for
    more); the timed frames are its last `steps` frames, so a short run times
    the settled pile rather than the free fall at the start.  The CPU baseline
    runs n_windows windows of `window` frames spread over the whole run, each
    from the GPU state at its first frame; every window ends inside the run."""
    steps = max(int(steps), 1)
    total = max(int(config_frames), steps)
    window = max(0, min(int(cpu_window), total // n_windows))
    starts = []
    if window:
        span = total - window
        starts = sorted({(k * span) // (n_windows - 1) for k in range(n_windows)}) if n_windows > 1 else [span]
    return {"total": total, "start": total - steps, "steps": steps, "window": window, "starts": starts}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def replay(w, init, s, total, want):
    """Untimed re-run of frames 0..total-1 from the initial state with
    per-frame statistics (host synchronisation every frame) and state
    snapshots at the frames in `want`."""
    w.bodies.upload_state(init, s)
    s.wait()
    stats, snaps = [], {}
    for f in range(total + 1):
        if f in want:
            snaps[f] = (w.bodies.download(s), w.fnv())
        if f < total:
            stats.append(w.step())
    return stats, snaps


def cpu_baseline(sc, snaps, starts, window, threads_alt: int = 4):
    """Time the CPU oracle (the reference algorithm restated in C++/OpenMP) on
    windows of `window` frames, each starting from the GPU trajectory's state
    at that frame (bit-identical to the oracle's own trajectory), so the
    sample covers both the falling and the settled phases without running
    every frame on the CPU.  The oracle's state hash at each window end is
    compared with the GPU's.  The last window is timed again on
    `threads_alt` threads."""
    from oracle import ref_cpu as rc

    rc.build()
    threads = rc.max_threads()
    ob, osh = rc.Bodies(sc), rc.Shapes(sc)
    ow = rc.World.for_scene(sc)

    def run(f0):
        st, _ = snaps[f0]
        for k in rc.BODY_FIELDS:
            ob.arr[k][:] = st[k]
        t0 = time.perf_counter()
        for _ in range(window):
            ow.frame(ob, osh, sc)
        return time.perf_counter() - t0, ob.fnv() == snaps[f0 + window][1]

    total, frames, equal = 0.0, 0, True
    for f0 in starts:
        dt, eq = run(f0)
        total += dt
        frames += window
        equal = equal and eq
    rc.set_threads(threads_alt)
    try:
        alt_s, eq = run(starts[-1])
    finally:
        rc.set_threads(threads)
    equal = equal and eq
    return {
        "value": round(sc.n * frames * sc.n_substeps / total, 1),
        "unit": "body-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{frames} frames of the same workload in {len(starts)} windows of {window} starting at frames "
                  f"{starts}, each from the GPU state at that frame; {total:.1f} s on {threads} OpenMP "
                  f"threads ({cpu_model()}); oracle state hash equal to the GPU at every window end: {equal}",
        "seconds": round(total, 2),
        "fnv_equal": equal,
        f"value_{threads_alt}_threads": round(sc.n * window * sc.n_substeps / alt_s, 1),
        f"sample_{threads_alt}_threads": f"the window at frame {starts[-1]} on {threads_alt} OpenMP threads "
                                         f"({alt_s:.1f} s)",
    }


def stage_kernels(capi, w, sc, s, n_frames: int) -> dict:
    """Per-kernel time and algorithmic bytes over n_frames frames from the
    world's current state (untimed: the stage events and the solver's kernel
    timing wait on the host every frame).  Returns {kernel: (ms per launch,
    bytes per launch)}; the broadphase's launches come from its own kernel
    timing, integrate and the narrowphase from events around their calls."""
    n = sc.n
    n_static = int(np.count_nonzero(sc.flags & 1))
    e = [capi.Event() for _ in range(4)]
    w.bp.set_timing(True)
    w.solver.set_timing(True)
    integ = narrow = 0.0
    P = C = 0.0
    try:
        for _ in range(n_frames):
            e[0].record(s)
            capi.integrate(s, w.bodies, sc.gravity, sc.dt, sc.n_substeps)
            e[1].record(s)
            w.bp.build_and_query(s, w.bodies, w.shapes, sc.bounds)
            pairs, count = w.bp.pairs_ptr()
            e[2].record(s)
            w.np.run_device_count(s, pairs, count, w.bp.max_pairs, w.bodies, w.shapes)
            e[3].record(s)
            w.solver.solve(s, w.np.contacts_view(), w.bodies, sc.dt, w.joints)
            s.wait()
            integ += e[0].elapsed_ms(e[1])
            narrow += e[2].elapsed_ms(e[3])
            P += min(w.bp.download_count(s), w.bp.max_pairs)
            C += w.np.download_count(s)
        bt = w.bp.timing()
        tm = w.solver.timing(s)
    finally:
        w.bp.set_timing(False)
        w.solver.set_timing(False)
    f = float(n_frames)
    P, C = P / f, C / f
    nb = float(max(bt["builds"], 1))
    ns = float(max(tm["solves"], 1))
    return {
        "integrate_kernel": (integ / f, 100.0 * (n - n_static) + 4.0 * n_static),
        "morton_sort_grid": (bt["sort_ms"] / nb, 48.0 * n),
        "karras_build": (bt["karras_ms"] / nb, 24.0 * n),
        "leaf_refit": (bt["refit_ms"] / nb, 108.0 * n),
        "traverse_count": (bt["traverse_ms"] / nb, 52.0 * n),
        "emit_pairs": (bt["emit_ms"] / nb, 8.0 * P),
        "narrowphase": (narrow / f, 84.0 * P + 36.0 * C),
        "assign_lists": (tm["assign_ms"] / ns, 12.0 * C + 16.0 * n),
        "sweep_flow": (tm["sweep_ms"] / ns, tm["sweep_bytes"] / ns),
    }


def roofline_of(kernels: dict) -> dict:
    """The dominant kernel (largest time per frame) with its achieved
    algorithmic GB/s against the HBM peak, and the whole frame's."""
    name = max(kernels, key=lambda k: kernels[k][0])
    ms, nbytes = kernels[name]
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    tot_ms = sum(v[0] for v in kernels.values())
    tot_b = sum(v[1] for v in kernels.values())
    return {
        "bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic_from_profiles(name),
        "avg_launch_ms": round(ms, 4), "algorithmic_bytes_per_launch": round(nbytes),
        "frame": {"kernel_ms": round(tot_ms, 4), "algorithmic_bytes": round(tot_b),
                  "achieved": round(tot_b / (tot_ms * 1e-3) / 1e9, 2) if tot_ms > 0 else 0.0},
        "kernels": {k: {"avg_launch_ms": round(v[0], 4), "algorithmic_bytes_per_launch": round(v[1]),
                        "achieved": round(v[1] / (v[0] * 1e-3) / 1e9, 2) if v[0] > 0 else 0.0}
                    for k, v in kernels.items()},
    }


def kinetic_energy(st: dict, sc) -> float:
    """Sum over dynamic bodies of m|v|^2/2 + w.(R I_body R^T w)/2 (float64)."""
    dyn = (sc.flags & 1) == 0
    m = sc.mass.astype(np.float64)[dyn]
    v = np.stack([st[k] for k in ("vel_x", "vel_y", "vel_z")], 1).astype(np.float64)[dyn]
    w = np.stack([st[k] for k in ("ang_x", "ang_y", "ang_z")], 1).astype(np.float64)[dyn]
    q = np.stack([st[k] for k in ("rot_w", "rot_x", "rot_y", "rot_z")], 1).astype(np.float64)[dyn]
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    a, b, c, d = q.T
    R = np.stack([1 - 2 * (c * c + d * d), 2 * (b * c - a * d), 2 * (b * d + a * c),
                  2 * (b * c + a * d), 1 - 2 * (b * b + d * d), 2 * (c * d - a * b),
                  2 * (b * d - a * c), 2 * (c * d + a * b), 1 - 2 * (b * b + c * c)], 1).reshape(-1, 3, 3)
    I = sc.inertia.astype(np.float64)[dyn].reshape(-1, 3, 3)
    wb = np.einsum("nji,nj->ni", R, w)          # body-frame angular velocity R^T w
    rot = 0.5 * np.einsum("ni,nij,nj->n", wb, I, wb)
    return float((0.5 * m * (v * v).sum(1) + rot).sum())


def run_driver(binary: str, prefix: str, frames: int, pitch: float | None = None) -> dict | None:
    """The C++ drop-in stress driver (the reference's main loop: a pair-count
    download every frame and a 7-array state download + trajectory write
    every second frame) as a child process; its fps and final state hash."""
    import subprocess
    cmd = [binary, prefix, str(frames)] + (["--pitch", str(pitch)] if pitch is not None else [])
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=180)
    except (OSError, subprocess.TimeoutExpired) as ex:
        return {"error": str(ex)}
    if r.returncode != 0:
        return {"error": f"rc {r.returncode}: {r.stdout[-400:]}"}
    try:
        m = json.load(open(prefix + ".metrics.json"))
        h = open(prefix + ".golden").read().strip()
    except (OSError, ValueError) as ex:
        return {"error": str(ex)}
    return {"fps": m["frames_per_second"], "wall_s": m["wall_clock_seconds"], "state_fnv1a64": h,
            "avg_contacts_last_frame": m.get("avg_contacts_per_frame")}


def stress_c1(capi, scenes, drivers: dict, frames: int = C1_FRAMES, settle_frames: int = 30) -> dict:
    """The reference's stress_test_blocks scene (1 024 boxes + ground, 1.0 m
    pitch as shipped, 300 frames, 1 substep, 10 iterations): the GPU without
    trajectory I/O (frames enqueued back to back) and the drop-in C++ driver
    with the reference's I/O (measured before this process touched the GPU);
    the CPU oracle on all host threads and on 4, without and with the same
    I/O; the solver-quality gates (largest penetration after each frame's
    solve, settled kinetic energy over the last frames against the oracle's)
    and the 0.85 m labelled variant."""
    import tempfile
    from dyphur_amd import formats
    from oracle import ref_cpu as rc

    out = {"workload": f"stress_test_blocks: 8x16x8 boxes (half 0.4 m, 1 kg) on a 1.0 m pitch + static ground, "
                       f"{frames} frames at 60 Hz, 1 substep, 10 iterations, friction 0.5",
           "penetration_note": "deepest contact left after each frame's solve (a second broadphase + narrowphase "
                               "over the solved state); the reference algorithm run by the oracle gives the same "
                               "depth bit for bit every frame, so the 5 mm bound is the reference solver's own "
                               "result on this scene",
           "published_fps": {"cpu_openmp_4_threads_xeon_e5_2667v4": 205.5, "cuda_rtx3060_warm": 233.0}}
    res = {}
    for label, pitch in (("pitch_1.0", 1.0), ("pitch_0.85", 0.85)):
        sc = scenes.stress_blocks(pitch=pitch)
        s = capi.Stream()
        w = capi.World(sc, s)
        init = w.bodies.host_arrays(sc)
        for _ in range(5):
            w.step_async()
        w.check()
        w.bodies.upload_state(init, s)
        s.wait()
        e0, e1 = capi.Event(), capi.Event()
        t0 = time.perf_counter()
        e0.record(s)
        for _ in range(frames):
            w.step_async()
        e1.record(s)
        s.wait()
        wall = time.perf_counter() - t0
        gpu_ms = e0.elapsed_ms(e1)
        w.check()
        h_gpu = w.fnv()
        # Gates: per frame, the deepest contact the narrowphase found before
        # the solve and the deepest one left after it (a second broadphase +
        # narrowphase over the solved state); kinetic energy over the last
        # settle_frames frames.
        w.bodies.upload_state(init, s)
        s.wait()
        pbp = capi.Broadphase(s, sc.n, sc.max_pairs)
        pnp = capi.Narrowphase(s, sc.max_contacts)
        pre, post, ke = [], [], []
        for f in range(frames):
            w.step()
            c = w.np.contacts(s)
            pre.append(float(c["depth"].max()) if c["n"] else 0.0)
            pbp.build_and_query(s, w.bodies, w.shapes, sc.bounds)
            n_p = pbp.download_count(s)
            pp, _ = pbp.pairs_ptr()
            pnp.run(s, pp, min(n_p, pbp.max_pairs), w.bodies, w.shapes)
            c2 = pnp.contacts(s)
            post.append(float(c2["depth"].max()) if c2["n"] else 0.0)
            if f >= frames - settle_frames:
                ke.append(kinetic_energy(w.state(), sc))
        r = {"gpu_fps_no_io": round(frames / wall, 1), "gpu_stream_ms_per_frame": round(gpu_ms / frames, 4),
             "gpu_body_substeps_per_s": round(sc.n * frames / wall, 1), "state_fnv1a64": f"{h_gpu:016x}",
             "max_penetration_m": {"after_solve": round(max(post), 6), "before_solve": round(max(pre), 6),
                                   "after_solve_last_frame": round(post[-1], 6)},
             "gate_penetration_5mm": bool(max(post) <= 0.005),
             "frames_deeper_than_5mm": int(sum(1 for x in post if x > 0.005)),
             "settled_ke_j": float(np.mean(ke)), "ke_frames": f"{frames - settle_frames}..{frames - 1}"}
        d = drivers.get(label)
        if d:
            r["driver_with_io"] = d
            if "fps" in d:
                r["driver_hash_equal"] = d["state_fnv1a64"] == f"{h_gpu:016x}"
        # Reference order on the GPU (forced serial), its settled energy.
        w.bodies.upload_state(init, s)
        s.wait()
        ke_s = []
        for f in range(frames):
            w.step(force_serial=True)
            if f >= frames - settle_frames:
                ke_s.append(kinetic_energy(w.state(), sc))
        r["settled_ke_reference_order_j"] = float(np.mean(ke_s))
        del pbp, pnp, w
        s.close()
        res[label] = (r, sc)
    out.update({k: v[0] for k, v in res.items()})

    # CPU oracle on the shipped scene: all threads and 4, without and with I/O.
    r, sc = res["pitch_1.0"]
    rc.build()
    threads = rc.max_threads()
    cpu = {"cores": threads, "cpu_model": cpu_model(), "kind": "port"}
    tmp = tempfile.mkdtemp(prefix="dyphur_c1_")
    for th in (threads, 4):
        rc.set_threads(th)
        for io in (False, True):
            ob, osh = rc.Bodies(sc), rc.Shapes(sc)
            ow = rc.World.for_scene(sc)
            snap = None
            tw = formats.TrajectoryWriter(os.path.join(tmp, "c1"), sc.n, frames) if io else None
            t0 = time.perf_counter()
            for f in range(frames):
                ow.frame(ob, osh, sc)
                if io and f % 2 == 0:
                    tw.write_frame({k: ob.arr[k].copy() for k in
                                    ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z")})
                if f == frames - settle_frames - 1:
                    snap = ob.copy()
            dt = time.perf_counter() - t0
            if tw:
                tw.close()
            key = f"fps_{th}_threads" + ("_with_io" if io else "")
            cpu[key] = round(frames / dt, 1)
            if th == threads and not io:
                cpu["state_fnv1a64"] = f"{ob.fnv():016x}"
                cpu["hash_equal_gpu"] = cpu["state_fnv1a64"] == r["state_fnv1a64"]
                # the oracle's settled energy: the last frames again from the snapshot
                kes = []
                for _ in range(settle_frames):
                    ow.frame(snap, osh, sc)
                    kes.append(kinetic_energy(snap.state(), sc))
                cpu["settled_ke_j"] = float(np.mean(kes))
    rc.set_threads(threads)
    cpu["body_substeps_per_s"] = round(sc.n * cpu[f"fps_{threads}_threads"], 1)
    out["cpu_oracle"] = cpu
    ke_g, ke_o = r["settled_ke_j"], cpu["settled_ke_j"]
    out["gate_settled_ke_within_5pct"] = bool(abs(ke_g - ke_o) <= 0.05 * abs(ke_o))
    out["gpu_over_cpu_no_io"] = round(r["gpu_fps_no_io"] / cpu[f"fps_{threads}_threads"], 2)
    return out


def frame_bytes(n_dyn, n_static, frames, body_solves, pairs, contacts, dyn_endpoints, iters):
    """Compulsory HBM bytes of reference frames, summed over the frames of a
    run (per stage: integrate 100 B per dynamic body + 4 per static; Morton
    keys, key sort, Karras 24 B per body each; leaf boxes 52; refit 56;
    traversal 52 per body + 8 per pair; narrowphase 76 per pair + 36 per
    contact; contact scan 8 per pair; colouring 12 per contact + 16 per body;
    solver, per iteration, 44 per contact + 132 per dynamic contact endpoint,
    and 32 per body renormalised).  body_solves: bodies x solves that ran;
    pairs, contacts, dyn_endpoints: totals over the frames."""
    n = n_dyn + n_static
    per_frame_n = 100.0 * n_dyn + 4.0 * n_static + (24 + 24 + 24 + 52 + 56 + 52 + 16) * n
    return (per_frame_n * frames + 32.0 * iters * body_solves + 8.0 * pairs + 76.0 * pairs + 36.0 * contacts
            + 8.0 * pairs + 12.0 * contacts + iters * (44.0 * contacts + 132.0 * dyn_endpoints))


def batched(capi, scenes, frames: int, plain_sample: int) -> dict:
    """4096 independent pile scenes stepped together: one launch for all
    frames (launch shape picked by the autotuner on a 10-frame trial from the
    initial state, which it restores), next to the plain per-scene loop of the
    drop-in API on a seeded sample of the same scenes."""
    t0 = time.perf_counter()
    scs = scenes.batched_piles()
    b = capi.Batch(scs)
    b.s.wait()
    shape, tune_ms = b.autotune(10)
    e0, e1 = capi.Event(), capi.Event()
    b.s.wait()
    w0 = time.perf_counter()
    e0.record(b.s)
    b.step(frames)
    e1.record(b.s)
    b.s.wait()
    wall = time.perf_counter() - w0
    gpu_ms = e0.elapsed_ms(e1)
    st = b.stats()  # raises on any device error flag
    n_bodies = b.n_bodies
    n_static = len(scs)
    n_per = np.diff(b.body_offset).astype(np.float64)
    nbytes = frame_bytes(n_bodies - n_static, n_static, frames, float((n_per * st["solves"]).sum()),
                         float(st["sum_pairs"].sum()), float(st["sum_contacts"].sum()),
                         float(st["sum_dyn_endpoints"].sum()), scs[0].n_iters)
    achieved = nbytes / (gpu_ms * 1e-3) / 1e9
    # The same frames captured as a HIP graph of single-frame launches.
    g_frames = 10
    gb = capi.Batch(scs)
    gb.set_shape(*shape)
    with capi.Graph.capture(gb.s) as g:
        for _ in range(g_frames):
            gb.step(1)
    gb.s.wait()
    e0.record(gb.s)
    for _ in range(frames // g_frames):
        g.launch(gb.s)
    e1.record(gb.s)
    gb.s.wait()
    graph_ms = e0.elapsed_ms(e1) / (g_frames * (frames // g_frames))
    same = all(np.array_equal(b.state()[k].view(np.uint32), gb.state()[k].view(np.uint32))
               for k in ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z"))
    # Plain per-scene loop: the drop-in frame (integrate, build_and_query,
    # device-count narrowphase, solve) scene after scene on one stream.
    rng = np.random.default_rng(20260003)
    idx = rng.choice(len(scs), min(plain_sample, len(scs)), replace=False)
    s = capi.Stream()
    worlds = [capi.World(scs[i], s) for i in idx]
    for w in worlds:
        w.step_async()
    s.wait()
    p_frames = 5
    p0 = time.perf_counter()
    for _ in range(p_frames):
        for w in worlds:
            w.step_async()
    s.wait()
    per_scene_frame = (time.perf_counter() - p0) / (p_frames * len(worlds))
    plain_frame = per_scene_frame * len(scs)
    value = n_bodies * frames / wall
    return {
        "workload": f"{len(scs)} pile scenes (boxes, lognormal body count median 64, sigma 0.75, "
                    f"clip [8, 512], own ground each), seed 20260003, {frames} frames at 60 Hz, "
                    f"1 substep, 10 iterations, all frames in one launch",
        "value": round(value, 1), "unit": "body-substeps/s", "bodies": n_bodies, "scenes": len(scs),
        "frames": frames, "ms_per_frame": round(1000.0 * wall / frames, 4),
        "stream_ms_per_frame": round(gpu_ms / frames, 4), "launches": 1,
        "shape": {"threads": shape[0], "wave_scene_max": shape[1]},
        "autotune_ms": {f"{t}/{m}": round(x, 3) for (t, m), x in zip(capi.Batch.TUNE_SHAPES, tune_ms)},
        "graph_ms_per_frame": round(graph_ms, 4), "graph_state_equal": bool(same),
        "plain_loop": {"ms_per_scene_frame": round(1000.0 * per_scene_frame, 4),
                       "sample_scenes": len(worlds), "sample_frames": p_frames,
                       "ms_per_frame_all_scenes": round(1000.0 * plain_frame, 2),
                       "value": round(n_bodies / plain_frame, 1)},
        "vs_plain_loop": round((n_bodies * frames / wall) / (n_bodies / plain_frame), 1),
        "load": {"pairs_mean": float(st["sum_pairs"].sum()) / frames,
                 "contacts_mean": float(st["sum_contacts"].sum()) / frames,
                 "serial_scenes_last_frame": int(st["used_serial"].sum())},
        "roofline": {"bound": "hbm", "kernel": "batch_frames", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                     "traffic": None, "algorithmic_bytes_per_launch": round(nbytes)},
        "setup_s": round(w0 - t0, 2),
    }


def articulated(capi, scenes, frames: int, check_frames: int = 20) -> dict:
    """Joint chains with contacts (1024 chains of 8 links over a ground,
    every joint type, limits, PD motors, compliance): the drop-in frame with
    a JointView, no host synchronisation inside the timed frames.  The oracle
    runs check_frames of the same scene on the host; its final state hash
    must equal the GPU's after as many frames."""
    sc = scenes.joint_chains(1024, 8, seed=20260006)
    n_joints = len(sc.joints["type"])
    warm = capi.World(sc)
    for _ in range(3):
        warm.step_async()
    warm.check()
    w = capi.World(sc)
    e0, e1 = capi.Event(), capi.Event()
    w.s.wait()
    t0 = time.perf_counter()
    e0.record(w.s)
    for _ in range(frames):
        w.step_async()
    e1.record(w.s)
    w.s.wait()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_ms(e1)
    last = w.check()  # raises on any device error flag
    # the solver's split, on further frames (its event timing waits for each solve)
    w.solver.set_timing(True)
    for _ in range(20):
        w.step_async()
    tm = w.solver.timing(w.s)
    # the oracle on the host, same frames from the same initial state
    from oracle import ref_cpu
    ref_cpu.build()
    ob, osh = ref_cpu.Bodies(sc), ref_cpu.Shapes(sc)
    ow = ref_cpu.World.for_scene(sc)
    js = ref_cpu.Joints(sc.joints)
    c0 = time.perf_counter()
    for _ in range(check_frames):
        ow.frame(ob, osh, sc, joints=js)
    cpu_s = time.perf_counter() - c0
    g = capi.World(sc)
    for _ in range(check_frames):
        g.step_async()
    g.check()
    same = g.fnv() == ob.fnv()
    return {
        "workload": f"joint chains: 1024 static bases x 8 links (boxes, spheres), {n_joints} joints "
                    f"(fixed, revolute, prismatic, ball; limits, PD motors, compliance) + contacts "
                    f"over a ground, seed 20260006, {frames} frames at 60 Hz, 10 iterations",
        "bodies": sc.n, "joints": n_joints, "frames": frames,
        "value": round(sc.n * frames / wall, 1), "unit": "body-substeps/s",
        "ms_per_frame": round(1000.0 * wall / frames, 4), "stream_ms_per_frame": round(gpu_ms / frames, 4),
        "solver_ms_per_frame": {"assign": round(tm["assign_ms"] / max(tm["solves"], 1), 4),
                                "sweep": round(tm["sweep_ms"] / max(tm["solves"], 1), 4)},
        "last_frame": {"contacts": last["n_contacts"], "classes": last["n_colors"],
                       "used_serial": last["used_serial"]},
        "cpu_baseline": {"value": round(sc.n * check_frames / cpu_s, 1), "unit": "body-substeps/s",
                         "cores": ref_cpu.lib().rc_max_threads(), "kind": "port",
                         "sample": f"the oracle's frames 0..{check_frames - 1} of the same scene"},
        "state_hash_equal_after_check_frames": bool(same),
    }


def differentiable(capi, scenes, frames: int = 120, every: int = 10) -> dict:
    """256 thrown 4x4x4 box stacks, loss sum |x_T - x*|^2.
    Forward alone (one launch of all frames) against forward with
    checkpoints + reverse sweep (recompute each checkpoint segment with a
    record, then the adjoints), both timed with HIP events on the stream."""
    scs = scenes.diff_batch(256)
    b = capi.Batch(scs)
    init = b.bodies.host_arrays(b.merged)
    e0, e1 = capi.Event(), capi.Event()
    b.step(1)
    b.bodies.upload_state(init, b.s)
    e0.record(b.s)
    b.step(frames)
    e1.record(b.s)
    e1.sync()
    fwd = e0.elapsed_ms(e1)
    d = capi.Diff(b, frames, every)
    b.bodies.upload_state(init, b.s)
    d.forward(frames)
    st = b.state()
    n = b.n_bodies
    adj = {k: np.zeros(n, np.float32) for k in ("pos_x", "pos_y", "pos_z")}
    for i, sc in enumerate(scs):
        lo = int(b.body_offset[i])
        for k, name in enumerate(("pos_x", "pos_y", "pos_z")):
            adj[name][lo:lo + sc.n - 1] = 2.0 * (st[name][lo:lo + sc.n - 1] - sc.meta["target"][:, k])
    d.backward(adj, download=False)  # warm-up
    b.bodies.upload_state(init, b.s)
    b.s.wait()
    e0.record(b.s)
    d.forward(frames)
    e1.record(b.s)
    e1.sync()
    dfw = e0.elapsed_ms(e1)
    e0.record(b.s)
    d.backward(adj, download=False)
    e1.record(b.s)
    e1.sync()
    bwd = e0.elapsed_ms(e1)
    return {
        "workload": f"256 scenes of 64 boxes (4x4x4 stack, half 0.2 m, 1 kg, 0.45 m pitch, 1 m above the "
                    f"ground) thrown with v0 ~ N(0,1), w0 ~ N(0,2), seed 20260004; {frames} frames at 60 Hz; "
                    f"loss sum |x_T - x*|^2; gradients w.r.t. v0 and inverse mass; checkpoint every {every} frames",
        "bodies": n, "frames": frames, "checkpoint_every": every,
        "forward_ms_per_frame": round(fwd / frames, 4),
        "forward_checkpointed_ms_per_frame": round(dfw / frames, 4),
        "adjoint_ms_per_frame": round(bwd / frames, 4),
        "fwd_adjoint_ms_per_frame": round((dfw + bwd) / frames, 4),
        "fwd_adjoint_over_fwd": round((dfw + bwd) / fwd, 2),
        "value": round(n * frames / ((dfw + bwd) * 1e-3), 1), "unit": "body-substeps/s (forward + adjoint)",
    }


def differentiable_at_scale(capi, scenes, frames: int = 240, every: int = 20) -> dict:
    """1 024 scenes of lognormal body counts (median 128, clipped to [16,
    1 024]) of mixed boxes and spheres at half scale, 4 substeps per frame:
    the forward alone against the checkpointed forward + the backward
    (recompute from each checkpoint with a record, then the reverse sweep),
    timed with HIP events on the stream.  Tapes hold 4 contacts per body
    (the workload peaks near 1.2) and 80 GiB each, so that one tape holds a
    20-frame segment; a step with more contacts would make the backward
    fail, not return a wrong gradient."""
    sub = 4
    scs = scenes.diff_at_scale()
    b = capi.Batch(scs)
    init = b.bodies.host_arrays(b.merged)
    e0, e1 = capi.Event(), capi.Event()
    b.step(1, n_substeps=sub)
    b.bodies.upload_state(init, b.s)
    e0.record(b.s)
    b.step(frames, n_substeps=sub)
    e1.record(b.s)
    e1.sync()
    fwd = e0.elapsed_ms(e1)
    d = capi.Diff(b, frames, every, n_substeps=sub, tape_contacts_per_body=4.0, tape_bytes=80 << 30)
    steps, step_bytes = d.tape_layout()
    b.bodies.upload_state(init, b.s)
    d.forward(frames)
    st = b.state()
    n = b.n_bodies
    adj = {k: np.zeros(n, np.float32) for k in ("pos_x", "pos_y", "pos_z")}
    for i, sc in enumerate(scs):
        lo = int(b.body_offset[i])
        for k, name in enumerate(("pos_x", "pos_y", "pos_z")):
            adj[name][lo:lo + sc.n - 1] = 2.0 * (st[name][lo:lo + sc.n - 1] - sc.meta["target"][:, k])
    d.backward(adj, download=False)  # warm-up
    b.bodies.upload_state(init, b.s)
    b.s.wait()
    e0.record(b.s)
    d.forward(frames)
    e1.record(b.s)
    e1.sync()
    dfw = e0.elapsed_ms(e1)
    e0.record(b.s)
    d.backward(adj, download=False)
    e1.record(b.s)
    e1.sync()
    bwd = e0.elapsed_ms(e1)
    counts = np.array([sc.n for sc in scs])
    return {
        "workload": f"{len(scs)} scenes of mixed boxes (2/3) and spheres (1/3) at half scale, lognormal body counts "
                    f"(median 128, sigma 0.75, clip [16, 1024]) + a ground each, seed 20260005; {frames} frames at "
                    f"60 Hz, {sub} substeps, 10 iterations; loss sum |x_T - x*|^2; gradients w.r.t. v0 and inverse "
                    f"mass; checkpoint every {every} frames",
        "bodies": n, "scenes": len(scs), "max_scene_bodies": int(counts.max()), "frames": frames, "substeps": sub,
        "checkpoint_every": every, "tape_substeps": steps, "tape_gib": round(steps * step_bytes / 2**30, 1),
        "forward_ms_per_frame": round(fwd / frames, 4),
        "forward_body_substeps_per_s": round(n * frames * sub / (fwd * 1e-3), 1),
        "forward_checkpointed_ms_per_frame": round(dfw / frames, 4),
        "adjoint_ms_per_frame": round(bwd / frames, 4),
        "fwd_adjoint_ms_per_frame": round((dfw + bwd) / frames, 4),
        "fwd_adjoint_over_fwd": round((dfw + bwd) / fwd, 2),
        "value": round(n * frames * sub / ((dfw + bwd) * 1e-3), 1), "unit": "body-substeps/s (forward + adjoint)",
    }


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs the frame step on one GPU only; --gpus {args.gpus} "
              f"is not supported (a multi-GPU run would shard the scene into Morton-range slabs "
              f"with a ghost-body exchange per substep).", file=sys.stderr)
        sys.exit(2)

    # The drop-in C++ stress driver first, in child processes, before this
    # process touches the GPU (a process's cooperative launches are refused
    # while another process on the device holds them).
    drivers = {}
    if args.c1_frames > 0:
        import tempfile
        binary = os.path.join(ROOT, "dyphur_amd", "bin", "stress_test_blocks")
        if os.path.exists(binary):
            tmp = tempfile.mkdtemp(prefix="dyphur_drv_")
            drivers["pitch_1.0"] = run_driver(binary, os.path.join(tmp, "p100"), args.c1_frames)
            drivers["pitch_0.85"] = run_driver(binary, os.path.join(tmp, "p085"), args.c1_frames, 0.85)

    from dyphur_amd import capi, scenes

    capi.load()
    if capi.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    plan = plan_frames(args.steps, C2_FRAMES, 0 if args.no_cpu else args.cpu_window)
    total, start, steps = plan["total"], plan["start"], plan["steps"]
    sc = scenes.random_mixed(args.bodies, seed=20260002, frames=total)
    s = capi.Stream()
    w = capi.World(sc, s)
    init = w.bodies.host_arrays(sc)

    def reset():
        w.bodies.upload_state(init, s)
        s.wait()

    # Warm-up frames (untimed), then restart from the initial state and
    # advance, untimed, to the first timed frame.
    for _ in range(args.warmup):
        w.step_async()
    w.check()
    reset()
    for _ in range(start):
        w.step_async()
    w.check()

    # Timed frames: no host synchronisation and no timing mode inside the loop.
    e0, e1 = capi.Event(), capi.Event()
    s.wait()
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(steps):
        w.step_async()
    e1.record(s)
    s.wait()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_ms(e1)
    w.check()  # raises on any device error flag of the timed frames
    final_hash = w.fnv()

    # Per-kernel times over the same frames, in a separate untimed pass.
    reset()
    for _ in range(start):
        w.step_async()
    w.check()
    kernels = stage_kernels(capi, w, sc, s, steps)
    if w.fnv() != final_hash:
        raise RuntimeError("the kernel-timing pass ended in a different state than the timed run")

    # Untimed replay of the whole run: per-frame load statistics, the CPU
    # windows' snapshots, and the same final state as the timed run.
    window, starts = plan["window"], plan["starts"]
    want = (set(starts) | {f + window for f in starts}) if window else set()
    stats, snaps = replay(w, init, s, total, want)
    replay_hash = w.fnv()
    if replay_hash != final_hash:
        raise RuntimeError(f"replay state {replay_hash:016x} differs from the timed run {final_hash:016x}")
    timed = stats[start:]
    n = sc.n
    value = n * steps * sc.n_substeps / wall
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * wall / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"single scene, {args.bodies} mixed boxes/spheres + static ground, seed 20260002, "
                        f"{total}-frame run, timed frames {start}..{total - 1} at 60 Hz, 1 substep, 10 iterations",
            "bodies": n, "frames": total, "timed_frames": [start, total - 1], "substeps": sc.n_substeps,
            "iterations": sc.n_iters, "seed": 20260002, "max_pairs": sc.max_pairs, "max_contacts": sc.max_contacts,
        },
        "stream_ms_per_step": round(gpu_ms / steps, 4),
        "load": {"timed_pairs_mean": float(np.mean([t["n_pairs"] for t in timed])),
                 "timed_contacts_mean": float(np.mean([t["n_contacts"] for t in timed])),
                 "run_pairs_mean": float(np.mean([t["n_pairs"] for t in stats])),
                 "run_contacts_mean": float(np.mean([t["n_contacts"] for t in stats])),
                 "timed_classes_mean": float(np.mean([t["n_colors"] for t in timed])),
                 "timed_assign_rounds_mean": float(np.mean([t["rounds"] for t in timed])),
                 "colors_max": int(max(t["n_colors"] for t in stats)),
                 "pairs_max": int(max(t["n_pairs"] for t in stats)),
                 "contacts_max": int(max(t["n_contacts"] for t in stats))},
        "final_state_fnv1a64": f"{final_hash:016x}",
    }
    # Roofline of the dominant kernel of the frame: per-kernel event times
    # over the timed frames' replay, algorithmic bytes per launch from the
    # per-stage figures and the frames' own pair and contact counts.
    out["roofline"] = roofline_of(kernels)
    if window:
        out["cpu_baseline"] = cpu_baseline(sc, snaps, starts, window)
        out["vs_cpu"] = round(out["value"] / out["cpu_baseline"]["value"], 2)
    # The other legs do not use the single-scene world: release it and its
    # stream (a process's streams share HIP's few hardware queues).
    del w, e0, e1, snaps
    s.close()
    del s
    if args.c1_frames > 0:
        out["stress_c1"] = stress_c1(capi, scenes, drivers, args.c1_frames)
    if args.batched_frames > 0:
        out["batched"] = batched(capi, scenes, args.batched_frames, args.plain_sample)
    if args.diff_frames > 0:
        out["differentiable"] = differentiable(capi, scenes, args.diff_frames)
    if args.diff_scale_frames > 0:
        out["differentiable_at_scale"] = differentiable_at_scale(capi, scenes, args.diff_scale_frames)
    if args.joint_frames > 0:
        out["articulated"] = articulated(capi, scenes, args.joint_frames)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
