"""Readers and writers of the reference drivers' files
(examples/common/scene_io.hpp:23-134, examples/stress_test_blocks/main.cpp:172-197,301-336).

  <prefix>.scene         u32 n_bodies, n_shapes; u32 body_shape_idx[n_bodies];
                         per shape 20 B: u32 type, f32 half_x, half_y, half_z, u32 0
  <prefix>.meshmap       text lines "<body_idx> <path>", '#' comments
  <prefix>.trajectory    u32 n_bodies, n_frames (simulation frames), then per
                         written frame n_bodies x 7 f32 (pos xyz, rot wxyz);
                         the drivers write every 2nd frame, ceil(n_frames / 2)
  <prefix>.golden        16 lower-case hex digits and a newline
  <prefix>.metrics.json  n_bodies, n_frames, simulation_seconds,
                         wall_clock_seconds, frames_per_second,
                         avg_contacts_per_frame, realtime_factor

The C++ equivalents are in dyphur_amd/cpp/include/io/scene_io.hpp.
"""
from __future__ import annotations

import json
from dataclasses import dataclass

import numpy as np

_SHAPE = np.dtype([("type", "<u4"), ("half", "<f4", (3,)), ("pad", "<u4")])
assert _SHAPE.itemsize == 20


@dataclass
class SceneFile:
    body_shape_idx: np.ndarray  # (n_bodies,) uint32
    shape_type: np.ndarray      # (n_shapes,) uint32
    shape_half: np.ndarray      # (n_shapes, 3) float32


def write_scene(prefix: str, body_shape_idx, shape_type, shape_half) -> None:
    bsi = np.ascontiguousarray(body_shape_idx, "<u4")
    rec = np.zeros(len(shape_type), _SHAPE)
    rec["type"] = shape_type
    rec["half"] = np.asarray(shape_half, "<f4").reshape(-1, 3)
    with open(prefix + ".scene", "wb") as f:
        f.write(np.array([len(bsi), len(rec)], "<u4").tobytes())
        f.write(bsi.tobytes())
        f.write(rec.tobytes())


def read_scene(prefix: str) -> SceneFile:
    raw = open(prefix + ".scene", "rb").read()
    n_bodies, n_shapes = np.frombuffer(raw, "<u4", 2)
    off = 8
    bsi = np.frombuffer(raw, "<u4", int(n_bodies), off).copy()
    off += 4 * int(n_bodies)
    rec = np.frombuffer(raw, _SHAPE, int(n_shapes), off)
    if off + rec.nbytes != len(raw):
        raise ValueError(f"{prefix}.scene: size {len(raw)} does not match its header")
    return SceneFile(bsi, rec["type"].copy(), rec["half"].copy())


def write_meshmap(prefix: str, entries) -> None:
    with open(prefix + ".meshmap", "w") as f:
        for idx, path in entries:
            f.write(f"{int(idx)} {path}\n")


def read_meshmap(prefix: str) -> dict:
    out = {}
    try:
        lines = open(prefix + ".meshmap").read().splitlines()
    except FileNotFoundError:
        return out
    for line in lines:
        if not line or line.startswith("#"):
            continue
        parts = line.split()
        if len(parts) >= 2:
            out[int(parts[0])] = parts[1]
    return out


class TrajectoryWriter:
    """Header at open, one record per write_frame (SoA arrays in)."""

    def __init__(self, prefix: str, n_bodies: int, n_frames: int):
        self.f = open(prefix + ".trajectory", "wb")
        self.n = n_bodies
        self.f.write(np.array([n_bodies, n_frames], "<u4").tobytes())

    def write_frame(self, state: dict) -> None:
        keys = ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z")
        rows = np.stack([np.asarray(state[k], "<f4") for k in keys], axis=1)
        assert rows.shape == (self.n, 7)
        self.f.write(rows.tobytes())

    def close(self) -> None:
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_trajectory(prefix: str):
    """Returns (n_bodies, n_frames, frames) with frames of shape
    (written frames, n_bodies, 7) float32."""
    raw = open(prefix + ".trajectory", "rb").read()
    n_bodies, n_frames = (int(v) for v in np.frombuffer(raw, "<u4", 2))
    body = raw[8:]
    rec = 7 * 4 * n_bodies
    if n_bodies == 0 or len(body) % rec:
        raise ValueError(f"{prefix}.trajectory: {len(body)} bytes is not a whole number of frames")
    frames = np.frombuffer(body, "<f4").reshape(len(body) // rec, n_bodies, 7)
    return n_bodies, n_frames, frames


def write_golden(prefix: str, value: int) -> None:
    with open(prefix + ".golden", "w") as f:
        f.write(f"{value:016x}\n")


def read_golden(prefix: str) -> int:
    return int(open(prefix + ".golden").read().strip(), 16)


METRIC_KEYS = ("n_bodies", "n_frames", "simulation_seconds", "wall_clock_seconds", "frames_per_second",
               "avg_contacts_per_frame", "realtime_factor")


def write_metrics(prefix: str, metrics: dict) -> None:
    with open(prefix + ".metrics.json", "w") as f:
        json.dump({k: metrics[k] for k in METRIC_KEYS}, f, indent=2)
        f.write("\n")


def read_metrics(prefix: str) -> dict:
    d = json.load(open(prefix + ".metrics.json"))
    missing = [k for k in METRIC_KEYS if k not in d]
    if missing:
        raise ValueError(f"{prefix}.metrics.json: missing {missing}")
    return d
