"""Synthetic scenes for the frame step, as host arrays of body parameters.

A ``Scene`` holds what ``BodyStore.add`` / ``ShapeStore.add`` receive in the
reference (position, rotation wxyz, velocities, mass, 3x3 inertia, shape
handle, flags; shape type and half extents), plus the run parameters the
driver uses.  The same arrays feed the HIP path and the CPU oracle, so both
see bit-identical inputs.

Scenes:
  * ``stress_blocks``: examples/stress_test_blocks/main.cpp:72-166 (the
    1 024-box stack over a static ground, 1.0 m pitch as shipped).
  * ``determinism_boxes``: core/tests/sim_determinism_test.cpp:248-349 (the
    8-box golden-hash scene and its 256-box variant).
  * ``random_mixed``: 2/3 boxes, 1/3 spheres, seeded uniform positions and
    Shoemake quaternions, density 1000 kg/m^3 (32 768 bodies by default).
  * ``batched_piles``: independent scenes with lognormal body counts, each with
    its own ground, stacked in jittered columns.
  * ``articulated_arm``: core/tests/sim_determinism_test.cpp:130-190 (static
    base and two revolute links with PD motors; the golden articulation hash).
  * ``joint_chains``: chains of boxes and spheres hanging from static bases,
    every joint type with seeded limits, motors and compliance, over a ground.
  * ``diff_at_scale``: 1 024 scenes for the batched + differentiable form at
    scale: lognormal body counts (median 128, clipped to [16, 1 024]), each
    scene random_mixed's 2/3 boxes and 1/3 spheres at half scale, 4 substeps
    per frame.

Randomness: PCG32 (XSH-RR, O'Neill 2014) seeded through splitmix64, so a seed
gives the same scene on every machine with this numpy.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32
SHAPE_BOX = 0
SHAPE_SPHERE = 1
FLAG_STATIC = 1


# ---------------------------------------------------------------------------
# PCG32 seeded by splitmix64
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def _splitmix64(state: int) -> tuple[int, int]:
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


class Pcg32:
    """PCG32 XSH-RR generator; ``seed`` is expanded with splitmix64."""

    _MULT = 6364136223846793005

    def __init__(self, seed: int):
        sm, s0 = _splitmix64(seed & _M64)
        _, s1 = _splitmix64(sm)
        self.inc = ((s1 << 1) | 1) & _M64
        self.state = 0
        self.next_u32()
        self.state = (self.state + s0) & _M64
        self.next_u32()

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * self._MULT + self.inc) & _M64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self) -> float:
        """Uniform in [0, 1) with 32 random bits (float64)."""
        return self.next_u32() * (1.0 / 4294967296.0)

    def uniform_range(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.uniform()

    def normal(self) -> float:
        """Standard normal by Box-Muller (one draw per call, float64)."""
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


# ---------------------------------------------------------------------------
# Scene container
# ---------------------------------------------------------------------------
@dataclass
class Scene:
    name: str
    # shapes
    shape_type: np.ndarray
    shape_half: np.ndarray          # (S, 3) float32
    # bodies
    pos: np.ndarray                 # (N, 3) float32
    rot: np.ndarray                 # (N, 4) float32, (w, x, y, z)
    vel: np.ndarray                 # (N, 3)
    ang: np.ndarray                 # (N, 3)
    mass: np.ndarray                # (N,)
    inertia: np.ndarray             # (N, 9) row-major body-frame inertia
    shape: np.ndarray               # (N,) uint32
    flags: np.ndarray               # (N,) uint32
    bounds: np.ndarray              # (6,) float32 min xyz, max xyz
    max_pairs: int
    max_contacts: int
    dt: float = 1.0 / 60.0
    n_substeps: int = 1
    n_iters: int = 10
    mu: float = 0.5
    gravity: tuple = (0.0, -9.81, 0.0)
    frames: int = 300
    meta: dict = field(default_factory=dict)
    # JointView arrays (body_parent, body_child, type, JOINT_F32 fields) or None
    joints: dict | None = None

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def n_shapes(self) -> int:
        return int(self.shape_type.shape[0])


def _alloc(n: int):
    rot = np.zeros((n, 4), F32)
    rot[:, 0] = 1.0
    eye = np.zeros((n, 9), F32)
    eye[:, [0, 4, 8]] = 1.0
    return dict(pos=np.zeros((n, 3), F32), rot=rot, vel=np.zeros((n, 3), F32),
                ang=np.zeros((n, 3), F32), mass=np.ones(n, F32), inertia=eye,
                shape=np.zeros(n, np.uint32), flags=np.zeros(n, np.uint32))


def _diag(v: float) -> np.ndarray:
    return np.array([v, 0, 0, 0, v, 0, 0, 0, v], F32)


# ---------------------------------------------------------------------------
# examples/stress_test_blocks/main.cpp:72-166
# ---------------------------------------------------------------------------
def stress_blocks(nx: int = 8, ny: int = 16, nz: int = 8, pitch: float = 1.0) -> Scene:
    n_dyn = nx * ny * nz
    n = n_dyn + 1
    box_h = F32(0.4)
    gnd_hy = F32(0.5)
    gnd_hx = F32(max(20.0, 0.5 * max(nx, nz) + 2.0))
    b = _alloc(n)
    m = F32(1.0)
    i_diag = m / F32(3.0) * (box_h * box_h + box_h * box_h)
    k = 0
    p = F32(pitch)
    for iy in range(ny):
        for ix in range(nx):
            for iz in range(nz):
                b["pos"][k] = (F32(ix - nx // 2) + F32(0.5)) * p, gnd_hy + F32(iy + 1) * p, \
                              (F32(iz - nz // 2) + F32(0.5)) * p
                b["inertia"][k] = _diag(i_diag)
                k += 1
    # static ground plate (mass 0, identity inertia, Static flag)
    b["mass"][n_dyn] = 0.0
    b["shape"][n_dyn] = 1
    b["flags"][n_dyn] = FLAG_STATIC
    bound_xz = float(gnd_hx) + 5.0
    bound_y = max(22.0, float(gnd_hy) + ny + 5.0)
    return Scene(
        name=f"stress_blocks_{nx}x{ny}x{nz}" + ("" if pitch == 1.0 else f"_pitch{pitch}"),
        shape_type=np.array([SHAPE_BOX, SHAPE_BOX], np.uint32),
        shape_half=np.array([[box_h] * 3, [gnd_hx, gnd_hy, gnd_hx]], F32),
        bounds=np.array([-bound_xz, -2.0, -bound_xz, bound_xz, bound_y, bound_xz], F32),
        max_pairs=max(65536, n_dyn * 48), max_contacts=max(16384, n_dyn * 16),
        frames=300, **b)


# ---------------------------------------------------------------------------
# core/tests/sim_determinism_test.cpp:248-349
# ---------------------------------------------------------------------------
def determinism_boxes(nx: int = 2, ny: int = 2, nz: int = 2) -> Scene:
    n_dyn = nx * ny * nz
    n = n_dyn + 1
    box_h = F32(0.4)
    gnd_hx, gnd_hy = F32(5.0), F32(0.5)
    b = _alloc(n)
    m = F32(1.0)
    i_diag = m / F32(3.0) * F32(2.0) * box_h * box_h
    k = 0
    for iy in range(ny):
        for ix in range(nx):
            for iz in range(nz):
                b["pos"][k] = (F32(ix - nx // 2) + F32(0.5)) * F32(1.0), \
                              gnd_hy + F32(iy + 1) * F32(1.0), \
                              (F32(iz - nz // 2) + F32(0.5)) * F32(1.0)
                b["inertia"][k] = _diag(i_diag)
                k += 1
    b["mass"][n_dyn] = 0.0
    b["shape"][n_dyn] = 1
    b["flags"][n_dyn] = FLAG_STATIC
    cap = n * 32 + 256
    return Scene(
        name=f"determinism_{nx}x{ny}x{nz}",
        shape_type=np.array([SHAPE_BOX, SHAPE_BOX], np.uint32),
        shape_half=np.array([[box_h] * 3, [gnd_hx, gnd_hy, gnd_hx]], F32),
        bounds=np.array([-20, -2, -20, 20, 30, 20], F32),
        max_pairs=cap, max_contacts=16384, frames=20, **b)


# ---------------------------------------------------------------------------
# One dynamic plate carrying a grid of spheres: every sphere contact shares the
# plate, so the contacts form one clique and Jones-Plassmann colouring runs out
# of its 64 colours (core/src/xpbd_solver.cpp:639-651 then falls back to the
# serial path).
# ---------------------------------------------------------------------------
def star_plate(nx: int = 20, nz: int = 15, pitch: float = 0.5, radius: float = 0.2) -> Scene:
    n_sph = nx * nz
    n = n_sph + 2
    gnd_hy = F32(0.5)
    plate_h = np.array([F32(0.5) * F32(nx) * F32(pitch) + F32(0.25), F32(0.25),
                        F32(0.5) * F32(nz) * F32(pitch) + F32(0.25)], F32)
    r = F32(radius)
    b = _alloc(n)
    plate_y = gnd_hy + plate_h[1] - F32(0.005)
    sph_y = plate_y + plate_h[1] + r - F32(0.005)
    k = 0
    for ix in range(nx):
        for iz in range(nz):
            b["pos"][k] = (F32(ix) - F32(0.5) * F32(nx - 1)) * F32(pitch), sph_y, \
                          (F32(iz) - F32(0.5) * F32(nz - 1)) * F32(pitch)
            b["inertia"][k] = _diag(F32(0.4) * r * r)
            b["shape"][k] = 0
            k += 1
    m_plate = F32(50.0)
    b["pos"][n_sph] = 0.0, plate_y, 0.0
    b["mass"][n_sph] = m_plate
    b["inertia"][n_sph] = np.array([m_plate / F32(3.0) * (plate_h[1] * plate_h[1] + plate_h[2] * plate_h[2]), 0, 0,
                                    0, m_plate / F32(3.0) * (plate_h[0] * plate_h[0] + plate_h[2] * plate_h[2]), 0,
                                    0, 0, m_plate / F32(3.0) * (plate_h[0] * plate_h[0] + plate_h[1] * plate_h[1])],
                                   F32)
    b["shape"][n_sph] = 1
    b["mass"][n_sph + 1] = 0.0
    b["shape"][n_sph + 1] = 2
    b["flags"][n_sph + 1] = FLAG_STATIC
    gnd = F32(max(float(plate_h[0]), float(plate_h[2])) + 2.0)
    return Scene(
        name=f"star_plate_{nx}x{nz}",
        shape_type=np.array([SHAPE_SPHERE, SHAPE_BOX, SHAPE_BOX], np.uint32),
        shape_half=np.array([[r, r, r], plate_h, [gnd, gnd_hy, gnd]], F32),
        bounds=np.array([-gnd - 2.0, -2.0, -gnd - 2.0, gnd + 2.0, 10.0, gnd + 2.0], F32),
        max_pairs=max(4096, n * 16), max_contacts=max(4096, n * 8), frames=10, **b)


# ---------------------------------------------------------------------------
# Random mixed boxes/spheres in a volume above a ground plate
# ---------------------------------------------------------------------------
def _shoemake(rng: Pcg32) -> tuple[float, float, float, float]:
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
    t2, t3 = 2.0 * math.pi * u2, 2.0 * math.pi * u3
    # (w, x, y, z)
    return b * math.cos(t3), a * math.sin(t2), a * math.cos(t2), b * math.sin(t3)


def random_mixed(n_dyn: int = 32768, seed: int = 20260002, scale: float = 1.0,
                 extent=(30.0, 40.0, 30.0), y0: float = 1.0, frames: int = 600) -> Scene:
    """Config: 2/3 boxes (half extents U[0.1,0.5]^3), 1/3 spheres (r U[0.1,0.4]),
    uniform positions in x,z in [-ex/2, ex/2], y in [y0, y0+ey], Shoemake
    rotations, density 1000 kg/m^3; boxes then spheres, then shuffled; one
    static ground (half 20 x 0.5 x 20) as the last body.  Every body has its
    own shape record."""
    rng = Pcg32(seed)
    n_box = (2 * n_dyn) // 3
    kinds = np.array([SHAPE_BOX] * n_box + [SHAPE_SPHERE] * (n_dyn - n_box), np.uint32)
    # Fisher-Yates shuffle with the same generator.
    for i in range(n_dyn - 1, 0, -1):
        j = rng.next_u32() % (i + 1)
        kinds[i], kinds[j] = kinds[j], kinds[i]
    n = n_dyn + 1
    b = _alloc(n)
    shape_type = np.zeros(n, np.uint32)
    shape_half = np.zeros((n, 3), F32)
    ex, ey, ez = extent
    for k in range(n_dyn):
        if kinds[k] == SHAPE_BOX:
            h = [rng.uniform_range(0.1, 0.5) * scale for _ in range(3)]
            hf = [F32(v) for v in h]
            m = 8000.0 * float(hf[0]) * float(hf[1]) * float(hf[2])
            ixx = m / 3.0 * (float(hf[1]) ** 2 + float(hf[2]) ** 2)
            iyy = m / 3.0 * (float(hf[0]) ** 2 + float(hf[2]) ** 2)
            izz = m / 3.0 * (float(hf[0]) ** 2 + float(hf[1]) ** 2)
            shape_type[k] = SHAPE_BOX
            shape_half[k] = hf
            b["inertia"][k] = np.array([ixx, 0, 0, 0, iyy, 0, 0, 0, izz], F32)
        else:
            r = F32(rng.uniform_range(0.1, 0.4) * scale)
            m = 4000.0 * math.pi * float(r) ** 3 / 3.0
            i = 0.4 * m * float(r) ** 2
            shape_type[k] = SHAPE_SPHERE
            shape_half[k] = (r, r, r)
            b["inertia"][k] = _diag(i)
        b["mass"][k] = m
        b["pos"][k] = (rng.uniform_range(-ex / 2, ex / 2) * scale,
                       (y0 + rng.uniform() * ey) * scale,
                       rng.uniform_range(-ez / 2, ez / 2) * scale)
        b["rot"][k] = _shoemake(rng)
        b["shape"][k] = k
    g = n_dyn
    shape_type[g] = SHAPE_BOX
    shape_half[g] = (20.0 * scale, 0.5, 20.0 * scale)
    b["mass"][g] = 0.0
    b["shape"][g] = g
    b["flags"][g] = FLAG_STATIC
    top = (y0 + ey) * scale + 5.0
    half_xz = 20.0 * scale
    return Scene(
        name=f"random_mixed_{n_dyn}",
        shape_type=shape_type, shape_half=shape_half,
        bounds=np.array([-half_xz, -2.0, -half_xz, half_xz, top, half_xz], F32),
        max_pairs=48 * n, max_contacts=16 * n, frames=frames,
        meta=dict(seed=seed, n_box=int(n_box), n_sphere=int(n_dyn - n_box)), **b)


# ---------------------------------------------------------------------------
# Batched: many independent scenes
# ---------------------------------------------------------------------------
def lognormal_counts(n_scenes: int, median: float, sigma: float, lo: int, hi: int,
                     rng: Pcg32) -> np.ndarray:
    out = np.zeros(n_scenes, np.int64)
    for s in range(n_scenes):
        v = int(round(math.exp(math.log(median) + sigma * rng.normal())))
        out[s] = min(max(v, lo), hi)
    return out


def pile_scene(n_dyn: int, rng: Pcg32, h_lo: float = 0.1, h_hi: float = 0.4,
               v_sigma: float = 0.5, gap: float = 0.05, per_column: int = 8,
               ground_half: float = 10.0) -> Scene:
    """One pile scene: boxes in a jittered column grid (<= per_column per
    column, 5 cm vertical gaps), N(0, v_sigma) initial linear velocity per
    axis, own static ground (half ground_half x 0.5 x ground_half) last."""
    n = n_dyn + 1
    b = _alloc(n)
    shape_type = np.zeros(n, np.uint32)
    shape_half = np.zeros((n, 3), F32)
    n_cols = (n_dyn + per_column - 1) // per_column
    side = max(1, int(math.ceil(math.sqrt(n_cols))))
    pitch = 2.0 * h_hi + 0.3
    col_top = [0.5] * n_cols
    for k in range(n_dyn):
        c = k % n_cols
        h = [F32(rng.uniform_range(h_lo, h_hi)) for _ in range(3)]
        m = 1000.0 * 8.0 * float(h[0]) * float(h[1]) * float(h[2])
        cx = (c % side - (side - 1) / 2.0) * pitch + rng.uniform_range(-0.05, 0.05)
        cz = (c // side - (side - 1) / 2.0) * pitch + rng.uniform_range(-0.05, 0.05)
        y = col_top[c] + gap + float(h[1])
        col_top[c] = y + float(h[1])
        shape_type[k] = SHAPE_BOX
        shape_half[k] = h
        b["pos"][k] = (cx, y, cz)
        b["mass"][k] = m
        b["inertia"][k] = np.array([m / 3.0 * (float(h[1]) ** 2 + float(h[2]) ** 2), 0, 0,
                                    0, m / 3.0 * (float(h[0]) ** 2 + float(h[2]) ** 2), 0,
                                    0, 0, m / 3.0 * (float(h[0]) ** 2 + float(h[1]) ** 2)], F32)
        b["vel"][k] = (v_sigma * rng.normal(), v_sigma * rng.normal(), v_sigma * rng.normal())
        b["shape"][k] = k
    g = n_dyn
    shape_type[g] = SHAPE_BOX
    shape_half[g] = (ground_half, 0.5, ground_half)
    b["mass"][g] = 0.0
    b["shape"][g] = g
    b["flags"][g] = FLAG_STATIC
    top = max(col_top) + 10.0
    return Scene(name=f"pile_{n_dyn}", shape_type=shape_type, shape_half=shape_half,
                 bounds=np.array([-ground_half - 5, -2, -ground_half - 5,
                                  ground_half + 5, top, ground_half + 5], F32),
                 max_pairs=max(256, 48 * n), max_contacts=max(64, 16 * n), frames=300, **b)


def batched_piles(n_scenes: int = 4096, seed: int = 20260003, median: float = 64.0,
                  sigma: float = 0.75, lo: int = 8, hi: int = 512) -> list[Scene]:
    rng = Pcg32(seed)
    counts = lognormal_counts(n_scenes, median, sigma, lo, hi, rng)
    return [pile_scene(int(c), rng) for c in counts]


def diff_stack(rng: Pcg32, n_side: int = 4, half: float = 0.2, pitch: float = 0.45, lift: float = 1.0,
               v_sigma: float = 1.0, w_sigma: float = 2.0, target_sigma: float = 0.5) -> Scene:
    """A differentiable-form scene: n_side^3 boxes (half extent `half`, 1 kg)
    on a `pitch` grid whose lowest layer is `lift` m above the ground's top,
    thrown with linear velocity N(0, v_sigma) and angular velocity
    N(0, w_sigma) per axis; the static ground (half 10 x 0.5 x 10) last.
    meta["target"] holds the loss targets x* = x0 + N(0, target_sigma)."""
    n_dyn = n_side ** 3
    n = n_dyn + 1
    b = _alloc(n)
    inertia = 1.0 / 3.0 * (2.0 * half * half)
    k = 0
    for iy in range(n_side):
        for ix in range(n_side):
            for iz in range(n_side):
                b["pos"][k] = ((ix - (n_side - 1) / 2.0) * pitch, 0.5 + lift + half + iy * pitch,
                               (iz - (n_side - 1) / 2.0) * pitch)
                b["inertia"][k] = _diag(inertia)
                k += 1
    for k in range(n_dyn):
        b["vel"][k] = (v_sigma * rng.normal(), v_sigma * rng.normal(), v_sigma * rng.normal())
        b["ang"][k] = (w_sigma * rng.normal(), w_sigma * rng.normal(), w_sigma * rng.normal())
    target = b["pos"][:n_dyn].astype(np.float64) + np.array(
        [[target_sigma * rng.normal() for _ in range(3)] for _ in range(n_dyn)])
    shape_type = np.zeros(2, np.uint32)
    shape_half = np.array([[half, half, half], [10.0, 0.5, 10.0]], F32)
    b["shape"][:n_dyn] = 0
    b["shape"][n_dyn] = 1
    b["mass"][n_dyn] = 0.0
    b["flags"][n_dyn] = FLAG_STATIC
    return Scene(name=f"diff_stack_{n_dyn}", shape_type=shape_type, shape_half=shape_half,
                 bounds=np.array([-15, -2, -15, 15, 12, 15], F32), max_pairs=max(256, 48 * n),
                 max_contacts=max(64, 16 * n), frames=120, meta={"target": target}, **b)


def diff_batch(n_scenes: int = 256, seed: int = 20260004) -> list[Scene]:
    rng = Pcg32(seed)
    return [diff_stack(rng) for _ in range(n_scenes)]


# ---------------------------------------------------------------------------
# Batched + differentiable at scale: random_mixed's boxes and spheres at half
# scale per scene, lognormal body counts, 4 substeps per frame.
# ---------------------------------------------------------------------------
def diff_mixed(n_dyn: int, rng: Pcg32, scale: float = 0.5, extent=(30.0, 40.0, 30.0), frames: int = 240,
               n_substeps: int = 4, target_sigma: float = 0.5) -> Scene:
    """random_mixed(n_dyn) at `scale` (its seed drawn from rng), with
    meta["target"] = x0 + N(0, target_sigma) for the dynamic bodies."""
    sc = random_mixed(n_dyn, seed=int(rng.next_u32()), scale=scale, extent=extent, frames=frames)
    sc.n_substeps = n_substeps
    tgt = sc.pos[:n_dyn].astype(np.float64).copy()
    for k in range(n_dyn):
        for a in range(3):
            tgt[k, a] += target_sigma * rng.normal()
    sc.meta["target"] = tgt
    return sc


def diff_at_scale(n_scenes: int = 1024, seed: int = 20260005, median: float = 128.0, sigma: float = 0.75,
                  lo: int = 16, hi: int = 1024) -> list[Scene]:
    rng = Pcg32(seed)
    counts = lognormal_counts(n_scenes, median, sigma, lo, hi, rng)
    return [diff_mixed(int(c), rng) for c in counts]


RUN_PARAMS = ("dt", "n_substeps", "n_iters", "mu", "gravity")


def concat(scenes: list[Scene], name: str = "batch", uniform=RUN_PARAMS) -> Scene:
    """Scenes laid end to end for the batched step: body arrays concatenated,
    each body's shape index offset into the concatenated catalog.  meta holds
    body_offset (n_scenes + 1) and the per-scene pair / contact capacities.
    The batch steps every scene with one set of run parameters, so each
    parameter named in `uniform` must be the same in every scene (ValueError
    otherwise)."""
    if not scenes:
        raise ValueError("no scenes")
    for key in uniform:
        v0 = tuple(np.atleast_1d(getattr(scenes[0], key)).tolist())
        for i, sc in enumerate(scenes[1:], 1):
            if tuple(np.atleast_1d(getattr(sc, key)).tolist()) != v0:
                raise ValueError(f"scene {i} has {key} = {getattr(sc, key)!r}, scene 0 has "
                                 f"{getattr(scenes[0], key)!r}: a batch steps every scene with the same {key}")
    body_off = np.zeros(len(scenes) + 1, np.uint32)
    body_off[1:] = np.cumsum([sc.n for sc in scenes])
    shape_off = np.zeros(len(scenes) + 1, np.int64)
    shape_off[1:] = np.cumsum([sc.n_shapes for sc in scenes])
    cat = lambda key: np.concatenate([getattr(sc, key) for sc in scenes])
    shape = np.concatenate([sc.shape.astype(np.int64) + shape_off[i] for i, sc in enumerate(scenes)])
    first = scenes[0]
    return Scene(name=name, shape_type=cat("shape_type"), shape_half=cat("shape_half"),
                 pos=cat("pos"), rot=cat("rot"), vel=cat("vel"), ang=cat("ang"), mass=cat("mass"),
                 inertia=cat("inertia"), shape=shape.astype(np.uint32), flags=cat("flags"),
                 bounds=first.bounds.copy(), max_pairs=int(sum(sc.max_pairs for sc in scenes)),
                 max_contacts=int(sum(sc.max_contacts for sc in scenes)), dt=first.dt,
                 n_substeps=first.n_substeps, n_iters=first.n_iters, mu=first.mu, gravity=first.gravity,
                 frames=first.frames,
                 meta=dict(body_offset=body_off,
                           max_pairs=np.array([sc.max_pairs for sc in scenes], np.uint32),
                           max_contacts=np.array([sc.max_contacts for sc in scenes], np.uint32)))


JOINT_F32 = ("anchor_px", "anchor_py", "anchor_pz", "anchor_cx", "anchor_cy", "anchor_cz",
             "axis_px", "axis_py", "axis_pz", "limit_lo", "limit_hi", "target_pos", "target_vel",
             "stiffness", "damping", "compliance_pos", "compliance_ang")
JOINT_FIXED, JOINT_REVOLUTE, JOINT_PRISMATIC, JOINT_BALL = 0, 1, 2, 3


def joint_arrays(n: int) -> dict:
    """n joints with JointParams' defaults (core/include/core/articulation.hpp:
    16-42): axis +y, limits disabled (lo 1 > hi -1), no motor, rigid."""
    d = {f: np.zeros(n, F32) for f in JOINT_F32}
    d["body_parent"] = np.zeros(n, np.uint32)
    d["body_child"] = np.zeros(n, np.uint32)
    d["type"] = np.zeros(n, np.uint8)
    d["axis_py"][:] = 1.0
    d["limit_lo"][:] = 1.0
    d["limit_hi"][:] = -1.0
    return d


# ---------------------------------------------------------------------------
# core/tests/sim_determinism_test.cpp:130-190: static base + two 1 kg links,
# revolute about x with PD motors (targets +0.5, -0.5 rad) and limits +-2.9.
# The reference runs integrate + solve only (no collision), 20 frames.
# ---------------------------------------------------------------------------
def articulated_arm() -> Scene:
    n = 3
    hy = F32(0.15)
    sep = F32(2.0) * hy
    i_diag = F32(1.0) / F32(3.0) * F32(2.0) * hy * hy
    b = _alloc(n)
    for i in (1, 2):
        b["pos"][i] = (0.0, F32(i) * sep, 0.0)
        b["inertia"][i] = _diag(i_diag)
    b["mass"][:] = (0.0, 1.0, 1.0)
    b["flags"][0] = FLAG_STATIC
    j = joint_arrays(2)
    j["body_parent"][:] = (0, 1)
    j["body_child"][:] = (1, 2)
    j["anchor_py"][:] = hy
    j["anchor_cy"][:] = -hy
    j["axis_px"][:], j["axis_py"][:] = 1.0, 0.0
    j["limit_lo"][:], j["limit_hi"][:] = -2.9, 2.9
    j["target_pos"][:] = (0.5, -0.5)
    j["stiffness"][:], j["damping"][:] = 200.0, 30.0
    j["type"][:] = JOINT_REVOLUTE
    return Scene(name="articulated_arm", shape_type=np.array([SHAPE_SPHERE], np.uint32),
                 shape_half=np.array([[0.05, 0.05, 0.05]], F32), bounds=np.array([-4, -4, -4, 4, 4, 4], F32),
                 max_pairs=32, max_contacts=16, frames=20, joints=j, **b)


# ---------------------------------------------------------------------------
# Chains hanging from static bases over a static ground: links are boxes
# (half 0.08 x 0.15 x 0.08) or spheres (r 0.12) joined parent-above-child by
# a seeded mix of the four joint types, with limits around the rest pose,
# PD motors and compliance on some, and small initial velocities.  Prismatic
# joints carry a holding motor and, only with prismatic_limits, limits: the
# reference's prismatic limit correction (core/src/xpbd_solver.cpp:338-345)
# moves the bodies along +-axis with the sign of the anchor-difference
# constraint while its coordinate is child minus parent, so an active limit
# grows instead of closing; it is reproduced as written.  Fixed and
# prismatic joints hang only from the static bases: the reference's
# orientation lock (solve_fixed_ang, core/src/xpbd_solver.cpp:234-252, also
# part of every prismatic joint) takes q_b conj(q_a), already a world-frame
# rotation, and rotates it by q_a once more, which skews its axis whenever
# the parent turns and lets chains of them diverge.  Chains
# stand 0.32 apart, so neighbours and the lowest links touch: contacts and
# joints share bodies.
# ---------------------------------------------------------------------------
def joint_chains(n_chains: int = 6, links: int = 5, seed: int = 20260005, spacing: float = 0.32,
                 prismatic_limits: bool = False) -> Scene:
    rng = Pcg32(seed)
    pitch = 0.34
    n = n_chains * (links + 1) + 1
    b = _alloc(n)
    side = max(1, int(math.ceil(math.sqrt(n_chains))))
    top = 0.6 + links * pitch
    shape_type = np.array([SHAPE_BOX, SHAPE_SPHERE, SHAPE_BOX, SHAPE_BOX], np.uint32)
    shape_half = np.array([[0.08, 0.15, 0.08], [0.12, 0.12, 0.12], [0.1, 0.1, 0.1],
                           [0.5 * side * spacing + 2.0, 0.5, 0.5 * side * spacing + 2.0]], F32)
    j = joint_arrays(n_chains * links)
    k = 0
    for c in range(n_chains):
        cx = (c % side - (side - 1) / 2.0) * spacing
        cz = (c // side - (side - 1) / 2.0) * spacing
        base = c * (links + 1)
        b["pos"][base] = (cx, top, cz)
        b["mass"][base] = 0.0
        b["flags"][base] = FLAG_STATIC
        b["shape"][base] = 2
        for li in range(links):
            i = base + 1 + li
            sph = rng.uniform() < 0.35
            b["shape"][i] = 1 if sph else 0
            b["pos"][i] = (cx + rng.uniform_range(-0.01, 0.01), top - (li + 1) * pitch, cz + rng.uniform_range(-0.01, 0.01))
            m = 1000.0 * (4.0 / 3.0 * math.pi * 0.12 ** 3 if sph else 8.0 * 0.08 * 0.15 * 0.08)
            b["mass"][i] = m
            if sph:
                b["inertia"][i] = _diag(0.4 * m * 0.12 * 0.12)
            else:
                b["inertia"][i] = np.array([m / 3.0 * (0.15 ** 2 + 0.08 ** 2), 0, 0, 0, m / 3.0 * (2 * 0.08 ** 2), 0,
                                            0, 0, m / 3.0 * (0.15 ** 2 + 0.08 ** 2)], F32)
            b["vel"][i] = (0.3 * rng.normal(), 0.1 * rng.normal(), 0.3 * rng.normal())
            b["ang"][i] = (0.5 * rng.normal(), 0.5 * rng.normal(), 0.5 * rng.normal())
            parent = base + li
            j["body_parent"][k], j["body_child"][k] = parent, i
            j["anchor_py"][k], j["anchor_cy"][k] = -0.5 * pitch, 0.5 * pitch  # midway between the centres
            t = int(rng.uniform() * 4.0) % 4
            if li > 0 and t == JOINT_FIXED:
                t = JOINT_BALL
            elif li > 0 and t == JOINT_PRISMATIC:
                t = JOINT_REVOLUTE
            j["type"][k] = t
            ax = int(rng.uniform() * 3.0) % 3
            axis = [0.0, 0.0, 0.0]
            axis[ax] = 1.0
            if t == JOINT_PRISMATIC:
                axis = [0.0, 1.0, 0.0]
            j["axis_px"][k], j["axis_py"][k], j["axis_pz"][k] = axis
            if t == JOINT_REVOLUTE and rng.uniform() < 0.7:
                w = rng.uniform_range(0.2, 1.2)
                j["limit_lo"][k], j["limit_hi"][k] = -w, w
            if t == JOINT_PRISMATIC and prismatic_limits and rng.uniform() < 0.7:
                rest = -pitch
                j["limit_lo"][k], j["limit_hi"][k] = rest - 0.04, rest + 0.04
            if t == JOINT_PRISMATIC or (t == JOINT_REVOLUTE and rng.uniform() < 0.6):
                j["stiffness"][k] = rng.uniform_range(0.0, 150.0) if t == JOINT_REVOLUTE else rng.uniform_range(150.0, 400.0)
                j["damping"][k] = rng.uniform_range(0.0, 20.0) if t == JOINT_REVOLUTE else rng.uniform_range(10.0, 30.0)
                j["target_pos"][k] = rng.uniform_range(-0.4, 0.4) if t == JOINT_REVOLUTE else -pitch + rng.uniform_range(-0.03, 0.03)
                j["target_vel"][k] = rng.uniform_range(-0.5, 0.5)
            if rng.uniform() < 0.4:
                j["compliance_pos"][k] = rng.uniform_range(1e-6, 1e-4)
                j["compliance_ang"][k] = rng.uniform_range(1e-6, 1e-4)
            k += 1
    g = n - 1
    b["mass"][g] = 0.0
    b["flags"][g] = FLAG_STATIC
    b["shape"][g] = 3
    b["pos"][g] = (0.0, -0.5 + 0.05, 0.0)
    half = float(shape_half[3, 0]) + 2.0
    return Scene(name=f"joint_chains_{n_chains}x{links}", shape_type=shape_type, shape_half=shape_half,
                 bounds=np.array([-half, -2.0, -half, half, top + 2.0, half], F32),
                 max_pairs=max(256, 48 * n), max_contacts=max(64, 16 * n), frames=60, joints=j, **b)


def two_body(shape_a: int, shape_b: int, half_a, half_b, pos_a, pos_b, mass=(1.0, 1.0),
             flags=(0, 0), rot_a=(1, 0, 0, 0), rot_b=(1, 0, 0, 0), vel_a=(0, 0, 0),
             vel_b=(0, 0, 0), inertia=(1.0, 1.0)) -> Scene:
    """Two bodies with their own shapes, as the reference's single-pair KATs build."""
    b = _alloc(2)
    b["pos"][:] = [pos_a, pos_b]
    b["rot"][:] = [rot_a, rot_b]
    b["vel"][:] = [vel_a, vel_b]
    b["mass"][:] = mass
    b["flags"][:] = flags
    b["shape"][:] = [0, 1]
    b["inertia"][0] = _diag(inertia[0])
    b["inertia"][1] = _diag(inertia[1])
    return Scene(name="two_body", shape_type=np.array([shape_a, shape_b], np.uint32),
                 shape_half=np.array([half_a, half_b], F32),
                 bounds=np.array([-4, -4, -4, 4, 4, 4], F32), max_pairs=32, max_contacts=16,
                 frames=1, **b)
