"""bench.py's frame planning and host-side helpers (no GPU).

The driver runs `bench.py --steps 20`: the timed frames must be the last 20
frames of the 600-frame run, and every CPU-baseline window must start and end
at frames the untimed replay snapshots (inside the run)."""
import math

import numpy as np
import pytest

import bench


@pytest.mark.parametrize("steps", [1, 5, 20, 150, 599, 600, 1000])
@pytest.mark.parametrize("window", [0, 1, 20, 300])
def test_plan_frames_stays_inside_the_run(steps, window):
    p = bench.plan_frames(steps, 600, window)
    assert p["total"] == max(600, steps)
    assert p["start"] + p["steps"] == p["total"]
    assert p["start"] >= 0 and p["steps"] == steps
    if window == 0:
        assert p["window"] == 0 and p["starts"] == []
        return
    assert 1 <= p["window"] <= window
    assert p["starts"] == sorted(set(p["starts"]))
    assert p["starts"][0] == 0
    for f0 in p["starts"]:
        assert 0 <= f0 and f0 + p["window"] <= p["total"]
    # windows spread over the run: the last one ends at the run's end
    assert p["starts"][-1] + p["window"] == p["total"]


def test_plan_frames_driver_command():
    # python bench.py --gpus 1 --steps 20 --warmup 5
    p = bench.plan_frames(20, 600, 20)
    assert (p["start"], p["total"]) == (580, 600)
    assert p["starts"] == [0, 193, 386, 580]


def test_roofline_picks_the_slowest_kernel():
    k = {"a": (0.1, 8e8), "b": (0.5, 1e9), "c": (0.2, 4e9)}
    r = bench.roofline_of(k)
    assert r["kernel"] == "b"
    assert math.isclose(r["achieved"], 1e9 / 0.5e-3 / 1e9, rel_tol=1e-6)
    assert math.isclose(r["frac"], r["achieved"] / 8000.0, rel_tol=1e-4)
    assert math.isclose(r["frame"]["kernel_ms"], 0.8, rel_tol=1e-9)


def test_kinetic_energy_closed_form():
    """Translational m v^2/2 plus rotational w.I_world.w/2 with I_world = R I R^T:
    a body turned 90 degrees about z spinning about world x sees its body-y
    moment."""
    from dyphur_amd import scenes
    sc = scenes.two_body(scenes.SHAPE_BOX, scenes.SHAPE_BOX, (0.5, 0.5, 0.5), (0.5, 0.5, 0.5),
                         (0, 2, 0), (0, 4, 0), mass=(2.0, 3.0))
    sc.inertia[0] = np.array([1, 0, 0, 0, 2, 0, 0, 0, 3], np.float32)
    c = math.cos(math.pi / 4)
    st = {k: np.zeros(sc.n, np.float32) for k in ("vel_x", "vel_y", "vel_z", "ang_x", "ang_y", "ang_z",
                                                   "rot_x", "rot_y")}
    st["rot_w"] = np.array([c, 1, 1][:sc.n], np.float32)
    st["rot_z"] = np.array([c, 0, 0][:sc.n], np.float32)
    st["vel_y"][0] = 3.0
    st["ang_x"][0] = 2.0
    st["vel_x"][1] = 1.0
    dyn = (sc.flags & 1) == 0
    want = 0.0
    if dyn[0]:
        want += 0.5 * 2.0 * 9.0 + 0.5 * 2.0 * 4.0  # world x = body -y after 90 degrees about z
    if dyn[1]:
        want += 0.5 * 3.0 * 1.0
    assert math.isclose(bench.kinetic_energy(st, sc), want, rel_tol=1e-6)


@pytest.mark.parametrize("field,value", [("dt", 1.0 / 120.0), ("n_substeps", 2), ("n_iters", 4), ("mu", 0.2),
                                         ("gravity", (0.0, -1.62, 0.0))])
def test_batch_rejects_scenes_with_different_run_parameters(field, value):
    """A batch steps all its scenes with one dt, substep count, iteration count,
    friction and gravity: a list that disagrees is refused before any device
    work (scenes.concat, which Batch calls first)."""
    from dyphur_amd import capi, scenes
    a, b = scenes.stress_blocks(2, 2, 2), scenes.stress_blocks(2, 3, 2)
    setattr(b, field, value)
    with pytest.raises(ValueError, match=field):
        scenes.concat([a, b])
    with pytest.raises(ValueError, match=field):
        capi.Batch([a, b])
    if field in ("n_iters", "mu"):  # Batch's explicit override applies one value to every scene
        uniform = [k for k in scenes.RUN_PARAMS if k != field]
        assert scenes.concat([a, b], uniform=uniform).n == a.n + b.n
