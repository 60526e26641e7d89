"""The drivers' file formats (examples/common/scene_io.hpp:23-134,
examples/stress_test_blocks/main.cpp:172-197,301-336): Python round trips,
byte-level layout, files written by the C++ header read back by Python, and
(on the GPU) the C++ stress driver's outputs against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from dyphur_amd import formats, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_scene_roundtrip_and_layout(tmp_path):
    p = str(tmp_path / "s")
    formats.write_scene(p, [0, 0, 1], [0, 1], [[0.4, 0.4, 0.4], [1.0, 0.5, 2.0]])
    raw = open(p + ".scene", "rb").read()
    assert len(raw) == 8 + 3 * 4 + 2 * 20
    assert np.frombuffer(raw, "<u4", 2).tolist() == [3, 2]
    # second shape record: type 1, half (1, .5, 2), pad 0
    rec = raw[8 + 12 + 20: 8 + 12 + 40]
    assert np.frombuffer(rec[:4], "<u4")[0] == 1
    assert np.frombuffer(rec[4:16], "<f4").tolist() == [1.0, 0.5, 2.0]
    assert np.frombuffer(rec[16:], "<u4")[0] == 0
    sf = formats.read_scene(p)
    assert sf.body_shape_idx.tolist() == [0, 0, 1] and sf.shape_type.tolist() == [0, 1]


def test_trajectory_golden_metrics_meshmap_roundtrip(tmp_path):
    p = str(tmp_path / "t")
    rng = np.random.default_rng(0)
    frames = []
    with formats.TrajectoryWriter(p, 5, 6) as w:
        for _ in range(3):
            st = {k: rng.standard_normal(5).astype(np.float32)
                  for k in ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z")}
            w.write_frame(st)
            frames.append(st)
    n, nf, got = formats.read_trajectory(p)
    assert (n, nf, got.shape) == (5, 6, (3, 5, 7))
    assert np.array_equal(got[2, :, 3], frames[2]["rot_w"])
    formats.write_golden(p, 0xCBCD209C3665C818)
    assert open(p + ".golden").read() == "cbcd209c3665c818\n"
    assert formats.read_golden(p) == 0xCBCD209C3665C818
    m = dict(n_bodies=5, n_frames=6, simulation_seconds=0.1, wall_clock_seconds=0.01,
             frames_per_second=600.0, avg_contacts_per_frame=3.0, realtime_factor=10.0)
    formats.write_metrics(p, m)
    assert formats.read_metrics(p) == m
    formats.write_meshmap(p, [(3, "a.obj"), (4, "b.obj")])
    with open(p + ".meshmap", "a") as f:
        f.write("# comment\n\n")
    assert formats.read_meshmap(p) == {3: "a.obj", 4: "b.obj"}
    assert formats.read_meshmap(p + "_absent") == {}


CPP_IO = r'''
#include <io/scene_io.hpp>
#include <cstdio>
using namespace dyphur;
int main(int, char **argv) {
    const std::string p = argv[1];
    uint32_t bsi[3] = {0, 1, 1};
    ShapeParams sh[2];
    sh[0].type = ShapeType::Sphere; sh[0].half_x = 0.25f;
    sh[1].half_x = 1.f; sh[1].half_y = 2.f; sh[1].half_z = 3.f;
    write_scene(p, 3, bsi, sh, 2);
    SceneFileDesc d = read_scene(p);
    if (d.n_bodies != 3 || d.shapes[1].half_z != 3.f || d.shapes[0].type != ShapeType::Sphere) return 3;
    TrajectoryWriter w(p, 3, 4);
    float a[3] = {1, 2, 3}, b[3] = {4, 5, 6}, c[3] = {7, 8, 9}, q[3] = {1, 1, 1}, z[3] = {0, 0, 0};
    w.write_frame(a, b, c, q, z, z, z);
    w.write_frame(c, b, a, q, z, z, z);
    write_golden(p, 0x0123456789abcdefull);
    RunMetrics m; m.n_bodies = 3; m.n_frames = 4; m.frames_per_second = 123.456;
    write_metrics(p, m);
    write_meshmap(p, {{2, "mesh.obj"}});
    return read_meshmap(p).at(2) == "mesh.obj" ? 0 : 4;
}
'''


def test_cpp_writer_python_reader(tmp_path):
    src = tmp_path / "io.cpp"
    src.write_text(CPP_IO)
    exe = tmp_path / "io"
    subprocess.run(["g++", "-std=c++20", "-O1", f"-I{ROOT}/include", f"-I{ROOT}/dyphur_amd/cpp/include",
                    str(src), "-o", str(exe)], check=True, capture_output=True, timeout=120)
    p = str(tmp_path / "x")
    assert subprocess.run([str(exe), p], timeout=30).returncode == 0
    sf = formats.read_scene(p)
    assert sf.body_shape_idx.tolist() == [0, 1, 1] and sf.shape_type.tolist() == [1, 0]
    assert sf.shape_half[1].tolist() == [1.0, 2.0, 3.0]
    n, nf, fr = formats.read_trajectory(p)
    assert (n, nf, fr.shape) == (3, 4, (2, 3, 7))
    assert fr[1, :, 0].tolist() == [7, 8, 9] and fr[0, 2].tolist() == [3, 6, 9, 1, 0, 0, 0]
    assert formats.read_golden(p) == 0x0123456789ABCDEF
    m = formats.read_metrics(p)
    assert m["n_bodies"] == 3 and abs(m["frames_per_second"] - 123.46) < 1e-9


@pytest.mark.gpu
def test_stress_driver_outputs_match_oracle(gpu, oracle, tmp_path):
    """The C++ stress driver (8x16x8 stack, 20 frames) writes the reference's
    four files; its golden hash and last trajectory frame equal the oracle's."""
    exe = os.path.join(ROOT, "dyphur_amd", "bin", "stress_test_blocks")
    p = str(tmp_path / "stress")
    print("[stress driver] 20 frames", flush=True)
    r = subprocess.run([exe, p, "20"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    sc = scenes.stress_blocks()
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    w = oracle.World.for_scene(sc)
    last = None
    for f in range(20):
        w.frame(b, s, sc)
        if f % 2 == 0:
            last = b.state()
    assert formats.read_golden(p) == b.fnv()
    n, nf, fr = formats.read_trajectory(p)
    assert (n, nf, fr.shape[0]) == (sc.n, 20, 10)
    for k, key in enumerate(("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z")):
        assert np.array_equal(fr[-1, :, k].view(np.uint32), last[key].view(np.uint32)), key
    sf = formats.read_scene(p)
    assert sf.body_shape_idx.tolist() == sc.shape.tolist()
    assert np.array_equal(sf.shape_half, sc.shape_half)
    m = formats.read_metrics(p)
    assert m["n_bodies"] == sc.n and m["n_frames"] == 20
