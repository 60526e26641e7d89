"""Differentiable batched form (dy_diff_*): gradients of a thrown-stack loss
with respect to initial velocities and inverse masses, from the gfx950
reverse sweep (float32), against a float64 host forward-mode reference of the
same frames (tests/test_adjoint.py pins that reference against central
differences and against the oracle's trajectory).

Loss L = sum over dynamic bodies of |x_T - x*|^2 (the thrown-stack loss).  Per
scene, a random direction u over (v0, inverse mass) of its dynamic bodies:
the GPU's u . grad L must equal the reference's dL/du within 1e-3 relative.
"""
import ctypes as C

import numpy as np
import pytest

from dyphur_amd import scenes
from tests import _hostmath
from tests.test_adjoint import P, _sim_args

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _gpu_grads(gpu, scs, frames, every):
    b = gpu.Batch(scs)
    d = gpu.Diff(b, frames, every)
    d.forward(frames)
    st = b.state()
    n = b.n_bodies
    adj = {k: np.zeros(n, np.float32) for k in ("pos_x", "pos_y", "pos_z")}
    for i, sc in enumerate(scs):
        sl = b.scene_slice(i)
        nd = sc.n - 1
        for k, name in enumerate(("pos_x", "pos_y", "pos_z")):
            x = st[name][sl][:nd].astype(np.float64)
            adj[name][sl.start:sl.start + nd] = 2.0 * (x - sc.meta["target"][:, k])
    g = d.backward(adj)
    return b, st, g


def _host_directional(hm, sc, frames, u_v, u_im, fn):
    """dL/du by forward mode over the same frames (fn: the float64 or the
    float reference); returns it and the final state (13 x n)."""
    a = _sim_args(hm, sc)
    n = sc.n
    st, sd = a["st"].copy(), np.zeros(13 * n)
    for k in range(3):
        sd[(7 + k) * n:(8 + k) * n] = u_v[:, k]
    im_dot = u_im.astype(np.float64)
    fn(n, P(st), P(sd), P(a["im"]), P(im_dot), P(a["I6"]), P(a["stype"]), P(a["hs"]), P(a["flags"]), P(a["g"]),
       sc.dt, frames, sc.n_iters, sc.mu, 0, sc.max_pairs, sc.max_contacts)
    nd = n - 1
    x = st.reshape(13, n)[0:3, :nd].T
    xd = sd.reshape(13, n)[0:3, :nd].T
    return float(np.sum(2.0 * (x - sc.meta["target"]) * xd)), st.reshape(13, n)


@pytest.fixture(scope="module")
def hm():
    L = _hostmath.lib()
    vp, d, i = C.c_void_p, C.c_double, C.c_int
    L.ha_sim_dir_d.argtypes = [i] + [vp] * 9 + [d, i, i, d, i, C.c_uint32, C.c_uint32]
    L.ha_sim_dir_f.argtypes = [i] + [vp] * 9 + [d, i, i, d, i, C.c_uint32, C.c_uint32]
    return L


def test_diff_forward_equals_batch_step(gpu):
    scs = scenes.diff_batch(6, seed=11)
    a = gpu.Batch(scs)
    a.step(23)
    b = gpu.Batch(scs)
    d = gpu.Diff(b, 40, 8)
    d.forward(23)
    sa, sb = a.state(), b.state()
    for k in gpu.MUTABLE_FIELDS:
        assert np.array_equal(bits(sa[k]), bits(sb[k])), k


@pytest.mark.parametrize("frames,every", [(30, 8), (60, 10)])
def test_diff_gradient_matches_forward_mode(gpu, hm, frames, every):
    """Against the float reference (the kernels' own trajectory, bit for bit):
    the reverse sweep's u . grad L equals forward mode within 1e-3.  Against
    the float64 reference the difference is the gradient's sensitivity to
    float32 rounding of the trajectory: bounded by 1e-2 over the 30 frames
    before the stacks reach the ground; after ground impact the derivative
    grows by orders of magnitude (contact events) and float32 and float64
    trajectories no longer share it, so it is only printed there."""
    scs = scenes.diff_batch(4, seed=20260004)
    print(f"[diff] {len(scs)} scenes frames={frames} checkpoint={every}", flush=True)
    b, st, g = _gpu_grads(gpu, scs, frames, every)
    rng = np.random.default_rng(5)
    e32, e64 = [], []
    for i, sc in enumerate(scs):
        sl = b.scene_slice(i)
        nd = sc.n - 1
        u_v = rng.normal(size=(nd, 3))
        u_im = np.concatenate([rng.normal(size=nd) * 0.2, [0.0]])
        gv = np.stack([g["vel_x"][sl][:nd], g["vel_y"][sl][:nd], g["vel_z"][sl][:nd]], 1).astype(np.float64)
        gpu_d = float(np.sum(gv * u_v) + np.sum(g["inv_mass"][sl].astype(np.float64) * u_im))
        uv = np.vstack([u_v, np.zeros((1, 3))])
        f32_d, f32_st = _host_directional(hm, sc, frames, uv, u_im, hm.ha_sim_dir_f)
        f64_d, _ = _host_directional(hm, sc, frames, uv, u_im, hm.ha_sim_dir_d)
        for k, name in enumerate(("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z")):
            assert np.array_equal(bits(f32_st[k].astype(np.float32)), bits(st[name][sl])), (i, name)
        e32.append(abs(gpu_d - f32_d) / max(abs(f32_d), 1e-12))
        e64.append(abs(gpu_d - f64_d) / max(abs(f64_d), 1e-12))
        print(f"[diff] scene {i}: gpu {gpu_d:.6e} f32-fwd {f32_d:.6e} ({e32[-1]:.1e}) "
              f"f64-fwd {f64_d:.6e} ({e64[-1]:.1e})", flush=True)
    assert max(e32) <= 1e-3, e32
    if frames <= 30:
        assert max(e64) <= 1e-2, e64


def test_diff_static_ground_gets_no_gradient(gpu):
    scs = scenes.diff_batch(2, seed=3)
    b, _, g = _gpu_grads(gpu, scs, 40, 10)
    for i, sc in enumerate(scs):
        ground = b.scene_slice(i).stop - 1
        for k in g:
            assert g[k][ground] == 0.0, k


def test_torch_autograd_rollout_matches_reverse_sweep(gpu, hm):
    torch = pytest.importorskip("torch")
    from dyphur_amd import autograd as ag
    scs = scenes.diff_batch(3, seed=77)
    frames = 24
    r = ag.Rollout(scs, frames, 8)
    v0, im0 = r.initial()
    vel = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    inv = torch.tensor(im0, dtype=torch.float64, requires_grad=True)
    x = ag.rollout(vel, inv, r)
    target = torch.zeros_like(x)
    offs = r.batch.body_offset
    for i, sc in enumerate(scs):
        target[int(offs[i]):int(offs[i]) + sc.n - 1] = torch.tensor(sc.meta["target"])
    dyn = torch.tensor(~r.static, dtype=torch.float64)[:, None]
    loss = (((x - target) * dyn) ** 2).sum()
    loss.backward()
    # the same numbers through the C ABI wrapper directly
    b, st, g = _gpu_grads(gpu, scs, frames, 8)
    assert np.array_equal(x.detach().numpy().astype(np.float32)[:, 0], st["pos_x"])
    assert np.allclose(vel.grad.numpy()[:, 0], g["vel_x"], rtol=1e-6, atol=1e-6)
    assert np.allclose(inv.grad.numpy(), g["inv_mass"], rtol=1e-6, atol=1e-6)
    # and a directional check against the float forward-mode reference
    rng = np.random.default_rng(1)
    sc = scs[0]
    nd = sc.n - 1
    u_v = rng.normal(size=(nd, 3))
    u_im = np.concatenate([rng.normal(size=nd) * 0.2, [0.0]])
    want, _ = _host_directional(hm, sc, frames, np.vstack([u_v, np.zeros((1, 3))]), u_im, hm.ha_sim_dir_f)
    got = float(np.sum(vel.grad.numpy()[:nd] * u_v) + np.sum(inv.grad.numpy()[:sc.n] * u_im))
    assert abs(got - want) <= 1e-3 * abs(want), (got, want)


def test_torch_rollout_with_substeps_equals_the_c_abi(gpu):
    """autograd.Rollout passes substeps and the tape options through: its
    gradients equal capi.Diff's over the same frames at 2 substeps."""
    torch = pytest.importorskip("torch")
    from dyphur_amd import autograd as ag
    rng = scenes.Pcg32(20260006)
    scs = [scenes.diff_mixed(n, rng, extent=(7.0, 7.0, 7.0), frames=4, n_substeps=2) for n in (30, 60)]
    r = ag.Rollout(scs, 4, 2, n_substeps=2, tape_contacts_per_body=8.0)
    v0, im0 = r.initial()
    vel = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    inv = torch.tensor(im0, dtype=torch.float64, requires_grad=True)
    x = ag.rollout(vel, inv, r)
    (x ** 2).sum().backward()
    b = gpu.Batch(scs)
    d = gpu.Diff(b, 4, 2, n_substeps=2, tape_contacts_per_body=8.0)
    d.forward(4)
    st = b.state()
    g = d.backward({k: (2.0 * st[k].astype(np.float64)).astype(np.float32) for k in ("pos_x", "pos_y", "pos_z")})
    assert np.array_equal(x.detach().numpy().astype(np.float32)[:, 1], st["pos_y"])
    assert np.array_equal(vel.grad.numpy().astype(np.float32)[:, 2], g["vel_z"])
    assert np.array_equal(inv.grad.numpy().astype(np.float32), g["inv_mass"])


# ---- at scale: substeps, a scene reversed from global memory, tape chunks ----
def _mixed_batch(frames):
    """Three scenes of mixed boxes and spheres at half scale (as the
    batched + differentiable workload at scale has them), packed densely so
    contacts start at once; the 700-body scene's state (78 KB) exceeds
    64 KiB of LDS.  2 substeps per frame."""
    rng = scenes.Pcg32(20260005)
    return [scenes.diff_mixed(n, rng, extent=(7.0, 7.0, 7.0), frames=frames, n_substeps=2) for n in (40, 700, 120)]


def _grads(gpu, scs, frames, every, lds_limit=0, **kw):
    b = gpu.Batch(scs)
    b.set_lds_limit(lds_limit)
    d = gpu.Diff(b, frames, every, n_substeps=scs[0].n_substeps, **kw)
    d.forward(frames)
    st = b.state()
    n = b.n_bodies
    adj = {k: np.zeros(n, np.float32) for k in ("pos_x", "pos_y", "pos_z")}
    for i, sc in enumerate(scs):
        sl = b.scene_slice(i)
        nd = sc.n - 1
        for k, name in enumerate(("pos_x", "pos_y", "pos_z")):
            x = st[name][sl][:nd].astype(np.float64)
            adj[name][sl.start:sl.start + nd] = 2.0 * (x - sc.meta["target"][:, k])
    return b, d, st, d.backward(adj)


@pytest.mark.parametrize("lds_limit", [0, 65536])
def test_diff_substeps_and_global_scene_match_forward_mode(gpu, hm, lds_limit):
    """Gradients through 2 substeps per frame against the host float forward
    mode over the same substeps (dt / 2, 2 x frames), whose trajectory
    equals the kernels' bit for bit.  The 700-body scene is recorded and
    reversed by a workgroup in LDS under the device's limit, and from global
    memory under a 64 KiB limit."""
    frames, every = 6, 3
    scs = _mixed_batch(frames)
    print(f"[diff] mixed scenes {[sc.n for sc in scs]}, {frames} frames x 2 substeps, lds limit {lds_limit}",
          flush=True)
    b, d, st, g = _grads(gpu, scs, frames, every, lds_limit=lds_limit)
    rng = np.random.default_rng(12)
    for i, sc in enumerate(scs):
        sl = b.scene_slice(i)
        nd = sc.n - 1
        u_v = rng.normal(size=(nd, 3))
        u_im = np.concatenate([rng.normal(size=nd) * 0.2, [0.0]])
        gv = np.stack([g["vel_x"][sl][:nd], g["vel_y"][sl][:nd], g["vel_z"][sl][:nd]], 1).astype(np.float64)
        gpu_d = float(np.sum(gv * u_v) + np.sum(g["inv_mass"][sl].astype(np.float64) * u_im))
        uv = np.vstack([u_v, np.zeros((1, 3))])
        sub = _substep_scene(sc)
        want, f32_st = _host_directional(hm, sub, frames * sc.n_substeps, uv, u_im, hm.ha_sim_dir_f)
        for k, name in enumerate(("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z")):
            assert np.array_equal(bits(f32_st[k].astype(np.float32)), bits(st[name][sl])), (i, name)
        err = abs(gpu_d - want) / max(abs(want), 1e-12)
        print(f"[diff] scene {i} ({sc.n} bodies): gpu {gpu_d:.6e} f32-fwd {want:.6e} ({err:.1e})", flush=True)
        assert err <= 1e-3, (i, gpu_d, want)


def test_diff_workgroup_and_wavefront_teams_give_the_same_bits(gpu):
    """A 300-body scene stepped by a 256-thread workgroup team and by a
    single-wavefront team records tapes that differ only in the order of
    contacts inside a class, which the reverse sweep does not see: the
    gradients agree bit for bit (a workgroup team must finish scattering
    the class order before the record gathers it)."""
    frames, every = 4, 2
    rng = scenes.Pcg32(20260005)
    scs = [scenes.diff_mixed(300, rng, extent=(7.0, 7.0, 7.0), frames=frames, n_substeps=2)]
    got = []
    for shape in ((256, 128), (64, 128)):
        b = gpu.Batch(scs)
        b.set_shape(*shape)
        d = gpu.Diff(b, frames, every, n_substeps=2)
        d.forward(frames)
        st = b.state()
        nd = scs[0].n - 1
        adj = {k: np.zeros(b.n_bodies, np.float32) for k in ("pos_x", "pos_y", "pos_z")}
        for k, name in enumerate(("pos_x", "pos_y", "pos_z")):
            adj[name][:nd] = 2.0 * (st[name][:nd].astype(np.float64) - scs[0].meta["target"][:, k])
        got.append(d.backward(adj))
    for k in got[0]:
        assert np.array_equal(bits(got[0][k]), bits(got[1][k])), k


def _substep_scene(sc):
    """The same scene stepped at the substep: dt / n_substeps per frame."""
    import copy
    s2 = copy.copy(sc)
    s2.dt = sc.dt / sc.n_substeps
    s2.n_substeps = 1
    return s2


def test_diff_tape_chunks_give_the_same_bits(gpu):
    """A tape of one frame (segments replayed into sub-checkpoints, then
    recomputed and reversed frame by frame) gives the one-tape-per-segment
    gradients bit for bit."""
    frames, every = 6, 3
    scs = _mixed_batch(frames)
    _, d1, _, g1 = _grads(gpu, scs, frames, every)
    steps, step_bytes = d1.tape_layout()
    assert steps == every * 2
    _, d2, _, g2 = _grads(gpu, scs, frames, every, tape_bytes=2 * step_bytes + 64)
    assert d2.tape_steps() == 2  # one frame of 2 substeps per tape
    for k in g1:
        assert np.array_equal(bits(g1[k]), bits(g2[k])), k


def test_diff_recorded_horizon_and_recompute_give_the_same_bits(gpu):
    """Three 2-frame segments: by default all three records fit the two
    tapes' budget, so the forward records them and the backward only
    reverses; with a budget of one segment per tape the backward recomputes
    each older segment from its checkpoint.  Same gradient bits."""
    frames, every = 6, 2
    scs = _mixed_batch(frames)
    _, d1, _, g1 = _grads(gpu, scs, frames, every)
    steps, step_bytes = d1.tape_layout()
    assert steps == every * 2
    _, d2, _, g2 = _grads(gpu, scs, frames, every, tape_bytes=steps * step_bytes)
    assert d2.tape_steps() == steps
    for k in g1:
        assert np.array_equal(bits(g1[k]), bits(g2[k])), k


def test_diff_overfull_reduced_tape_is_an_error(gpu):
    """A tape that holds fewer contacts than a step produced: the backward
    reports it instead of returning a wrong gradient."""
    scs = _mixed_batch(4)
    with pytest.raises(gpu.DyError) as e:
        _grads(gpu, scs, 4, 2, tape_contacts_per_body=0.01)
    assert e.value.status == -1


def _loss64(hm, sc, frames, v_shift, im_shift):
    """The loss after `frames` frames of the float64 host simulation from the
    scene's initial state with shifted initial velocities and inverse masses."""
    from tests.test_adjoint import _sim_args
    a = _sim_args(hm, sc)
    n = sc.n
    st = a["st"].copy()
    for k in range(3):
        st[(7 + k) * n:(8 + k) * n] += v_shift[:, k]
    im = a["im"] + im_shift
    hm.ha_sim_d(n, P(st), P(im), P(a["I6"]), P(a["stype"]), P(a["hs"]), P(a["flags"]), P(a["g"]),
                sc.dt, frames, sc.n_iters, sc.mu, 0, sc.max_pairs, sc.max_contacts)
    x = st.reshape(13, n)[0:3, :n - 1].T
    return float(np.sum((x - sc.meta["target"]) ** 2))


@pytest.mark.parametrize("frames,bar", [(10, 1e-3), (20, 1e-3), (30, 1e-2)])
def test_gradient_against_fp64_central_differences(gpu, hm, frames, bar):
    """The reverse sweep's directional derivative against float64 central
    differences of the float64 simulation's loss, on active-set-stable
    perturbations: the difference quotient at step eps = 1e-6 and at 2.5e-7
    must agree within 1e-6 relative (a perturbation that switches a contact
    on or off between them is not differentiable there and is left out).
    The float64 forward mode is held to the same difference quotients within
    1e-6.  Bar: 1e-3 relative through 20 frames (measured: at most 9.5e-5);
    at 30 frames 1e-2 (measured: 2.8e-3 for one stack of eight, the others
    within 3e-4), where the float32 trajectory's derivative has separated
    from the float64 one through its contacts (rounding amplified by contact
    events; the float forward mode of the same float32 trajectory agrees
    with the GPU within 1e-3, test_diff_gradient_matches_forward_mode).
    Beyond about 30 frames of these thrown stacks the difference quotients
    stop converging at any step size (the loss is not differentiable at that
    resolution), so no longer horizon is checked this way."""
    scs = scenes.diff_batch(8, seed=20260004)
    b, st, g = _gpu_grads(gpu, scs, frames, 5)
    rng = np.random.default_rng(5)
    stable, errs = 0, []
    for i, sc in enumerate(scs):
        sl = b.scene_slice(i)
        nd = sc.n - 1
        u_v = rng.normal(size=(nd, 3))
        u_im = np.concatenate([rng.normal(size=nd) * 0.2, [0.0]])
        uv = np.vstack([u_v, np.zeros((1, 3))])
        gv = np.stack([g["vel_x"][sl][:nd], g["vel_y"][sl][:nd], g["vel_z"][sl][:nd]], 1).astype(np.float64)
        gpu_d = float(np.sum(gv * u_v) + np.sum(g["inv_mass"][sl].astype(np.float64) * u_im))
        fd = []
        for eps in (1e-6, 2.5e-7):
            fd.append((_loss64(hm, sc, frames, eps * uv, eps * u_im) -
                       _loss64(hm, sc, frames, -eps * uv, -eps * u_im)) / (2 * eps))
        spread = abs(fd[0] - fd[1]) / max(abs(fd[1]), 1e-12)
        f64_d, _ = _host_directional(hm, sc, frames, uv, u_im, hm.ha_sim_dir_d)
        err = abs(gpu_d - fd[1]) / max(abs(fd[1]), 1e-12)
        print(f"[fd] {frames} frames scene {i}: gpu {gpu_d:+.6e} fd {fd[1]:+.6e} (spread {spread:.1e}) "
              f"f64 forward {f64_d:+.6e} -> gpu error {err:.1e}", flush=True)
        if spread > 1e-6:
            continue
        stable += 1
        errs.append(err)
        assert abs(f64_d - fd[1]) <= 1e-6 * abs(fd[1]), (i, f64_d, fd)
        assert err <= bar, (i, gpu_d, fd[1])
    assert stable >= 4, f"only {stable} of 8 perturbations are active-set-stable"
