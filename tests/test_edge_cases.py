"""Edge cases of the substep on the GPU.

* One particle: with a single particle every grid node it touches carries its velocity,
  so G2P returns it exactly (partition of unity), C = 0 (the two sub-grids' first
  moments cancel) and F stays I: the motion is symplectic Euler, v_k = v_0 + k g dt and
  x_K = x_0 + dt sum v_k (the first substep advects with dt = 0).
* Block-list capacity: a scene whose layout outgrows the engine's block capacity stops
  with CKMPM_ERR_CAPACITY instead of writing past its buffers, and the engine can still
  be queried and destroyed.
"""
import numpy as np
import pytest

from ckmpm_amd import scenes

pytestmark = pytest.mark.gpu


def _one_particle(v0, gravity=-9.8, dt=1e-4):
    dx = 1.0 / 64
    lam, mu = scenes.lame(1e4, 0.3)
    vol = dx ** 3 / 8
    x = np.array([[0.503, 0.497, 0.51]], np.float32)
    m = scenes.ModelSpec(scenes.Material(scenes.FIXED_COROTATED, lam, mu), 1000 * vol, vol, x,
                         np.array([v0], np.float32))
    return scenes.Scene("one", (16, 16, 16), (0, 0, 0), dx, gravity, dt, 50, [m])


def test_single_particle_follows_symplectic_euler():
    from ckmpm_amd.engine import engine_from_scene

    v0 = np.array([0.3, 0.0, -0.2])
    sc = _one_particle(v0)
    steps = 50
    print(f"one particle, {steps} substeps", flush=True)
    eng = engine_from_scene(sc)
    for k in range(steps):
        eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
    eng.Synchronize()
    g = np.array([0.0, sc.gravity, 0.0])
    x = sc.models[0].x[0].astype(np.float64)
    for j in range(1, steps):
        x = x + sc.dt * (v0 + j * sc.dt * g)
    got = eng.positions(0)[0].astype(np.float64)
    np.testing.assert_allclose(got, x, rtol=0, atol=1e-5 * sc.dx)
    st = eng.global_state(0)
    np.testing.assert_allclose(st[3:12, 0], np.eye(3).reshape(9), rtol=0, atol=1e-6)   # F = I
    assert eng.counters().flags == 0
    eng.close()


def test_capacity_overflow_is_reported():
    from ckmpm_amd.engine import engine_from_scene
    from ckmpm_amd.native import CkmpmError

    # two spheres expanding radially at up to 30 m/s (0.19 cells per substep): the layout
    # grows with their surface; the inner faces close at 60 m/s, 0.16 m apart, so 20
    # substeps end before they meet (a translating body keeps its block count constant)
    sc = scenes.colliding_spheres(speed=0.0)
    for m, c in zip(sc.models, ((0.3, 0.5, 0.5), (0.7, 0.5, 0.5))):
        m.v = (30.0 * (m.x.astype(np.float64) - np.array(c)) / 0.12).astype(np.float32)
    probe = engine_from_scene(sc)
    _, (A, Nb, Ex) = probe.layout(0)
    probe.close()
    print(f"capacity overflow: setup layout {A, Nb, Ex}, capacity {Ex + 8}", flush=True)
    eng = engine_from_scene(sc, capacity=Ex + 8)
    with pytest.raises(CkmpmError, match="CAPACITY"):
        for k in range(20):
            eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
            eng.Synchronize()
    c = eng.counters()
    assert c.exterior_blocks <= Ex + 8 and c.neighbor_blocks <= Ex + 8
    eng.close()


def test_frame_loop_without_frame_rate_is_refused():
    """ckmpm_run_frames with sim.fps = 0 returns CKMPM_ERR_INVALID at once (a frame of
    1 / fps would never end) and leaves the engine usable."""
    from ckmpm_amd.engine import MPMConfig, MPMDomain, MPMEngine, MPMMaterial, MPMModel
    from ckmpm_amd.native import ERR_INVALID, CkmpmError

    sc = _one_particle(np.array([0.1, 0.0, 0.0]))
    m = sc.models[0]
    eng = MPMEngine()
    eng.Initialize(MPMDomain(sc.dim, sc.offset), MPMConfig(dx=sc.dx, gravity=sc.gravity, dt=sc.dt, fps=0,
                                                           max_active_block_count=sc.block_capacity()))
    mat = MPMMaterial(m.material.kind, m.material.lam, m.material.mu)
    eng.InitializeParticle([MPMModel(m.mass, m.volume, m.x, m.v, mat)])
    eng.InitialSetup()
    with pytest.raises(CkmpmError) as ei:
        eng.Simulate(1)
    assert ei.value.code == ERR_INVALID
    eng.Substep(0.0, sc.dt)
    eng.Synchronize()
    assert eng.counters().substeps == 1
    eng.close()
