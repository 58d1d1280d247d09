"""GPU adjoint (ckmpm_adjoint_*, ckmpm_amd.autodiff) against the FP64 reference.

The FP64 reference (oracle/adjoint_ref.py) is pinned to the C++ oracle and to central
differences by tests/test_adjoint_ref.py.  Here the FP32 GPU forward must reproduce its
final positions, and the GPU gradients w.r.t. per-particle log E and initial velocity
(through three checkpointed segments) must match its autograd gradients.
"""
import numpy as np
import pytest
import torch

from ckmpm_amd import scenes

NU = 0.3


def _case(seed=6, kind=scenes.FIXED_COROTATED):
    sc = scenes.cantilever(seed=seed, scale=4)
    m = sc.models[0]
    m.material.kind = kind
    rng = scenes.Splitmix(seed + 100)
    m.v = (0.3 * rng.normal(3 * m.x.shape[0]).reshape(-1, 3)).astype(np.float32)
    logE = np.log(m.mu_p.astype(np.float64) * 2 * (1 + NU))
    return sc, logE


def _reference(sc, logE, v0, steps, target):
    from oracle.adjoint_ref import DenseMPM, centre_of_mass_loss

    m = sc.models[0]
    sim = DenseMPM(sc.dim, sc.dx, sc.gravity, m.mass, m.volume, m.material.kind, sc.mass_clamp)
    lE = torch.tensor(logE, requires_grad=True)
    v = torch.tensor(v0.astype(np.float64), requires_grad=True)
    E = torch.exp(lE)
    lam, mu = E * NU / ((1 + NU) * (1 - 2 * NU)), E / (2 * (1 + NU))
    x, _, _ = sim.run(torch.tensor(m.x.astype(np.float64)), v, lam, mu, steps, sc.dt)
    L = centre_of_mass_loss(x, target)
    gE, gv = torch.autograd.grad(L, [lE, v])
    return x.detach().numpy(), gE.numpy(), gv.numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [scenes.FIXED_COROTATED, scenes.NEO_HOOKEAN])
def test_gpu_gradients_match_fp64_reference(kind):
    from ckmpm_amd.autodiff import simulate

    sc, logE0 = _case(kind=kind)
    steps, segment = 12, 5
    v00 = sc.models[0].v.copy()
    logE = torch.tensor(logE0, requires_grad=True)
    v0 = torch.tensor(v00.astype(np.float64), requires_grad=True)
    x = simulate(logE, v0, sc, steps, segment)
    target = x.detach().mean(0) + 1e-3
    L = ((x.mean(0) - target) ** 2).sum()
    L.backward()
    x_ref, gE_ref, gv_ref = _reference(sc, logE0, v00, steps, target)
    err_x = np.abs(x.detach().numpy() - x_ref).max()
    eE, ev = _rel(logE.grad.numpy(), gE_ref), _rel(v0.grad.numpy(), gv_ref)
    print(f"max |x - x_ref| = {err_x / sc.dx:.2e} dx; grad rel err logE {eE:.2e}, v0 {ev:.2e}")
    assert err_x <= 1e-4 * sc.dx
    assert np.abs(gE_ref).max() > 0 and np.abs(gv_ref).max() > 0
    assert eE < 1e-3
    assert ev < 1e-3


@pytest.mark.gpu
def test_checkpoint_restore_replays_the_forward():
    from ckmpm_amd.engine import engine_from_scene

    sc, _ = _case(seed=7)
    eng = engine_from_scene(sc, adjoint_segment=4)
    for k in range(3):
        eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
    ck = eng.checkpoint()
    for k in range(4):
        eng.Substep(sc.dt, sc.dt)
    eng.Synchronize()
    x1 = eng.positions_input_order(0)
    keys1, c1 = eng.layout(0)
    eng.restore(ck)
    eng.adjoint_record(True)
    for k in range(4):
        eng.Substep(sc.dt, sc.dt)
    eng.adjoint_record(False)
    eng.Synchronize()
    x2 = eng.positions_input_order(0)
    keys2, c2 = eng.layout(0)
    assert c1 == c2
    np.testing.assert_array_equal(keys1, keys2)
    # global coordinates (1 ulp near 0.45 is 1e-6 dx here), order-dependent FP32 flush
    np.testing.assert_allclose(x2, x1, rtol=0, atol=1e-5 * sc.dx)
    eng.free_checkpoint(ck)
    eng.close()


@pytest.mark.gpu
def test_rewound_engine_matches_a_new_one():
    """simulate() keeps its engine for the next call on the same scene and rewinds it to
    the setup state with the new velocities and Young's moduli (ckmpm_adjoint_rewind),
    reusing its checkpoint memory: positions and gradients equal those of a freshly built
    engine, whichever inputs the engine last ran with."""
    from ckmpm_amd import autodiff

    sc, logE_a = _case(seed=8)
    steps, segment = 12, 5
    rng = np.random.default_rng(3)
    v_a = sc.models[0].v.astype(np.float64)
    v_b = v_a + 0.1 * rng.standard_normal(v_a.shape)
    logE_b = logE_a + 0.2 * rng.standard_normal(logE_a.shape)
    target = torch.tensor([0.3, 0.5, 0.5], dtype=torch.float64)

    def run(logE, v):
        lE = torch.tensor(logE, requires_grad=True)
        vv = torch.tensor(v, requires_grad=True)
        x = autodiff.simulate(lE, vv, sc, steps, segment)
        ((x.mean(0) - target) ** 2).sum().backward()
        return x.detach().numpy(), lE.grad.numpy(), vv.grad.numpy()

    autodiff.clear_engine_cache()
    b_new = run(logE_b, v_b)  # built for b
    autodiff.clear_engine_cache()
    a_new = run(logE_a, v_a)  # built for a
    b_rew = run(logE_b, v_b)  # a's engine rewound to b
    a_rew = run(logE_a, v_a)  # and back to a
    autodiff.clear_engine_cache()
    # positions in global coordinates near 0.45 (1 ulp = 3e-8 = 1e-6 dx here); the P2G
    # flush's FP32 global atomics are order-dependent, so allow ~10 ulps
    for new, rew in ((a_new, a_rew), (b_new, b_rew)):
        np.testing.assert_allclose(rew[0], new[0], rtol=0, atol=1e-5 * sc.dx)
        assert _rel(rew[1], new[1]) < 1e-5 and _rel(rew[2], new[2]) < 1e-5
    # different inputs give different results (the rewind took the new ones)
    assert np.abs(a_new[0] - b_new[0]).max() > 1e-3 * sc.dx


@pytest.mark.gpu
@pytest.mark.parametrize("segment", [4, 5, 12, 20])
def test_checkpointed_gradients_match_the_full_tape(segment):
    """The forward records the last segment (steps 8..11 for segment 4, 10..11 for 5, the
    whole run for 12 and 20) and the backward replays only the segments before it: the
    gradients equal those of the run recorded whole (segment=None)."""
    from ckmpm_amd import autodiff

    sc, logE0 = _case(seed=9)
    steps = 12
    v00 = sc.models[0].v.astype(np.float64)
    target = torch.tensor([0.3, 0.5, 0.5], dtype=torch.float64)

    def run(seg):
        lE = torch.tensor(logE0, requires_grad=True)
        vv = torch.tensor(v00, requires_grad=True)
        x = autodiff.simulate(lE, vv, sc, steps, seg)
        ((x.mean(0) - target) ** 2).sum().backward()
        return x.detach().numpy(), lE.grad.numpy(), vv.grad.numpy()

    full = run(None)
    ck = run(segment)
    autodiff.clear_engine_cache()
    np.testing.assert_allclose(ck[0], full[0], rtol=0, atol=1e-5 * sc.dx)
    assert np.abs(full[1]).max() > 0 and np.abs(full[2]).max() > 0
    assert _rel(ck[1], full[1]) < 1e-5 and _rel(ck[2], full[2]) < 1e-5


@pytest.mark.gpu
def test_second_backward_and_foreign_checkpoint_are_refused():
    """A second backward through one simulate() call raises instead of replaying freed
    checkpoints; restoring a checkpoint into another engine returns CKMPM_ERR_INVALID."""
    from ckmpm_amd.autodiff import clear_engine_cache, simulate
    from ckmpm_amd.engine import engine_from_scene
    from ckmpm_amd.native import ERR_INVALID, CkmpmError

    sc, logE0 = _case()
    logE = torch.tensor(logE0, requires_grad=True)
    v0 = torch.tensor(sc.models[0].v.astype(np.float64), requires_grad=True)
    x = simulate(logE, v0, sc, 6, 3)
    L = (x.mean(0) ** 2).sum()
    L.backward(retain_graph=True)
    g = logE.grad.clone()
    with pytest.raises(RuntimeError, match="second time"):
        L.backward()
    assert torch.isfinite(g).all()
    clear_engine_cache()
    a = engine_from_scene(sc, adjoint_segment=4)
    b = engine_from_scene(sc, adjoint_segment=4)
    ck = a.checkpoint()
    with pytest.raises(CkmpmError) as ei:
        b.restore(ck)
    assert ei.value.code == ERR_INVALID
    a.restore(ck)
    a.free_checkpoint(ck)
    a.close()
    b.close()
