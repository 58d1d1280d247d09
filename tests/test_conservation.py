"""Per-substep conservation metric (ckmpm_set_conservation / ckmpm_get_conservation;
the reference's CollectConservationMetric, mpm_algorithm.cuh:1083-1357, and its
13-float conservationMetric file records, mpm_engine.cuh:494-603).

CPU: the oracle's metric on an imposed affine grid velocity field a + w x node equals
the closed form (G2P reproduces affine fields exactly: partition of unity on each
sub-grid and cancelling first moments across the two): linear sum v = N a + w x sum x,
angular sum x x v = sum x x (a + w x x).  A dropped sub-grid, a missing 1/2 or a wrong
cross-product sign fails it.
GPU: in lockstep with the oracle (the GPU state imported before every substep), each
substep's record equals the oracle's metric; records carry the simulated time of the
frame loop's timestamp; a full log stops the next substep with CKMPM_ERR_CAPACITY.
"""
import numpy as np
import pytest

from ckmpm_amd import scenes
from oracle.pyoracle import Oracle


def _node_positions(orc, sc):
    """Global positions of every grid node of the current layout's Nb blocks,
    (Nb, 128, 3): grid g node of cell index c sits at (c - 1/4 + g/2) dx."""
    keys = orc.keys(0)[: orc.counts(0)[1]].astype(np.int64)
    D = sc.dim
    b = np.stack([keys // (D[1] * D[2]), (keys // D[2]) % D[1], keys % D[2]], 1)
    n = np.arange(128)
    g, loc = n >> 6, np.stack([(n >> 4) & 3, (n >> 2) & 3, n & 3], 1)
    cell = 4 * b[:, None, :] + loc[None, :, :]
    return (cell - 0.25 + 0.5 * g[None, :, None]) * sc.dx


@pytest.mark.parametrize("case", ["spheres", "mixed"])
def test_oracle_metric_on_affine_field_matches_closed_form(case):
    sc = scenes.colliding_spheres() if case == "spheres" else scenes.mixed_dam_sand(n_fluid=4000, n_sand=3000,
                                                                                     scale=8)
    assert tuple(sc.offset) == (0, 0, 0)
    orc = Oracle.from_scene(sc)
    orc.setup()
    a = np.array([0.3, -1.2, 0.7])
    w = np.array([2.0, -0.5, 1.5])
    P = _node_positions(orc, sc)
    V = a + np.cross(w, P)                  # (Nb, 128, 3)
    orc.set_grid_v(np.ascontiguousarray(V.transpose(0, 2, 1)))
    for m in range(len(sc.models)):
        x = orc.state(m)[:3].T              # (n, 3) global FP64 positions
        v = a + np.cross(w, x)
        want = np.concatenate([v.sum(0), np.cross(x, v).sum(0)])
        got = orc.conservation(m)
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11 * np.abs(want).max())
    # no field, no momentum
    orc.set_grid_v(np.zeros_like(V.transpose(0, 2, 1)))
    assert np.abs(orc.conservation(0)).max() == 0.0


# ------------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("case", ["spheres", "mixed"])
def test_gpu_metric_matches_oracle_in_lockstep(case):
    from ckmpm_amd.engine import engine_from_scene
    from tests.test_gpu_parity import _import

    sc = scenes.colliding_spheres() if case == "spheres" else scenes.mixed_dam_sand(n_fluid=9_000, n_sand=5_000,
                                                                                     scale=8)
    steps = 6
    print(f"conservation lockstep {case}: {sc.n_particles} particles, {steps} substeps", flush=True)
    eng = engine_from_scene(sc)
    eng.set_conservation(steps)
    orc = Oracle.from_scene(sc)
    orc.setup()
    nm = len(sc.models)
    want = np.zeros((steps, nm, 6))
    t = [0.0]
    for k in range(steps):
        dt = 0.0 if k == 0 else sc.dt
        _import(eng, orc, nm)
        orc.physics(dt, sc.dt)
        for m in range(nm):
            want[k, m] = orc.conservation(m)
        eng.Substep(dt, sc.dt)
        t.append(t[-1] + sc.dt)
    eng.Synchronize()
    for m in range(nm):
        rec = eng.conservation(m)
        assert rec.shape == (steps, 13)
        np.testing.assert_allclose(rec[:, 0], np.float32(t[:steps]), rtol=1e-6)
        assert (rec[:, 1:7] == 0).all()  # Eulerian entries: never filled, as the reference
        n = eng.n(m)
        vmax = max(float(np.abs(want[:, m, :3]).max()) / n, float(eng.counters().max_speed), 1e-3)
        # FP32 G2P per particle against FP64: allow 1e-5 of the sum of |v| (and |x| <= 1)
        tol = 1e-5 * n * vmax
        err = np.abs(rec[:, 7:].astype(np.float64) - want[:, m]).max()
        print(f"  model {m}: max |gpu - oracle| {err:.3e}, tolerance {tol:.3e}", flush=True)
        assert err <= tol
    eng.close()


@pytest.mark.gpu
def test_gpu_metric_log_capacity_and_frame_loop_time():
    from ckmpm_amd.engine import engine_from_scene
    from ckmpm_amd.native import CkmpmError

    sc = scenes.colliding_spheres()
    print("conservation log capacity, colliding spheres", flush=True)
    eng = engine_from_scene(sc)
    eng.set_conservation(3)
    for k in range(3):
        eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
    with pytest.raises(CkmpmError, match="conservation log full"):
        eng.Substep(sc.dt, sc.dt)
    assert eng.conservation(1).shape == (3, 13)
    eng.set_conservation(1000)  # clears the log
    assert eng.conservation(0).shape == (0, 13)
    eng.Simulate(1)
    n = eng.counters().substeps - 3
    rec = eng.conservation(0)
    assert rec.shape == (n, 13)
    # the frame loop's timestamps: strictly increasing from 3 dt, all inside one frame
    assert rec[0, 0] == pytest.approx(3 * sc.dt, rel=1e-6)
    assert (np.diff(rec[:, 0]) > 0).all() and rec[-1, 0] < 3 * sc.dt + 1 / 48
    # colliding spheres: mirror images, so the net x momentum per unit mass nearly cancels
    lx = rec[:, 7] + eng.conservation(1)[:, 7]
    assert np.abs(lx).max() <= 1e-3 * eng.n(0)
    eng.set_conservation(0)
    eng.Substep(sc.dt, sc.dt)  # recording off: no capacity limit
    eng.close()
