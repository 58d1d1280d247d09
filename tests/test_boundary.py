"""Grid boundary descriptions (ckmpm_set_boundary): the reference's
ProcessGridCellVelocity scene hook (mpm_algorithm.cuh:795-799) as data.

CPU, the oracle, against a run without walls on identical input (same scene, same
rasterized grid, one grid update):
* slip walls: nodes within 8 cells of a face lose exactly their component along the
  normalized sum of the face normals (corners included); other nodes are unchanged;
* sticky walls given as data equal the reference's default block rule (8 cells = 2
  blocks), node for node;
* regions: fixed gives 0, velocity gives the value, rotation gives w x (cell dx - c)
  (orthogonal to w and to the lever arm, |w| |r_perp| long); later regions win; nodes
  outside every region are unchanged.
GPU: a box pressed into a corner with slip walls and fixed / velocity / rotation regions
runs in lockstep with the oracle; the adjoint refuses custom boundaries.
"""
import numpy as np
import pytest

from ckmpm_amd import scenes
from oracle.pyoracle import Oracle


def _corner_scene(boundary, seed=40, dim=12):
    """A box of FC particles over cells [4, 16)^3, inside the 8-cell margins of the three
    low faces, drifting into the corner."""
    dx = 1.0 / (4 * dim)
    rng = scenes.Splitmix(seed)
    g = np.stack(np.meshgrid(*[np.arange(4, 16)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = scenes._stratified_cells(rng, g.astype(np.float64), dx, 2).astype(np.float32)
    v = (0.3 * rng.normal(3 * x.shape[0]).reshape(-1, 3) - 0.5).astype(np.float32)
    vol = dx ** 3 / 8
    lam, mu = scenes.lame(2e5, 0.3)
    m = scenes.ModelSpec(scenes.Material(scenes.FIXED_COROTATED, lam, mu), 1000 * vol, vol, x, v)
    sc = scenes.Scene("corner", (dim, dim, dim), (0, 0, 0), dx, -9.8, 2e-5, 10, [m])
    sc.boundary = boundary
    return sc


def _regions(dx):
    return [
        scenes.Region((4, 10, 4), (6, 16, 16), scenes.REGION_FIXED),
        scenes.Region((12, 4, 4), (16, 16, 16), scenes.REGION_ROTATION, (0.0, 0.0, 20.0), (14 * dx, 10 * dx, 10 * dx)),
        scenes.Region((4, 14, 4), (16, 16, 6), scenes.REGION_VELOCITY, (0.1, 0.0, 0.0)),
    ]


def _grid_v(boundary):
    sc = _corner_scene(boundary)
    o = Oracle.from_scene(sc)
    o.setup()
    o.physics(0.0, sc.dt)
    keys = o.keys(0)[: o.counts(0)[1]].astype(np.int64)
    D = sc.dim
    b = np.stack([keys // (D[1] * D[2]), (keys // D[2]) % D[1], keys % D[2]], 1)
    n = np.arange(128)
    cell = 4 * b[:, None, :] + np.stack([(n >> 4) & 3, (n >> 2) & 3, n & 3], 1)[None]
    return o.grid_v().transpose(0, 2, 1), cell, o.grid_m()[:, 0] > 0, sc   # V (Nb, 128, 3)


def test_slip_walls_remove_the_normal_component():
    V0, cell, has, _ = _grid_v(scenes.Boundary(scenes.WALLS_NONE))
    V1, _, _, _ = _grid_v(scenes.Boundary(scenes.WALLS_SLIP))
    nrm = np.where(cell < 8, 1.0, 0.0) - np.where(cell >= 40, 1.0, 0.0)
    on = (nrm != 0).any(-1) & has
    assert on.sum() > 1000 and ((nrm != 0).sum(-1)[on] >= 2).any()       # faces and edges / corners
    nh = nrm[on] / np.linalg.norm(nrm[on], axis=-1, keepdims=True)
    assert np.abs((V1[on] * nh).sum(-1)).max() < 1e-12 * np.abs(V0).max()
    np.testing.assert_allclose(V0[on] - V1[on], (V0[on] * nh).sum(-1, keepdims=True) * nh, atol=1e-13)
    np.testing.assert_array_equal(V1[~on], V0[~on])


def test_sticky_walls_as_data_equal_the_default_block_rule():
    Vd, _, has, _ = _grid_v(None)
    Vs, _, _, _ = _grid_v(scenes.Boundary(scenes.WALLS_STICKY))
    np.testing.assert_array_equal(Vs, Vd)
    assert (np.abs(Vd[has]).sum(-1) == 0).sum() > 1000                     # the margin is zero


def test_regions_in_order_then_walls():
    V0, cell, has, sc = _grid_v(scenes.Boundary(scenes.WALLS_NONE))
    regs = _regions(sc.dx)
    V1, _, _, _ = _grid_v(scenes.Boundary(scenes.WALLS_NONE, regs))
    inside = [np.all((cell >= np.array(r.lo)) & (cell < np.array(r.hi)), -1) & has for r in regs]
    last = np.full(cell.shape[:2], -1)
    for i, m in enumerate(inside):
        last[m] = i
        assert m.sum() > 50
    np.testing.assert_array_equal(V1[(last < 0)], V0[(last < 0)])
    assert np.abs(V1[last == 0]).max() == 0.0
    np.testing.assert_allclose(V1[last == 2], np.broadcast_to(np.float32([0.1, 0, 0]), V1[last == 2].shape), atol=0)
    r = regs[1]
    # node positions with the FP32 dx the engine and the oracle use
    q = cell[last == 1] * np.float64(np.float32(sc.dx)) - np.float32(r.center)
    w = np.float32(r.value).astype(np.float64)
    vr = V1[last == 1]
    assert np.abs(vr @ w).max() < 1e-10 and np.abs((vr * q).sum(-1)).max() < 1e-10
    rperp = np.linalg.norm(q - np.outer(q @ w, w) / (w @ w), axis=-1)
    np.testing.assert_allclose(np.linalg.norm(vr, axis=-1), np.linalg.norm(w) * rperp, rtol=1e-10, atol=1e-12)
    # walls after the regions: with slip walls the fixed region stays zero on the wall
    V2, _, _, _ = _grid_v(scenes.Boundary(scenes.WALLS_SLIP, regs))
    assert np.abs(V2[last == 0]).max() == 0.0


# ------------------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_boundary_lockstep():
    from tests.test_gpu_parity import _lockstep

    sc = _corner_scene(None)
    sc.boundary = scenes.Boundary(scenes.WALLS_SLIP, _regions(sc.dx))
    print("boundary lockstep: corner box, slip walls, 3 regions, 6 substeps", flush=True)
    _lockstep(sc, 6, lambda k: (0.0 if k == 0 else sc.dt, sc.dt), (2e-7, 2e-5))


@pytest.mark.gpu
def test_gpu_adjoint_refuses_custom_boundaries():
    from ckmpm_amd.engine import engine_from_scene
    from ckmpm_amd.native import CkmpmError

    sc = scenes.cantilever(seed=6, scale=4)
    eng = engine_from_scene(sc, adjoint_segment=4)
    with pytest.raises(CkmpmError, match="default boundary"):
        eng.set_boundary(scenes.Boundary(scenes.WALLS_SLIP))
    eng.set_boundary(None)  # the default stays allowed
    eng.close()


@pytest.mark.gpu
def test_gpu_invalid_descriptions_are_rejected():
    import ctypes

    from ckmpm_amd import native
    from ckmpm_amd.engine import engine_from_scene

    sc = _corner_scene(None)
    eng = engine_from_scene(sc)
    lib = native.load()
    bad = []
    d = native.BoundaryDesc()
    d.walls = 7
    bad.append(d)
    d = native.BoundaryDesc()
    d.nregions = native.MAX_REGIONS + 1
    bad.append(d)
    d = native.BoundaryDesc()
    d.nregions = 1
    d.regions[0].kind = 9
    bad.append(d)
    for b in bad:
        assert lib.ckmpm_set_boundary(eng._h, ctypes.byref(b)) == native.ERR_INVALID
    # the engine still runs with its default boundary
    eng.Substep(0.0, sc.dt)
    eng.Synchronize()
    eng.close()
    # an unknown material kind is refused at add_model
    sc.models[0].material.kind = 6
    with pytest.raises(native.CkmpmError, match="material kind"):
        engine_from_scene(sc)
