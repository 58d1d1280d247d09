"""Batched scenes (ckmpm_batch_*).

The graph replays (one graph over all engine streams, one per engine, or fused launches
whose workgroups are split by scene) must give each engine the same result as stepping
it alone with ckmpm_substep: same block lists and orderings (integer, bit-exact), same
particle state up to the order of the floating-point atomics (the only nondeterminism).
"""
import numpy as np
import pytest

from ckmpm_amd import scenes


def _small_batch():
    # four scenes, one per material, at 64^3 and 128^3
    out = []
    for i, sc in enumerate(scenes.batch_scenes(count=24, seed=7)):
        if sc.models[0].material.kind == len(out) % 4 and sc.n_particles < 120_000:
            out.append(sc)
        if len(out) == 4:
            break
    assert len(out) == 4
    return out


def test_batch_scenes_are_seeded_and_within_spec():
    a = scenes.batch_scenes(count=8, seed=4)
    b = scenes.batch_scenes(count=8, seed=4)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.models[0].x, y.models[0].x)
    for sc in scenes.batch_scenes():
        g = sc.dim[0] * 4
        assert g in (64, 128, 256)
        x = sc.models[0].x
        assert 15_000 <= x.shape[0] <= 510_000
        cell = np.floor(x / sc.dx)
        assert cell.min() >= 8 and cell.max() < g - 8   # clear of the sticky band


def _graphs_captured(runs, fused):
    """Graphs a batch captures over runs of the given substep counts: the first substep of
    a batch runs eagerly (dt = 0), then one graph of two substeps per buffer parity, and in
    fused mode one graph of a single substep per parity for odd remainders (the other
    modes run those eagerly)."""
    done, graphs = 0, set()
    for n in runs:
        if done == 0 and n:
            done, n = 1, n - 1
        if n // 2:
            graphs.add(("pair", done & 1))
            done += 2 * (n // 2)
        if n & 1:
            if fused:
                graphs.add(("single", done & 1))
            done += 1
    return len(graphs)


@pytest.mark.gpu
@pytest.mark.parametrize("substeps", [7, 10])
@pytest.mark.parametrize("mode", [2, 3, 4])
def test_graph_batch_matches_separate_engines(substeps, mode):
    from ckmpm_amd.engine import MPMBatch, engine_from_scene

    scs = _small_batch()
    alone = [engine_from_scene(sc) for sc in scs]
    for e, sc in zip(alone, scs):
        for k in range(substeps):
            e.Substep(0.0 if k == 0 else sc.dt, sc.dt)
        e.Synchronize()
    together = [engine_from_scene(sc) for sc in scs]
    batch = MPMBatch(together)
    batch.Run(substeps, mode)
    batch.Synchronize()
    per_capture = len(scs) if mode == MPMBatch.ENGINE_GRAPHS else 1
    fused = mode == MPMBatch.FUSED
    assert batch.captures == per_capture * _graphs_captured([substeps], fused)
    for a, b, sc in zip(alone, together, scs):
        ka, ca = a.layout(0)
        kb, cb = b.layout(0)
        assert ca == cb
        np.testing.assert_array_equal(ka, kb)
        pa, oa = a.ordering(0)
        pb, ob = b.ordering(0)
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(oa, ob)
        np.testing.assert_allclose(b.state(0), a.state(0), rtol=0, atol=1e-5)
    # later runs reuse the graphs of their substep parity (one per parity and kind at most)
    batch.Run(4, mode)
    batch.Run(5, mode)
    batch.Run(4, mode)
    batch.Synchronize()
    assert batch.captures == per_capture * _graphs_captured([substeps, 4, 5, 4], fused)
    batch.close()


@pytest.mark.gpu
def test_sequential_and_stream_modes_match_graph():
    from ckmpm_amd.engine import MPMBatch, engine_from_scene

    scs = _small_batch()
    res = {}
    for mode in (MPMBatch.SEQUENTIAL, MPMBatch.STREAMS, MPMBatch.GRAPH, MPMBatch.ENGINE_GRAPHS, MPMBatch.FUSED):
        engs = [engine_from_scene(sc) for sc in scs]
        b = MPMBatch(engs)
        b.Run(6, mode)
        b.Synchronize()
        res[mode] = [e.state(0) for e in engs]
        b.close()
    for m in (MPMBatch.STREAMS, MPMBatch.GRAPH, MPMBatch.ENGINE_GRAPHS, MPMBatch.FUSED):
        for x, y in zip(res[MPMBatch.SEQUENTIAL], res[m]):
            np.testing.assert_allclose(y, x, rtol=0, atol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("serial", [False, True])
def test_fused_batch_variants_match_separate_engines(serial):
    """The fused launches' variants: a custom boundary (second grid-update instantiation),
    two models of one engine in one G2P2G launch (two fixed-corotated spheres, with the
    corner box's model: three entries), two kinds of one engine (fluid + sand) and
    512-thread G2P2G workgroups (the fluid)."""
    from ckmpm_amd.engine import MPMBatch, engine_from_scene
    from tests.test_boundary import _corner_scene, _regions

    def scenes_():
        corner = _corner_scene(None)
        corner.boundary = scenes.Boundary(scenes.WALLS_SLIP, _regions(corner.dx))
        return [corner, scenes.colliding_spheres(), scenes.mixed_dam_sand(n_fluid=9_000, n_sand=5_000, scale=8)]

    def engines():
        engs = [engine_from_scene(sc) for sc in scenes_()]
        engs[2].set_tuning(0, 512)
        return engs

    substeps = 9
    print(f"fused variants: corner (custom boundary), spheres (2 FC models), dam+sand; {substeps} substeps", flush=True)
    alone = engines()
    for e, sc in zip(alone, scenes_()):
        for k in range(substeps):
            e.Substep(0.0 if k == 0 else sc.dt, sc.dt)
        e.Synchronize()
    together = engines()
    batch = MPMBatch(together)
    batch.serial_g2p2g = serial
    batch.Run(substeps, MPMBatch.FUSED)
    batch.Synchronize()
    assert batch.captures == _graphs_captured([substeps], True)
    for a, b in zip(alone, together):
        ka, ca = a.layout(0)
        kb, cb = b.layout(0)
        assert ca == cb
        np.testing.assert_array_equal(ka, kb)
        for m in range(len(a.models)):
            pa, oa = a.ordering(m)
            pb, ob = b.ordering(m)
            np.testing.assert_array_equal(pa, pb)
            np.testing.assert_array_equal(oa, ob)
            np.testing.assert_allclose(b.state(m), a.state(m), rtol=0, atol=1e-5)
        assert b.counters().flags == 0
    batch.close()


@pytest.mark.gpu
def test_batch_rejects_an_engine_twice():
    from ckmpm_amd.engine import MPMBatch, engine_from_scene
    from ckmpm_amd.native import CkmpmError

    e = engine_from_scene(scenes.colliding_spheres())
    with pytest.raises(CkmpmError, match="ckmpm_batch_create"):
        MPMBatch([e, e])
    e.close()


@pytest.mark.gpu
def test_fused_batch_of_64_scenes_matches_separate_engines_and_the_oracle():
    """The 64-scene workload (20k-500k particles per scene, every material, grids
    64^3-256^3) in fused mode: 6 substeps (the first eager, two fused pairs, one fused
    single substep) equal separate engines (block lists and orderings bit-exact, state
    within 1e-5); then the smallest scene of each material is imported into the oracle
    before each of two more fused single substeps (one per buffer parity) and compared
    with it in lockstep: grid within 1e-5 per substep, state, and the rebin of the GPU's
    FP32 positions bit-exact."""
    from ckmpm_amd.engine import MPMBatch, engine_from_scene
    from oracle.pyoracle import Oracle
    from tests.test_gpu_parity import _compare_layout, _compare_ordering, _grid_close, _import

    scs = scenes.batch_scenes()
    assert len(scs) == 64
    together = [engine_from_scene(sc) for sc in scs]
    batch = MPMBatch(together)
    substeps = 6
    print(f"64 scenes, {batch.n_particles} particles: {substeps} fused substeps", flush=True)
    batch.Run(substeps, MPMBatch.FUSED)
    batch.Synchronize()
    assert batch.captures == _graphs_captured([substeps], True) == 2   # a pair graph, a single-substep graph
    for i, (b, sc) in enumerate(zip(together, scs)):
        a = engine_from_scene(sc)
        for k in range(substeps):
            a.Substep(0.0 if k == 0 else sc.dt, sc.dt)
        a.Synchronize()
        ka, ca = a.layout(0)
        kb, cb = b.layout(0)
        assert ca == cb, i
        np.testing.assert_array_equal(ka, kb, err_msg=f"scene {i}")
        pa, oa = a.ordering(0)
        pb, ob = b.ordering(0)
        np.testing.assert_array_equal(pa, pb, err_msg=f"scene {i}")
        np.testing.assert_array_equal(oa, ob, err_msg=f"scene {i}")
        np.testing.assert_allclose(b.state(0), a.state(0), rtol=0, atol=1e-5, err_msg=f"scene {i}")
        assert b.counters().flags == 0, i
        a.close()
    print("separate engines agree", flush=True)
    picks = {}
    for i, sc in enumerate(scs):
        k = sc.models[0].material.kind
        if k not in picks or sc.n_particles < scs[picks[k]].n_particles:
            picks[k] = i
    assert sorted(picks) == [scenes.FIXED_COROTATED, scenes.NEO_HOOKEAN, scenes.DRUCKER_PRAGER, scenes.FLUID]
    orcs = {}
    for i in picks.values():
        orcs[i] = Oracle.from_scene(scs[i])
        orcs[i].setup()
    for step in range(2):
        for i, o in orcs.items():
            _import(together[i], o, 1)
            o.physics(scs[i].dt, scs[i].dt)
        batch.Run(1, MPMBatch.FUSED)
        batch.Synchronize()
        for i, o in orcs.items():
            e = together[i]
            print(f"oracle lockstep, fused single substep {step}: scene {i} "
                  f"(kind {scs[i].models[0].material.kind}, {scs[i].n_particles} particles)", flush=True)
            _grid_close(e.grid_m(), o.grid_mnext())
            sg, so = e.global_state(0), o.state_next(0)
            np.testing.assert_allclose(sg[:3], so[:3], rtol=0, atol=2e-7)
            if sg.shape[0] > 3:
                np.testing.assert_allclose(sg[3:], so[3:], rtol=0, atol=2e-5)
            o.rebin([e.state(0)[:3].T.copy()])
            _compare_layout(e, o, 0)
            _compare_layout(e, o, 1)
            _compare_ordering(e, o, 1)
            assert o.flags == 0
    assert batch.captures == _graphs_captured([substeps, 1, 1], True) == 3
    batch.close()


@pytest.mark.gpu
def test_single_engine_fused_graphs_match_plain_launches_at_8m():
    """The benchmark's timed path: the 8M-particle mixed scene advanced by one-engine
    batches of fused launches replayed as hipGraphs, against plain launches."""
    from ckmpm_amd.engine import MPMBatch, engine_from_scene

    sc = scenes.mixed_dam_sand()
    substeps = 7
    print(f"8M mixed scene, {substeps} substeps: plain launches vs fused hipGraphs", flush=True)
    a = engine_from_scene(sc)
    for k in range(substeps):
        a.Substep(0.0 if k == 0 else sc.dt, sc.dt)
    a.Synchronize()
    b = engine_from_scene(sc)
    b.Substep(0.0, sc.dt)
    batch = MPMBatch([b])
    batch.Run(substeps - 1, MPMBatch.FUSED)
    batch.Synchronize()
    ka, ca = a.layout(0)
    kb, cb = b.layout(0)
    assert ca == cb
    np.testing.assert_array_equal(ka, kb)
    for m in range(len(sc.models)):
        pa, oa = a.ordering(m)
        pb, ob = b.ordering(m)
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(oa, ob)
        np.testing.assert_allclose(b.state(m), a.state(m), rtol=0, atol=1e-5)
    assert a.counters().flags == 0 and b.counters().flags == 0
    batch.close()
    a.close()
    b.close()


@pytest.mark.gpu
def test_batch_autotune_runs_every_candidate_and_keeps_one():
    """The fused mode's tuner: each kind's G2P2G workgroup size, then side-by-side or
    serial G2P2G launches, each candidate timed over replays of the fused graph (the
    configurations it chooses among are checked against separate engines in
    test_fused_batch_variants_match_separate_engines)."""
    from ckmpm_amd.engine import MPMBatch, engine_from_scene

    engs = [engine_from_scene(sc) for sc in (scenes.colliding_spheres(),
                                              scenes.mixed_dam_sand(n_fluid=9_000, n_sand=5_000, scale=8))]
    batch = MPMBatch(engs)
    batch.Run(1, MPMBatch.FUSED)
    print("batch autotune: spheres (FC) + dam (fluid, sand), 1 timed pair per candidate", flush=True)
    batch.autotune(1)
    batch.Run(3, MPMBatch.FUSED)
    batch.Synchronize()
    # 3 kinds x 2 sizes + 2 launch orders, 2 (1 + 1) substeps each
    assert all(e.counters().substeps == 1 + 4 * (2 * 3 + 2) + 3 for e in engs)
    for e in engs:
        assert e.counters().flags == 0
        for m in range(len(e.models)):
            assert e.tuning(m) in (256, 512)
            assert np.isfinite(e.state(m)).all()
    # one size per kind: the two fixed-corotated spheres share theirs
    assert engs[0].tuning(0) == engs[0].tuning(1)
    assert isinstance(batch.serial_g2p2g, bool)
    batch.close()
