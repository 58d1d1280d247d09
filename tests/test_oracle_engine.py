"""Pins the FP64 oracle's full substep with invariants the mathematics fixes.

* mass and momentum: P2G deposits exactly 2 sum m_p on the two grids and, without
  gravity or boundary contact, conserves 2 sum m_p v_p (the two grids' first moments
  cancel, so the affine and stress terms carry no net momentum);
* uniform translation: C = 0, x += dt u exactly, F stays I (a dropped 1/2 on the
  two-grid average or a wrong node offset breaks it);
* free fall: v_k = k g dt, x_k = x_0 + g dt^2 k (k + 1) / 2 (symplectic Euler);
* mirror symmetry of the colliding spheres about x = 1/2 (a swapped sub-grid or a
  transposed matrix breaks it);
* stress direction: a compressed block pushes its grid momentum outward (a sign error
  in the stress term pulls it inward);
* rebin and setup: independent numpy restatements of the canonical orderings.
"""
import dataclasses

import numpy as np
import pytest

from ckmpm_amd import scenes
from oracle.pyoracle import Oracle

SLOT = (1 << 27) - 1


def _box(kind, lo_cell, n_cells, dx, v, gravity=0.0, dim=16, seed=4):
    rng = scenes.Splitmix(seed)
    g = np.stack(np.meshgrid(*[np.arange(lo_cell[a], lo_cell[a] + n_cells) for a in range(3)], indexing="ij"), -1)
    x = scenes._stratified_cells(rng, g.reshape(-1, 3).astype(np.float64), dx).astype(np.float32)
    lam, mu = scenes.lame(1e4, 0.3)
    mat = scenes.Material(kind, lam, mu, bulk=1e4, gamma=7.0, viscosity=0.0)
    vol = dx ** 3 / 8
    m = scenes.ModelSpec(mat, 1000 * vol, vol, x, np.array([v], np.float32))
    return scenes.Scene("box", (dim, dim, dim), (0, 0, 0), dx, gravity, 1e-4, 10, [m])


def _totals(o):
    M = o.grid_m()
    return M[:, 0].sum(), M[:, 1:].sum(axis=(0, 2))


def test_mass_and_momentum_conservation():
    sc = scenes.colliding_spheres(speed=0.7)
    o = Oracle.from_scene(sc)
    o.setup()
    m_tot = sum(m.mass * m.x.shape[0] for m in sc.models)
    p_tot = sum(m.mass * m.x.shape[0] * m.v[0].astype(np.float64) for m in sc.models)
    mass, mom = _totals(o)
    assert abs(mass - 2 * m_tot) < 1e-12 * m_tot
    np.testing.assert_allclose(mom, 2 * p_tot, atol=1e-12 * m_tot)
    for k in range(6):
        o.substep(0.0 if k == 0 else sc.dt, sc.dt)
        mass, mom = _totals(o)
        assert abs(mass - 2 * m_tot) < 1e-12 * m_tot
        np.testing.assert_allclose(mom, 2 * p_tot, atol=1e-11 * m_tot)
    assert o.flags == 0


@pytest.mark.parametrize("kind", [scenes.FIXED_COROTATED, scenes.NEO_HOOKEAN, scenes.DRUCKER_PRAGER, scenes.FLUID])
def test_uniform_translation(kind):
    u = np.array([0.3, -0.2, 0.1])
    sc = _box(kind, (24, 24, 24), 6, 1 / 64, u)
    o = Oracle.from_scene(sc)
    o.setup()
    x0 = o.state(0)[:3].copy()
    steps = 8
    for k in range(steps):
        o.substep(0.0 if k == 0 else sc.dt, sc.dt)
    # state is in processing order; follow identities through the orderings
    st = o.state(0)
    ids = np.arange(x0.shape[1])
    o2 = Oracle.from_scene(sc)
    o2.setup()
    for k in range(steps):
        ids = ids[o2.perm(0) & SLOT]
        o2.substep(0.0 if k == 0 else sc.dt, sc.dt)
    disp = st[:3] - x0[:, ids]
    np.testing.assert_allclose(disp, (steps - 1) * sc.dt * u[:, None] * np.ones_like(disp), atol=1e-13)
    if kind == scenes.FLUID:
        np.testing.assert_allclose(st[3], 1.0, atol=1e-12)
    else:
        np.testing.assert_allclose(st[3:12], np.eye(3).reshape(9, 1) * np.ones((9, st.shape[1])), atol=1e-12)


def test_free_fall_is_symplectic_euler():
    g = -9.8
    sc = _box(scenes.NEO_HOOKEAN, (24, 30, 24), 6, 1 / 64, (0.0, 0.0, 0.0), gravity=g)
    o = Oracle.from_scene(sc)
    o.setup()
    y0 = o.state(0)[1].copy()
    ids = np.arange(y0.shape[0])
    K = 12
    for k in range(K + 1):   # one priming substep (dt = 0), then K substeps
        ids = ids[o.perm(0) & SLOT]
        o.substep(0.0 if k == 0 else sc.dt, sc.dt)
    y = o.state(0)[1]
    np.testing.assert_allclose(y - y0[ids], g * sc.dt ** 2 * K * (K + 1) / 2, rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(o.max_v2, (g * sc.dt * K) ** 2, rtol=1e-10)


def test_colliding_spheres_mirror_symmetry():
    sc = scenes.colliding_spheres(speed=1.0)
    # make sphere 2 the exact mirror image of sphere 1 about x = 1/2
    m0 = sc.models[0]
    m0.x = (np.round(m0.x.astype(np.float64) * 2 ** 23) / 2 ** 23).astype(np.float32)  # 1 - x exact
    x1 = m0.x.astype(np.float64).copy()
    x1[:, 0] = 1.0 - x1[:, 0]
    sc.models[1].x = x1.astype(np.float32)
    assert np.array_equal(sc.models[1].x[:, 0].astype(np.float64), 1.0 - m0.x[:, 0].astype(np.float64))
    o = Oracle.from_scene(sc)
    o.setup()
    orders = [o.setup_order(m).astype(np.int64) for m in range(2)]
    for k in range(20):
        for m in range(2):
            orders[m] = orders[m][o.perm(m) & SLOT]
        o.substep(0.0 if k == 0 else sc.dt, sc.dt)
    xs = []
    for m in range(2):
        st = o.state(m)
        x = np.empty((3, st.shape[1]))
        x[:, orders[m]] = st[:3]
        xs.append(x)
    np.testing.assert_allclose(xs[1][0], 1.0 - xs[0][0], atol=1e-12)
    np.testing.assert_allclose(xs[1][1:], xs[0][1:], atol=1e-12)
    _, mom = _totals(o)
    assert abs(mom[0]) < 1e-12


def test_compressed_block_pushes_outward():
    sc = _box(scenes.FIXED_COROTATED, (24, 24, 24), 8, 1 / 64, (0.0, 0.0, 0.0))
    o = Oracle.from_scene(sc)
    o.setup()
    st = o.state(0).copy()
    st[3:12] = (0.98 * np.eye(3)).reshape(9, 1)
    o.set_state(0, st)
    o.physics(0.0, sc.dt)
    M = o.grid_mnext()
    keys = o.keys(0)
    dim = sc.dim
    # node positions of the grid-0 / grid-1 nodes of each block
    radial = 0.0
    centre = st[:3].mean(axis=1)
    for bi in range(M.shape[0]):
        k = int(keys[bi])
        b = np.array([k // (dim[1] * dim[2]), (k // dim[2]) % dim[1], k % dim[2]])
        for n in range(128):
            g, l = divmod(n, 64)
            c = 4 * b + np.array([l // 16, (l // 4) % 4, l % 4])
            xn = (c + (0.25 if g else -0.25)) * sc.dx
            radial += M[bi, 1:, n] @ (xn - centre)
    assert radial > 0.0


def _expected_rebin(o, nmodels, dim, dx):
    """Canonical rebin restated in numpy from the oracle's pre-rebin state."""
    keys = o.keys(0)
    A, Nb, Ex = o.counts(0)
    index = {int(k): i for i, k in enumerate(keys)}
    coords = lambda k: np.array([k // (dim[1] * dim[2]), (k // dim[2]) % dim[1], k % dim[2]])
    targets, moves = [], []
    occupied = np.zeros(Ex, bool)
    for m in range(nmodels):
        st = o.state_next(m)
        off = o.offsets(m)
        tg = np.zeros(st.shape[1], np.int64)
        mv = np.zeros(st.shape[1], np.int64)
        for j in range(A):
            b = coords(int(keys[j]))
            for t in range(off[j], off[j + 1]):
                u = (st[:3, t] - 4 * b * dx).astype(np.float32)
                cell = np.floor((u * np.float32(1 / dx)).astype(np.float32) + np.float32(0.25)).astype(np.int64)
                d = np.clip(np.floor_divide(cell - 2, 4), -1, 1)
                nb = b + d
                tg[t] = index[int((nb[0] * dim[1] + nb[1]) * dim[2] + nb[2])]
                mv[t] = (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1)
        occupied[tg] = True
        targets.append(tg)
        moves.append(mv)
    newidx = np.cumsum(occupied) - 1
    out = []
    for m in range(nmodels):
        ni = newidx[targets[m]]
        order = np.argsort(ni.astype(np.int64) * 27 + moves[m], kind="stable")
        perm = order | (moves[m][order] << 27)
        counts = np.bincount(ni, minlength=int(occupied.sum()))
        off = np.concatenate([[0], np.cumsum(counts)])
        out.append((perm.astype(np.uint32), off.astype(np.uint32)))
    return keys[occupied], out


def test_rebin_matches_numpy_restatement():
    sc = scenes.colliding_spheres(speed=3.0)   # fast, so particles cross blocks
    o = Oracle.from_scene(sc)
    o.setup()
    crossed = 0
    for k in range(12):
        o.physics(0.0 if k == 0 else 4 * sc.dt, sc.dt)
        akeys, exp = _expected_rebin(o, 2, sc.dim, sc.dx)
        o.rebin()
        A = o.counts(0)[0]
        np.testing.assert_array_equal(o.keys(0)[:A], akeys)
        for m in range(2):
            perm, off = o.perm(m), o.offsets(m)
            np.testing.assert_array_equal(perm, exp[m][0])
            np.testing.assert_array_equal(off, exp[m][1])
            assert np.array_equal(np.sort(perm & SLOT), np.arange(perm.shape[0]))
            crossed += int(((perm >> 27) != 13).sum())
    assert crossed > 100
    assert o.flags == 0


def _wide_spheres():
    """The colliding spheres 600 blocks along x of a 1024 x 512 x 512-block domain (2^28
    blocks, keys above 1.5e8: beyond a 32-bit block << 6 | cell key)."""
    import copy

    sc = scenes.colliding_spheres()
    big = copy.copy(sc)
    big.dim = (1024, 512, 512)
    big.models = []
    for m in sc.models:
        mm = copy.copy(m)
        mm.x = (m.x + np.array([600 * 4 * sc.dx, 0.0, 0.0], np.float32)).astype(np.float32)
        big.models.append(mm)
    return big


@pytest.mark.parametrize("wide", [False, True])
def test_setup_order_matches_numpy_restatement(wide):
    """Setup order = sort by (block, rank of the particle among the earlier input
    particles of its binning cell, cell within the block)."""
    sc = _wide_spheres() if wide else scenes.mixed_dam_sand(n_fluid=9000, n_sand=5000, scale=8)
    o = Oracle.from_scene(sc)
    o.setup()
    dim = sc.dim
    inv_dx = np.float32(1.0 / sc.dx)
    for mi, m in enumerate(sc.models):
        x = m.x.astype(np.float32)
        cell = np.floor((x * inv_dx).astype(np.float32) + np.float32(0.25)).astype(np.int64)
        b = np.floor_divide(cell - 2, 4)
        assert (b >= 0).all() and (b < np.array(dim)).all()
        key = (b[:, 0] * dim[1] + b[:, 1]) * dim[2] + b[:, 2]
        cb = ((cell[:, 0] - 2) % 4) * 16 + ((cell[:, 1] - 2) % 4) * 4 + (cell[:, 2] - 2) % 4
        ck = key * 64 + cb
        # rank among earlier particles with the same cell key
        o1 = np.argsort(ck, kind="stable")
        sk = ck[o1]
        start = np.r_[0, np.flatnonzero(np.diff(sk)) + 1]
        run = np.repeat(start, np.diff(np.r_[start, sk.size]))
        rank = np.empty_like(ck)
        rank[o1] = np.arange(sk.size) - run
        expected = np.lexsort((cb, rank, key))
        got = o.setup_order(mi)
        np.testing.assert_array_equal(got, expected)
        if wide:
            assert key.min() > (1 << 26)
        # consecutive particles of a block sit in distinct cells while the block has
        # unused cells left for that rank
        assert (np.diff(ck[got]) != 0).mean() > 0.95


def test_oracle_frame_loop_dt_rules():
    """The oracle's frame loop (mpm_engine.cuh:514-876): the first substep moves nothing
    (dt = 0); the time advances by each substep's next_dt, which is at most the default dt
    and ends the frame exactly; at 0.1 m/s the CFL bound dx cfl / max|v| is far and the
    steps are the default dt, at 10 m/s it sets them."""
    from ckmpm_amd import scenes

    fps, cfl = 240, 0.5
    for speed, cfl_active in ((0.1, False), (10.0, True)):
        sc = dataclasses.replace(scenes.colliding_spheres(radius=0.06, speed=speed), dt=1e-3)
        o = Oracle.from_scene(sc)
        o.setup()
        state = None
        for frame in range(2):
            state, dts = o.run_frames(1, fps, cfl, sc.dt, sc.dx, state)
            if frame == 0:
                assert dts[0] == 0.0
            steps = dts[1:] + [state[1]]   # the time increments of this frame
            assert abs(sum(steps) - 1.0 / fps) < 1e-12
            assert abs(state[0] - (frame + 1) / fps) < 1e-12
            assert all(0.0 < d <= sc.dt for d in steps)
            if cfl_active:
                vmax = speed * 1.05
                assert all(d < 0.9 * sc.dt for d in steps) and min(steps[:-1]) > 0.5 * sc.dx * cfl / vmax
            else:
                assert steps[:-1] == [sc.dt] * (len(steps) - 1)


def _rigid_box(rigid_count, velocity=(0.4, 0.0, 0.0)):
    """A von Mises box (12^3 blocks, 2^3 particles per cell) at rest under gravity whose
    first rigid_count particles are rigid."""
    from ckmpm_amd import scenes

    dim = 12
    dx = 1.0 / (4 * dim)
    rng = scenes.Splitmix(44)
    g = np.stack(np.meshgrid(*[np.arange(16, 30)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = scenes._stratified_cells(rng, g.astype(np.float64), dx, 2).astype(np.float32)
    lam, mu = scenes.lame(2e5, 0.3)
    vol = dx ** 3 / 8
    m = scenes.ModelSpec(scenes.Material(scenes.VON_MISES, lam, mu, yield_stress=30.0), 1000 * vol, vol, x,
                         np.zeros((1, 3), np.float32), rigid_count=rigid_count, rigid_velocity=velocity)
    return scenes.Scene("rigid_box", (dim, dim, dim), (0, 0, 0), dx, -9.8, 2e-5, 10, [m])


def test_oracle_rigid_particles_move_at_their_velocity():
    """Rigid particles (mpm_algorithm.cuh:1609-1624): they take no velocity from the grid
    and move by dt * velocity exactly, their F stays I; the other particles fall under
    gravity; without rigid particles every particle falls."""
    from ckmpm_amd.native import SLOT_MASK

    sc = _rigid_box(600)
    o = Oracle.from_scene(sc)
    o.setup()
    ids = o.setup_order(0).astype(np.int64)
    steps = 6
    for k in range(steps + 1):
        ids = ids[o.perm(0) & SLOT_MASK]
        o.substep(0.0 if k == 0 else sc.dt, sc.dt)
    st = o.state(0)
    x = np.empty((3, st.shape[1]))
    x[:, ids] = st[:3]
    F = np.empty((9, st.shape[1]))
    F[:, ids] = st[3:12]
    flag = np.empty(st.shape[1])
    flag[ids] = st[12]          # von Mises state: x, F, then the rigid flag
    x0 = sc.models[0].x.T.astype(np.float64)
    rigid = np.arange(st.shape[1]) < 600
    assert (flag[rigid] == 1).all() and (flag[~rigid] == 0).all()
    np.testing.assert_allclose(x[:, rigid] - x0[:, rigid],
                               np.outer([0.4, 0.0, 0.0], np.ones(rigid.sum())) * steps * sc.dt, rtol=0, atol=1e-12)
    np.testing.assert_allclose(F[:, rigid], np.tile(np.eye(3).reshape(9, 1), (1, rigid.sum())), rtol=0, atol=0)
    assert (x[1, ~rigid] - x0[1, ~rigid]).mean() < -0.4 * 9.8 * (steps * sc.dt) ** 2
    # the flag is the state's last attribute
    assert st.shape[0] == 13
