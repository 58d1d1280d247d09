"""Pins the FP64 oracle's unit functions against mathematics that fixes them.

The reference ships no golden vectors for this path, so each check is chosen to fail on
a plausible mistake: an off-by-one in the +-0.25 cell shift, a swapped sub-grid, a
missing dx factor, a transposed SVD factor, a wrong stress sign.
"""
import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg

from oracle import pyoracle
from oracle.pyoracle import Oracle

FC, NH, DP, FL = 0, 1, 2, 3


def _oracle(dim=(16, 16, 16), dx=1.0 / 64):
    return Oracle(dim, dx)


# ------------------------------------------------------------------ binning
def test_binning_exact_at_quarter_boundaries():
    """cell = floor(x/dx + 1/4) exactly at and next to every boundary (dx = 1/64)."""
    o = _oracle()
    xs = []
    for k in range(2, 63):
        b = np.float32((k - 0.25) / 64)             # exactly representable boundary
        xs += [b, np.nextafter(b, np.float32(0)), np.nextafter(b, np.float32(1))]
    x = np.array([[v, v, v] for v in xs], np.float32)
    blk, cell = o.bin_f32(x)
    for i, v in enumerate(xs):
        exact = math.floor(Fraction(float(v)) * 64 + Fraction(1, 4))
        assert cell[i, 0] == exact, (v, cell[i, 0], exact)
        assert blk[i, 0] == (exact - 2) // 4


def test_binning_matches_float32_arithmetic_non_power_of_two():
    """For dx = 1/100 the multiply rounds; the oracle must follow FP32 semantics."""
    o = Oracle((32, 32, 32), 1.0 / 100)
    rng = np.random.default_rng(7)
    x = rng.uniform(0.02, 1.2, size=(20000, 3)).astype(np.float32)
    blk, cell = o.bin_f32(x)
    inv = np.float32(1.0 / np.float64(np.float32(1.0 / 100)))
    ref = np.floor((x * inv).astype(np.float32) + np.float32(0.25)).astype(np.int64)
    np.testing.assert_array_equal(cell, ref)
    np.testing.assert_array_equal(blk, np.floor_divide(ref - 2, 4))


# ------------------------------------------------------------------ weights / stencil
def test_weight_closed_form():
    for t in np.linspace(0, 1, 41)[:-1]:
        assert abs(pyoracle.weight0(t) - (1 - t + math.sin(2 * math.pi * t) / (2 * math.pi))) < 1e-15
    # C1 continuity across a cell: node 0 of cell c at t -> 1 equals node 1 of cell c - 1
    assert abs(pyoracle.weight0(1 - 1e-12)) < 1e-11


@pytest.mark.parametrize("g", [0, 1])
def test_stencil_geometry(g):
    """x/dx = c + t - 1/4 on grid 0 and c + t + 1/4 on grid 1, t in [0, 1), weights sum to 1."""
    o = _oracle()
    dx = 1.0 / 64
    rng = np.random.default_rng(1)
    for x in rng.uniform(0.2, 0.8, size=(200, 3)):
        c, t, w = o.stencil(x, g)
        assert np.all((t >= 0) & (t < 1))
        np.testing.assert_allclose(c + t + (-0.25 if g == 0 else 0.25), x / dx, rtol=0, atol=1e-12)
        np.testing.assert_allclose(w.sum(axis=1), 1.0, atol=1e-15)
        W = np.einsum("i,j,k->ijk", w[0], w[1], w[2])
        assert abs(W.sum() - 1.0) < 1e-14


def test_first_moments_cancel_and_D_closed_form():
    """sum_g sum_i w (x_i - x_p) = 0, and D = dx^2 (diag(sum_g m2) + 2 s s^T off-diagonal)."""
    o = _oracle()
    dx = 1.0 / 64
    rng = np.random.default_rng(2)
    for x in rng.uniform(0.2, 0.8, size=(200, 3)):
        first = np.zeros(3)
        D = np.zeros((3, 3))
        m2 = np.zeros(3)
        for g in (0, 1):
            c, t, w = o.stencil(x, g)
            for d in np.ndindex(2, 2, 2):
                wt = w[0, d[0]] * w[1, d[1]] * w[2, d[2]]
                r = (np.array(d) - t) * dx
                first += wt * r
                D += wt * np.outer(r, r)
            m2 += t * (1 - t) - np.sin(2 * np.pi * t) / (2 * np.pi) * (1 - 2 * t)
        np.testing.assert_allclose(first, 0.0, atol=1e-13 * dx)
        # grid-0 offset t0 gives s = sin(2 pi t0) / (2 pi); grid 1 sits half a cell away
        _, t0, _ = o.stencil(x, 0)
        s = np.sin(2 * np.pi * t0) / (2 * np.pi)
        ref = 2 * np.outer(s, s) * dx * dx
        ref[np.diag_indices(3)] = m2 * dx * dx
        np.testing.assert_allclose(D, ref, rtol=1e-12, atol=1e-20)


# ------------------------------------------------------------------ SVD
def _svd_cases():
    rng = np.random.default_rng(3)
    cases = [np.eye(3), np.diag([2.0, 1.0, 0.5]), np.diag([1.0, 1.0, 1.0]) * 0.7, np.diag([1.0, -1.0, 1.0]),
             np.diag([3.0, 3.0, 1.0]), np.outer([1, 2, 3], [0.5, -1, 2]), np.zeros((3, 3))]
    cases += [np.eye(3) + 0.05 * rng.standard_normal((3, 3)) for _ in range(20)]
    cases += [rng.standard_normal((3, 3)) for _ in range(20)]
    R = scipy.linalg.expm(np.array([[0, -0.3, 0.2], [0.3, 0, -0.1], [-0.2, 0.1, 0]]))
    cases += [R, R @ np.diag([1.2, 0.9, -0.8])]
    return cases


@pytest.mark.parametrize("F", _svd_cases())
def test_signed_svd(F):
    U, s, V = pyoracle.svd(F)
    np.testing.assert_allclose(U @ np.diag(s) @ V.T, F, atol=1e-12 * max(1.0, np.abs(F).max()))
    np.testing.assert_allclose(U.T @ U, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-12)
    assert abs(np.linalg.det(U) - 1) < 1e-12 and abs(np.linalg.det(V) - 1) < 1e-12
    assert s[0] >= s[1] - 1e-14 and s[1] >= abs(s[2]) - 1e-14
    np.testing.assert_allclose(np.sort(np.abs(s))[::-1], np.linalg.svd(F, compute_uv=False), atol=1e-12)
    d = np.linalg.det(F)
    if abs(d) > 1e-9:
        assert np.sign(s[2]) == np.sign(d)


# ------------------------------------------------------------------ stresses
def _energy_fc(F, lam, mu):
    sv = np.linalg.svd(F, compute_uv=False)
    if np.linalg.det(F) < 0:
        sv[-1] = -sv[-1]
    J = np.prod(sv)
    return mu * np.sum((sv - 1) ** 2) + 0.5 * lam * (J - 1) ** 2


def _energy_nh(F, lam, mu):
    J = np.linalg.det(F)
    return 0.5 * mu * (np.trace(F.T @ F) - 3) - mu * np.log(J) + 0.5 * lam * np.log(J) ** 2


def _first_pk_fd(psi, F, h=1e-6):
    P = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            Fp, Fm = F.copy(), F.copy()
            Fp[i, j] += h
            Fm[i, j] -= h
            P[i, j] = (psi(Fp) - psi(Fm)) / (2 * h)
    return P


@pytest.mark.parametrize("kind,energy", [(FC, _energy_fc), (NH, _energy_nh)])
def test_elastic_stress_is_energy_gradient(kind, energy):
    """tau = (d psi / dF) F^T for the fixed-corotated and Neo-Hookean energies."""
    lam, mu = 5769.23, 3846.15
    rng = np.random.default_rng(11)
    for _ in range(12):
        F = np.eye(3) + 0.15 * rng.standard_normal((3, 3))
        if np.linalg.det(F) <= 0.2:
            continue
        _, _, tau = pyoracle.stress(kind, [lam, mu, 0, 0, 0, 1, 30], F)
        P = _first_pk_fd(lambda G: energy(G, lam, mu), F)
        np.testing.assert_allclose(tau, P @ F.T, rtol=1e-6, atol=1e-6 * mu)


def test_fixed_corotated_polar_form():
    lam, mu = 1000.0, 700.0
    rng = np.random.default_rng(12)
    for _ in range(10):
        F = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        if np.linalg.det(F) <= 0.1:
            continue
        R, _ = scipy.linalg.polar(F)
        J = np.linalg.det(F)
        ref = 2 * mu * (F - R) @ F.T + lam * J * (J - 1) * np.eye(3)
        _, _, tau = pyoracle.stress(FC, [lam, mu, 0, 0, 0, 1, 30], F)
        np.testing.assert_allclose(tau, ref, atol=1e-9 * mu)


def _rot(seed):
    a = np.random.default_rng(seed).standard_normal(3) * 0.4
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return scipy.linalg.expm(K)


def test_drucker_prager_branches():
    lam, mu = 201923.0, 134615.0
    params = [lam, mu, 0, 0, 0, 1, 30]
    alpha = math.sqrt(2 / 3) * 2 * 0.5 / (3 - 0.5)
    U, V = _rot(1), _rot(2)
    # elastic: mild compression, small deviator -> Hencky stress, F unchanged
    sig = np.array([0.995, 0.99, 0.992])
    F = U @ np.diag(sig) @ V.T
    Fo, lj, tau = pyoracle.stress(DP, params, F)
    eps = np.log(sig)
    ref = U @ np.diag(2 * mu * eps + lam * eps.sum()) @ U.T
    np.testing.assert_allclose(Fo, F, atol=1e-12)
    np.testing.assert_allclose(tau, ref, rtol=1e-9, atol=1e-6)
    assert lj == 0.0
    # plastic: strong shear under compression -> lands on the cone
    sig = np.array([1.05, 0.9, 0.96])
    F = U @ np.diag(sig) @ V.T
    Fo, lj, tau = pyoracle.stress(DP, params, F)
    H = np.log(np.linalg.svd(Fo, compute_uv=False))
    dev = H - H.mean()
    assert abs(np.linalg.norm(dev) + (3 * lam + 2 * mu) / (2 * mu) * H.sum() * alpha) < 1e-10
    assert abs(H.sum() - np.log(sig).sum()) < 1e-12   # volume preserved by the projection
    # tension: F -> U V^T, no stress, logJp accumulates the volumetric strain
    sig = np.array([1.02, 1.01, 1.03])
    F = U @ np.diag(sig) @ V.T
    Fo, lj, tau = pyoracle.stress(DP, params, F, logJp=0.0)
    np.testing.assert_allclose(Fo, U @ V.T, atol=1e-12)
    np.testing.assert_allclose(tau, 0.0, atol=1e-9)
    assert abs(lj - np.log(sig).sum()) < 1e-12


def test_drucker_prager_elastic_branch_clamps_and_uninverts():
    """Elastic branch (deltaGamma <= 0) of mpm_material.cuh:417-443: F is rebuilt as
    U diag(exp eps) V^T with eps = log max(|s|, 1e-4) (:385).  Built here from the
    rotations the test constructs F with: an inverted F = R1 diag(s0, s1, -s2) R2^T comes
    back as R1 diag(s0, s1, s2) R2^T, and a singular value below 1e-4 comes back as 1e-4.
    The stress is R1 diag(2 mu eps + lambda tr eps) R1^T on the rebuilt F.  With mu = 0
    no branch runs (:412) and F is returned unchanged."""
    lam, mu = 201923.0, 134615.0
    params = [lam, mu, 0, 0, 0, 1, 30]
    R1, R2 = _rot(3), _rot(4)
    cases = [(np.array([0.998, 0.995, -0.99]), np.array([0.998, 0.995, 0.99])),
             (np.array([1.0, 0.99, 1e-5]), np.array([1.0, 0.99, 1e-4])),
             (np.array([1.0, 0.99, -3e-5]), np.array([1.0, 0.99, 1e-4]))]
    for s, ns in cases:
        F = R1 @ np.diag(s) @ R2.T
        Fo, lj, tau = pyoracle.stress(DP, params, F, logJp=-0.5)
        assert lj == 0.0   # compressive branch with mu != 0 resets logJp
        np.testing.assert_allclose(Fo, R1 @ np.diag(ns) @ R2.T, atol=1e-12)
        assert np.linalg.det(Fo) > 0
        eps = np.log(ns)
        np.testing.assert_allclose(tau, R1 @ np.diag(2 * mu * eps + lam * eps.sum()) @ R1.T, rtol=1e-9,
                                   atol=1e-9 * lam * abs(eps).max())
    # the same inverted F is elastic under the sign-blind Hencky measure: the deviator is
    # small, so a projection (which would change the ratios of the singular values) must
    # not happen
    s, ns = cases[0]
    Fo, _, _ = pyoracle.stress(DP, params, R1 @ np.diag(s) @ R2.T)
    sv = np.linalg.svd(Fo, compute_uv=False)
    np.testing.assert_allclose(sv, np.sort(ns)[::-1], atol=1e-12)
    # mu = 0: neither branch of the compressive case runs, F stays inverted
    F = R1 @ np.diag(cases[0][0]) @ R2.T
    Fo, lj, _ = pyoracle.stress(DP, [lam, 0.0, 0, 0, 0, 1, 30], F, logJp=-0.5)
    np.testing.assert_allclose(Fo, F, atol=0)
    assert lj == -0.5


def test_clamped_sand_grid_sensitivity_to_one_ulp():
    """The FP32 conditioning behind the GPU bar of the substep after a clamp
    (test_gpu_parity.py, CLAMPED_NEXT_SUBSTEP_TOL).  In the FP64 oracle alone: sand with
    inverted and near-singular particles takes one substep (F rebuilt, smallest singular
    value 1e-4), F is stored in FP32 as E = F - I, and the next substep runs twice, once
    with one ulp added to or taken from the stored E of the particles at the clamp.  The
    grid momentum moves by more than 1e-5 of its largest value (so the usual bar cannot
    hold for any FP32 engine) and by less than the 1e-4 bar."""
    from tests.test_gpu_parity import CLAMPED_NEXT_SUBSTEP_TOL, _clamped_sand_scene

    sc, picks, Fs = _clamped_sand_scene()
    orc = Oracle.from_scene(sc)
    orc.setup()
    st = orc.state(0)
    for p, F in zip(picks, Fs):
        E = (F - np.eye(3)).astype(np.float32).astype(np.float64)
        st[3:12, p] = (E + np.eye(3)).reshape(9)
    orc.set_state(0, st)
    orc.substep(0.0, sc.dt)
    st1 = orc.state(0)
    M1 = orc.grid_m().astype(np.float32)
    E32 = (st1[3:12] - np.eye(3).reshape(9, 1)).astype(np.float32)
    F32 = (E32.astype(np.float64) + np.eye(3).reshape(9, 1)).T.reshape(-1, 3, 3)
    clamped = np.linalg.svd(F32, compute_uv=False)[:, 2] < 1e-3
    assert clamped.sum() == 64

    def next_grid(E):
        o = Oracle.from_scene(sc)
        o.setup()
        o.substep(0.0, sc.dt)   # the same layout and ordering as orc
        s = st1.copy()
        s[3:12] = E.astype(np.float64) + np.eye(3).reshape(9, 1)
        o.set_state(0, s)
        o.set_grid_m(M1)
        o.physics(sc.dt, sc.dt)
        return o.grid_mnext()[:, 1:4]

    base = next_grid(E32)
    rng = np.random.default_rng(1)
    Ep = E32.copy()
    step = rng.integers(-1, 2, size=Ep[:, clamped].shape).astype(np.float32)
    Ep[:, clamped] += step * np.spacing(np.abs(Ep[:, clamped]))
    moved = np.abs(next_grid(Ep) - base).max() / np.abs(base).max()
    assert 1e-5 < moved < CLAMPED_NEXT_SUBSTEP_TOL, moved


def test_fluid_stress():
    K, gamma, nu = 1e5, 7.0, 0.1
    params = [0, 0, K, gamma, nu, 1, 30]
    # compressed at rest: isotropic compression with J^-gamma pressure
    for J in (0.97, 1.0, 1.02):
        _, _, tau = pyoracle.stress(FL, params, np.eye(3), C=np.zeros((3, 3)), J=J)
        p = K * (J ** -gamma - 1)
        np.testing.assert_allclose(tau, -J * p * np.eye(3), rtol=1e-12, atol=1e-9)
        assert (np.diag(tau) < 0).all() if J < 1 else True
    # shear flow at J = 1: tau = nu (C + C^T), symmetric, linear in C
    C = np.array([[0.0, 2.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.0, 0.0]])
    _, _, tau = pyoracle.stress(FL, params, np.eye(3), C=C, J=1.0)
    np.testing.assert_allclose(tau, nu * (C + C.T), atol=1e-12)


# ------------------------------------------------------------------ layouts
def test_layout_extension_brute_force():
    dim = (12, 10, 9)
    o = Oracle(dim, 1.0 / 64)
    rng = np.random.default_rng(5)
    coords = {tuple(rng.integers(0, dim)) for _ in range(40)}
    coords |= {(0, 0, 0), (11, 9, 8)}           # domain corners exercise the clipping
    akeys = [(x * dim[1] + y) * dim[2] + z for (x, y, z) in coords]
    rng.shuffle(akeys)
    keys, (A, Nb, Ex) = o.build_layout(np.array(akeys, np.uint32))

    def ext(offsets):
        out = set()
        for (x, y, z) in coords:
            for (a, b, c) in offsets:
                p = (x + a, y + b, z + c)
                if all(0 <= p[i] < dim[i] for i in range(3)):
                    out.add((p[0] * dim[1] + p[1]) * dim[2] + p[2])
        return out

    nb = ext(list(np.ndindex(2, 2, 2)))
    ex = ext([(a - 1, b - 1, c - 1) for (a, b, c) in np.ndindex(3, 3, 3)])
    assert A == len(akeys) and list(keys[:A]) == akeys
    assert list(keys[A:Nb]) == sorted(nb - set(akeys))
    assert list(keys[Nb:Ex]) == sorted(ex - nb - set(akeys))
