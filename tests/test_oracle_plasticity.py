"""Pins of the oracle's von Mises and non-associated Cam-clay models (FP64,
oracle/ckmpm_oracle.cpp StressVonMises / StressCamClay; reference mpm_material.cuh:479-671,
mpm_algorithm.cuh:329-520).  Each check uses a route independent of the oracle's formulas:

von Mises (StVK-Hencky elasticity, radial return in log-strain space):
* elastic: tau = 2 mu eps + lambda tr(eps) I with eps = 1/2 logm(F F^T) (scipy), F kept;
* plastic: the returned F has |dev eps'| = yield / (2 mu), tr eps' = tr eps, dev eps'
  parallel to dev eps, the rotation of F (scipy polar) kept, and tau is the Hencky stress
  of the returned F; a pure rotation gives tau = 0.
Cam-clay:
* elastic: the principal Kirchhoff stresses are the derivatives of the model's energy
  psi = mu/2 (J^(-2/3) tr b - 3) + kappa/2 ((J^2 - 1)/2 - log J) with respect to the log
  principal stretches (central differences), F and logJp kept;
* compression / tension cap: the returned F is an isotropic stretch whose pressure is p0
  (or p_min) exactly, and logJp grows by log(J / J_new) only with hardening on;
* shear return: the returned stretches lie on the yield surface at the trial pressure,
  keep tr(s^2) and the direction of dev(s^2);
* hardening on the shear branch: the increment equals 9 log(J / J_fake), J_fake from the
  ray-ellipse intersection found here by bisection (the reference applies the update
  inside a 9-iteration loop, mpm_material.cuh:612-640); zero with hardening off.
"""
import numpy as np
import pytest
from scipy.linalg import logm, polar

from oracle.pyoracle import stress

VM, CC = 4, 5
LAM, MU = 5769.23, 3846.15


def _vm_params(yield_stress):
    return [LAM, MU, 0, 0, 0, 1, 30, yield_stress]


def _cc_params(beta=0.5, xi=0.8, M=2.36, hardening=True, lam=LAM, mu=MU):
    return [lam, mu, 0, 0, 0, 1, 30, 0, 0.0, beta, xi, M, float(hardening)]


def _rot(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def _hencky(F):
    return 0.5 * np.real(logm(F @ F.T))


def _hencky_tau(F, lam=LAM, mu=MU):
    e = _hencky(F)
    return 2 * mu * e + lam * np.trace(e) * np.eye(3)


@pytest.mark.parametrize("seed", range(4))
def test_von_mises_elastic_branch_is_hencky(seed):
    R, Q = _rot(seed), _rot(seed + 10)
    F = R @ np.diag([1.001, 0.9995, 1.0002]) @ Q.T
    dev = _hencky(F) - np.trace(_hencky(F)) / 3 * np.eye(3)
    ys = 2 * MU * np.linalg.norm(dev) * 1.5     # 50% above the current deviatoric radius
    Fo, _, tau = stress(VM, _vm_params(ys), F)
    np.testing.assert_allclose(Fo, F, rtol=0, atol=1e-13)
    np.testing.assert_allclose(tau, _hencky_tau(F), rtol=0, atol=1e-9 * np.abs(tau).max())


@pytest.mark.parametrize("seed", range(4))
def test_von_mises_plastic_return(seed):
    R, Q = _rot(seed), _rot(seed + 20)
    F = R @ np.diag([1.05, 0.97, 1.01]) @ Q.T
    e = _hencky(F)
    tr = np.trace(e)
    dev = e - tr / 3 * np.eye(3)
    ys = 2 * MU * np.linalg.norm(dev) * 0.4     # well inside: the return is active
    Fo, _, tau = stress(VM, _vm_params(ys), F)
    e2 = _hencky(Fo)
    dev2 = e2 - np.trace(e2) / 3 * np.eye(3)
    assert np.linalg.norm(dev2) == pytest.approx(ys / (2 * MU), rel=1e-10)
    assert np.trace(e2) == pytest.approx(tr, abs=1e-12)
    np.testing.assert_allclose(dev2 / np.linalg.norm(dev2), dev / np.linalg.norm(dev), atol=1e-10)
    np.testing.assert_allclose(polar(Fo)[0], polar(F)[0], atol=1e-10)   # same rotation
    np.testing.assert_allclose(tau, _hencky_tau(Fo), rtol=0, atol=1e-9 * np.abs(tau).max())
    # a pure rotation carries no stress
    _, _, t0 = stress(VM, _vm_params(ys), R)
    assert np.abs(t0).max() < 1e-9


# ------------------------------------------------------------------------------ Cam-clay
def _kappa(lam=LAM, mu=MU):
    return 2 * mu / 3 + lam


def _p0(logJp, xi=0.8):
    return _kappa() * (1e-5 + np.sinh(xi * max(-logJp, 0.0)))


def _psi(logs, mu=MU, kappa=None):
    kappa = _kappa() if kappa is None else kappa
    s = np.exp(logs)
    J = s.prod()
    return mu / 2 * (J ** (-2 / 3) * (s ** 2).sum() - 3) + kappa / 2 * ((J * J - 1) / 2 - np.log(J))


def test_cam_clay_elastic_stress_is_the_energy_derivative():
    logJp = -0.05                                 # p0 = kappa (1e-5 + sinh(0.04)) > 0
    s = np.array([1.0004, 0.9998, 1.0001])        # small strain: inside the yield surface
    R, Q = _rot(3), _rot(4)
    F = R @ np.diag(s) @ Q.T
    Fo, lj, tau = stress(CC, _cc_params(), F, logJp=logJp)
    np.testing.assert_allclose(Fo, F, atol=1e-13)
    assert lj == logJp
    h = 1e-6
    g = np.array([(_psi(np.log(s) + h * np.eye(3)[i]) - _psi(np.log(s) - h * np.eye(3)[i])) / (2 * h)
                  for i in range(3)])
    np.testing.assert_allclose(tau, R @ np.diag(g) @ R.T, rtol=0, atol=1e-5 * np.abs(g).max() + 1e-6)
    # a pure rotation: zero pressure, zero deviator, nothing changes
    Fo, lj, tau = stress(CC, _cc_params(), R, logJp=logJp)
    np.testing.assert_allclose(Fo, R, atol=1e-13)
    assert lj == logJp and np.abs(tau).max() < 1e-8


@pytest.mark.parametrize("side", ["compression", "tension"])
@pytest.mark.parametrize("hardening", [True, False])
def test_cam_clay_cap_projection(side, hardening):
    logJp, beta = -0.01, 0.5
    p0 = _p0(logJp)
    pcap = p0 if side == "compression" else -beta * p0
    s = 0.97 if side == "compression" else 1.03       # J = s^3 well beyond the cap
    J = s ** 3
    assert (-_kappa() / 2 * (J * J - 1) > p0) == (side == "compression")
    R = _rot(7)
    Fo, lj, tau = stress(CC, _cc_params(beta=beta, hardening=hardening), R * s, logJp=logJp)
    Jn = np.linalg.det(Fo)
    np.testing.assert_allclose(Fo, Jn ** (1 / 3) * R, atol=1e-12)      # isotropic stretch, same rotation
    assert -_kappa() / 2 * (Jn * Jn - 1) == pytest.approx(pcap, rel=1e-10)
    np.testing.assert_allclose(tau, -pcap * np.eye(3), rtol=0, atol=1e-8 * abs(pcap))
    assert lj == pytest.approx(logJp + (np.log(J / Jn) if hardening else 0.0), abs=1e-13)


def _shear_case():
    s = np.array([1.02, 0.985, 0.9985])
    return s, -0.02


def test_cam_clay_shear_return_lands_on_the_yield_surface():
    beta, M = 0.5, 2.36
    s, logJp = _shear_case()
    J = s.prod()
    kappa, p0 = _kappa(), _p0(logJp)
    pmin = -beta * p0
    ptr = -kappa / 2 * (J * J - 1)
    assert pmin < ptr < p0
    R, Q = _rot(11), _rot(12)
    Fo, _, _ = stress(CC, _cc_params(beta=beta, M=M, hardening=False), R @ np.diag(s) @ Q.T, logJp=logJp)
    ns = np.linalg.svd(Fo, compute_uv=False)
    ns = ns[np.argsort(-ns)]
    order = np.argsort(-s)
    b, bn = s[order] ** 2, ns ** 2
    assert bn.sum() == pytest.approx(b.sum(), rel=1e-12)                       # tr(s^2) kept
    dv, dn = b - b.mean(), bn - bn.mean()
    np.testing.assert_allclose(dn / np.linalg.norm(dn), dv / np.linalg.norm(dv), atol=1e-9)
    sh = MU * J ** (-2 / 3) * dn                                                # at the trial J
    y = 1.5 * (1 + 2 * beta) * (sh ** 2).sum() + M * M * (ptr - pmin) * (ptr - p0)
    assert abs(y) <= 1e-9 * M * M * abs((ptr - pmin) * (ptr - p0))
    # the trial state was outside (y >= 1e-4), so a return happened
    sh0 = MU * J ** (-2 / 3) * dv
    assert 1.5 * (1 + 2 * beta) * (sh0 ** 2).sum() + M * M * (ptr - pmin) * (ptr - p0) >= 1e-4


@pytest.mark.parametrize("hardening", [True, False])
def test_cam_clay_shear_hardening_increment(hardening):
    beta, M = 0.5, 2.36
    s, logJp = _shear_case()
    J = s.prod()
    kappa, p0 = _kappa(), _p0(logJp)
    pmin = -beta * p0
    ptr = -kappa / 2 * (J * J - 1)
    b = s ** 2
    q = np.sqrt(1.5 * ((MU * J ** (-2 / 3) * (b - b.mean())) ** 2).sum())
    # ray from the ellipse centre (pc, 0) through (ptr, -q): bisection for the surface
    pc = (1 - beta) * p0 / 2
    d = np.array([ptr - pc, -q])
    d /= np.linalg.norm(d)

    def f(t):
        p, qq = pc + t * d[0], t * d[1]
        return (1 + 2 * beta) * qq * qq + M * M * (p - pmin) * (p - p0)

    lo, hi = 0.0, 1.0
    while f(hi) < 0:
        hi *= 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if f(mid) < 0 else (lo, mid)
    pf = pc + 0.5 * (lo + hi) * d[0]
    Jf = np.sqrt(abs(1 - 2 * pf / kappa))
    _, lj, _ = stress(CC, _cc_params(beta=beta, M=M, hardening=hardening), np.diag(s), logJp=logJp)
    want = logJp + (9 * np.log(J / Jf) if hardening else 0.0)
    assert lj == pytest.approx(want, abs=1e-11)


@pytest.mark.parametrize("kind", [VM, CC])
def test_gpu_parity_box_scenes_reach_the_plastic_branches(kind):
    """The random-velocity boxes that the GPU lockstep test runs for these models do
    leave the elastic range within its 6 substeps: von Mises particles end on the yield
    radius, Cam-clay particles change logJp (cap or shear hardening)."""
    from oracle.pyoracle import Oracle
    from tests.test_gpu_parity import _box_scene

    sc = _box_scene(kind, seed=10 + kind)
    orc = Oracle.from_scene(sc)
    orc.setup()
    for k in range(6):
        orc.substep(0.0 if k == 0 else sc.dt, sc.dt)
    st = orc.state(0)
    F = st[3:12].T.reshape(-1, 3, 3)
    mat = sc.models[0].material
    if kind == VM:
        s = np.linalg.svd(F, compute_uv=False)
        e = np.log(s)
        dev = np.linalg.norm(e - e.mean(1, keepdims=True), axis=1)
        on_radius = np.abs(dev - mat.yield_stress / (2 * mat.mu)) < 1e-9
        assert (dev <= mat.yield_stress / (2 * mat.mu) + 1e-9).all()
        assert on_radius.mean() > 0.05, on_radius.mean()
    else:
        changed = np.abs(st[12] - mat.alpha0) > 1e-12
        assert changed.mean() > 0.05, changed.mean()
