"""Pins the FP64 differentiable reference (oracle/adjoint_ref.py).

* forward: equal to the C++ oracle's substeps (positions and F) on a scene with
  per-particle Lame parameters and per-particle initial velocities;
* gradients: torch.autograd through it equals central finite differences (eps = 1e-3
  on log E and on v0) of the same FP64 forward, for the centre-of-mass loss.
A wrong sign or a dropped term in the reference's forward fails the first check; the
second pins what autograd computes.
"""
import math

import numpy as np
import pytest
import torch

from ckmpm_amd import scenes
from oracle.adjoint_ref import DenseMPM, centre_of_mass_loss
from oracle.pyoracle import Oracle

SLOT = (1 << 27) - 1
NU = 0.3


def _small(seed=6, v_scale=0.3):
    sc = scenes.cantilever(seed=seed, scale=4)   # 32^3 cells, bar 19 x 2 x 2 cells
    m = sc.models[0]
    rng = scenes.Splitmix(seed + 100)
    m.v = (v_scale * rng.normal(3 * m.x.shape[0]).reshape(-1, 3)).astype(np.float32)
    return sc


def _dense(sc):
    m = sc.models[0]
    return DenseMPM(sc.dim, sc.dx, sc.gravity, m.mass, m.volume, m.material.kind, sc.mass_clamp)


def _params(sc):
    m = sc.models[0]
    x0 = torch.tensor(m.x.astype(np.float64))
    v0 = torch.tensor(m.v.astype(np.float64))
    E = m.mu_p.astype(np.float64) * 2 * (1 + NU)
    return x0, v0, torch.tensor(np.log(E))


def _lame(logE):
    E = torch.exp(logE)
    return E * NU / ((1 + NU) * (1 - 2 * NU)), E / (2 * (1 + NU))


def test_forward_matches_cpp_oracle():
    sc = _small()
    steps = 6
    o = Oracle.from_scene(sc)
    o.setup()
    ids = o.setup_order(0).astype(np.int64)
    for k in range(steps):
        ids = ids[o.perm(0) & SLOT]
        o.substep(0.0 if k == 0 else sc.dt, sc.dt)
    assert o.flags == 0
    st = o.state(0)
    x_o = np.empty((st.shape[1], 3))
    F_o = np.empty((st.shape[1], 9))
    x_o[ids] = st[:3].T
    F_o[ids] = st[3:12].T
    x0, v0, logE = _params(sc)
    # the oracle stores the FP32-rounded per-particle parameters
    lam = torch.tensor(sc.models[0].lam_p.astype(np.float64))
    mu = torch.tensor(sc.models[0].mu_p.astype(np.float64))
    x, F, _ = _dense(sc).run(x0, v0, lam, mu, steps, sc.dt)
    np.testing.assert_allclose(x.numpy(), x_o, rtol=0, atol=1e-12)
    np.testing.assert_allclose(F.reshape(-1, 9).numpy(), F_o, rtol=0, atol=1e-10)


def _loss(sim, x0, v0, logE, steps, dt, target):
    lam, mu = _lame(logE)
    x, _, _ = sim.run(x0, v0, lam, mu, steps, dt)
    return centre_of_mass_loss(x, target)


@pytest.mark.parametrize("wrt", ["logE", "v0"])
def test_gradients_match_central_differences(wrt):
    sc = _small()
    sim = _dense(sc)
    x0, v0, logE = _params(sc)
    steps = 12
    # target 1e-3 away from the simulated centre of mass: a small loss keeps its
    # rounding far below the finite difference
    with torch.no_grad():
        lam, mu = _lame(logE)
        xc = sim.run(x0, v0, lam, mu, steps, sc.dt)[0].mean(0)
    target = tuple(float(c) + 1e-3 for c in xc)
    logE.requires_grad_(True)
    v0.requires_grad_(True)
    L = _loss(sim, x0, v0, logE, steps, sc.dt, target)
    gE, gv = torch.autograd.grad(L, [logE, v0])
    rng = np.random.default_rng(3)
    eps = 1e-3
    with torch.no_grad():
        if wrt == "logE":
            u = torch.tensor(rng.standard_normal(logE.shape[0]))
            Lp = _loss(sim, x0, v0, logE + eps * u, steps, sc.dt, target)
            Lm = _loss(sim, x0, v0, logE - eps * u, steps, sc.dt, target)
            analytic = float((gE * u).sum())
        else:
            u = torch.tensor(rng.standard_normal(v0.shape))
            Lp = _loss(sim, x0, v0 + eps * u, logE, steps, sc.dt, target)
            Lm = _loss(sim, x0, v0 - eps * u, logE, steps, sc.dt, target)
            analytic = float((gv * u).sum())
    fd = float(Lp - Lm) / (2 * eps)
    assert abs(analytic) > 0
    assert abs(fd - analytic) <= 1e-4 * abs(analytic) + 1e-14, (fd, analytic)
