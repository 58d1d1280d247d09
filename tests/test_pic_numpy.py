"""The NumPy restatement of the reference's Python CK-MPM demo (oracle/pic_numpy.py), the
CPU baseline beside the oracle.  Pinned by what the transfer must preserve: a uniform
translation with gravity off keeps every particle's velocity and leaves F = I (the
weights of each sub-grid sum to one and their gradients to zero), the grid holds twice
the particle mass (two sub-grids), and a body at rest falls as v = g dt per substep."""
import numpy as np

from oracle.pic_numpy import PicDemo


def _block(n=2000, seed=0):
    rng = np.random.default_rng(seed)
    return (0.4 + 0.2 * rng.random((n, 3))).astype(np.float32)


def test_uniform_translation_is_kept():
    p = PicDemo(_block(), gravity=0.0)
    v = np.array([0.3, -0.2, 0.1], np.float32)
    p.v[:] = v
    p.substep(1e-4)
    np.testing.assert_allclose(p.v, np.broadcast_to(v, p.v.shape), rtol=0, atol=2e-6)
    np.testing.assert_allclose(p.F, np.broadcast_to(np.eye(3), p.F.shape), rtol=0, atol=1e-6)


def test_grid_mass_and_free_fall():
    x = _block()
    p = PicDemo(x)
    sten = p._stencils()
    total = sum(float((s[1] * p.mass).sum()) for s in sten)
    np.testing.assert_allclose(total, 2 * x.shape[0] * float(p.mass), rtol=1e-5)
    for k in range(3):
        p.substep(1e-4)
    np.testing.assert_allclose(p.v[:, 1], 3 * 1e-4 * -4.0, rtol=1e-4)
    np.testing.assert_allclose(p.v[:, [0, 2]], 0.0, atol=1e-9)
