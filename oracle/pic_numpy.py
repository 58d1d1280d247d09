"""TEST INFRASTRUCTURE ONLY: a CPU baseline, never part of the product path.

NumPy restatement of the reference's Python CK-MPM demo substep (python-src/python_demo.py,
method 0 of `substep`, :101-215), used as a timed CPU baseline beside the FP64 oracle.  The
demo itself needs Taichi on CUDA (:1-4) and is not run anywhere.

One substep, vectorised over particles (FP32 like the demo's default_fp):
  P2G (:106-132): fixed-corotated first Piola stress times volume,
      PF = -V0 (2 mu (F - R) F^T + lambda (J - 1) J I) (:82-89);
      on each sub-grid w (sign -1, +1): cell = int(x n - 0.25 sign), offset t = x n - cell
      - 0.25 sign, weights w0 = 1 - t + sin(2 pi t)/(2 pi), w1 = 1 - w0 (:63-70); per node I
      of the 2^3 stencil: m += w m_p, p += w m_p v + dt PF grad_I with
      grad_I = n (w_y w_z, w_x w_z, w_x w_y) * sgn(t - I) (cos(2 pi |t - I|) - 1) (:72-76);
  grid update (:153-171): v = p / m where m > 1e-8 m_p, v_y += dt g, zero within 8 cells
      of the faces; max |v|;
  G2P (:175-193): v_p = 0.5 sum w v_i, F <- (I + 0.5 dt sum v_i grad_I^T) F, x += dt v_p.
The sparse Taichi grid becomes the set of touched nodes (np.unique over the node ids).
"""
from __future__ import annotations

import numpy as np

TWO_PI = np.float32(2.0 * np.pi)


class PicDemo:
    """The demo's constants (:7-24): 256^3 grid, rho 1000, 8 particles per cell volume,
    gravity -4, E = 5e3, nu = 0.4."""

    def __init__(self, x: np.ndarray, n_grid: int = 256, E: float = 5e3, nu: float = 0.4, gravity: float = -4.0):
        self.n = n_grid
        self.dx = 1.0 / n_grid
        self.vol = np.float32(self.dx ** 3 / 8)
        self.mass = np.float32(self.vol * 1000.0)
        self.mu = np.float32(E / (2 * (1 + nu)))
        self.lam = np.float32(E * nu / ((1 + nu) * (1 - 2 * nu)))
        self.g = np.float32(gravity)
        self.x = np.ascontiguousarray(x, np.float32)
        self.v = np.zeros_like(self.x)
        self.F = np.tile(np.eye(3, dtype=np.float32), (self.x.shape[0], 1, 1))

    def _stencils(self):
        """Per sub-grid: node ids (N, 8), weights (N, 8), gradients (N, 8, 3)."""
        n = self.n
        out = []
        for w, sign in ((0, -1.0), (1, 1.0)):
            xs = self.x * np.float32(n)
            cell = (xs - np.float32(0.25 * sign)).astype(np.int32)  # C truncation, as ti.cast
            t = xs - (cell + np.float32(0.25 * sign))
            w0 = np.float32(1) - t + np.sin(TWO_PI * t) / TWO_PI
            st = np.stack([w0, np.float32(1) - w0], -1)              # (N, 3, 2)
            ids, wts, grads = [], [], []
            for I in range(8):
                o = ((I >> 2) & 1, (I >> 1) & 1, I & 1)
                kx, ky, kz = st[:, 0, o[0]], st[:, 1, o[1]], st[:, 2, o[2]]
                dw = t - np.array(o, np.float32)
                gr = np.sign(dw) * (np.cos(TWO_PI * np.abs(dw)) - np.float32(1))
                grads.append(np.float32(n) * np.stack([ky * kz, kx * kz, kx * ky], -1) * gr)
                wts.append(kx * ky * kz)
                c = cell + np.array(o, np.int32)
                ids.append(((c[:, 0].astype(np.int64) * n + c[:, 1]) * n + c[:, 2]) * 2 + w)
            out.append((np.stack(ids, 1), np.stack(wts, 1), np.stack(grads, 1)))
        return out

    def substep(self, dt: float) -> float:
        dt = np.float32(dt)
        U, S, Vt = np.linalg.svd(self.F)
        R = U @ Vt
        J = S.prod(1)
        PF = -self.vol * (np.float32(2) * self.mu * (self.F - R) @ self.F.transpose(0, 2, 1)
                          + (self.lam * (J - 1) * J)[:, None, None] * np.eye(3, dtype=np.float32))
        sten = self._stencils()
        ids = np.concatenate([s[0] for s in sten], 1)                # (N, 16)
        wts = np.concatenate([s[1] for s in sten], 1)
        grads = np.concatenate([s[2] for s in sten], 1)              # (N, 16, 3)
        nodes, inv = np.unique(ids.ravel(), return_inverse=True)
        inv = inv.reshape(ids.shape)
        m_nodes = np.bincount(inv.ravel(), (wts * self.mass).ravel(), minlength=nodes.shape[0])
        mom = self.mass * wts[..., None] * self.v[:, None, :] + dt * np.einsum("pab,pib->pia", PF, grads)
        p_nodes = np.stack([np.bincount(inv.ravel(), mom[..., a].ravel(), minlength=nodes.shape[0])
                            for a in range(3)], 1)
        v_nodes = np.zeros_like(p_nodes)
        live = m_nodes > self.mass * 1e-8
        v_nodes[live] = p_nodes[live] / m_nodes[live, None]
        v_nodes[:, 1] += dt * self.g
        cell = nodes // 2
        cz, cy, cx = cell % self.n, (cell // self.n) % self.n, cell // (self.n * self.n)
        oob = np.zeros(nodes.shape[0], bool)
        for c in (cx, cy, cz):
            oob |= (c < 8) | (c >= self.n - 8)
        v_nodes[oob] = 0.0
        vmax = float(np.sqrt((v_nodes ** 2).sum(1)).max()) if nodes.shape[0] else 0.0
        vi = v_nodes[inv].astype(np.float32)                         # (N, 16, 3)
        vp = np.float32(0.5) * (wts[..., None] * vi).sum(1)
        cov = np.einsum("pia,pib->pab", vi, grads)
        self.F = (np.eye(3, dtype=np.float32) + np.float32(0.5) * dt * cov) @ self.F
        self.x = self.x + dt * vp
        self.v = vp
        return vmax
