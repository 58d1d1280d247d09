"""FP64 differentiable restatement of the CK-MPM substep (TEST INFRASTRUCTURE ONLY).

The reverse-mode reference for the adjoint kernels: the same arithmetic as the C++
oracle's Rasterize / GridUpdate / Physics (oracle/ckmpm_oracle.cpp, itself following
mpm_algorithm.cuh:738-947,1365-1845 and mpm_material.cuh:318-357), written with torch
float64 tensors on a dense node grid so that torch.autograd differentiates it.  The
block machinery (lists, orderings, LDS windows) does not change the arithmetic and is
absent; a node of sub-grid g at integer coordinate n belongs to grid block n // 4.

Fixed-corotated and Neo-Hookean only (the materials with per-particle Lame parameters).
FC's rotation is the polar factor from Newton's iteration R <- (R + R^-T) / 2, which is
differentiable where the SVD's U and V are not (F = I has equal singular values).

Pinned by tests/test_adjoint_ref.py: its forward equals the C++ oracle's (positions,
F, grid) on the same scene, and its gradients equal central finite differences.
"""
from __future__ import annotations

import math

import torch

TWO_PI = 2.0 * math.pi
FIXED_COROTATED, NEO_HOOKEAN = 0, 1


def _stencil(x: torch.Tensor, inv_dx: float, g: int):
    """Base node, offsets t and 1-D weights [N, 3, 2] of sub-grid g (mpm_kernel.cuh:91-114)."""
    u = x * inv_dx + (0.25 if g == 0 else -0.25)
    c = torch.floor(u.detach())
    t = u - c
    w0 = 1.0 - t + torch.sin(TWO_PI * t) / TWO_PI
    return c.long(), t, torch.stack([w0, 1.0 - w0], -1)


def _polar(F: torch.Tensor, iters: int = 14) -> torch.Tensor:
    R = F
    for _ in range(iters):
        R = 0.5 * (R + torch.linalg.inv(R).transpose(-1, -2))
    return R


def kirchhoff(kind: int, F: torch.Tensor, lam: torch.Tensor, mu: torch.Tensor) -> torch.Tensor:
    J = torch.linalg.det(F)
    eye = torch.eye(3, dtype=F.dtype)
    if kind == FIXED_COROTATED:
        # tau = 2 mu (F - R) F^T + lambda J (J - 1) I  (mpm_material.cuh:318-357)
        R = _polar(F)
        return (2.0 * mu[:, None, None] * (F - R) @ F.transpose(-1, -2)
                + (lam * J * (J - 1.0))[:, None, None] * eye)
    # tau = mu (F F^T - I) + lambda ln J I
    return mu[:, None, None] * (F @ F.transpose(-1, -2) - eye) + (lam * torch.log(J))[:, None, None] * eye


class DenseMPM:
    """One model on a dense pair of sub-grids of (4 dim + 2)^3 nodes each."""

    def __init__(self, dim, dx: float, gravity: float, mass: float, volume: float, kind: int, mass_clamp: float = 0.0):
        self.dim = tuple(int(d) for d in dim)
        self.dx = float(dx)
        self.inv_dx = 1.0 / self.dx
        self.gravity = float(gravity)
        self.mass = float(mass)
        self.volume = float(volume)
        self.kind = int(kind)
        self.clamp = float(mass_clamp)
        self.n = tuple(4 * d + 2 for d in self.dim)     # nodes per axis per sub-grid
        nn = self.n[0] * self.n[1] * self.n[2]
        self.nodes = 2 * nn
        # whole grid blocks within 2 blocks of the domain edge are zeroed (mpm_algorithm.cuh:795-805)
        keep = []
        for a in range(3):
            b = torch.arange(self.n[a]) // 4
            keep.append(((b >= 2) & (b + 2 < self.dim[a])).double())
        k = keep[0][:, None, None] * keep[1][None, :, None] * keep[2][None, None, :]
        self.bc = torch.cat([k.reshape(-1), k.reshape(-1)])

    def _nodes(self, x: torch.Tensor):
        """Flat node index and weight of the 16 nodes of each particle, and offsets x_i - x_p."""
        idx, wts, off = [], [], []
        nx, ny, nz = self.n
        for g in range(2):
            c, t, w = _stencil(x, self.inv_dx, g)
            for o in range(8):
                d = ((o >> 2) & 1, (o >> 1) & 1, o & 1)
                ci = [c[:, a] + d[a] for a in range(3)]
                idx.append(g * nx * ny * nz + (ci[0] * ny + ci[1]) * nz + ci[2])
                wts.append(w[:, 0, d[0]] * w[:, 1, d[1]] * w[:, 2, d[2]])
                off.append(torch.stack([(d[a] - t[:, a]) * self.dx for a in range(3)], -1))
        return torch.stack(idx, 1), torch.stack(wts, 1), torch.stack(off, 1)   # [N,16], [N,16], [N,16,3]

    def rasterize(self, x0: torch.Tensor, v0: torch.Tensor):
        """mpm_algorithm.cuh:871-947: m w and m w v0 on both grids."""
        idx, w, _ = self._nodes(x0)
        m = torch.zeros(self.nodes, dtype=torch.float64).index_add(0, idx.reshape(-1), (w * self.mass).reshape(-1))
        p = torch.zeros(self.nodes, 3, dtype=torch.float64).index_add(
            0, idx.reshape(-1), (w[..., None] * self.mass * v0[:, None, :]).reshape(-1, 3))
        return m, p

    def grid_update(self, m: torch.Tensor, p: torch.Tensor, dt: float) -> torch.Tensor:
        live = m > self.clamp
        safe = torch.where(live, m, torch.ones_like(m))
        v = p / safe[:, None]
        v = v + torch.tensor([0.0, self.gravity * dt, 0.0], dtype=torch.float64)
        return v * (live.double() * self.bc)[:, None]

    def g2p2g(self, x, F, V, lam, mu, dt: float, next_dt: float):
        idx, w, off = self._nodes(x)
        vi = V[idx]                                                  # [N,16,3]
        v = 0.5 * (w[..., None] * vi).sum(1)
        B = (w[..., None, None] * vi[..., :, None] * off[..., None, :]).sum(1)
        D = (w[..., None, None] * off[..., :, None] * off[..., None, :]).sum(1)
        C = B @ torch.linalg.inv(D)
        x = x + dt * v
        F = (torch.eye(3, dtype=torch.float64) + dt * C) @ F
        tau = kirchhoff(self.kind, F, lam, mu)
        idx, w, off = self._nodes(x)
        Dn = 0.5 * (w[..., None, None] * off[..., :, None] * off[..., None, :]).sum(1)
        K = C * self.mass - (next_dt * self.volume) * (torch.linalg.inv(Dn) @ tau)
        mom = w[..., None] * (self.mass * v[:, None, :] + (K[:, None, :, :] @ off[..., None]).squeeze(-1))
        m = torch.zeros(self.nodes, dtype=torch.float64).index_add(0, idx.reshape(-1), (w * self.mass).reshape(-1))
        p = torch.zeros(self.nodes, 3, dtype=torch.float64).index_add(0, idx.reshape(-1), mom.reshape(-1, 3))
        return x, F, m, p

    def run(self, x0, v0, lam, mu, steps: int, dt: float):
        """Setup + `steps` substeps; the first uses dt = 0 (mpm_engine.cuh:479).  Returns x_T, F_T."""
        m, p = self.rasterize(x0, v0)
        x = x0
        F = torch.eye(3, dtype=torch.float64).expand(x0.shape[0], 3, 3)
        for k in range(steps):
            d = 0.0 if k == 0 else dt
            V = self.grid_update(m, p, d)
            x, F, m, p = self.g2p2g(x, F, V, lam, mu, d, dt)
        return x, F, (m, p)


def centre_of_mass_loss(x: torch.Tensor, target) -> torch.Tensor:
    """Squared distance of the centre of mass (equal masses) to a target point."""
    t = torch.as_tensor(target, dtype=torch.float64)
    return ((x.mean(0) - t) ** 2).sum()
