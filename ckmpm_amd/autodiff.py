"""Differentiable CK-MPM simulation behind torch.autograd.Function.

``simulate(log_E, v0, scene, steps)`` runs `steps` substeps of a one-model fixed-
corotated (or Neo-Hookean) scene on the GPU with per-particle Young's modulus
E = exp(log_E) (Poisson ratio fixed) and per-particle initial velocity v0, and returns
the final positions in input order.  Its backward replays the run from checkpoints
taken every `segment` substeps, recording each segment's substeps and undoing them with
the adjoint kernels (ckmpm_adjoint_*), then maps (lambda-bar, mu-bar) to log_E-bar.
The forward records the last segment as it runs, so the backward undoes that segment
at once and replays only the earlier ones.
With segment=None the forward records every substep (a tape of all of them, which
HBM holds for scenes of this size) and the backward undoes them without replaying.
All arithmetic is the library's (FP32 kernels); torch only carries the tensors.

The engine of a call is kept after its backward and reused by the next call on the
same scene (positions, domain, material and tape length): it is rewound to its setup
state with the new velocities and Lame parameters (ckmpm_adjoint_rewind) instead of
being rebuilt, and checkpoint memory is reused, so a gradient loop allocates nothing
after its first iteration.
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from .engine import engine_from_scene
from .scenes import Scene


class _EngineCache:
    """The engine of the last finished call, keyed by everything its setup depends on
    except the velocities and Lame parameters."""

    def __init__(self):
        self.key = None
        self.x = None
        self.eng = None

    @staticmethod
    def key_of(scene: Scene, tape: int):
        m = scene.models[0]
        mat = m.material
        return (tuple(scene.dim), tuple(scene.offset), scene.dx, scene.gravity, scene.mass_clamp, scene.dt,
                scene.block_capacity(), mat.kind, mat.lam, mat.mu, mat.J0, m.mass, m.volume, m.x.shape[0], tape)

    def take(self, scene: Scene, tape: int):
        """The cached engine if it matches (removed from the cache), else None."""
        key = self.key_of(scene, tape)
        if self.eng is not None and self.key == key and np.array_equal(self.x, scene.models[0].x):
            eng, self.eng = self.eng, None
            return eng
        return None

    def give(self, scene: Scene, tape: int, eng):
        self.clear()
        self.key, self.x, self.eng = self.key_of(scene, tape), np.array(scene.models[0].x, copy=True), eng

    def clear(self):
        if self.eng is not None:
            self.eng.close()
        self.key = self.x = self.eng = None


_CACHE = _EngineCache()


def clear_engine_cache():
    """Release the engine kept for the next simulate() call."""
    _CACHE.clear()


def _lame_factors(nu: float):
    return nu / ((1 + nu) * (1 - 2 * nu)), 1.0 / (2 * (1 + nu))


class _Simulate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_E, v0, scene: Scene, steps: int, segment: int, nu: float):
        fl, fm = _lame_factors(nu)
        E = torch.exp(log_E.detach().double()).cpu().numpy()
        sc = copy.copy(scene)
        m = copy.copy(scene.models[0])
        m.lam_p = (E * fl).astype(np.float32)
        m.mu_p = (E * fm).astype(np.float32)
        m.v = v0.detach().cpu().numpy().astype(np.float32).reshape(-1, 3)
        sc.models = [m]
        full = segment is None
        tape = steps if full else segment
        eng = _CACHE.take(sc, tape)
        if eng is not None:
            eng.adjoint_rewind(m.v, m.lam_p, m.mu_p)
        else:
            eng = engine_from_scene(sc, adjoint_segment=tape)
        cks = []
        # first substep of the recorded tail: the whole run, or the last segment
        tail = 0 if full else (steps - 1) // segment * segment
        for k in range(steps):
            if k == tail:
                eng.adjoint_record(True)
            elif not full and k % segment == 0:
                cks.append(eng.checkpoint())
            eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
        eng.adjoint_record(False)
        eng.Synchronize()
        x = eng.positions_input_order(0)
        ctx.eng, ctx.cks, ctx.sc, ctx.steps, ctx.segment, ctx.tape = eng, cks, sc, steps, segment, tape
        ctx.lam, ctx.mu = m.lam_p.astype(np.float64), m.mu_p.astype(np.float64)
        return torch.from_numpy(x.astype(np.float64)).to(log_E.device)

    @staticmethod
    def backward(ctx, gx):
        if ctx.eng is None:
            # the first backward released the checkpoints and handed the engine on
            raise RuntimeError("simulate(): backward through the same call a second time is not supported "
                               "(retain_graph / repeated autograd.grad); call simulate() again")
        eng, cks, sc, steps, seg = ctx.eng, ctx.cks, ctx.sc, ctx.steps, ctx.segment
        ctx.eng, ctx.cks = None, None
        eng.adjoint_reset()
        eng.adjoint_seed(0, gx.detach().cpu().numpy().astype(np.float32))
        tail = 0 if seg is None else (steps - 1) // seg * seg
        for _ in range(tail, steps):
            eng.adjoint_substep()
        for s in reversed(range(len(cks))):
            eng.restore(cks[s])
            eng.adjoint_record(True)
            for k in range(s * seg, min((s + 1) * seg, steps)):
                eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
            eng.adjoint_record(False)
            for _ in range(s * seg, min((s + 1) * seg, steps)):
                eng.adjoint_substep()
        lb, mb, vb = eng.adjoint_gradients(0)
        for c in cks:
            eng.free_checkpoint(c)
        _CACHE.give(sc, ctx.tape, eng)
        # lambda and mu are both proportional to E: d/dlogE = lambda-bar lambda + mu-bar mu
        g_logE = torch.from_numpy(lb.astype(np.float64) * ctx.lam + mb.astype(np.float64) * ctx.mu)
        g_v0 = torch.from_numpy(vb.astype(np.float64))
        return g_logE.to(gx.device), g_v0.to(gx.device), None, None, None, None


def simulate(log_E: torch.Tensor, v0: torch.Tensor, scene: Scene, steps: int, segment=25,
             nu: float = 0.3) -> torch.Tensor:
    """Final positions (n, 3, float64, input order) of `steps` substeps of `scene`'s
    model with Young's modulus exp(log_E) and initial velocities v0 (n, 3)."""
    return _Simulate.apply(log_E, v0, scene, steps, segment, nu)
