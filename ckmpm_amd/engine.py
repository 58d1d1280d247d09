"""Python mirror of the reference's engine API over the C-ABI.

Names follow the reference (include/mpm_engine.cuh, mpm_material.cuh, mpm_model.h,
mpm_domain.h): ``MPMDomain``, ``MPMMaterial``, ``MPMModel``, ``MPMConfig`` and
``MPMEngine`` with ``Initialize`` / ``InitializeParticle`` / ``InitialSetup`` /
``Simulate``.  Unlike the reference there is no singleton (any number of engines) and
no call exits the process: errors raise ``CkmpmError``.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import List, Optional, Sequence

import numpy as np

from . import native
from .native import CkmpmError, _f32, _ptr, _f32p, _u32p
from .scenes import CAM_CLAY, DRUCKER_PRAGER, FIXED_COROTATED, FLUID, NEO_HOOKEAN, VON_MISES, Scene

NATTR = {FIXED_COROTATED: 12, NEO_HOOKEAN: 12, DRUCKER_PRAGER: 13, FLUID: 4, VON_MISES: 12, CAM_CLAY: 13}


@dataclasses.dataclass
class MPMDomain:
    """MPMDomain<MPMDomainRange<...>, MPMDomainOffset<...>> (mpm_domain.h:53-61), in blocks."""
    dim: Sequence[int]
    offset: Sequence[int] = (0, 0, 0)


@dataclasses.dataclass
class MPMConfig:
    """MPMConfigBase getters (mpm_config.h:17-123)."""
    dx: float
    gravity: float = 0.0
    mass_clamp: float = 0.0
    cfl: float = 0.5
    dt: float = 1e-4
    fps: int = 48
    max_active_block_count: int = 1 << 16


@dataclasses.dataclass
class MPMMaterial:
    """MPMMaterial<Model>::parameter_ (mpm_material.cuh:62-175)."""
    kind: int
    lam: float = 0.0
    mu: float = 0.0
    bulk: float = 0.0
    gamma: float = 0.0
    viscosity: float = 0.0
    J0: float = 1.0
    friction_angle_deg: float = 30.0
    yield_stress: float = 0.0       # VonMisesMaterialParameter (mpm_material.cuh:236-253)
    alpha0: float = 0.0             # NonAssociatedCamClayMaterialParameter (mpm_material.cuh:189-214)
    beta: float = 0.0
    hardening_factor: float = 0.0
    m: float = 0.0
    hardening_on: bool = True


@dataclasses.dataclass
class MPMModel:
    """MPMModel (mpm_model.h:10-60): constant particle mass / volume, positions, 1 or n velocities."""
    mass: float
    volume: float
    positions: np.ndarray
    velocities: np.ndarray
    material: MPMMaterial
    lam_p: Optional[np.ndarray] = None
    mu_p: Optional[np.ndarray] = None


class MPMEngine:
    def __init__(self):
        self._lib = native.load()
        self._h = ctypes.c_void_p()
        self.models: List[MPMModel] = []

    # -- lifecycle ------------------------------------------------------------------
    def _check(self, rc: int, where: str):
        if rc != native.OK:
            msg = self._lib.ckmpm_last_error(self._h).decode() if self._h else ""
            raise CkmpmError(rc, where, msg)

    def Initialize(self, domain: MPMDomain, config: MPMConfig):
        d = native.DomainDesc((ctypes.c_int32 * 3)(*domain.dim), (ctypes.c_int32 * 3)(*domain.offset))
        s = native.SimDesc(config.dx, config.gravity, config.mass_clamp, config.cfl, config.dt, config.fps,
                           config.max_active_block_count)
        self.config = config
        self.domain = domain
        rc = self._lib.ckmpm_create(ctypes.byref(d), ctypes.byref(s), ctypes.byref(self._h))
        self._check(rc, "ckmpm_create")

    def InitializeParticle(self, models: Sequence[MPMModel]):
        for m in models:
            x = _f32(m.positions).reshape(-1, 3)
            v = _f32(m.velocities).reshape(-1, 3)
            mat = m.material
            md = native.MaterialDesc(mat.kind, mat.lam, mat.mu, mat.bulk, mat.gamma, mat.viscosity, mat.J0,
                                     mat.friction_angle_deg, mat.yield_stress, mat.alpha0, mat.beta,
                                     mat.hardening_factor, mat.m, int(mat.hardening_on))
            lp = None if m.lam_p is None else _f32(m.lam_p)
            mp = None if m.mu_p is None else _f32(m.mu_p)
            idx = ctypes.c_int32()
            rc = self._lib.ckmpm_add_model(self._h, ctypes.byref(md), m.mass, m.volume, _ptr(x, _f32p), x.shape[0],
                                           _ptr(v, _f32p), v.shape[0], _ptr(lp, _f32p), _ptr(mp, _f32p),
                                           ctypes.byref(idx))
            self._check(rc, "ckmpm_add_model")
            self.models.append(m)

    def set_rigid(self, model: int, count: int, velocity) -> None:
        """The first `count` particles of `model` (input order) are rigid and move at
        `velocity`; the count is fixed before InitialSetup, the velocity may change
        between substeps (GetRigidParticleCount / GetRigidParticleVelocity)."""
        v = np.ascontiguousarray(velocity, np.float32).reshape(3)
        self._check(self._lib.ckmpm_set_rigid(self._h, int(model), int(count), _ptr(v, _f32p)), "ckmpm_set_rigid")

    def InitialSetup(self):
        self._check(self._lib.ckmpm_initial_setup(self._h), "ckmpm_initial_setup")

    def Substep(self, dt: float, next_dt: float):
        self._check(self._lib.ckmpm_substep(self._h, float(dt), float(next_dt)), "ckmpm_substep")

    def Simulate(self, frames: int):
        """Frame loop with the reference dt rules (mpm_engine.cuh:514-876)."""
        self._check(self._lib.ckmpm_run_frames(self._h, int(frames)), "ckmpm_run_frames")

    def Synchronize(self):
        self._check(self._lib.ckmpm_synchronize(self._h), "ckmpm_synchronize")

    def ClearFlags(self):
        self._check(self._lib.ckmpm_clear_flags(self._h), "ckmpm_clear_flags")

    def counters(self) -> native.Counters:
        c = native.Counters()
        self._check(self._lib.ckmpm_get_counters(self._h, ctypes.byref(c)), "ckmpm_get_counters")
        return c

    def autotune(self, substeps: int = 3):
        """Pick each model's G2P2G workgroup size by timing real substeps (the scene advances)."""
        self._check(self._lib.ckmpm_autotune(self._h, int(substeps)), "ckmpm_autotune")

    def set_tuning(self, model: int, threads: int):
        self._check(self._lib.ckmpm_set_tuning(self._h, model, int(threads)), "ckmpm_set_tuning")

    def tuning(self, model: int) -> int:
        t = ctypes.c_int32()
        self._check(self._lib.ckmpm_get_tuning(self._h, model, ctypes.byref(t)), "ckmpm_get_tuning")
        return int(t.value)

    def set_profiling(self, enable: bool):
        self._check(self._lib.ckmpm_set_profiling(self._h, int(enable)), "set_profiling")

    def phase_times(self):
        """(grid update, G2P2G, rebin) device milliseconds summed since the last call, substeps."""
        ms = (ctypes.c_double * 3)()
        n = ctypes.c_uint64()
        self._check(self._lib.ckmpm_get_phase_times(self._h, ms, ctypes.byref(n)), "get_phase_times")
        return tuple(ms), int(n.value)

    def set_boundary(self, boundary):
        """Grid boundary (scenes.Boundary; None: the default sticky walls)."""
        if boundary is None:
            self._check(self._lib.ckmpm_set_boundary(self._h, None), "ckmpm_set_boundary")
            return
        if len(boundary.regions) > native.MAX_REGIONS:
            raise CkmpmError(native.ERR_INVALID, "ckmpm_set_boundary", "too many regions")
        d = native.BoundaryDesc()
        d.walls = boundary.walls
        d.nregions = len(boundary.regions)
        for i, r in enumerate(boundary.regions):
            g = d.regions[i]
            g.lo[:] = [int(c) for c in r.lo]
            g.hi[:] = [int(c) for c in r.hi]
            g.kind = r.kind
            g.value[:] = [float(c) for c in r.value]
            g.center[:] = [float(c) for c in r.center]
        self._check(self._lib.ckmpm_set_boundary(self._h, ctypes.byref(d)), "ckmpm_set_boundary")

    def set_conservation(self, capacity: int):
        """Record the per-substep conservation metric (up to `capacity` substeps; 0: off)."""
        self._check(self._lib.ckmpm_set_conservation(self._h, int(capacity)), "ckmpm_set_conservation")

    def conservation(self, model: int) -> np.ndarray:
        """(n, 13) float32 records of the reference's conservationMetric file: time,
        Eulerian linear / angular momentum (zero), Lagrangian linear / angular momentum."""
        n = ctypes.c_uint64()
        self._check(self._lib.ckmpm_get_conservation(self._h, model, None, 0, ctypes.byref(n)), "get_conservation")
        out = np.zeros((n.value, 13), np.float32)
        if n.value:
            self._check(self._lib.ckmpm_get_conservation(self._h, model, _ptr(out, _f32p), n.value, ctypes.byref(n)),
                        "ckmpm_get_conservation")
        return out

    def close(self):
        if self._h:
            self._lib.ckmpm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- adjoint ------------------------------------------------------------------
    def adjoint_enable(self, segment: int):
        """Before InitialSetup: keep a tape of up to `segment` substeps and particle identities."""
        self._check(self._lib.ckmpm_adjoint_enable(self._h, int(segment)), "ckmpm_adjoint_enable")

    def checkpoint(self) -> ctypes.c_void_p:
        c = ctypes.c_void_p()
        self._check(self._lib.ckmpm_checkpoint_save(self._h, ctypes.byref(c)), "ckmpm_checkpoint_save")
        return c

    def restore(self, ck: ctypes.c_void_p):
        self._check(self._lib.ckmpm_checkpoint_restore(self._h, ck), "ckmpm_checkpoint_restore")

    def free_checkpoint(self, ck: ctypes.c_void_p):
        self._lib.ckmpm_checkpoint_free(ck)

    def adjoint_rewind(self, v: np.ndarray, lam: np.ndarray, mu: np.ndarray):
        """Back to the state right after InitialSetup (one model) with new initial
        velocities (1 or n rows) and per-particle Lame parameters; allocates nothing."""
        vv = _f32(v).reshape(-1, 3)
        la, mu_ = _f32(lam).reshape(-1), _f32(mu).reshape(-1)
        self._check(self._lib.ckmpm_adjoint_rewind(self._h, _ptr(vv, _f32p), vv.shape[0], _ptr(la, _f32p),
                                                   _ptr(mu_, _f32p)), "ckmpm_adjoint_rewind")

    def adjoint_record(self, on: bool):
        self._check(self._lib.ckmpm_adjoint_record(self._h, int(on)), "ckmpm_adjoint_record")

    def adjoint_reset(self):
        self._check(self._lib.ckmpm_adjoint_reset(self._h), "ckmpm_adjoint_reset")

    def adjoint_seed(self, model: int, xbar: np.ndarray):
        xb = _f32(xbar).reshape(-1)
        self._check(self._lib.ckmpm_adjoint_seed(self._h, model, _ptr(xb, _f32p)), "ckmpm_adjoint_seed")

    def adjoint_substep(self):
        self._check(self._lib.ckmpm_adjoint_substep(self._h), "ckmpm_adjoint_substep")

    def adjoint_gradients(self, model: int):
        n = self.n(model)
        lb = np.empty(n, np.float32)
        mb = np.empty(n, np.float32)
        vb = np.empty((n, 3), np.float32)
        self._check(self._lib.ckmpm_adjoint_gradients(self._h, model, _ptr(lb, _f32p), _ptr(mb, _f32p),
                                                      _ptr(vb, _f32p)), "ckmpm_adjoint_gradients")
        return lb, mb, vb

    def positions_input_order(self, model: int) -> np.ndarray:
        out = np.empty((self.n(model), 3), np.float32)
        self._check(self._lib.ckmpm_get_positions_input_order(self._h, model, _ptr(out, _f32p)),
                    "ckmpm_get_positions_input_order")
        return out

    # -- state access ---------------------------------------------------------------
    def n(self, model: int) -> int:
        n = ctypes.c_uint64()
        na = ctypes.c_int32()
        self._check(self._lib.ckmpm_model_info(self._h, model, ctypes.byref(n), ctypes.byref(na)), "model_info")
        return int(n.value)

    def nattr(self, model: int) -> int:
        n = ctypes.c_uint64()
        na = ctypes.c_int32()
        self._check(self._lib.ckmpm_model_info(self._h, model, ctypes.byref(n), ctypes.byref(na)), "model_info")
        return int(na.value)

    def positions(self, model: int) -> np.ndarray:
        n = self.n(model)
        out = np.zeros((n, 3), np.float32)
        got = ctypes.c_uint64()
        self._check(self._lib.ckmpm_get_positions(self._h, model, _ptr(out, _f32p), n, ctypes.byref(got)),
                    "get_positions")
        return out

    def layout(self, which: int = 0):
        counts = np.zeros(3, np.uint32)
        self._check(self._lib.ckmpm_export_layout(self._h, which, None, 0, _ptr(counts, _u32p)), "export_layout")
        keys = np.zeros(int(counts[2]), np.uint32)
        self._check(self._lib.ckmpm_export_layout(self._h, which, _ptr(keys, _u32p), keys.shape[0],
                                                  _ptr(counts, _u32p)), "export_layout")
        return keys, tuple(int(c) for c in counts)

    def ordering(self, model: int, which: int = 0):
        """(perm, offsets) of the current (0) or previous (1) ordering; perm = slot | move << 27."""
        A = self.layout(which)[1][0]
        perm = np.zeros(self.n(model), np.uint32)
        off = np.zeros(A + 1, np.uint32)
        self._check(self._lib.ckmpm_export_ordering(self._h, model, which, _ptr(perm, _u32p), _ptr(off, _u32p),
                                                    A + 1), "export_ordering")
        return perm, off

    def global_state(self, model: int) -> np.ndarray:
        """State in the reference's variables, FP64: global positions (exact origin + u),
        F = I + E for a solid, J = 1 + q for a fluid."""
        st = self.state(model).astype(np.float64)
        keys, (A, _, _) = self.layout(1)
        _, off = self.ordering(model, 1)
        dim = self.domain.dim
        blk = np.repeat(np.arange(A), np.diff(off.astype(np.int64)))
        k = keys[blk].astype(np.int64)
        coords = np.stack([k // (dim[1] * dim[2]), (k // dim[2]) % dim[1], k % dim[2]])
        st[:3] += 4.0 * coords * np.float64(np.float32(self.config.dx))
        if self.models[model].material.kind == FLUID:
            st[3] += 1.0                     # J = 1 + q
        else:
            st[[3, 7, 11]] += 1.0            # F = I + E
        return st

    def state(self, model: int) -> np.ndarray:
        n, na = self.n(model), self.nattr(model)
        out = np.zeros((na, n), np.float32)
        self._check(self._lib.ckmpm_export_state(self._h, model, _ptr(out, _f32p)), "export_state")
        return out

    def set_state(self, model: int, soa: np.ndarray):
        s = _f32(soa)
        self._check(self._lib.ckmpm_import_state(self._h, model, _ptr(s, _f32p)), "import_state")

    def grid_m(self) -> np.ndarray:
        nb = self.layout(1)[1][1]
        out = np.zeros((nb, 4, 128), np.float32)
        self._check(self._lib.ckmpm_export_grid_m(self._h, _ptr(out, _f32p), nb), "export_grid_m")
        return out

    def grid_v(self) -> np.ndarray:
        nb = self.layout(1)[1][1]
        out = np.zeros((nb, 3, 128), np.float32)
        self._check(self._lib.ckmpm_export_grid_v(self._h, _ptr(out, _f32p), nb), "export_grid_v")
        return out


class MPMBatch:
    """Independent engines advanced together (ckmpm_batch_*).

    mode 0 runs the scenes one after another with plain launches (the sequential
    baseline), mode 1 interleaves them on their own streams, mode 2 replays a hipGraph
    of two substeps of every engine, mode 3 one hipGraph per engine, and mode 4 a
    hipGraph in which each kernel of a substep is one launch for all scenes.  The
    engines stay owned by the caller."""

    SEQUENTIAL, STREAMS, GRAPH, ENGINE_GRAPHS, FUSED = 0, 1, 2, 3, 4

    def __init__(self, engines: Sequence[MPMEngine]):
        self._lib = native.load()
        self.engines = list(engines)
        arr = (ctypes.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        self._h = ctypes.c_void_p()
        rc = self._lib.ckmpm_batch_create(arr, len(self.engines), ctypes.byref(self._h))
        if rc != native.OK:
            raise CkmpmError(rc, "ckmpm_batch_create", "")

    def _check(self, rc: int, where: str):
        if rc != native.OK:
            msg = self._lib.ckmpm_batch_last_error(self._h)
            raise CkmpmError(rc, where, msg.decode() if msg else "")

    def Run(self, substeps: int, mode: int = 4):
        self._check(self._lib.ckmpm_batch_run(self._h, int(substeps), int(mode)), "ckmpm_batch_run")

    def Synchronize(self):
        self._check(self._lib.ckmpm_batch_synchronize(self._h), "ckmpm_batch_synchronize")

    @property
    def captures(self) -> int:
        c = ctypes.c_uint64()
        self._check(self._lib.ckmpm_batch_captures(self._h, ctypes.byref(c)), "ckmpm_batch_captures")
        return int(c.value)

    def autotune(self, pairs: int = 3):
        """Times the fused mode's candidates (each material kind's G2P2G workgroup size,
        then side-by-side or serial G2P2G launches) and keeps the fastest; advances the
        engines 2 (pairs + 1) substeps per candidate."""
        self._check(self._lib.ckmpm_batch_autotune(self._h, int(pairs)), "ckmpm_batch_autotune")

    @property
    def serial_g2p2g(self) -> bool:
        v = ctypes.c_int32()
        self._check(self._lib.ckmpm_batch_get_serial(self._h, ctypes.byref(v)), "ckmpm_batch_get_serial")
        return bool(v.value)

    @serial_g2p2g.setter
    def serial_g2p2g(self, on: bool):
        self._check(self._lib.ckmpm_batch_set_serial(self._h, int(bool(on))), "ckmpm_batch_set_serial")

    @property
    def n_particles(self) -> int:
        return int(sum(e.n(m) for e in self.engines for m in range(len(e.models))))

    def close(self):
        if self._h:
            self._lib.ckmpm_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def engine_from_scene(scene: Scene, capacity: Optional[int] = None, adjoint_segment: int = 0,
                      fps: Optional[int] = None, cfl: Optional[float] = None) -> MPMEngine:
    """Build and set up an engine for a synthetic scene (adjoint_segment > 0: with an
    adjoint tape of that many substeps; fps and cfl: the frame loop's rate and CFL number
    instead of MPMConfig's defaults)."""
    eng = MPMEngine()
    cfg = MPMConfig(dx=scene.dx, gravity=scene.gravity, mass_clamp=scene.mass_clamp, dt=scene.dt,
                    max_active_block_count=capacity or scene.block_capacity())
    if fps is not None:
        cfg.fps = int(fps)
    if cfl is not None:
        cfg.cfl = float(cfl)
    eng.Initialize(MPMDomain(scene.dim, scene.offset), cfg)
    models = []
    for m in scene.models:
        q = m.material
        mat = MPMMaterial(q.kind, q.lam, q.mu, q.bulk, q.gamma, q.viscosity, q.J0, q.friction_deg, q.yield_stress,
                          q.alpha0, q.beta, q.hardening_factor, q.m, q.hardening_on)
        models.append(MPMModel(m.mass, m.volume, m.x, m.v, mat, m.lam_p, m.mu_p))
    eng.InitializeParticle(models)
    for i, m in enumerate(scene.models):
        if m.rigid_count:
            eng.set_rigid(i, m.rigid_count, m.rigid_velocity)
    if scene.boundary is not None:
        eng.set_boundary(scene.boundary)
    if adjoint_segment:
        eng.adjoint_enable(adjoint_segment)
    eng.InitialSetup()
    return eng
