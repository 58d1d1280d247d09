"""ctypes wrapper of the FP64 host restatement (oracle/ckmpm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product path never imports it.

How each part of the restatement is pinned (the reference ships no golden vectors,
KATs or fixtures for this path, so the pins are mathematical):

* FP32 binning / block keys (mpm_algorithm.cuh:59-67,595-600): exhaustive checks at the
  +-0.25 cell boundaries against exact rational arithmetic, and brute force over random
  positions against numpy float32 (tests/test_oracle_kernel.py).
* Weights (mpm_kernel.cuh:91-114): closed form; partition of unity per grid; the two
  grids' first moments cancel; D closed form per axis and off-diagonal 2 s_a s_b dx^2.
* SVD (mpm_math.cuh convention): reconstruction, orthogonality, det U = det V = +1,
  ordering s0 >= s1 >= |s2|, agreement of |s| with numpy.linalg.svd, inverted and
  degenerate inputs.
* Stresses (mpm_material.cuh:318-477 + Neo-Hookean): FC against 2 mu (F - R) F^T +
  lambda J (J - 1) I with R from scipy's polar decomposition; NH closed form; DP: elastic
  branch equals the Hencky closed form, the projection lands on the yield surface and
  the tension branch returns F = U V^T, tau = 0; fluid closed form.
* Layout extension (mpm_algorithm.cuh:657-695): brute-force set construction.
* P2G / G2P / grid update (mpm_algorithm.cuh:738-869,1365-1845): total grid mass equals
  2 sum m_p, total momentum 2 sum m_p v_p, a uniform velocity field gives C = 0 and
  x += dt v exactly, free fall matches symplectic Euler, the colliding spheres keep
  mirror symmetry and zero net x-momentum, and a sign error in the stress term is
  caught by the energy/force direction test.
* Rebin (canonical order): brute-force stable sort and permutation invariants.
* Conservation metric (mpm_algorithm.cuh:1083-1357): on an imposed affine grid velocity
  field a + w x node, G2P is exact (partition of unity, cancelling first moments), so
  the sums equal N a + w x sum x and sum x x (a + w x x) in closed form; with a = 0 and
  w = 0 the metric is zero, and a wrong sign, a dropped grid or a missing 1/2 fails
  the closed form (tests/test_conservation.py).
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libckmpm_oracle.so")
_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [_i32p, _i32p, ctypes.c_float, ctypes.c_double, ctypes.c_double, ctypes.c_uint32]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_add_model.restype = ctypes.c_int
        L.orc_add_model.argtypes = [ctypes.c_void_p, ctypes.c_int, _f64p, ctypes.c_double, ctypes.c_double, _f32p,
                                    ctypes.c_uint32, _f32p, ctypes.c_uint32, _f32p, _f32p]
        for name in ("orc_initial_setup", "orc_flags"):
            getattr(L, name).restype = ctypes.c_uint32
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_clear_flags.argtypes = [ctypes.c_void_p]
        L.orc_set_rigid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, _f64p]
        L.orc_set_rigid.restype = None
        L.orc_grid_update.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_grid_update.restype = ctypes.c_uint32
        L.orc_particles.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_particles.restype = ctypes.c_uint32
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = None
        L.orc_max_threads.argtypes = []
        L.orc_max_threads.restype = ctypes.c_int
        for name in ("orc_physics", "orc_substep"):
            getattr(L, name).restype = ctypes.c_uint32
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_rebin.restype = ctypes.c_uint32
        L.orc_rebin.argtypes = [ctypes.c_void_p, ctypes.POINTER(_f32p)]
        L.orc_max_v2.restype = ctypes.c_double
        L.orc_max_v2.argtypes = [ctypes.c_void_p]
        L.orc_layout_counts.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p]
        L.orc_layout_keys.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p]
        L.orc_model_size.restype = ctypes.c_uint32
        L.orc_model_size.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_model_attrs.restype = ctypes.c_int
        L.orc_model_attrs.argtypes = [ctypes.c_void_p, ctypes.c_int]
        for name in ("orc_get_perm", "orc_get_offsets"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p]
        for name in ("orc_get_state", "orc_get_state_next"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int, _f64p]
        for name in ("orc_get_grid_m", "orc_get_grid_mnext", "orc_get_grid_v"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _f64p]
        L.orc_set_layout.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_set_ordering.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p, _u32p]
        L.orc_set_state_f32.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p]
        L.orc_set_state_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, _f64p]
        L.orc_get_setup_order.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p]
        L.orc_set_grid_m_f32.argtypes = [ctypes.c_void_p, _f32p]
        L.orc_conservation.argtypes = [ctypes.c_void_p, ctypes.c_int, _f64p]
        L.orc_set_grid_v.argtypes = [ctypes.c_void_p, _f64p]
        L.orc_set_boundary.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _i32p, _i32p,
                                       _f64p]
        L.orc_svd.argtypes = [_f64p, _f64p, _f64p, _f64p]
        L.orc_stress.argtypes = [ctypes.c_int, _f64p, _f64p, ctypes.c_double, _f64p, ctypes.c_double, _f64p, _f64p, _f64p]
        L.orc_weight0.restype = ctypes.c_double
        L.orc_weight0.argtypes = [ctypes.c_double]
        L.orc_stencil.argtypes = [ctypes.c_void_p, _f64p, ctypes.c_int, _i32p, _f64p, _f64p]
        L.orc_bin_f32.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32, _i32p, _i32p]
        L.orc_build_layout.restype = ctypes.c_uint32
        L.orc_build_layout.argtypes = [ctypes.c_void_p, _u32p, ctypes.c_uint32, _u32p, _u32p]
        _u8p = ctypes.POINTER(ctypes.c_uint8)
        _u64p = ctypes.POINTER(ctypes.c_uint64)
        L.orc_murmur3_32.restype = ctypes.c_uint32
        L.orc_murmur3_32.argtypes = [_u8p, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_hash_seeds.argtypes = [_u32p]
        L.orc_hash_buckets.restype = ctypes.c_uint32
        L.orc_hash_buckets.argtypes = [ctypes.c_uint32]
        L.orc_hash_build.restype = ctypes.c_uint32
        L.orc_hash_build.argtypes = [_i32p, ctypes.c_uint32, _u32p, _u32p, _u32p, ctypes.c_uint32, _u64p, _u64p,
                                     ctypes.c_uint32]
        L.orc_hash_find.argtypes = [_i32p, ctypes.c_uint32, _u32p, _u64p, _u64p, ctypes.c_uint32, _u32p,
                                    ctypes.c_uint32, _u32p]
        _lib = L
    return _lib


# ----------------------------------------------------------------- unit helpers
def svd(F: np.ndarray):
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    U, s, V = np.zeros(9), np.zeros(3), np.zeros(9)
    lib().orc_svd(_p(F, _f64p), _p(U, _f64p), _p(s, _f64p), _p(V, _f64p))
    return U.reshape(3, 3), s, V.reshape(3, 3)


def _params13(params) -> np.ndarray:
    """lam, mu, bulk, gamma, viscosity, J0, friction (deg), yield, alpha0, beta,
    hardening factor, M, hardening on; missing trailing entries: 0 (hardening on: 1)."""
    pr = np.array([0, 0, 0, 0, 0, 1, 30, 0, 0, 0, 0, 0, 1], np.float64)
    p = np.asarray(params, np.float64)
    pr[: p.shape[0]] = p
    return pr


def stress(kind: int, params, F, logJp=0.0, C=None, J=1.0):
    pr = _params13(params)
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    C = np.zeros(9) if C is None else np.ascontiguousarray(C, np.float64).reshape(9)
    Fo, lj, tau = np.zeros(9), np.zeros(1), np.zeros(9)
    lib().orc_stress(kind, _p(pr, _f64p), _p(F, _f64p), float(logJp), _p(C, _f64p), float(J), _p(Fo, _f64p),
                     _p(lj, _f64p), _p(tau, _f64p))
    return Fo.reshape(3, 3), float(lj[0]), tau.reshape(3, 3)


def weight0(t: float) -> float:
    return lib().orc_weight0(float(t))


class Oracle:
    """One FP64 engine instance (mirrors the HIP engine's state layout)."""

    def __init__(self, dim, dx: float, gravity: float = 0.0, mass_clamp: float = 0.0, capacity: int = 1 << 30,
                 offset=(0, 0, 0)):
        d = np.ascontiguousarray(dim, np.int32)
        o = np.ascontiguousarray(offset, np.int32)
        self.h = lib().orc_create(_p(d, _i32p), _p(o, _i32p), float(dx), float(gravity), float(mass_clamp), int(capacity))
        self.nmodels = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @classmethod
    def from_scene(cls, scene) -> "Oracle":
        o = cls(scene.dim, scene.dx, scene.gravity, scene.mass_clamp, scene.block_capacity(), scene.offset)
        for i, m in enumerate(scene.models):
            o.add_model(m.material.kind, m.material.params(), m.mass, m.volume, m.x, m.v, m.lam_p, m.mu_p)
            if getattr(m, "rigid_count", 0):
                o.set_rigid(i, m.rigid_count, m.rigid_velocity)
        if getattr(scene, "boundary", None) is not None:
            o.set_boundary(scene.boundary)
        return o

    def set_boundary(self, b):
        """Boundary with .walls and .regions (each .lo, .hi, .kind, .value, .center)."""
        n = len(b.regions)
        box = np.zeros(max(6 * n, 1), np.int32)
        kinds = np.zeros(max(n, 1), np.int32)
        vals = np.zeros(max(6 * n, 1), np.float64)
        for i, r in enumerate(b.regions):
            box[6 * i:6 * i + 6] = list(r.lo) + list(r.hi)
            kinds[i] = r.kind
            # the engine holds value and center in FP32: use the same numbers
            vals[6 * i:6 * i + 6] = np.float32(list(r.value) + list(r.center))
        lib().orc_set_boundary(self.h, 1, int(b.walls), n, _p(box, _i32p), _p(kinds, _i32p), _p(vals, _f64p))

    def add_model(self, kind, params, mass, volume, x, v, lam_p=None, mu_p=None) -> int:
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        pr = _params13(params)
        lp = None if lam_p is None else np.ascontiguousarray(lam_p, np.float32)
        mp = None if mu_p is None else np.ascontiguousarray(mu_p, np.float32)
        idx = lib().orc_add_model(self.h, int(kind), _p(pr, _f64p), float(mass), float(volume), _p(x, _f32p),
                                  x.shape[0], _p(v, _f32p), v.shape[0], None if lp is None else _p(lp, _f32p),
                                  None if mp is None else _p(mp, _f32p))
        self.nmodels += 1
        return idx

    def set_rigid(self, m: int, count: int, velocity) -> None:
        v = np.ascontiguousarray(velocity, np.float64).reshape(3)
        lib().orc_set_rigid(self.h, int(m), int(count), _p(v, _f64p))

    def setup(self) -> int:
        return lib().orc_initial_setup(self.h)

    def physics(self, dt, next_dt) -> int:
        return lib().orc_physics(self.h, float(dt), float(next_dt))

    def rebin(self, positions: Optional[List[np.ndarray]] = None) -> int:
        if positions is None:
            return lib().orc_rebin(self.h, None)
        keep = [np.ascontiguousarray(p, np.float32) for p in positions]
        arr = (_f32p * len(keep))(*[_p(k, _f32p) for k in keep])
        return lib().orc_rebin(self.h, arr)

    def substep(self, dt, next_dt) -> int:
        return lib().orc_substep(self.h, float(dt), float(next_dt))

    def run_frames(self, frames: int, fps: float, cfl: float, dt_default: float, dx: float, state=None):
        """The reference's frame loop (mpm_engine.cuh:514-876): per frame of 1/fps, substeps
        until the frame's end; each substep's grid update runs with the current dt, then
        next_dt = min(dt_default, time left in the frame, dx cfl / max|v|) and the particles
        move with (dt, next_dt); the first substep of a run has dt = 0 (:115,178,552-557).
        `state` = (time, dt) carried between calls; returns (state, dt of each substep)."""
        t, cur = state if state is not None else (0.0, 0.0)
        dts = []
        for _ in range(frames):
            frame_end = t + 1.0 / fps
            while t < frame_end:
                lib().orc_grid_update(self.h, cur)
                vmax = math.sqrt(self.max_v2)
                if math.isinf(vmax):
                    raise FloatingPointError("velocity reached infinity")
                nxt = min(dt_default, frame_end - t)
                if vmax > 0.0:
                    nxt = min(nxt, dx * cfl / vmax)
                lib().orc_particles(self.h, cur, nxt)
                self.rebin()
                dts.append(cur)
                cur = nxt
                t += cur
        return (t, cur), dts

    @property
    def flags(self) -> int:
        return lib().orc_flags(self.h)

    def clear_flags(self):
        lib().orc_clear_flags(self.h)

    @property
    def max_v2(self) -> float:
        return lib().orc_max_v2(self.h)

    def counts(self, which: int = 0):
        c = np.zeros(3, np.uint32)
        lib().orc_layout_counts(self.h, which, _p(c, _u32p))
        return tuple(int(v) for v in c)

    def keys(self, which: int = 0) -> np.ndarray:
        A, Nb, Ex = self.counts(which)
        k = np.zeros(Ex, np.uint32)
        lib().orc_layout_keys(self.h, which, _p(k, _u32p))
        return k

    def n(self, m: int) -> int:
        return lib().orc_model_size(self.h, m)

    def nattr(self, m: int) -> int:
        return lib().orc_model_attrs(self.h, m)

    def perm(self, m: int) -> np.ndarray:
        p = np.zeros(self.n(m), np.uint32)
        lib().orc_get_perm(self.h, m, _p(p, _u32p))
        return p

    def offsets(self, m: int) -> np.ndarray:
        o = np.zeros(self.counts()[0] + 1, np.uint32)
        lib().orc_get_offsets(self.h, m, _p(o, _u32p))
        return o

    def state(self, m: int) -> np.ndarray:
        s = np.zeros(self.nattr(m) * self.n(m))
        lib().orc_get_state(self.h, m, _p(s, _f64p))
        return s.reshape(self.nattr(m), self.n(m))

    def state_next(self, m: int) -> np.ndarray:
        s = np.zeros(self.nattr(m) * self.n(m))
        lib().orc_get_state_next(self.h, m, _p(s, _f64p))
        return s.reshape(self.nattr(m), self.n(m))

    def grid_m(self) -> np.ndarray:
        nb = self.counts(1)[1]
        g = np.zeros(nb * 512)
        lib().orc_get_grid_m(self.h, _p(g, _f64p))
        return g.reshape(nb, 4, 128)

    def grid_mnext(self) -> np.ndarray:
        nb = self.counts(0)[1]
        g = np.zeros(nb * 512)
        lib().orc_get_grid_mnext(self.h, _p(g, _f64p))
        return g.reshape(nb, 4, 128)

    def grid_v(self) -> np.ndarray:
        nb = self.counts(0)[1]
        g = np.zeros(nb * 384)
        lib().orc_get_grid_v(self.h, _p(g, _f64p))
        return g.reshape(nb, 3, 128)

    def set_grid_v(self, V: np.ndarray):
        nb = self.counts(0)[1]
        g = np.ascontiguousarray(V, np.float64).reshape(-1)
        assert g.shape[0] == nb * 384
        lib().orc_set_grid_v(self.h, _p(g, _f64p))

    def conservation(self, m: int) -> np.ndarray:
        """Lagrangian linear and angular momentum per unit mass of model m on the current
        grid velocities (after physics(), before rebin()): 6 values."""
        out = np.zeros(6)
        lib().orc_conservation(self.h, m, _p(out, _f64p))
        return out

    # lockstep import
    def set_layout(self, which, keys, A, Nb, Ex):
        k = np.ascontiguousarray(keys, np.uint32)
        lib().orc_set_layout(self.h, which, _p(k, _u32p), A, Nb, Ex)

    def set_ordering(self, m, perm, off):
        p = np.ascontiguousarray(perm, np.uint32)
        o = np.ascontiguousarray(off, np.uint32)
        lib().orc_set_ordering(self.h, m, _p(p, _u32p), _p(o, _u32p))

    def set_state(self, m, soa):
        """State SoA in storage order with global positions (FP64)."""
        s = np.ascontiguousarray(soa, np.float64)
        lib().orc_set_state_f64(self.h, m, _p(s, _f64p))

    def setup_order(self, m) -> np.ndarray:
        o = np.zeros(self.n(m), np.uint32)
        lib().orc_get_setup_order(self.h, m, _p(o, _u32p))
        return o

    def set_grid_m(self, M_f32):
        g = np.ascontiguousarray(M_f32, np.float32)
        lib().orc_set_grid_m_f32(self.h, _p(g, _f32p))

    def bin_f32(self, x: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        n = x.shape[0]
        b = np.zeros((n, 3), np.int32)
        c = np.zeros((n, 3), np.int32)
        lib().orc_bin_f32(self.h, _p(x, _f32p), n, _p(b, _i32p), _p(c, _i32p))
        return b, c

    def stencil(self, x, g: int):
        x = np.ascontiguousarray(x, np.float64)
        c = np.zeros(3, np.int32)
        t = np.zeros(3)
        w = np.zeros(6)
        lib().orc_stencil(self.h, _p(x, _f64p), g, _p(c, _i32p), _p(t, _f64p), _p(w, _f64p))
        return c, t, w.reshape(3, 2)

    def build_layout(self, akeys) -> tuple:
        a = np.ascontiguousarray(akeys, np.uint32)
        out = np.zeros(int(a.shape[0]) * 27 + 27, np.uint32)
        cnt = np.zeros(3, np.uint32)
        ex = lib().orc_build_layout(self.h, _p(a, _u32p), a.shape[0], _p(out, _u32p), _p(cnt, _u32p))
        return out[:ex].copy(), tuple(int(v) for v in cnt)


# ----------------------------------------------------------------- sparse block table
def murmur3_32(data: bytes, seed: int) -> int:
    """MurmurHash3_x86_32 of a byte string."""
    b = np.frombuffer(bytes(data), np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
    return int(lib().orc_murmur3_32(_p(b, ctypes.POINTER(ctypes.c_uint8)), len(data), seed & 0xFFFFFFFF))


def hash_seeds() -> np.ndarray:
    out = np.zeros(3, np.uint32)
    lib().orc_hash_seeds(_p(out, _u32p))
    return out


class HashOracle:
    """Deterministic three-choice bucketed hash (the oracle side of ckmpm_hash_*)."""

    def __init__(self, dim, max_keys: int):
        self.dim = np.ascontiguousarray(dim, np.int32)
        self.buckets = int(lib().orc_hash_buckets(max_keys))
        self.seeds = hash_seeds()
        self.stash_cap = 256 + max_keys // 64
        self.slots = np.full(self.buckets * 16, np.iinfo(np.uint64).max, np.uint64)
        self.stash = np.zeros(self.stash_cap, np.uint64)
        self.nstash = 0

    def build(self, keys, values) -> int:
        k = np.ascontiguousarray(keys, np.uint32)
        v = np.ascontiguousarray(values, np.uint32)
        u64 = ctypes.POINTER(ctypes.c_uint64)
        ns = lib().orc_hash_build(_p(self.dim, _i32p), self.buckets, _p(self.seeds, _u32p), _p(k, _u32p),
                                  _p(v, _u32p), k.shape[0], _p(self.slots, u64), _p(self.stash, u64), self.stash_cap)
        self.nstash = int(ns)
        return self.nstash

    def find(self, keys) -> np.ndarray:
        q = np.ascontiguousarray(keys, np.uint32)
        out = np.zeros(q.shape[0], np.uint32)
        u64 = ctypes.POINTER(ctypes.c_uint64)
        lib().orc_hash_find(_p(self.dim, _i32p), self.buckets, _p(self.seeds, _u32p), _p(self.slots, u64),
                            _p(self.stash, u64), min(self.nstash, self.stash_cap), _p(q, _u32p), q.shape[0],
                            _p(out, _u32p))
        return out
