"""Writes tests/golden/tiny_mixed.npz: FP64-oracle vectors for a tiny two-material scene.

The reference ships no vectors for this path (its tests only write frames for visual
inspection), so these fixtures come from the FP64 restatement in oracle/, after it has
passed its own pins (tests/test_oracle_*.py).  They freeze the canonical orderings and
the FP64 trajectory so that a change to either is caught without re-deriving it, and
let the GPU parity test compare against stored vectors.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from ckmpm_amd import scenes  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiny_mixed.npz")
STEPS = 6


def tiny_scene() -> scenes.Scene:
    dx = 1.0 / 64
    rng = scenes.Splitmix(11)
    # fixed-corotated ball moving towards a resting fluid box
    c = np.array([0.40, 0.50, 0.50])
    g = np.stack(np.meshgrid(*[np.arange(18, 34)] * 3, indexing="ij"), -1).reshape(-1, 3)
    xb = scenes._stratified_cells(rng, g.astype(np.float64), dx)
    xb = xb[np.linalg.norm(xb - c, axis=1) <= 0.06].astype(np.float32)
    lam, mu = scenes.lame(2e4, 0.3)
    vol = dx ** 3 / 8
    ball = scenes.ModelSpec(scenes.Material(scenes.FIXED_COROTATED, lam, mu), 1000 * vol, vol, xb,
                            np.array([[2.0, 0.3, -0.1]], np.float32))
    g = np.stack(np.meshgrid(np.arange(36, 42), np.arange(26, 38), np.arange(28, 36), indexing="ij"), -1).reshape(-1, 3)
    xf = scenes._stratified_cells(rng, g.astype(np.float64), dx).astype(np.float32)
    fluid = scenes.ModelSpec(scenes.Material(scenes.FLUID, bulk=4e4, gamma=7.0, viscosity=0.1), 1000 * vol, vol, xf,
                             np.zeros((1, 3), np.float32))
    return scenes.Scene("tiny_mixed", (16, 16, 16), (0, 0, 0), dx, -9.8, 5e-5, STEPS, [ball, fluid])


def run(scene):
    o = Oracle.from_scene(scene)
    o.setup()
    out = {"setup_keys": o.keys(0), "setup_counts": np.array(o.counts(0))}
    for m in range(len(scene.models)):
        out[f"setup_order_{m}"] = o.setup_order(m)
    for k in range(STEPS):
        o.substep(0.0 if k == 0 else scene.dt, scene.dt)
        out[f"keys_{k}"] = o.keys(0)
        out[f"counts_{k}"] = np.array(o.counts(0))
        for m in range(len(scene.models)):
            out[f"perm_{k}_{m}"] = o.perm(m)
            out[f"off_{k}_{m}"] = o.offsets(m)
    for m in range(len(scene.models)):
        out[f"state_{m}"] = o.state(m)
    out["grid_m"] = o.grid_m()
    out["prev_keys"] = o.keys(1)
    out["flags"] = np.array([o.flags])
    return out


if __name__ == "__main__":
    vec = run(tiny_scene())
    np.savez_compressed(OUT, **vec)
    print(f"wrote {OUT}: {sum(v.size for v in vec.values())} values")
