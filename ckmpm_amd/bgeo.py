"""Per-frame particle export in Houdini's classic binary geometry format (.bgeo).

The reference writes every model's positions after each frame as
model_<i>_particle_frame_<f>.bgeo, one "position" attribute, through partio
(include/mpm_engine.cuh:446-452, include/mpm_io.h:93-131; partio is a git submodule
of the reference and is not vendored).  partio stores the "position" attribute as the
points' own P, so such a file holds, all numbers big-endian:

  magic 'Bgeo' (int32), 'V' (1 byte), version 5 (int32),
  counts (int32 each): points, primitives, point groups, primitive groups, point
  attributes, vertex attributes, primitive attributes, detail attributes
  (all zero but the points),
  per point x y z w with w = 1 (float32),
  0x00 0xff (the empty extra section).

`ckmpm::WriteBgeoPositions` (include/ckmpm.hpp) writes the same bytes from C++.
"""
from __future__ import annotations

import os

import numpy as np

MAGIC = 0x4267656F  # 'Bgeo'
VERSION = 5
HEADER_BYTES = 4 + 1 + 4 + 8 * 4
TRAILER = b"\x00\xff"


def encode_positions(xyz) -> bytes:
    """The .bgeo bytes of an (n, 3) array of positions."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    n = xyz.shape[0]
    if n > 0x7FFFFFFF:
        raise ValueError("too many points for a bgeo file")
    head = np.array([MAGIC], ">u4").tobytes() + b"V" + np.array([VERSION, n, 0, 0, 0, 0, 0, 0, 0], ">u4").tobytes()
    pts = np.empty((n, 4), ">f4")
    pts[:, :3] = xyz
    pts[:, 3] = 1.0
    return head + pts.tobytes() + TRAILER


def write_positions(path: str | os.PathLike, xyz) -> None:
    with open(path, "wb") as f:
        f.write(encode_positions(xyz))


def decode_positions(data: bytes) -> np.ndarray:
    """Positions (n, 3) float32 of a points-only .bgeo file as `encode_positions` writes."""
    if len(data) < HEADER_BYTES + len(TRAILER):
        raise ValueError("not a bgeo file: too short")
    magic = int(np.frombuffer(data, ">u4", 1, 0)[0])
    if magic != MAGIC or data[4:5] != b"V":
        raise ValueError("not a bgeo file: bad magic")
    version, n, *rest = (int(v) for v in np.frombuffer(data, ">u4", 9, 5))
    if version != VERSION:
        raise ValueError(f"bgeo version {version}, expected {VERSION}")
    if any(rest):
        raise ValueError("only point positions are supported (primitives, groups or attributes present)")
    if len(data) != HEADER_BYTES + 16 * n + len(TRAILER) or data[-2:] != TRAILER:
        raise ValueError("bgeo file size does not match its point count")
    pts = np.frombuffer(data, ">f4", 4 * n, HEADER_BYTES).reshape(n, 4)
    return pts[:, :3].astype(np.float32)


def read_positions(path: str | os.PathLike) -> np.ndarray:
    with open(path, "rb") as f:
        return decode_positions(f.read())
