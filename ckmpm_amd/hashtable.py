"""Sparse block table over the C-ABI (ckmpm_hash_*).

Mirrors BucketedCuckooHashTable (include/gpu_hash_table.cuh:274-552): Allocate
(:343-354) -> the constructor, CooperativeInsert (:378-461) -> Insert (a whole key set
at once, deterministic placement), CooperativeFind (:463-513) -> Find.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import native
from .native import CkmpmError, _u32p

EMPTY = 0xFFFFFFFF


class BlockHashTable:
    def __init__(self, dim: Sequence[int], max_keys: int):
        self._lib = native.load()
        d = native.DomainDesc()
        for a in range(3):
            d.dim[a] = int(dim[a])
            d.offset[a] = 0
        h = ctypes.c_void_p()
        rc = self._lib.ckmpm_hash_create(ctypes.byref(d), int(max_keys), ctypes.byref(h))
        if rc != native.OK:
            raise CkmpmError(rc, "ckmpm_hash_create", "")
        self._h = h
        self.dim = tuple(int(v) for v in dim)
        self.max_keys = int(max_keys)

    def _check(self, rc: int, where: str):
        if rc != native.OK:
            msg = self._lib.ckmpm_hash_last_error(self._h)
            raise CkmpmError(rc, where, msg.decode() if msg else "")

    def Insert(self, keys, values=None) -> None:
        k = np.ascontiguousarray(keys, np.uint32)
        v = np.arange(k.shape[0], dtype=np.uint32) if values is None else np.ascontiguousarray(values, np.uint32)
        self._check(self._lib.ckmpm_hash_build(self._h, k.ctypes.data_as(_u32p), v.ctypes.data_as(_u32p),
                                               k.shape[0]), "ckmpm_hash_build")

    def Find(self, keys) -> np.ndarray:
        q = np.ascontiguousarray(keys, np.uint32)
        out = np.empty(q.shape[0], np.uint32)
        self._check(self._lib.ckmpm_hash_find(self._h, q.ctypes.data_as(_u32p), q.shape[0],
                                              out.ctypes.data_as(_u32p)), "ckmpm_hash_find")
        return out

    def InsertDevice(self, keys_ptr: int, values_ptr: int, n: int) -> None:
        self._check(self._lib.ckmpm_hash_build_device(self._h, ctypes.c_void_p(keys_ptr), ctypes.c_void_p(values_ptr),
                                                      int(n)), "ckmpm_hash_build_device")

    # lookup forms of FindDevice
    SCALAR, COOPERATIVE, LDS = 0, 1, 2

    def FindDevice(self, keys_ptr: int, n: int, out_ptr: int, cooperative=True) -> None:
        """Enqueue a lookup of n device keys; call Synchronize before reading out_ptr.
        cooperative: True / COOPERATIVE (a 16-lane group per probe), False / SCALAR (a
        thread per key) or LDS (the table's keys staged in LDS; tables of at most 2560
        buckets)."""
        self._check(self._lib.ckmpm_hash_find_device(self._h, ctypes.c_void_p(keys_ptr), int(n),
                                                     ctypes.c_void_p(out_ptr), int(cooperative)),
                    "ckmpm_hash_find_device")

    def Synchronize(self) -> None:
        self._check(self._lib.ckmpm_hash_synchronize(self._h), "ckmpm_hash_synchronize")

    def info(self) -> native.HashInfo:
        inf = native.HashInfo()
        self._check(self._lib.ckmpm_hash_info(self._h, ctypes.byref(inf)), "ckmpm_hash_info")
        return inf

    def export(self):
        inf = self.info()
        slots = np.empty(inf.buckets * inf.slots_per_bucket, np.uint64)
        stash = np.empty(max(inf.stash_capacity, 1), np.uint64)
        ns = ctypes.c_uint32()
        u64 = ctypes.POINTER(ctypes.c_uint64)
        self._check(self._lib.ckmpm_hash_export(self._h, slots.ctypes.data_as(u64), slots.shape[0],
                                                stash.ctypes.data_as(u64), stash.shape[0], ctypes.byref(ns)),
                    "ckmpm_hash_export")
        return slots, stash[:ns.value].copy()

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ckmpm_hash_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NOALLOC_MAX_KEYS = 1 << 18   # CKMPM_HASH_NOALLOC_MAX_KEYS


def table_sizes(max_keys: int):
    """(slot words, stash words) of a table created for max_keys pairs."""
    buckets = (int(max_keys) * 10 + 7 * 16 - 1) // (7 * 16)
    return buckets * 16, 256 + int(max_keys) // 64


def build_noalloc(dim: Sequence[int], max_keys: int, keys_ptr: int, values_ptr: int, n: int, slots_ptr: int,
                  stash_ptr: int, nstash_ptr: int, stream: int = 0) -> int:
    """ckmpm_hash_build_noalloc on device arrays (the table is the caller's memory, sized by
    table_sizes); returns the status code."""
    lib = native.load()
    d = native.DomainDesc()
    for a in range(3):
        d.dim[a] = int(dim[a])
        d.offset[a] = 0
    vp = ctypes.c_void_p
    return lib.ckmpm_hash_build_noalloc(ctypes.byref(d), int(max_keys), vp(keys_ptr), vp(values_ptr), int(n),
                                        vp(slots_ptr), vp(stash_ptr), vp(nstash_ptr), vp(stream))
