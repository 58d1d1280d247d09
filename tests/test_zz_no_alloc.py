"""The hash build that allocates nothing (ckmpm_hash_build_noalloc) against the build of
a created table (ckmpm_hash_create + ckmpm_hash_build_device) on the same device arrays:
slots, stash and stash count equal word for word, on two calls in a row, the second
with a smaller input (a table that needs its stash); sizes beyond the scratch are
refused before anything runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_keys(dim, n, seed):
    rng = np.random.default_rng(seed)
    return rng.choice(int(np.prod(dim)), n, replace=False).astype(np.uint32)


def _compare(dim, max_keys, keys, vals):
    import torch
    from ckmpm_amd import native
    from ckmpm_amd.hashtable import BlockHashTable, build_noalloc, table_sizes

    n = keys.shape[0]
    dk = torch.from_numpy(keys.view(np.int32)).cuda()
    dv = torch.from_numpy(vals.view(np.int32)).cuda()
    ref = BlockHashTable(dim, max_keys)
    ref.InsertDevice(dk.data_ptr(), dv.data_ptr(), n)
    slots_ref, stash_ref = ref.export()
    ref.close()
    nslots, nstash_cap = table_sizes(max_keys)
    assert nslots == slots_ref.shape[0]
    # the caller's table starts as garbage: the build must write all of it
    slots = torch.full((nslots,), 0x5A5A5A5A5A5A5A5A, dtype=torch.int64, device="cuda")
    stash = torch.full((nstash_cap,), -7, dtype=torch.int64, device="cuda")
    nstash = torch.full((1,), 12345, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = build_noalloc(dim, max_keys, dk.data_ptr(), dv.data_ptr(), n, slots.data_ptr(), stash.data_ptr(),
                       nstash.data_ptr())
    assert rc == native.OK, rc
    ns = int(nstash.cpu()[0])
    np.testing.assert_array_equal(slots.cpu().numpy().view(np.uint64), slots_ref)
    assert ns == stash_ref.shape[0]
    np.testing.assert_array_equal(stash.cpu().numpy().view(np.uint64)[:ns], stash_ref)
    return ns


def test_noalloc_build_equals_created_table_twice():
    from oracle.pyoracle import HashOracle
    from tests.test_hash import _stash_keys

    dim = (128, 128, 128)
    keys = _random_keys(dim, 25000, seed=21)
    vals = np.arange(keys.shape[0], dtype=np.uint32)[::-1].copy()
    print("noalloc build 1: 25000 keys, max_keys 40000", flush=True)
    _compare(dim, 40000, keys, vals)
    # second call, smaller: 24 keys chosen to overflow their buckets into the stash
    dim2 = (16, 16, 16)
    o = HashOracle(dim2, 40)
    keys2 = _stash_keys(dim2, o.buckets, o.seeds, 24)
    vals2 = np.arange(24, dtype=np.uint32) + 100
    print("noalloc build 2: 24 keys, max_keys 40 (stash)", flush=True)
    assert _compare(dim2, 40, keys2, vals2) == 8


def test_noalloc_build_refuses_sizes_beyond_its_scratch():
    import torch
    from ckmpm_amd import native
    from ckmpm_amd.hashtable import NOALLOC_MAX_KEYS, build_noalloc

    k = torch.zeros(4, dtype=torch.int32, device="cuda")
    t = torch.zeros(16, dtype=torch.int64, device="cuda")
    # more than the scratch holds, and more keys than the table is sized for: refused
    # on the host before any kernel runs
    assert build_noalloc((64, 64, 64), NOALLOC_MAX_KEYS + 1, k.data_ptr(), k.data_ptr(), 4, t.data_ptr(),
                         t.data_ptr(), k.data_ptr()) == native.ERR_CAPACITY
    assert build_noalloc((64, 64, 64), 2, k.data_ptr(), k.data_ptr(), 4, t.data_ptr(), t.data_ptr(),
                         k.data_ptr()) == native.ERR_CAPACITY
