"""Sparse block table (ckmpm_hash_*, oracle HashOracle).

Pinned without the reference (its table is never exercised by its own tests):
* the hash is MurmurHash3_x86_32: the oracle matches the algorithm's published test
  vectors and scikit-learn's independent implementation on int3 keys;
* the seeds are std::mt19937(2)'s first three outputs: numpy's MT19937 with the same
  init_genrand seeding gives the same words;
* the build is checked against an independent numpy restatement (bit-exact slots),
  against a dict (Find of every present key; absent keys miss), for order invariance,
  and on an adversarial key set that forces the stash.
GPU: the HIP build reproduces the oracle's slots and stash bit for bit; both Find
kernels agree with the oracle.
"""
import numpy as np
import pytest
from sklearn.utils import murmurhash3_32 as sk_murmur

from oracle.pyoracle import HashOracle, hash_seeds, murmur3_32

EMPTY = 0xFFFFFFFF
U64_EMPTY = np.uint64(0xFFFFFFFFFFFFFFFF)

# MurmurHash3_x86_32 test vectors (key bytes, seed, hash) as published with the algorithm
KAT = [
    (b"", 0, 0x00000000), (b"", 1, 0x514E28B7), (b"", 0xFFFFFFFF, 0x81F16F39),
    (b"\x00\x00\x00\x00", 0, 0x2362F9DE), (b"\xff\xff\xff\xff", 0, 0x76293B50),
    (b"\x21\x43\x65\x87", 0, 0xF55B516B), (b"\x21\x43\x65\x87", 0x5082EDEE, 0x2362F9DE),
    (b"\x21\x43\x65", 0, 0x7E4A8634), (b"\x21\x43", 0, 0xA0F7B07A), (b"\x21", 0, 0x72661CF4),
    (b"\x00\x00\x00", 0, 0x85F0B427), (b"\x00\x00", 0, 0x30F4C306), (b"\x00", 0, 0x514E28B7),
    (b"aaaa", 0x9747B28C, 0x5A97808A), (b"aaa", 0x9747B28C, 0x283E0130), (b"aa", 0x9747B28C, 0x5D211726),
    (b"a", 0x9747B28C, 0x7FA09EA6), (b"abcd", 0x9747B28C, 0xF0478627), (b"abc", 0x9747B28C, 0xC84A62DD),
    (b"ab", 0x9747B28C, 0x74875592), (b"Hello, world!", 0x9747B28C, 0x24884CBA),
    (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD),
]


@pytest.mark.parametrize("data,seed,expected", KAT)
def test_murmur3_known_answers(data, seed, expected):
    assert murmur3_32(data, seed) == expected


def test_murmur3_int3_matches_sklearn():
    rng = np.random.default_rng(0)
    for _ in range(200):
        c = rng.integers(-2 ** 31, 2 ** 31, size=3).astype(np.int32)
        seed = int(rng.integers(0, 2 ** 32))
        b = c.tobytes()
        assert murmur3_32(b, seed) == sk_murmur(b, seed=seed, positive=True)


def test_seeds_are_mt19937_seed_2():
    bg = np.random.MT19937()
    bg._legacy_seeding(2)   # init_genrand(2), the std::mt19937(2) constructor
    np.testing.assert_array_equal(hash_seeds(), bg.random_raw(3).astype(np.uint32))


def _coords(keys, dim):
    k = np.asarray(keys, np.int64)
    return np.stack([k // (dim[1] * dim[2]), (k // dim[2]) % dim[1], k % dim[2]], 1).astype(np.int32)


def _numpy_build(dim, buckets, seeds, keys, vals, stash_cap):
    """Independent restatement of the deterministic build."""
    order = np.argsort(keys, kind="stable")
    keys, vals = np.asarray(keys, np.uint64)[order], np.asarray(vals, np.uint64)[order]
    c = _coords(keys, dim)
    hb = np.array([[sk_murmur(c[i].tobytes(), seed=int(s), positive=True) % buckets for i in range(len(keys))]
                   for s in seeds], np.int64)
    slots = np.full(buckets * 16, U64_EMPTY, np.uint64)
    fill = np.zeros(buckets, np.int64)
    todo = np.arange(len(keys))
    for r in range(3):
        nxt = []
        used = {}
        for i in todo:
            b = hb[r, i]
            rank = used.get(b, 0)
            used[b] = rank + 1
            if rank < 16 - fill[b]:
                slots[b * 16 + fill[b] + rank] = (keys[i] << np.uint64(32)) | vals[i]
            else:
                nxt.append(i)
        for b, cnt in used.items():
            fill[b] = min(16, fill[b] + cnt)
        todo = np.array(nxt, np.int64)
    stash = ((keys[todo] << np.uint64(32)) | vals[todo]).astype(np.uint64)
    return slots, stash


def _random_keys(dim, n, seed):
    rng = np.random.default_rng(seed)
    return rng.choice(dim[0] * dim[1] * dim[2], size=n, replace=False).astype(np.uint32)


@pytest.mark.parametrize("n,max_keys", [(1, 1), (700, 1000), (1000, 1000), (3000, 20000)])
def test_build_matches_numpy_restatement_and_dict(n, max_keys):
    dim = (64, 32, 48)
    keys = _random_keys(dim, n, seed=n)
    vals = np.random.default_rng(1).integers(0, 2 ** 31, n).astype(np.uint32)
    h = HashOracle(dim, max_keys)
    ns = h.build(keys, vals)
    slots, stash = _numpy_build(dim, h.buckets, h.seeds, keys, vals, h.stash_cap)
    np.testing.assert_array_equal(h.slots, slots)
    assert ns == stash.shape[0]
    np.testing.assert_array_equal(h.stash[:ns], stash)
    np.testing.assert_array_equal(h.find(keys), vals)
    absent = np.setdiff1d(np.arange(dim[0] * dim[1] * dim[2], dtype=np.uint32), keys)[::7]
    assert (h.find(absent) == EMPTY).all()
    # occupied slots are a prefix of each bucket
    occ = (h.slots != U64_EMPTY).reshape(-1, 16)
    assert (np.diff(occ.astype(np.int8), axis=1) <= 0).all()
    assert occ.sum() + ns == n


def test_build_is_order_invariant():
    dim = (40, 40, 40)
    keys = _random_keys(dim, 5000, seed=3)
    vals = keys * 3 + 1
    a, b = HashOracle(dim, 6000), HashOracle(dim, 6000)
    a.build(keys, vals)
    p = np.random.default_rng(9).permutation(keys.shape[0])
    b.build(keys[p], vals[p])
    np.testing.assert_array_equal(a.slots, b.slots)
    assert a.nstash == b.nstash


def _stash_keys(dim, buckets, seeds, want):
    """Keys whose three buckets are all bucket 0: more than 16 of them must spill."""
    allk = np.arange(dim[0] * dim[1] * dim[2], dtype=np.uint32)
    c = _coords(allk, dim)
    ok = np.ones(allk.shape[0], bool)
    for s in seeds:
        ok &= np.array([sk_murmur(c[i].tobytes(), seed=int(s), positive=True) % buckets == 0
                        for i in range(allk.shape[0])])
    sel = allk[ok]
    assert sel.shape[0] >= want
    return sel[:want]


def test_stash_path():
    dim = (16, 16, 16)
    h = HashOracle(dim, 40)
    assert h.buckets == 4
    keys = _stash_keys(dim, h.buckets, h.seeds, 24)
    ns = h.build(keys, np.arange(24, dtype=np.uint32))
    assert ns == 8
    np.testing.assert_array_equal(h.find(keys), np.arange(24, dtype=np.uint32))
    np.testing.assert_array_equal(h.stash[:8] >> np.uint64(32), np.sort(keys)[16:])


# ------------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("n,max_keys", [(700, 1000), (25000, 40000)])
def test_gpu_table_matches_oracle(n, max_keys):
    from ckmpm_amd.hashtable import BlockHashTable

    dim = (128, 128, 128)
    keys = _random_keys(dim, n, seed=11)
    vals = np.arange(n, dtype=np.uint32)[::-1].copy()
    o = HashOracle(dim, max_keys)
    o.build(keys, vals)
    g = BlockHashTable(dim, max_keys)
    g.Insert(keys, vals)
    slots, stash = g.export()
    np.testing.assert_array_equal(slots, o.slots)
    np.testing.assert_array_equal(stash, o.stash[:o.nstash])
    q = np.concatenate([keys, _random_keys(dim, 5000, seed=12)])
    np.testing.assert_array_equal(g.Find(q), o.find(q))
    inf = g.info()
    occ = (o.slots != U64_EMPTY).reshape(-1, 16).sum(1)
    np.testing.assert_array_equal(np.array(inf.fill_histogram), np.bincount(occ, minlength=17))
    assert inf.occupied_slots + inf.stash == n
    g.close()


@pytest.mark.gpu
def test_gpu_stash_and_both_find_kernels():
    import torch
    from ckmpm_amd.hashtable import BlockHashTable

    dim = (16, 16, 16)
    o = HashOracle(dim, 40)
    keys = _stash_keys(dim, o.buckets, o.seeds, 24)
    vals = np.arange(24, dtype=np.uint32) + 100
    o.build(keys, vals)
    g = BlockHashTable(dim, 40)
    g.Insert(keys, vals)
    slots, stash = g.export()
    np.testing.assert_array_equal(slots, o.slots)
    np.testing.assert_array_equal(stash, o.stash[:8])
    q = np.arange(16 * 16 * 16, dtype=np.uint32)
    expect = o.find(q)
    dq = torch.from_numpy(q.view(np.int32)).cuda()
    for mode in (g.COOPERATIVE, g.SCALAR, g.LDS):
        out = torch.empty_like(dq)
        g.FindDevice(dq.data_ptr(), q.shape[0], out.data_ptr(), cooperative=mode)
        g.Synchronize()
        np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), expect, err_msg=f"find mode {mode}")
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("nkeys", [1000, 28000])
def test_gpu_lds_find_matches_the_oracle(nkeys):
    """The LDS-probed find against the oracle's find, on a table near the LDS limit (2560
    buckets hold 28.6k keys at the creation load) and a small one, with 2M queries: every
    key, keys that are absent and keys outside the domain."""
    import torch
    from ckmpm_amd.hashtable import BlockHashTable
    from ckmpm_amd.native import CkmpmError

    dim = (128, 128, 128)
    rng = np.random.default_rng(nkeys)
    keys = rng.choice(dim[0] * dim[1] * dim[2], nkeys, replace=False).astype(np.uint32)
    vals = rng.integers(0, 2 ** 31, nkeys).astype(np.uint32)
    o = HashOracle(dim, nkeys)
    o.build(keys, vals)
    g = BlockHashTable(dim, nkeys)
    g.Insert(keys, vals)
    q = np.concatenate([keys, rng.integers(0, dim[0] * dim[1] * dim[2] + 5000, 2_000_000 - nkeys).astype(np.uint32)])
    print(f"LDS find: {nkeys} keys, {o.buckets} buckets, {q.shape[0]} queries", flush=True)
    expect = o.find(q)
    dq = torch.from_numpy(q.view(np.int32)).cuda()
    out = torch.empty_like(dq)
    g.FindDevice(dq.data_ptr(), q.shape[0], out.data_ptr(), cooperative=g.LDS)
    g.Synchronize()
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), expect)
    assert (expect[:nkeys] == vals).all()
    g.close()
    big = BlockHashTable(dim, 30000)  # 2679 buckets: beyond the LDS-probed form
    with pytest.raises(CkmpmError, match="too large"):
        big.FindDevice(dq.data_ptr(), 16, out.data_ptr(), cooperative=big.LDS)
    big.close()


@pytest.mark.gpu
def test_gpu_rejects_duplicates_and_outside_keys():
    from ckmpm_amd.hashtable import BlockHashTable
    from ckmpm_amd.native import CkmpmError

    g = BlockHashTable((8, 8, 8), 64)
    with pytest.raises(CkmpmError, match="duplicate"):
        g.Insert(np.array([3, 5, 3], np.uint32))
    with pytest.raises(CkmpmError, match="outside"):
        g.Insert(np.array([3, 512], np.uint32))
    g.Insert(np.array([7, 1], np.uint32))
    np.testing.assert_array_equal(g.Find(np.array([1, 7, 2], np.uint32)), [1, 0, EMPTY])
    g.close()


@pytest.mark.gpu
def test_gpu_layout_hash_occupancy_every_substep():
    """Neo-Hookean drop: every substep, the HIP build of the hash over the engine's
    exterior block list has exactly the oracle's slots, stash and occupancy for the same
    list (the lists themselves equal the oracle's bit for bit in the parity tests)."""
    from ckmpm_amd import scenes
    from ckmpm_amd.engine import engine_from_scene
    from ckmpm_amd.hashtable import BlockHashTable

    sc = scenes.jelly_drop(side_cells=12)
    eng = engine_from_scene(sc)
    cap = sc.block_capacity()
    g = BlockHashTable(sc.dim, cap)
    ho = HashOracle(sc.dim, cap)
    for k in range(8):
        eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
        eng.Synchronize()
        keys_g, (A, Nb, Ex) = eng.layout(0)
        vals = np.arange(Ex, dtype=np.uint32)
        g.Insert(keys_g[:Ex], vals)
        ho.build(keys_g[:Ex], vals)
        slots, stash = g.export()
        np.testing.assert_array_equal(slots, ho.slots, err_msg=f"substep {k}")
        np.testing.assert_array_equal(stash, ho.stash[:ho.nstash])
        occ = (ho.slots != U64_EMPTY).reshape(-1, 16).sum(1)
        np.testing.assert_array_equal(np.array(g.info().fill_histogram), np.bincount(occ, minlength=17))
    g.close()
