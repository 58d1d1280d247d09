"""Committed oracle vectors (tests/golden/tiny_mixed.npz, written by make_golden.py).

CPU: the oracle reproduces them (orderings bit-exact, FP64 values to 1e-12).
GPU: the HIP engine reproduces the orderings bit-exactly and the final state / grid
within the parity tolerances, straight from the stored vectors.
"""
import os

import numpy as np
import pytest

from tests.golden import make_golden

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiny_mixed.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_oracle_reproduces_golden(golden):
    got = make_golden.run(make_golden.tiny_scene())
    assert set(got) == set(golden)
    for k, v in golden.items():
        if v.dtype.kind in "ui":
            np.testing.assert_array_equal(got[k], v, err_msg=k)
        else:
            np.testing.assert_allclose(got[k], v, rtol=1e-12, atol=1e-15, err_msg=k)
    assert int(golden["flags"][0]) == 0


@pytest.mark.gpu
def test_gpu_matches_golden(golden):
    from ckmpm_amd.engine import engine_from_scene

    sc = make_golden.tiny_scene()
    eng = engine_from_scene(sc)
    keys, counts = eng.layout(0)
    np.testing.assert_array_equal(keys, golden["setup_keys"])
    assert counts == tuple(golden["setup_counts"])
    for k in range(make_golden.STEPS):
        eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
        keys, counts = eng.layout(0)
        np.testing.assert_array_equal(keys, golden[f"keys_{k}"], err_msg=f"step {k}")
        for m in range(len(sc.models)):
            perm, off = eng.ordering(m)
            np.testing.assert_array_equal(perm, golden[f"perm_{k}_{m}"], err_msg=f"step {k} model {m}")
            np.testing.assert_array_equal(off, golden[f"off_{k}_{m}"], err_msg=f"step {k} model {m}")
    eng.Synchronize()
    for m in range(len(sc.models)):
        st = eng.global_state(m)
        ref = golden[f"state_{m}"]
        np.testing.assert_allclose(st[:3], ref[:3], rtol=0, atol=1e-4 * sc.dx)
        np.testing.assert_allclose(st[3:], ref[3:], rtol=0, atol=1e-4)
    Mg, Mo = eng.grid_m().astype(np.float64), golden["grid_m"]
    for sl in (slice(0, 1), slice(1, 4)):
        assert np.abs(Mg[:, sl] - Mo[:, sl]).max() <= 1e-4 * np.abs(Mo[:, sl]).max()
