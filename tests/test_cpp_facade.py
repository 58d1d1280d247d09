"""C++ facade (include/ckmpm.hpp) and the ckmpm_scene driver built on it.

CPU: the header compiles warning-free on its own; its raw float32 xyz reader (the format
of the reference's particle assets) round-trips what the writer wrote and rejects a
ragged file; the driver prints its usage and rejects bad command lines before it
touches a device.
GPU: the driver, given scenes as command lines and positions as .bin assets, ends with
the same particle positions as the Python engine on the same scenes (fixed-corotated,
Neo-Hookean, Drucker-Prager, fluid, von Mises and Cam-clay models; substep and frame
loops).
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
DRIVER = os.path.join(ROOT, "ckmpm_amd", "ckmpm_scene")


def _cxx(tmp_path, src, name):
    p = tmp_path / f"{name}.cpp"
    p.write_text(src)
    exe = tmp_path / name
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", INCLUDE, str(p), "-o", str(exe)],
                   check=True, capture_output=True)
    return exe


def test_header_compiles_alone(tmp_path):
    p = tmp_path / "h.cpp"
    p.write_text('#include "ckmpm.hpp"\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I", INCLUDE, str(p)],
                   check=True)


def test_raw_positions_round_trip_and_reject_ragged(tmp_path):
    # host-only use of the header (no library symbol is referenced)
    exe = _cxx(tmp_path, r'''
#include <cstdio>
#include "ckmpm.hpp"
int main(int, char** argv) {
    std::vector<float> xyz = ckmpm::ReadRawPositions(argv[1]);
    for (float& v : xyz) v = 2.0f * v + 1.0f;
    ckmpm::WriteRawPositions(argv[2], xyz);
    try { ckmpm::ReadRawPositions(argv[3]); } catch (const ckmpm::Error& e) {
        std::printf("%d %zu\n", e.code(), xyz.size()); return 0; }
    return 3;
}''', "raw")
    x = np.random.default_rng(0).random((1000, 3), dtype=np.float32)
    x.tofile(tmp_path / "a.bin")
    (tmp_path / "bad.bin").write_bytes(b"\0" * 13)
    r = subprocess.run([str(exe), str(tmp_path / "a.bin"), str(tmp_path / "b.bin"), str(tmp_path / "bad.bin")],
                       capture_output=True, text=True, check=True)
    assert r.stdout.split() == ["1", "3000"]  # CKMPM_ERR_INVALID for the 13-byte file
    np.testing.assert_array_equal(np.fromfile(tmp_path / "b.bin", np.float32).reshape(-1, 3), 2 * x + 1)


def test_driver_usage_and_argument_errors(tmp_path):
    assert os.path.exists(DRIVER), "ckmpm_scene is built with the library (ckmpm_amd/build.py)"
    r = subprocess.run([DRIVER, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("usage: ckmpm_scene")
    bad = [
        ([], "--domain and --dx are required"),
        (["--domain", "8,8"], "needs 3 comma-separated numbers"),
        (["--domain", "8,8,8", "--dx", "0.03"], "no model"),
        (["--domain", "8,8,8", "--dx", "x"], "needs a number"),
        (["--domain", "8,8,8", "--dx", "0.03", "--material", "wood", "--positions", "p.bin"], "must be fc, nh, dp"),
        (["--domain", "8,8,8", "--dx", "0.03", "--material", "fc", "--mass", "1", "--volume", "1",
          "--positions", "p.bin"], "a solid needs"),
        (["--bogus", "1"], "unknown option"),
        (["--walls", "soft"], "--walls must be"),
        (["--region", "fixed,1,2,3"], "needs 6 comma-separated numbers"),
        (["--region", "spin,1,2,3,4,5,6"], "--region kind must be"),
        (["--domain", "8,8,8", "--dx", "0.03", "--material", "fluid", "--bulk", "1", "--mass", "1", "--volume", "1",
          "--positions", "ok.bin", "--substeps", "1", "--export", "."], "use it with --frames"),
    ]
    (tmp_path / "ok.bin").write_bytes(b"\0" * 24)   # two particles at the origin
    (tmp_path / "odd.bin").write_bytes(b"\0" * 20)  # not whole xyz triples
    for args, msg in bad:
        r = subprocess.run([DRIVER, *args], capture_output=True, text=True, cwd=tmp_path)
        assert r.returncode == 2 and msg in r.stderr, (args, r.stderr)
    r = subprocess.run([DRIVER, "--domain", "8,8,8", "--dx", "0.03", "--material", "fluid", "--bulk", "1e4",
                        "--gamma", "7", "--mass", "1", "--volume", "1", "--positions", "odd.bin", "--substeps", "1"],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "whole number of float32 xyz" in r.stderr


# ------------------------------------------------------------------------------- GPU
_KIND = {0: "fc", 1: "nh", 2: "dp", 3: "fluid", 4: "vm", 5: "nacc"}


def _command(sc, tmp_path):
    args = [DRIVER, "--domain", ",".join(map(str, sc.dim)), "--offset", ",".join(map(str, sc.offset)),
            "--dx", repr(sc.dx), "--dt", repr(sc.dt), "--gravity", repr(sc.gravity),
            "--mass-clamp", repr(sc.mass_clamp), "--capacity", str(sc.block_capacity())]
    for i, m in enumerate(sc.models):
        mat = m.material
        args += ["--material", _KIND[mat.kind], "--mass", repr(m.mass), "--volume", repr(m.volume)]
        if mat.kind == 3:
            args += ["--bulk", repr(mat.bulk), "--gamma", repr(mat.gamma), "--viscosity", repr(mat.viscosity),
                     "--J0", repr(mat.J0)]
        else:
            args += ["--lambda", repr(mat.lam), "--mu", repr(mat.mu), "--friction", repr(mat.friction_deg),
                     "--yield", repr(mat.yield_stress), "--alpha0", repr(mat.alpha0), "--beta", repr(mat.beta),
                     "--xi", repr(mat.hardening_factor), "--M", repr(mat.m), "--hardening", str(int(mat.hardening_on))]
        v = np.asarray(m.v, np.float32).reshape(-1, 3)
        if v.shape[0] == 1:
            args += ["--velocity", ",".join(repr(float(c)) for c in v[0])]
        else:
            v.tofile(tmp_path / f"v{i}.bin")
            args += ["--velocities", str(tmp_path / f"v{i}.bin")]
        np.asarray(m.x, np.float32).tofile(tmp_path / f"x{i}.bin")
        args += ["--positions", str(tmp_path / f"x{i}.bin")]
    b = sc.boundary
    if b is not None:
        args += ["--walls", ["sticky", "slip", "none"][b.walls]]
        for r in b.regions:
            kind = ["fixed", "velocity", "rotation"][r.kind]
            vals = list(r.lo) + list(r.hi) + ([] if r.kind == 0 else list(r.value)) + (list(r.center) if r.kind == 2 else [])
            args += ["--region", kind + "," + ",".join(repr(float(x)) if i >= 6 else str(int(x)) for i, x in enumerate(vals))]
    return args + ["--out", str(tmp_path / "out")]


def _scenes():
    from ckmpm_amd import scenes

    nh = scenes.colliding_spheres(seed=3)
    nh.models[1].material.kind = scenes.NEO_HOOKEAN
    from tests.test_boundary import _corner_scene, _regions
    from tests.test_gpu_parity import _box_scene  # boxes with random velocities: plastic flow

    corner = _corner_scene(None)
    corner.boundary = scenes.Boundary(scenes.WALLS_SLIP, _regions(corner.dx))

    return [("spheres_fc", scenes.colliding_spheres()), ("spheres_fc_nh", nh),
            ("box_vm", _box_scene(scenes.VON_MISES, seed=30)), ("box_nacc", _box_scene(scenes.CAM_CLAY, seed=31)),
            ("corner_boundary", corner),
            ("fluid_sand", scenes.mixed_dam_sand(n_fluid=9_000, n_sand=5_000, scale=8))]


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["spheres_fc", "spheres_fc_nh", "box_vm", "box_nacc", "corner_boundary", "fluid_sand"])
def test_driver_matches_python_engine(case, tmp_path):
    from ckmpm_amd.engine import engine_from_scene

    sc = dict(_scenes())[case]
    steps = 12
    print(f"driver case {case}, {sc.n_particles} particles, {steps} substeps", flush=True)
    r = subprocess.run(_command(sc, tmp_path) + ["--substeps", str(steps), "--conservation", str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f'"substeps": {steps}' in r.stdout and '"flags": 0' in r.stdout, r.stdout
    eng = engine_from_scene(sc)
    eng.set_conservation(steps)
    for k in range(steps):
        eng.Substep(0.0 if k == 0 else sc.dt, sc.dt)
    eng.Synchronize()
    for i in range(len(sc.models)):
        got = np.fromfile(tmp_path / f"out{i}.bin", np.float32).reshape(-1, 3)
        want = eng.positions(i)
        assert got.shape == want.shape
        # FP32 global atomics of the P2G flush add in any order: allow a few ulps
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-5 * sc.dx)
        # the reference's conservation file: 13 floats per substep
        rec = np.fromfile(tmp_path / f"conservationMetric_model_{i}.bin", np.float32).reshape(-1, 13)
        ref = eng.conservation(i)
        assert rec.shape == (steps, 13)
        # times: the driver steps with the FP32 Config::dt, this engine with the FP64 scene dt
        np.testing.assert_allclose(rec[:, 0], ref[:, 0], rtol=1e-6)
        assert (rec[:, 1:7] == 0).all()
        np.testing.assert_allclose(rec[:, 7:], ref[:, 7:], rtol=0, atol=1e-5 * max(np.abs(ref[:, 7:]).max(), 1e-3))
    eng.close()


@pytest.mark.gpu
def test_driver_frame_loop_matches_simulate(tmp_path):
    from ckmpm_amd import scenes
    from ckmpm_amd.engine import engine_from_scene

    sc = scenes.colliding_spheres()
    print("driver frame loop, colliding spheres, 1 frame", flush=True)
    r = subprocess.run(_command(sc, tmp_path) + ["--frames", "1", "--export", str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    from ckmpm_amd import bgeo
    for i in range(len(sc.models)):  # the per-frame export of the last frame is the final state
        final = np.fromfile(tmp_path / f"out{i}.bin", np.float32)
        np.testing.assert_array_equal(np.fromfile(tmp_path / f"model_{i}_particle_frame_0.txt", np.float32), final)
        np.testing.assert_array_equal(bgeo.read_positions(tmp_path / f"model_{i}_particle_frame_0.bgeo"),
                                      final.reshape(-1, 3))
    eng = engine_from_scene(sc)
    eng.Simulate(1)
    eng.Synchronize()
    n = eng.counters().substeps
    assert f'"substeps": {n}' in r.stdout, (n, r.stdout)
    for i in range(len(sc.models)):
        got = np.fromfile(tmp_path / f"out{i}.bin", np.float32).reshape(-1, 3)
        np.testing.assert_allclose(got, eng.positions(i), rtol=0, atol=1e-4 * sc.dx)
    eng.close()
