"""The drop-in boundary end to end from C: the reference's own scene
(source.cu's argument values, tests/c/patch_antenna_main.c) through
marchingLoop_mi355x, the library's counterpart of marchingLoop, on the GPU.

Its port voltage and current must equal, bit for bit, those of the same scene
built through the Python binding (scenes.patch_antenna, itself bitwise equal
to the FP32 oracle in tests/test_loads.py), and its results1003.dat /
results1004.dat must be byte-identical to fdtd_save_port's from that run.
"""
import os
import subprocess

import numpy as np
import pytest

from fdtd_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEPS = 300


COMPILERS = {"c": ["gcc", "-std=c99", "-x", "c"], "c++": ["g++", "-std=c++17", "-x", "c++"]}


def build_driver(tmp_path, lang="c"):
    """The C main against include/ and libfdtd_amd.so, built as C (gcc) or
    as C++ (g++, the language nvcc compiles source.cu in)."""
    from fdtd_amd import api
    libdir = os.path.dirname(api.LIB_PATH)
    exe = tmp_path / ("patch_antenna_" + lang.replace("+", "p"))
    cmd = COMPILERS[lang] + ["-O2", "-Wall", "-Wextra", "-Wno-unused-parameter", "-Werror",
                             "-I", os.path.join(ROOT, "include"),
                             os.path.join(ROOT, "tests", "c", "patch_antenna_main.c"), "-x", "none",
                             "-L", libdir, "-lfdtd_amd", "-Wl,-rpath," + libdir, "-lm", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("lang", sorted(COMPILERS))
def test_c_driver_builds(tmp_path, lang):
    """CPU: the C main compiles and links against the library as C and C++."""
    from fdtd_amd import api
    api.load_library()
    assert build_driver(tmp_path, lang).exists()


@pytest.mark.parametrize("lang", sorted(COMPILERS))
def test_driver_dt_passes_the_courant_check(tmp_path, lang):
    """source.cu's dt at Courant factor 1, evaluated by C's or C++'s rules
    (C++ takes the float sqrt for c, which puts dt a few float ulps above the
    double limit), passes fdtd_create's stability check.  Run for one step:
    with a device it completes; on a host without one, fdtd_create gets past
    validation and stops at the device (FDTD_E_DEVICE = -5, not
    FDTD_E_INVALID = -1)."""
    exe = build_driver(tmp_path, lang)
    r = subprocess.run([str(exe), "1", str(tmp_path / "vi.raw")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0 or "marchingLoop_mi355x: -5" in r.stderr, r.stdout + r.stderr


@pytest.mark.gpu
def test_c_driver_matches_python_binding(engine, tmp_path):
    exe = build_driver(tmp_path)
    cdir = tmp_path / "c_run"
    cdir.mkdir()
    raw = cdir / "vi.raw"
    print(f"[c-driver] patch antenna through marchingLoop_mi355x, {STEPS} steps", flush=True)
    r = subprocess.run([str(exe), str(STEPS), str(raw)], cwd=cdir, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    print(f"[c-driver] {r.stdout.strip()}", flush=True)
    vi = np.fromfile(raw, dtype=np.float32).reshape(2, STEPS)

    sc = scenes.patch_antenna(steps=STEPS)
    assert r.stdout.split()[1:4] == [str(sc.nx), str(sc.ny), str(sc.nz)]
    sim = engine.Simulation.from_scene(sc, "f32")
    sim.run(STEPS)
    v, i = sim.port()
    pdir = tmp_path / "py_run"
    pdir.mkdir()
    sim.save_port(str(pdir / "results1003.dat"), str(pdir / "results1004.dat"))
    sim.close()

    assert np.max(np.abs(v)) > 0 and np.max(np.abs(i)) > 0
    np.testing.assert_array_equal(vi[0], v)
    np.testing.assert_array_equal(vi[1], i)
    for name in ("results1003.dat", "results1004.dat"):
        c_bytes = (cdir / name).read_bytes()
        assert c_bytes == (pdir / name).read_bytes(), name
        assert len(c_bytes.splitlines()) == STEPS

    # the same main built as C++: c and dt differ in the last float bits, so
    # the series agree to FP32 round-off, not bit for bit
    exe = build_driver(tmp_path, "c++")
    xdir = tmp_path / "cpp_run"
    xdir.mkdir()
    print(f"[c-driver] the same scene built as C++, {STEPS} steps", flush=True)
    r = subprocess.run([str(exe), str(STEPS), str(xdir / "vi.raw")], cwd=xdir, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    vx = np.fromfile(xdir / "vi.raw", dtype=np.float32).reshape(2, STEPS).astype(np.float64)
    for got, want in ((vx[0], v), (vx[1], i)):
        assert np.linalg.norm(got - want) <= 1e-4 * np.linalg.norm(want)
