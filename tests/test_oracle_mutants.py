"""The oracle's pins must catch plausible mistakes.

Each case plants one mistake in a copy of oracle/fdtd_oracle.cpp (on the
CPU, in a temporary directory), builds it, and runs the pin tests against
that build in a subprocess: at least one pin has to fail.  The committed
oracle is never modified."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "fdtd_oracle.cpp")
FLAGS = ["-O3", "-march=x86-64-v3", "-mf16c", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared"]

MUTANTS = {
    # Ex curl coefficient divided by dz instead of dy (transposed cell size)
    "cell_size": ("s->Cexhz[x] = (two * dt) / (den * dy);", "s->Cexhz[x] = (two * dt) / (den * dz);"),
    # CPML psi_eyx driven by the negated difference (sign error)
    "psi_sign": ("s->a_ex[i] * (Hz[x] - Hz[x - 1]);", "s->a_ex[i] * (Hz[x - 1] - Hz[x]);"),
    # Ex edge average over the wrong neighbour cell (i-1 instead of j-1)
    "avg_index": ("(((m_eps[0][x] + m_eps[0][x - NX]) + m_eps[0][x - NXY]) + m_eps[0][x - NX - NXY])",
                  "(((m_eps[0][x] + m_eps[0][x - 1]) + m_eps[0][x - NXY]) + m_eps[0][x - 1 - NXY])"),
    # H update drops the Ez term of Hx
    "dropped_term": ("Hx[x] = (s->Chxh[x] * Hx[x] + s->Chxey[x] * (Ey[x + NXY] - Ey[x])) + s->Chxez[x] * (Ez[x + NX] - Ez[x]);",
                     "Hx[x] = (s->Chxh[x] * Hx[x] + s->Chxey[x] * (Ey[x + NXY] - Ey[x]));"),
    # CPML: kappa dropped from a's denominator (sigma + alpha for sigma + kappa alpha)
    "cpml_a_kappa": ("double dne = kpe * (sge + kpe * ale)", "double dne = kpe * (sge + ale)"),
    # CPML: kappa dropped from b's exponent (sigma for sigma / kappa)
    "cpml_b_kappa": ("std::exp((-ddt / e0) * ((sge / kpe) + ale))", "std::exp((-ddt / e0) * ((sge) + ale))"),
    # CPML: alpha graded the wrong way (largest at the outer face)
    "cpml_alpha_grading": ("double ale = amin + (amax - amin) * (1.0 - re);",
                           "double ale = amin + (amax - amin) * (re);"),
    # CPML: the kappa division of Ceyhz left out
    "cpml_kappa_division": ("s->Ceyhz[x] = s->Ceyhz[x] / kex[i];", "s->Ceyhz[x] = s->Ceyhz[x];"),
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_catch_planted_mistake(name, tmp_path):
    old, new = MUTANTS[name]
    text = open(SRC).read()
    assert text.count(old) == 1, f"mutation site for {name} not found"
    src = tmp_path / "fdtd_oracle.cpp"
    src.write_text(text.replace(old, new))
    lib = tmp_path / "liboracle.so"
    subprocess.run(["g++"] + FLAGS + ["-o", str(lib), str(src)], check=True)
    env = dict(os.environ, ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py"),
                        "-k", "not reflection"], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "failed" in r.stdout


# ---- port reductions (oracle/port.py) -------------------------------------------

PORT_SRC = os.path.join(ROOT, "oracle", "port.py")
PORT_MUTANTS = {
    # z-directed loop one plane too high
    "loop_plane": ('                if k == ke - 1:\n                    acc("yp"', '                if k == ke:\n                    acc("yp"'),
    # one line of the loop with the wrong sign
    "loop_sign": ('- f32(dx) * sums["xn"])\n    else:', '+ f32(dx) * sums["xn"])\n    else:'),
    # x-directed loop's z-line one row inside the box
    "loop_row": ('acc("zp", H(HZ, cnx - 2, cny - 1, k))', 'acc("zp", H(HZ, cnx - 2, cny - 2, k))'),
    # y-directed loop's x-line on the wrong row
    "loop_col": ('acc("xp", H(HX, i, cny - 2, k))', 'acc("xp", H(HX, i, cny - 1, k))'),
    # voltage with the sign of -dz dropped
    "volt_sign": ("csuf = f32(-f32(dz))", "csuf = f32(f32(dz))"),
    # cross-section count off by one
    "volt_area": ("(_je - _js + f32(1)))", "(_je - _js + f32(2)))"),
}


@pytest.mark.parametrize("name", sorted(PORT_MUTANTS))
def test_port_pins_catch_planted_mistake(name, tmp_path):
    old, new = PORT_MUTANTS[name]
    text = open(PORT_SRC).read()
    assert text.count(old) == 1, f"mutation site for {name} not found"
    pkg = tmp_path / "oracle"
    pkg.mkdir()
    for f in ("__init__.py", "oracle.py"):
        (pkg / f).write_text(open(os.path.join(ROOT, "oracle", f)).read())
    (pkg / "port.py").write_text(text.replace(old, new))
    env = dict(os.environ, PYTHONPATH=f"{tmp_path}{os.pathsep}{ROOT}",
               ORACLE_LIB=os.path.join(ROOT, "oracle", "liboracle.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_port.py")], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "failed" in r.stdout


# ---- plates and resistors (oracle/fdtd_oracle.cpp, pinned by tests/test_loads.py) ----

LOAD_MUTANTS = {
    # z-plate Ex edges taken with i <= ie (one edge past the plate)
    "plate_z_range": ("if (i >= is && i < ie && j >= js && j <= je && k == ks) se[0][x]",
                      "if (i >= is && i <= ie && j >= js && j <= je && k == ks) se[0][x]"),
    # x-plate Ez edges stop one row short
    "plate_x_range": ("if (i == is && j >= js && j <= je && k >= ks && k < ke) se[2][x]",
                      "if (i == is && j >= js && j < je && k >= ks && k < ke) se[2][x]"),
    # y-plate Ex edges get the z sigma
    "plate_y_sigma": ("if (i >= is && i < ie && j == js && k >= ks && k <= ke) se[0][x] = (R)sg[0];",
                      "if (i >= is && i < ie && j == js && k >= ks && k <= ke) se[0][x] = (R)sg[2];"),
    # resistor factor with dx in place of dz (transposed cell size)
    "resistor_cell": ("const R rfactor = (dt * dz) / (r_comp * dx * dy);\n        for (int k = 0; k < NZ; ++k)",
                      "const R rfactor = (dt * dx) / (r_comp * dx * dy);\n        for (int k = 0; k < NZ; ++k)"),
    # resistor loss entering Ce with the wrong sign
    "resistor_sign": ("s->Ceze[x]  = (two * eps[2][x] * eps_0 - dt * se[2][x] - rfactor) / D;\n"
                      "                    s->Cezhy[x] = (two * dt / dx) / D;\n"
                      "                    s->Cezhx[x] = ((R)-2 * dt / dy) / D;\n                }\n    }",
                      "s->Ceze[x]  = (two * eps[2][x] * eps_0 - dt * se[2][x] + rfactor) / D;\n"
                      "                    s->Cezhy[x] = (two * dt / dx) / D;\n"
                      "                    s->Cezhx[x] = ((R)-2 * dt / dy) / D;\n                }\n    }"),
    # sector radius measured from a ray end instead of the apex
    "sector_radius": ("dist = ((px1 - pontox) * (px1 - pontox)) + ((py1 - pontoy) * (py1 - pontoy));",
                      "dist = ((px0 - pontox) * (px0 - pontox)) + ((py0 - pontoy) * (py0 - pontoy));"),
    # sector: the first side test keeps the other half-plane
    "sector_side": ("rvetorial2 = pvetorial2d(ax, ay, zx1, zy1);\n    rescalar = rvetorial1 * rvetorial2;\n"
                    "    if (rescalar < 0) return -1;",
                    "rvetorial2 = pvetorial2d(ax, ay, zx1, zy1);\n    rescalar = rvetorial1 * rvetorial2;\n"
                    "    if (rescalar > 0) return -1;"),
    # sector Ey edges stop one column short
    "sector_ey_range": ("if (i >= 0 && i <= max_x && j >= 0 && j < max_y && kz == k)",
                        "if (i >= 0 && i < max_x && j >= 0 && j < max_y && kz == k)"),
    # x-source rows bounded by je (the reference bounds them by ie)
    "xsource_rows": ("i >= b[0] && i < b[3] && j >= b[1] && j <= b[3] && k >= b[2] && k <= b[5]",
                     "i >= b[0] && i < b[3] && j >= b[1] && j <= b[4] && k >= b[2] && k <= b[5]"),
    # x-source resistance not split over the parallel columns
    "xsource_rcomp": ("R r_comp = (((R)1 + _je) - _js) * (((R)1 + _ke) - _ks) / (_ie - _is);",
                      "R r_comp = (R)1 / (_ie - _is);"),
    # x-source factor with dz in place of dx
    "xsource_cell": ("const R rfactor = (dt * dx) / (r_comp * dz * dy);",
                     "const R rfactor = (dt * dz) / (r_comp * dz * dy);"),
    # x-source Ez injection over ks..ke inclusive
    "xsource_inject": ("for (int k = b[2]; k < b[5]; ++k)\n                for (int j = b[1]; j <= b[4]; ++j)\n"
                       "                    for (int i = b[0]; i <= b[3]; ++i) {\n                        i64 x = s->id(i, j, k);\n"
                       "                        s->src_of[x] = q;\n                        s->Cezj[x] = (R)(float)1.2e-38;",
                       "for (int k = b[2]; k <= b[5]; ++k)\n                for (int j = b[1]; j <= b[4]; ++j)\n"
                       "                    for (int i = b[0]; i <= b[3]; ++i) {\n                        i64 x = s->id(i, j, k);\n"
                       "                        s->src_of[x] = q;\n                        s->Cezj[x] = (R)(float)1.2e-38;"),
}


@pytest.mark.parametrize("name", sorted(LOAD_MUTANTS))
def test_load_pins_catch_planted_mistake(name, tmp_path):
    old, new = LOAD_MUTANTS[name]
    text = open(SRC).read()
    assert text.count(old) == 1, f"mutation site for {name} not found"
    src = tmp_path / "fdtd_oracle.cpp"
    src.write_text(text.replace(old, new))
    lib = tmp_path / "liboracle.so"
    subprocess.run(["g++"] + FLAGS + ["-o", str(lib), str(src)], check=True)
    env = dict(os.environ, ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_loads.py")], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "failed" in r.stdout
