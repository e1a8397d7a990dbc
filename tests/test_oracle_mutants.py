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
