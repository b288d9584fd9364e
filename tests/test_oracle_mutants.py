"""Planted-mistake check of the oracle pins (CPU).

Each case copies the oracle's sources to a temporary directory, plants one
plausible mistake (a coefficient, a dropped term), builds the copy and runs
the known-answer pins of tests/test_oracle_ctu_pins.py against it through
PC_ORACLE_SO.  The pins must fail for every mistake and pass for the
unchanged copy: a pin that a mistake survives pins nothing.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PINS = os.path.join(ROOT, "tests", "test_oracle_ctu_pins.py")

# (name, file, text, replacement, expected count)
MUTANTS = [
    # CTU corner coupling dt/(3 dx) -> dt/(2 dx) (Source/Godunov.cpp:88-91)
    ("cdtd_half", "oracle_hydro.hpp", "1.0 / 3.0 * dt / d", "1.0 / 2.0 * dt / d", 3),
    # transdo pressure: (Gamma1 - 1) -> Gamma1 (Source/Godunov.H:589-629)
    ("transdo_gamma1", "oracle_hydro.hpp",
     "cdtdx * (dAup + pav * dAu * (c - 1.0))", "cdtdx * (dAup + pav * dAu * c)", 2),
    # transdd pressure: (Gamma1 - 1) -> Gamma1 (Source/Godunov.H:783-786)
    ("transdd_gamma1", "oracle_hydro.hpp", "dAux * (c - 1.0)", "dAux * c", 1),
    # transdd half-step coefficient dt/(2 dx) -> dt/(3 dx) (Source/Godunov.cpp:85-87)
    ("hdtd_third", "oracle_hydro.hpp", "0.5 * dt / d", "1.0 / 3.0 * dt / d", 3),
    # pc_estdt_hydro: dx/(c + |u|) -> dx/(c + |u|/2) (Source/Timestep.H:73-84)
    ("estdt_half_u", "oracle_hydro.hpp", "(c + std::abs(u", "(c + 0.5 * std::abs(u", 3),
    # pc_pdivu: 1/2 -> 1/4 (Source/Godunov.H:1043)
    ("pdivu_quarter", "oracle_hydro.hpp", "pdivu(i, j, k) =\n        0.5 *",
     "pdivu(i, j, k) =\n        0.25 *", 1),
    # computeNewDt without the change_max limit (Source/PeleC.cpp:996-1010)
    ("no_change_max", "pc_oracle.cpp",
     "if (prm.fixed_dt <= 0.0) dtn = mn(dtn, prm.change_max * dt);",
     "(void)dt;", 1),
    # stop_time window 0.001 dt -> 0.01 dt (Source/PeleC.cpp:1019)
    ("stop_eps", "pc_oracle.cpp", "const double dt_eps = 0.001 * dt0;",
     "const double dt_eps = 0.01 * dt0;", 1),
]


def _build(tmp, name, mutation):
    d = os.path.join(tmp, name)
    shutil.copytree(os.path.join(ROOT, "oracle"), os.path.join(d, "oracle"),
                    ignore=shutil.ignore_patterns("*.so", "_ref"))
    shutil.copytree(os.path.join(ROOT, "include"), os.path.join(d, "include"))
    if mutation is not None:
        fname, old, new, count = mutation
        p = os.path.join(d, "oracle", fname)
        src = open(p).read()
        assert src.count(old) == count, (name, src.count(old))
        open(p, "w").write(src.replace(old, new))
    so = os.path.join(d, "liboracle.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared",
                           "-ffp-contract=off", "-mfma", "-fopenmp", "-o", so,
                           os.path.join(d, "oracle", "pc_oracle.cpp")])
    return so


def _run_pins(so):
    env = dict(os.environ, PC_ORACLE_SO=so, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p",
                        "no:cacheprovider", PINS], env=env, cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-600:]


def test_every_planted_mistake_fails_a_pin(tmp_path):
    cases = [("unchanged", None)] + [(m[0], m[1:]) for m in MUTANTS]
    with ThreadPoolExecutor(max_workers=4) as ex:
        sos = list(ex.map(lambda c: _build(str(tmp_path), *c), cases))
        results = list(ex.map(_run_pins, sos))
    rc0, out0 = results[0]
    assert rc0 == 0, "pins fail on the unchanged oracle copy:\n" + out0
    survivors = [c[0] for c, (rc, _) in zip(cases[1:], results[1:]) if rc == 0]
    assert not survivors, f"mistakes that pass every pin: {survivors}"
    for (name, _), (rc, out) in zip(cases[1:], results[1:]):
        # a failing test, not a crash or a collection error
        assert rc == 1, (name, rc, out)
