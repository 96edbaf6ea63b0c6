"""The field-solve and Jz pins have teeth: each plausible slip in the oracle's
restatement of field.f90:125-171 and :292-298 -- a flipped sign of one term,
a wrong coefficient, a shifted index -- is planted in a copy of the oracle's
C source, compiled on the CPU, and must fail at least one pin of
test_oracle_field.py.  The shipped oracle passes all of them (that file)."""
import os
import subprocess

import pytest

import test_oracle_field as pins
from oracle import Oracle, OracleError

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                   "wm_oracle.c")

# (name, text in wm_oracle.c, replacement)
MUTANTS = [
    ("gkl1 J sign", "UF(1, i, j + 1) + f3 * (-UJ(3, i, j - 1) + UJ(3, i, j))) -",
     "UF(1, i, j + 1) - f3 * (-UJ(3, i, j - 1) + UJ(3, i, j))) -"),
    ("gkl2 J sign", "UF(2, i, j + 1) - f3 * (-UJ(3, i - 1, j) + UJ(3, i, j))) +",
     "UF(2, i, j + 1) + f3 * (-UJ(3, i - 1, j) + UJ(3, i, j))) +"),
    ("gkl3 Jx sign", "f3 * (-UJ(2, i - 1, j) + UJ(2, i, j) + UJ(1, i, j - 1) - UJ(1, i, j))",
     "f3 * (-UJ(2, i - 1, j) + UJ(2, i, j) - UJ(1, i, j - 1) + UJ(1, i, j))"),
    ("gkl3 Jy index", "f3 * (-UJ(2, i - 1, j) + UJ(2, i, j)",
     "f3 * (-UJ(2, i, j - 1) + UJ(2, i, j)"),
    ("gkl1 E sign", "f1 * (-UF(6, i, j - 1) + UF(6, i, j));",
     "f1 * (UF(6, i, j - 1) - UF(6, i, j));"),
    ("gkl2 E sign", "f1 * (-UF(6, i - 1, j) + UF(6, i, j));",
     "f1 * (UF(6, i - 1, j) - UF(6, i, j));"),
    ("gkl3 Ex sign", "UF(5, i, j) + UF(4, i, j - 1) - UF(4, i, j));",
     "UF(5, i, j) - UF(4, i, j - 1) + UF(4, i, j));"),
    ("gkl Laplacian", "UF(2, i - 1, j) - 4. * UF(2, i, j)", "UF(2, i - 1, j) - 3. * UF(2, i, j)"),
    ("gkl f3 term dropped", "UF(3, i, j + 1) +\n                f3 *", "UF(3, i, j + 1) +\n                0.0 *"),
    ("dEx gfac sign", "gfac * (-DF(3, i, j) + DF(3, i, j + 1))", "gfac * (DF(3, i, j) - DF(3, i, j + 1))"),
    ("dEy B index", "(-UF(3, i, j) + UF(3, i + 1, j))", "(-UF(3, i, j) + UF(3, i, j + 1))"),
    ("dEz J sign", "4. * OR_PI * delt * UJ(3, i, j);", "-4. * OR_PI * delt * UJ(3, i, j);"),
    ("dEz Bx sign", "UF(1, i, j) - UF(1, i, j + 1))) -", "-UF(1, i, j) + UF(1, i, j + 1))) -"),
    ("dEz dBy sign", "gfac * (-DF(2, i, j) + DF(2, i + 1, j) + DF(1, i, j) -",
     "gfac * (DF(2, i, j) - DF(2, i + 1, j) + DF(1, i, j) -"),
    ("dEx J coefficient", "4. * OR_PI * delt * UJ(1, i, j);", "2. * OR_PI * delt * UJ(1, i, j);"),
    ("Jz third -> quarter", "const double fac = 1.0 / 3.0;", "const double fac = 1.0 / 4.0;"),
    ("Jz half term", "0.5 * ds[ip + 2][0] * s0[jp + 2][1]", "0.4 * ds[ip + 2][0] * s0[jp + 2][1]"),
    ("Jz cross term transposed", "fac * ds[ip + 2][0] * ds[jp + 2][1]",
     "fac * ds[jp + 2][0] * ds[ip + 2][1]"),
]


def _checks(factory):
    """(deviation, tolerance) of every pin; an error counts as a failure."""
    out = []
    runs = [
        (lambda: pins.check_single_mode(factory, 1, 1, 1, 0.501), 1e-11),
        (lambda: pins.check_single_mode(factory, 2, 3, -2, 0.6), 1e-11),
        (lambda: pins.check_single_mode(factory, 3, 2, 1, 0.501, with_fields=False), 1e-11),
        (lambda: pins.check_vacuum_mode(factory, "TE", 3, -1, 0.501, -1), 1e-11),
        (lambda: pins.check_vacuum_mode(factory, "TM", 1, 2, 0.5, 1), 1e-11),
        (lambda: pins.check_uniform_current(factory), 1e-15),
        (lambda: pins.check_div_b(factory, steps=4), 1e-13),
    ] + [(lambda a=a: pins.check_moving_jz(factory, *a), 1e-14) for a in pins.JZ_MOVES]
    for fn, tol in runs:
        try:
            out.append((fn(), tol))
        except (OracleError, AssertionError, FloatingPointError, ValueError):
            out.append((float("inf"), tol))
    return out


def _build(tmp_path, name, old, new):
    src = open(SRC).read()
    assert src.count(old) == 1, f"mutant '{name}': text not found once in wm_oracle.c"
    d = tmp_path / name.replace(" ", "_")
    d.mkdir()
    c = d / "wm_oracle.c"
    c.write_text(src.replace(old, new))
    (d / "wm_oracle.h").write_text(open(SRC[:-1] + "h").read())
    lib = d / "libmut.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off",
                    "-fopenmp", "-o", str(lib), str(c), "-lm"], check=True)
    return str(lib)


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_slip_fails_a_pin(tmp_path, name, old, new):
    lib = _build(tmp_path, name, old, new)
    res = _checks(lambda p: Oracle(p, lib=lib))
    assert any(not (dev < tol) for dev, tol in res), f"mutant '{name}' passes every pin"


def test_shipped_oracle_passes_the_same_checks():
    res = _checks(Oracle)
    assert all(dev < tol for dev, tol in res), res
