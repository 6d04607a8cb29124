"""Mutation check of the oracle's pinning tests.

Each mutant is the oracle built with one plausible mistake planted
(-DORC_MUTANT=n, see fauxgl_oracle.c); the pinning tests in
test_oracle_pins.py, run against that build, must fail.  This shows the
pins constrain the oracle beyond restating its own formulas.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

MUTANTS = {
    1: ("pow: mantissa renormalisation loses its exponent step", "pow"),
    2: ("Reflect adds instead of subtracting", "phong"),
    3: ("Color.NRGBA rounds instead of truncating", "nrgba"),
    4: ("LookAt's up vector is z x x reversed", "look_at"),
    5: ("Barycentric u = 1 - v + w", "barycentric"),
    6: ("BilinearSample swaps the (x0,y1) and (x1,y0) texels", "bilinear"),
    7: ("edge() uses the wrong vertex in its second product", "coverage or depth_is"),
    8: ("depth test uses < instead of <=", "counters"),
    9: ("blend drops the 0x101 factor", "blend"),
    10: ("drawClippedTriangle never swaps v0/v2", "windings"),
    11: ("the near clip plane tests against P instead of N", "clip_polygon"),
    12: ("tent taps start one source pixel early", "resolve"),
    13: ("tent horizontal pass does not premultiply by alpha", "resolve"),
    14: ("box resolve truncates instead of rounding", "resolve"),
    15: ("pow: the fractional part uses log10 for log", "pow"),
    16: ("pow: the fold yf > 1/2 drops its yi + 1", "pow"),
}


@pytest.mark.parametrize("n", sorted(MUTANTS))
def test_mutant_is_caught(n, tmp_path):
    desc, sel = MUTANTS[n]
    lib = tmp_path / f"liboracle_m{n}.so"
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=gnu11", "-D_GNU_SOURCE",
                    f"-DORC_MUTANT={n}", "-shared", "-o", str(lib), os.path.join(ORACLE_DIR, "fauxgl_oracle.c"),
                    "-lm", "-lpthread"], check=True)
    env = dict(os.environ, ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py"), "-k", sel],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode != 0, f"mutant {n} ({desc}) survived the pinning tests:\n{r.stdout[-2000:]}"
    assert "passed" in r.stdout or "failed" in r.stdout
