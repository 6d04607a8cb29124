"""The golden scenes (tests/golden) rendered by the gfx950 library must
reproduce the frozen RasterizeInfo and buffer hashes bit for bit."""
import json
import os
import sys

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import scenes as S  # noqa: E402

EXPECTED = json.load(open(os.path.join(HERE, "expected.json")))

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", S.SCENES)
def test_gpu_reproduces_golden(name):
    print(f"[case] golden {name}", flush=True)
    W, H, ops = S.build(name, S.load_inputs())
    color, depth, infos, _ = S.run_gpu(W, H, ops)
    got = S.summary(W, H, color, depth, infos)
    assert got == EXPECTED[name], {k: (got[k], EXPECTED[name][k]) for k in got if got[k] != EXPECTED[name][k]}
