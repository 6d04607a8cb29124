"""Golden fixtures (tests/golden): scenes built from three of the
reference's example assets and the seeded sphere-and-cube scene, with the
oracle's RasterizeInfo and buffer hashes frozen in expected.json.  This CPU
test re-renders them with the oracle (a change to the oracle shows up
here); test_gpu_golden.py renders them on the GPU."""
import json
import os
import sys

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import scenes as S  # noqa: E402

EXPECTED = json.load(open(os.path.join(HERE, "expected.json")))


def test_fixture_lists_every_scene():
    assert sorted(EXPECTED) == sorted(S.SCENES)


def test_inputs_are_the_assets():
    z = S.load_inputs()
    assert z["cube_pos"].shape == (12, 3, 3)       # examples/cube.stl
    assert z["square_pos"].shape == (2, 3, 3)      # examples/square.obj, fan-triangulated quad
    assert z["hello_pos"].shape[1:] == (3, 3) and len(z["hello_pos"]) > 1000  # examples/hello.stl
    assert abs(z["hello_pos"]).max() == pytest.approx(1.0)  # BiUnitCube


@pytest.mark.parametrize("name", S.SCENES)
def test_oracle_reproduces_golden(name):
    W, H, ops = S.build(name, S.load_inputs())
    color, depth, infos, _ = S.run_oracle(W, H, ops)
    assert S.summary(W, H, color, depth, infos) == EXPECTED[name]
