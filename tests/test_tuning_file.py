"""The tile-shape tuning table's file form (fgl_tune_save / fgl_tune_load /
fgl_tune_clear) and the list of compiled shapes: host-side only, no GPU."""
import os

import pytest

from fauxgl_amd import _native as N
from fauxgl_amd import api as A

HEADER = "fauxgl_amd tile tuning 1\n"


def test_compiled_shapes():
    shapes = A.TileShapes()
    assert shapes[0] == (32, 32)  # the default comes first
    assert len(set(shapes)) == len(shapes) >= 2
    for w, h in shapes:
        assert w * h <= 65536  # tile-local pixel indices are 16-bit


def test_round_trip(tmp_path):
    A.ClearTuning()
    src = tmp_path / "in.txt"
    src.write_text(HEADER + "# comment\n"
                   "7680 4320 0 20 64 32 1000000 9000000 24000000 0.912000\n"
                   "1920 1080 0 20 32 16 800000 3000000 2000000 0.300000\n")
    A.LoadTuning(src)
    out = tmp_path / "out.txt"
    A.SaveTuning(out)
    lines = [l for l in out.read_text().splitlines() if l and not l.startswith("#")]
    assert lines[0] == HEADER.strip()
    rows = sorted(tuple(l.split()[:9]) for l in lines[1:])
    assert rows == sorted([("7680", "4320", "0", "20", "64", "32", "1000000", "9000000", "24000000"),
                           ("1920", "1080", "0", "20", "32", "16", "800000", "3000000", "2000000")])
    # loading merges; clearing empties
    A.ClearTuning()
    A.SaveTuning(out)
    assert [l for l in out.read_text().splitlines() if l and not l.startswith("#")] == [HEADER.strip()]
    A.ClearTuning()


@pytest.mark.parametrize("body", [
    "1920 1080 0 20 48 48 1 1 1 0.1\n",          # not a compiled shape
    "1920 1080 0 20 32 32 1 1\n",                # short line
    "-5 1080 0 20 32 32 1 1 1 0.1\n",           # bad size
    "1920 1080 7 20 32 32 1 1 1 0.1\n",         # bad primitive kind
])
def test_rejects_bad_files(tmp_path, body):
    A.ClearTuning()
    good = tmp_path / "good.txt"
    good.write_text(HEADER + "640 480 0 10 32 16 1 2 3 0.5\n")
    A.LoadTuning(good)
    bad = tmp_path / "bad.txt"
    bad.write_text(HEADER + body)
    with pytest.raises(N.FglError) as e:
        A.LoadTuning(bad)
    assert e.value.code == N.FGL_E_INVALID
    # a rejected file changes nothing
    out = tmp_path / "out.txt"
    A.SaveTuning(out)
    rows = [l for l in out.read_text().splitlines() if l and not l.startswith("#")][1:]
    assert [r.split()[:6] for r in rows] == [["640", "480", "0", "10", "32", "16"]]
    A.ClearTuning()


def test_missing_and_foreign_files(tmp_path):
    with pytest.raises(N.FglError) as e:
        A.LoadTuning(tmp_path / "nope.txt")
    assert e.value.code == N.FGL_E_IO
    other = tmp_path / "other.txt"
    other.write_text("something else\n")
    with pytest.raises(N.FglError) as e:
        A.LoadTuning(other)
    assert e.value.code == N.FGL_E_INVALID
    with pytest.raises(N.FglError) as e:
        A.SaveTuning(os.path.join(tmp_path, "no_such_dir", "t.txt"))
    assert e.value.code == N.FGL_E_IO
