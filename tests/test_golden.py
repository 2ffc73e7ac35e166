"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

Two kinds of check, neither of which retypes the oracle's formulas:
* the fixtures satisfy what the reference's grid build fixes independently of any arithmetic
  choice -- the cell key closed form of count.comp:31-33 recomputed in numpy, counts equal to
  the key histogram, offsets equal to its exclusive scan (linearScan.comp:18-27), a sorted order
  that is a stable sort by key (reorder.comp:32-45 with the order defined as previous-slot), the
  sorted records equal to the gathered input, the density pass touching only density/pressure
  (density.comp:126-135) and the force pass passing them through (update.comp:229-231);
* the CPU restatement still reproduces every array bit for bit (a regression pin).
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden as G  # noqa: E402

CASES = sorted(json.loads((GOLDEN / "manifest.json").read_text())["cases"])


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("name", CASES)
def test_manifest_digests(name):
    m = json.loads((GOLDEN / "manifest.json").read_text())["cases"][name]
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        for k, want in m["sha256_16"].items():
            got = hashlib.sha256(_bits(z[k]).tobytes()).hexdigest()[:16]
            assert got == want, k


@pytest.mark.parametrize("name", CASES)
def test_grid_build_invariants(name):
    d = G.load(name)
    p = G.params_from_fixture(d["params"])
    n, R = int(p.num_particles), int(p.grid_res)
    pos = d["input"]["position"]
    # count.comp:31-33: clamp(int(p / binSize), 0, R-1); x fastest
    c = np.clip(np.trunc(pos / np.float32(p.bin_size)), 0, R - 1).astype(np.int64)
    keys = (c[:, 2] * R + c[:, 1]) * R + c[:, 0]
    assert np.array_equal(d["keys"], keys.astype(np.uint32))
    assert np.array_equal(d["counts"], np.bincount(keys, minlength=R ** 3).astype(np.uint32))
    assert np.array_equal(d["offsets"][1:], np.cumsum(d["counts"])[:-1].astype(np.uint32))
    assert d["offsets"][0] == 0 and int(d["counts"].sum()) == n
    order = d["sorted"].astype(np.int64)
    assert np.array_equal(np.sort(order), np.arange(n))
    assert np.array_equal(order, np.argsort(keys, kind="stable"))
    assert _bits(d["sorted_records"]).tobytes() == _bits(d["input"][order]).tobytes()
    # every slot range of a cell holds exactly that cell's particles
    for cell in np.nonzero(d["counts"])[0][:50]:
        lo, hi = d["offsets"][cell], d["offsets"][cell] + d["counts"][cell]
        assert np.all(keys[order[lo:hi]] == cell)


@pytest.mark.parametrize("name", CASES)
def test_pass_outputs_touch_only_their_fields(name):
    d = G.load(name)
    s, a, u = d["sorted_records"], d["after_density"], d["after_update"]
    for f in ("position", "velocity"):
        assert _bits(a[f]).tobytes() == _bits(s[f]).tobytes()
    for f in ("density", "pressure"):
        assert _bits(u[f]).tobytes() == _bits(a[f]).tobytes()
    assert np.all(np.isfinite(u["position"])) and np.all(a["density"] > 0)
    p = G.params_from_fixture(d["params"])
    lo, hi = np.float32(p.border), np.float32(p.size) - np.float32(p.border)
    for rec in (u, d["trajectory"]):
        assert np.all(rec["position"] >= lo) and np.all(rec["position"] <= hi)
        assert np.all(np.abs(rec["velocity"]) <= np.float32(p.max_speed))


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name):
    d = G.load(name)
    p = G.params_from_fixture(d["params"])
    got = G.run_case(p, d["input"], float(d["dt"][0]), int(d["trajectory_steps"][0]))
    for k in ("keys", "counts", "offsets", "sorted", "sorted_records", "after_density", "after_update",
              "trajectory"):
        assert _bits(got[k]).tobytes() == _bits(d[k]).tobytes(), k
