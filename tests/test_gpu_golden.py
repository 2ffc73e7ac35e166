"""The GPU path, called through the C ABI, reproduces the golden fixtures bit for bit: every
pass of one step (cell keys, counts, offsets, stable order, sorted records, density/pressure,
force + integration) and the records after the fixture's whole-step trajectory."""
import sys
from pathlib import Path

import numpy as np
import pytest

from watercube_amd import Fluid

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden as G  # noqa: E402

pytestmark = pytest.mark.gpu
CASES = ["dam_break_2k", "reference_default_2k", "mouse_ray_2k"]


def _same(a, b):
    return np.ascontiguousarray(a).view(np.uint8).tobytes() == np.ascontiguousarray(b).view(np.uint8).tobytes()


@pytest.mark.parametrize("name", CASES)
def test_gpu_reproduces_fixture(name):
    d = G.load(name)
    p = G.params_from_fixture(d["params"])
    dt = float(d["dt"][0])
    print(f"golden case {name}: {p.num_particles} particles, R={p.grid_res}", flush=True)
    with Fluid(p, d["input"]) as f:
        f.sort_.run()
        assert _same(f.sort_.getCellKeys(), d["keys"])
        assert _same(f.sort_.getCountBuffer(), d["counts"])
        assert _same(f.sort_.getOffsetBuffer(), d["offsets"])
        assert _same(f.sort_.getSortedBuffer(), d["sorted"])
        assert _same(f.particles(2), d["sorted_records"])
        f.runDensityProg()
        assert _same(f.particles(2), d["after_density"])
        f.runUpdateProg(dt)
        assert _same(f.particles(1), d["after_update"])
        f.step(dt, int(d["trajectory_steps"][0]) - 1)
        assert _same(f.particles(1), d["trajectory"])
