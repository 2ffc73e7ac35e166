"""GPU parity: the HIP path through the C ABI against the CPU restatement, pass by pass and over
whole trajectories.  Integer outputs (cell keys, counts, offsets, sorted order) must match bit
for bit; so must every float, because both sides pin the same binary32 arithmetic."""
import numpy as np
import pytest

from watercube_amd import Fluid, scenes
from watercube_amd import _lib as L

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _same_bits(a: np.ndarray, b: np.ndarray) -> bool:
    return a.tobytes() == b.tobytes()


def _first_diff(a: np.ndarray, b: np.ndarray, field: str) -> str:
    x, y = a[field], b[field]
    bad = np.nonzero((x.view(np.uint32) != y.view(np.uint32)).reshape(x.shape[0], -1).any(axis=1))[0]
    if bad.size == 0:
        return f"{field}: identical"
    i = bad[0]
    return f"{field}: {bad.size} records differ, first slot {i}: gpu={x[i]} oracle={y[i]}"


def _check_records(gpu, ref, what):
    if not _same_bits(gpu, ref):
        msgs = [_first_diff(gpu, ref, f) for f in ("position", "density", "velocity", "pressure")]
        pytest.fail(f"{what}: " + "; ".join(msgs))


def _run_sort_pass(p, parts, oracle):
    st = oracle.OracleState(p, parts)
    st.sort_run()
    with Fluid(p, parts) as f:
        f.sort_.run()
        keys = f.sort_.getCellKeys()
        counts = f.sort_.getCountBuffer()
        offsets = f.sort_.getOffsetBuffer()
        order = f.sort_.getSortedBuffer()
        buf2 = f.particles(2)
    return st, keys, counts, offsets, order, buf2


@pytest.mark.parametrize("name", ["dam_break_16k", "block_1m", "reference_default"])
def test_sort_pass_bit_exact(name, oracle):
    p, parts, dt, _ = scenes.workload(name)
    _log(f"sort pass {name} n={p.num_particles} R={p.grid_res}")
    st, keys, counts, offsets, order, buf2 = _run_sort_pass(p, parts, oracle)
    n = p.num_particles
    assert np.array_equal(keys, st.keys[:n]), "cell keys differ"
    assert np.array_equal(counts, st.counts), "counts differ"
    assert np.array_equal(offsets, st.offsets), "offsets differ"
    assert np.array_equal(order, st.sorted[:n]), "sorted order differs"
    _check_records(buf2, st.buf2, "reordered buffer")


def test_sort_of_shuffled_input_is_stable(oracle):
    """Input in random order (the first frame of a scene is not cell-sorted): the within-cell order
    must still be ascending input index, i.e. numpy's stable argsort of the cell keys."""
    p, parts, _, _ = scenes.workload("dam_break_16k")
    rng = np.random.default_rng(7)
    parts = parts[rng.permutation(parts.shape[0])]
    st, keys, counts, offsets, order, buf2 = _run_sort_pass(p, parts, oracle)
    assert np.array_equal(order, np.argsort(keys, kind="stable").astype(np.uint32))
    assert np.array_equal(order, st.sorted[: p.num_particles])
    _check_records(buf2, st.buf2, "reordered buffer")


@pytest.mark.parametrize("name", ["dam_break_16k", "block_1m", "reference_default"])
def test_density_and_update_pass_bit_exact(name, oracle):
    p, parts, dt, _ = scenes.workload(name)
    _log(f"density/update pass {name}")
    st = oracle.OracleState(p, parts)
    st.sort_run()
    st.density()
    with Fluid(p, parts) as f:
        f.sort_.run()
        f.runDensityProg()
        _check_records(f.particles(2), st.buf2, "density pass")
        st.update(dt)
        f.runUpdateProg(dt)
        _check_records(f.particles(1), st.buf1, "update pass")


@pytest.mark.parametrize("name,steps", [("dam_break_16k", 100), ("reference_default", 50)])
def test_trajectory_bit_exact(name, steps, oracle):
    p, parts, dt, _ = scenes.workload(name)
    _log(f"trajectory {name} {steps} steps")
    st = oracle.OracleState(p, parts)
    st.step(dt, steps)
    with Fluid(p, parts) as f:
        f.step(dt, steps)
        f.synchronize()
        _check_records(f.particles(1), st.buf1, f"{steps}-step trajectory")


def test_graph_step_equals_eager_passes(oracle):
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    with Fluid(p, parts) as a, Fluid(p, parts) as b:
        a.step(dt, 5)
        for _ in range(5):
            b.sort_.run()
            b.runDensityProg()
            b.runUpdateProg(dt)
        _check_records(a.particles(1), b.particles(1), "graph vs eager")
