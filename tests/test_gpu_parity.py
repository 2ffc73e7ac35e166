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


def test_every_tuning_variant_is_bit_identical(oracle):
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    st = oracle.OracleState(p, parts)
    st.step(dt, 5)
    nv = L.lib().wc_num_variants()
    for lists in (0, 1):
        for v in range(nv):
            _log(f"variant {v}: {L.variants()[v]}, neighbour lists {lists}")
            with Fluid(p, parts) as f:
                f.set_tuning(v, v, lists)
                f.step(dt, 5)
                _check_records(f.particles(1), st.buf1, f"variant {v} {L.variants()[v]} lists {lists}")


def test_autotuner_keeps_state_and_results(oracle):
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    st = oracle.OracleState(p, parts)
    st.step(dt, 4)
    with Fluid(p, parts) as f:
        f.step(dt, 2)
        before = f.particles(1)
        info = f.tune(dt, reps=2)
        _log(f"tuned: {info['density_variant']} {info['update_variant']}")
        assert f.particles(1).tobytes() == before.tobytes()
        assert all(x > 0 for x in info["density_ms"]) and all(x > 0 for x in info["update_ms"])
        assert all(x > 0 for x in info["density_list_ms"]) and all(x > 0 for x in info["update_list_ms"])
        assert f.get_tuning() == (info["density_variant"], info["update_variant"], info["neighbour_list"])
        f.step(dt, 2)
        _check_records(f.particles(1), st.buf1, "after tuning")


def _params_for(n, preset=L.WC_PRESET_DAM_BREAK_16K):
    p = L.preset(preset)
    p.num_particles = n
    return p


def _records(pos, vel=None):
    a = np.zeros(len(pos), L.PARTICLE_DTYPE)
    a["position"] = pos
    if vel is not None:
        a["velocity"] = vel
    return a


def _trajectory_case(p, parts, dt, steps, oracle, what):
    _log(f"case {what}: n={p.num_particles} steps={steps}")
    st = oracle.OracleState(p, parts)
    st.step(dt, steps)
    with Fluid(p, parts) as f:
        f.step(dt, steps)
        f.synchronize()
        _check_records(f.particles(1), st.buf1, what)
        if p.num_particles:
            assert np.array_equal(f.sort_.getCountBuffer(), st.counts)
            assert np.array_equal(f.sort_.getOffsetBuffer(), st.offsets)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 65, 257])
def test_tiny_scenes(n, oracle):
    rng = np.random.default_rng(n)
    parts = _records(rng.uniform(0.3, 0.36, (n, 3)).astype(np.float32))
    _trajectory_case(_params_for(n), parts, 2e-4, 10, oracle, f"tiny n={n}")


def test_particles_outside_the_cube(oracle):
    """Positions below 0, above size, and exactly on cell boundaries: the cell hash clamps and the
    boundary pass pulls them back, identically on both sides."""
    p = _params_for(0)
    b = float(p.bin_size)
    pos = [[-0.2, 0.5, 0.5], [1.3, 0.5, 0.5], [0.5, -1e-3, 2.0], [b, 2 * b, 3 * b], [b, 2 * b, 3 * b + 1e-7],
           [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.5, 0.5, 0.5], [0.5 + 0.01, 0.5, 0.5]]
    p.num_particles = len(pos)
    _trajectory_case(p, _records(np.float32(pos), np.float32([[3, -4, 5]] * len(pos))), 2e-4, 5, oracle, "outside")


def _check_records_nan_aware(gpu, ref, what):
    """NaN where the oracle has NaN (payload and sign may differ: the two sides' NaN-producing
    instructions pick different canonical NaNs), the same bits everywhere else."""
    for f in ("position", "density", "velocity", "pressure"):
        x, y = gpu[f], ref[f]
        nx, ny = np.isnan(x), np.isnan(y)
        assert np.array_equal(nx, ny), f"{what}: {f} NaN at {np.count_nonzero(nx)} vs {np.count_nonzero(ny)} places"
        assert np.array_equal(x[~nx].view(np.uint32), y[~ny].view(np.uint32)), f"{what}: {f} differs"


@pytest.mark.parametrize("lists", [0, 1])
def test_nan_and_infinite_inputs(lists, oracle):
    """A NaN position poisons the density and pressure of every particle within reach (a NaN
    distance passes the cutoff test, as `dist >= h` is false, density.comp:117); clamping the speed
    with min/max then drops the NaN from the velocity (update.comp:199). Infinite positions hash
    to the last cell and are pulled back by the wall clamp. Every step must agree with the oracle,
    and cell counts, offsets and order exactly."""
    rng = np.random.default_rng(5)
    n0 = 300
    pos = rng.uniform(0.01, 0.07, (n0, 3)).astype(np.float32)
    extra_pos = np.float32([[np.nan, 0.03, 0.03], [np.inf, 0.5, 0.5], [0.04, -np.inf, 0.04], [0.5, 0.5, 0.5],
                            [0.05, 0.05, 0.05]])
    extra_vel = np.float32([[0, 0, 0], [0, 0, 0], [0, 0, 0], [np.nan, 1, 1], [np.inf, -np.inf, 0]])
    parts = _records(np.concatenate([pos, extra_pos]))
    parts["velocity"][n0:] = extra_vel
    p = _params_for(parts.size)
    _log(f"case nan/inf inputs lists={lists}: n={parts.size}")
    st = oracle.OracleState(p, parts)
    with Fluid(p, parts) as f:
        f.set_tuning(0, 0, lists)
        for s in range(6):
            st.step(2e-4, 1)
            f.step(2e-4, 1)
            f.synchronize()
            _check_records_nan_aware(f.particles(1), st.buf1, f"step {s + 1}")
            assert np.array_equal(f.sort_.getCountBuffer(), st.counts)
            assert np.array_equal(f.sort_.getOffsetBuffer(), st.offsets)
            assert np.array_equal(f.sort_.getSortedBuffer(), st.sorted[: parts.size])
        assert np.isnan(st.buf1["density"]).sum() > 200  # the NaN did spread


def test_few_particles_on_a_fine_grid(oracle):
    """3,000 particles on the 250^3 grid of the 17.8M-particle scene: 15.6M cells, nearly all
    empty, so the scan and the cell ranges run over long empty stretches; two clusters and a
    few loners.  Bit-exact against the oracle, in both neighbour-list modes."""
    p = L.preset(L.WC_PRESET_BLOCK_SPHERE_17M)
    rng = np.random.default_rng(17)
    h = float(p.kernel_radius)
    pos = np.concatenate([rng.uniform(0.2, 0.2 + 6 * h, (1500, 3)), rng.uniform(0.7, 0.7 + 8 * h, (1490, 3)),
                          rng.uniform(0.0, 1.0, (10, 3))]).astype(np.float32)
    parts = _records(pos)
    p.num_particles = parts.size
    _log(f"case fine grid: n={parts.size} R={p.grid_res}")
    st = oracle.OracleState(p, parts)
    st.step(2.5e-5, 4)
    for lists in (0, 1):
        with Fluid(p, parts) as f:
            f.set_tuning(2, 2, lists)
            f.step(2.5e-5, 4)
            f.synchronize()
            _check_records(f.particles(1), st.buf1, f"fine grid lists={lists}")
            assert np.array_equal(f.sort_.getCountBuffer(), st.counts)
            assert np.array_equal(f.sort_.getOffsetBuffer(), st.offsets)


def test_dense_cluster_uses_fallback_path(oracle):
    """400 particles packed into a quarter-cell: every particle has ~399 neighbours, more than the
    neighbour list holds, so the force pass walks the cells again; results must not change."""
    p = _params_for(400)
    h = float(p.kernel_radius)
    rng = np.random.default_rng(9)
    parts = _records(rng.uniform(0.5, 0.5 + h / 4, (400, 3)).astype(np.float32))
    _trajectory_case(p, parts, 2e-5, 3, oracle, "dense cluster")
    # with neighbour lists: every list overflows (399 > 64) and each particle searches again
    st = oracle.OracleState(p, parts)
    st.step(2e-5, 3)
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, 1)
        f.step(2e-5, 3)
        _check_records(f.particles(1), st.buf1, "dense cluster, neighbour lists")


def _cell_box(p, cell, margin=0.05):
    """The inside of grid cell `cell` = (x, y, z), shrunk by `margin` of the cell size."""
    b = float(p.bin_size)
    lo = np.array(cell, np.float64) * b + margin * b
    return lo, lo + (1 - 2 * margin) * b


def _crowded_sort_check(p, parts, oracle, what):
    st, keys, counts, offsets, order, buf2 = _run_sort_pass(p, parts, oracle)
    n = p.num_particles
    assert np.array_equal(keys, st.keys[:n]), f"{what}: keys"
    assert np.array_equal(counts, st.counts) and np.array_equal(offsets, st.offsets), f"{what}: cells"
    assert np.array_equal(order, np.argsort(keys, kind="stable").astype(np.uint32)), f"{what}: not stable"
    assert np.array_equal(order, st.sorted[:n]), f"{what}: order"
    _check_records_nan_aware(buf2, st.buf2, f"{what}: sorted records")
    return counts


@pytest.mark.parametrize("lists", [0, 1])
def test_crowded_cells_keep_input_order(lists, oracle):
    """Cells of 3,000, 700 and 257 particles -- above the 256 the gather ranks directly, so their
    input indices are sorted by the crowded-cell path -- among 5,000 scattered particles, in
    shuffled input order: the sorted order must be numpy's stable argsort of the keys, and two
    whole steps (every crowded particle overflows its neighbour list) must match the oracle."""
    p = _params_for(0)
    rng = np.random.default_rng(31)
    pos = [rng.uniform(0.05, 0.95, (5000, 3))]
    for cell, m in (((16, 16, 16), 3000), ((4, 20, 9), 700), ((30, 2, 17), 257)):
        lo, hi = _cell_box(p, cell)
        pos.append(rng.uniform(lo, hi, (m, 3)))
    pos = np.concatenate(pos).astype(np.float32)
    parts = _records(pos[rng.permutation(len(pos))])
    p.num_particles = parts.size
    _log(f"crowded cells: n={parts.size}, lists {lists}")
    counts = _crowded_sort_check(p, parts, oracle, "crowded cells")
    assert sorted(counts)[-3:] == [257, 700, 3000]
    st = oracle.OracleState(p, parts)
    st.step(2e-5, 2)
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, lists)
        f.step(2e-5, 2)
        _check_records(f.particles(1), st.buf1, f"crowded cells, two steps, lists {lists}")


def test_very_crowded_cells_and_nan_positions(oracle):
    """20,000 particles in one cell (sorted in 4,096-entry chunks, then three merge passes) and
    12,000 NaN positions, which all hash to cell 0, among 3,000 others, in shuffled order: the sort
    pass bit for bit, then one whole step (NaNs in the same places, every other value the same
    bits)."""
    p = _params_for(0)
    rng = np.random.default_rng(47)
    lo, hi = _cell_box(p, (12, 25, 7))
    pos = np.concatenate([rng.uniform(lo, hi, (20000, 3)), rng.uniform(0.05, 0.95, (3000, 3)),
                          np.full((12000, 3), np.nan)]).astype(np.float32)
    parts = _records(pos[rng.permutation(len(pos))])
    p.num_particles = parts.size
    _log(f"very crowded cells: n={parts.size}")
    counts = _crowded_sort_check(p, parts, oracle, "very crowded cells")
    assert counts.max() == 20000 and counts[0] >= 12000
    st = oracle.OracleState(p, parts)
    st.step(2e-6, 1)
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, 1)
        f.step(2e-6, 1)
        f.synchronize()
        _check_records_nan_aware(f.particles(1), st.buf1, "very crowded cells, one step")


def test_mouse_force_trajectory(oracle):
    p, parts, dt, _ = scenes.workload("reference_default")
    p.mouse_force = 1
    p.camera_position[:] = (0.3, 0.2, -2.0)
    p.mouse_ray_direction[:] = (0.0, 0.05, 0.998749)
    _trajectory_case(p, parts, dt, 20, oracle, "mouse force")


def test_rotating_gravity_and_mouse_ray_between_steps(oracle):
    """Fluid::update sets gravity (rotated with the fluid, Fluid.cpp:259-263) and the mouse ray
    before every step.  Changing them between captured graph steps must take effect at once, as
    it does in the oracle when its parameters change between steps."""
    from watercube_amd.fluid import gravity_direction

    p, parts, dt, _ = scenes.workload("reference_default")
    st = oracle.OracleState(p.copy(), parts)
    quats = [(1.0, 0.0, 0.0, 0.0), (0.9238795, 0.3826834, 0.0, 0.0), (0.7071068, 0.0, 0.0, 0.7071068)]
    with Fluid(p, parts) as f:
        for i, q in enumerate(quats):
            f.updateGravity(q, 9.81)
            if i == 2:
                f.params.mouse_force = 1
                f.params.camera_position[:] = (0.3, 0.2, -2.0)
                f.params.mouse_ray_direction[:] = (0.0, 0.05, 0.998749)
                L.check(L.lib().wc_set_params(f.ctx, L.C.byref(f.params)), "wc_set_params")
            d = gravity_direction(q)
            for a in range(3):
                st.params.gravity[a] = float(np.float32(d[a]) * np.float32(9.81))
            if i == 2:
                st.params.mouse_force = 1
                st.params.camera_position[:] = (0.3, 0.2, -2.0)
                st.params.mouse_ray_direction[:] = (0.0, 0.05, 0.998749)
            f.step(dt, 4)
            st.step(dt, 4)
            f.synchronize()
            _check_records(f.particles(1), st.buf1, f"after orientation {i}")


def test_long_trajectory_dam_break(oracle):
    """1000 steps of the 16,384-particle dam break: bit-identical, so the relative error bound of
    1e-3 after 1000 steps holds with error 0."""
    p, parts, dt, steps = scenes.workload("dam_break_16k")
    _trajectory_case(p, parts, dt, steps, oracle, "dam break 1000 steps")


def test_block_1m_properties_after_run(oracle):
    """Full-size invariants after 200 steps of the 1,000,000-particle block: counts sum to N,
    offsets are their exclusive scan, the sorted order is a permutation whose cells ascend, every
    slot range holds only its cell, all values finite and inside the cube."""
    p, parts, dt, _ = scenes.workload("block_1m")
    with Fluid(p, parts) as f:
        f.step(dt, 200)
        f.sort_.run()
        f.synchronize()
        counts = f.sort_.getCountBuffer()
        offsets = f.sort_.getOffsetBuffer()
        order = f.sort_.getSortedBuffer()
        keys_in = f.sort_.getCellKeys()
        buf1 = f.particles(1)
        buf2 = f.particles(2)
    n = p.num_particles
    assert counts.sum() == n
    assert np.array_equal(offsets, np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32))
    assert np.array_equal(np.sort(order), np.arange(n, dtype=np.uint32))
    k_sorted = keys_in[order]
    assert np.all(np.diff(k_sorted.astype(np.int64)) >= 0)
    assert np.array_equal(np.repeat(np.arange(len(counts)), counts), k_sorted)
    # stable: inside a cell, input order ascends
    same = np.diff(k_sorted.astype(np.int64)) == 0
    assert np.all(np.diff(order.astype(np.int64))[same] > 0)
    assert buf2.tobytes() == buf1[order].tobytes()
    pos = buf1["position"]
    assert np.isfinite(pos).all() and (pos >= p.border).all() and (pos <= 1 - p.border).all()
    # and one more step from this state matches the CPU restatement bit for bit
    st = oracle.OracleState(p, buf1)
    st.step(dt, 1)
    with Fluid(p, buf1) as g:
        g.step(dt, 1)
        _check_records(g.particles(1), st.buf1, "block_1m step from step-200 state")


def test_cpp_harness_matches_oracle(tmp_path, oracle):
    import subprocess
    from pathlib import Path

    exe = Path(L.__file__).resolve().parent / "wc_harness"
    out = tmp_path / "state.bin"
    r = subprocess.run([str(exe), "dam_break_16k", "100", "--out", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    _log(r.stdout.strip())
    got = np.fromfile(out, L.PARTICLE_DTYPE)
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    st = oracle.OracleState(p, parts)
    st.step(dt, 100)
    _check_records(got, st.buf1, "C++ harness 100 steps")


def test_block_sphere_17m_bit_exact_and_invariants(oracle):
    """The 17,777,216-particle scene (256^3 block plus a falling 1M ball; R = 250, 15.6 M cells):
    the grid build at scale (look-back scan over 3,815 tiles) and one whole step bit-exact
    against the CPU restatement; then, after the rest of the workload's 200 steps on the GPU, the
    cell invariants of the next sort and one further step bit-exact again."""
    p, parts, dt, _ = scenes.workload("block_sphere_17m")
    n = int(p.num_particles)
    _log(f"block_sphere_17m n={n} R={p.grid_res}: sort + one step")
    st = oracle.OracleState(p, parts)
    st.sort_run()
    with Fluid(p, parts) as f:
        f.sort_.run()
        assert np.array_equal(f.sort_.getOffsetBuffer(), st.offsets), "offsets differ"
        assert np.array_equal(f.sort_.getSortedBuffer(), st.sorted[:n]), "sorted order differs"
        f.runDensityProg()
        st.density()
        _check_records(f.particles(2), st.buf2, "17m density pass")
        f.runUpdateProg(dt)
        st.update(dt)
        _check_records(f.particles(1), st.buf1, "17m update pass")
        _log("block_sphere_17m: 199 more steps")
        f.step(dt, 199)
        f.sort_.run()
        f.synchronize()
        counts = f.sort_.getCountBuffer()
        offsets = f.sort_.getOffsetBuffer()
        order = f.sort_.getSortedBuffer()
        keys_in = f.sort_.getCellKeys()
        buf1 = f.particles(1)
    assert int(counts.sum()) == n
    assert np.array_equal(offsets[1:], np.cumsum(counts, dtype=np.uint64)[:-1].astype(np.uint32))
    k_sorted = keys_in[order].astype(np.int64)
    assert np.all(np.diff(k_sorted) >= 0)
    same = np.diff(k_sorted) == 0
    assert np.all(np.diff(order.astype(np.int64))[same] > 0)
    pos = buf1["position"]
    assert np.isfinite(pos).all() and (pos >= p.border).all() and (pos <= p.size - p.border).all()
    _log("block_sphere_17m: one more step against the oracle")
    st = oracle.OracleState(p, buf1)
    st.step(dt, 1)
    with Fluid(p, buf1) as g:
        g.step(dt, 1)
        _check_records(g.particles(1), st.buf1, "17m step from the step-200 state")


def test_resume_from_snapshot_equals_uninterrupted_run(tmp_path, oracle):
    """40 steps, snapshot, 40 more steps from the snapshot (Python and the C++ harness) equal 80
    uninterrupted steps of the CPU restatement bit for bit."""
    import subprocess
    from pathlib import Path

    from watercube_amd import snapshot

    p, parts, dt, _ = scenes.workload("dam_break_16k")
    st = oracle.OracleState(p, parts)
    st.step(dt, 80)
    with Fluid(p, parts) as f:
        f.step(dt, 40)
        snapshot.save(tmp_path / "a", p, f.particles(1), step=40, dt=dt)
    q, rec, meta = snapshot.load(tmp_path / "a")
    with Fluid(q, rec) as g:
        g.step(meta["dt"], 40)
        _check_records(g.particles(1), st.buf1, "resumed (Python)")
    exe = Path(L.__file__).resolve().parent / "wc_harness"
    _log("harness resume from a snapshot")
    r = subprocess.run([str(exe), "dam_break_16k", "40", "--in", str(tmp_path / "a.bin"), "--out",
                        str(tmp_path / "b")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    q2, rec2, meta2 = snapshot.load(tmp_path / "b")
    assert meta2["step"] == 80 and bytes(q2) == bytes(q)
    _check_records(rec2, st.buf1, "resumed (C++ harness)")


def test_neighbour_lists_hold_every_accepted_neighbour(oracle):
    """With neighbour lists, the count stored for each sorted particle equals the number of other
    particles with !(d2 >= cut2) -- the density pass's acceptance test, in the same float32
    arithmetic -- counted by brute force over all pairs (2,048 particles, every one inside the
    grid, so the 27 cells hold every neighbour)."""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
    import make_golden as G

    d = G.load("dam_break_2k")
    p = G.params_from_fixture(d["params"])
    with Fluid(p, d["input"]) as f:
        f.set_tuning(2, 2, 1)
        f.sort_.run()
        f.runDensityProg()
        got = f.neighbour_counts()
        pos = f.particles(2)["position"]
    f32 = np.float32
    h = f32(p.kernel_radius)
    cut2 = f32(h * h)
    while np.sqrt(np.nextafter(cut2, f32(0))) >= h:
        cut2 = np.nextafter(cut2, f32(0))
    while np.sqrt(cut2) < h:
        cut2 = np.nextafter(cut2, f32(np.inf))
    want = np.empty(len(pos), np.uint32)
    for i in range(len(pos)):
        r = pos[i] - pos  # float32
        d2 = (r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2]
        ok = ~(d2 >= cut2)
        ok[i] = False
        want[i] = ok.sum()
    want = np.where(want > 64, 0xFFFFFFFF, want).astype(np.uint32)
    assert np.array_equal(got, want)
    assert 20 < float(want[want != 0xFFFFFFFF].mean()) < 40


def test_graph_replays_interleaved_with_eager_work(oracle):
    """The captured step is replayed many times with eager passes and host copies between the
    replays; every replay's result is checked against the CPU restatement (no replay may see
    stale scan state, counters or clocks from the one before)."""
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    st = oracle.OracleState(p, parts)
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, 1)
        for r in range(6):
            f.step(dt, 1)          # graph replay
            st.step(dt, 1)
            _check_records(f.particles(1), st.buf1, f"replay {r}")
            f.sort_.run()          # eager work on the same stream between replays
            f.runDensityProg()
            f.synchronize()
            # the eager sort of buffer 1 is the sort the next step will do
            nxt = oracle.OracleState(p, st.buf1)
            nxt.sort_run()
            assert np.array_equal(f.sort_.getOffsetBuffer(), nxt.offsets), f"eager sort after replay {r}"
        f.step(dt, 3)
        st.step(dt, 3)
        _check_records(f.particles(1), st.buf1, "three more replays")


@pytest.mark.parametrize("lists", [0, 1])
def test_force_pass_division_modes(lists, oracle):
    """The force pass shares one reciprocal between a neighbour's quotients only when the density
    pass found every operand 0 or within [2^-40, 2^40]; otherwise it divides as IEEE does.  Both
    paths must give the oracle's bits: an in-range scene with velocity components of -0 (whose
    quotients' zero sign differs between the paths), and the same scene with one velocity
    component of 1e-30, which switches the whole step to IEEE division."""
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    neg0 = parts.copy()
    neg0["velocity"][::2, 0] = np.float32(-0.0)
    neg0["velocity"][1::3, 2] = np.float32(-0.0)
    tiny = neg0.copy()
    tiny["velocity"][100, 1] = np.float32(1e-30)
    for what, state in (("-0 velocities", neg0), ("a 1e-30 velocity", tiny)):
        _log(f"division modes: {what}, lists {lists}")
        st = oracle.OracleState(p, state)
        st.step(dt, 20)
        with Fluid(p, state) as f:
            f.set_tuning(2, 2, lists)
            f.step(dt, 20)
            _check_records(f.particles(1), st.buf1, f"{what}, lists {lists}")


@pytest.mark.parametrize("lists", [0, 1])
def test_force_pass_right_after_a_sort(lists, oracle):
    """A force pass straight after a sort, with no density pass between them (the densities are
    then the ones the records carried in): it must not rely on the last density pass's operand
    check or neighbour lists, which describe an older sorted buffer, and so divides as IEEE does
    and searches for its neighbours."""
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    st = oracle.OracleState(p, parts)
    st.step(dt, 3)
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, lists)
        f.step(dt, 3)  # leaves a density pass's state behind
        # move the particles, then sort and update only
        moved = f.particles(1)
        moved["position"][:, 0] += np.float32(0.001)
        stm = oracle.OracleState(p, moved)
        f.upload(moved)
        f.sort_.run()
        f.runUpdateProg(dt)
        stm.sort_run()
        stm.update(dt)
        _check_records(f.particles(1), stm.buf1, "update after sort without density")


def _oracle_after_upload(oracle, p, st_prev, state):
    """The oracle's view of an upload: both buffers hold the new records (Fluid.cpp:146, 155),
    counts and offsets are still those of the last sort."""
    st = oracle.OracleState(p, state)
    st.counts[:] = st_prev.counts
    st.offsets[:] = st_prev.offsets
    return st


@pytest.mark.parametrize("lists", [0, 1])
def test_update_right_after_an_upload(lists, oracle):
    """runUpdateProg straight after new records were uploaded, with no sort or density pass in
    between: the pass must find each particle's cell from its new position, as update.comp:141
    does, with the last sort's cell ranges, and must neither read the neighbour lists of the
    last density pass (they describe the old positions) nor trust its division-operand check.
    Every particle moves by its own random offset, so cells and neighbour sets change."""
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    rng = np.random.default_rng(11)
    st = oracle.OracleState(p, parts)
    st.step(dt, 3)
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, lists)
        f.step(dt, 3)
        moved = f.particles(1)
        _check_records(moved, st.buf1, "before the upload")
        moved["position"] += rng.uniform(-0.6, 0.6, moved["position"].shape).astype(np.float32) * np.float32(
            p.bin_size)
        moved["velocity"] += rng.uniform(-0.5, 0.5, moved["velocity"].shape).astype(np.float32)
        _log(f"update after an upload, lists {lists}")
        ref = _oracle_after_upload(oracle, p, st, moved)
        f.upload(moved)
        f.runUpdateProg(dt)
        ref.update(dt)
        _check_records(f.particles(1), ref.buf1, f"update after an upload, lists {lists}")


@pytest.mark.parametrize("lists", [0, 1])
def test_density_and_update_after_an_upload_with_a_tiny_velocity(lists, oracle):
    """An upload whose records carry a velocity of 1e-35, then runDensityProg and runUpdateProg
    without a sort: the density pass checks positions, densities and pressures but not
    velocities, so the upload itself must leave the force pass dividing as IEEE does."""
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    st = oracle.OracleState(p, parts)
    st.step(dt, 2)
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, lists)
        f.step(dt, 2)
        state = f.particles(1)
        state["velocity"][77, 0] = np.float32(1e-35)
        state["velocity"][78, 2] = np.float32(-3e-38)
        _log(f"density + update after an upload with a tiny velocity, lists {lists}")
        ref = _oracle_after_upload(oracle, p, st, state)
        f.upload(state)
        f.runDensityProg()
        ref.density()
        _check_records(f.particles(2), ref.buf2, f"density after the upload, lists {lists}")
        f.runUpdateProg(dt)
        ref.update(dt)
        _check_records(f.particles(1), ref.buf1, f"update after the upload, lists {lists}")


PARAM_CASES = {
    # densities near 1e-13, below the 2^-40 operand range: the force pass divides as IEEE does
    "tiny mass": dict(particle_mass=1e-16),
    # every pressure 0: the pressure term never applies, quotients of 0
    "zero pressure": dict(stiffness=0.0, rest_pressure=0.0),
    "negative rest pressure": dict(rest_pressure=-50.0),
    "no viscosity": dict(viscosity=0.0),
    "linear eos, stiff": dict(eos=L.WC_EOS_LINEAR, stiffness=1e6),
    # pressures above 2^40: IEEE division again
    "huge stiffness": dict(stiffness=1e13),
    "no gravity, no walls": dict(gravity=(0.0, 0.0, 0.0), wall_density=0, wall_force=0),
}


@pytest.mark.parametrize("case", sorted(PARAM_CASES))
def test_parameter_extremes_bit_exact(case, oracle):
    """Unusual constants move densities, pressures and forces across the force pass's division
    guard in both directions; every combination must still give the oracle's bits, with and
    without neighbour lists."""
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    for k, v in PARAM_CASES[case].items():
        if isinstance(v, tuple):
            for i, x in enumerate(v):
                getattr(p, k)[i] = x
        else:
            setattr(p, k, v)
    st = oracle.OracleState(p, parts)
    st.step(dt, 10)
    for lists in (0, 1):
        _log(f"parameter case {case}, lists {lists}")
        with Fluid(p, parts) as f:
            f.set_tuning(2, 2, lists)
            f.step(dt, 10)
            _check_records(f.particles(1), st.buf1, f"{case}, lists {lists}")


def test_advect_pass_bit_exact(oracle):
    """Fluid::runAdvectProg (advect.comp:20-59) on buffer 1, after a few plain steps and with
    velocities that push particles through every wall."""
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    rng = np.random.default_rng(3)
    st = oracle.OracleState(p, parts)
    st.step(dt, 5)
    with Fluid(p, parts) as f:
        f.step(dt, 5)
        state = f.particles(1)
        _check_records(state, st.buf1, "before advect")
        state["velocity"] = rng.uniform(-60.0, 60.0, (state.shape[0], 3)).astype(np.float32)
        f.upload(state)
        ref = oracle.OracleState(p, state)
        for _ in range(3):
            f.runAdvectProg(2e-3)
            ref.advect(2e-3)
        _check_records(f.particles(1), ref.buf1, "advect")
