"""GPU checks of the learned closure: network kernels against the CPU restatement
(oracle/mlp_oracle.py), the feature kernel against a numpy 32-nearest-neighbour reference, and
the closure step's plain limits (zero network = plain large steps; a judge that always
rejects = the plain small-dt trajectory), bit for bit."""
import ctypes as C

import numpy as np
import pytest

from oracle import mlp_oracle as MO
from watercube_amd import Fluid, scenes
from watercube_amd import _lib as L
from watercube_amd.closure import ClosureNet, DeviceArray

pytestmark = pytest.mark.gpu


def _bf16_bits(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _random_batch(rng, m):
    x = np.zeros((m, 208), np.float32)
    x[:, :195] = rng.normal(0, 1, (m, 195))
    xb = _bf16_bits(x)
    t = np.zeros((m, 4), np.float32)
    t[:, :3] = rng.normal(0, 0.5, (m, 3))
    return xb, t


def test_network_forward_and_gradient_match_cpu():
    rng = np.random.default_rng(0)
    m = 300
    net = ClosureNet(max_batch=m, seed=3)
    p = net.params().astype(np.float64)
    p[MO.layout()["b1"][0]:MO.layout()["b1"][0] + 128] = rng.normal(0, 0.1, 128)  # nonzero biases too
    net.set_params(p.astype(np.float32))
    xb, t = _random_batch(rng, m)
    xd = DeviceArray((m, 208), np.uint16).upload(xb)
    td = DeviceArray((m, 4), np.float32).upload(t)
    yd = DeviceArray((m, 4), np.float32)
    X = MO.bf16_bits_to_f64(xb)
    y_ref, loss_ref, g_ref = MO.forward_backward(p.astype(np.float32).astype(np.float64), X, t[:, :3], bf16=True)
    net.forward(xd, yd, m)
    y_gemm = yd.download()[:, :3]
    net.predict(xd, yd, m)
    y_fused = yd.download()[:, :3]
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"forward rel err gemm {rel(y_gemm, y_ref):.2e} fused {rel(y_fused, y_ref):.2e}", flush=True)
    assert rel(y_gemm, y_ref) < 1e-2
    assert rel(y_fused, y_ref) < 1e-2
    g, loss = net.gradient(xd, td, m)
    assert abs(loss - loss_ref) / loss_ref < 1e-2
    Lo = MO.layout()
    for k in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4"):
        off, n = Lo[k]
        e = rel(g[off:off + n], g_ref[off:off + n])
        print(f"grad {k}: rel err {e:.2e}", flush=True)
        assert e < 3e-2, k
    # padding receives no gradient
    w1 = g[Lo["w1"][0]:Lo["w1"][0] + Lo["w1"][1]].reshape(208, 128)
    assert np.all(w1[195:] == 0)


def test_training_fits_a_teacher_network():
    rng = np.random.default_rng(1)
    m = 4096
    teacher = ClosureNet(max_batch=m, seed=11)
    student = ClosureNet(max_batch=m, seed=12)
    xb, t = _random_batch(rng, m)
    xd = DeviceArray((m, 208), np.uint16).upload(xb)
    yd = DeviceArray((m, 4), np.float32)
    teacher.forward(xd, yd, m)
    td = DeviceArray((m, 4), np.float32).upload(yd.download())
    _, first = student.gradient(xd, td, m)
    for i in range(300):
        student.train_step(xd.ptr, td.ptr, m, 3e-3)
    _, last = student.gradient(xd, td, m)
    print(f"teacher fit: loss {first:.4f} -> {last:.5f}", flush=True)
    assert last < 0.1 * first


def test_timed_training_steps_are_training_steps():
    """wc_mlp_profile_train runs one warm-up step and `reps` timed steps of the same kernels as
    wc_mlp_train_step: its parameters afterwards equal those of reps + 1 plain training steps.  A
    batch under 1,024 rows keeps every GEMM unsplit, so each gradient is summed in one order."""
    rng = np.random.default_rng(2)
    m = 1000
    a = ClosureNet(max_batch=m, seed=5)
    b = ClosureNet(max_batch=m, seed=5)
    xb, t = _random_batch(rng, m)
    xd = DeviceArray((m, 208), np.uint16).upload(xb)
    td = DeviceArray((m, 4), np.float32).upload(t)
    ms = a.profile_train(xd, td, m, lr=1e-3, reps=3)
    for _ in range(4):
        b.train_step(xd.ptr, td.ptr, m, 1e-3)
    pa, pb = a.params(), b.params()
    print(f"timed training step: {ms:.4f} ms at batch {m}", flush=True)
    assert ms > 0
    np.testing.assert_array_equal(pa, pb)
    assert np.abs(pa - ClosureNet(max_batch=m, seed=5).params()).max() > 1e-4


def _enumeration(oracle, p, pos, counts, offsets, t):
    R = p.grid_res
    cx, cy, cz = oracle.cell(p, pos[t])
    out = []
    for b in range(27):
        nx, ny, nz = cx + b // 9 - 1, cy + (b // 3) % 3 - 1, cz + b % 3 - 1
        if min(nx, ny, nz) < 0 or max(nx, ny, nz) >= R:
            continue
        k = (nz * R + ny) * R + nx
        out += list(range(offsets[k], offsets[k] + counts[k]))
    return out


@pytest.mark.parametrize("lists", [0, 1])
@pytest.mark.parametrize("scene", ["dam_break_16k", "dense_h3"])
def test_features_match_numpy_reference(scene, lists, oracle):
    """dam_break_16k has about 33 neighbours inside h per particle (the 32-of-n selection at the
    end of the walk); dense_h3 (h = 3 spacings) about 113, so the kernel's list also overflows
    and is cut back to the 32 nearest during the walk."""
    if scene == "dam_break_16k":
        p, parts, dt, _ = scenes.workload("dam_break_16k")
    else:
        p, parts = scenes.scaled_block((24, 24, 24), 0.01, seed=5, h_over_s=3.0)
        dt = 2e-4
    K = 4.0
    print(f"features {scene}: {p.num_particles} particles, neighbour lists {lists}", flush=True)
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, lists)
        f.step(dt, 30)
        f.sort_.run()
        f.runDensityProg()
        s = f.particles(2)
        counts, offsets = f.sort_.getCountBuffer(), f.sort_.getOffsetBuffer()
        rows = np.arange(0, p.num_particles, 37, dtype=np.uint32)
        xd = DeviceArray((len(rows), 208), np.uint16)
        L.check(L.lib().wc_closure_features(f.ctx, C.c_float(K), L.u32ptr(rows), len(rows), xd.ptr), "features")
        got = xd.download()
    f32 = np.float32
    h = f32(p.kernel_radius)
    cut2 = f32(h * h)
    while np.sqrt(np.nextafter(cut2, f32(0))) >= h:
        cut2 = np.nextafter(cut2, f32(0))
    while np.sqrt(cut2) < h:
        cut2 = np.nextafter(cut2, f32(np.inf))
    inv_h = f32(1.0) / h
    pos, vel, rho = s["position"], s["velocity"], s["density"]
    for r, t in enumerate(rows):
        cand = []
        for e, u in enumerate(_enumeration(oracle, p, pos, counts, offsets, t)):
            if u == t:
                continue
            d = pos[t] - pos[u]
            d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
            if not (d2 >= cut2):
                cand.append((d2, e, u))
        # the 32 nearest (ties: earlier in enumeration order), listed in enumeration order
        cand.sort(key=lambda c: (c[0].view(np.uint32), c[1]))
        cand = sorted(cand[:32], key=lambda c: c[1])
        want = np.zeros(208, np.float32)
        for k, (_, _, u) in enumerate(cand):
            want[6 * k:6 * k + 3] = (pos[u] - pos[t]) * inv_h
            want[6 * k + 3:6 * k + 6] = (vel[u] - vel[t]) * f32(1.0)
        want[192] = (rho[t] / f32(p.rest_density) - f32(1.0)) * f32(10.0)
        want[193] = f32(K) * f32(0.25)
        want[194] = f32(len(cand)) * f32(1.0 / 32)
        assert np.array_equal(got[r], _bf16_bits(want)), f"row {r} (slot {t}) differs"


@pytest.mark.parametrize("scene", ["dam_break_16k", "dense_h3", "h23", "ragged"])
def test_inference_from_lists_equals_features_then_network(scene):
    """wc_closure_infer forms each row's values inside the inference kernel (short lists read
    straight, longer ones through the selection's slots); it must give the bits of
    wc_closure_features followed by wc_mlp_predict.  dense_h3's lists all overflow; h23 (h = 2.3
    spacings) lists about 50 neighbours, so most waves select by bisection; ragged has 9,177
    particles, so the last 128-row tile is partial."""
    if scene == "dam_break_16k":
        p, parts, dt, _ = scenes.workload("dam_break_16k")
    elif scene == "dense_h3":
        p, parts = scenes.scaled_block((24, 24, 24), 0.01, seed=5, h_over_s=3.0)
        dt = 2e-4
    elif scene == "h23":
        p, parts = scenes.scaled_block((24, 24, 24), 0.01, seed=9, h_over_s=2.3)
        dt = 2e-4
    else:
        p, parts = scenes.scaled_block((23, 21, 19), 0.01, seed=7, h_over_s=2.0)
        dt = 2e-4
    n = p.num_particles
    print(f"closure inference from lists {scene}: {n} particles", flush=True)
    net = ClosureNet(max_batch=64)
    net.set_params((np.random.default_rng(12).normal(0, 0.05, ClosureNet.num_params())).astype(np.float32))
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, 1)
        f.step(dt, 30)
        f.sort_.run()
        f.runDensityProg()
        x = DeviceArray((n, 208), np.uint16)
        y1 = DeviceArray((n, 4), np.float32)
        y2 = DeviceArray((n, 4), np.float32)
        L.check(L.lib().wc_closure_infer(f.ctx, net.h, C.c_float(4.0), y1.ptr), "wc_closure_infer")
        L.check(L.lib().wc_closure_features(f.ctx, C.c_float(4.0), None, 0, x.ptr), "wc_closure_features")
        L.check(L.lib().wc_mlp_predict(net.h, x.ptr, y2.ptr, n), "wc_mlp_predict")
        a, b = y1.download(), y2.download()
    bad = np.nonzero((a.view(np.uint32) != b.view(np.uint32)).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}"
    assert np.abs(a).max() > 0


def test_lattice_ties_select_like_the_walk(oracle):
    """An unjittered lattice at exactly representable coordinates (spacing 1/64, h = 2.3
    spacings): squared distances come in shells of exactly equal values, so the 32 nearest of a
    list of 34 or 38 (the per-lane selection) or of 41-56 (the wave's) are decided by the tie rule,
    the earlier entry in enumeration order kept.  The fused inference's selection must give the
    bits of the walk's features followed by the network, and the walk's rows must equal the numpy
    reference's on every row that selects."""
    p, _ = scenes.scaled_block((14, 13, 12), 1.0 / 64, seed=0, h_over_s=2.3)
    parts = scenes.block(14, 13, 12, 1.0 / 64, 0.0, 0, origin=(0.25, 0.25, 0.25))
    n = p.num_particles
    net = ClosureNet(max_batch=64)
    net.set_params((np.random.default_rng(13).normal(0, 0.05, ClosureNet.num_params())).astype(np.float32))
    K = 4.0
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, 1)
        f.sort_.run()
        f.runDensityProg()
        listed = f.neighbour_counts()
        x = DeviceArray((n, 208), np.uint16)
        y1 = DeviceArray((n, 4), np.float32)
        y2 = DeviceArray((n, 4), np.float32)
        L.check(L.lib().wc_closure_infer(f.ctx, net.h, C.c_float(K), y1.ptr), "wc_closure_infer")
        L.check(L.lib().wc_closure_features(f.ctx, C.c_float(K), None, 0, x.ptr), "wc_closure_features")
        L.check(L.lib().wc_mlp_predict(net.h, x.ptr, y2.ptr, n), "wc_mlp_predict")
        a, b, got = y1.download(), y2.download(), x.download()
        s = f.particles(2)
        counts, offsets = f.sort_.getCountBuffer(), f.sort_.getOffsetBuffer()
    kinds = {m: int((listed == m).sum()) for m in np.unique(listed)}
    print(f"lattice: {n} particles, listed neighbours {kinds}", flush=True)
    assert ((listed > 32) & (listed <= 40)).sum() > 0 and ((listed > 40) & (listed <= 64)).sum() > 0
    bad = np.nonzero((a.view(np.uint32) != b.view(np.uint32)).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}"
    f32 = np.float32
    h = f32(p.kernel_radius)
    inv_h = f32(1.0) / h
    pos, rho = s["position"], s["density"]
    for t in np.nonzero(listed > 32)[0]:
        cand = []
        for e, u in enumerate(_enumeration(oracle, p, pos, counts, offsets, t)):
            d = pos[t] - pos[u]
            d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
            if u != t and np.sqrt(d2) < h:
                cand.append((d2, e, u))
        assert len(cand) == listed[t]
        cand.sort(key=lambda c: (c[0].view(np.uint32), c[1]))
        cand = sorted(cand[:32], key=lambda c: c[1])
        want = np.zeros(208, np.float32)
        for k, (_, _, u) in enumerate(cand):
            want[6 * k:6 * k + 3] = (pos[u] - pos[t]) * inv_h
        want[192] = (rho[t] / f32(p.rest_density) - f32(1.0)) * f32(10.0)
        want[193] = f32(K) * f32(0.25)
        want[194] = f32(len(cand)) * f32(1.0 / 32)
        assert np.array_equal(got[t], _bf16_bits(want)), f"slot {t} ({listed[t]} listed) differs"


@pytest.mark.parametrize("scene", ["dam_break_16k", "ragged"])
def test_closure_run_identical_with_and_without_lists(scene):
    """A learned run with a random network: with neighbour lists the macro step's closure rows come
    from the lists (short lists straight, longer ones by selection), without them from the
    searching feature kernel; the trajectories and the judge's statistics must be identical."""
    if scene == "dam_break_16k":
        p, parts, dt, _ = scenes.workload("dam_break_16k")
    else:
        p, parts = scenes.scaled_block((23, 21, 19), 0.01, seed=7, h_over_s=2.0)
        dt = 2e-4
    K, macro = 4, 6
    net = ClosureNet(max_batch=64)
    net.set_params((np.random.default_rng(11).normal(0, 0.05, ClosureNet.num_params())).astype(np.float32))
    res = []
    for lists in (0, 1):
        with Fluid(p, parts) as f:
            f.set_tuning(2, 2, lists)
            st = L.wc_closure_stats()
            L.check(L.lib().wc_closure_run(f.ctx, net.h, C.c_float(dt), K, macro, C.c_float(0.01), C.byref(st)),
                    "wc_closure_run")
            res.append((f.particles(1).tobytes(), st.macro_steps, st.fallbacks, st.max_density_jump))
    assert res[0] == res[1]


def _zero_net():
    net = ClosureNet(max_batch=64)
    net.set_params(np.zeros(ClosureNet.num_params(), np.float32))
    return net


@pytest.mark.parametrize("lists", [0, 1])
def test_zero_network_equals_plain_large_steps(lists):
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    K, macro = 4, 10
    with Fluid(p, parts) as a:
        a.step(dt * K, macro)
        want = a.particles(1)
    net = _zero_net()  # keep the handle alive for the whole run
    with Fluid(p, parts) as b:
        b.set_tuning(2, 2, lists)
        st = L.wc_closure_stats()
        L.check(L.lib().wc_closure_run(b.ctx, net.h, C.c_float(dt), K, macro, C.c_float(1e9), C.byref(st)),
                "wc_closure_run")
        assert st.macro_steps == macro and st.fallbacks == 0
        assert b.particles(1).tobytes() == want.tobytes()


@pytest.mark.parametrize("lists", [0, 1])
def test_rejecting_judge_reproduces_plain_small_steps(lists, oracle):
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    K, macro = 4, 6
    ref = oracle.OracleState(p, parts)
    ref.step(dt, K * macro)
    net = _zero_net()
    with Fluid(p, parts) as b:
        b.set_tuning(2, 2, lists)
        st = L.wc_closure_stats()
        L.check(L.lib().wc_closure_run(b.ctx, net.h, C.c_float(dt), K, macro, C.c_float(-1.0), C.byref(st)),
                "wc_closure_run")
        assert st.fallbacks == macro
        assert b.particles(1).tobytes() == ref.buf1.tobytes()


def test_rejecting_judge_restores_identities():
    """Identities tracked through a learned run whose judge rejects every step: each rejected
    step restores the state, identities included, that the force pass saved, so the run ends
    with the particles and identities of plain small steps."""
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    K, macro = 4, 5
    with Fluid(p, parts) as a:
        a.track_ids(True)
        a.step(dt, K * macro)
        want, want_ids = a.particles(1), a.ids(1)
    net = _zero_net()
    with Fluid(p, parts) as b:
        b.set_tuning(2, 2, 1)
        b.track_ids(True)
        st = L.wc_closure_stats()
        L.check(L.lib().wc_closure_run(b.ctx, net.h, C.c_float(dt), K, macro, C.c_float(-1.0), C.byref(st)),
                "wc_closure_run")
        assert st.fallbacks == macro
        assert b.particles(1).tobytes() == want.tobytes()
        assert np.array_equal(b.ids(1), want_ids)


def test_closure_entry_points_reject_what_they_cannot_run():
    """Closure inputs formed from the lists need the lists; the learned closure runs on
    single-scene contexts only.  Both are refused with WC_ERR_INVALID_ARG, not run."""
    from watercube_amd import batch

    p, parts, dt, _ = scenes.workload("dam_break_16k")
    net = _zero_net()
    y = DeviceArray((p.num_particles, 4), np.float32)
    with Fluid(p, parts) as f:
        f.set_tuning(2, 2, 0)
        assert L.lib().wc_closure_infer(f.ctx, net.h, C.c_float(4.0), y.ptr) == L.WC_ERR_INVALID_ARG
    with batch.SceneBatch(batch.scenes_256(seed=1, num_scenes=2, n_min=4096, n_max=8192)) as b:
        x = DeviceArray((b.total_particles, 208), np.uint16)
        assert L.lib().wc_closure_features(b.ctx, C.c_float(4.0), None, 0, x.ptr) == L.WC_ERR_INVALID_ARG
        assert L.lib().wc_closure_infer(b.ctx, net.h, C.c_float(4.0), y.ptr) == L.WC_ERR_INVALID_ARG
        st = L.wc_closure_stats()
        assert L.lib().wc_closure_run(b.ctx, net.h, C.c_float(dt), 4, 1, C.c_float(0.01),
                                      C.byref(st)) == L.WC_ERR_INVALID_ARG


def test_identities_follow_particles():
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    with Fluid(p, parts) as f:
        f.track_ids(True)
        f.step(dt, 7)
        ids = f.ids(1)
        assert np.array_equal(np.sort(ids), np.arange(p.num_particles))
        f.sort_.run()
        order = f.sort_.getSortedBuffer()
        assert np.array_equal(f.ids(2), ids[order])


@pytest.mark.parametrize("lists", [0, 1])
def test_neighbour_averaged_correction_of_a_constant_field(lists):
    """A network whose output is the constant c: averaged over neighbours, the correction of a
    particle whose neighbours all lie 2h or more from the walls is c again (up to rounding), while
    within 3h of a wall the neighbours there count with zero and pull it down."""
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    K = 2
    Lo = MO.layout()
    w = np.zeros(ClosureNet.num_params(), np.float32)
    c = np.array([0.5, -0.3, 0.2], np.float32)
    w[Lo["b4"][0]:Lo["b4"][0] + 3] = c
    net = ClosureNet(max_batch=64)
    net.set_params(w)
    out = {}
    for smooth in (0, 1):
        with Fluid(p, parts) as f:
            f.set_tuning(2, 2, lists)
            L.check(L.lib().wc_closure_set_smoothing(f.ctx, smooth), "smoothing")
            st = L.wc_closure_stats()
            L.check(L.lib().wc_closure_run(f.ctx, net.h, C.c_float(dt), K, 1, C.c_float(1e30), C.byref(st)),
                    "wc_closure_run")
            out[smooth] = f.particles(1)
    h = float(p.kernel_radius)
    # both runs sorted the same initial state, so slot i holds the same particle in both
    v_own, v_avg = out[0]["velocity"], out[1]["velocity"]
    x = out[0]["position"]
    dist_wall = np.minimum(x, 1.0 - x).min(axis=1)
    scale = float(np.abs(c).max()) * 9.81 * K * dt
    deep = dist_wall > 3.2 * h
    band = (dist_wall > 2.2 * h) & (dist_wall < 2.8 * h)
    assert deep.sum() > 100 and band.sum() > 100
    assert np.abs(v_avg[deep] - v_own[deep]).max() < 1e-3 * scale
    assert np.abs(v_avg[band] - v_own[band]).max() > 0.05 * scale


def test_frozen_force_integrator_matches_its_formula():
    """wc_closure_run with no network and the substep integrator advances each particle as K
    symplectic Euler steps of dt under the force frozen at the start: v = v0 + (K dt) a,
    x = x0 + dt (K v0 + (K (K+1)/2) dt a), in float32 in that association (checked bit for bit
    where no wall clamp applies; a from wc_closure_update's plain acceleration)."""
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    K = 4
    n = p.num_particles
    f32 = np.float32
    with Fluid(p, parts) as a:
        a.step(dt, 20)
        state = a.particles(1)
    acc = DeviceArray((n, 4), np.float32)
    with Fluid(p, state) as f:
        f.sort_.run()
        f.runDensityProg()
        s0 = f.particles(2)
        L.check(L.lib().wc_closure_update(f.ctx, C.c_float(dt), None, C.c_float(0.0), acc.ptr), "closure update")
        f.synchronize()
        a_plain = acc.download()[:, :3]
    with Fluid(p, state) as g:
        L.check(L.lib().wc_closure_set_integrator(g.ctx, 1), "integrator")
        st = L.wc_closure_stats()
        L.check(L.lib().wc_closure_run(g.ctx, None, C.c_float(dt), K, 1, C.c_float(1e30), C.byref(st)), "run")
        out = g.particles(1)
    dt_large = f32(dt) * f32(K)
    kf = f32(K)
    dts = dt_large / kf
    s = (kf * (kf + f32(1.0))) * f32(0.5)
    x0, v0 = s0["position"], s0["velocity"]
    want_x = x0 + dts * (kf * v0 + (s * dts) * a_plain)
    want_v = v0 + a_plain * dt_large
    free = np.all((want_x > f32(p.border)) & (want_x < f32(p.size) - f32(p.border)), axis=1)
    assert free.sum() > 0.9 * n
    assert np.array_equal(out["position"][free].view(np.uint32), want_x[free].view(np.uint32))
    assert np.array_equal(out["velocity"][free].view(np.uint32), want_v[free].view(np.uint32))
