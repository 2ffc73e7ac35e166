"""The CPU restatement and the scene generators under AddressSanitizer + UndefinedBehaviorSanitizer.

oracle/sanitize.mk builds both (gcc/g++, host code only) with a small command-line driver.  Each
case below runs one piece of them on the edge cases the GPU suite uses -- empty and tiny scenes,
particles outside the cube and on cell faces, NaN and infinite inputs, a cluster whose every
particle has hundreds of neighbours, a grid with far more cells than particles, the reference's
own scene with its wall terms, the never-dispatched advect pass -- and must (a) finish without a
sanitizer report and (b) give the same bits as the unsanitized builds the other tests use.  The
test is skipped where the sanitizer runtime cannot be linked."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from watercube_amd import _lib as L
from watercube_amd import scenes

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "oracle" / "_san" / "wc_sanitize"
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=87", OMP_NUM_THREADS="4")


@pytest.fixture(scope="module")
def exe():
    r = subprocess.run(["make", "-s", "-f", str(ROOT / "oracle" / "sanitize.mk")], capture_output=True, text=True)
    if r.returncode != 0 or not EXE.exists():
        pytest.skip(f"sanitizer build unavailable: {r.stderr.strip()[-300:]}")
    return EXE


def _run(exe, *args):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, env=ENV, timeout=300)
    report = r.stderr
    assert r.returncode == 0, f"sanitized run failed ({r.returncode}):\n{report[-3000:]}"
    assert "runtime error" not in report and "AddressSanitizer" not in report, report[-3000:]
    return r.stdout


def _same(a: np.ndarray, b: np.ndarray, what: str):
    """Bit-identical, except that NaNs only need to sit in the same places (their payloads follow
    operand order, which the two optimisation levels may swap)."""
    fa, fb = a.view(np.float32).ravel(), b.view(np.float32).ravel()
    na, nb = np.isnan(fa), np.isnan(fb)
    assert np.array_equal(na, nb), f"{what}: NaNs in different places"
    assert np.array_equal(fa[~na].view(np.uint32), fb[~nb].view(np.uint32)), f"{what}: values differ"


def _step_case(exe, tmp_path, oracle, p, parts, dt, steps, what):
    n = int(p.num_particles)
    (tmp_path / "p.bin").write_bytes(bytes(p))
    np.ascontiguousarray(parts, L.PARTICLE_DTYPE).tofile(tmp_path / "in.bin")
    _run(exe, "step", tmp_path / "p.bin", tmp_path / "in.bin", n, repr(float(dt)), steps, tmp_path / "out.bin")
    raw = (tmp_path / "out.bin").read_bytes()
    cells = int(p.grid_res) ** 3
    buf1 = np.frombuffer(raw, L.PARTICLE_DTYPE, n)
    rest = np.frombuffer(raw, np.uint32, offset=32 * n)
    counts, offsets, order = rest[:cells], rest[cells:2 * cells], rest[2 * cells:2 * cells + n]
    st = oracle.OracleState(p, parts)
    st.step(float(np.float32(dt)), steps)
    _same(buf1, st.buf1, what)
    assert np.array_equal(counts, st.counts) and np.array_equal(offsets, st.offsets), f"{what}: cells differ"
    assert np.array_equal(order, st.sorted[:n]), f"{what}: order differs"


def _params(n, preset=L.WC_PRESET_DAM_BREAK_16K):
    p = L.preset(preset)
    p.num_particles = n
    return p


def _records(pos, vel=None):
    a = np.zeros(len(pos), L.PARTICLE_DTYPE)
    a["position"] = pos
    if vel is not None:
        a["velocity"] = vel
    return a


def test_step_dam_break_and_tiny_scenes(exe, tmp_path, oracle):
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    _step_case(exe, tmp_path, oracle, p, parts, dt, 8, "dam break 16k")
    rng = np.random.default_rng(1)
    for n in (0, 1, 2, 65):
        _step_case(exe, tmp_path, oracle, _params(n), _records(rng.uniform(0.3, 0.36, (n, 3)).astype(np.float32)),
                   2e-4, 5, f"tiny n={n}")


def test_step_outside_the_cube_and_non_finite(exe, tmp_path, oracle):
    p = _params(0)
    b = float(p.bin_size)
    pos = np.float32([[-0.2, 0.5, 0.5], [1.3, 0.5, 0.5], [0.5, -1e-3, 2.0], [b, 2 * b, 3 * b], [0.0, 0.0, 0.0],
                      [1.0, 1.0, 1.0], [np.nan, 0.03, 0.03], [np.inf, 0.5, 0.5], [0.04, -np.inf, 0.04],
                      [1e30, -1e30, 3e38], [0.05, 0.05, 0.05], [0.051, 0.05, 0.05]])
    vel = np.float32([[3, -4, 5]] * 10 + [[np.nan, 1, 1], [np.inf, -np.inf, 0]])
    p.num_particles = len(pos)
    _step_case(exe, tmp_path, oracle, p, _records(pos, vel), 2e-4, 6, "outside the cube, NaN and infinity")


def test_step_dense_cluster_and_fine_grid(exe, tmp_path, oracle):
    p = _params(400)
    h = float(p.kernel_radius)
    rng = np.random.default_rng(9)
    _step_case(exe, tmp_path, oracle, p, _records(rng.uniform(0.5, 0.5 + h / 4, (400, 3)).astype(np.float32)),
               2e-5, 3, "dense cluster")
    q = L.preset(L.WC_PRESET_BLOCK_SPHERE_17M)
    pos = np.concatenate([rng.uniform(0.2, 0.2 + 6 * float(q.kernel_radius), (500, 3)),
                          rng.uniform(0.0, 1.0, (10, 3))]).astype(np.float32)
    q.num_particles = len(pos)
    _step_case(exe, tmp_path, oracle, q, _records(pos), 2.5e-5, 2, "250^3 grid, 510 particles")


def test_step_reference_default_scene(exe, tmp_path, oracle):
    """The reference's own constants: cubic EOS, wall density and wall force, g = -900."""
    p, parts, dt, _ = scenes.workload("reference_default")
    _step_case(exe, tmp_path, oracle, p, parts, dt, 2, "reference default 80k")


def test_advect(exe, tmp_path, oracle):
    p, parts, _, _ = scenes.workload("dam_break_16k")
    rng = np.random.default_rng(3)
    parts = parts.copy()
    parts["velocity"] = rng.uniform(-60.0, 60.0, (parts.shape[0], 3)).astype(np.float32)
    (tmp_path / "p.bin").write_bytes(bytes(p))
    parts.tofile(tmp_path / "in.bin")
    _run(exe, "advect", tmp_path / "p.bin", tmp_path / "in.bin", parts.shape[0], "0.002", 3, tmp_path / "out.bin")
    st = oracle.OracleState(p, parts)
    for _ in range(3):
        st.advect(2e-3)
    _same(np.fromfile(tmp_path / "out.bin", L.PARTICLE_DTYPE), st.buf1, "advect")


@pytest.mark.parametrize("args", [
    ("block", 32, 32, 16, 0, 0, 0, 0.015, 0.0015, 0, 0, 0, 0),
    ("block", 0, 5, 5, 0, 0, 0, 0.01, 0.001, 0, 0, 0, 1),
    ("block", 7, 1, 3, 0.75, 0, 0.75, 0.01, 0.001, 0, -2, 0, 7919),
    ("sphere", 5000, 0.5, 0.6, 0.5, 0.008, 0.0008, 0, -2, 0, 3),
    ("sphere", 1, 0.5, 0.5, 0.5, 0.01, 0.0, 0, 0, 0, 4),
    ("sphere", 0, 0.5, 0.5, 0.5, 0.01, 0.0, 0, 0, 0, 4),
    ("sphere", 200000, 0.5, 0.9, 0.5, 0.008, 0.0008, 0, 0, 0, 5),  # leaves the cube: an error
    ("reference", 80000, 0.01, 0, None),
    ("reference", 0, 0.01, 0, None),
    ("reference", 1001, 0.02, 12, None),
])
def test_scene_generators(exe, tmp_path, args):
    import ctypes as C

    out = tmp_path / "scene.bin"
    argv = [a for a in args if a is not None]
    rc = int(_run(exe, *argv, out).strip())
    kind, n = args[0], None
    if kind == "block":
        nx, ny, nz = args[1:4]
        n = nx * ny * nz
        want = np.zeros(n, L.PARTICLE_DTYPE)
        got_rc = L.lib().wc_scene_block(want.ctypes.data, nx, ny, nz, L.f3(args[4:7]), args[7], args[8],
                                        L.f3(args[9:12]), C.c_uint64(args[12]))
    elif kind == "sphere":
        n = args[1]
        want = np.zeros(n, L.PARTICLE_DTYPE)
        got_rc = L.lib().wc_scene_sphere(want.ctypes.data, n, L.f3(args[2:5]), args[5], args[6], L.f3(args[7:10]),
                                         C.c_uint64(args[10]))
    else:
        n = args[1]
        want = np.zeros(n, L.PARTICLE_DTYPE)
        got_rc = L.lib().wc_scene_reference(want.ctypes.data, n, args[2], C.c_uint64(args[3]))
    assert rc == got_rc
    if rc == 0:
        got = np.fromfile(out, L.PARTICLE_DTYPE)
        assert got.tobytes() == want.tobytes()
