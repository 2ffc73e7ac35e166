"""Batched scenes: several independent scenes in one context, each pass one launch over all of
them.  Every scene must evolve exactly as it does alone (bit for bit against the CPU
restatement run scene by scene), whatever the other scenes hold."""
import numpy as np
import pytest

from watercube_amd import batch
from watercube_amd import _lib as L

pytestmark = pytest.mark.gpu


def _small_batch():
    out = []
    for i, (kind, n) in enumerate([("block", 3000), ("ball", 2500), ("two_columns", 4000), ("block", 600),
                                   ("ball", 9000), ("two_columns", 1200)]):
        parts, s = batch.fill(kind, n, seed=100 + i)
        p = batch.scene_params(len(parts), s)
        dt = float(np.float32(2e-4 * s / 0.015))
        out.append((p, parts, dt, kind))
    # an empty scene in the middle is skipped by every pass
    p0 = batch.scene_params(0, 0.02)
    out.insert(3, (p0, np.zeros(0, L.PARTICLE_DTYPE), 1e-4, "empty"))
    return out


def test_batch_matches_each_scene_alone(oracle):
    scenes = _small_batch()
    steps = 20
    with batch.SceneBatch(scenes) as b:
        b.step(steps)
        b.synchronize()
        lat = b.latency_ms()
        for s, (p, parts, dt, kind) in enumerate(scenes):
            print(f"scene {s} {kind} n={len(parts)} R={p.grid_res} latency {lat[s]:.3f} ms", flush=True)
            st = oracle.OracleState(p, parts)
            st.step(dt, steps)
            got = b.scene_particles(s)
            assert got.tobytes() == st.buf1.tobytes(), f"scene {s} ({kind}) differs from the CPU restatement"
            if len(parts):
                counts, offsets = b.scene_cells(s)
                assert np.array_equal(counts, st.counts) and np.array_equal(offsets, st.offsets)
                assert lat[s] > 0
            else:
                assert lat[s] == 0


def test_batch_of_ragged_tiny_scenes(oracle):
    """Scenes of 1, 2, 63, 64, 65, 129, 257 and 4,097 particles: every pass maps its workgroups to
    scenes by each scene's first block, and these sizes put scene boundaries inside and at the
    edges of waves and workgroups.  Each scene must still match the CPU restatement run alone."""
    parts_all, s = batch.fill("block", 5000, seed=21)
    scenes = []
    for i, n in enumerate((1, 2, 63, 64, 65, 129, 257, 4097)):
        parts = np.ascontiguousarray(parts_all[(7 * i) % 200:(7 * i) % 200 + n])
        p = batch.scene_params(len(parts), s)
        scenes.append((p, parts, float(np.float32(2e-4 * s / 0.015)), f"n={n}"))
    steps = 10
    with batch.SceneBatch(scenes) as b:
        b.step(steps)
        b.synchronize()
        for i, (p, parts, dt, what) in enumerate(scenes):
            print(f"scene {i} {what}", flush=True)
            st = oracle.OracleState(p, parts)
            st.step(dt, steps)
            assert b.scene_particles(i).tobytes() == st.buf1.tobytes(), f"scene {what} differs"
            counts, offsets = b.scene_cells(i)
            assert np.array_equal(counts, st.counts) and np.array_equal(offsets, st.offsets)


def test_batch_tuning_keeps_results(oracle):
    scenes = _small_batch()
    with batch.SceneBatch(scenes) as b:
        b.step(3)
        info = b.tune(reps=2)
        print(f"batch tuning {info['density_variant']} {info['update_variant']}", flush=True)
        b.step(2)
        for s in (0, 4):
            p, parts, dt, _ = scenes[s]
            st = oracle.OracleState(p, parts)
            st.step(dt, 5)
            assert b.scene_particles(s).tobytes() == st.buf1.tobytes()


def test_batch_with_one_out_of_range_scene(oracle):
    """The force pass's division mode is one flag for the whole batch: a single velocity of 1e-30
    in one scene switches every scene's step to IEEE division.  Each scene must still match the
    CPU restatement bit for bit, the altered one and the others alike."""
    scenes = _small_batch()
    p, parts, dt, kind = scenes[2]
    parts = parts.copy()
    parts["velocity"][5, 0] = np.float32(1e-30)
    scenes[2] = (p, parts, dt, kind)
    with batch.SceneBatch(scenes) as b:
        b.step(4)
        b.synchronize()
        for s, (p, parts, dt, kind) in enumerate(scenes):
            st = oracle.OracleState(p, parts)
            st.step(dt, 4)
            assert b.scene_particles(s).tobytes() == st.buf1.tobytes(), f"scene {s} ({kind}) differs"
