"""Snapshot format (watercube_amd/snapshot.py): the saved state reads back with identical bits,
corruption is detected, and the text dumps follow util::printParticles / Sort::printGrids
(src/core/util.cpp:112-126, src/core/Sort.cpp:237-249)."""
import json

import numpy as np
import pytest

from watercube_amd import _lib as L
from watercube_amd import scenes, snapshot


def _state():
    p, parts, dt, _ = scenes.workload("dam_break_16k")
    parts = parts.copy()
    parts["density"] = np.linspace(900, 1100, len(parts), dtype=np.float32)
    parts["velocity"][:, 1] = -0.25
    return p, parts, dt


def test_round_trip_keeps_every_bit(tmp_path):
    p, parts, dt = _state()
    snapshot.save(tmp_path / "s", p, parts, step=17, dt=dt)
    q, rec, meta = snapshot.load(tmp_path / "s.bin")
    assert bytes(q) == bytes(p)
    assert rec.tobytes() == np.ascontiguousarray(parts).tobytes()
    assert meta["step"] == 17 and meta["particles"] == p.num_particles
    assert (tmp_path / "s.bin").stat().st_size == 32 * p.num_particles


def test_corruption_is_detected(tmp_path):
    p, parts, dt = _state()
    snapshot.save(tmp_path / "s", p, parts, step=0, dt=dt)
    raw = bytearray((tmp_path / "s.bin").read_bytes())
    raw[1000] ^= 1
    (tmp_path / "s.bin").write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        snapshot.load(tmp_path / "s")
    (tmp_path / "s.bin").write_bytes(bytes(raw[:-32]))
    with pytest.raises(ValueError, match="bytes for"):
        snapshot.load(tmp_path / "s")


def test_header_without_checksum_is_accepted(tmp_path):
    """The C++ harness writes the same header without the optional checksum."""
    p, parts, dt = _state()
    snapshot.save(tmp_path / "s", p, parts, step=3, dt=dt)
    meta = json.loads((tmp_path / "s.json").read_text())
    for k in ("sha256", "params"):
        del meta[k]
    (tmp_path / "s.json").write_text(json.dumps(meta))
    q, rec, m = snapshot.load(tmp_path / "s")
    assert bytes(q) == bytes(p) and m["step"] == 3


def test_text_dumps():
    p, parts, _ = _state()
    text = snapshot.format_particles(parts, p, n=3).splitlines()
    assert text[0] == "-----Particles-----" and len(text) == 4
    cell = np.trunc(parts["position"][2] / np.float32(p.bin_size)).astype(int)
    assert text[3].endswith(f"binCoord=<{cell[0]}, {cell[1]}, {cell[2]}>")
    assert "d=900.000000" in text[1]
    g = snapshot.format_grids(np.array([2, 0, 1]), np.array([0, 2, 2])).splitlines()
    assert g == ["counts: 2, 0, 1, ", "offsets: 0, 2, 2, "]


def test_particle_count_mismatch_is_refused(tmp_path):
    p, parts, dt = _state()
    with pytest.raises(ValueError):
        snapshot.save(tmp_path / "s", p, parts[:-1], dt=dt)
