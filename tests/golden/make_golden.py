"""Writes the golden fixtures of the SPH step (tests/golden/*.npz + manifest.json).

Each fixture is one small scene run through the CPU restatement of the reference's passes
(oracle/wc_oracle.c): the input records, then the outputs of every pass of one step --
cell key per particle (count.comp:31-33), per-cell counts and exclusive offsets
(linearScan.comp:15-28), the stable sorted order and sorted records (reorder.comp:32-45),
density/pressure in place (density.comp:81-137), the force + integration output
(update.comp:134-232) -- and the records after a short trajectory of whole steps
(Fluid::update, src/core/Fluid.cpp:342-354).

The reference ships no golden vectors of its own (it has no tests), so these files pin the
restatement against later change and give the GPU path fixed data to reproduce bit for bit.
Scene inputs come from the library's host-side generators (no GPU needed).  Regenerate with
``python tests/golden/make_golden.py``; the loaders use ``numpy.load`` without pickles.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from oracle import wc_oracle as O  # noqa: E402
from watercube_amd import _lib as L  # noqa: E402
from watercube_amd import scenes as S  # noqa: E402

FP_SEMANTICS = ("binary32; correctly rounded / and sqrt; subnormals kept; no FMA contraction; "
                "pow(x,3) = (x*x)*x; length = sqrt((x*x+y*y)+z*z); left-to-right association")
TRAJ_STEPS = 20


def _dam_break_2k():
    p = L.preset(L.WC_PRESET_DAM_BREAK_16K)
    parts = S.block(16, 16, 8, 0.015, 0.0015, seed=7)
    p.num_particles = len(parts)
    return p, parts, 2e-4


def _reference_default_2k():
    # the reference's own constants (cubic EOS, wall density and wall force on, g = -900)
    p = L.preset(L.WC_PRESET_REFERENCE_DEFAULT)
    parts = S.reference_scene(2048, p.particle_radius, seed=3)
    p.num_particles = len(parts)
    dt = float(np.float32(np.float32(1.0 / 60.0) * np.float32(p.time_scale)))
    return p, parts, dt


def _mouse_ray_2k():
    # dam-break constants with the mouse-ray repulsion (update.comp:105-132) switched on and a
    # ray through the block; particles start moving so viscosity and the clamp are exercised
    p = L.preset(L.WC_PRESET_DAM_BREAK_16K)
    parts = S.block(16, 8, 16, 0.015, 0.0015, seed=11, velocity=(0.4, -0.3, 0.2))
    p.num_particles = len(parts)
    p.mouse_force = 1
    for i, v in enumerate((0.12, 0.06, -1.0)):
        p.camera_position[i] = v
    d = np.array([0.0, 0.05, 1.0])
    d /= np.linalg.norm(d)
    for i, v in enumerate(d):
        p.mouse_ray_direction[i] = float(v)
    return p, parts, 2e-4


CASES = {
    "dam_break_2k": _dam_break_2k,
    "reference_default_2k": _reference_default_2k,
    "mouse_ray_2k": _mouse_ray_2k,
}


def run_case(p, parts, dt, traj_steps=TRAJ_STEPS) -> dict:
    """Every pass output of one step, then the records after `traj_steps` whole steps."""
    st = O.OracleState(p, parts)
    out = {"params": np.frombuffer(bytes(p), np.uint8).copy(), "input": st.buf1.copy(),
           "dt": np.array([dt], np.float32)}
    st.sort_run()
    out["keys"] = st.keys.copy()
    out["counts"] = st.counts.copy()
    out["offsets"] = st.offsets.copy()
    out["sorted"] = st.sorted.copy()
    out["sorted_records"] = st.buf2.copy()
    st.density(threads=1)
    out["after_density"] = st.buf2.copy()
    st.update(dt, threads=1)
    out["after_update"] = st.buf1.copy()
    st.step(dt, traj_steps - 1, threads=1)
    out["trajectory"] = st.buf1.copy()
    out["trajectory_steps"] = np.array([traj_steps], np.uint32)
    return out


def params_from_fixture(blob: np.ndarray) -> L.wc_params:
    p = L.wc_params()
    raw = bytes(np.asarray(blob, np.uint8))
    if len(raw) != C.sizeof(p):
        raise ValueError(f"parameter block of {len(raw)} bytes, expected {C.sizeof(p)}")
    C.memmove(C.byref(p), raw, len(raw))
    return p


def load(name: str) -> dict:
    with np.load(HERE / f"{name}.npz", allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    d["input"] = d["input"].view(O.PARTICLE_DTYPE).reshape(-1)
    for k in ("sorted_records", "after_density", "after_update", "trajectory"):
        d[k] = d[k].view(O.PARTICLE_DTYPE).reshape(-1)
    return d


def _digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()[:16]


def main() -> int:
    manifest = {"fp_semantics": FP_SEMANTICS, "generator": "tests/golden/make_golden.py",
                "oracle": "oracle/wc_oracle.c", "cases": {}}
    for name, make in CASES.items():
        p, parts, dt = make()
        d = run_case(p, parts, dt)
        arrays = {k: (v.view(np.uint8) if v.dtype == O.PARTICLE_DTYPE else v) for k, v in d.items()}
        np.savez_compressed(HERE / f"{name}.npz", **arrays)
        manifest["cases"][name] = {
            "particles": int(p.num_particles), "grid_res": int(p.grid_res), "dt": float(d["dt"][0]),
            "trajectory_steps": TRAJ_STEPS, "params": p.as_dict(),
            "sha256_16": {k: _digest(v) for k, v in d.items()},
        }
        print(f"{name}: {p.num_particles} particles, R={p.grid_res}, dt={d['dt'][0]:.6g}", flush=True)
    (HERE / "manifest.json").write_text(json.dumps(manifest, indent=1, sort_keys=True) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
