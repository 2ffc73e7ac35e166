"""Snapshots of a simulation state: readback, dump and resume.

The reference's only observability is reading its buffers back and printing them:
``util::getParticles`` / ``util::getUints`` (src/core/util.cpp:42-75), ``util::printParticles``
(src/core/util.cpp:112-126) and ``Sort::printGrids`` (src/core/Sort.cpp:237-249).  Here a state
is saved as two files:

``<stem>.bin``   the particle buffer as raw 32-byte records, exactly the reference's
                 ``struct Particle`` (src/core/util.h:29-35): float32 little-endian
                 px py pz density vx vy vz pressure;
``<stem>.json``  the simulation parameters (every field of ``wc_params``, plus its raw bytes
                 in hex so that a resumed run sees the same bits), the particle count, the step
                 index, dt and a SHA-256 of the records (optional: the C++ harness, which
                 writes and reads the same format, leaves it out).

Buffer 1 (the one the next step sorts from) is what is saved, so a run resumed from a snapshot
takes exactly the steps the original run would have taken next.

    python -m watercube_amd.snapshot show STEM [--n 10] [--grid]
    python -m watercube_amd.snapshot run WORKLOAD STEPS --out STEM [--from STEM]
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

from . import _lib as L

FORMAT = "watercube-snapshot-1"
FIELDS = ["px", "py", "pz", "density", "vx", "vy", "vz", "pressure"]


def _paths(stem) -> tuple[Path, Path]:
    stem = Path(stem)
    if stem.suffix in (".bin", ".json"):
        stem = stem.with_suffix("")
    return stem.with_suffix(".bin"), stem.with_suffix(".json")


def save(stem, params: L.wc_params, particles: np.ndarray, step: int = 0, dt: float | None = None) -> Path:
    """Write `particles` (structured records, ``PARTICLE_DTYPE``) and `params` as a snapshot."""
    rec = np.ascontiguousarray(particles, dtype=L.PARTICLE_DTYPE)
    if rec.shape[0] != params.num_particles:
        raise ValueError(f"{rec.shape[0]} records for a {params.num_particles}-particle parameter block")
    bin_path, json_path = _paths(stem)
    raw = rec.view(np.uint8).tobytes()
    bin_path.write_bytes(raw)
    meta = {"format": FORMAT, "particles": int(rec.shape[0]), "record_bytes": 32, "fields": FIELDS,
            "step": int(step), "dt": None if dt is None else float(dt),
            "params": params.as_dict(), "params_hex": bytes(params).hex(),
            "sha256": hashlib.sha256(raw).hexdigest()}
    json_path.write_text(json.dumps(meta, indent=1) + "\n")
    return bin_path


def load(stem) -> tuple[L.wc_params, np.ndarray, dict]:
    """Read a snapshot: (parameters with the saved bits, records, metadata)."""
    bin_path, json_path = _paths(stem)
    meta = json.loads(json_path.read_text())
    if meta.get("format") != FORMAT:
        raise ValueError(f"{json_path}: not a {FORMAT} file")
    raw = bin_path.read_bytes()
    if len(raw) != 32 * meta["particles"]:
        raise ValueError(f"{bin_path}: {len(raw)} bytes for {meta['particles']} records")
    if meta.get("sha256") and hashlib.sha256(raw).hexdigest() != meta["sha256"]:
        raise ValueError(f"{bin_path}: checksum differs from {json_path}")
    blob = bytes.fromhex(meta["params_hex"])
    p = L.wc_params()
    if len(blob) != C.sizeof(p):
        raise ValueError(f"{json_path}: parameter block of {len(blob)} bytes, expected {C.sizeof(p)}")
    C.memmove(C.byref(p), blob, len(blob))
    rec = np.frombuffer(raw, dtype=L.PARTICLE_DTYPE).copy()
    return p, rec, meta


def format_particles(particles: np.ndarray, params: L.wc_params, n: int = 10, debug=None) -> str:
    """The first n records one per line, as util::printParticles prints them: position, velocity,
    density, pressure and the cell coordinate ivec3(position / binSize) (src/core/util.cpp:112-126)."""
    lines = ["-----Particles-----"]
    bin_size = np.float32(params.bin_size)
    for i in range(min(n, particles.shape[0])):
        r = particles[i]
        pos, vel = r["position"], r["velocity"]
        cell = np.trunc(pos / bin_size).astype(np.int64)
        s = (f"p=<{pos[0]:.6f}, {pos[1]:.6f}, {pos[2]:.6f}> v=<{vel[0]:.6f}, {vel[1]:.6f}, {vel[2]:.6f}> "
             f"d={float(r['density']):.6f}, pr={float(r['pressure']):.6f} binCoord=<{cell[0]}, {cell[1]}, {cell[2]}>")
        if debug is not None:
            s += f" debug={int(debug[i])}"
        lines.append(s)
    return "\n".join(lines)


def format_grids(counts: np.ndarray, offsets: np.ndarray, limit: int | None = None) -> str:
    """Per-cell counts and offsets as Sort::printGrids prints them (src/core/Sort.cpp:237-249)."""
    k = len(counts) if limit is None else min(limit, len(counts))
    c = "counts: " + "".join(f"{int(v)}, " for v in counts[:k])
    o = "offsets: " + "".join(f"{int(v)}, " for v in offsets[:k])
    return c + "\n" + o


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("show", help="print a snapshot's header and first records")
    a.add_argument("stem")
    a.add_argument("--n", type=int, default=10)
    a.add_argument("--grid", action="store_true", help="also sort it on the GPU and print the first cells")
    b = sub.add_parser("run", help="advance a workload (or a snapshot) and save the state")
    b.add_argument("workload")
    b.add_argument("steps", type=int)
    b.add_argument("--out", required=True)
    b.add_argument("--from", dest="src", default=None, help="resume from this snapshot instead")
    args = ap.parse_args(argv)

    if args.cmd == "show":
        p, rec, meta = load(args.stem)
        print(json.dumps({k: meta.get(k) for k in ("format", "particles", "step", "dt", "sha256")}))
        print(format_particles(rec, p, args.n))
        if args.grid:
            from .fluid import Fluid

            with Fluid(p, rec) as f:
                f.sort_.run()
                print(format_grids(f.sort_.getCountBuffer(), f.sort_.getOffsetBuffer(), limit=64))
        return 0

    from . import scenes
    from .fluid import Fluid

    if args.src:
        p, parts, meta = load(args.src)
        dt, step0 = meta["dt"], meta["step"]
    else:
        p, parts, dt, _ = scenes.workload(args.workload)
        step0 = 0
    with Fluid(p, parts) as f:
        f.step(dt, args.steps)
        state = f.particles(1)
    save(args.out, p, state, step=step0 + args.steps, dt=dt)
    print(json.dumps({"out": str(_paths(args.out)[0]), "step": step0 + args.steps}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
