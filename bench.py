"""Benchmark of the SPH step on one MI355X: particle-steps/s on the 1,000,000-particle block.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one Fluid::update of the reference (sort + density + force/integrate) over all
particles, launched as one captured HIP graph.  The particle state is resident in HBM before the
timed region starts.  Prints ONE JSON line.

Besides the throughput the line carries
  roofline     the dominant kernel's algorithmic bytes per launch over its average duration,
               measured with HIP events on the context's stream (eager launches, same state),
               against the 8 TB/s HBM3E peak; `traffic` is the HBM bytes per launch that the
               committed rocprofv3 FETCH_SIZE / WRITE_SIZE profile of that kernel measured;
  cpu_baseline the CPU restatement of the same step on this host's cores, on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
WORKLOAD = "block_1m"


def algorithmic_bytes(n: int, ncells: int) -> dict:
    """Bytes each kernel must move once per launch (every array touched once, neighbour
    gathers counted once)."""
    return {
        "hash": 16 * n + 8 * n + 4 * ncells,            # read pos+rho, write key+rank, histogram
        "scan": 4 * ncells + 4 * ncells + 8 * ncells,   # read counts, zero counts, write (offset,count)
        "place": 8 * n + 8 * ncells + 4 * n,            # read key+rank, offsets, write slot map
        "gather": 4 * n + 4 * n + 8 * ncells + 32 * n + 32 * n + 8 * n,  # map, key, ranges, record in/out, order+key
        "density": 16 * n + 4 * n + 8 * ncells + 8 * n,  # pos, cell key, ranges, write rho+p
        "update": 32 * n + 4 * n + 8 * ncells + 32 * n,  # record in, cell key, ranges, record out
    }


def read_traffic(kernel: str) -> float | None:
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if present."""
    f = ROOT / "profiles" / f"traffic_{WORKLOAD}.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        v = d["kernels"][kernel]["hbm_bytes_per_launch"]
        return float(v)
    except Exception:
        return None


def cpu_baseline(p, parts, dt, budget_s: float = 12.0, max_steps: int = 60) -> dict:
    from oracle import wc_oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    threads = max(1, min(threads, 16, os.cpu_count() or 1))
    st = wc_oracle.OracleState(p, parts)
    done, elapsed = 0, 0.0
    while done < max_steps and elapsed < budget_s:
        t0 = time.perf_counter()
        st.step_timed(dt, 1, threads)
        elapsed += time.perf_counter() - t0
        done += 1
    n = int(p.num_particles)
    return {
        "value": n * done / elapsed,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{done} steps of {WORKLOAD} ({n} particles, dt={dt}) from the same initial state, "
                  f"C restatement -O3 with OpenMP over particles for density/update, {elapsed:.1f} s",
    }


def batched_section(steps: int, warmup: int = 10, latency_samples: int = 20) -> dict:
    """256 independent scenes (log-uniform 4,096..262,144 particles) advanced as one batch:
    aggregate particle-steps/s over `steps` graph-launched steps, and per-scene step latency
    (device clock from the step's first pass to the scene's last particle) over single steps."""
    from watercube_amd import batch

    scenes = batch.scenes_256(seed=0)
    with batch.SceneBatch(scenes) as b:
        n = b.total_particles
        tuning = b.tune(reps=2)
        b.step(warmup)
        b.synchronize()
        t0 = time.perf_counter()
        b.step(steps)
        b.synchronize()
        elapsed = time.perf_counter() - t0
        lat = []
        for _ in range(latency_samples):
            b.step(1)
            b.synchronize()
            lat.append(b.latency_ms())
    lat = np.concatenate(lat)
    kinds = {}
    for _, _, _, k in scenes:
        kinds[k] = kinds.get(k, 0) + 1
    return {
        "workload": "scenes_256",
        "scenes": len(scenes),
        "particles": n,
        "fills": kinds,
        "steps": steps,
        "value": n * steps / elapsed,
        "unit": "particle-steps/s",
        "ms_per_step": elapsed * 1e3 / steps,
        "scene_latency_ms_p50": float(np.percentile(lat, 50)),
        "scene_latency_ms_p99": float(np.percentile(lat, 99)),
        "tuning": [tuning["density_variant"], tuning["update_variant"]],
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch-steps", type=int, default=500, help="steps of the 256-scene batch (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the step runs on a single MI355X)",
              file=sys.stderr)
        return 2

    from watercube_amd import Fluid, scenes

    p, parts, dt, _ = scenes.workload(WORKLOAD)
    n = int(p.num_particles)
    ncells = int(p.grid_res) ** 3
    print(f"# bench: {WORKLOAD} n={n} R={p.grid_res} dt={dt} warmup={args.warmup} steps={args.steps}",
          file=sys.stderr, flush=True)

    with Fluid(p, parts) as f:
        tuning = f.tune(dt, reps=5)  # untimed: picks the fastest variant of each neighbour pass
        print(f"# tuned variants: density {tuning['variants'][tuning['density_variant']]} "
              f"update {tuning['variants'][tuning['update_variant']]}", file=sys.stderr, flush=True)
        print("# variant ms density: " + " ".join(f"{x:.4f}" for x in tuning["density_ms"]) +
              " | update: " + " ".join(f"{x:.4f}" for x in tuning["update_ms"]), file=sys.stderr, flush=True)
        f.step(dt, args.warmup)
        f.synchronize()
        t0 = time.perf_counter()
        f.step(dt, args.steps)
        f.synchronize()
        elapsed = time.perf_counter() - t0
        print(f"# timed {args.steps} steps in {elapsed:.4f} s", file=sys.stderr, flush=True)

        # per-kernel durations: HIP events on the context's stream around each launch
        prof = f.profile(dt, args.profile_steps)
        f.synchronize()
        state = f.particles(1)
    if not np.isfinite(state["position"]).all():
        print("bench: non-finite positions after the run", file=sys.stderr)
        return 1

    per_launch_ms = {k: v / args.profile_steps for k, v in prof.items()}
    algo = algorithmic_bytes(n, ncells)
    dom = max(per_launch_ms, key=per_launch_ms.get)
    achieved = algo[dom] / (per_launch_ms[dom] * 1e-3) / 1e9
    traffic = read_traffic(dom)
    print("# per-launch ms: " + ", ".join(f"{k}={v:.4f}" for k, v in per_launch_ms.items()), file=sys.stderr,
          flush=True)

    out = {
        "metric": "particle-steps/s at 1M particles: plain vs learned large-dt, ≤1% density err",
        "value": n * args.steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "particles": n, "grid_res": int(p.grid_res), "dt": dt,
                   "method": "plain", "kernel_radius": float(p.kernel_radius),
                   "tuning": {"density": tuning["variants"][tuning["density_variant"]],
                              "update": tuning["variants"][tuning["update_variant"]]}},
        "simulated_s_per_s": dt * args.steps / elapsed,
        "pass_ms": per_launch_ms,
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": algo[dom],
        },
    }
    if args.batch_steps > 0:
        out["batched"] = batched_section(args.batch_steps)
        print(f"# batched: {out['batched']}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, parts, dt)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
