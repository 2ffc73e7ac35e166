"""Fixed workload for profiler runs (rocprofv3 kernel traces and PMC counters).

    python3 -m watercube_amd.profile_workload [--workload block_1m|block_1m_h23|scenes_256] [--steps 20]
        [--mode graph|eager]

Uploads the workload's initial state (the upload's record-split kernel `k_split` reads 32 bytes
and writes 32 bytes per particle: a plain copy of known size, used to calibrate the byte
counters), runs a few warm-up steps, then `--steps` steps either as the production captured
graph or as eager per-pass launches, and prints one line per stage.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def batch_workload(args) -> int:
    """The 256-scene batch (log-uniform 4,096..262,144 particles, seed 0), stepped as the bench
    steps it: one launch per pass over every scene, graph-captured.  Only kernel durations are
    meaningful here: each scene's upload is its own record-split copy, of its own size."""
    from watercube_amd import batch

    scenes = batch.scenes_256(seed=0)
    with batch.SceneBatch(scenes) as b:
        print(f"profile_workload: scenes_256 n={b.total_particles} scenes={b.num_scenes}", flush=True)
        if args.tuning == "auto":
            t = b.tune(reps=3)
            print(f"profile_workload: tuning {t['density_variant']},{t['update_variant']},{t['neighbour_list']}",
                  flush=True)
        else:
            import ctypes as C

            from watercube_amd import _lib as L

            d, u, *rest = (int(x) for x in args.tuning.split(","))
            t = L.wc_tuning()
            L.check(L.lib().wc_get_tuning(b.ctx, C.byref(t)), "wc_get_tuning")
            t.density_variant, t.update_variant, t.neighbour_list = d, u, (rest[0] if rest else 0)
            L.check(L.lib().wc_set_tuning(b.ctx, C.byref(t)), "wc_set_tuning")
        b.step(args.warmup)
        b.synchronize()
        t0 = time.perf_counter()
        b.step(args.steps)
        b.synchronize()
        print(f"profile_workload: {args.steps} graph steps in {(time.perf_counter() - t0) * 1e3:.2f} ms", flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="block_1m")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mode", choices=["graph", "eager"], default="graph")
    ap.add_argument("--tuning", default="auto",
                    help="'auto' runs the autotuner first; 'D,U[,L]' fixes the density / force-pass variants "
                         "(and L = 1 the neighbour lists)")
    ap.add_argument("--closure-k", type=int, default=0,
                    help="K > 0: after the plain steps, run --steps judged learned steps of K*dt "
                         "(wc_closure_run with the default weights)")
    args = ap.parse_args()

    from watercube_amd import Fluid, scenes

    if args.workload == "scenes_256":
        return batch_workload(args)
    if args.workload == "block_1m_h23":
        # the 1M block's lattice with h = 2.3 spacings: about 50 listed neighbours per particle
        p, parts = scenes.scaled_block((100, 100, 100), 0.008, seed=3, h_over_s=2.3)
        dt = 1e-4
    else:
        p, parts, dt, _ = scenes.workload(args.workload)
    print(f"profile_workload: {args.workload} n={p.num_particles} copy-calibration bytes/dispatch "
          f"read={32 * p.num_particles} write={32 * p.num_particles}", flush=True)
    with Fluid(p, parts) as f:
        if args.tuning == "auto":
            f.tune(dt, reps=3)
        else:
            f.set_tuning(*(int(x) for x in args.tuning.split(",")))
        f.step(dt, args.warmup)
        f.synchronize()
        t0 = time.perf_counter()
        if args.mode == "graph":
            f.step(dt, args.steps)
        else:
            for _ in range(args.steps):
                f.sort_.run()
                f.runDensityProg()
                f.runUpdateProg(dt)
        f.synchronize()
        dt_wall = time.perf_counter() - t0
        print(f"profile_workload: {args.steps} {args.mode} steps in {dt_wall * 1e3:.2f} ms", flush=True)
        if args.closure_k > 0:
            import ctypes as C

            from watercube_amd import _lib as L
            from watercube_amd.closure import DEFAULT_WEIGHTS, ClosureNet

            net = ClosureNet(max_batch=1024).load(DEFAULT_WEIGHTS)
            st = L.wc_closure_stats()
            t0 = time.perf_counter()
            L.check(L.lib().wc_closure_run(f.ctx, net.h, C.c_float(dt), args.closure_k, args.steps,
                                           C.c_float(0.01), C.byref(st)), "wc_closure_run")
            wall = time.perf_counter() - t0
            print(f"profile_workload: {st.macro_steps} learned steps of {args.closure_k} dt in {wall * 1e3:.2f} ms "
                  f"(device {st.wall_ms:.2f} ms, {st.fallbacks} fallbacks)", flush=True)
            net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
