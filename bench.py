"""Benchmark of the SPH step on one MI355X: particle-steps/s on the 1,000,000-particle block.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one Fluid::update of the reference (sort + density + force/integrate) over all
particles, launched as one captured HIP graph.  The particle state is resident in HBM before the
timed region starts.  Prints ONE JSON line.

Besides the throughput the line carries
  roofline     the dominant kernel's algorithmic bytes per launch over its average duration,
               measured with HIP events on the context's stream (eager launches, same state),
               against the 8 TB/s HBM3E peak; `traffic` is the HBM bytes per launch that the
               committed rocprofv3 FETCH_SIZE / WRITE_SIZE profile of that kernel measured, and
               `valu` its VALU issue rate (SQ_INSTS_VALU of the committed profile over the live
               duration) against one wave64 instruction per CU per cycle;
  pass_ms      the average duration of each of the six kernels of the step, and pass_gbs the
               same passes' algorithmic and measured HBM bytes per launch over it (GB/s);
  batched      256 independent scenes of 4,096..262,144 particles stepped together;
  scale        the 17,777,216-particle scene;
  learned      plain and learned steps of K*dt against plain steps of dt (density errors,
               simulated time per second), and the closure network's matrix-core utilisation;
  cpu_baseline the CPU restatement of the same step on this host's cores, on a bounded sample
               (and, in single_thread, on one core).
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
    """Bytes each pass must move once per launch: every array touched once, each neighbour
    gather counted once, nothing the implementation adds (keys kept beside the sorted records,
    neighbour lists).  The reorder is one pass here (place + gather)."""
    return {
        "hash": 16 * n + 4 * n + 4 * ncells + 4 * ncells,  # read positions, write keys, histogram and its clearing
        "scan": 8 * ncells,                                 # read counts, write offsets
        "scatter": 36 * n + 4 * ncells + 32 * n,            # read key + record, offsets; write the record
        "density": 16 * n + 8 * ncells + 8 * n,             # positions, cell ranges; write density + pressure
        "update": 32 * n + 8 * ncells + 32 * n,             # records, cell ranges; write records
    }


def overhead_bytes(n: int, list_entries: float) -> dict:
    """Bytes the neighbour passes move beyond the algorithmic ones by this build's design: the
    sorted cell key each pass reads, and with neighbour lists the list entries and counts the
    density pass writes and the force pass reads back."""
    lists = int(round(4 * list_entries * n)) + 4 * n if list_entries else 0
    return {"density": 4 * n + lists, "update": 4 * n + lists}


PROFILE_FROM_STEP = 520  # the committed counter profiles cover steps 521-570 of the workload
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 2.4e9  # one wave64 VALU instruction per CU per cycle (no packed math)


def _profile_entry(kernel: str) -> dict:
    """The committed profile's entry for pass `kernel` (its kernel is k_<pass>)."""
    f = ROOT / "profiles" / f"traffic_{WORKLOAD}.json"
    if not f.exists():
        return {}
    try:
        return dict(json.loads(f.read_text())["kernels"].get(f"k_{kernel}", {}))
    except (ValueError, KeyError):
        return {}


def read_traffic(kernel: str) -> float | None:
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if present."""
    v = _profile_entry(kernel).get("hbm_bytes_per_launch")
    return None if v is None else float(v)


def read_valu(kernel: str) -> float | None:
    """VALU wave instructions per launch of `kernel` (SQ_INSTS_VALU, committed summary)."""
    v = _profile_entry(kernel).get("valu_wave_instr")
    return None if v is None else float(v)


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
    # the same restatement on one core, over a few more steps of the same run
    one_done, one_elapsed = 0, 0.0
    while one_done < 5 and one_elapsed < budget_s / 2:
        t0 = time.perf_counter()
        st.step_timed(dt, 1, 1)
        one_elapsed += time.perf_counter() - t0
        one_done += 1
    return {
        "value": n * done / elapsed,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {done} steps of {WORKLOAD}, C restatement -O3, OpenMP over particles, {elapsed:.1f} s",
        "single_thread": {"value": n * one_done / one_elapsed, "cores": 1,
                          "sample": f"next {one_done} steps on one core, {one_elapsed:.1f} s"},
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
        "tuning": [tuning["density_variant"], tuning["update_variant"], tuning["neighbour_list"]],
    }


MFMA_BF16_PEAK_TFLOPS = 2500.0


def scale_section(steps: int, warmup: int = 5) -> dict:
    """The 17,777,216-particle scene (256^3 block plus a falling 1M ball, R = 250): graph-launched
    steps per second and the per-pass durations, with the variants the autotuner picks there."""
    from watercube_amd import Fluid, scenes

    p, parts, dt, _ = scenes.workload("block_sphere_17m")
    n = int(p.num_particles)
    with Fluid(p, parts) as f:
        tuning = f.tune(dt, reps=2)
        f.step(dt, warmup)
        f.synchronize()
        t0 = time.perf_counter()
        f.step(dt, steps)
        f.synchronize()
        elapsed = time.perf_counter() - t0
        prof = f.profile(dt, 5)
    return {"workload": "block_sphere_17m", "particles": n, "grid_res": int(p.grid_res), "steps": steps,
            "value": n * steps / elapsed, "unit": "particle-steps/s", "ms_per_step": elapsed * 1e3 / steps,
            "pass_ms": {k: v / 5 for k, v in prof.items()},
            "tuning": [tuning["variants"][tuning["density_variant"]], tuning["variants"][tuning["update_variant"]],
                       tuning["neighbour_list"]]}


def learned_section(p, parts, dt, weights: Path, horizon: int = 500, reps: int = 20) -> dict:
    """Large steps on the 1,000,000-particle block over `horizon` small steps of dt: the plain dt
    run (the reference), and for K = 2, 4, 8 each method of closure.METHODS -- plain, frozen-force
    substeps, learned, learned on frozen-force substeps; the learned ones judged, with the
    network's correction averaged over neighbours and each particle's own -- plus the exact
    correction every network is bounded by.  For each: simulated time per wall second, the peak
    density's excess over the reference run (the density error the methods are ranked by), the
    judge's measure and per-particle errors (closure.evaluate).  Then the fastest method within 1 %
    peak-density error, and the matrix-core utilisation of the inference kernel (116,224 FLOP per
    particle over its measured duration, against the dense bf16 peak) and of a training step
    (298,752 FLOP per row: forward, weight gradients, input gradients of layers 2-4)."""
    import ctypes as C

    from watercube_amd import Fluid
    from watercube_amd import _lib as L
    from watercube_amd import closure

    log = lambda m: print(f"# learned: {m}", file=sys.stderr, flush=True)  # noqa: E731
    net = closure.ClosureNet(max_batch=1024).load(weights)
    ks = (2, 4, 8)
    res = closure.evaluate(net, WORKLOAD, horizon=horizon, ks=ks, judge_tol=0.01, smooth=True, exact=True, log=log)
    raw = closure.evaluate(net, WORKLOAD, horizon=horizon, ks=ks, judge_tol=0.01, smooth=False, log=log)
    for K in ks:
        for m in ("learned", "learned_frozen"):
            res[f"K{K}"][f"{m}_own"] = {k: v for k, v in raw[f"K{K}"][m].items()}
            res[f"K{K}"][f"{m}_own"]["one_step_ratio"] = raw[f"K{K}"]["one_step"]["ratio_to_plain_large"][m]
    n = int(p.num_particles)
    x = closure.DeviceArray((n, closure.NFEAT), np.uint16)
    y = closure.DeviceArray((n, 4), np.float32)
    with Fluid(p, parts) as f:
        f.sort_.run()
        f.runDensityProg()
        f.synchronize()
        L.check(L.lib().wc_closure_features(f.ctx, C.c_float(4.0), None, n, x.ptr), "wc_closure_features")
    # the learned macro step's form: inputs formed inside the inference kernel from the lists
    with Fluid(p, parts) as f:
        f.set_tuning(4, 8, 1)
        f.sort_.run()
        f.runDensityProg()
        f.synchronize()
        fused_s = []
        for _ in range(6):
            t0 = time.perf_counter()
            L.check(L.lib().wc_closure_infer(f.ctx, net.h, C.c_float(4.0), y.ptr), "wc_closure_infer")
            fused_s.append(time.perf_counter() - t0)
    fused_ms = 1e3 * float(np.median(fused_s[1:]))
    ms = net.profile_predict(x, y, n, reps=reps)
    x.free()
    y.free()
    flops = closure.mlp_flops_per_particle() * n
    tflops = flops / (ms * 1e-3) / 1e12
    res["inference"] = {"particles": n, "ms": ms, "tflops": tflops, "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                        "mfma_utilisation": tflops / MFMA_BF16_PEAK_TFLOPS,
                        "flop_per_particle": closure.mlp_flops_per_particle()}
    fused_tflops = flops / (fused_ms * 1e-3) / 1e12
    res["inference_from_lists"] = {"particles": n, "ms": fused_ms, "clock": "host, per call, selection included",
                                   "tflops": fused_tflops, "mfma_utilisation": fused_tflops / MFMA_BF16_PEAK_TFLOPS}
    # training: forward, backward and Adam through the layer-by-layer GEMM kernels, on random rows
    rng = np.random.default_rng(0)
    res["training"] = {}
    for batch in (8192, 65536):
        tnet = closure.ClosureNet(max_batch=batch)
        xt = closure.DeviceArray((batch, closure.NFEAT), np.uint16)
        tt = closure.DeviceArray((batch, 4), np.float32)
        xf = rng.standard_normal((batch, closure.NFEAT)).astype(np.float32)
        xf[:, 195:] = 0.0
        xt.upload((xf.view(np.uint32) >> 16).astype(np.uint16))
        tt.upload(rng.standard_normal((batch, 4)).astype(np.float32) * 0.1)
        t_ms = tnet.profile_train(xt, tt, batch, lr=1e-4, reps=reps)
        tnet.close()
        xt.free()
        tt.free()
        t_tflops = closure.mlp_train_flops_per_particle() * batch / (t_ms * 1e-3) / 1e12
        res["training"][str(batch)] = {"ms_per_step": t_ms, "tflops": t_tflops,
                                       "mfma_utilisation": t_tflops / MFMA_BF16_PEAK_TFLOPS}
    res["training"]["flop_per_row"] = closure.mlp_train_flops_per_particle()
    res["weights"] = str(weights.relative_to(ROOT)) if weights.is_relative_to(ROOT) else str(weights)

    # candidates: (method, K, entry); the plain dt run is the reference (no error by definition)
    cands = [("plain", 1, {"peak_excess": 0.0, **res["plain"], "speedup_vs_plain": 1.0})]
    for K in ks:
        for m, e in res[f"K{K}"].items():
            if isinstance(e, dict) and "speedup_vs_plain" in e:
                cands.append((m, K, e))

    def best(learned_only=False):
        ok = [c for c in cands if abs(c[2]["peak_excess"]) <= 0.01 and (not learned_only or c[0].startswith("learned"))]
        if not ok:
            return None
        m, K, e = max(ok, key=lambda c: c[2]["sim_s_per_s"])
        return {"method": m, "K": K, "particle_steps_per_s": e["sim_s_per_s"] / dt * n,
                "speedup_vs_plain": e["speedup_vs_plain"], "peak_excess": e["peak_excess"]}

    # the fastest of all methods, and of the learned ones alone, within 1 % peak-density error
    res["best_within_1pct"] = {"all": best(), "learned": best(learned_only=True)}
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch-steps", type=int, default=500, help="steps of the 256-scene batch (0 skips it)")
    ap.add_argument("--weights", default=str(ROOT / "watercube_amd" / "weights" / "closure.npy"),
                    help="closure network weights for the learned section (skipped if absent)")
    ap.add_argument("--learned-horizon", type=int, default=500)
    ap.add_argument("--scale-steps", type=int, default=50, help="steps of the 17.8M-particle scene (0 skips it)")
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
              f"update {tuning['variants'][tuning['update_variant']]} neighbour lists {tuning['neighbour_list']}",
              file=sys.stderr, flush=True)
        print("# variant ms density: " + " ".join(f"{x:.4f}" for x in tuning["density_ms"]) +
              " | update: " + " ".join(f"{x:.4f}" for x in tuning["update_ms"]), file=sys.stderr, flush=True)
        print("# with lists, density: " + " ".join(f"{x:.4f}" for x in tuning["density_list_ms"]) +
              " | update: " + " ".join(f"{x:.4f}" for x in tuning["update_list_ms"]), file=sys.stderr, flush=True)
        f.step(dt, args.warmup)
        f.synchronize()
        t0 = time.perf_counter()
        f.step(dt, args.steps)
        f.synchronize()
        elapsed = time.perf_counter() - t0
        print(f"# timed {args.steps} steps in {elapsed:.4f} s", file=sys.stderr, flush=True)

        # per-kernel durations: HIP events on the context's stream around each launch, over the
        # steps the committed counter profiles cover (steps 521-570), so both describe one state
        done = args.warmup + args.steps
        if done < PROFILE_FROM_STEP:
            f.step(dt, PROFILE_FROM_STEP - done)
            done = PROFILE_FROM_STEP
        prof = f.profile(dt, args.profile_steps)
        f.synchronize()
        state = f.particles(1)
        list_entries = 0.0
        if tuning["neighbour_list"]:
            cnt = f.neighbour_counts()
            list_entries = float(np.where(cnt == 0xFFFFFFFF, 0, cnt).mean())
    if not np.isfinite(state["position"]).all():
        print("bench: non-finite positions after the run", file=sys.stderr)
        return 1
    prof_steps = f"{done + 1}-{done + args.profile_steps}"

    per_launch_ms = {k: v / args.profile_steps for k, v in prof.items()}
    print("# per-launch ms: " + ", ".join(f"{k}={v:.4f}" for k, v in per_launch_ms.items()), file=sys.stderr,
          flush=True)
    # the reorder is two kernels here (place, gather); its algorithmic bytes are counted once
    pass_ms = {"hash": per_launch_ms["hash"], "scan": per_launch_ms["scan"],
               "scatter": per_launch_ms["place"] + per_launch_ms["gather"],
               "density": per_launch_ms["density"], "update": per_launch_ms["update"]}
    algo = algorithmic_bytes(n, ncells)
    extra = overhead_bytes(n, list_entries)
    dom = max(per_launch_ms, key=per_launch_ms.get)
    dom_s = per_launch_ms[dom] * 1e-3
    achieved = algo[dom] / dom_s / 1e9

    def measured(kernels):
        vals = [read_traffic(k) for k in kernels]
        return None if any(v is None for v in vals) else sum(vals)

    pass_kernels = {"hash": ["hash"], "scan": ["scan"], "scatter": ["place", "gather"], "density": ["density"],
                    "update": ["update"]}
    # per pass: algorithmic bytes and rocprofv3-measured HBM bytes (committed summary of the same
    # steps) over the live per-launch duration, in GB/s
    pass_gbs = {k: [round(algo[k] / (v * 1e-3) / 1e9, 1),
                    None if measured(pass_kernels[k]) is None else round(measured(pass_kernels[k]) / (v * 1e-3) / 1e9, 1)]
                for k, v in pass_ms.items() if v > 0}
    print(f"# per-pass GB/s [algorithmic, measured]: {json.dumps(pass_gbs)}", file=sys.stderr, flush=True)
    traffic = measured([dom])
    valu = read_valu(dom)
    valu_frac = None if valu is None else valu / dom_s / VALU_PEAK_WAVE_INSTR_PER_S
    out = {
        "metric": "particle-steps/s at 1M particles: plain vs learned large-dt, \u22641% density err",
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
                              "update": tuning["variants"][tuning["update_variant"]],
                              "neighbour_list": tuning["neighbour_list"]}},
        "simulated_s_per_s": dt * args.steps / elapsed,
        "pass_ms": per_launch_ms,
        "pass_ms_steps": prof_steps,
        "roofline": {
            "kernel": dom,
            # the density pass is bound by its instruction issue, not by HBM: `frac` is its
            # algorithmic bytes over the HBM peak, `valu_frac` its VALU issue rate over one wave64
            # instruction per CU per cycle
            "bound": "valu",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "hbm_frac": achieved / HBM_PEAK_GBS,
            "valu_frac": valu_frac,
            "traffic": traffic,
            "traffic_source": f"profiles/traffic_{WORKLOAD}.json (steps 521-570)",
            "duration_us": per_launch_ms[dom] * 1e3,
            "duration_steps": prof_steps,
            "algorithmic_bytes": algo[dom],
            "overhead_bytes": extra.get(dom),
            "list_entries": list_entries,
            "valu_wave_instr": valu,
        },
    }
    if args.batch_steps > 0:
        bt = batched_section(args.batch_steps)
        print(f"# batched: {bt}", file=sys.stderr, flush=True)
        out["batched"] = {k: bt[k] for k in ("workload", "particles", "value", "ms_per_step", "scene_latency_ms_p50",
                                             "scene_latency_ms_p99")}
    if args.scale_steps > 0:
        sc = scale_section(args.scale_steps)
        print(f"# scale: {sc}", file=sys.stderr, flush=True)
        out["scale"] = {k: sc[k] for k in ("workload", "particles", "value", "ms_per_step")}
    weights = Path(args.weights)
    if weights.exists():
        learned = learned_section(p, parts, dt, weights, horizon=args.learned_horizon)
        # the whole comparison goes to stderr; the line keeps its summary
        print(f"# learned: {json.dumps(learned)}", file=sys.stderr, flush=True)
        out["learned"] = learned_summary(learned)
    else:
        print(f"# learned section skipped: no weights at {weights}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, parts, dt)
    print(json.dumps(_rounded(out)), flush=True)
    return 0


def _rounded(x, digits: int = 4):
    """Floats to `digits` significant digits (the line stays short enough for a log tail); the
    headline value and step time keep full precision."""
    if isinstance(x, float):
        return float(f"{x:.{digits}g}")
    if isinstance(x, dict):
        return {k: (v if k in ("value", "ms_per_step") and isinstance(v, float) else _rounded(v, digits))
                for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_rounded(v, digits) for v in x]
    return x


def learned_summary(res: dict) -> dict:
    """The learned comparison's headline figures: per K and method [peak-density excess over the
    plain dt run, speed-up over it, fallbacks]; the exact-correction bound's excess; the one-step
    velocity-error ratio of the learned step to the plain large step; the fastest methods within
    1 % peak-density error; the network's matrix-core use in inference and in training (the whole
    comparison goes to stderr)."""
    out = {"horizon_steps": res.get("horizon_steps")}
    for key, entry in res.items():
        if not (key.startswith("K") and isinstance(entry, dict)):
            continue
        k = {}
        for m, e in entry.items():
            if m == "one_step":
                k["one_step_vel_ratio"] = e["ratio_to_plain_large"]["learned"]["vel"]
            elif m == "exact_correction":
                k[m] = e["peak_excess"]
            elif isinstance(e, dict) and "peak_excess" in e:
                k[m] = [e["peak_excess"], e["speedup_vs_plain"], e.get("fallbacks", 0)]
        out[key] = k
    best = res.get("best_within_1pct") or {}
    out["best_within_1pct"] = {k: (None if v is None else [v["method"], v["K"], v["speedup_vs_plain"]])
                               for k, v in best.items()}
    for key in ("inference", "inference_from_lists"):
        if key in res:
            out[f"{key}_mfma_utilisation"] = res[key]["mfma_utilisation"]
    if "training" in res:
        out["training_mfma_utilisation"] = {b: e["mfma_utilisation"] for b, e in res["training"].items()
                                            if isinstance(e, dict)}
    return out


if __name__ == "__main__":
    sys.exit(main())
