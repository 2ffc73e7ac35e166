"""Summaries of rocprofv3 output for the SPH step's kernels.

    python3 -m watercube_amd.profile_summary --trace DIR [--fetch DIR --write DIR] [--sq DIR] \
        --workload block_1m --command "..." --out profiles/

Reads the CSV files rocprofv3 writes (``*_kernel_stats.csv`` of a ``--kernel-trace --stats`` run,
``*_counter_collection.csv`` of ``--pmc`` runs) and writes short text summaries plus
``traffic_<workload>.json``, the HBM bytes per launch of each kernel that ``bench.py`` reports as
``roofline.traffic``.

Byte counters are calibrated against the record-split copy kernel ``k_split`` that every upload
runs: it reads and writes exactly 32 bytes per particle, so the factor that maps its counter
reading onto 32 N bytes maps every other kernel's reading onto bytes.  SQ instruction counters
are chip totals; ``SQ_INSTS_VALU`` over the launch's duration against 256 CUs x 2.4 GHz (one
wave64 VALU instruction per CU per cycle without packed math) gives the VALU issue utilisation.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import sys
from pathlib import Path

SIMDS = 256 * 4
CLOCK_HZ = 2.4e9
VALU_WAVE_INSTR_PER_S = 256 * CLOCK_HZ
STEP_KERNELS = ("k_hash", "k_scan", "k_place", "k_gather", "k_density", "k_update")


def _one(d: Path, pattern: str) -> Path:
    hits = sorted(Path(d).rglob(pattern))
    if not hits:
        raise FileNotFoundError(f"no {pattern} under {d}")
    return hits[0]


def kernel_stats(trace_dir: Path, last: int = 0) -> dict:
    """Per-kernel call count and durations: the stats file, or with `last` > 0 the last `last`
    dispatches of each kernel in the per-dispatch trace."""
    out = {}
    if last <= 0:
        with open(_one(trace_dir, "*kernel_stats.csv")) as fh:
            for r in csv.DictReader(fh):
                out[r["Name"]] = {"calls": int(r["Calls"]), "avg_ns": float(r["AverageNs"]),
                                  "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"]),
                                  "percent": float(r["Percentage"])}
        return out
    durs = collections.defaultdict(list)
    with open(_one(trace_dir, "*kernel_trace.csv")) as fh:
        rows = sorted(csv.DictReader(fh), key=lambda r: int(r["Start_Timestamp"]))
    for r in rows:
        durs[r["Kernel_Name"]].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    total = sum(sum(v[-last:]) for v in durs.values())
    for k, v in durs.items():
        v = v[-last:]
        out[k] = {"calls": len(v), "avg_ns": sum(v) / len(v), "min_ns": float(min(v)), "max_ns": float(max(v)),
                  "percent": 100.0 * sum(v) / max(total, 1)}
    return out


def counters(pmc_dir: Path) -> dict:
    """Mean value per dispatch of every counter, per kernel."""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    with open(_one(pmc_dir, "*counter_collection.csv")) as fh:
        for r in csv.DictReader(fh):
            acc[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in acc.items()}


def summarise(trace: Path, fetch: Path | None, write: Path | None, sq: Path | None, workload: str, particles: int,
              command: str, out: Path, kernels=STEP_KERNELS, last: int = 0) -> dict:
    ks = kernel_stats(trace, last)
    out.mkdir(parents=True, exist_ok=True)
    traffic = {"workload": workload, "particles": particles, "command": command, "kernels": {}}
    if fetch is not None and write is not None:
        traffic = _traffic(ks, counters(fetch), counters(write), workload, particles, command, out)
    _kernel_table(ks, workload, particles, command, out)
    if sq is not None:
        _sq_table(ks, counters(sq), traffic, workload, particles, command, out, kernels)
        if fetch is not None and write is not None:
            (out / f"traffic_{workload}.json").write_text(json.dumps(traffic, indent=1) + "\n")
    return traffic


def _traffic(ks, f, w, workload, particles, command, out) -> dict:
    copy_bytes = 32.0 * particles
    f_factor = copy_bytes / (f["k_split"]["FETCH_SIZE"] * 1024.0)
    w_factor = copy_bytes / (w["k_split"]["WRITE_SIZE"] * 1024.0)
    traffic = {"workload": workload, "particles": particles,
               "calibration": {"kernel": "k_split", "bytes_read": copy_bytes, "bytes_written": copy_bytes,
                               "fetch_factor": f_factor, "write_factor": w_factor},
               "command": command, "kernels": {}}
    lines = [f"# HBM traffic per launch, {workload} ({particles} particles)", "",
             f"Command: `{command}`", "",
             f"Calibration: `k_split` moves {copy_bytes / 1e6:.1f} MB each way; FETCH_SIZE reads "
             f"{f['k_split']['FETCH_SIZE']:.0f} KiB (factor {f_factor:.3f}), WRITE_SIZE reads "
             f"{w['k_split']['WRITE_SIZE']:.0f} KiB (factor {w_factor:.3f}).", "",
             "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | read MB | written MB | HBM MB per launch |",
             "|---|---|---|---|---|---|"]
    for k in STEP_KERNELS:
        if k not in f or k not in w:
            continue
        rd = f[k]["FETCH_SIZE"] * 1024.0 * f_factor
        wr = w[k]["WRITE_SIZE"] * 1024.0 * w_factor
        traffic["kernels"][k] = {"fetch_kib": f[k]["FETCH_SIZE"], "write_kib": w[k]["WRITE_SIZE"],
                                 "read_bytes": rd, "write_bytes": wr, "hbm_bytes_per_launch": rd + wr}
        lines.append(f"| {k} | {f[k]['FETCH_SIZE']:.0f} | {w[k]['WRITE_SIZE']:.0f} | {rd / 1e6:.2f} | "
                     f"{wr / 1e6:.2f} | {(rd + wr) / 1e6:.2f} |")
    (out / f"traffic_{workload}.json").write_text(json.dumps(traffic, indent=1) + "\n")
    (out / f"traffic_{workload}.md").write_text("\n".join(lines) + "\n")
    return traffic


def _kernel_table(ks, workload, particles, command, out) -> None:
    lines = [f"# Kernel durations, {workload} ({particles} particles)", "", f"Command: `{command}`", "",
             "| kernel | calls | average us | min us | max us | % of GPU time |", "|---|---|---|---|---|---|"]
    for k, v in sorted(ks.items(), key=lambda kv: -kv[1]["percent"]):
        lines.append(f"| {k} | {v['calls']} | {v['avg_ns'] / 1e3:.2f} | {v['min_ns'] / 1e3:.2f} | "
                     f"{v['max_ns'] / 1e3:.2f} | {v['percent']:.2f} |")
    (out / f"kernels_{workload}.md").write_text("\n".join(lines) + "\n")


def _sq_table(ks, s, traffic, workload, particles, command, out, kernels) -> None:
    lines = [f"# SQ instruction counters per launch, {workload} ({particles} particles)", "",
             f"Command: `{command}`", "",
             "Chip totals per dispatch.  VALU utilisation = SQ_INSTS_VALU / (duration x 256 CUs x 2.4 GHz), "
             "duration from the kernel trace above.", "",
             "| kernel | waves | VALU | SALU | VMEM rd | LDS | VALU per wave | VALU utilisation |",
             "|---|---|---|---|---|---|---|---|"]
    for k in kernels:
        if k not in s or k not in ks:
            continue
        c = s[k]
        util = c["SQ_INSTS_VALU"] / (ks[k]["avg_ns"] * 1e-9 * VALU_WAVE_INSTR_PER_S)
        traffic["kernels"].setdefault(k, {})["valu_wave_instr"] = c["SQ_INSTS_VALU"]
        traffic["kernels"][k]["valu_utilisation"] = util
        lines.append(f"| {k} | {c['SQ_WAVES']:.0f} | {c['SQ_INSTS_VALU']:.4g} | {c['SQ_INSTS_SALU']:.4g} | "
                     f"{c['SQ_INSTS_VMEM_RD']:.4g} | {c['SQ_INSTS_LDS']:.4g} | "
                     f"{c['SQ_INSTS_VALU'] / max(c['SQ_WAVES'], 1):.0f} | {util:.2f} |")
    mfma = [k for k in kernels if k in s and k in ks and s[k].get("SQ_INSTS_MFMA", 0) > 0
            and "SQ_VALU_MFMA_BUSY_CYCLES" in s[k]]
    if mfma:
        lines += ["", "Matrix cores: SQ_VALU_MFMA_BUSY_CYCLES is summed over the 1,024 SIMDs, so busy = it / "
                  "(duration x 1,024 x 2.4 GHz).  Transcendental instructions are SQ_INSTS_VALU_TRANS_F32.", "",
                  "| kernel | MFMA instructions | MFMA per wave | matrix cores busy | transcendental per wave |",
                  "|---|---|---|---|---|"]
        for k in mfma:
            c = s[k]
            busy = c["SQ_VALU_MFMA_BUSY_CYCLES"] / (ks[k]["avg_ns"] * 1e-9 * SIMDS * CLOCK_HZ)
            trans = c.get("SQ_INSTS_VALU_TRANS_F32")
            trans_s = f"{trans / max(c['SQ_WAVES'], 1):.0f}" if trans is not None else "-"
            lines.append(f"| {k} | {c['SQ_INSTS_MFMA']:.4g} | {c['SQ_INSTS_MFMA'] / max(c['SQ_WAVES'], 1):.0f} | "
                         f"{busy:.2f} | {trans_s} |")
    (out / f"sq_{workload}.md").write_text("\n".join(lines) + "\n")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True, type=Path)
    ap.add_argument("--fetch", type=Path, default=None)
    ap.add_argument("--write", type=Path, default=None)
    ap.add_argument("--kernels", default=",".join(STEP_KERNELS), help="kernels of the SQ table")
    ap.add_argument("--sq", type=Path, default=None)
    ap.add_argument("--workload", default="block_1m")
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--command", required=True)
    ap.add_argument("--out", type=Path, default=Path("profiles"))
    ap.add_argument("--last", type=int, default=0, help="average only each kernel's last N dispatches of the trace")
    a = ap.parse_args(argv)
    t = summarise(a.trace, a.fetch, a.write, a.sq, a.workload, a.particles, a.command, a.out,
                  kernels=tuple(a.kernels.split(",")), last=a.last)
    print(json.dumps({k: round(v.get("hbm_bytes_per_launch", 0) / 1e6, 2) for k, v in t["kernels"].items()}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
