"""Benchmark of the bit-plane CA step on one MI355X.

Workload: the plain-GPU configuration (one grid at a time): Game of Life (Bernoulli 0.5) and
WireWorld (80/15/3/2 % empty/conductor/head/tail), 16384 x 16384 cells, seeded i.i.d.,
K generations timed after W warm-up generations, with the simple traverser and the temporal
traverser at depths 1/2/4/8 (linear layout, 64-bit words). A "step" is one generation of the
whole grid. `value` is the best Game-of-Life variant in GCUPS (cell updates per second / 1e9)
with the grid resident in HBM; every measured variant is re-run on a 2048^2 subset and
compared bit for bit with the CPU oracle.

Prints one JSON line (see the fields below); per-variant results are in "variants".
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "2026-cellato-smpt-journal_amd"))
sys.path.insert(0, str(REPO / "tests"))

HBM_PEAK_GBPS = 8000.0
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated against a
# plain copy kernel); see profiles/stream_kernel_gol_cyclic_16384_hbm_traffic.md
TRAFFIC_FILE = REPO / "profiles" / "hbm_traffic_16384.json"


def measured_traffic(automaton, traverser, T, size):
    try:
        data = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    for e in data["entries"]:
        if (e["automaton"], e["traverser"], e["temporal_steps"], e["size"]) == (automaton, traverser, T, size):
            return int(e["traffic_bytes"])
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="generations timed per variant")
    ap.add_argument("--warmup", type=int, default=50, help="untimed generations per variant")
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--automata", default="game-of-life,wire")
    ap.add_argument("--depths", default="1,2,4,8", help="temporal depths to measure")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-check", action="store_true", help="skip the 2048^2 bit-exactness check")
    ap.add_argument("--no-simple", action="store_true", help="measure the temporal depths only")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also time a plain device copy of this many MiB (the calibration kernel for "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def time_variant(C, torch, rule, dev, kind, T, steps, warmup, chunk):
    """GCUPS and per-launch average (ms) of one traverser variant; the grid stays on the device."""
    stream = torch.cuda.current_stream().cuda_stream
    W, H = dev.width, dev.height

    def advance(n, t0):
        if n <= 0:
            return
        if kind == "simple":
            in_b = C._lib.simple(rule.id, 0, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), W, H, t0, n, chunk, stream)
        else:
            in_b = C._lib.temporal(rule.id, 0, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), W, H, t0, n, T, chunk, stream)
        if in_b:
            dev.cur = 1 - dev.cur

    advance(warmup, 0)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    advance(steps, warmup)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    launches = steps if kind == "simple" else -(-steps // T)
    return W * H * steps / (ms * 1e-3) / 1e9, ms, ms / launches, launches


def check_subset(C, O, rule_name, kind, T):
    from gpu_util import run_gpu
    v = C.iid_values(rule_name, 2048, 2048, 11)
    want = O.run_bitplanes(rule_name, v, 64)
    import numpy as np
    return bool(np.array_equal(run_gpu(rule_name, v, 0, kind, T, 64), want))


def cpu_baseline(C, O):
    """The oracle's bit_planes evaluator on the CPU simple traverser (the reference's CPU path,
    single-threaded), timed on a bounded sample of the same workload."""
    import numpy as np
    n = 16384
    v = C.iid_values("game-of-life", n, n, 7)
    steps = 4
    t = time.perf_counter()
    O.run_bitplanes("game-of-life", v, steps, threads=1)
    one = n * n * steps / (time.perf_counter() - t) / 1e9
    cores = os.cpu_count() or 1
    threads = min(cores, 16)
    t = time.perf_counter()
    O.run_bitplanes("game-of-life", v, 16, threads=threads)
    many = n * n * 16 / (time.perf_counter() - t) / 1e9
    vs = np.ascontiguousarray(v[:1024, :1024])
    t = time.perf_counter()
    O.run_standard("game-of-life", vs, 4, threads=1)
    std = 1024 * 1024 * 4 / (time.perf_counter() - t) / 1e9
    return {"value": round(one, 3), "unit": "GCUPS", "cores": 1, "kind": "port",
            "sample": "game-of-life 16384^2 i.i.d. p=0.5, 4 generations, bit_planes evaluator, 1 thread "
                      "(oracle restatement of the reference CPU simple traverser)",
            "all_cores": {"value": round(many, 3), "threads": threads, "host_cpus": cores,
                          "sample": "same grid, 16 generations, rows split over threads"},
            "standard_evaluator_1core": {"value": round(std, 4), "sample": "1024^2, 4 generations"}}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (each box has one MI355X)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import cellato_amd as C
    import oracle_lib as O

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    variants = []
    best = None
    depths = [int(d) for d in a.depths.split(",") if d]
    for name in a.automata.split(","):
        rule = C.rule(name)
        v = C.iid_values(name, a.size, a.size, 2026)
        planes = C.grid.pack(rule.index_of(v), rule.planes, 0)
        del v
        dev = C.DeviceGrid(rule, a.size, a.size, 0)
        for kind, T in ([] if a.no_simple else [("simple", 1)]) + [("temporal", d) for d in depths]:
            dev.upload(planes)
            dev.cur = 0
            gcups, ms, ms_launch, launches = time_variant(C, torch, rule, dev, kind, T, a.steps, a.warmup, a.chunk)
            algo_bytes = 2 * rule.planes * a.size * a.size / 8   # P bits read + P bits written per cell and launch
            rec = {"automaton": name, "traverser": kind, "temporal_steps": T, "gcups": round(gcups, 1),
                   "ms_total": round(ms, 3), "ms_per_launch": round(ms_launch, 5), "launches": launches,
                   "algorithmic_GBps": round(algo_bytes / (ms_launch * 1e-3) / 1e9, 1)}
            if not a.no_check:
                rec["bit_exact_2048_subset"] = check_subset(C, O, name, kind, T)
            print(json.dumps(rec), file=sys.stderr, flush=True)
            variants.append(rec)
            if name == "game-of-life" and (best is None or gcups > best["gcups"]):
                best = rec
        del dev
        torch.cuda.empty_cache()

    if a.copy_mib:
        n = a.copy_mib * 2**20 // 8
        src = torch.empty(n, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(20):
            dst.copy_(src)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / 20
        copy_rec = {"copy_MiB": a.copy_mib, "GBps": round(2 * a.copy_mib * 2**20 / ms / 1e6, 1)}
        print(json.dumps(copy_rec), file=sys.stderr, flush=True)
        del src, dst

    g = best or variants[0]
    achieved = g["algorithmic_GBps"]
    simple_gol = next((r for r in variants if r["automaton"] == g["automaton"] and r["traverser"] == "simple"), None)
    out = {
        "metric": "GCUPS (cell updates/s) at 16384^2 bit-exact; batched/served vs plain GPU; p99",
        "value": g["gcups"],
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(g["ms_total"] / a.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"{g['automaton']} {a.size}x{a.size}, i.i.d. seeded, {a.steps} generations, "
                               f"plain GPU path (one grid), linear bit planes, 64-bit words",
                   "traverser": g["traverser"], "temporal_steps": g["temporal_steps"]},
        "bit_exact": all(r.get("bit_exact_2048_subset", True) for r in variants),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": measured_traffic(g["automaton"], g["traverser"], g["temporal_steps"], a.size),
                     "kernel": f"stream_kernel ({g['traverser']}, T={g['temporal_steps']})",
                     "bytes_per_launch": 2 * C.rule(g["automaton"]).planes * a.size * a.size // 8},
        "roofline_simple": None if simple_gol is None else {
            "bound": "hbm", "achieved": simple_gol["algorithmic_GBps"], "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(simple_gol["algorithmic_GBps"] / HBM_PEAK_GBPS, 4)},
        "variants": variants,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(C, O)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
