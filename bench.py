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
# HBM bytes and VALU instructions per launch measured with rocprofv3 (FETCH_SIZE / WRITE_SIZE
# calibrated against a plain copy kernel, SQ_INSTS_VALU); see profiles/stream_kernel_hbm_traffic.md
# and profiles/stream_kernel_16384_valu.md
TRAFFIC_FILE = REPO / "profiles" / "hbm_traffic.json"
# VALU issue peak: one wave64 instruction per SIMD every 4 cycles, 256 CUs x 4 SIMDs, 2.4 GHz
VALU_PEAK_INSTR_PER_S = 256 * 4 * 2.4e9 / 4


def measured_launch(automaton, traverser, T, size, key="traffic_bytes", layout="linear"):
    try:
        data = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    for e in data["entries"]:
        if (e["automaton"], e["layout"], e["traverser"], e["temporal_steps"], e["size"]) == \
                (automaton, layout, traverser, T, size):
            return e[key] if key in e else None
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
    ap.add_argument("--layout", choices=["linear", "tiled"], default="linear",
                    help="bit-plane layout of the variants phase (the headline is linear)")
    ap.add_argument("--rounds", type=int, default=5, help="timed rounds per variant, each from the initial state")
    ap.add_argument("--roofline-launches", type=int, default=200,
                    help="back-to-back launches timed for the headline kernel's average launch duration")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-check", action="store_true", help="skip the 2048^2 bit-exactness check")
    ap.add_argument("--no-simple", action="store_true", help="measure the temporal depths only")
    ap.add_argument("--no-batch", action="store_true", help="skip the batching measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--batch-grids", type=int, default=512)
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--tick-passes", type=int, default=16, help="passes per live job and serving tick")
    ap.add_argument("--serve-baseline-limit", type=float, default=45.0,
                    help="seconds after which the one-at-a-time baseline stops admitting requests")
    ap.add_argument("--no-multiplane", action="store_true", help="skip the multi-plane rules")
    ap.add_argument("--multiplane", default="cyclic,excitable,fluid,critters")
    ap.add_argument("--multiplane-steps", type=int, default=400)
    ap.add_argument("--no-bit-array", action="store_true", help="skip the bit_array layout")
    ap.add_argument("--bit-array-steps", type=int, default=128)
    ap.add_argument("--phase", choices=["variants", "multiplane", "bit_array", "batch", "serve", "cpu", "table"],
                    default=None,
                    help="run one phase in this process (bench.py runs each phase in a child process)")
    ap.add_argument("--tuner-cache", default=None, help="autotuner measurements shared by the phases")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also time a plain device copy of this many MiB (the calibration kernel for "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def time_variant(C, torch, rule, dev, init, kind, T, steps, warmup, rounds, chunk):
    """One traverser variant timed as the reference times a suite (experiment_manager.hpp:61-148):
    every round starts from the same initial state (`init`, copied in outside the timed region),
    `warmup` generations run untimed first, then `rounds` timed rounds of `steps` generations,
    each bracketed by HIP events on the launch stream, and each round's result is checksummed on
    the device (16 tiles); the rounds must agree. The grid stays on the device throughout."""
    stream = torch.cuda.current_stream()
    W, H = dev.width, dev.height

    def reset():
        dev.buf[0].copy_(init)
        dev.cur = 0

    def advance(n, t0):
        if n <= 0:
            return
        if kind == "simple":
            in_b = C._lib.simple(rule.id, dev.layout, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), W, H, t0, n, chunk,
                                 stream.cuda_stream)
        else:
            in_b = C._lib.temporal(rule.id, dev.layout, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), W, H, t0, n, T,
                                   chunk, stream.cuda_stream)
        if in_b:
            dev.cur = 1 - dev.cur

    reset()
    advance(warmup, 0)
    torch.cuda.synchronize()
    times, sums = [], []
    for _ in range(rounds):
        reset()
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        advance(steps, 0)
        stop.record(stream)
        stop.synchronize()
        times.append(start.elapsed_time(stop))
        sums.append(dev.checksum_device())
    import numpy as np
    ms = float(np.mean(times))
    launches = steps if kind == "simple" else -(-steps // T)
    return {"gcups": W * H * steps / (ms * 1e-3) / 1e9, "ms_mean": ms, "ms_std": float(np.std(times)),
            "ms_rounds": [round(t, 4) for t in times], "ms_per_launch": ms / launches, "launches": launches,
            "checksum_consistent": all(x == sums[0] for x in sums), "checksum": sums[0]}


def time_launches(C, torch, rule, dev, kind, T, launches, chunk):
    """Average duration of one launch of a variant's kernel: `launches` launches back to back on
    one stream, bracketed by HIP events (the launch-to-launch time the kernel trace reports as its
    average duration; short windows carry the slower first launches)."""
    stream = torch.cuda.current_stream()
    W, H = dev.width, dev.height
    n = launches * (1 if kind == "simple" else T)

    def advance(k):
        if kind == "simple":
            in_b = C._lib.simple(rule.id, dev.layout, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), W, H, 0, k, chunk,
                                 stream.cuda_stream)
        else:
            in_b = C._lib.temporal(rule.id, dev.layout, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), W, H, 0, k, T,
                                   chunk, stream.cuda_stream)
        if in_b:
            dev.cur = 1 - dev.cur

    advance(4 * (1 if kind == "simple" else T))
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    advance(n)
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / launches


_SUBSET_WANT = {}


def check_subset(C, O, rule_name, kind, T, layout=0):
    """The variant on a 2048^2 subset (seeded i.i.d., 64 generations) against the CPU standard
    evaluator (one oracle run per automaton, shared by its variants)."""
    import numpy as np
    from gpu_util import run_gpu
    v = C.iid_values(rule_name, 2048, 2048, 11)
    if rule_name not in _SUBSET_WANT:
        _SUBSET_WANT[rule_name] = O.run_standard(rule_name, v, 64, threads=min(16, os.cpu_count() or 1))
    return bool(np.array_equal(run_gpu(rule_name, v, layout, kind, T, 64), _SUBSET_WANT[rule_name]))


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


def make_tuner(cache_path=None):
    """Depth tuner guided by the cost model, trained on the GPU on the committed MI355X
    dataset; measures the model's two best depths per (automaton, size) bucket (measurements
    kept in `cache_path`, shared by the phases of one bench run)."""
    import numpy as np
    from cellato_amd.autotune import Tuner, features
    from cellato_amd.costmodel import CostModel
    from cellato_amd.tuning_eval import load
    rows, _ = load()
    X = np.array([features(r["rule"], r["width"], r["height"], r["T"], r.get("chunk", 0)) for r in rows], np.float32)
    y = np.log(np.array([r["ns_per_cell"] for r in rows], np.float32))
    t = time.perf_counter()
    model = CostModel(X.shape[1], hidden=(64, 64), seed=0)
    model.fit(X, y, epochs=600)
    train_s = time.perf_counter() - t
    return Tuner(cache_path=cache_path, model=model, budget=2, target_cells=1 << 26), train_s


def bench_batch(C, O, torch, tuner, n):
    """Batching configuration: n grids of five automata, sides log-uniform 64..4096, steps
    16..1024, one graph launch; against the plain per-grid launch loop."""
    import numpy as np
    from cellato_amd.batch import Batch, PlainLoop, batch_workload
    from cellato_amd.rules import device_random_values
    jobs = batch_workload(n, seed=2026)
    cu = sum(j.cells * j.steps for j in jobs)

    def timed(fn):
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        return s.elapsed_time(e) / 1e3

    t = time.perf_counter()
    depths, winner, measured = tuner.batch_depths(jobs)
    tune_s = time.perf_counter() - t
    for j, T in zip(jobs, depths):
        j.temporal_steps = T
    b = Batch(jobs)
    b.run()                          # warm-up replay, then back to the initial states
    b.reset()
    batch_s = timed(b.run)
    stats = b.stats()
    sums_batch = [b.checksum(i) for i in range(len(jobs))]
    small = sorted(range(len(jobs)), key=lambda i: jobs[i].cells * jobs[i].steps)[:8]
    oracle_ok = all(
        np.array_equal(b.result(i), O.run_standard(jobs[i].rule,
                       device_random_values(jobs[i].rule, jobs[i].width, jobs[i].height, jobs[i].seed), jobs[i].steps))
        for i in small)
    b.close()
    del b
    torch.cuda.empty_cache()
    pl = PlainLoop(jobs)
    plain_s = timed(pl.run)
    sums_plain = [pl.grids[i].checksum_device() for i in range(len(jobs))]
    pl.reset()
    for j in jobs:                   # the plain loop with each grid's tuned depth
        tuner.pick(j.rule, j.width, j.height)
    plain_t_s = timed(lambda: pl.run(tuner=tuner))
    del pl
    torch.cuda.empty_cache()
    return {"workload": f"{n} grids, automata uniform over game-of-life/brian/forest-fire/traffic/maze, square sides "
                        "log-uniform 64..4096 (multiple of 64), steps uniform 16..1024, seeded i.i.d. states made on "
                        "the device, one seed",
            "value": round(cu / batch_s / 1e9, 1), "unit": "GCUPS", "cell_updates": cu,
            "ms": round(batch_s * 1e3, 3), "graph_launches": 1, **stats,
            "depths": winner,
            "depths_per_automaton": {k: sorted({T for j, T in zip(jobs, depths) if j.rule == k})
                                     for k in sorted({j.rule for j in jobs})},
            "depth_candidates_gcups": {k: round(v, 1) for k, v in measured.items()},
            "plain_loop_gcups": round(cu / plain_s / 1e9, 1),
            "speedup_vs_plain_loop": round(plain_s / batch_s, 2),
            "plain_loop_tuned_depth_gcups": round(cu / plain_t_s / 1e9, 1),
            "bit_exact_vs_plain_all_grids": sums_batch == sums_plain,
            "bit_exact_vs_cpu_oracle_8_smallest": bool(oracle_ok),
            "tuning_seconds": round(tune_s, 2)}


def bench_serve(C, O, torch, tuner, n, rate, limit, tick_passes):
    """Serving configuration: Poisson arrivals, all ten automata, sides 128..8192, steps
    64..2048, continuous batching under a 200 GB HBM budget; against serving one request at
    a time with the plain simple traverser. Every result is compared with the one-at-a-time
    result by on-device checksums, and the four smallest with the CPU oracle."""
    import numpy as np
    from cellato_amd.serve import Server, serving_workload, serve_one_at_a_time
    from cellato_amd.rules import device_random_values
    reqs = serving_workload(n=n, rate=rate, seed=7)
    srv = Server(hbm_budget=200e9, default_T=4, tick_passes=tick_passes, keep_results=True)
    rep = srv.run(reqs)
    s = rep.summary()
    sample = sorted(reqs, key=lambda r: r.cell_updates)[:4]
    exact = all(np.array_equal(srv.results[r.id].to_plane_grid().to_standard(),
                               O.run_standard(r.rule, device_random_values(r.rule, r.width, r.height, r.seed), r.steps))
                for r in sample)
    sums = {i: g.checksum_device() for i, g in srv.results.items()}
    srv.results.clear()
    torch.cuda.empty_cache()
    for r in reqs:                   # tune outside the served time
        tuner.pick(r.rule, r.width, r.height)
    tuned = Server(hbm_budget=200e9, tuner=tuner, tick_passes=tick_passes).run(reqs).summary()
    torch.cuda.empty_cache()
    base, base_results = serve_one_at_a_time(reqs, keep_results=True, time_limit=limit)
    same = all(sums[i] == g.checksum_device() for i, g in base_results.items())
    base_results.clear()
    torch.cuda.empty_cache()
    b = base.summary()
    done = len(base.latencies)
    # requests the baseline did not reach wait at least until it stopped
    first = min(r.arrival for r in reqs)
    censored = [max(0.0, base.makespan + first - r.arrival) for r in reqs if r.id not in base.latencies]
    lat_all = np.array(list(base.latencies.values()) + censored)
    return {"workload": f"{n} requests, Poisson {rate:g}/s, automaton uniform over the ten, square side log-uniform "
                        "128..8192, steps uniform 64..2048, seeded i.i.d. states made on the device, HBM budget 200 GB",
            "p50_s": round(s["p50_s"], 4), "p99_s": round(s["p99_s"], 4), "gcups": round(s["gcups"], 1),
            "makespan_s": round(s["makespan_s"], 3), "temporal_steps": 4, "tick_passes": tick_passes,
            "ticks": s["ticks"], "max_live": s["max_live"], "peak_GB": round(s["peak_GB"], 2),
            "bit_exact_vs_one_at_a_time": bool(same), "compared": len(sums) if done == n else done,
            "bit_exact_vs_cpu_oracle_4_smallest": bool(exact),
            "tuned_depth_per_automaton_and_size": {"gcups": round(tuned["gcups"], 1), "p50_s": round(tuned["p50_s"], 4),
                                                   "p99_s": round(tuned["p99_s"], 4)},
            "one_at_a_time": {"completed": done, "of": n, "gcups": round(b["gcups"], 1),
                              "p50_s": round(float(np.percentile(lat_all, 50)), 4),
                              "p99_s": round(float(np.percentile(lat_all, 99)), 4),
                              "latency_note": "requests not served within the time limit count with their "
                                              "waiting time so far (a lower bound)" if done < n else ""},
            "gcups_vs_one_at_a_time": round(s["gcups"] / b["gcups"], 2) if b["gcups"] else None,
            "p99_vs_one_at_a_time": round(float(np.percentile(lat_all, 99)) / s["p99_s"], 2) if s["p99_s"] else None}


def valu_roofline(g, size, ms_per_launch, layout="linear"):
    """Temporal launches are bound by VALU issue, not HBM: wave64 VALU instructions per launch
    (rocprofv3 SQ_INSTS_VALU, a fixed count for the configuration) over the live average
    launch duration, against the issue peak."""
    ins = measured_launch(g["automaton"], g["traverser"], g["temporal_steps"], size, "valu_wave_instructions", layout)
    if not ins:
        return None
    rate = ins / (ms_per_launch * 1e-3)
    return {"achieved": round(rate / 1e9, 1), "peak": round(VALU_PEAK_INSTR_PER_S / 1e9, 1),
            "unit": "G wave64 VALU instructions/s", "frac": round(rate / VALU_PEAK_INSTR_PER_S, 4),
            "instructions_per_launch": ins}


def phase_variants(a):
    """The plain-GPU configuration: every (automaton, traverser, depth) variant at a.size^2, timed
    as warm-up plus a.rounds rounds from the same initial state; then the headline variant's
    kernel alone over a.roofline_launches back-to-back launches for its average launch duration."""
    import numpy as np
    import torch
    import cellato_amd as C
    import oracle_lib as O

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    variants = []
    best = None
    depths = [int(d) for d in a.depths.split(",") if d]
    grids = {}
    layout = 1 if a.layout == "tiled" else 0
    for name in a.automata.split(","):
        rule = C.rule(name)
        dev = C.DeviceGrid(rule, a.size, a.size, layout)
        if a.size <= 16384:
            v = C.iid_values(name, a.size, a.size, 2026)
            dev.upload(C.grid.pack(rule.index_of(v), rule.planes, layout))
            del v
        else:   # larger grids: the same distribution drawn on the device
            dev.fill_random(2026)
        init = dev.buf[0].clone()
        for kind, T in ([] if a.no_simple else [("simple", 1)]) + [("temporal", d) for d in depths]:
            t = time_variant(C, torch, rule, dev, init, kind, T, a.steps, a.warmup, a.rounds, a.chunk)
            algo_bytes = 2 * rule.planes * a.size * a.size / 8   # P bits read + P bits written per cell and launch
            rec = {"automaton": name, "traverser": kind, "temporal_steps": T, "gcups": round(t["gcups"], 1),
                   "rounds": a.rounds, "ms_mean": round(t["ms_mean"], 4), "ms_std": round(t["ms_std"], 4),
                   "ms_rounds": t["ms_rounds"], "ms_per_launch": round(t["ms_per_launch"], 5),
                   "launches_per_round": t["launches"], "checksum_consistent": t["checksum_consistent"],
                   "algorithmic_GBps": round(algo_bytes / (t["ms_per_launch"] * 1e-3) / 1e9, 1)}
            # average launch duration over back-to-back launches (what a kernel trace reports)
            ms_b2b = time_launches(C, torch, rule, dev, kind, T, a.roofline_launches, a.chunk)
            rec["us_per_launch_b2b"] = round(ms_b2b * 1e3, 3)
            rec["algorithmic_GBps_b2b"] = round(algo_bytes / (ms_b2b * 1e-3) / 1e9, 1)
            if layout:
                rec["layout"] = "tiled"
            if not a.no_check:
                rec["bit_exact_2048_subset"] = check_subset(C, O, name, kind, T, layout)
            print(json.dumps(rec), file=sys.stderr, flush=True)
            variants.append(rec)
            if name == "game-of-life" and (best is None or rec["gcups"] > best["gcups"]):
                best = rec
        grids[name] = (rule, dev)
        del init
        torch.cuda.empty_cache()

    g = best or variants[0]
    rule, dev = grids[g["automaton"]]
    bytes_per_launch = 2 * rule.planes * a.size * a.size // 8
    ms_launch, achieved = g["us_per_launch_b2b"] * 1e-3, g["algorithmic_GBps_b2b"]
    valu = valu_roofline(g, a.size, ms_launch, a.layout)
    simple_gol = next((r for r in variants if r["automaton"] == g["automaton"] and r["traverser"] == "simple"), None)
    simple_roof = None
    if simple_gol is not None:
        gbps_s = simple_gol["algorithmic_GBps_b2b"]
        simple_roof = {"bound": "hbm", "achieved": gbps_s, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                       "frac": round(gbps_s / HBM_PEAK_GBPS, 4), "kernel": "stream_kernel (simple)",
                       "us_per_launch": simple_gol["us_per_launch_b2b"], "launches_timed": a.roofline_launches,
                       "bytes_per_launch": bytes_per_launch,
                       "traffic": measured_launch(g["automaton"], "simple", 1, a.size, layout=a.layout)}
    for name, (r, d) in grids.items():
        del d
    grids.clear()
    torch.cuda.empty_cache()

    copy_rec = None
    if a.copy_mib:
        # calibration kernel of the FETCH_SIZE / WRITE_SIZE counters: a plain elementwise pass
        # that reads a.copy_mib MiB and writes a.copy_mib MiB (torch's bitwise_not, a kernel
        # nothing else in the run launches, so that its counters are not mixed with others)
        n = a.copy_mib * 2**20 // 8
        src = torch.zeros(n, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.bitwise_not(src, out=dst)
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(20):
            torch.bitwise_not(src, out=dst)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / 20
        copy_rec = {"copy_MiB": a.copy_mib, "kernel": "torch bitwise_not (read + write)",
                    "GBps": round(2 * a.copy_mib * 2**20 / ms / 1e6, 1)}
        print(json.dumps(copy_rec), file=sys.stderr, flush=True)
        del src, dst

    out = {
        "metric": "GCUPS (cell updates/s) at 16384^2 bit-exact; batched/served vs plain GPU; p99",
        "value": g["gcups"],
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "rounds": a.rounds,
        "ms_per_step": round(g["ms_mean"] / a.steps, 5),
        "ms_per_step_std": round(g["ms_std"] / a.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"{g['automaton']} {a.size}x{a.size}, i.i.d. seeded, {a.steps} generations per round, "
                               f"{a.rounds} rounds from the same initial state after {a.warmup} warm-up generations, "
                               f"plain GPU path (one grid), {a.layout} bit planes, 64-bit words",
                   "traverser": g["traverser"], "temporal_steps": g["temporal_steps"]},
        "bit_exact": all(r.get("bit_exact_2048_subset", True) for r in variants),
        "checksum_consistent": all(r["checksum_consistent"] for r in variants),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": measured_launch(g["automaton"], g["traverser"], g["temporal_steps"], a.size,
                                                layout=a.layout),
                     "kernel": f"stream_kernel ({g['traverser']}, T={g['temporal_steps']})",
                     "bytes_per_launch": bytes_per_launch,
                     "us_per_launch": round(ms_launch * 1e3, 3), "launches_timed": a.roofline_launches,
                     # at T > 1 generations per launch the kernel is limited by VALU issue, not HBM
                     "limiter": "valu" if g["temporal_steps"] > 1 and g["traverser"] == "temporal" else "hbm",
                     "valu": valu,
                     # the bytes a one-generation-per-launch traverser would move for the same cell
                     # updates (P/4 per update), over the same time: what temporal blocking gains
                     "single_step_equivalent_GBps": round(g["gcups"] * rule.planes / 4, 1),
                     "single_step_equivalent_frac": round(g["gcups"] * rule.planes / 4 / HBM_PEAK_GBPS, 4)},
        "roofline_simple": simple_roof,
        "variants": variants,
    }
    if copy_rec:
        out["copy"] = copy_rec
    return out


# the reference's best published H100 figures (GCUPS) for the multi-plane rules at 16384^2,
# any traverser (results/H100/_h100.csv: cyclic :2717, excitable :1362, fluid :2290,
# critters :1664)
H100_BEST = {"cyclic": 1505, "excitable": 3146, "fluid": 9197, "critters": 4680}


def phase_multiplane(a):
    """Multi-plane rules on the plain GPU path: Cyclic (5 planes), Greenberg-Hastings (3),
    HPP fluid (4), Critters (1, Margolus), 16384^2, seeded i.i.d., temporal traverser at
    depths 2/4/6/8, each timed as warm-up plus a.rounds rounds from the same initial state; per
    automaton the best depth, re-run bit-exact on a 2048^2 subset against the standard
    evaluator. A depth whose kernel would spill registers runs at the deepest one that does
    not (`temporal_steps_used`)."""
    import torch
    import cellato_amd as C
    import oracle_lib as O
    from cellato_amd import _lib
    res = {}
    for name in a.multiplane.split(","):
        rule = C.rule(name)
        dev = C.DeviceGrid(rule, a.size, a.size, 0)
        dev.upload(C.grid.pack(rule.index_of(C.iid_values(name, a.size, a.size, 2026)), rule.planes, 0))
        init = dev.buf[0].clone()
        runs, used, std, consistent = {}, {}, {}, True
        for T in (2, 4, 6, 8):
            t = time_variant(C, torch, rule, dev, init, "temporal", T, a.multiplane_steps, 10, a.rounds, 0)
            runs[T] = round(t["gcups"], 1)
            std[T] = round(t["ms_std"] / t["ms_mean"], 4)
            used[T] = _lib.kernel_info(rule.id, 0, a.size, T)[0]
            consistent &= t["checksum_consistent"]
        bestT = max(runs, key=runs.get)
        rec = {"planes": rule.planes, "gcups": runs[bestT], "temporal_steps": bestT,
               "gcups_by_depth": {f"T{T}": g for T, g in runs.items()},
               "temporal_steps_used": {f"T{T}": u for T, u in used.items()},
               "rel_std_by_depth": {f"T{T}": v for T, v in std.items()}, "checksum_consistent": consistent,
               "bit_exact_2048_subset": check_subset(C, O, name, "temporal", bestT),
               "reference_h100_best_gcups": H100_BEST.get(name)}
        print(json.dumps({name: rec}), file=sys.stderr, flush=True)
        res[name] = rec
        del dev, init
        torch.cuda.empty_cache()
    return {"multi_plane": {"workload": f"16384^2, seeded i.i.d., {a.multiplane_steps} generations per round, "
                                         f"{a.rounds} rounds per depth, plain GPU path, linear bit planes",
                            "automata": res}}


# the reference's best published H100 GCUPS of its bit_array evaluator at 16384^2 x 128
# (results/H100/_h100.csv, best configuration per word size)
H100_BIT_ARRAY = {("game-of-life", 32): 285.8, ("game-of-life", 64): 225.6, ("wire", 32): 307.0, ("wire", 64): 285.9}


def phase_bit_array(a):
    """The bit_array layout (cells packed as P-bit indices, one generation per launch), a
    comparison point: Game of Life and WireWorld at a.size^2 with 32- and 64-bit words, seeded
    i.i.d., a.rounds rounds of a.bit_array_steps generations from the same initial state, words
    equal across rounds; bit-exact on a 2048^2 subset against the standard evaluator."""
    import numpy as np
    import torch
    import cellato_amd as C
    import oracle_lib as O
    from cellato_amd import _lib
    res = {}
    stream = torch.cuda.current_stream()
    for name in ("game-of-life", "wire"):
        rule = C.rule(name)
        idx = rule.index_of(C.iid_values(name, a.size, a.size, 2026))
        sub = C.iid_values(name, 2048, 2048, 11)
        want = O.run_standard(name, sub, 64, threads=min(16, os.cpu_count() or 1))
        for wb in (32, 64):
            rows = 1024   # pack in bands of rows to bound the host memory
            words = np.concatenate([C.grid.pack_bit_array(idx[r:r + rows], rule.planes, wb) for r in range(0, a.size, rows)])
            init = torch.from_numpy(words.view(np.int64 if wb == 64 else np.int32)).to("cuda")
            buf = [init.clone(), torch.zeros_like(init)]

            def run(steps, t0=0):
                return _lib.bit_array(rule.id, wb, buf[0].data_ptr(), buf[1].data_ptr(), a.size, a.size, t0, steps,
                                      stream.cuda_stream)
            run(4)
            torch.cuda.synchronize()
            times, finals = [], []
            for _ in range(a.rounds):
                buf[0].copy_(init)
                torch.cuda.synchronize()
                s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s0.record(stream)
                in_b = run(a.bit_array_steps)
                s1.record(stream)
                s1.synchronize()
                times.append(s0.elapsed_time(s1))
                finals.append(buf[1 if in_b else 0].clone())
            same = all(torch.equal(f, finals[0]) for f in finals)
            del finals
            ms = float(np.mean(times))
            # bit-exact on the 2048^2 subset
            sw = torch.from_numpy(C.grid.pack_bit_array(rule.index_of(sub), rule.planes, wb).view(
                np.int64 if wb == 64 else np.int32)).to("cuda")
            sb = torch.zeros_like(sw)
            in_b = _lib.bit_array(rule.id, wb, sw.data_ptr(), sb.data_ptr(), 2048, 2048, 0, 64, stream.cuda_stream)
            out = (sb if in_b else sw).cpu().numpy().view(np.uint64 if wb == 64 else np.uint32)
            got = rule.value_of(C.grid.unpack_bit_array(out, 2048, 2048, rule.planes, wb))
            rec = {"gcups": round(a.size * a.size * a.bit_array_steps / (ms * 1e-3) / 1e9, 1), "ms_mean": round(ms, 3),
                   "ms_std": round(float(np.std(times)), 4), "rounds_had_same_result": bool(same),
                   "bit_exact_2048_subset": bool(np.array_equal(got, want)),
                   "reference_h100_best_gcups": H100_BIT_ARRAY.get((name, wb))}
            res[f"{name} {wb}-bit"] = rec
            print(json.dumps({name: {wb: rec}}), file=sys.stderr, flush=True)
            del init, buf, sw, sb
            torch.cuda.empty_cache()
    return {"bit_array": {"workload": f"{a.size}^2, seeded i.i.d., {a.bit_array_steps} generations per round, "
                                      f"{a.rounds} rounds, simple traverser, bit_array layout", "runs": res}}


def phase_batch(a):
    import torch
    import cellato_amd as C
    import oracle_lib as O
    from cellato_amd.tuning_eval import evaluate
    tuner, train_s = make_tuner(a.tuner_cache)
    gains = {h: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in evaluate(budget=2, holdout=h).items()}
             for h in ("size", "automaton")}
    out = {"tuner": {"model": f"MLP {'-'.join(str(d) for d in tuner.model.dims)} trained on the GPU (csrc/mlp.hip) on "
                              "cellato_amd/data/depth_tuning_mi355x.json", "train_seconds": round(train_s, 2),
                     "measured_per_bucket": tuner.budget, "held_out_fraction_of_best": gains}}
    out["batched"] = bench_batch(C, O, torch, tuner, a.batch_grids)
    print(json.dumps({"batched": out["batched"]}), file=sys.stderr, flush=True)
    return out


def phase_serve(a):
    import torch
    import cellato_amd as C
    import oracle_lib as O
    tuner, _ = make_tuner(a.tuner_cache)
    out = {"served": bench_serve(C, O, torch, tuner, a.serve_requests, a.serve_rate, a.serve_baseline_limit,
                                 a.tick_passes)}
    print(json.dumps(out), file=sys.stderr, flush=True)
    return out


def phase_cpu(a):
    import cellato_amd as C
    import oracle_lib as O
    return {"cpu_baseline": cpu_baseline(C, O)}


def phase_table(a):
    """Every automaton, both layouts, simple and temporal depths 2/4/6/8 at a.size^2 (seeded
    i.i.d. states made on the device, 200 generations each): the best GCUPS per automaton and
    layout. Not part of the default run (`--phase table`)."""
    import torch
    import cellato_amd as C
    from cellato_amd import _lib
    out = {}
    stream = torch.cuda.current_stream().cuda_stream
    for r in C.all_rules():
        for layout in (0, 1):
            dev = C.DeviceGrid(r, a.size, a.size, layout)
            dev.fill_random(7)
            best = {}
            for kind, T in (("simple", 1), ("temporal", 2), ("temporal", 4), ("temporal", 6), ("temporal", 8)):
                def adv(n, t0):
                    if kind == "simple":
                        in_b = _lib.simple(r.id, layout, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), a.size, a.size, t0, n, 0, stream)
                    else:
                        in_b = _lib.temporal(r.id, layout, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), a.size, a.size, t0, n, T, 0, stream)
                    if in_b:
                        dev.cur = 1 - dev.cur
                adv(8, 0)
                torch.cuda.synchronize()
                s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s0.record()
                adv(200, 8)
                s1.record()
                torch.cuda.synchronize()
                best[f"{kind}{T if kind == 'temporal' else ''}"] = round(a.size * a.size * 200 / (s0.elapsed_time(s1) * 1e-3) / 1e9, 1)
            out.setdefault(r.name, {})["linear" if layout == 0 else "tiled"] = best
            print(json.dumps({r.name: out[r.name]}), file=sys.stderr, flush=True)
            del dev
            torch.cuda.empty_cache()
    return {"table": out}


PHASES = {"variants": phase_variants, "multiplane": phase_multiplane, "bit_array": phase_bit_array,
          "batch": phase_batch, "serve": phase_serve, "cpu": phase_cpu, "table": phase_table}


def run_phase(name, argv):
    """Run one phase in a child process (its own GPU context; a crash there cannot take the
    whole bench line with it). Returns (result dict or None, error dict or None)."""
    import collections
    import subprocess
    cmd = [sys.executable, "-X", "faulthandler", "-u", str(Path(__file__).resolve()), "--phase", name, *argv]
    p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    tail = collections.deque(maxlen=40)
    for line in p.stderr:           # progress lines pass straight through
        sys.stderr.write(line)
        sys.stderr.flush()
        tail.append(line.rstrip())
    out = p.stdout.read()
    rc = p.wait()
    lines = [l for l in out.strip().splitlines() if l.startswith("{")]
    if rc == 0 and lines:
        return json.loads(lines[-1]), None
    return None, {"phase": name, "exit_status": rc, "stderr_tail": list(tail)[-15:]}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (each box has one MI355X)",
              file=sys.stderr)
        sys.exit(2)
    if a.phase:
        print(json.dumps(PHASES[a.phase](a)), flush=True)
        return
    import tempfile
    argv = [x for x in sys.argv[1:]]
    with tempfile.TemporaryDirectory() as tmp:
        argv += ["--tuner-cache", str(Path(tmp) / "tuner_cache.json")]
        out, err = run_phase("variants", argv)
        errors = [err] if err else []
        if out is None:   # no headline: still one JSON line, saying what failed
            out = {"metric": "GCUPS (cell updates/s) at 16384^2 bit-exact; batched/served vs plain GPU; p99",
                   "value": None, "unit": "GCUPS", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                   "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                   "dtype": "u64", "data": "synthetic", "config": {"workload": "game-of-life 16384x16384"}}
        for name, skip in (("multiplane", a.no_multiplane), ("bit_array", a.no_bit_array), ("batch", a.no_batch),
                           ("serve", a.no_serve), ("cpu", a.no_cpu)):
            if skip:
                continue
            res, err = run_phase(name, argv)
            if res:
                out.update(res)
            if err:
                errors.append(err)
        if errors:
            out["phase_errors"] = errors
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
