"""The reference-compatible command line (2026-cellato-smpt-journal_amd/cellato): the
reference's CSV columns and messages, its seeded initial states (equal to the oracle's
restatement, itself pinned by the published checksums), and on the GPU the published
16384^2 checksums reproduced from the command line."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

REPO = Path(__file__).resolve().parent.parent
CLI = REPO / "2026-cellato-smpt-journal_amd" / "cellato"
DUMP = REPO / "tests" / "cpp" / "init_dump"

# the reference's header: run_params.hpp:73-75 + experiment_report.hpp:82-84
HEADER = ("automaton,device,traverser,evaluator,layout,reference_impl,x_size,y_size,steps,rounds,warmup_rounds,"
          "x_tile_size,y_tile_size,cuda_block_size_x,cuda_block_size_y,seed,word_size,temporal_steps,"
          "temporal_tile_size_y,average_time_ms,average_time_per_cell_ps,std_time_ms,rounds_had_same_checksums,checksum")


def cli(*args, timeout=300):
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_csv_header_is_the_references():
    r = cli("--print_csv_header")
    assert r.returncode == 0 and r.stdout.strip() == HEADER


def test_help_and_missing_options():
    r = cli("--help")
    assert r.returncode == 0 and "Usage" in r.stdout
    r = cli("--automaton", "wire", "--device", "HIP", "--traverser", "temporal", "--evaluator", "bit_planes",
            "--layout", "bit_planes", "--x_size", "64", "--y_size", "64", "--steps", "4", "--word_size", "64")
    assert r.returncode == 1 and "Missing required option: temporal_steps" in r.stderr
    r = cli("--automaton", "wire", "--device", "HIP", "--traverser", "simple", "--evaluator", "bit_planes",
            "--layout", "bit_planes", "--x_size", "64", "--y_size", "64", "--steps", "4")
    assert r.returncode == 1 and "Missing required option: word_size" in r.stderr


@pytest.mark.parametrize("args", [
    ["--device", "CPU", "--traverser", "simple", "--evaluator", "standard", "--layout", "standard"],
    ["--device", "HIP", "--traverser", "simple", "--evaluator", "bit_planes", "--layout", "bit_planes", "--word_size", "32"],
    ["--device", "HIP", "--traverser", "spacial_blocking", "--evaluator", "standard", "--layout", "standard"],
])
def test_unserved_combinations_say_so(args):
    r = cli("--automaton", "game-of-life", "--x_size", "64", "--y_size", "64", "--steps", "2", *args)
    assert r.returncode == 0 and r.stdout == ""
    assert "No suitable test suite found for the given parameters." in r.stderr


def test_reference_impl_is_not_served():
    r = cli("--automaton", "game-of-life", "--reference_impl", "baseline", "--x_size", "64", "--y_size", "64",
            "--steps", "2")
    assert r.returncode == 0 and "No suitable reference implementation found" in r.stderr


@pytest.mark.parametrize("rule", ["game-of-life", "wire", "forest-fire", "excitable", "brian", "maze", "traffic",
                                  "fluid", "critters", "cyclic"])
def test_initial_states_equal_the_oracle(rule):
    for w, h, seed in [(64, 48, 42), (333, 211, 7), (1024, 512, 42)]:
        r = subprocess.run([str(DUMP), rule, str(w), str(h), str(seed)], capture_output=True, timeout=60)
        assert r.returncode == 0
        got = np.frombuffer(r.stdout, np.int32).reshape(h, w)
        assert np.array_equal(got, O.init_reference(rule, w, h, seed)), (rule, w, h, seed)


GOLD = json.loads((REPO / "tests" / "golden" / "reference_checksums.json").read_text())


def parse(line):
    return dict(zip(HEADER.split(","), line.strip().split(",")))


@pytest.mark.gpu
@pytest.mark.parametrize("rule,evaluator,traverser,T,word,size", [
    ("game-of-life", "bit_planes", "temporal", 4, 64, (16384, 16384)),
    ("wire", "tiled_bit_planes", "simple", 0, 64, (16384, 16384)),
    ("critters", "bit_planes", "temporal", 3, 64, (16384, 16384)),
    ("brian", "bit_array", "simple", 0, 32, (16384, 16384)),
    ("excitable", "bit_array", "simple", 0, 64, (16128, 16644)),   # 21 cells per 64-bit word
])
def test_published_checksum_from_the_command_line(rule, evaluator, traverser, T, word, size):
    W, H = size
    g = next(e for e in GOLD if e["automaton"] == rule and (e["x_size"], e["y_size"], e["steps"]) == (W, H, 128))
    args = ["--automaton", rule, "--device", "CUDA", "--traverser", traverser, "--evaluator", evaluator,
            "--layout", evaluator, "--x_size", W, "--y_size", H, "--steps", 128, "--word_size", word,
            "--seed", g["seed"], "--rounds", 2, "--warmup_rounds", 1]
    if traverser == "temporal":
        args += ["--temporal_steps", T, "--temporal_tile_size_y", 32]
    r = cli(*args)
    assert r.returncode == 0, r.stderr[-2000:]
    row = parse(r.stdout.strip().splitlines()[-1])
    assert row["checksum"] == "-".join(str(v) for v in g["checksum"]), g["source"]
    assert row["rounds_had_same_checksums"] == "true"
    assert row["automaton"] == rule and row["device"] == "CUDA" and int(row["rounds"]) == 2
    assert float(row["average_time_ms"]) > 0
