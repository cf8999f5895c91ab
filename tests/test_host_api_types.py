"""The C++ host API's types, checked with the host compiler (no GPU): the reference's five GPU
bit-level test-suite type expressions (include/experiments/test_suites.hpp:86-225) compile against
include/cellato with only the traversers' namespace changed, and misuse fails at compile time
with a message (an evaluator of one layout on a grid of the other, a halo radius the kernels do
not carry)."""
import shutil
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent

PRELUDE = """
#include "cellato/automata/all.hpp"
#include "cellato/evaluators/bit_array.hpp"
#include "cellato/evaluators/bit_planes.hpp"
#include "cellato/evaluators/tiled_bit_planes.hpp"
#include "cellato/memory/bit_planes_grid.hpp"
#include "cellato/traversers/hip/traversers.hpp"
namespace grids = cellato::memory::grids;
namespace evaluators = cellato::evaluators;
namespace traversers = cellato::traversers::hip;
using word = unsigned long;
using ca = wire::config;
using algorithm_t = ca::algorithm;
using state_dictionary_t = ca::state_dictionary;
using lin_grid = grids::bit_planes::grid<word, state_dictionary_t>;
using til_grid = grids::tiled_bit_planes::grid<word, state_dictionary_t>;
using lin_eval = evaluators::bit_planes::evaluator<word, state_dictionary_t, algorithm_t>;
using til_eval = evaluators::tiled_bit_planes::evaluator<word, state_dictionary_t, algorithm_t>;
using ba_grid = grids::bit_array::grid<state_dictionary_t, unsigned int>;
using ba_eval = evaluators::bit_array::evaluator<ba_grid, algorithm_t>;
"""


def compile_snippet(body, tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "snippet.cpp"
    src.write_text(PRELUDE + body)
    r = subprocess.run([cxx, "-std=c++20", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        f"-I{REPO / 'include'}", str(src)], capture_output=True, text=True)
    return r.returncode, r.stderr


def test_reference_suite_shapes_compile(tmp_path):
    body = """
template <class T> void use() { T t; (void)t; static_assert(T::is_CUDA); }
void f() {
    use<traversers::simple::traverser<lin_eval, lin_grid>>();
    use<traversers::simple::traverser<til_eval, til_grid>>();
    use<traversers::temporal::traverser<til_eval, til_grid, ca::average_halo_radius>>();
    use<traversers::temporal::traverser<lin_eval, lin_grid, ca::average_halo_radius>>();
    use<traversers::simple::traverser<ba_eval, ba_grid>>();
    static_assert(ba_grid::cells_per_word == 16 && ba_grid::bits_per_cell == 2 && ba_grid::cell_mask == 3u);
    static_assert(grids::bit_array::grid<state_dictionary_t>::cells_per_word == 16);   // 32-bit words by default
    // the grid members the suites and the reference traversers read
    static_assert(lin_grid::x_word_tile_size == 64 && lin_grid::y_word_tile_size == 1);
    static_assert(til_grid::x_word_tile_size == 8 && til_grid::y_word_tile_size == 8);
    static_assert(lin_grid::word_store_bits == 64 && lin_grid::needed_bits == 2);
    using cuda_grid_t = std::invoke_result<decltype(&lin_grid::to_cuda), lin_grid>::type;
    cuda_grid_t g; g.free_cuda_memory();
    // implicit conversion from a standard grid (the experiment manager returns one as grid_t)
    auto conv = [](const grids::standard::grid<wire::wire_cell_state>& s) -> lin_grid { return s; };
    (void)conv;
}
"""
    rc, err = compile_snippet(body, tmp_path)
    assert rc == 0, err[-3000:]


def test_layout_mismatch_is_a_compile_error(tmp_path):
    rc, err = compile_snippet("traversers::simple::traverser<til_eval, lin_grid> t;\n", tmp_path)
    assert rc != 0 and "evaluator and grid layouts differ" in err
    rc, err = compile_snippet("traversers::temporal::traverser<lin_eval, til_grid, 1.0> t;\n", tmp_path)
    assert rc != 0 and "evaluator and grid layouts differ" in err


def test_halo_radius_beyond_one_cell_is_a_compile_error(tmp_path):
    rc, err = compile_snippet("traversers::temporal::traverser<lin_eval, lin_grid, 2.0> t;\n", tmp_path)
    assert rc != 0 and "halo of one cell per generation" in err


def test_bit_array_grid_mismatch_is_a_compile_error(tmp_path):
    rc, err = compile_snippet("traversers::simple::traverser<lin_eval, ba_grid> t;\n", tmp_path)
    assert rc != 0
