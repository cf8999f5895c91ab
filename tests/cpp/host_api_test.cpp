// GPU test of the C++ host API (include/cellato/traversers/hip/traversers.hpp).
//
// The five GPU bit-level test-suite shapes of the reference (include/experiments/test_suites.hpp:
// 86-225: bit_array, bit_planes, tiled_bit_planes, temporal_tiled_bit_planes,
// temporal_linear_bit_planes) are bound below with the same type expressions, the traversers' namespace alias pointing at
// cellato::traversers::hip instead of cellato::traversers::cuda. Each is driven the way the
// reference's experiment manager drives a suite (experiment_manager.hpp:61-97): standard grid
// -> grid_t by implicit conversion -> init -> run -> fetch_result -> to_standard -> checksum,
// and compared with the CPU oracle's standard evaluator on the same seeded grid.
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "cellato/automata/all.hpp"
#include "cellato/evaluators/bit_array.hpp"
#include "cellato/evaluators/bit_planes.hpp"
#include "cellato/evaluators/tiled_bit_planes.hpp"
#include "cellato/memory/bit_array_grid.hpp"
#include "cellato/memory/bit_planes_grid.hpp"
#include "cellato/traversers/hip/traversers.hpp"
#include "standard.hpp"   // oracle (the checker)

namespace suites {
namespace grids = cellato::memory::grids;
namespace evaluators = cellato::evaluators;
namespace traversers = cellato::traversers::hip;   // the reference: cellato::traversers::cuda

template <typename store_word_type>
struct using_ {
    template <typename cellular_automaton>
    struct bit_array {
        using original_cell_t = typename cellular_automaton::cell_state;
        using grid_store_word_t = store_word_type;
        using algorithm_t = typename cellular_automaton::algorithm;
        using state_dictionary_t = typename cellular_automaton::state_dictionary;
        using grid_t = grids::bit_array::grid<state_dictionary_t, grid_store_word_t>;
        using evaluator_t = evaluators::bit_array::evaluator<grid_t, algorithm_t>;
        using traverser_t = traversers::simple::traverser<evaluator_t, grid_t>;
        static constexpr const char* what = sizeof(store_word_type) == 8 ? "bit_array simple 64" : "bit_array simple 32";
    };

    template <typename cellular_automaton>
    struct bit_planes {
        using original_cell_t = typename cellular_automaton::cell_state;
        using algorithm_t = typename cellular_automaton::algorithm;
        using state_dictionary_t = typename cellular_automaton::state_dictionary;
        using grid_t = grids::bit_planes::grid<store_word_type, state_dictionary_t>;
        using evaluator_t = evaluators::bit_planes::evaluator<store_word_type, state_dictionary_t, algorithm_t>;
        using traverser_t = traversers::simple::traverser<evaluator_t, grid_t>;
        static constexpr const char* what = "bit_planes simple";
    };

    template <typename cellular_automaton>
    struct tiled_bit_planes {
        using original_cell_t = typename cellular_automaton::cell_state;
        using algorithm_t = typename cellular_automaton::algorithm;
        using state_dictionary_t = typename cellular_automaton::state_dictionary;
        using grid_t = grids::tiled_bit_planes::grid<store_word_type, state_dictionary_t>;
        using evaluator_t = evaluators::tiled_bit_planes::evaluator<store_word_type, state_dictionary_t, algorithm_t>;
        using traverser_t = traversers::simple::traverser<evaluator_t, grid_t>;
        static constexpr const char* what = "tiled_bit_planes simple";
    };

    template <typename cellular_automaton>
    struct temporal_tiled_bit_planes {
        using original_cell_t = typename cellular_automaton::cell_state;
        using algorithm_t = typename cellular_automaton::algorithm;
        using state_dictionary_t = typename cellular_automaton::state_dictionary;
        using grid_t = grids::tiled_bit_planes::grid<store_word_type, state_dictionary_t>;
        using evaluator_t = evaluators::tiled_bit_planes::evaluator<store_word_type, state_dictionary_t, algorithm_t>;
        using traverser_t = traversers::temporal::traverser<evaluator_t, grid_t, cellular_automaton::average_halo_radius>;
        static constexpr const char* what = "tiled_bit_planes temporal";
    };

    template <typename cellular_automaton>
    struct temporal_linear_bit_planes {
        using original_cell_t = typename cellular_automaton::cell_state;
        using algorithm_t = typename cellular_automaton::algorithm;
        using state_dictionary_t = typename cellular_automaton::state_dictionary;
        using grid_t = grids::bit_planes::grid<store_word_type, state_dictionary_t>;
        using evaluator_t = evaluators::bit_planes::evaluator<store_word_type, state_dictionary_t, algorithm_t>;
        using traverser_t = traversers::temporal::traverser<evaluator_t, grid_t, cellular_automaton::average_halo_radius>;
        static constexpr const char* what = "bit_planes temporal";
    };
};
} // namespace suites

using namespace cellato;
using word = std::uint64_t;

// One round of a suite: initial state -> grid_t (implicit conversion) -> traverser -> checksum,
// run in two calls of run() to cover a start step other than 0.
template <template <class> class Suite, class Cfg>
bool check(int W, int H, int steps, int T) {
    using S = Suite<Cfg>;
    using state_t = typename S::original_cell_t;
    using dict = typename S::state_dictionary_t;
    using grid_t = typename S::grid_t;
    std::mt19937 rng(123);
    std::vector<state_t> cells(static_cast<size_t>(W) * H);
    for (auto& c : cells) c = dict::index_to_state(static_cast<int>(rng() % dict::number_of_values));
    memory::grids::standard::grid<state_t> start(std::vector<state_t>(cells), W, H);

    grid_t grid = start;
    typename S::traverser_t t;
    run::run_params p;
    p.temporal_steps = T;
    p.word_size = static_cast<int>(sizeof(typename grid_t::store_type) * 8);   // as --word_size sets it
    t.init(grid, p);
    t.run(steps / 2);
    t.run(steps - steps / 2);
    grid_t result = t.fetch_result();
    const std::string got = result.to_standard().get_checksum();

    std::vector<int32_t> ref(cells.size());
    for (size_t i = 0; i < cells.size(); ++i) ref[i] = static_cast<int32_t>(cells[i]);
    oracle::standard::run<Cfg>(ref.data(), W, H, 0, steps, 4);
    std::vector<state_t> refv(ref.size());
    for (size_t i = 0; i < ref.size(); ++i) refv[i] = static_cast<state_t>(ref[i]);
    const std::string want = memory::grids::standard::grid<state_t>(std::move(refv), W, H).get_checksum();
    std::printf("%-12s %-26s T=%d %dx%d %s\n", Cfg::name, S::what, T, W, H, got == want ? "ok" : "MISMATCH");
    return got == want;
}

template <class C> using ba64 = suites::using_<word>::bit_array<C>;
template <class C> using ba32 = suites::using_<std::uint32_t>::bit_array<C>;
template <class C> using lin = suites::using_<word>::bit_planes<C>;
template <class C> using til = suites::using_<word>::tiled_bit_planes<C>;
template <class C> using ttil = suites::using_<word>::temporal_tiled_bit_planes<C>;
template <class C> using tlin = suites::using_<word>::temporal_linear_bit_planes<C>;

int main() {
    bool ok = true;
    ok &= check<ba64, game_of_life::config>(1024, 48, 13, 0);
    ok &= check<ba32, wire::config>(512, 40, 11, 0);
    ok &= check<ba64, critters::config>(640, 36, 13, 0);
    ok &= check<ba32, cyclic::config>(600, 30, 9, 0);      // 6 cells per 32-bit word
    ok &= check<ba64, excitable::config>(630, 33, 9, 0);   // 21 cells per 64-bit word
    ok &= check<lin, game_of_life::config>(8192, 96, 13, 0);
    ok &= check<lin, excitable::config>(4160, 40, 9, 0);
    ok &= check<til, wire::config>(520, 40, 13, 0);
    ok &= check<til, traffic::config>(1032, 48, 11, 0);
    ok &= check<ttil, cyclic::config>(1024, 64, 13, 3);
    ok &= check<ttil, critters::config>(2048, 72, 13, 4);
    ok &= check<tlin, critters::config>(4160, 80, 13, 4);
    ok &= check<tlin, fire::config>(192, 77, 13, 8);
    ok &= check<tlin, cyclic::config>(4096, 64, 13, 8);   // runs as passes of 4 (deeper would spill)
    ok &= check<tlin, fluid::config>(640, 50, 13, 5);
    // reference error behaviour: invalid temporal depth throws std::runtime_error
    try {
        tlin<maze::config>::traverser_t t;
        run::run_params p;
        p.temporal_steps = 99;
        std::vector<maze::maze_cell_state> c(64 * 4, maze::maze_cell_state::empty);
        t.init(tlin<maze::config>::grid_t(4, 64, c.data()), p);
        ok = false;
        std::printf("invalid temporal depth accepted\n");
    } catch (const std::runtime_error&) {
        std::printf("invalid temporal depth rejected: ok\n");
    }
    std::printf(ok ? "ALL OK\n" : "FAILED\n");
    return ok ? 0 : 1;
}
