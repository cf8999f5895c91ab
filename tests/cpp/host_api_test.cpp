// GPU test of the C++ host API (include/cellato/traversers/hip/traversers.hpp), written the
// way the reference's experiment manager drives a traverser (experiment_manager.hpp:61-90):
// standard grid -> bit-plane grid -> init -> run -> fetch_result -> to_standard -> checksum,
// compared with the CPU oracle's standard evaluator on the same seeded grid.
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "cellato/automata/all.hpp"
#include "cellato/evaluators/bit_planes.hpp"
#include "cellato/traversers/hip/traversers.hpp"
#include "standard.hpp"   // oracle (the checker)

using namespace cellato;
using word = std::uint64_t;

template <class Cfg, template <class, class, memory::grids::device> class Grid, template <class, class> class Trav>
bool check(int W, int H, int steps, int T, const char* what) {
    using state_t = typename Cfg::cell_state;
    using dict = typename Cfg::state_dictionary;
    std::mt19937 rng(123);
    std::vector<state_t> cells(static_cast<size_t>(W) * H);
    for (auto& c : cells) c = dict::index_to_state(static_cast<int>(rng() % dict::number_of_values));
    memory::grids::standard::grid<state_t> start(std::vector<state_t>(cells), W, H);

    using grid_t = Grid<word, dict, memory::grids::device::CPU>;
    using eval_t = evaluators::bit_planes::evaluator<word, dict, typename Cfg::algorithm>;
    Trav<eval_t, grid_t> t;
    run::run_params p;
    p.temporal_steps = T;
    t.init(grid_t(start), p);
    t.run(steps / 2);
    t.run(steps - steps / 2);
    const std::string got = t.fetch_result().to_standard().get_checksum();

    std::vector<int32_t> ref(cells.size());
    for (size_t i = 0; i < cells.size(); ++i) ref[i] = static_cast<int32_t>(cells[i]);
    oracle::standard::run<Cfg>(ref.data(), W, H, 0, steps, 4);
    std::vector<state_t> refv(ref.size());
    for (size_t i = 0; i < ref.size(); ++i) refv[i] = static_cast<state_t>(ref[i]);
    const std::string want = memory::grids::standard::grid<state_t>(std::move(refv), W, H).get_checksum();
    std::printf("%-12s %-28s %s\n", Cfg::name, what, got == want ? "ok" : "MISMATCH");
    return got == want;
}

template <class E, class G> using simple_t = traversers::hip::simple::traverser<E, G>;
template <class E, class G> using temporal_t = traversers::hip::temporal::traverser<E, G>;

int main() {
    bool ok = true;
    ok &= check<game_of_life::config, memory::grids::bit_planes::grid, simple_t>(8192, 96, 13, 0, "linear simple");
    ok &= check<critters::config, memory::grids::bit_planes::grid, temporal_t>(4160, 80, 13, 4, "linear temporal T=4");
    ok &= check<cyclic::config, memory::grids::tiled_bit_planes::grid, temporal_t>(1024, 64, 13, 3, "tiled temporal T=3");
    ok &= check<wire::config, memory::grids::tiled_bit_planes::grid, simple_t>(520, 40, 13, 0, "tiled simple");
    ok &= check<fire::config, memory::grids::bit_planes::grid, temporal_t>(192, 77, 13, 8, "linear temporal T=8");
    // reference error behaviour: invalid temporal depth throws std::runtime_error
    try {
        temporal_t<evaluators::bit_planes::evaluator<word, maze::config::state_dictionary, maze::config::algorithm>,
                   memory::grids::bit_planes::grid<word, maze::config::state_dictionary>> t;
        run::run_params p;
        p.temporal_steps = 99;
        std::vector<maze::maze_cell_state> c(64 * 4, maze::maze_cell_state::empty);
        t.init(memory::grids::bit_planes::grid<word, maze::config::state_dictionary>(4, 64, c.data()), p);
        ok = false;
        std::printf("invalid temporal depth accepted\n");
    } catch (const std::runtime_error&) {
        std::printf("invalid temporal depth rejected: ok\n");
    }
    std::printf(ok ? "ALL OK\n" : "FAILED\n");
    return ok ? 0 : 1;
}
