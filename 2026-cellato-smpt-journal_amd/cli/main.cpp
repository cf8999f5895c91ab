// `cellato`: the reference's command line (src/main.cpp:131-222 options, :238-330 dispatch)
// for the HIP traversers. Same option names, required options and messages, so that the
// reference's sweep scripts run unchanged; one CSV line per experiment on stdout in the
// reference's columns (include/experiments/experiment_report.hpp:82-93), a readable report
// on stderr. Runs the bit_planes and tiled_bit_planes evaluators with the simple and temporal
// traversers, and the bit_array evaluator (32- or 64-bit words) with the simple traverser, for
// all ten automata on the GPU (--device HIP, or CUDA as the reference's scripts spell it);
// other combinations (CPU paths, the standard evaluator, --reference_impl) print the
// reference's "No suitable ..." message.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "cellato/automata/all.hpp"
#include "cellato/evaluators/bit_array.hpp"
#include "cellato/evaluators/bit_planes.hpp"
#include "cellato/evaluators/tiled_bit_planes.hpp"
#include "cellato/experiments/experiment_report.hpp"
#include "cellato/experiments/initial_state.hpp"
#include "cellato/experiments/run_params.hpp"
#include "cellato/memory/bit_array_grid.hpp"
#include "cellato/memory/bit_planes_grid.hpp"
#include "cellato/memory/standard_grid.hpp"
#include "cellato/traversers/hip/traversers.hpp"

namespace {

using cellato::run::run_params;

// --name value tokens; a flag is a name without a value
class options {
public:
    options(int argc, char** argv) : _t(argv + 1, argv + argc) {}
    bool has(const std::string& name) const { return std::find(_t.begin(), _t.end(), "--" + name) != _t.end(); }
    std::string get(const std::string& name) const {
        auto it = std::find(_t.begin(), _t.end(), "--" + name);
        return it != _t.end() && it + 1 != _t.end() ? *(it + 1) : std::string();
    }
    int num(const std::string& name, int fallback) const {
        if (!has(name)) return fallback;
        try {
            return std::stoi(get(name));
        } catch (const std::exception&) {
            std::cerr << "Invalid number for --" << name << ": '" << get(name) << "'" << std::endl;
            std::exit(1);
        }
    }

private:
    std::vector<std::string> _t;
};

void print_usage() {
    std::cout << "Usage: ./cellato [options]\n"
                 "Options:\n"
                 "  --automaton <name>              game-of-life, forest-fire, wire, excitable, brian, maze,\n"
                 "                                  traffic, fluid, critters, cyclic\n"
                 "  --device <name>                 HIP (CUDA is accepted as the same device)\n"
                 "  --traverser <name>              simple, temporal\n"
                 "  --evaluator <name>              bit_planes, tiled_bit_planes, bit_array\n"
                 "  --layout <name>                 bit_planes, tiled_bit_planes, bit_array (same as the evaluator)\n"
                 "  --x_size <number>               X size of the grid (cells)\n"
                 "  --y_size <number>               Y size of the grid (cells)\n"
                 "  --steps <number>                Number of generations\n"
                 "  --word_size <number>            64 (required with bit_planes evaluators)\n"
                 "  --temporal_steps <number>       Generations per launch of the temporal traverser (1..8)\n"
                 "  --temporal_tile_size_y <number> Required with --traverser temporal; recorded only\n"
                 "  --chunk_rows <number>           Rows a wave streams per task (0 = automatic)\n"
                 "  --rounds <number>               Measured rounds (default 1)\n"
                 "  --warmup_rounds <number>        Unmeasured rounds before them (default 0)\n"
                 "  --seed <number>                 Random seed of the initial state (default 42)\n"
                 "  --x_tile_size, --y_tile_size, --cuda_block_size_x, --cuda_block_size_y <number>\n"
                 "                                  Accepted and recorded in the CSV line, not used\n"
                 "  --print                         Print the run parameters\n"
                 "  --print_csv_header              Print the CSV header\n"
                 "  --help                          Show this help message\n";
}

run_params read_params(const options& o) {
    run_params p;
    if (o.has("help")) {
        p.help = true;
        return p;
    }
    if (o.has("print_csv_header")) {
        p.print_csv_header = true;
        return p;
    }
    std::vector<std::string> required{"automaton", "device", "traverser", "evaluator", "layout", "x_size", "y_size", "steps"};
    if (o.get("evaluator") == "bit_planes" || o.get("evaluator") == "bit_array" || o.get("evaluator") == "tiled_bit_planes")
        required.push_back("word_size");
    if (o.get("traverser") == "temporal") {
        required.push_back("temporal_steps");
        required.push_back("temporal_tile_size_y");
    }
    if (o.has("reference_impl"))
        for (const char* gone : {"device", "traverser", "evaluator", "layout"})
            required.erase(std::remove(required.begin(), required.end(), gone), required.end());
    for (const auto& r : required)
        if (!o.has(r)) {
            std::cerr << "Missing required option: " << r << std::endl;
            std::exit(1);
        }
    p.automaton = o.get("automaton");
    p.device = o.get("device");
    p.traverser = o.get("traverser");
    p.evaluator = o.get("evaluator");
    p.layout = o.get("layout");
    p.reference_impl = o.has("reference_impl") ? o.get("reference_impl") : "none";
    p.x_size = o.num("x_size", 0);
    p.y_size = o.num("y_size", 0);
    p.steps = o.num("steps", 0);
    p.word_size = o.num("word_size", 0);
    p.x_tile_size = o.num("x_tile_size", 0);
    p.y_tile_size = o.num("y_tile_size", 0);
    p.temporal_steps = o.num("temporal_steps", 0);
    p.temporal_tile_size_y = o.num("temporal_tile_size_y", 0);
    p.rounds = o.num("rounds", 1);
    p.warmup_rounds = o.num("warmup_rounds", 0);
    p.seed = o.num("seed", 42);
    p.print = o.has("print");
    p.cuda_block_size_x = o.num("cuda_block_size_x", 32);
    p.cuda_block_size_y = o.num("cuda_block_size_y", 8);
    p.chunk_rows = o.num("chunk_rows", 0);
    return p;
}

void hip_ok(hipError_t e, const char* what) { cellato::memory::grids::hip_check(e, what); }

enum class layout_kind { linear, tiled, bit_array };

// grid, evaluator and traverser of a suite, as the reference's test_suites.hpp:86-225 bind them
template <class config, layout_kind L, bool Temporal, class Word>
struct suite_types;
template <class config, bool Temporal, class Word>
struct suite_types<config, layout_kind::linear, Temporal, Word> {
    using dict_t = typename config::state_dictionary;
    using grid_t = cellato::memory::grids::bit_planes::grid<std::uint64_t, dict_t>;
    using eval_t = cellato::evaluators::bit_planes::evaluator<std::uint64_t, dict_t, typename config::algorithm>;
    using traverser_t = std::conditional_t<Temporal, cellato::traversers::hip::temporal::traverser<eval_t, grid_t>,
                                           cellato::traversers::hip::simple::traverser<eval_t, grid_t>>;
};
template <class config, bool Temporal, class Word>
struct suite_types<config, layout_kind::tiled, Temporal, Word> {
    using dict_t = typename config::state_dictionary;
    using grid_t = cellato::memory::grids::tiled_bit_planes::grid<std::uint64_t, dict_t>;
    using eval_t = cellato::evaluators::tiled_bit_planes::evaluator<std::uint64_t, dict_t, typename config::algorithm>;
    using traverser_t = std::conditional_t<Temporal, cellato::traversers::hip::temporal::traverser<eval_t, grid_t>,
                                           cellato::traversers::hip::simple::traverser<eval_t, grid_t>>;
};
template <class config, bool Temporal, class Word>
struct suite_types<config, layout_kind::bit_array, Temporal, Word> {
    using dict_t = typename config::state_dictionary;
    using grid_t = cellato::memory::grids::bit_array::grid<dict_t, Word>;
    using eval_t = cellato::evaluators::bit_array::evaluator<grid_t, typename config::algorithm>;
    using traverser_t = cellato::traversers::hip::simple::traverser<eval_t, grid_t>;
};

// one experiment: rounds of init -> run(steps) (timed with HIP events) -> fetch -> checksum
template <int id, layout_kind L, bool Temporal, class Word = std::uint64_t>
cellato::run::experiment_report run_experiment(const run_params& p) {
    namespace g = cellato::memory::grids;
    using config = typename cellato::automata::by_id<id>::config;
    using cell_t = typename config::cell_state;
    using types = suite_types<config, L, Temporal, Word>;
    using grid_t = typename types::grid_t;
    using traverser_t = typename types::traverser_t;

    const auto values = cellato::experiments::initial_state(p.automaton, p.x_size, p.y_size, p.seed);
    std::vector<cell_t> cells(values.size());
    std::transform(values.begin(), values.end(), cells.begin(), [](std::int32_t v) { return static_cast<cell_t>(v); });
    const grid_t initial(g::standard::grid<cell_t>(std::move(cells), p.x_size, p.y_size));

    cellato::run::experiment_report report;
    report.params = p;
    hipEvent_t start, stop;
    hip_ok(hipEventCreate(&start), "hipEventCreate");
    hip_ok(hipEventCreate(&stop), "hipEventCreate");
    for (int round = 0; round < p.warmup_rounds + p.rounds; ++round) {
        const bool warm = round < p.warmup_rounds;
        std::cerr << (warm ? "\nWarmup round: " : "\nRound: ") << (warm ? round : round - p.warmup_rounds) << "\n";
        traverser_t t;
        t.init(initial, p);
        hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_ok(hipEventRecord(start, nullptr), "hipEventRecord");
        t.run(p.steps);
        hip_ok(hipEventRecord(stop, nullptr), "hipEventRecord");
        hip_ok(hipEventSynchronize(stop), "hipEventSynchronize");
        float ms = 0.0f;
        hip_ok(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
        const std::string sum = t.fetch_result().get_checksum();
        if (!warm) {
            report.execution_times_ms.push_back(ms);
            report.checksums.push_back(sum);
        }
    }
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    return report;
}

template <int id>
bool run_for(const run_params& p) {
    if (p.automaton != cellato::automata::by_id<id>::config::name) return false;
    const bool temporal = p.traverser == "temporal";
    cellato::run::experiment_report r;
    if (p.evaluator == "bit_array")
        r = p.word_size == 32 ? run_experiment<id, layout_kind::bit_array, false, std::uint32_t>(p)
                              : run_experiment<id, layout_kind::bit_array, false, std::uint64_t>(p);
    else if (p.evaluator == "tiled_bit_planes")
        r = temporal ? run_experiment<id, layout_kind::tiled, true>(p) : run_experiment<id, layout_kind::tiled, false>(p);
    else
        r = temporal ? run_experiment<id, layout_kind::linear, true>(p) : run_experiment<id, layout_kind::linear, false>(p);
    std::cout << r.csv_line() << std::endl;
    r.pretty_print(std::cerr);
    return true;
}

// the combinations the HIP traversers serve (the reference's test_suites.hpp:86-225 on_cuda
// suites: bit_planes / tiled_bit_planes with 64-bit words, simple or temporal; bit_array with
// 32- or 64-bit words, simple)
bool served(const run_params& p) {
    const bool gpu = p.device == "HIP" || p.device == "CUDA";
    if (p.evaluator == "bit_array" && p.layout == "bit_array")
        return gpu && p.traverser == "simple" && (p.word_size == 32 || p.word_size == 64);
    const bool evaluator = (p.evaluator == "bit_planes" && p.layout == "bit_planes") ||
                           (p.evaluator == "tiled_bit_planes" && p.layout == "tiled_bit_planes");
    return gpu && evaluator && (p.traverser == "simple" || p.traverser == "temporal") && p.word_size == 64;
}

} // namespace

int main(int argc, char** argv) {
    const options o(argc, argv);
    run_params p = read_params(o);
    if (p.help) {
        print_usage();
        return 0;
    }
    if (p.print_csv_header) {
        std::cout << cellato::run::experiment_report::csv_header() << std::endl;
        return 0;
    }
    if (p.print) p.print_to(std::cerr);
    if (p.reference_impl != "none") {
        std::cerr << "No suitable reference implementation found for the given parameters." << std::endl;
        return 0;
    }
    if (!served(p)) {
        std::cerr << "No suitable test suite found for the given parameters." << std::endl;
        return 0;
    }
    try {
        bool ran = false;
        [&]<int... ids>(std::integer_sequence<int, ids...>) {
            ((ran = ran || run_for<ids>(p)), ...);
        }(std::make_integer_sequence<int, cellato::automata::RULE_COUNT>{});
        if (!ran) std::cerr << "No suitable test suite found for the given parameters." << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
