// Run parameters of an experiment. Field names, defaults and the CSV columns follow the
// reference's include/experiments/run_params.hpp:10-97, so that a command line written for
// the reference's `cellato` binary means the same here and its CSV lines line up with the
// published ones. `chunk_rows` (rows a wave streams per task, 0 = the library's choice) is
// the only addition; the CUDA block and tile sizes are accepted and recorded, not used.
#ifndef CELLATO_EXPERIMENTS_RUN_PARAMS_HPP
#define CELLATO_EXPERIMENTS_RUN_PARAMS_HPP

#include <ostream>
#include <string>

namespace cellato::run {

struct run_params {
    std::string automaton = "game-of-life";
    std::string device = "HIP";
    std::string traverser = "simple";
    std::string evaluator = "bit_planes";
    std::string layout = "bit_planes";
    std::string reference_impl = "none";
    int x_size = 0;
    int y_size = 0;
    int steps = 0;
    int word_size = 64;
    int x_tile_size = 0;
    int y_tile_size = 0;
    int temporal_steps = 0;
    int temporal_tile_size_y = 0;
    int rounds = 1;
    int warmup_rounds = 0;
    int seed = 42;
    bool print = false;
    bool help = false;
    bool print_csv_header = false;
    int cuda_block_size_x = 32;
    int cuda_block_size_y = 8;
    int chunk_rows = 0;

    // the reference's column order (run_params.hpp:73-75)
    static std::string csv_header() {
        return "automaton,device,traverser,evaluator,layout,reference_impl,x_size,y_size,steps,rounds,warmup_rounds,"
               "x_tile_size,y_tile_size,cuda_block_size_x,cuda_block_size_y,seed,word_size,temporal_steps,"
               "temporal_tile_size_y";
    }

    std::string csv_line() const {
        const auto n = [](int v) { return std::to_string(v); };
        std::string s = automaton + "," + device + "," + traverser + "," + evaluator + "," + layout + ",";
        if (reference_impl != "none") s += reference_impl;
        for (int v : {x_size, y_size, steps, rounds, warmup_rounds, x_tile_size, y_tile_size, cuda_block_size_x,
                      cuda_block_size_y, seed, word_size, temporal_steps, temporal_tile_size_y})
            s += "," + n(v);
        return s;
    }

    void print_to(std::ostream& os) const {
        os << "Run Parameters:\n"
           << "  Automaton: " << automaton << "\n  Device: " << device << "\n  Traverser: " << traverser
           << "\n  Evaluator: " << evaluator << "\n  Layout: " << layout << "\n  Reference Implementation: "
           << reference_impl << "\n  Grid Size: (" << x_size << ", " << y_size << ")\n  Steps: " << steps
           << "\n  Word Size: " << word_size << "\n  Rounds: " << rounds << "\n  Warmup Rounds: " << warmup_rounds
           << "\n  Seed: " << seed << "\n  Temporal Steps: " << temporal_steps
           << "\n  Temporal Tile Size Y: " << temporal_tile_size_y << "\n  Chunk Rows: " << chunk_rows << "\n";
    }
};

} // namespace cellato::run

#endif
