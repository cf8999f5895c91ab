// Run parameters read by the traversers. Field names and meaning follow the reference's
// include/experiments/run_params.hpp:10-43; `chunk_rows` (rows a wave streams per task,
// 0 = library default) replaces the CUDA block/tile sizes, which have no meaning here.
#ifndef CELLATO_EXPERIMENTS_RUN_PARAMS_HPP
#define CELLATO_EXPERIMENTS_RUN_PARAMS_HPP

#include <string>

namespace cellato::run {

struct run_params {
    std::string automaton = "game-of-life";
    std::string device = "HIP";
    std::string traverser = "simple";
    std::string evaluator = "bit_planes";
    std::string layout = "bit_planes";
    int x_size = 0;
    int y_size = 0;
    int steps = 0;
    int word_size = 64;
    int temporal_steps = 0;
    int temporal_tile_size_y = 0;
    int rounds = 1;
    int warmup_rounds = 0;
    int seed = 42;
    int chunk_rows = 0;
};

} // namespace cellato::run

#endif
