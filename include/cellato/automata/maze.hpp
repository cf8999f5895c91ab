// Maze (B3/S12345), Moore neighbourhood, 2 states / 1 plane.
// Rule: src/maze/algorithm.hpp:29-35; dictionary: src/maze/config.hpp:22-23.
#ifndef CELLATO_AUTOMATA_MAZE_HPP
#define CELLATO_AUTOMATA_MAZE_HPP
#include "../core/ast.hpp"
#include "../memory/state_dictionary.hpp"

namespace maze {
using namespace cellato::ast;

enum class maze_cell_state { empty, wall };

using empty = state_constant<maze_cell_state::empty>;
using wall  = state_constant<maze_cell_state::wall>;

using wall_count = count_neighbors<wall, moore_8_neighbors>;

using maze_algorithm =
    if_<p<wall_count, equals, constant<3>>>::then_<wall>::
    elif_<p<p<current_state, equals, wall>, and_, p<wall_count, less_than, constant<6>>>>::then_<wall>::
    else_<empty>;

struct config {
    static constexpr char name[] = "maze";
    static constexpr double average_halo_radius = 1.0;
    using algorithm = maze_algorithm;
    using cell_state = maze_cell_state;
    using state_dictionary = cellato::memory::grids::state_dictionary<cell_state::empty, cell_state::wall>;
};
} // namespace maze
#endif
