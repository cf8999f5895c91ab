// Conway's Game of Life (B3/S23), Moore neighbourhood, 2 states / 1 plane.
// Rule as defined by the reference in src/game_of_life/algorithm.hpp:35-49, dictionary
// order from src/game_of_life/config.hpp:21-22.
#ifndef CELLATO_AUTOMATA_GAME_OF_LIFE_HPP
#define CELLATO_AUTOMATA_GAME_OF_LIFE_HPP
#include "../core/ast.hpp"
#include "../memory/state_dictionary.hpp"

namespace game_of_life {
using namespace cellato::ast;

enum class gol_cell_state { dead, alive };

using alive = state_constant<gol_cell_state::alive>;
using dead  = state_constant<gol_cell_state::dead>;

using alive_count = count_neighbors<alive, moore_8_neighbors>;
using survives = p<p<alive_count, equals, constant<2>>, or_, p<alive_count, equals, constant<3>>>;
using born     = p<alive_count, equals, constant<3>>;

using gol_algorithm =
    if_<p<current_state, equals, alive>>::then_<
        if_<survives>::then_<alive>::else_<dead>
    >::else_<
        if_<born>::then_<alive>::else_<dead>
    >;

struct config {
    static constexpr char name[] = "game-of-life";
    static constexpr double average_halo_radius = 1.0;
    using algorithm = gol_algorithm;
    using cell_state = gol_cell_state;
    using state_dictionary = cellato::memory::grids::state_dictionary<cell_state::dead, cell_state::alive>;
};
} // namespace game_of_life
#endif
