// Critters: reversible Margolus-block rule, 2 states / 1 plane. The 2x2 block offset
// alternates with the time step. Rule: src/critters/algorithm.hpp:16-31;
// dictionary: src/critters/config.hpp:21-22.
#ifndef CELLATO_AUTOMATA_CRITTERS_HPP
#define CELLATO_AUTOMATA_CRITTERS_HPP
#include "../core/ast.hpp"
#include "../memory/state_dictionary.hpp"

namespace critters {
using namespace cellato::ast;

enum class critters_cell_state { dead, alive };

using dead  = state_constant<critters_cell_state::dead>;
using alive = state_constant<critters_cell_state::alive>;

using alive_count = count_neighbors<alive, margolus_alternating_neighborhood>;

using toggled_rotated = if_<p<margolus_180_neighbor, equals, alive>>::then_<dead>::else_<alive>;
using toggled         = if_<p<current_state, equals, alive>>::then_<dead>::else_<alive>;

using critters_algorithm =
    if_<p<alive_count, equals, constant<3>>>::then_<toggled_rotated>::
    elif_<p<alive_count, equals, constant<2>>>::then_<current_state>::
    else_<toggled>;

struct config {
    static constexpr char name[] = "critters";
    static constexpr double average_halo_radius = 1.0;
    using algorithm = critters_algorithm;
    using cell_state = critters_cell_state;
    using state_dictionary = cellato::memory::grids::state_dictionary<cell_state::dead, cell_state::alive>;
};
} // namespace critters
#endif
