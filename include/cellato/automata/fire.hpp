// Forest fire, von Neumann neighbourhood, 4 states / 2 planes.
// Rule: src/fire/algorithm.hpp:39-57. NOTE the dictionary (src/fire/config.hpp:22-24) lists
// fire before ash, so plane indices differ from the enum values (ash=2, fire=3).
#ifndef CELLATO_AUTOMATA_FIRE_HPP
#define CELLATO_AUTOMATA_FIRE_HPP
#include "../core/ast.hpp"
#include "../memory/state_dictionary.hpp"

namespace fire {
using namespace cellato::ast;

enum class fire_cell_state { empty, tree, ash, fire };

using empty = state_constant<fire_cell_state::empty>;
using tree  = state_constant<fire_cell_state::tree>;
using ash   = state_constant<fire_cell_state::ash>;
using fire  = state_constant<fire_cell_state::fire>;

using fire_count = count_neighbors<fire, von_neumann_4_neighbors>;
using has_fire_neighbors = p<fire_count, greater_than, constant<0>>;

using fire_algorithm =
    if_<p<current_state, equals, fire>>::then_<ash>::
    elif_<p<current_state, equals, ash>>::then_<
        if_<has_fire_neighbors>::then_<ash>::else_<empty>
    >::
    elif_<p<current_state, equals, tree>>::then_<
        if_<has_fire_neighbors>::then_<fire>::else_<tree>
    >::
    else_<empty>;

struct config {
    static constexpr char name[] = "forest-fire";
    static constexpr double average_halo_radius = 1.0;
    using algorithm = fire_algorithm;
    using cell_state = fire_cell_state;
    using state_dictionary = cellato::memory::grids::state_dictionary<
        cell_state::empty, cell_state::tree, cell_state::fire, cell_state::ash>;
};
} // namespace fire
#endif
