// Brian's Brain, Moore neighbourhood, 3 states / 2 planes.
// Rule: src/brian/algorithm.hpp:31-44; dictionary: src/brian/config.hpp:21-22.
#ifndef CELLATO_AUTOMATA_BRIAN_HPP
#define CELLATO_AUTOMATA_BRIAN_HPP
#include "../core/ast.hpp"
#include "../memory/state_dictionary.hpp"

namespace brian {
using namespace cellato::ast;

enum class brian_cell_state { dead, dying, alive };

using dead  = state_constant<brian_cell_state::dead>;
using dying = state_constant<brian_cell_state::dying>;
using alive = state_constant<brian_cell_state::alive>;

using has_two_alive_neighbors = p<count_neighbors<alive, moore_8_neighbors>, equals, constant<2>>;

using brian_algorithm =
    if_<p<current_state, equals, dead>>::then_<
        if_<has_two_alive_neighbors>::then_<alive>::else_<dead>
    >::
    elif_<p<current_state, equals, alive>>::then_<dying>::
    else_<dead>;

struct config {
    static constexpr char name[] = "brian";
    static constexpr double average_halo_radius = 1.0;
    using algorithm = brian_algorithm;
    using cell_state = brian_cell_state;
    using state_dictionary = cellato::memory::grids::state_dictionary<
        cell_state::dead, cell_state::dying, cell_state::alive>;
};
} // namespace brian
#endif
