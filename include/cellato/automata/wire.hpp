// WireWorld, Moore neighbourhood, 4 states / 2 planes.
// Rule: src/wire/algorithm.hpp:48-59; dictionary: src/wire/config.hpp:21-23.
#ifndef CELLATO_AUTOMATA_WIRE_HPP
#define CELLATO_AUTOMATA_WIRE_HPP
#include "../core/ast.hpp"
#include "../memory/state_dictionary.hpp"

namespace wire {
using namespace cellato::ast;

enum class wire_cell_state { empty, electron_head, electron_tail, conductor };

using empty         = state_constant<wire_cell_state::empty>;
using electron_head = state_constant<wire_cell_state::electron_head>;
using electron_tail = state_constant<wire_cell_state::electron_tail>;
using conductor     = state_constant<wire_cell_state::conductor>;

using head_count = count_neighbors<electron_head, moore_8_neighbors>;
using one_or_two_heads = p<p<head_count, equals, constant<1>>, or_, p<head_count, equals, constant<2>>>;

using wire_algorithm =
    if_<p<current_state, equals, electron_head>>::then_<electron_tail>::
    elif_<p<current_state, equals, electron_tail>>::then_<conductor>::
    elif_<p<current_state, equals, conductor>>::then_<
        if_<one_or_two_heads>::then_<electron_head>::else_<conductor>
    >::
    else_<empty>;

struct config {
    static constexpr char name[] = "wire";
    static constexpr double average_halo_radius = 1.0;
    using algorithm = wire_algorithm;
    using cell_state = wire_cell_state;
    using state_dictionary = cellato::memory::grids::state_dictionary<
        cell_state::empty, cell_state::electron_head, cell_state::electron_tail, cell_state::conductor>;
};
} // namespace wire
#endif
