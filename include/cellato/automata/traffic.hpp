// Biham-Middleton-Levine traffic, 3 states / 2 planes. Even steps: red cars move right;
// odd steps: blue cars move down. Rule: src/traffic/algorithm.hpp:15-61;
// dictionary: src/traffic/config.hpp:21-22.
#ifndef CELLATO_AUTOMATA_TRAFFIC_HPP
#define CELLATO_AUTOMATA_TRAFFIC_HPP
#include "../core/ast.hpp"
#include "../memory/state_dictionary.hpp"

namespace traffic {
using namespace cellato::ast;

enum class traffic_cell_state { empty, red_car, blue_car };

using empty = state_constant<traffic_cell_state::empty>;

// One half-step of the rule: cars of kind `movable_car` advance from `incoming_neighbor`
// towards `outgoing_neighbor` when the target cell is empty.
template <typename movable_car, typename stationary_car, typename incoming_neighbor, typename outgoing_neighbor>
struct one_direction {
    using algorithm =
        typename if_<p<current_state, equals, movable_car>>::template then_<
            typename if_<p<outgoing_neighbor, equals, empty>>::template then_<empty>::template else_<movable_car>
        >::template elif_<p<current_state, equals, empty>>::template then_<
            typename if_<p<incoming_neighbor, equals, movable_car>>::template then_<movable_car>::template else_<empty>
        >::template else_<stationary_car>;
};

using red_car  = state_constant<traffic_cell_state::red_car>;
using blue_car = state_constant<traffic_cell_state::blue_car>;

using up    = neighbor_at<0, -1>;
using down  = neighbor_at<0, 1>;
using left  = neighbor_at<-1, 0>;
using right = neighbor_at<1, 0>;

using traffic_algorithm = alternate_algorithms<
    one_direction<red_car, blue_car, left, right>::algorithm,
    one_direction<blue_car, red_car, up, down>::algorithm>;

struct config {
    static constexpr char name[] = "traffic";
    static constexpr double average_halo_radius = 1.0;
    using algorithm = traffic_algorithm;
    using cell_state = traffic_cell_state;
    using state_dictionary = cellato::memory::grids::state_dictionary<
        cell_state::empty, cell_state::red_car, cell_state::blue_car>;
};
} // namespace traffic
#endif
