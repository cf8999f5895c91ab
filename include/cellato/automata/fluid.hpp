// HPP lattice gas, 4 direction bits / 4 planes, von Neumann neighbourhood.
// Rule: src/fluid/algorithm.hpp:11-60; dictionary int_based<4> (src/fluid/config.hpp:21).
#ifndef CELLATO_AUTOMATA_FLUID_HPP
#define CELLATO_AUTOMATA_FLUID_HPP
#include "../core/ast.hpp"
#include "../memory/state_dictionary.hpp"

namespace fluid {
using namespace cellato::ast;

using fluid_cell_state = int;

constexpr int TOP_bit = 0,    TOP = 1 << TOP_bit;
constexpr int BOTTOM_bit = 1, BOTTOM = 1 << BOTTOM_bit;
constexpr int LEFT_bit = 2,   LEFT = 1 << LEFT_bit;
constexpr int RIGHT_bit = 3,  RIGHT = 1 << RIGHT_bit;

using above = neighbor_at<0, -1>;
using below = neighbor_at<0, 1>;
using west  = neighbor_at<-1, 0>;
using east  = neighbor_at<1, 0>;

using vertical_collision   = p<has_bit_set<above, TOP_bit>,  bit_and_, has_bit_set<below, BOTTOM_bit>>;
using horizontal_collision = p<has_bit_set<west,  LEFT_bit>, bit_and_, has_bit_set<east,  RIGHT_bit>>;

using combined_vertical_incoming =
    p<p<above, bit_and_, constant<TOP>>, bit_or_, p<below, bit_and_, constant<BOTTOM>>>;
using combined_horizontal_incoming =
    p<p<west, bit_and_, constant<LEFT>>, bit_or_, p<east, bit_and_, constant<RIGHT>>>;

using just_vertical_collision =
    p<vertical_collision, bit_and_, p<combined_horizontal_incoming, equals, constant<0>>>;
using just_horizontal_collision =
    p<horizontal_collision, bit_and_, p<combined_vertical_incoming, equals, constant<0>>>;

using vertical_result =
    if_<just_vertical_collision>::then_<state_constant<LEFT | RIGHT>>::else_<combined_vertical_incoming>;
using horizontal_result =
    if_<just_horizontal_collision>::then_<state_constant<TOP | BOTTOM>>::else_<combined_horizontal_incoming>;

using fluid_algorithm = p<vertical_result, bit_or_, horizontal_result>;

struct config {
    static constexpr char name[] = "fluid";
    static constexpr double average_halo_radius = 1.0;
    using algorithm = fluid_algorithm;
    using cell_state = fluid_cell_state;
    using state_dictionary = cellato::memory::grids::int_based_state_dictionary<4>;
};
} // namespace fluid
#endif
