// Cyclic cellular automaton with 32 states / 5 planes: a cell advances to (s+1) mod 32 when
// any Moore neighbour already holds that state. Rule: src/cyclic/algorithm.hpp:11-23;
// dictionary int_based<5> (src/cyclic/config.hpp:21).
#ifndef CELLATO_AUTOMATA_CYCLIC_HPP
#define CELLATO_AUTOMATA_CYCLIC_HPP
#include "../core/ast.hpp"
#include "../memory/state_dictionary.hpp"

namespace cyclic {
using namespace cellato::ast;

using cyclic_cell_state = int;

constexpr int BITS = 5;
constexpr int STATES = 1 << BITS;

using successor = p<p<current_state, plus, constant<1>>, modulo, constant<STATES>>;
using successor_count = count_neighbors<successor, moore_8_neighbors>;

using cyclic_algorithm =
    if_<p<successor_count, greater_than, constant<0>>>::then_<successor>::else_<current_state>;

struct config {
    static constexpr char name[] = "cyclic";
    static constexpr double average_halo_radius = 1.0;
    using algorithm = cyclic_algorithm;
    using cell_state = cyclic_cell_state;
    using state_dictionary = cellato::memory::grids::int_based_state_dictionary<BITS>;
};
} // namespace cyclic
#endif
