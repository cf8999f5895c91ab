// Greenberg-Hastings excitable medium, Moore neighbourhood, 8 states / 3 planes.
// Rule: src/excitable/algorithm.hpp:57-71; dictionary: src/excitable/config.hpp:21-25.
#ifndef CELLATO_AUTOMATA_EXCITABLE_HPP
#define CELLATO_AUTOMATA_EXCITABLE_HPP
#include "../core/ast.hpp"
#include "../memory/state_dictionary.hpp"

namespace excitable {
using namespace cellato::ast;

enum class ghm_cell_state {
    quiescent = 0, excited = 1,
    refractory_1 = 2, refractory_2 = 3, refractory_3 = 4,
    refractory_4 = 5, refractory_5 = 6, refractory_6 = 7
};

using quiescent    = state_constant<ghm_cell_state::quiescent>;
using excited      = state_constant<ghm_cell_state::excited>;
using refractory_1 = state_constant<ghm_cell_state::refractory_1>;
using refractory_2 = state_constant<ghm_cell_state::refractory_2>;
using refractory_3 = state_constant<ghm_cell_state::refractory_3>;
using refractory_4 = state_constant<ghm_cell_state::refractory_4>;
using refractory_5 = state_constant<ghm_cell_state::refractory_5>;
using refractory_6 = state_constant<ghm_cell_state::refractory_6>;

template <typename S> using is = p<current_state, equals, S>;

using has_excited_neighbors =
    p<count_neighbors<excited, moore_8_neighbors>, greater_than, constant<0>>;

using ghm_algorithm =
    if_<is<quiescent>>::then_<
        if_<has_excited_neighbors>::then_<excited>::else_<quiescent>
    >::
    elif_<is<excited>>::then_<refractory_1>::
    elif_<is<refractory_1>>::then_<refractory_2>::
    elif_<is<refractory_2>>::then_<refractory_3>::
    elif_<is<refractory_3>>::then_<refractory_4>::
    elif_<is<refractory_4>>::then_<refractory_5>::
    elif_<is<refractory_5>>::then_<refractory_6>::
    else_<quiescent>;

struct config {
    static constexpr char name[] = "excitable";
    static constexpr double average_halo_radius = 1.0;
    using algorithm = ghm_algorithm;
    using cell_state = ghm_cell_state;
    using state_dictionary = cellato::memory::grids::state_dictionary<
        ghm_cell_state::quiescent, ghm_cell_state::excited,
        ghm_cell_state::refractory_1, ghm_cell_state::refractory_2,
        ghm_cell_state::refractory_3, ghm_cell_state::refractory_4,
        ghm_cell_state::refractory_5, ghm_cell_state::refractory_6>;
};
} // namespace excitable
#endif
