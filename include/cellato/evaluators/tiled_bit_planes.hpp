// The reference's evaluator for the tiled bit-plane layout, by name:
// cellato::evaluators::tiled_bit_planes::evaluator<word, dictionary, Expression>
// (include/evaluators/tiled_bit_planes.hpp:29 there), as the test suites bind it
// (include/experiments/test_suites.hpp:158,186). The rule is lowered once, by the bit-plane
// evaluator; the 8x8-block neighbour logic is in the window the HIP kernels pass to it, so
// this type differs from bit_planes::evaluator only in the layout it declares, which the HIP
// traversers check against the grid's.
#ifndef CELLATO_EVALUATORS_TILED_BIT_PLANES_HPP
#define CELLATO_EVALUATORS_TILED_BIT_PLANES_HPP

#include "bit_planes.hpp"

namespace cellato::evaluators::tiled_bit_planes {

template <typename cell_row_type, typename state_dictionary_type, typename Expression>
struct evaluator : cellato::evaluators::bit_planes::evaluator<cell_row_type, state_dictionary_type, Expression> {
    static constexpr int layout = 1;
};

} // namespace cellato::evaluators::tiled_bit_planes

#endif
