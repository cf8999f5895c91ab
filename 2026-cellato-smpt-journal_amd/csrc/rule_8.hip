// Kernel instantiations for automaton id 8 of include/cellato/automata/all.hpp.
#include "cellato/automata/all.hpp"
#include "dispatch.hpp"

namespace cellato::hip {
const rule_kernels& kernels_rule_8() {
    // tiled temporal passes run as rows from depth 4 on (Critters, 16384^2, us per launch:
    // rows 39.2 / 75.7 vs blocks 40.0 / 83.2 at depths 4 / 8 (32.8 vs 32.6 at 3))
    static const rule_kernels k = make_rule_kernels<cellato::automata::by_id<8>::config, MAX_T, MAX_T, 4>();
    return k;
}
} // namespace cellato::hip
