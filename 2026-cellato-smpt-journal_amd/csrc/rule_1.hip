// Kernel instantiations for automaton id 1 of include/cellato/automata/all.hpp.
#include "cellato/automata/all.hpp"
#include "dispatch.hpp"

namespace cellato::hip {
const rule_kernels& kernels_rule_1() {
    // tiled temporal passes run as rows from depth 5 on (WireWorld, 16384^2, us per launch:
    // rows 39.3 / 51.4 / 55.9 vs blocks 40.4 / 54.0 / 63.4 at depths 5 / 7 / 8; 36.1 vs 33.3 at 4)
    static const rule_kernels k = make_rule_kernels<cellato::automata::by_id<1>::config, MAX_T, MAX_T, 5>();
    return k;
}
} // namespace cellato::hip
