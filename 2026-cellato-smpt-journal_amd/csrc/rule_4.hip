// Kernel instantiations for automaton id 4 of include/cellato/automata/all.hpp.
#include "cellato/automata/all.hpp"
#include "dispatch.hpp"

namespace cellato::hip {
const rule_kernels& kernels_rule_4() {
    // tiled temporal passes run as rows from depth 5 on (Brian's Brain, 16384^2, us per launch:
    // rows 38.2 / 49.1 vs blocks 39.1 / 52.3 at depths 5 / 7; 35.4 vs 33.4 at 4)
    static const rule_kernels k = make_rule_kernels<cellato::automata::by_id<4>::config, MAX_T, MAX_T, 5>();
    return k;
}
} // namespace cellato::hip
