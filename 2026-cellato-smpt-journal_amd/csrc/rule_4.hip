// Kernel instantiations for automaton id 4 of include/cellato/automata/all.hpp.
#include "cellato/automata/all.hpp"
#include "dispatch.hpp"

namespace cellato::hip {
const rule_kernels& kernels_rule_4() {
    // tiled temporal passes run as rows from depth 8 on (Brian's Brain, 16384^2, us per launch:
    // rows 53.7 / 59.9 vs blocks 53.6 / 63.1 at depths 7 / 8)
    static const rule_kernels k = make_rule_kernels<cellato::automata::by_id<4>::config, MAX_T, MAX_T, 8>();
    return k;
}
} // namespace cellato::hip
