// Kernel instantiations for automaton id 2 of include/cellato/automata/all.hpp.
#include "cellato/automata/all.hpp"
#include "dispatch.hpp"

namespace cellato::hip {
const rule_kernels& kernels_rule_2() {
    // tiled temporal passes run as rows from depth 3 on (Forest fire, 16384^2, us per launch:
    // rows 32.0 / 32.8 / 44.7 vs blocks 32.5 / 37.4 / 60.7 at depths 3 / 4 / 8)
    static const rule_kernels k = make_rule_kernels<cellato::automata::by_id<2>::config, MAX_T, MAX_T, 3>();
    return k;
}
} // namespace cellato::hip
