// Kernel instantiations for automaton id 8 of include/cellato/automata/all.hpp.
#include "cellato/automata/all.hpp"
#include "dispatch.hpp"

namespace cellato::hip {
const rule_kernels& kernels_rule_8() {
    // tiled temporal passes run as rows from depth 3 on (Critters, 16384^2, us per launch:
    // rows 26.9 / 30.7 vs blocks 28.3 / 32.4 at depths 3 / 4; 21.9 vs 21.1 at 2)
    static const rule_kernels k = make_rule_kernels<cellato::automata::by_id<8>::config, MAX_T, MAX_T, 3>();
    return k;
}
} // namespace cellato::hip
