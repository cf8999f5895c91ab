// Kernel instantiations for automaton id 0 of include/cellato/automata/all.hpp.
#include "cellato/automata/all.hpp"
#include "dispatch.hpp"

namespace cellato::hip {
const rule_kernels& kernels_rule_0() {
    // tiled temporal passes run as rows from depth 3 on (Game of Life, 16384^2, us per launch:
    // rows 22.0 / 24.3 / 44.0 vs blocks 22.8 / 27.5 / 54.6 at depths 3 / 4 / 8)
    static const rule_kernels k = make_rule_kernels<cellato::automata::by_id<0>::config, MAX_T, MAX_T, 3>();
    return k;
}
} // namespace cellato::hip
