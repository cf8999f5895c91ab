// Kernel instantiations for automaton id 5 of include/cellato/automata/all.hpp.
#include "cellato/automata/all.hpp"
#include "dispatch.hpp"

namespace cellato::hip {
const rule_kernels& kernels_rule_5() {
    // tiled temporal passes run as rows from depth 3 on (Maze, 16384^2, us per launch:
    // rows 24.7 / 27.6 / 50.9 vs blocks 25.8 / 31.2 / 62.2 at depths 3 / 4 / 8)
    static const rule_kernels k = make_rule_kernels<cellato::automata::by_id<5>::config, MAX_T, MAX_T, 3>();
    return k;
}
} // namespace cellato::hip
