// Kernel instantiations for automaton id 9 of include/cellato/automata/all.hpp.
#include "cellato/automata/all.hpp"
#include "dispatch.hpp"

namespace cellato::hip {
const rule_kernels& kernels_rule_9() {
    // cyclic (5 planes): linear kernels deeper than 4 generations would spill to scratch
    static const rule_kernels k = make_rule_kernels<cellato::automata::by_id<9>::config, 4>();
    return k;
}
} // namespace cellato::hip
