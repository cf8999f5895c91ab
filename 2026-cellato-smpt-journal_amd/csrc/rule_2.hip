// Kernel instantiations for automaton id 2 of include/cellato/automata/all.hpp.
#include "cellato/automata/all.hpp"
#include "dispatch.hpp"

namespace cellato::hip {
const rule_kernels& kernels_rule_2() {
    static const rule_kernels k = make_rule_kernels<cellato::automata::by_id<2>::config>();
    return k;
}
} // namespace cellato::hip
