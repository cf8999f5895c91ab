// Build-time instantiation table: for every automaton, the stream kernel instantiated for
// each (layout, lane-exchange, halo, T) the host launcher can pick. Plays the role of the
// reference's explicit-instantiation glue (src/_shared/cuda_instantiation_template.cuh:19-193,
// included from traversers/cuda/simple.cu:101-114 and temporal.cu) and of its run-time ->
// compile-time dispatcher (include/utils/static_dispatcher.hpp:27-92): here the host code
// indexes a table instead of searching option sets.
#ifndef CELLATO_HIP_DISPATCH_HPP
#define CELLATO_HIP_DISPATCH_HPP

#include <cstdint>

#include "bit_array.hpp"
#include "stream.hpp"

namespace cellato::hip {

constexpr int MAX_T = 8;   // generations per launch of the temporal traverser

enum simple_variant : int {
    SIMPLE_SEGMENT = 0,    // rows wider than 64 columns: DPP, end lanes load the outside column
    SIMPLE_RING64 = 1,     // rows of exactly 64 columns
    SIMPLE_GROUPS = 2,     // rows of fewer than 64 columns, several rows per wave
    SIMPLE_VARIANTS = 3
};

enum temporal_variant : int {
    TEMPORAL_SEGMENT = 0,  // rows wider than 64 columns: 62 owned + 2 halo lanes per wave
    TEMPORAL_RING64 = 1,
    TEMPORAL_GROUPS = 2,
    TEMPORAL_VARIANTS = 3
};

struct rule_kernels {
    const char* name;
    int planes;
    int nstates;                 // states of the dictionary
    unsigned traits;             // rule_trait flags of the rule's AST
    int state_values[32];        // enum value of each dictionary index (for checksums)
    const void* simple[2][SIMPLE_VARIANTS];                   // [layout][variant]
    const void* temporal[2][TEMPORAL_VARIANTS][MAX_T + 1];    // [layout][variant][T], T >= 1
    const void* temporal_rows[2][MAX_T + 1];                  // tiled planes processed as rows: [segments / ring of 64][T], T >= 2
    const void* batch[MAX_T + 1];                             // linear layout, [T], T >= 1
    const void* bit_array[2];                                 // bit_array layout, [0] 32-bit, [1] 64-bit words
};

// MaxT: the deepest temporal kernel instantiated for this layout; deeper entries stay null and
// the planner steps a pass down to the deepest depth that exists and uses no scratch memory.
template <class Eval, int L, int MaxT = MAX_T>
constexpr void fill_layout(rule_kernels& k) {
    k.simple[L][SIMPLE_SEGMENT] = reinterpret_cast<const void*>(&stream_kernel<Eval, L, 1, XCH_SEGMENT, HALO_LOAD>);
    k.simple[L][SIMPLE_RING64] = reinterpret_cast<const void*>(&stream_kernel<Eval, L, 1, XCH_RING64, HALO_NONE>);
    k.simple[L][SIMPLE_GROUPS] = reinterpret_cast<const void*>(&stream_kernel<Eval, L, 1, XCH_PERMUTE, HALO_NONE>);
    k.temporal[L][TEMPORAL_SEGMENT][0] = nullptr;
    k.temporal[L][TEMPORAL_RING64][0] = nullptr;
    k.temporal[L][TEMPORAL_GROUPS][0] = nullptr;
    for (int v = 0; v < TEMPORAL_VARIANTS; ++v)
        for (int t = MaxT + 1; t <= MAX_T; ++t) k.temporal[L][v][t] = nullptr;
    [&]<int... Ts>(std::integer_sequence<int, Ts...>) {
        ((k.temporal[L][TEMPORAL_SEGMENT][Ts + 1] =
              reinterpret_cast<const void*>(&stream_kernel<Eval, L, Ts + 1, XCH_SEGMENT, HALO_LANES>)), ...);
        ((k.temporal[L][TEMPORAL_RING64][Ts + 1] =
              reinterpret_cast<const void*>(&stream_kernel<Eval, L, Ts + 1, XCH_RING64, HALO_NONE>)), ...);
        ((k.temporal[L][TEMPORAL_GROUPS][Ts + 1] =
              reinterpret_cast<const void*>(&stream_kernel<Eval, L, Ts + 1, XCH_PERMUTE, HALO_NONE>)), ...);
    }(std::make_integer_sequence<int, MaxT>{});
}

// Temporal kernels that process tiled planes as rows (TILED_ROWS), for depths FromT .. MaxT of
// rules of at most two planes (the LDS staging areas take 32 KiB per plane and block); null
// elsewhere. The planner takes them for rows of at least 64 columns of 64 cells.
template <class Eval, int FromT, int MaxT>
constexpr void fill_rows(rule_kernels& k) {
    for (int v = 0; v < 2; ++v)
        for (int t = 0; t <= MAX_T; ++t) k.temporal_rows[v][t] = nullptr;
    if constexpr (Eval::planes <= 2 && FromT >= 2 && MaxT >= FromT) {
        [&]<int... Ts>(std::integer_sequence<int, Ts...>) {
            ((k.temporal_rows[0][Ts + FromT] =
                  reinterpret_cast<const void*>(&stream_kernel<Eval, TILED_ROWS, Ts + FromT, XCH_SEGMENT, HALO_LANES>)), ...);
            ((k.temporal_rows[1][Ts + FromT] =
                  reinterpret_cast<const void*>(&stream_kernel<Eval, TILED_ROWS, Ts + FromT, XCH_RING64, HALO_NONE>)), ...);
        }(std::make_integer_sequence<int, MaxT - FromT + 1>{});
    }
}

// MaxLinearT / MaxTiledT: the deepest depth whose kernels of that layout fit the register file
// without scratch memory (every ring slot also holds the neighbours' edge bits: halves of the
// neighbouring words in the linear layout, an edge word per plane in the tiled one, so a rule
// with many planes runs out of VGPRs at a lower depth). RowsFromT: the shallowest depth from
// which the tiled layout's temporal passes run as rows (TILED_ROWS), measured per automaton
// at 16384^2 as the depth from which they beat the block kernels; above MAX_T: never.
template <class Config, int MaxLinearT = MAX_T, int MaxTiledT = MAX_T, int RowsFromT = MAX_T + 1>
rule_kernels make_rule_kernels() {
    using Eval = cellato::evaluators::bit_planes::evaluator<word, typename Config::state_dictionary, typename Config::algorithm>;
    rule_kernels k{};
    k.name = Config::name;
    k.planes = Config::state_dictionary::needed_bits;
    k.nstates = Config::state_dictionary::number_of_values;
    k.traits = cellato::evaluators::bit_planes::rule_traits<typename Config::algorithm>::value;
    for (int i = 0; i < k.nstates && i < 32; ++i)
        k.state_values[i] = static_cast<int>(Config::state_dictionary::index_to_state(i));
    fill_layout<Eval, LINEAR, MaxLinearT>(k);
    fill_layout<Eval, TILED, MaxTiledT>(k);
    fill_rows<Eval, RowsFromT, MaxLinearT>(k);
    k.bit_array[0] = reinterpret_cast<const void*>(
        &bit_array_kernel<typename Config::state_dictionary, typename Config::algorithm, std::uint32_t>);
    k.bit_array[1] = reinterpret_cast<const void*>(
        &bit_array_kernel<typename Config::state_dictionary, typename Config::algorithm, std::uint64_t>);
    for (int t = 0; t <= MAX_T; ++t) k.batch[t] = nullptr;
    [&]<int... Ts>(std::integer_sequence<int, Ts...>) {
        ((k.batch[Ts + 1] = reinterpret_cast<const void*>(&batch_kernel<Eval, LINEAR, Ts + 1>)), ...);
    }(std::make_integer_sequence<int, MaxLinearT>{});
    return k;
}

// one per automaton, defined in rule_<id>.hip
const rule_kernels& kernels_of_rule(int id);
const rule_kernels& kernels_rule_0();
const rule_kernels& kernels_rule_1();
const rule_kernels& kernels_rule_2();
const rule_kernels& kernels_rule_3();
const rule_kernels& kernels_rule_4();
const rule_kernels& kernels_rule_5();
const rule_kernels& kernels_rule_6();
const rule_kernels& kernels_rule_7();
const rule_kernels& kernels_rule_8();
const rule_kernels& kernels_rule_9();

} // namespace cellato::hip

#endif
