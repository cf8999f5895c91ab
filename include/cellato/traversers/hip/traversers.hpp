// The HIP traversers as C++ templates with the reference's interface
// (include/traversers/cuda/simple.hpp:25-106, temporal.hpp:24-150):
//
//   simple::traverser<evaluator, grid> t;                          (simple.hpp:25-28)
//   temporal::traverser<evaluator, grid, average_halo_radius> t;   (temporal.hpp:24-27)
//   t.init(grid, params);  t.run(steps[, callback]);  grid result = t.fetch_result();
//
// so a test suite of the reference (include/experiments/test_suites.hpp:116-225) binds them
// with its namespace alias changed from cellato::traversers::cuda to cellato::traversers::hip.
// The evaluator is bit_planes::evaluator for the linear grid, tiled_bit_planes::evaluator for
// the tiled one, bit_array::evaluator<grid, algorithm> for the bit_array grid (checked at
// compile time: a mismatch would compute a different automaton).
//
// Each call goes through the C ABI of include/cellato_hip.h (libcellato_hip.so), which holds
// the gfx950 kernels of the automata registered in include/cellato/automata/all.hpp; the
// evaluator type selects the rule. Errors surface as the reference's exception types:
// std::runtime_error for invalid parameters, cellato::memory::grids::hip_error for HIP.
#ifndef CELLATO_TRAVERSERS_HIP_TRAVERSERS_HPP
#define CELLATO_TRAVERSERS_HIP_TRAVERSERS_HPP

#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "../../../cellato_hip.h"
#include "../../automata/all.hpp"
#include "../../evaluators/bit_array.hpp"
#include "../../evaluators/bit_planes.hpp"
#include "../../evaluators/tiled_bit_planes.hpp"
#include "../../experiments/run_params.hpp"
#include "../../memory/bit_array_grid.hpp"
#include "../../memory/bit_planes_grid.hpp"

namespace cellato::traversers::hip {

// rule id (C ABI) of an evaluator's expression; fails to compile for an unregistered rule
template <class Expression>
constexpr int rule_id_of() {
    int id = -1;
    [&]<int... ids>(std::integer_sequence<int, ids...>) {
        ((id = std::is_same_v<Expression, typename cellato::automata::by_id<ids>::config::algorithm> ? ids : id), ...);
    }(std::make_integer_sequence<int, cellato::automata::RULE_COUNT>{});
    return id;
}

inline void check(int rc) {
    if (rc == CELLATO_E_HIP) throw cellato::memory::grids::hip_error(cellato_hip_last_error());
    if (rc < 0) throw std::runtime_error(cellato_hip_last_error());
}

// average_halo_radius: cells of halo one generation needs (the reference sizes its temporal
// tiles with it, temporal.hpp:47); the kernels here carry a one-cell halo per generation,
// which covers every rule whose neighbour reads stay within distance 1.
template <typename evaluator_type, typename grid_type, bool Temporal, double average_halo_radius = 1.0>
class traverser_base {
    using evaluator_t = evaluator_type;
    using grid_t = grid_type;
    static constexpr int P = grid_t::needed_bits;
    static constexpr int rule = rule_id_of<typename evaluator_t::expression>();
    static_assert(rule >= 0, "this rule is not compiled into libcellato_hip.so (include/cellato/automata/all.hpp)");
    static_assert(evaluator_t::layout == grid_t::layout,
                  "evaluator and grid layouts differ: bit_planes::evaluator goes with bit_planes::grid, "
                  "tiled_bit_planes::evaluator with tiled_bit_planes::grid");
    static_assert(average_halo_radius > 0.0 && average_halo_radius <= 1.0,
                  "the HIP kernels carry a halo of one cell per generation");
    using device_grid_t = std::invoke_result_t<decltype(&grid_t::to_cuda), grid_t>;

public:
    static constexpr bool is_CUDA = true;   // the reference's name for "runs on the GPU"

    struct no_callback {};

    void init(grid_t grid, const cellato::run::run_params& params) {
        if (params.word_size != 0 && params.word_size != 64) throw std::runtime_error("the HIP traversers use 64-bit words");
        if constexpr (Temporal) {
            if (params.temporal_steps < 1 || params.temporal_steps > cellato_max_temporal_steps())
                throw std::runtime_error("Invalid temporal steps for the temporal traverser.");
        }
        _params = params;
        _cur = grid.to_cuda();
        _other = grid.to_cuda();
        _t = 0;
        _ready = true;
    }

    template <typename callback = no_callback>
    void run(int steps, callback&& cb = no_callback{}) {
        if (!_ready) throw std::runtime_error("run() before init()");
        if constexpr (std::is_same_v<std::decay_t<callback>, no_callback>) {
            advance(steps);
        } else {
            auto host = _cur.to_cpu();
            cb(0, host);
            for (int k = 0; k < steps; ++k) {
                advance(1);
                auto h = _cur.to_cpu();
                cb(k + 1, h);
            }
        }
    }

    grid_t fetch_result() {
        cellato::memory::grids::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        grid_t r = _cur.to_cpu();
        _cur.free_cuda_memory();
        _other.free_cuda_memory();
        _ready = false;
        return r;
    }

private:
    void advance(int steps) {
        auto a = _cur.data();
        auto b = _other.data();
        const int w = static_cast<int>(_cur.x_size_original()), h = static_cast<int>(_cur.y_size_original());
        int in_b = 0;
        if constexpr (Temporal)
            check(cellato_hip_temporal(rule, grid_t::layout, a.data(), b.data(), w, h, _t, steps, _params.temporal_steps,
                                       _params.chunk_rows, nullptr, &in_b));
        else
            check(cellato_hip_simple(rule, grid_t::layout, a.data(), b.data(), w, h, _t, steps, _params.chunk_rows,
                                     nullptr, &in_b));
        if (in_b) std::swap(_cur, _other);
        _t += steps;
    }

    cellato::run::run_params _params;
    device_grid_t _cur, _other;
    int _t = 0;
    bool _ready = false;
};

// The simple traverser over the bit_array layout (the reference binds it as
// traversers::simple::traverser<evaluators::bit_array::evaluator<grid_t, algorithm_t>, grid_t>,
// include/experiments/test_suites.hpp:86-113): one device array of words per buffer, one
// generation per launch, through cellato_hip_bit_array.
template <typename evaluator_type, typename grid_type>
class bit_array_traverser {
    using evaluator_t = evaluator_type;
    using grid_t = grid_type;
    static constexpr int rule = rule_id_of<typename evaluator_t::expression>();
    static_assert(rule >= 0, "this rule is not compiled into libcellato_hip.so (include/cellato/automata/all.hpp)");
    static_assert(evaluator_t::layout == grid_t::layout,
                  "evaluators::bit_array::evaluator goes with grids::bit_array::grid");
    static constexpr int word_bits = static_cast<int>(sizeof(typename grid_t::store_type) * 8);
    using device_grid_t = std::invoke_result_t<decltype(&grid_t::to_cuda), grid_t>;

public:
    static constexpr bool is_CUDA = true;
    struct no_callback {};

    void init(grid_t grid, const cellato::run::run_params& params) {
        if (params.word_size != 0 && params.word_size != word_bits)
            throw std::runtime_error("word_size does not match the grid's word type");
        _cur = grid.to_cuda();
        _other = grid.to_cuda();
        _t = 0;
        _ready = true;
    }

    template <typename callback = no_callback>
    void run(int steps, callback&& cb = no_callback{}) {
        if (!_ready) throw std::runtime_error("run() before init()");
        if constexpr (std::is_same_v<std::decay_t<callback>, no_callback>) {
            advance(steps);
        } else {
            auto host = _cur.to_cpu();
            cb(0, host);
            for (int k = 0; k < steps; ++k) {
                advance(1);
                auto h = _cur.to_cpu();
                cb(k + 1, h);
            }
        }
    }

    grid_t fetch_result() {
        cellato::memory::grids::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        grid_t r = _cur.to_cpu();
        _cur.free_cuda_memory();
        _other.free_cuda_memory();
        _ready = false;
        return r;
    }

private:
    void advance(int steps) {
        int in_b = 0;
        check(cellato_hip_bit_array(rule, word_bits, _cur.raw_data(), _other.raw_data(),
                                    static_cast<int>(_cur.x_size_logical()), static_cast<int>(_cur.y_size_logical()), _t,
                                    steps, nullptr, &in_b));
        if (in_b) std::swap(_cur, _other);
        _t += steps;
    }

    device_grid_t _cur, _other;
    int _t = 0;
    bool _ready = false;
};

template <typename grid_type, typename = void>
struct is_bit_array_grid : std::false_type {};
template <typename grid_type>
struct is_bit_array_grid<grid_type, std::void_t<decltype(grid_type::is_bit_array)>> : std::bool_constant<grid_type::is_bit_array> {};

namespace simple {
template <typename evaluator_type, typename grid_type>
using traverser = std::conditional_t<is_bit_array_grid<grid_type>::value, bit_array_traverser<evaluator_type, grid_type>,
                                     traverser_base<evaluator_type, grid_type, false>>;
}
namespace temporal {
template <typename evaluator_type, typename grid_type, double average_halo_radius = 1.0>
using traverser = traverser_base<evaluator_type, grid_type, true, average_halo_radius>;
}

} // namespace cellato::traversers::hip

#endif
