// ORACLE — test infrastructure only. liboracle.so is loaded by tests/, by
// __graft_entry__.smoke() and by bench.py's CPU-baseline leg, always as the checker (or the
// timed CPU baseline), never by the product library.
//
// Plain-C entry points over the CPU restatements in this directory:
//   standard.hpp   — standard evaluator + CPU simple traverser + 16-tile checksum
//   bit_planes.hpp — bit_planes evaluator on the CPU (linear layout, 64-bit words)
//   init.hpp       — the reference's seeded initial states
//   layouts.hpp    — linear / tiled bit-plane packing
// Rule ids are the order of include/cellato/automata/all.hpp.
#include <cstdint>
#include <cstring>
#include <vector>

#include "cellato/automata/all.hpp"
#include "standard.hpp"
#include "bit_planes.hpp"
#include "init.hpp"
#include "layouts.hpp"

namespace {

using namespace cellato::automata;

template <class F>
int dispatch(int rule, F&& f) {
    switch (rule) {
        case GAME_OF_LIFE: f.template operator()<GAME_OF_LIFE>(); return 0;
        case WIRE:         f.template operator()<WIRE>(); return 0;
        case FOREST_FIRE:  f.template operator()<FOREST_FIRE>(); return 0;
        case EXCITABLE:    f.template operator()<EXCITABLE>(); return 0;
        case BRIAN:        f.template operator()<BRIAN>(); return 0;
        case MAZE:         f.template operator()<MAZE>(); return 0;
        case TRAFFIC:      f.template operator()<TRAFFIC>(); return 0;
        case FLUID:        f.template operator()<FLUID>(); return 0;
        case CRITTERS:     f.template operator()<CRITTERS>(); return 0;
        case CYCLIC:       f.template operator()<CYCLIC>(); return 0;
        default: return -1;
    }
}

template <class Cfg>
std::vector<int32_t> enum_values() {
    using D = typename Cfg::state_dictionary;
    std::vector<int32_t> v;
    for (int i = 0; i < D::number_of_values; ++i) v.push_back(static_cast<int32_t>(D::index_to_state(i)));
    return v;
}

} // namespace

extern "C" {

int oracle_rule_count(void) { return RULE_COUNT; }

const char* oracle_rule_name(int rule) {
    const char* n = nullptr;
    dispatch(rule, [&]<int id>() { n = by_id<id>::config::name; });
    return n;
}

int oracle_rule_planes(int rule) {
    int p = -1;
    dispatch(rule, [&]<int id>() { p = by_id<id>::config::state_dictionary::needed_bits; });
    return p;
}

// enum value of each dictionary index; returns the number of states (<= 32).
int oracle_rule_states(int rule, int32_t* values) {
    int n = -1;
    dispatch(rule, [&]<int id>() {
        auto v = enum_values<typename by_id<id>::config>();
        n = static_cast<int>(v.size());
        std::memcpy(values, v.data(), v.size() * sizeof(int32_t));
    });
    return n;
}

// The reference's initial state for `rule` (enum values). Rules whose reference generator
// ignores the seed use 12345 (grid_utils.hpp:20), as the reference does.
int oracle_init_reference(int rule, int W, int H, int seed, int32_t* cells) {
    const size_t n = static_cast<size_t>(W) * H;
    std::memset(cells, 0, n * sizeof(int32_t));
    const unsigned s = static_cast<unsigned>(seed);
    using oracle::init::categorical;
    switch (rule) {
        case GAME_OF_LIFE: categorical(cells, n, {{1, 0.2}, {0, 1.0 - 0.2}}, s); return 0;       // gol data_init:17-23
        case BRIAN:        categorical(cells, n, {{2, 0.2}, {0, 1.0 - 0.2}}, s); return 0;       // brian data_init:17-23
        case EXCITABLE:                                                                          // excitable data_init:19-36
            categorical(cells, n, {{0, .70}, {1, .15}, {2, .05}, {3, .03}, {4, .03}, {5, .02}, {6, .01}, {7, .01}}, s);
            return 0;
        case FOREST_FIRE:  categorical(cells, n, {{0, .20}, {1, .79}, {3, .01}, {2, .00}}, 12345); return 0;  // fire data_init:16-27
        case MAZE:         categorical(cells, n, {{0, .80}, {1, .20}}, 12345); return 0;                      // maze data_init:16-27
        case TRAFFIC:      categorical(cells, n, {{0, .50}, {1, .25}, {2, .25}}, 12345); return 0;            // traffic data_init:16-27
        case CRITTERS:     categorical(cells, n, {{0, .50}, {1, .50}}, 12345); return 0;                      // critters data_init:16-25
        case FLUID: {                                                                                         // fluid data_init:16-39
            std::vector<std::pair<int32_t, double>> p;
            for (int v = 0; v < 16; ++v) p.push_back({v, 1.0 / 16.0});
            categorical(cells, n, p, 12345);
            return 0;
        }
        case CYCLIC: {                                                                                        // cyclic data_init:18-27
            std::vector<std::pair<int32_t, double>> p;
            for (int v = 0; v < 32; ++v) p.push_back({v, 1.0 / 32});
            categorical(cells, n, p, 12345);
            return 0;
        }
        case WIRE: oracle::init::wire(cells, W, H, s); return 0;
        default: return -1;
    }
}

// Standard evaluator, in place, cells = enum values.
int oracle_run_standard(int rule, int32_t* cells, int W, int H, int t0, int steps, int threads) {
    if (W <= 0 || H <= 0 || steps < 0) return -2;
    return dispatch(rule, [&]<int id>() {
        oracle::standard::run<typename by_id<id>::config>(cells, W, H, t0, steps, threads);
    });
}

// bit_planes evaluator on the CPU, in place, linear 64-bit planes back to back.
int oracle_run_bitplanes(int rule, uint64_t* planes, int W, int H, int t0, int steps, int threads) {
    if (W <= 0 || W % 64 || H <= 0 || steps < 0) return -2;
    return dispatch(rule, [&]<int id>() {
        oracle::bit_planes::run<typename by_id<id>::config, uint64_t>(planes, W / 64, H, t0, steps, threads);
    });
}

// Same evaluator over 8-bit words (the reference's unit tests use uint8_t words).
int oracle_run_bitplanes_u8(int rule, uint8_t* planes, int W, int H, int t0, int steps) {
    if (W <= 0 || W % 8 || H <= 0 || steps < 0) return -2;
    return dispatch(rule, [&]<int id>() {
        oracle::bit_planes::run<typename by_id<id>::config, uint8_t>(planes, W / 8, H, t0, steps, 1);
    });
}

void oracle_checksum(const int32_t* cells, int W, int H, uint64_t* out16) {
    oracle::standard::checksum(cells, W, H, out16);
}

void oracle_pack_linear(const int32_t* idx, int W, int H, int P, uint64_t* planes) { oracle::layouts::pack_linear(idx, W, H, P, planes); }
void oracle_unpack_linear(const uint64_t* planes, int W, int H, int P, int32_t* idx) { oracle::layouts::unpack_linear(planes, W, H, P, idx); }
void oracle_pack_tiled(const int32_t* idx, int W, int H, int P, uint64_t* planes) { oracle::layouts::pack_tiled(idx, W, H, P, planes); }
void oracle_unpack_tiled(const uint64_t* planes, int W, int H, int P, int32_t* idx) { oracle::layouts::unpack_tiled(planes, W, H, P, idx); }
void oracle_pack_bit_array(const int32_t* idx, int W, int H, int P, int word_bits, uint64_t* words) {
    oracle::layouts::pack_bit_array(idx, W, H, P, word_bits, words);
}
void oracle_unpack_bit_array(const uint64_t* words, int W, int H, int P, int word_bits, int32_t* idx) {
    oracle::layouts::unpack_bit_array(words, W, H, P, word_bits, idx);
}

} // extern "C"
