// The ten automata, in a fixed order that defines the numeric rule ids of the C ABI
// (include/cellato_hip.h). `for_each_automaton` visits (id, config) pairs at compile time.
#ifndef CELLATO_AUTOMATA_ALL_HPP
#define CELLATO_AUTOMATA_ALL_HPP
#include "game_of_life.hpp"
#include "wire.hpp"
#include "fire.hpp"
#include "excitable.hpp"
#include "brian.hpp"
#include "maze.hpp"
#include "traffic.hpp"
#include "fluid.hpp"
#include "critters.hpp"
#include "cyclic.hpp"
#include <utility>

namespace cellato::automata {

enum rule_id : int {
    GAME_OF_LIFE = 0, WIRE = 1, FOREST_FIRE = 2, EXCITABLE = 3, BRIAN = 4,
    MAZE = 5, TRAFFIC = 6, FLUID = 7, CRITTERS = 8, CYCLIC = 9, RULE_COUNT = 10
};

template <int id> struct by_id;
template <> struct by_id<GAME_OF_LIFE> { using config = ::game_of_life::config; };
template <> struct by_id<WIRE>         { using config = ::wire::config; };
template <> struct by_id<FOREST_FIRE>  { using config = ::fire::config; };
template <> struct by_id<EXCITABLE>    { using config = ::excitable::config; };
template <> struct by_id<BRIAN>        { using config = ::brian::config; };
template <> struct by_id<MAZE>         { using config = ::maze::config; };
template <> struct by_id<TRAFFIC>      { using config = ::traffic::config; };
template <> struct by_id<FLUID>        { using config = ::fluid::config; };
template <> struct by_id<CRITTERS>     { using config = ::critters::config; };
template <> struct by_id<CYCLIC>       { using config = ::cyclic::config; };

template <class F, int... ids>
constexpr void for_each_id(F&& f, std::integer_sequence<int, ids...>) { (f.template operator()<ids>(), ...); }

} // namespace cellato::automata
#endif
