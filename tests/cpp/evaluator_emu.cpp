// CPU run of the product's bit-plane evaluator (include/cellato/evaluators/bit_planes.hpp)
// through a window that reads a whole grid from host memory: the same evaluator code the
// gfx950 kernels inline, checked against the oracle's standard evaluator for every rule,
// both layouts and both time-step parities. No GPU needed.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "cellato/automata/all.hpp"
#include "cellato/evaluators/bit_planes.hpp"
#include "layouts.hpp"
#include "standard.hpp"

// One-plane rules made of the cell and one Moore count (Game of Life, Maze) take the
// evaluator's table path when the window has row sums (the row-sum cases below run it); the
// other rules keep the node-by-node evaluation.
namespace bp = cellato::evaluators::bit_planes;
template <int id>
constexpr bool table_path = cellato::automata::by_id<id>::config::state_dictionary::needed_bits == 1 &&
                            bp::cc_rule<typename cellato::automata::by_id<id>::config::state_dictionary,
                                        typename cellato::automata::by_id<id>::config::algorithm>::ok;
static_assert(table_path<cellato::automata::GAME_OF_LIFE> && table_path<cellato::automata::MAZE>);
static_assert(!table_path<cellato::automata::WIRE> && !table_path<cellato::automata::CRITTERS> &&
              !table_path<cellato::automata::BRIAN> && !table_path<cellato::automata::CYCLIC>);
// Rules of several planes whose neighbourhood enters only through predicates of neighbour
// counts, and that compare the cell with two states or more, take the predicate-table path
// (every window); the others do not.
template <int id>
constexpr bool predicate_path = cellato::automata::by_id<id>::config::state_dictionary::needed_bits > 1 &&
    bp::pt_rule<typename cellato::automata::by_id<id>::config::state_dictionary,
                typename cellato::automata::by_id<id>::config::algorithm,
                cellato::automata::by_id<id>::config::state_dictionary::needed_bits>::ok;
static_assert(predicate_path<cellato::automata::WIRE> && predicate_path<cellato::automata::EXCITABLE> &&
              predicate_path<cellato::automata::FOREST_FIRE> && predicate_path<cellato::automata::BRIAN>);
static_assert(!predicate_path<cellato::automata::TRAFFIC> && !predicate_path<cellato::automata::FLUID> &&
              !predicate_path<cellato::automata::CYCLIC> && !predicate_path<cellato::automata::GAME_OF_LIFE>);
// WireWorld reads its head count only at conductors, Greenberg-Hastings its excited count only
// at resting cells and Brian's Brain its count of firing cells only at off cells: their
// predicates count through windows that include the cell itself
template <int id, int i>
constexpr bool self_free_pred = [] {
    using C = typename cellato::automata::by_id<id>::config;
    using D = typename C::state_dictionary;
    using R = bp::pt_rule<D, typename C::algorithm, D::needed_bits>;
    using Pr = typename bp::tl_at<typename R::preds, i>::type;
    return R::template self_free<i>(typename bp::pt_counted<D, Pr>::type{});
}();
static_assert(self_free_pred<cellato::automata::WIRE, 0> && self_free_pred<cellato::automata::EXCITABLE, 0> &&
              self_free_pred<cellato::automata::BRIAN, 0>);
// Greenberg-Hastings reads its excited count only as "some neighbour is excited": its row sums
// may be kept as 3-cell ORs; WireWorld and Brian's Brain compare their counts with 1, 2
template <int id, long long K>
constexpr bool or_sums = bp::evaluator<std::uint64_t, typename cellato::automata::by_id<id>::config::state_dictionary,
                                       typename cellato::automata::by_id<id>::config::algorithm>::template row_sum_or_only<K>;
static_assert(or_sums<cellato::automata::EXCITABLE, 1> && !or_sums<cellato::automata::WIRE, 1> &&
              !or_sums<cellato::automata::BRIAN, 2> && !or_sums<cellato::automata::GAME_OF_LIFE, 1>);
// forest fire reads its burning von Neumann neighbours only at trees and ash: its fire mask is
// kept per row as (3-cell OR, mask); the windows of this test keep none, so the masks path runs
// Critters takes the Margolus table path; its table over (cell, vertical, horizontal, diagonal
// partner) bits, written out from the rule by hand: count 3 -> dead if the diagonal (180-degree)
// partner is alive else alive, count 2 -> the cell, otherwise the cell toggled
template <int id>
using mg_of = bp::mg_rule<typename cellato::automata::by_id<id>::config::state_dictionary,
                          typename cellato::automata::by_id<id>::config::algorithm>;
static_assert(mg_of<cellato::automata::CRITTERS>::ok && mg_of<cellato::automata::CRITTERS>::table() == 0x3BD);
static_assert(!mg_of<cellato::automata::GAME_OF_LIFE>::ok && !mg_of<cellato::automata::MAZE>::ok &&
              !mg_of<cellato::automata::TRAFFIC>::ok);
template <int id>
using vn_sums = typename bp::evaluator<std::uint64_t, typename cellato::automata::by_id<id>::config::state_dictionary,
                                       typename cellato::automata::by_id<id>::config::algorithm>::vn_sum_targets;
static_assert(std::is_same_v<vn_sums<cellato::automata::FOREST_FIRE>, bp::klist<2>> &&
              std::is_same_v<vn_sums<cellato::automata::WIRE>, bp::klist<>>);

using namespace cellato::evaluators::bit_planes;
using u64 = std::uint64_t;

template <int P, bool Tiled>
struct host_window {
    const u64* planes;    // P planes of wx * h words
    int wx, h;            // words per row, rows of words
    int x, y;             // word position
    int t;
    int yglobal;          // cell row of the word (linear layout)

    u64 at(int p, int X, int Y) const {
        X = (X % wx + wx) % wx;
        Y = (Y % h + h) % h;
        return planes[static_cast<size_t>(p) * wx * h + static_cast<size_t>(Y) * wx + X];
    }
    u64 yodd() const { return Tiled ? 0xFF00FF00FF00FF00ULL : ((yglobal & 1) ? ~0ULL : 0ULL); }

    // word of plane p at word column X shifted by dy cell rows
    template <int dy>
    u64 vrow(int p, int X) const {
        if constexpr (!Tiled) return at(p, X, y + dy);
        else if constexpr (dy == 0) return at(p, X, y);
        else if constexpr (dy < 0) return (at(p, X, y) << 8) | (at(p, X, y - 1) >> 56);
        else return (at(p, X, y) >> 8) | (at(p, X, y + 1) << 56);
    }
    template <int dx>
    static u64 hsh(u64 own, u64 side) {
        if constexpr (dx == 0) return own;
        if constexpr (!Tiled) {
            if constexpr (dx < 0) return (own << 1) | (side >> 63);
            else return (own >> 1) | (side << 63);
        } else {
            constexpr u64 C0 = 0x0101010101010101ULL, C7 = 0x8080808080808080ULL;
            if constexpr (dx < 0) return ((own << 1) & ~C0) | ((side >> 7) & C0);
            else return ((own >> 1) & ~C7) | ((side << 7) & C7);
        }
    }
    template <int dx, int dy>
    vint<P> nb() const {
        vint<P> r;
        for (int p = 0; p < P; ++p) r.p[p] = hsh<dx>(vrow<dy>(p, x), vrow<dy>(p, x + dx));
        return r;
    }
    template <int dx, int dy, class F>
    u64 nbf(F f) const {
        vint<P> o, s;
        for (int p = 0; p < P; ++p) { o.p[p] = vrow<dy>(p, x); s.p[p] = vrow<dy>(p, x + dx); }
        return hsh<dx>(f(o), f(s));
    }
};

// The window of a streaming caller that keeps per-row 3-cell sums of the counted states (the
// kernels' ring slots do, in both layouts): sums formed here from the grid, row by row.
template <int P, class Targets, bool Tiled>
struct host_window_rows : host_window<P, Tiled> {
    static constexpr bool row_sums = true;
    template <long long K, int dy>
    void rsum(u64& s, u64& c) const {
        static_assert(kindex<K, Targets>::value >= 0, "not a counted state");
        const is_index<K> f{};
        fa(this->template nbf<-1, dy>(f), this->template nbf<0, dy>(f), this->template nbf<1, dy>(f), s, c);
    }
};

template <class Cfg, bool Tiled, bool Rows = false>
bool check(int W, int H, int t0, int steps) {
    using dict = typename Cfg::state_dictionary;
    constexpr int P = dict::needed_bits;
    using E = evaluator<u64, dict, typename Cfg::algorithm>;
    std::mt19937 rng(7 + t0);
    std::vector<int32_t> vals(static_cast<size_t>(W) * H), idx(vals.size());
    for (size_t i = 0; i < vals.size(); ++i) {
        idx[i] = static_cast<int32_t>(rng() % dict::number_of_values);
        vals[i] = static_cast<int32_t>(dict::index_to_state(idx[i]));
    }
    const int wx = Tiled ? W / 8 : W / 64, h = Tiled ? H / 8 : H;
    std::vector<u64> a(static_cast<size_t>(P) * wx * h), b(a.size());
    if (Tiled) oracle::layouts::pack_tiled(idx.data(), W, H, P, a.data());
    else oracle::layouts::pack_linear(idx.data(), W, H, P, a.data());
    for (int k = 0; k < steps; ++k) {
        for (int Y = 0; Y < h; ++Y)
            for (int X = 0; X < wx; ++X) {
                host_window<P, Tiled> w{a.data(), wx, h, X, Y, t0 + k, Y};
                vint<P> r;
                if constexpr (Rows) r = E::evaluate(host_window_rows<P, typename E::row_sum_targets, Tiled>{w});
                else r = E::evaluate(w);
                for (int p = 0; p < P; ++p) b[static_cast<size_t>(p) * wx * h + static_cast<size_t>(Y) * wx + X] = r.p[p];
            }
        a.swap(b);
    }
    oracle::standard::run<Cfg>(vals.data(), W, H, t0, steps, 1);
    std::vector<int32_t> got(vals.size());
    if (Tiled) oracle::layouts::unpack_tiled(a.data(), W, H, P, got.data());
    else oracle::layouts::unpack_linear(a.data(), W, H, P, got.data());
    for (size_t i = 0; i < got.size(); ++i)
        if (static_cast<int32_t>(dict::index_to_state(got[i])) != vals[i]) {
            std::printf("%s %s%s t0=%d: first mismatch at cell %zu\n", Cfg::name, Tiled ? "tiled" : "linear",
                        Rows ? " (row sums)" : "", t0, i);
            return false;
        }
    return true;
}

int main() {
    bool ok = true;
    int n = 0;
    cellato::automata::for_each_id([&]<int id>() {
        using Cfg = typename cellato::automata::by_id<id>::config;
        for (int t0 : {0, 1}) {
            ok &= check<Cfg, false>(192, 22, t0, 5);
            ok &= check<Cfg, true>(48, 24, t0, 5);
            ok &= check<Cfg, false, true>(192, 22, t0, 5);
            ok &= check<Cfg, true, true>(48, 24, t0, 5);
            n += 4;
        }
    }, std::make_integer_sequence<int, cellato::automata::RULE_COUNT>{});
    std::printf("%d cases: %s\n", n, ok ? "ALL OK" : "FAILED");
    return ok ? 0 : 1;
}
