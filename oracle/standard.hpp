// ORACLE — test infrastructure only. Nothing in the product links or calls this file; only
// tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it (through
// liboracle.so), and only as the checker.
//
// CPU restatement of the reference's *standard* evaluator (include/evaluators/standard.hpp)
// and CPU simple traverser (include/traversers/cpu/simple.hpp:42-88): one cell per call,
// plain integer semantics, torus indexing of include/memory/interface.hpp:28-33.
// Cells hold the enum *values* of the rule's state type (as the reference standard grid does).
#ifndef CELLATO_ORACLE_STANDARD_HPP
#define CELLATO_ORACLE_STANDARD_HPP

#include <cstdint>
#include <thread>
#include <vector>
#include <algorithm>

#include "cellato/core/ast.hpp"

namespace oracle::standard {

using namespace cellato::ast;

// point_in_grid restated (interface.hpp:41-60) over an int32 cell array.
struct point {
    const int32_t* grid;
    int W, H;        // grid size in cells
    int x, y;        // cell position
    int t;           // time step
    int idx(int xx, int yy) const {                 // interface.hpp:28-33
        const int xr = (xx + W) % W;
        const int yr = (yy + H) % H;
        return yr * W + xr;
    }
};

template <class Cfg, class E>
struct eval;

// constant / state_constant: the literal itself (standard.hpp:31-43)
template <class Cfg, auto V>
struct eval<Cfg, constant<V>> {
    static auto at(const point&) { return V; }
};
template <class Cfg, auto V>
struct eval<Cfg, state_constant<V>> {
    static auto at(const point&) { return V; }
};

// neighbor_at: the torus-wrapped cell (standard.hpp:143-150)
template <class Cfg, int dx, int dy>
struct eval<Cfg, neighbor_at<dx, dy>> {
    using state_t = typename Cfg::cell_state;
    static state_t at(const point& s) {
        return static_cast<state_t>(s.grid[s.idx(s.x + dx, s.y + dy)]);
    }
};

// if_then_else: C ? A : B (standard.hpp:45-52)
template <class Cfg, class C, class A, class B>
struct eval<Cfg, if_then_else<C, A, B>> {
    static auto at(const point& s) {
        return eval<Cfg, C>::at(s) ? eval<Cfg, A>::at(s) : eval<Cfg, B>::at(s);
    }
};

#define ORACLE_BINARY(NODE, EXPR)                                                  \
    template <class Cfg, class L, class R>                                         \
    struct eval<Cfg, NODE<L, R>> {                                                 \
        static auto at(const point& s) {                                           \
            auto l = eval<Cfg, L>::at(s);                                          \
            auto r = eval<Cfg, R>::at(s);                                          \
            return EXPR;                                                           \
        }                                                                          \
    };

ORACLE_BINARY(bit_and_, l & r)        // standard.hpp:54-60
ORACLE_BINARY(plus, l + r)            // standard.hpp:62-68
ORACLE_BINARY(modulo, l % r)          // standard.hpp:70-76
ORACLE_BINARY(bit_or_, l | r)         // standard.hpp:78-84
ORACLE_BINARY(equals, l == r)         // standard.hpp:111-117
ORACLE_BINARY(greater_than, l > r)    // standard.hpp:119-125
ORACLE_BINARY(less_than, l < r)       // standard.hpp:127-133
ORACLE_BINARY(not_equals, l != r)     // standard.hpp:135-141
#undef ORACLE_BINARY

// logical and/or short-circuit exactly like && / || (standard.hpp:95-109)
template <class Cfg, class L, class R>
struct eval<Cfg, and_<L, R>> {
    static bool at(const point& s) { return eval<Cfg, L>::at(s) && eval<Cfg, R>::at(s); }
};
template <class Cfg, class L, class R>
struct eval<Cfg, or_<L, R>> {
    static bool at(const point& s) { return eval<Cfg, L>::at(s) || eval<Cfg, R>::at(s); }
};
template <class Cfg, class V>
struct eval<Cfg, not_<V>> {                                           // standard.hpp:87-92
    static bool at(const point& s) { return !eval<Cfg, V>::at(s); }
};

// has_bit_set (standard.hpp:152-158)
template <class Cfg, class V, int b>
struct eval<Cfg, has_bit_set<V, b>> {
    static bool at(const point& s) {
        auto v = eval<Cfg, V>::at(s);
        return (v & (decltype(v))(1 << b)) != 0;
    }
};

// alternate_algorithms: even step -> first rule (standard.hpp:160-169)
template <class Cfg, class Even, class Odd>
struct eval<Cfg, alternate_algorithms<Even, Odd>> {
    static auto at(const point& s) {
        if (s.t % 2 == 0) return eval<Cfg, Even>::at(s);
        return eval<Cfg, Odd>::at(s);
    }
};

// Moore count, neighbour order of standard.hpp:177-180
template <class Cfg, class C>
struct eval<Cfg, count_neighbors<C, moore_8_neighbors>> {
    static int at(const point& s) {
        static constexpr int dx[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
        static constexpr int dy[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
        using state_t = typename Cfg::cell_state;
        const auto target = eval<Cfg, C>::at(s);
        int n = 0;
        for (int i = 0; i < 8; ++i)
            n += static_cast<state_t>(s.grid[s.idx(s.x + dx[i], s.y + dy[i])]) == target;
        return n;
    }
};

// von Neumann count, order of standard.hpp:256-259
template <class Cfg, class C>
struct eval<Cfg, count_neighbors<C, von_neumann_4_neighbors>> {
    static int at(const point& s) {
        static constexpr int dx[4] = {0, 0, 1, -1};
        static constexpr int dy[4] = {1, -1, 0, 0};
        using state_t = typename Cfg::cell_state;
        const auto target = eval<Cfg, C>::at(s);
        int n = 0;
        for (int i = 0; i < 4; ++i)
            n += static_cast<state_t>(s.grid[s.idx(s.x + dx[i], s.y + dy[i])]) == target;
        return n;
    }
};

// Margolus block count (standard.hpp:185-216): the 2x2 block containing (x, y); on even
// steps blocks start at even (x, y), on odd steps at odd (x, y).
template <class Cfg, auto V>
struct eval<Cfg, count_neighbors<state_constant<V>, margolus_alternating_neighborhood>> {
    static int at(const point& s) {
        const int tp = s.t % 2, xp = s.x % 2, yp = s.y % 2;
        int xs[2], ys[2];
        if (tp == 0) {
            if (xp == 0) { xs[0] = 0;  xs[1] = 1; } else { xs[0] = -1; xs[1] = 0; }
            if (yp == 0) { ys[0] = 0;  ys[1] = 1; } else { ys[0] = -1; ys[1] = 0; }
        } else {
            if (xp == 0) { xs[0] = -1; xs[1] = 0; } else { xs[0] = 0;  xs[1] = 1; }
            if (yp == 0) { ys[0] = -1; ys[1] = 0; } else { ys[0] = 0;  ys[1] = 1; }
        }
        using state_t = typename Cfg::cell_state;
        int n = 0;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                n += static_cast<state_t>(s.grid[s.idx(s.x + xs[i], s.y + ys[j])]) == V;
        return n;
    }
};

// The diagonally opposite cell of the Margolus block (standard.hpp:218-248).
template <class Cfg>
struct eval<Cfg, margolus_180_neighbor> {
    static auto at(const point& s) {
        const int tp = s.t % 2, xp = s.x % 2, yp = s.y % 2;
        int dx, dy;
        if (tp == 0) { dx = xp == 0 ? 1 : -1;  dy = yp == 0 ? 1 : -1; }
        else         { dx = xp == 0 ? -1 : 1;  dy = yp == 0 ? -1 : 1; }
        using state_t = typename Cfg::cell_state;
        return static_cast<state_t>(s.grid[s.idx(s.x + dx, s.y + dy)]);
    }
};

// One generation over the whole torus (cpu/simple.hpp:60-85), rows [y0, y1).
template <class Cfg>
void step_rows(const int32_t* in, int32_t* out, int W, int H, int t, int y0, int y1) {
    using rule = eval<Cfg, typename Cfg::algorithm>;
    point s{in, W, H, 0, 0, t};
    for (int y = y0; y < y1; ++y) {
        s.y = y;
        for (int x = 0; x < W; ++x) {
            s.x = x;
            out[y * W + x] = static_cast<int32_t>(rule::at(s));
        }
    }
}

// Advance `cells` in place by `steps` generations starting at time step t0.
// threads == 1 is the reference's single-threaded loop; threads > 1 splits rows.
template <class Cfg>
void run(int32_t* cells, int W, int H, int t0, int steps, int threads) {
    std::vector<int32_t> tmp(static_cast<size_t>(W) * H);
    int32_t* cur = cells;
    int32_t* nxt = tmp.data();
    threads = std::max(1, std::min(threads, H));
    for (int k = 0; k < steps; ++k) {
        const int t = t0 + k;
        if (threads == 1) {
            step_rows<Cfg>(cur, nxt, W, H, t, 0, H);
        } else {
            std::vector<std::thread> pool;
            for (int i = 0; i < threads; ++i) {
                const int y0 = static_cast<int>(static_cast<long long>(H) * i / threads);
                const int y1 = static_cast<int>(static_cast<long long>(H) * (i + 1) / threads);
                pool.emplace_back(step_rows<Cfg>, cur, nxt, W, H, t, y0, y1);
            }
            for (auto& th : pool) th.join();
        }
        std::swap(cur, nxt);
    }
    if (cur != cells) std::copy(cur, cur + static_cast<size_t>(W) * H, cells);
}

// 16-tile checksum of enum values (standard_grid.hpp:179-213): 4x4 tiles of
// ceil(W/4) x ceil(H/4) cells, row-major tile order, uint64 sums.
inline void checksum(const int32_t* cells, int W, int H, uint64_t out[16]) {
    const int tw = (W + 3) / 4, th = (H + 3) / 4;
    for (int ty = 0; ty < 4; ++ty)
        for (int tx = 0; tx < 4; ++tx) {
            uint64_t sum = 0;
            const int x0 = tx * tw, y0 = ty * th;
            const int x1 = std::min(x0 + tw, W), y1 = std::min(y0 + th, H);
            for (int y = y0; y < y1; ++y)
                for (int x = x0; x < x1; ++x)
                    sum += static_cast<uint64_t>(static_cast<int64_t>(cells[static_cast<size_t>(y) * W + x]));
            out[ty * 4 + tx] = sum;
        }
}

} // namespace oracle::standard

#endif
