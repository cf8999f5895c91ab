// Seeded initial states of the reference's experiments, so that a run started from the
// command line reproduces the reference's checksums: categorical i.i.d. draws
// (include/memory/grid_utils.hpp:15-76, with the tables of src/<automaton>/data_init.hpp) and
// WireWorld's circuit generator (src/wire/data_init.hpp:11-216). The draws go through
// libstdc++'s std::mt19937, uniform distributions and std::shuffle in the reference's order;
// that sequence of draws is what has to match, the code around it is this project's.
// Automata whose reference generator ignores --seed use the reference default 12345.
#ifndef CELLATO_EXPERIMENTS_INITIAL_STATE_HPP
#define CELLATO_EXPERIMENTS_INITIAL_STATE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cellato::experiments {

struct weighted_state {
    std::int32_t value;   // enum value of the state
    double p;
};

// Per cell, row-major, one uniform double r in [0, 1): the first state whose running sum of
// probabilities is >= r. A draw above the last sum (rounding) leaves the cell at 0.
inline void draw_states(std::vector<std::int32_t>& cells, const std::vector<weighted_state>& table, unsigned seed) {
    std::vector<double> upto(table.size());
    double run = 0.0;
    for (std::size_t k = 0; k < table.size(); ++k) upto[k] = run += table[k].p;
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> unit(0.0, 1.0);
    for (auto& c : cells) {
        const double r = unit(gen);
        const auto it = std::find_if(upto.begin(), upto.end(), [r](double u) { return r <= u; });
        if (it != upto.end()) c = table[static_cast<std::size_t>(it - upto.begin())].value;
    }
}

// WireWorld circuits: concentric 2:1 rectangles every 4 cells around the centre, 10 random
// straight segments, 5 random circles, 8 crossings, then on every 20th conductor cell (in a
// shuffled order) an electron head with its tail on a random conductor neighbour.
class wire_canvas {
public:
    static constexpr std::int32_t EMPTY = 0, HEAD = 1, TAIL = 2, CONDUCTOR = 3;

    wire_canvas(std::vector<std::int32_t>& cells, int w, int h) : _c(cells), _w(w), _h(h) {}

    void draw(unsigned seed) {
        std::fill(_c.begin(), _c.end(), EMPTY);
        std::mt19937 gen;
        gen.seed(seed);
        rectangles();
        std::uniform_int_distribution<int> pick_x(0, _w - 1), pick_y(0, _h - 1), pick_len(10, 30), pick_r(5, 15);
        for (int i = 0; i < 10; ++i) {
            const int x0 = pick_x(gen), y0 = pick_y(gen);
            int dx = static_cast<int>(gen() % 3) - 1;
            const int dy = static_cast<int>(gen() % 3) - 1;
            if (dx == 0 && dy == 0) dx = 1;
            const int len = pick_len(gen);
            for (int j = 0; j < len; ++j) conduct_if_inside(x0 + j * dx, y0 + j * dy);
        }
        for (int i = 0; i < 5; ++i) {
            const int ox = pick_x(gen), oy = pick_y(gen), r = pick_r(gen);
            for (int deg = 0; deg < 360; deg += 5) {
                const double a = deg * M_PI / 180.0;
                conduct_if_inside(ox + static_cast<int>(r * std::cos(a)), oy + static_cast<int>(r * std::sin(a)));
            }
        }
        for (int i = 0; i < 8; ++i) {
            const int x = pick_x(gen), y = pick_y(gen);
            if (x <= 5 || x >= _w - 5 || y <= 5 || y >= _h - 5) continue;
            for (int d = -4; d <= 4; ++d) conduct(x + d, y);
            for (int d = -4; d <= 4; ++d)
                if (d) conduct(x, y + d);
        }
        electrons(gen);
    }

private:
    bool inside(int x, int y) const { return x >= 0 && x < _w && y >= 0 && y < _h; }
    std::int32_t& at(int x, int y) { return _c[static_cast<std::size_t>(y) * _w + x]; }
    void conduct(int x, int y) {
        at(x, y) = CONDUCTOR;
        _laid.emplace_back(x, y);
    }
    void conduct_if_inside(int x, int y) {
        if (inside(x, y)) conduct(x, y);
    }

    void rectangles() {
        const int cx = _w / 2, cy = _h / 2;
        for (int hw = 4; hw <= std::max(cx, cy); hw += 4) {
            const int hh = hw / 2;
            if (cx + hw >= _w || cx - hw < 0 || cy + hh >= _h || cy - hh < 0) return;
            for (int x = cx - hw; x <= cx + hw; ++x) conduct(x, cy - hh);
            for (int x = cx - hw; x <= cx + hw; ++x) conduct(x, cy + hh);
            for (int y = cy - hh + 1; y < cy + hh; ++y) conduct(cx - hw, y);
            for (int y = cy - hh + 1; y < cy + hh; ++y) conduct(cx + hw, y);
        }
    }

    void electrons(std::mt19937& gen) {
        if (_laid.empty()) return;
        std::shuffle(_laid.begin(), _laid.end(), gen);
        for (std::size_t i = 0; i < _laid.size(); i += 20) {
            const auto [x, y] = _laid[i];
            at(x, y) = HEAD;
            std::vector<std::pair<int, int>> around{{0, 1}, {1, 0}, {0, -1}, {-1, 0}};
            std::shuffle(around.begin(), around.end(), gen);
            bool tailed = false;
            for (const auto& [dx, dy] : around) {
                if (!inside(x + dx, y + dy) || at(x + dx, y + dy) != CONDUCTOR) continue;
                at(x + dx, y + dy) = TAIL;
                tailed = true;
                break;
            }
            if (!tailed) at(x, y) = CONDUCTOR;
        }
    }

    std::vector<std::int32_t>& _c;
    int _w, _h;
    std::vector<std::pair<int, int>> _laid;   // conductor cells in the order they were laid
};

// Enum values of the reference's initial state of `automaton` (its command-line name).
inline std::vector<std::int32_t> initial_state(const std::string& automaton, int width, int height, int seed) {
    std::vector<std::int32_t> cells(static_cast<std::size_t>(width) * static_cast<std::size_t>(height), 0);
    const unsigned own = static_cast<unsigned>(seed), fixed = 12345u;
    const auto uniform = [](int n) {
        std::vector<weighted_state> t;
        for (int v = 0; v < n; ++v) t.push_back({v, 1.0 / n});
        return t;
    };
    if (automaton == "game-of-life") draw_states(cells, {{1, 0.2}, {0, 1.0 - 0.2}}, own);          // alive, dead
    else if (automaton == "brian") draw_states(cells, {{2, 0.2}, {0, 1.0 - 0.2}}, own);            // alive, dead
    else if (automaton == "excitable")
        draw_states(cells, {{0, .70}, {1, .15}, {2, .05}, {3, .03}, {4, .03}, {5, .02}, {6, .01}, {7, .01}}, own);
    else if (automaton == "forest-fire") draw_states(cells, {{0, .20}, {1, .79}, {3, .01}, {2, .00}}, fixed);  // enum ash 2, fire 3
    else if (automaton == "maze") draw_states(cells, {{0, .80}, {1, .20}}, fixed);
    else if (automaton == "traffic") draw_states(cells, {{0, .50}, {1, .25}, {2, .25}}, fixed);
    else if (automaton == "critters") draw_states(cells, {{0, .50}, {1, .50}}, fixed);
    else if (automaton == "fluid") draw_states(cells, uniform(16), fixed);
    else if (automaton == "cyclic") draw_states(cells, uniform(32), fixed);
    else if (automaton == "wire") wire_canvas(cells, width, height).draw(own);
    else throw std::invalid_argument("unknown automaton " + automaton);
    return cells;
}

} // namespace cellato::experiments

#endif
