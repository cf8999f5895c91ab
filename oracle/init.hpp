// ORACLE — test infrastructure only (see oracle/standard.hpp for the rules of use).
//
// Restatement of the reference's seeded initial-state generators, needed to reproduce the
// reference's published checksums. The random draws go through libstdc++'s std::mt19937,
// uniform_real_distribution<double>, uniform_int_distribution<int> and std::shuffle exactly
// as the reference calls them, so the sequence of draws (not the code) is what must match.
#ifndef CELLATO_ORACLE_INIT_HPP
#define CELLATO_ORACLE_INIT_HPP

#include <cmath>
#include <cstdint>
#include <random>
#include <utility>
#include <vector>
#include <algorithm>

namespace oracle::init {

// include/memory/grid_utils.hpp:15-57: per cell one uniform double r in [0,1); the cell gets
// the first state whose cumulative probability is >= r (cells past the last threshold keep 0).
inline void categorical(int32_t* cells, size_t n, const std::vector<std::pair<int32_t, double>>& probs,
                        unsigned seed) {
    std::vector<double> cum;
    double acc = 0.0;
    for (auto& pr : probs) { acc += pr.second; cum.push_back(acc); }
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (size_t i = 0; i < n; ++i) {
        const double r = u(rng);
        for (size_t k = 0; k < cum.size(); ++k)
            if (r <= cum[k]) { cells[i] = probs[k].first; break; }
    }
}

// src/wire/data_init.hpp:11-216: concentric 2:1 rectangles every 4 cells around the centre,
// 10 random line segments, 5 random circles, 8 crossover bridges, then an electron
// (head + tail on a random free conductor neighbour) on every 20th shuffled conductor cell.
inline void wire(int32_t* cells, int W, int H, unsigned seed) {
    enum : int32_t { EMPTY = 0, HEAD = 1, TAIL = 2, COND = 3 };
    std::fill(cells, cells + static_cast<size_t>(W) * H, EMPTY);
    std::mt19937 gen;
    gen.seed(seed);
    std::vector<std::pair<int, int>> wires;   // conductor cells in placement order
    auto put = [&](int x, int y) {
        cells[static_cast<size_t>(y) * W + x] = COND;
        wires.push_back({x, y});
    };

    const int cx = W / 2, cy = H / 2;
    for (int half_w = 4; half_w <= std::max(cx, cy); half_w += 4) {
        const int half_h = half_w / 2;
        if (cx + half_w >= W || cx - half_w < 0 || cy + half_h >= H || cy - half_h < 0) break;
        for (int x = cx - half_w; x <= cx + half_w; ++x) put(x, cy - half_h);
        for (int x = cx - half_w; x <= cx + half_w; ++x) put(x, cy + half_h);
        for (int y = cy - half_h + 1; y < cy + half_h; ++y) put(cx - half_w, y);
        for (int y = cy - half_h + 1; y < cy + half_h; ++y) put(cx + half_w, y);
    }

    std::uniform_int_distribution<int> xd(0, W - 1), yd(0, H - 1), lend(10, 30), radd(5, 15);
    auto inside = [&](int x, int y) { return x >= 0 && x < W && y >= 0 && y < H; };

    for (int i = 0; i < 10; ++i) {
        const int x1 = xd(gen);
        const int y1 = yd(gen);
        int dx = static_cast<int>(gen() % 3) - 1;
        const int dy = static_cast<int>(gen() % 3) - 1;
        if (dx == 0 && dy == 0) dx = 1;
        const int len = lend(gen);
        for (int j = 0; j < len; ++j)
            if (inside(x1 + j * dx, y1 + j * dy)) put(x1 + j * dx, y1 + j * dy);
    }

    for (int i = 0; i < 5; ++i) {
        const int ox = xd(gen);
        const int oy = yd(gen);
        const int r = radd(gen);
        for (int deg = 0; deg < 360; deg += 5) {
            const double a = deg * M_PI / 180.0;
            const int x = ox + static_cast<int>(r * std::cos(a));
            const int y = oy + static_cast<int>(r * std::sin(a));
            if (inside(x, y)) put(x, y);
        }
    }

    for (int i = 0; i < 8; ++i) {
        const int x = xd(gen);
        const int y = yd(gen);
        if (x > 5 && x < W - 5 && y > 5 && y < H - 5) {
            for (int d = -4; d <= 4; ++d) put(x + d, y);
            for (int d = -4; d <= 4; ++d)
                if (d != 0) put(x, y + d);
        }
    }

    if (wires.empty()) return;
    std::shuffle(wires.begin(), wires.end(), gen);
    for (size_t i = 0; i < wires.size(); i += 20) {
        const auto [x, y] = wires[i];
        const size_t at = static_cast<size_t>(y) * W + x;
        cells[at] = HEAD;
        std::vector<std::pair<int, int>> dirs = {{0, 1}, {1, 0}, {0, -1}, {-1, 0}};
        std::shuffle(dirs.begin(), dirs.end(), gen);
        bool tail = false;
        for (auto [dx, dy] : dirs) {
            const int nx = x + dx, ny = y + dy;
            if (!inside(nx, ny)) continue;
            int32_t& c = cells[static_cast<size_t>(ny) * W + nx];
            if (c == COND) { c = TAIL; tail = true; break; }
        }
        if (!tail) cells[at] = COND;
    }
}

} // namespace oracle::init

#endif
