// ORACLE — test infrastructure only (see oracle/standard.hpp for the rules of use).
//
// Packing of standard grids (one dictionary index per cell) into bit planes, restating
//   linear: include/memory/bit_planes_grid.hpp:234-270 (word (X, y) holds cells
//           64X .. 64X+63 of row y, cell x at bit x % 64; plane b = bit b of the index), and
//   tiled:  include/memory/tiled_bit_planes_grid.hpp:243-288 (word (X, Y) holds the 8x8 block
//           at (8X, 8Y); cell (x, y) at bit (y % 8) * 8 + x % 8), and
//   bit_array: include/memory/bit_array_grid.hpp:293-312 (cell i of the row-major grid in word
//           i / cpw at bits [(i % cpw) * P, + P) as its index, cpw = word bits / P).
// Planes are stored back to back: plane b starts at b * words_per_plane.
#ifndef CELLATO_ORACLE_LAYOUTS_HPP
#define CELLATO_ORACLE_LAYOUTS_HPP

#include <cstdint>
#include <cstddef>

namespace oracle::layouts {

inline void pack_linear(const int32_t* idx, int W, int H, int P, uint64_t* planes) {
    const int Wx = W / 64;
    const size_t plane = static_cast<size_t>(Wx) * H;
    for (int b = 0; b < P; ++b)
        for (int y = 0; y < H; ++y)
            for (int X = 0; X < Wx; ++X) {
                uint64_t w = 0;
                for (int i = 0; i < 64; ++i)
                    w |= static_cast<uint64_t>((idx[static_cast<size_t>(y) * W + X * 64 + i] >> b) & 1) << i;
                planes[b * plane + static_cast<size_t>(y) * Wx + X] = w;
            }
}

inline void unpack_linear(const uint64_t* planes, int W, int H, int P, int32_t* idx) {
    const int Wx = W / 64;
    const size_t plane = static_cast<size_t>(Wx) * H;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int32_t v = 0;
            for (int b = 0; b < P; ++b)
                v |= static_cast<int32_t>((planes[b * plane + static_cast<size_t>(y) * Wx + x / 64] >> (x % 64)) & 1) << b;
            idx[static_cast<size_t>(y) * W + x] = v;
        }
}

inline void pack_tiled(const int32_t* idx, int W, int H, int P, uint64_t* planes) {
    const int Wx = W / 8, Hy = H / 8;
    const size_t plane = static_cast<size_t>(Wx) * Hy;
    for (int b = 0; b < P; ++b)
        for (int Y = 0; Y < Hy; ++Y)
            for (int X = 0; X < Wx; ++X) {
                uint64_t w = 0;
                for (int r = 0; r < 8; ++r)
                    for (int c = 0; c < 8; ++c)
                        w |= static_cast<uint64_t>((idx[static_cast<size_t>(8 * Y + r) * W + 8 * X + c] >> b) & 1) << (r * 8 + c);
                planes[b * plane + static_cast<size_t>(Y) * Wx + X] = w;
            }
}

inline void unpack_tiled(const uint64_t* planes, int W, int H, int P, int32_t* idx) {
    const int Wx = W / 8, Hy = H / 8;
    const size_t plane = static_cast<size_t>(Wx) * Hy;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int32_t v = 0;
            const size_t w = static_cast<size_t>(y / 8) * Wx + x / 8;
            const int bit = (y % 8) * 8 + x % 8;
            for (int b = 0; b < P; ++b) v |= static_cast<int32_t>((planes[b * plane + w] >> bit) & 1) << b;
            idx[static_cast<size_t>(y) * W + x] = v;
        }
}

// bit_array words of `word_bits` bits (32 or 64), returned in uint64 slots (one word each)
inline void pack_bit_array(const int32_t* idx, int W, int H, int P, int word_bits, uint64_t* words) {
    const int cpw = word_bits / P;
    const size_t n = static_cast<size_t>(W) * H, nw = (n + cpw - 1) / cpw;
    for (size_t w = 0; w < nw; ++w) words[w] = 0;
    for (size_t i = 0; i < n; ++i)
        words[i / cpw] |= static_cast<uint64_t>(idx[i] & ((1 << P) - 1)) << ((i % cpw) * P);
}

inline void unpack_bit_array(const uint64_t* words, int W, int H, int P, int word_bits, int32_t* idx) {
    const int cpw = word_bits / P;
    const size_t n = static_cast<size_t>(W) * H;
    for (size_t i = 0; i < n; ++i)
        idx[i] = static_cast<int32_t>((words[i / cpw] >> ((i % cpw) * P)) & ((1u << P) - 1));
}

} // namespace oracle::layouts

#endif
