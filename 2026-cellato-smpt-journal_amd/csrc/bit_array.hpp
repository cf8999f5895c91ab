// One generation of a grid in the bit_array layout (cells packed as P-bit dictionary indices,
// cells_per_word = word bits / P consecutive cells of a row per word; the reference's
// bit_array evaluator under its simple traverser, include/evaluators/bit_array.hpp:73-107 and
// traversers/cuda/simple.cu:20-43 there).
//
// One lane per word. The lane reads the 3x3 words around its own (the row above, its row and
// the row below, each with the word to the left and right, torus-wrapped; neighbouring lanes
// read overlapping words, which the L1 / L2 serve), then evaluates each of its cells with the
// rule's per-cell integer interpreter (cellato/evaluators/bit_array.hpp), the loop over cells
// unrolled so that every neighbour read is a constant shift of one of the nine words, and
// stores the packed word. No temporal blocking: the reference has none for this layout.
#ifndef CELLATO_HIP_BIT_ARRAY_HPP
#define CELLATO_HIP_BIT_ARRAY_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "cellato/evaluators/bit_array.hpp"
#include "stream.hpp"   // xcd_block, BLOCK_THREADS

namespace cellato::hip {

struct bit_array_args {
    const void* in;
    void* out;
    int wx;          // words per row
    int h;           // rows
    int t;           // time step of the generation
    long long words; // wx * h
};

// The cell context of the per-cell interpreter: cell J of the middle word of a 3x3 block of
// words (w[row][column], row 0 above, column 0 to the left).
template <class Word, int P, int J>
struct ba_cell {
    static constexpr int CPW = static_cast<int>(sizeof(Word) * 8) / P;
    static constexpr Word MASK = (Word(1) << P) - 1;
    const Word (&w)[3][3];
    int xodd, yodd, t;

    template <int dx, int dy>
    __device__ __forceinline__ int nb() const {
        constexpr int pos = J + dx;
        constexpr int col = pos < 0 ? 0 : (pos >= CPW ? 2 : 1);
        constexpr int cell = pos < 0 ? CPW - 1 : (pos >= CPW ? 0 : pos);
        return static_cast<int>((w[dy + 1][col] >> (cell * P)) & MASK);
    }
};

template <class Dict, class Expr, class Word>
__global__ void __launch_bounds__(BLOCK_THREADS) bit_array_kernel(bit_array_args a) {
    constexpr int P = Dict::needed_bits;
    constexpr int CPW = static_cast<int>(sizeof(Word) * 8) / P;
    constexpr Word MASK = (Word(1) << P) - 1;
    const long long blk = xcd_block(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    const long long i = blk * BLOCK_THREADS + threadIdx.x;
    if (i >= a.words) return;
    const int y = static_cast<int>(i / a.wx);
    const int x = static_cast<int>(i - static_cast<long long>(y) * a.wx);
    const int ys[3] = {y == 0 ? a.h - 1 : y - 1, y, y + 1 == a.h ? 0 : y + 1};
    const int xs[3] = {x == 0 ? a.wx - 1 : x - 1, x, x + 1 == a.wx ? 0 : x + 1};
    const Word* in = static_cast<const Word*>(a.in);
    Word w[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) w[r][c] = in[static_cast<long long>(ys[r]) * a.wx + xs[c]];
    // parity of the first cell's x (cells_per_word may be odd)
    const int x0odd = static_cast<int>((static_cast<long long>(x) * CPW) & 1);
    Word out = 0;
    [&]<int... J>(std::integer_sequence<int, J...>) {
        ((out |= (static_cast<Word>(evaluators::bit_array::cell_eval<Dict, Expr>::at(
                      ba_cell<Word, P, J>{w, x0odd ^ (J & 1), y & 1, a.t})) & MASK)
                 << (J * P)),
         ...);
    }(std::make_integer_sequence<int, CPW>{});
    static_cast<Word*>(a.out)[i] = out;
}

} // namespace cellato::hip

#endif
