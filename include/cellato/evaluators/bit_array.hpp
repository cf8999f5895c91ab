// Evaluator of the bit_array layout: the reference's evaluators::bit_array::evaluator<grid, Expression>
// (include/evaluators/bit_array.hpp:73-107 there). A word of the layout packs
// cells_per_word = word bits / P consecutive cells of a row, cell j holding its dictionary index
// in bits [j * P, (j + 1) * P). Every cell is evaluated on its own with plain integer
// semantics over dictionary indices (state literals are their index, a neighbour read is the
// neighbour's index), and the results are packed back into a word.
//
// The evaluation here is a compile-time interpreter of the rule's AST for one cell, over a
// cell context that supplies the neighbour indices, the cell's coordinate parities and the
// time step. The word loop and the neighbour extraction live in the HIP kernel
// (2026-cellato-smpt-journal_amd/csrc/bit_array.hpp); this header is also compiled for the host.
#ifndef CELLATO_EVALUATORS_BIT_ARRAY_HPP
#define CELLATO_EVALUATORS_BIT_ARRAY_HPP

#include <type_traits>

#include "../core/ast.hpp"

#if defined(__HIPCC__)
#define CELLATO_BA_HD __host__ __device__ __forceinline__
#else
#define CELLATO_BA_HD inline
#endif

namespace cellato::evaluators::bit_array {

using namespace cellato::ast;

// One cell's value of expression E. C is the cell context:
//   C::template nb<dx, dy>()  dictionary index of the cell at (x + dx, y + dy) on the torus
//   c.xodd, c.yodd            parities of the cell's x and y
//   c.t                       time step of the generation being computed
template <class Dict, class E>
struct cell_eval;

template <class Dict, auto V>
struct cell_eval<Dict, constant<V>> {
    template <class C>
    CELLATO_BA_HD static int at(const C&) { return static_cast<int>(V); }
};

template <class Dict, auto V>
struct cell_eval<Dict, state_constant<V>> {
    static constexpr int index = static_cast<int>(Dict::state_to_index(V));   // resolved at compile time
    template <class C>
    CELLATO_BA_HD static int at(const C&) { return index; }
};

template <class Dict, int dx, int dy>
struct cell_eval<Dict, neighbor_at<dx, dy>> {
    template <class C>
    CELLATO_BA_HD static int at(const C& c) { return c.template nb<dx, dy>(); }
};

template <class Dict, class Cond, class A, class B>
struct cell_eval<Dict, if_then_else<Cond, A, B>> {
    template <class C>
    CELLATO_BA_HD static int at(const C& c) {
        return cell_eval<Dict, Cond>::at(c) ? cell_eval<Dict, A>::at(c) : cell_eval<Dict, B>::at(c);
    }
};

#define CELLATO_BA_BINARY(NODE, EXPR)                                                       \
    template <class Dict, class L, class R>                                                 \
    struct cell_eval<Dict, NODE<L, R>> {                                                    \
        template <class C>                                                                  \
        CELLATO_BA_HD static int at(const C& c) {                                           \
            const int l = cell_eval<Dict, L>::at(c);                                        \
            const int r = cell_eval<Dict, R>::at(c);                                        \
            return EXPR;                                                                    \
        }                                                                                   \
    };

CELLATO_BA_BINARY(bit_and_, l & r)
CELLATO_BA_BINARY(bit_or_, l | r)
CELLATO_BA_BINARY(plus, l + r)
CELLATO_BA_BINARY(modulo, l % r)
CELLATO_BA_BINARY(equals, l == r)
CELLATO_BA_BINARY(not_equals, l != r)
CELLATO_BA_BINARY(greater_than, l > r)
CELLATO_BA_BINARY(less_than, l < r)
CELLATO_BA_BINARY(and_, (l != 0) & (r != 0))
CELLATO_BA_BINARY(or_, (l != 0) | (r != 0))
#undef CELLATO_BA_BINARY

template <class Dict, class V>
struct cell_eval<Dict, not_<V>> {
    template <class C>
    CELLATO_BA_HD static int at(const C& c) { return cell_eval<Dict, V>::at(c) == 0; }
};

template <class Dict, class V, int b>
struct cell_eval<Dict, has_bit_set<V, b>> {
    template <class C>
    CELLATO_BA_HD static int at(const C& c) { return (cell_eval<Dict, V>::at(c) >> b) & 1; }
};

// even time steps evaluate the first rule, odd ones the second
template <class Dict, class Even, class Odd>
struct cell_eval<Dict, alternate_algorithms<Even, Odd>> {
    template <class C>
    CELLATO_BA_HD static int at(const C& c) {
        return (c.t & 1) ? cell_eval<Dict, Odd>::at(c) : cell_eval<Dict, Even>::at(c);
    }
};

template <class Dict, class K>
struct cell_eval<Dict, count_neighbors<K, moore_8_neighbors>> {
    template <class C>
    CELLATO_BA_HD static int at(const C& c) {
        const int k = cell_eval<Dict, K>::at(c);
        return (c.template nb<-1, -1>() == k) + (c.template nb<0, -1>() == k) + (c.template nb<1, -1>() == k) +
               (c.template nb<-1, 0>() == k) + (c.template nb<1, 0>() == k) + (c.template nb<-1, 1>() == k) +
               (c.template nb<0, 1>() == k) + (c.template nb<1, 1>() == k);
    }
};

template <class Dict, class K>
struct cell_eval<Dict, count_neighbors<K, von_neumann_4_neighbors>> {
    template <class C>
    CELLATO_BA_HD static int at(const C& c) {
        const int k = cell_eval<Dict, K>::at(c);
        return (c.template nb<0, -1>() == k) + (c.template nb<-1, 0>() == k) + (c.template nb<1, 0>() == k) +
               (c.template nb<0, 1>() == k);
    }
};

// Margolus blocks: the 2x2 block containing (x, y) starts at even coordinates on even steps
// and odd ones on odd steps; per axis the cell is the block's first or second column (row).
// first(par, t): whether a cell of coordinate parity par is the block's first on step t.
CELLATO_BA_HD constexpr bool margolus_first(int par, int t) { return ((par ^ t) & 1) == 0; }

template <class Dict, class K>
struct cell_eval<Dict, count_neighbors<K, margolus_alternating_neighborhood>> {
    template <class C>
    CELLATO_BA_HD static int at(const C& c) {
        const int k = cell_eval<Dict, K>::at(c);
        const bool fx = margolus_first(c.xodd, c.t), fy = margolus_first(c.yodd, c.t);
        // the block's other column / row is at +1 for the first, -1 for the second
        const int self = c.template nb<0, 0>() == k;
        const int h = fx ? (c.template nb<1, 0>() == k) : (c.template nb<-1, 0>() == k);
        const int v = fy ? (c.template nb<0, 1>() == k) : (c.template nb<0, -1>() == k);
        const int d = fx ? (fy ? c.template nb<1, 1>() == k : c.template nb<1, -1>() == k)
                         : (fy ? c.template nb<-1, 1>() == k : c.template nb<-1, -1>() == k);
        return self + h + v + d;
    }
};

// the diagonally opposite cell of the cell's Margolus block
template <class Dict>
struct cell_eval<Dict, margolus_180_neighbor> {
    template <class C>
    CELLATO_BA_HD static int at(const C& c) {
        const bool fx = margolus_first(c.xodd, c.t), fy = margolus_first(c.yodd, c.t);
        return fx ? (fy ? c.template nb<1, 1>() : c.template nb<1, -1>())
                  : (fy ? c.template nb<-1, 1>() : c.template nb<-1, -1>());
    }
};

// The reference's evaluator name and parameters: evaluator<grid_t, Expression>. The grid type
// supplies the dictionary and the word type; the HIP traversers launch the kernel that applies
// `cell` to every cell of a word.
template <typename bit_array_grid_t, typename Expression>
struct evaluator {
    using grid_type = bit_array_grid_t;
    using dictionary = typename bit_array_grid_t::states_dict_t;
    using expression = Expression;
    static constexpr int layout = 2;   // the bit_array layout (the HIP traversers check the grid's)

    template <class C>
    CELLATO_BA_HD static int cell(const C& c) {
        return cell_eval<dictionary, Expression>::at(c);
    }
};

} // namespace cellato::evaluators::bit_array

#endif
