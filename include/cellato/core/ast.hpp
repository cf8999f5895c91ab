// Rule vocabulary of Cellato: a cellular-automaton rule is a *type* built from the node
// templates below. Nothing here has run-time state; evaluators interpret the type.
//
// The node names, their template parameters and the `if_<>::then_<>::elif_<>::else_<>`
// sugar are the reference's interface (include/core/ast.hpp:4-152), kept identical so that
// the reference's rule headers (src/<automaton>/algorithm.hpp) compile against this file
// unchanged.
#ifndef CELLATO_AST_HPP
#define CELLATO_AST_HPP

namespace cellato::ast {

// ---- leaves -----------------------------------------------------------------------------

// An integer literal (e.g. a neighbour count to compare with).        ref ast.hpp:6-10
template <auto V>
struct constant {
    using type = decltype(V);
    static constexpr type value = V;
};

// A cell state literal; evaluators map it through the state dictionary. ref ast.hpp:12-16
template <auto V>
struct state_constant {
    using type = decltype(V);
    static constexpr type value = V;
};

// The state of the cell at (x + dx, y + dy) on the torus.                ref ast.hpp:18-22
template <int dx, int dy>
struct neighbor_at {
    static constexpr int x_offset = dx;
    static constexpr int y_offset = dy;
};

using current_state = neighbor_at<0, 0>;                              // ref ast.hpp:150

// ---- neighbourhood tags ------------------------------------------------------------------
struct moore_8_neighbors {};                   // the 8 cells around (x, y)
struct von_neumann_4_neighbors {};             // (x, y±1) and (x±1, y)
struct margolus_alternating_neighborhood {};   // the 2x2 block of (x, y); offset flips each step
struct margolus_180_neighbor {};               // the diagonally opposite cell of that block

// Number of cells of the neighbourhood whose state equals `Constant`.   ref ast.hpp:29-33
template <typename Constant, typename Neighborhood>
struct count_neighbors {
    using state_constant = Constant;
    using neighborhood = Neighborhood;
};

// Whether bit `bit` of the (integer) state `Value` is set.               ref ast.hpp:35-39
template <typename Value, int bit>
struct has_bit_set {
    using value = Value;
    static constexpr int bit_position = bit;
};

// ---- control and arithmetic --------------------------------------------------------------
template <typename Condition, typename Then, typename Else>
struct if_then_else {
    using condition = Condition;
    using then_expr = Then;
    using else_expr = Else;
};

template <typename L, typename R> struct and_        { using left = L; using right = R; };
template <typename L, typename R> struct or_         { using left = L; using right = R; };
template <typename V>             struct not_        { using value = V; };
template <typename L, typename R> struct bit_and_    { using left = L; using right = R; };
template <typename L, typename R> struct bit_or_     { using left = L; using right = R; };
template <typename L, typename R> struct plus        { using left = L; using right = R; };
template <typename L, typename R> struct modulo      { using left = L; using right = R; };
template <typename L, typename R> struct equals      { using left = L; using right = R; };
template <typename L, typename R> struct not_equals  { using left = L; using right = R; };
template <typename L, typename R> struct greater_than{ using left = L; using right = R; };
template <typename L, typename R> struct less_than   { using left = L; using right = R; };

// Even time steps evaluate the first rule, odd time steps the second.  ref ast.hpp:83-85
template <typename... Rules>
struct alternate_algorithms {};

// ---- if_ sugar ---------------------------------------------------------------------------
// if_<C1>::then_<A>::elif_<C2>::then_<B>::else_<E>  ==  if_then_else<C1, A, if_then_else<C2, B, E>>
// (ref ast.hpp:117-144). `chain` folds the collected (condition, branch) pairs right to left.
template <typename... Parts>
struct __unpacked_if;

template <typename Last>
struct __unpacked_if<Last> {
    using nested_if_then_else = Last;
};

template <typename Cond, typename Branch, typename... Rest>
struct __unpacked_if<Cond, Branch, Rest...> {
    using nested_if_then_else =
        if_then_else<Cond, Branch, typename __unpacked_if<Rest...>::nested_if_then_else>;
};

template <typename Condition, typename... Collected>
struct if_ {
    template <typename Branch>
    struct then_ {
        template <typename Else>
        using else_ = typename __unpacked_if<Condition, Collected..., Branch, Else>::nested_if_then_else;

        template <typename NextCondition>
        using elif_ = if_<Condition, Collected..., Branch, NextCondition>;
    };
};

// Infix helper: p<A, equals, B> == equals<A, B>.                         ref ast.hpp:147-148
template <typename L, template <typename, typename> class Op, typename R>
using p = Op<L, R>;

} // namespace cellato::ast

#endif // CELLATO_AST_HPP
