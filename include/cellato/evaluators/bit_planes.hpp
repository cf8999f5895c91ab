// Bit-plane evaluator: lowers a rule type (cellato::ast) into bitwise logic over 64-bit
// words, 64 cells per word and one word per bit plane. Same role and name as the
// reference's evaluators/bit_planes.hpp (`evaluator<word, dictionary, Expression>`), but
// the evaluation does not read memory itself. It reads a *window*: the caller (a HIP
// kernel or a CPU loop) holds the words around the word being evaluated and answers
//
//   template <int dx, int dy> vint<P> nb() const        value of neighbour (dx, dy), all planes
//   template <int dx, int dy, class F> word nbf(F) const F(value) for a per-cell F, computed
//                                                         before the shift (cheaper for P > 1)
//   word yodd() const                                     cells whose global row is odd
//   int t                                                 time step of the generation computed
//
// so one evaluator serves the linear (64x1 cells per word) and the tiled (8x8) layouts and
// any register/LDS staging. Results are exact: a value is a bit-sliced unsigned integer
// wide enough for every intermediate, so the result equals the reference's standard
// evaluator (include/evaluators/standard.hpp) on the dictionary index of every cell.
#ifndef CELLATO_EVALUATORS_BIT_PLANES_HPP
#define CELLATO_EVALUATORS_BIT_PLANES_HPP

#include <cstdint>
#include <type_traits>
#include <utility>

#include "../core/ast.hpp"

#if defined(__HIPCC__)
#define CELLATO_HD __host__ __device__ __forceinline__
#else
#define CELLATO_HD inline
#endif

namespace cellato::evaluators::bit_planes {

using word = std::uint64_t;
using namespace cellato::ast;

// ---- bit-sliced unsigned integers ---------------------------------------------------------
template <int N>
struct vint {
    static_assert(N >= 1, "a value has at least one plane");
    static constexpr int bits = N;
    word p[N];
};

template <class T> struct is_vint : std::false_type {};
template <int N> struct is_vint<vint<N>> : std::true_type {};

template <class T> struct is_state_literal : std::false_type {};
template <auto V> struct is_state_literal<cellato::ast::state_constant<V>> : std::true_type {};

constexpr int bit_width(long long v) { return v <= 0 ? 0 : 1 + bit_width(v >> 1); }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <long long V>
CELLATO_HD vint<cmax(1, bit_width(V))> make_const() {
    constexpr int N = cmax(1, bit_width(V));
    vint<N> r;
#pragma unroll
    for (int i = 0; i < N; ++i) r.p[i] = ((V >> i) & 1) ? ~word(0) : word(0);
    return r;
}

template <int N>
CELLATO_HD word plane(const vint<N>& a, int i) { return i < N ? a.p[i] : word(0); }

template <int K, int N>
CELLATO_HD vint<K> resize(const vint<N>& a) {
    vint<K> r;
#pragma unroll
    for (int i = 0; i < K; ++i) r.p[i] = plane(a, i);
    return r;
}

// A mask (one word, bit = condition of that cell) used as a 1-plane value.
CELLATO_HD vint<1> as_vint(word m) { vint<1> r; r.p[0] = m; return r; }
template <int N> CELLATO_HD vint<N> as_vint(const vint<N>& v) { return v; }

// Three-input boolean function `lut` (truth table indexed by a*4 + b*2 + c) per bit; one
// v_bitop3_b32 per 32-bit half on gfx950. With a compile-time constant input the plain
// expression is kept, so that the compiler can fold it.
template <unsigned lut>
CELLATO_HD word lop3(word a, word b, word c) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (__builtin_constant_p(a) || __builtin_constant_p(b) || __builtin_constant_p(c)) {
        word r = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if ((lut >> i) & 1) r |= ((i & 4) ? a : ~a) & ((i & 2) ? b : ~b) & ((i & 1) ? c : ~c);
        return r;
    }
    const unsigned lo = __builtin_amdgcn_bitop3_b32(static_cast<unsigned>(a), static_cast<unsigned>(b), static_cast<unsigned>(c), lut);
    const unsigned hi = __builtin_amdgcn_bitop3_b32(static_cast<unsigned>(a >> 32), static_cast<unsigned>(b >> 32), static_cast<unsigned>(c >> 32), lut);
    return (word(hi) << 32) | lo;
#else
    word r = 0;
    for (int i = 0; i < 8; ++i)
        if ((lut >> i) & 1) r |= ((i & 4) ? a : ~a) & ((i & 2) ? b : ~b) & ((i & 1) ? c : ~c);
    return r;
#endif
}

// full and half adders on words (sum = xor3, carry = majority)
CELLATO_HD void fa(word a, word b, word c, word& s, word& cy) {
    s = lop3<0x96>(a, b, c);
    cy = lop3<0xE8>(a, b, c);
}
CELLATO_HD void ha(word a, word b, word& s, word& cy) { s = a ^ b; cy = a & b; }

// exact sum: one plane wider than the wider operand
template <int A, int B>
CELLATO_HD vint<cmax(A, B) + 1> add(const vint<A>& a, const vint<B>& b) {
    constexpr int R = cmax(A, B);
    vint<R + 1> r;
    word carry = 0;
#pragma unroll
    for (int i = 0; i < R; ++i) fa(plane(a, i), plane(b, i), carry, r.p[i], carry);
    r.p[R] = carry;
    return r;
}

// a == b per cell: every plane agrees; from the second plane on, one three-input op per plane
// (r & ~(x ^ y))
template <int A, int B>
CELLATO_HD word eq(const vint<A>& a, const vint<B>& b) {
    word r = ~(plane(a, 0) ^ plane(b, 0));
#pragma unroll
    for (int i = 1; i < cmax(A, B); ++i) r = lop3<0x90>(r, plane(a, i), plane(b, i));
    return r;
}

// a == K per cell: the AND of the planes taken as they are or negated, three literals per
// three-input op (the truth table of the op has the one row that matches K set)
template <long long K, int A>
CELLATO_HD word eq_const(const vint<A>& a) {
    if constexpr (bit_width(K) > A) {
        return word(0);
    } else if constexpr (A == 1) {
        return (K & 1) ? a.p[0] : ~a.p[0];
    } else if constexpr (A == 2) {
        return ((K & 1) ? a.p[0] : ~a.p[0]) & ((K & 2) ? a.p[1] : ~a.p[1]);
    } else {
        constexpr unsigned row = 4 * (K & 1) + 2 * ((K >> 1) & 1) + ((K >> 2) & 1);
        word r = lop3<1u << row>(a.p[0], a.p[1], a.p[2]);
#pragma unroll
        for (int i = 3; i < A; i += 2) {
            const bool k0 = (K >> i) & 1, k1 = (K >> (i + 1)) & 1;
            if (i + 1 < A) {
                if (k0 && k1) r = lop3<0x80>(r, a.p[i], a.p[i + 1]);
                else if (k0) r = lop3<0x40>(r, a.p[i], a.p[i + 1]);
                else if (k1) r = lop3<0x20>(r, a.p[i], a.p[i + 1]);
                else r = lop3<0x10>(r, a.p[i], a.p[i + 1]);
            } else {
                r &= k0 ? a.p[i] : ~a.p[i];
            }
        }
        return r;
    }
}

// a > b per cell, scanning from the most significant plane
template <int A, int B>
CELLATO_HD word gt(const vint<A>& a, const vint<B>& b) {
    word res = 0, undecided = ~word(0);
#pragma unroll
    for (int i = cmax(A, B) - 1; i >= 0; --i) {
        const word x = plane(a, i), y = plane(b, i);
        res |= undecided & x & ~y;
        undecided &= ~(x ^ y);
    }
    return res;
}

template <int A, int B, class Op>
CELLATO_HD vint<cmax(A, B)> zip(const vint<A>& a, const vint<B>& b, Op op) {
    vint<cmax(A, B)> r;
#pragma unroll
    for (int i = 0; i < cmax(A, B); ++i) r.p[i] = op(plane(a, i), plane(b, i));
    return r;
}

// per-cell select: m ? a : b
template <int A, int B>
CELLATO_HD vint<cmax(A, B)> select(word m, const vint<A>& a, const vint<B>& b) {
    vint<cmax(A, B)> r;
#pragma unroll
    for (int i = 0; i < cmax(A, B); ++i) r.p[i] = (plane(a, i) & m) | (plane(b, i) & ~m);
    return r;
}
// the same as one three-input op per plane (a rule's if-then-else of computed values: the
// compiler does not always find that form by itself)
template <int A, int B>
CELLATO_HD vint<cmax(A, B)> select_lut(word m, const vint<A>& a, const vint<B>& b) {
    vint<cmax(A, B)> r;
#pragma unroll
    for (int i = 0; i < cmax(A, B); ++i) r.p[i] = lop3<0xCA>(m, plane(a, i), plane(b, i));
    return r;
}
CELLATO_HD word select(word m, word a, word b) { return (a & m) | (b & ~m); }

// Population count of the 8 Moore masks as a 4-plane value (carry-save adder tree). The
// masks come in row order: m[0..2] = row above (x-1, x, x+1), m[3], m[4] = left and right,
// m[5..7] = row below. The 3-cell sums of the rows above and below are formed by identical
// expressions for consecutive output rows, so a streaming caller computes each once.
CELLATO_HD vint<4> count8(const word (&m)[8]) {
    word s1, c1, s2, c2, s3, c3, c4, s5, c5, c6;
    vint<4> r;
    fa(m[0], m[1], m[2], s1, c1);
    fa(m[5], m[6], m[7], s2, c2);
    ha(m[3], m[4], s3, c3);
    fa(s1, s2, s3, r.p[0], c4);
    fa(c1, c2, c3, s5, c5);
    ha(s5, c4, r.p[1], c6);
    ha(c5, c6, r.p[2], r.p[3]);
    return r;
}

CELLATO_HD vint<3> count4(const word (&m)[4]) {
    word s1, c1, c2;
    vint<3> r;
    fa(m[0], m[1], m[2], s1, c1);
    ha(s1, m[3], r.p[0], c2);
    ha(c1, c2, r.p[1], r.p[2]);
    return r;
}

// x-partner inside 2-cell pairs (cells 2j <-> 2j+1 of a word); valid for both layouts since
// every word starts at an even x.
CELLATO_HD word swap_pairs(word v) {
    // per 32-bit half (pairs never straddle the halves): even cells from the shifted-down
    // word, odd cells from the shifted-up one, one select
    const auto half = [](std::uint32_t x) -> std::uint32_t { return ((x >> 1) & 0x55555555u) | ((x << 1) & 0xAAAAAAAAu); };
    return (word(half(static_cast<std::uint32_t>(v >> 32))) << 32) | half(static_cast<std::uint32_t>(v));
}

// ---- per-row sums of the counted states (streaming callers) --------------------------------
// A caller that streams rows can work out once per row, for every state the rule counts over
// the Moore neighbourhood, the 3-cell horizontal sum (cells x-1, x, x+1) of the mask "cell
// is in that state" as a sum bit s and a carry c, and hand it to the three output rows that
// read the row (as the row above, its own row, the row below). Such a window declares
//
//   static constexpr bool row_sums = true;
//   template <long long K, int dy> void rsum(word& s, word& c) const   // K = dictionary index
//
// and the Moore counts below are formed from those sums: the own row's sum minus the cell
// itself, plus the sums above and below, gives the count as b0 + 2 (s1 + k + 2 m1).
// Comparisons of a count with a constant are answered from (b0, k, s1, m1) without forming
// the count.

template <long long... K>
struct klist {
    static constexpr int size = sizeof...(K);
};

// union of lists, first occurrence order
template <class A, class B> struct kjoin;
template <long long... A>
struct kjoin<klist<A...>, klist<>> { using type = klist<A...>; };
template <long long... A, long long B0, long long... B>
struct kjoin<klist<A...>, klist<B0, B...>> {
    using type = typename kjoin<std::conditional_t<((A == B0) || ...), klist<A...>, klist<A..., B0>>, klist<B...>>::type;
};
template <class... L> struct kjoin_all { using type = klist<>; };
template <class L0, class... L>
struct kjoin_all<L0, L...> { using type = typename kjoin<L0, typename kjoin_all<L...>::type>::type; };

// position of K in a list (-1 if absent)
template <long long K, class L> struct kindex;
template <long long K>
struct kindex<K, klist<>> { static constexpr int value = -1; };
template <long long K, long long H, long long... T>
struct kindex<K, klist<H, T...>> {
    static constexpr int value = K == H ? 0 : (kindex<K, klist<T...>>::value < 0 ? -1 : 1 + kindex<K, klist<T...>>::value);
};

template <class W>
constexpr bool has_row_sums() {
    if constexpr (requires { W::row_sums; }) return W::row_sums;
    else return false;
}
// the von Neumann "any" states whose row sums a window keeps as (3-cell OR, mask), and whether
// its counts include the cell itself
template <class W>
struct win_vn { using type = klist<>; };
template <class W>
    requires requires { typename W::vn_targets; }
struct win_vn<W> { using type = typename W::vn_targets; };
template <class W>
constexpr bool includes_cell() {
    if constexpr (requires { W::includes_cell; }) return W::includes_cell;
    else return false;
}

// the Moore count (0..8) of a counted state around a cell, in adder form
struct moore_sum {
    word b0, k, s1, m1;   // count = b0 + 2 (s1 + k + 2 m1)
};

// q = s1 + k + 2 m1 (0..4): q == j, q >= j. Each is one three-input function of (s1, k, m1),
// written as its truth table (index 4 s1 + 2 k + m1) so that it stays one bitop3 per half.
template <int j>
CELLATO_HD word q_is(const moore_sum& a) {
    if constexpr (j == 0) return lop3<0x01>(a.s1, a.k, a.m1);        // (0,0,0)
    else if constexpr (j == 1) return lop3<0x14>(a.s1, a.k, a.m1);   // (1,0,0) (0,1,0)
    else if constexpr (j == 2) return lop3<0x42>(a.s1, a.k, a.m1);   // (1,1,0) (0,0,1)
    else if constexpr (j == 3) return lop3<0x28>(a.s1, a.k, a.m1);   // (1,0,1) (0,1,1)
    else if constexpr (j == 4) return lop3<0x80>(a.s1, a.k, a.m1);   // (1,1,1)
    else return word(0);
}
template <int j>
CELLATO_HD word q_ge(const moore_sum& a) {
    if constexpr (j <= 0) return ~word(0);
    else if constexpr (j == 1) return lop3<0xFE>(a.s1, a.k, a.m1);   // s1 | k | m1
    else if constexpr (j == 2) return lop3<0xEA>(a.s1, a.k, a.m1);   // m1 | (s1 & k)
    else if constexpr (j == 3) return lop3<0xA8>(a.s1, a.k, a.m1);   // m1 & (s1 | k)
    else if constexpr (j == 4) return lop3<0x80>(a.s1, a.k, a.m1);
    else return word(0);
}
// count == K, count > K
template <long long K>
CELLATO_HD word count_is(const moore_sum& a) {
    if constexpr (K < 0 || K > 8) return word(0);
    else if constexpr (K % 2) return a.b0 & q_is<static_cast<int>((K - 1) / 2)>(a);
    else return ~a.b0 & q_is<static_cast<int>(K / 2)>(a);
}
template <long long K>
CELLATO_HD word count_gt(const moore_sum& a) {
    if constexpr (K < 0) return ~word(0);
    else if constexpr (K >= 8) return word(0);
    else if constexpr (K % 2) return q_ge<static_cast<int>((K + 1) / 2)>(a);
    else return q_ge<static_cast<int>(K / 2 + 1)>(a) | (a.b0 & q_ge<static_cast<int>(K / 2)>(a));
}
// the count as a 4-plane value
CELLATO_HD vint<4> count_value(const moore_sum& a) {
    const word cy = a.s1 & a.k;
    vint<4> r;
    r.p[0] = a.b0;
    r.p[1] = a.s1 ^ a.k;
    r.p[2] = a.m1 ^ cy;
    r.p[3] = a.m1 & cy;
    return r;
}

// ---- evaluator ------------------------------------------------------------------------------
template <class Dict, class E>
struct ev;

template <class Dict, class E, class Win>
CELLATO_HD auto eval(const Win& w) { return ev<Dict, E>::at(w); }

// dictionary index of a state literal, evaluated at compile time outside any function so
// that device code never references the (host-only) throwing lookup
template <class Dict, auto V>
inline constexpr long long index_of = Dict::state_to_index(V);

// per-cell "state index == K" on a P-plane value
template <long long K>
struct is_index {
    template <int N>
    CELLATO_HD word operator()(const vint<N>& v) const { return eq_const<K>(v); }
};

// the Moore count of dictionary index K from a window with row sums: the own row's 3-cell sum
// (sm, cm) less the cell itself (m), then the three rows added column by column
template <long long K, class Win>
CELLATO_HD moore_sum moore_sum_of(const Win& w) {
    word st, ct, sm, cm, sb, cb;
    w.template rsum<K, -1>(st, ct);
    w.template rsum<K, 0>(sm, cm);
    w.template rsum<K, 1>(sb, cb);
    const word m = w.template nbf<0, 0>(is_index<K>{});
    // (sm + 2 cm) - m, never below 0: e0 = sm ^ m, e1 = cm & (sm | ~m)
    const word e0 = sm ^ m, e1 = lop3<0xD0>(cm, sm, m);
    moore_sum r;
    fa(st, e0, sb, r.b0, r.k);
    fa(ct, e1, cb, r.s1, r.m1);
    return r;
}

// the states a rule counts over the Moore neighbourhood (dictionary indices): what a
// streaming caller keeps row sums of
template <class Dict, class E>
struct row_targets { using type = klist<>; };
template <class Dict, template <class...> class N, class... A>
struct row_targets<Dict, N<A...>> { using type = typename kjoin_all<typename row_targets<Dict, A>::type...>::type; };
template <class Dict, auto V>
struct row_targets<Dict, count_neighbors<state_constant<V>, moore_8_neighbors>> { using type = klist<index_of<Dict, V>>; };

template <class Dict, auto V>
struct ev<Dict, constant<V>> {
    template <class Win> CELLATO_HD static auto at(const Win&) { return make_const<static_cast<long long>(V)>(); }
};

template <class Dict, auto V>
struct ev<Dict, state_constant<V>> {
    static constexpr long long index = index_of<Dict, V>;
    template <class Win> CELLATO_HD static auto at(const Win&) { return make_const<index>(); }
};

template <class Dict, int dx, int dy>
struct ev<Dict, neighbor_at<dx, dy>> {
    static_assert(dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1, "neighbour offsets are limited to the 3x3 block");
    template <class Win> CELLATO_HD static auto at(const Win& w) { return w.template nb<dx, dy>(); }
};

template <class Dict, class Even, class Odd>
struct ev<Dict, alternate_algorithms<Even, Odd>> {
    template <class Win>
    CELLATO_HD static auto at(const Win& w) {
        using A = decltype(as_vint(eval<Dict, Even>(w)));
        using B = decltype(as_vint(eval<Dict, Odd>(w)));
        constexpr int N = cmax(A::bits, B::bits);
        if ((w.t & 1) == 0) return resize<N>(as_vint(eval<Dict, Even>(w)));
        return resize<N>(as_vint(eval<Dict, Odd>(w)));
    }
};

template <class Dict, class C, class A, class B>
struct ev<Dict, if_then_else<C, A, B>> {
    template <class Win>
    CELLATO_HD static auto at(const Win& w) {
        const word c = eval<Dict, C>(w);
        auto a = eval<Dict, A>(w);
        auto b = eval<Dict, B>(w);
        if constexpr (std::is_same_v<decltype(a), word> && std::is_same_v<decltype(b), word>)
            return select(c, a, b);
        else
            return select_lut(c, as_vint(a), as_vint(b));
    }
};

template <class Dict, class L, class R>
struct ev<Dict, bit_and_<L, R>> {
    template <class Win>
    CELLATO_HD static auto at(const Win& w) {
        auto l = eval<Dict, L>(w);
        auto r = eval<Dict, R>(w);
        if constexpr (std::is_same_v<decltype(l), word> && std::is_same_v<decltype(r), word>)
            return word(l & r);
        else
            return zip(as_vint(l), as_vint(r), [](word x, word y) { return x & y; });
    }
};

template <class Dict, class L, class R>
struct ev<Dict, bit_or_<L, R>> {
    template <class Win>
    CELLATO_HD static auto at(const Win& w) {
        auto l = eval<Dict, L>(w);
        auto r = eval<Dict, R>(w);
        if constexpr (std::is_same_v<decltype(l), word> && std::is_same_v<decltype(r), word>)
            return word(l | r);
        else
            return zip(as_vint(l), as_vint(r), [](word x, word y) { return x | y; });
    }
};

// (c ? a : b) | s per bit as one three-input op when a or b is a compile-time constant (a
// rule's constant branch), instead of a select and an OR
CELLATO_HD word select_or(word c, word a, word b, word s) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (__builtin_constant_p(a) && a == 0) return lop3<0xAE>(c, b, s);             // (~c & b) | s
    if (__builtin_constant_p(a) && a == ~word(0)) return lop3<0xFE>(c, b, s);      // c | b | s
    if (__builtin_constant_p(b) && b == 0) return lop3<0xEA>(c, a, s);             // (c & a) | s
    if (__builtin_constant_p(b) && b == ~word(0)) return lop3<0xEF>(c, a, s);      // ~c | a | s
#endif
    return (a & c) | (b & ~c) | s;
}

// an OR whose left operand is an if-then-else (HPP: the vertical and horizontal results)
template <class Dict, class C, class A, class B, class R>
struct ev<Dict, bit_or_<if_then_else<C, A, B>, R>> {
    template <class Win>
    CELLATO_HD static auto at(const Win& w) {
        const word c = eval<Dict, C>(w);
        auto a = eval<Dict, A>(w);
        auto b = eval<Dict, B>(w);
        auto r = eval<Dict, R>(w);
        if constexpr (std::is_same_v<decltype(a), word> && std::is_same_v<decltype(b), word> &&
                      std::is_same_v<decltype(r), word>) {
            return select_or(c, a, b, r);
        } else {
            const auto va = as_vint(a);
            const auto vb = as_vint(b);
            const auto vr = as_vint(r);
            constexpr int N = cmax(cmax(decltype(va)::bits, decltype(vb)::bits), decltype(vr)::bits);
            vint<N> out;
#pragma unroll
            for (int i = 0; i < N; ++i) out.p[i] = select_or(c, plane(va, i), plane(vb, i), plane(vr, i));
            return out;
        }
    }
};

template <class Dict, class L, class R>
struct ev<Dict, plus<L, R>> {
    template <class Win>
    CELLATO_HD static auto at(const Win& w) { return add(as_vint(eval<Dict, L>(w)), as_vint(eval<Dict, R>(w))); }
};

template <class Dict, class L, int M>
struct ev<Dict, modulo<L, constant<M>>> {
    static_assert(M > 0 && (M & (M - 1)) == 0, "modulo is supported for powers of two");
    template <class Win>
    CELLATO_HD static auto at(const Win& w) { return resize<cmax(1, bit_width(M) - 1)>(as_vint(eval<Dict, L>(w))); }
};

template <class Dict, class V>
struct ev<Dict, not_<V>> {
    template <class Win> CELLATO_HD static word at(const Win& w) { return ~word(eval<Dict, V>(w)); }
};
template <class Dict, class L, class R>
struct ev<Dict, and_<L, R>> {
    template <class Win> CELLATO_HD static word at(const Win& w) { return word(eval<Dict, L>(w)) & word(eval<Dict, R>(w)); }
};
template <class Dict, class L, class R>
struct ev<Dict, or_<L, R>> {
    template <class Win> CELLATO_HD static word at(const Win& w) { return word(eval<Dict, L>(w)) | word(eval<Dict, R>(w)); }
};

template <class Dict, class L, class R>
struct ev<Dict, equals<L, R>> {
    template <class Win>
    CELLATO_HD static word at(const Win& w) { return eq(as_vint(eval<Dict, L>(w)), as_vint(eval<Dict, R>(w))); }
};
// equality of a neighbour with a state literal: per-cell test before the shift
template <class Dict, int dx, int dy, auto V>
struct ev<Dict, equals<neighbor_at<dx, dy>, state_constant<V>>> {
    template <class Win>
    CELLATO_HD static word at(const Win& w) {
        return w.template nbf<dx, dy>(is_index<index_of<Dict, V>>{});
    }
};
template <class Dict, class L, class R>
struct ev<Dict, not_equals<L, R>> {
    template <class Win> CELLATO_HD static word at(const Win& w) { return ~eq(as_vint(eval<Dict, L>(w)), as_vint(eval<Dict, R>(w))); }
};
template <class Dict, class L, class R>
struct ev<Dict, greater_than<L, R>> {
    template <class Win> CELLATO_HD static word at(const Win& w) { return gt(as_vint(eval<Dict, L>(w)), as_vint(eval<Dict, R>(w))); }
};
template <class Dict, class L, class R>
struct ev<Dict, less_than<L, R>> {
    template <class Win> CELLATO_HD static word at(const Win& w) { return gt(as_vint(eval<Dict, R>(w)), as_vint(eval<Dict, L>(w))); }
};

template <class Dict, class V, int b>
struct ev<Dict, has_bit_set<V, b>> {
    template <class Win> CELLATO_HD static word at(const Win& w) { return plane(as_vint(eval<Dict, V>(w)), b); }
};
// a bit of a neighbour: shift only that plane
template <class Dict, int dx, int dy, int b>
struct ev<Dict, has_bit_set<neighbor_at<dx, dy>, b>> {
    struct bit_of { template <int N> CELLATO_HD word operator()(const vint<N>& v) const { return plane(v, b); } };
    template <class Win> CELLATO_HD static word at(const Win& w) { return w.template nbf<dx, dy>(bit_of{}); }
};

// ---- neighbourhood counts -------------------------------------------------------------------
// Masks of "neighbour (dx, dy) matches the target", in the 8 Moore positions.
template <class Dict, class C>
struct moore_masks {
    template <class Win>
    CELLATO_HD static void get(const Win& w, word (&m)[8]) {
        if constexpr (is_state_literal<C>::value) {
            constexpr long long k = index_of<Dict, C::value>;
            const is_index<k> f{};
            m[0] = w.template nbf<-1, -1>(f); m[1] = w.template nbf<0, -1>(f); m[2] = w.template nbf<1, -1>(f);
            m[3] = w.template nbf<-1, 0>(f);                                   m[4] = w.template nbf<1, 0>(f);
            m[5] = w.template nbf<-1, 1>(f);  m[6] = w.template nbf<0, 1>(f);  m[7] = w.template nbf<1, 1>(f);
        } else {
            const auto target = as_vint(eval<Dict, C>(w));
            m[0] = eq(w.template nb<-1, -1>(), target); m[1] = eq(w.template nb<0, -1>(), target);
            m[2] = eq(w.template nb<1, -1>(), target);  m[3] = eq(w.template nb<-1, 0>(), target);
            m[4] = eq(w.template nb<1, 0>(), target);   m[5] = eq(w.template nb<-1, 1>(), target);
            m[6] = eq(w.template nb<0, 1>(), target);   m[7] = eq(w.template nb<1, 1>(), target);
        }
    }
};

template <class Dict, class C>
struct vn_masks {
    template <class Win>
    CELLATO_HD static void get(const Win& w, word (&m)[4]) {
        if constexpr (is_state_literal<C>::value) {
            constexpr long long k = index_of<Dict, C::value>;
            const is_index<k> f{};
            m[0] = w.template nbf<0, 1>(f); m[1] = w.template nbf<0, -1>(f);
            m[2] = w.template nbf<1, 0>(f); m[3] = w.template nbf<-1, 0>(f);
        } else {
            const auto target = as_vint(eval<Dict, C>(w));
            m[0] = eq(w.template nb<0, 1>(), target); m[1] = eq(w.template nb<0, -1>(), target);
            m[2] = eq(w.template nb<1, 0>(), target); m[3] = eq(w.template nb<-1, 0>(), target);
        }
    }
};

template <class Dict, class C>
struct ev<Dict, count_neighbors<C, moore_8_neighbors>> {
    template <class Win>
    CELLATO_HD static vint<4> at(const Win& w) {
        if constexpr (is_state_literal<C>::value && has_row_sums<Win>()) {
            return count_value(moore_sum_of<index_of<Dict, C::value>>(w));
        } else {
            word m[8];
            moore_masks<Dict, C>::get(w, m);
            return count8(m);
        }
    }
};

// a Moore count of a state compared with a constant: from the row sums when the window has
// them, without forming the count
template <class Dict, auto V, auto K>
struct ev<Dict, equals<count_neighbors<state_constant<V>, moore_8_neighbors>, constant<K>>> {
    template <class Win>
    CELLATO_HD static word at(const Win& w) {
        if constexpr (has_row_sums<Win>())
            return count_is<static_cast<long long>(K)>(moore_sum_of<index_of<Dict, V>>(w));
        else
            return eq(ev<Dict, count_neighbors<state_constant<V>, moore_8_neighbors>>::at(w), make_const<static_cast<long long>(K)>());
    }
};
template <class Dict, auto V, auto K>
struct ev<Dict, not_equals<count_neighbors<state_constant<V>, moore_8_neighbors>, constant<K>>> {
    template <class Win>
    CELLATO_HD static word at(const Win& w) {
        return ~ev<Dict, equals<count_neighbors<state_constant<V>, moore_8_neighbors>, constant<K>>>::at(w);
    }
};
template <class Dict, auto V, auto K>
struct ev<Dict, less_than<count_neighbors<state_constant<V>, moore_8_neighbors>, constant<K>>> {
    template <class Win>
    CELLATO_HD static word at(const Win& w) {
        if constexpr (has_row_sums<Win>())
            return ~count_gt<static_cast<long long>(K) - 1>(moore_sum_of<index_of<Dict, V>>(w));
        else
            return gt(make_const<static_cast<long long>(K)>(), ev<Dict, count_neighbors<state_constant<V>, moore_8_neighbors>>::at(w));
    }
};

template <class Dict, class C>
struct ev<Dict, count_neighbors<C, von_neumann_4_neighbors>> {
    template <class Win>
    CELLATO_HD static vint<3> at(const Win& w) {
        word m[4];
        vn_masks<Dict, C>::get(w, m);
        return count4(m);
    }
};

// "any neighbour matches": count > 0 without building the count
template <class Dict, class C, auto Z>
struct ev<Dict, greater_than<count_neighbors<C, moore_8_neighbors>, constant<Z>>> {
    template <class Win>
    CELLATO_HD static word at(const Win& w) {
        if constexpr (is_state_literal<C>::value && has_row_sums<Win>() && Z == 0) {
            // any neighbour: a row sum above or below is not 0, or the own row holds more than the cell
            constexpr long long K = index_of<Dict, C::value>;
            word st, ct, sm, cm, sb, cb;
            w.template rsum<K, -1>(st, ct);
            w.template rsum<K, 0>(sm, cm);
            w.template rsum<K, 1>(sb, cb);
            return st | ct | sb | cb | cm | (sm & ~w.template nbf<0, 0>(is_index<K>{}));
        } else if constexpr (is_state_literal<C>::value && has_row_sums<Win>()) {
            return count_gt<static_cast<long long>(Z)>(moore_sum_of<index_of<Dict, C::value>>(w));
        } else if constexpr (Z == 0) {
            word m[8];
            moore_masks<Dict, C>::get(w, m);
            return m[0] | m[1] | m[2] | m[3] | m[4] | m[5] | m[6] | m[7];
        } else {
            return gt(ev<Dict, count_neighbors<C, moore_8_neighbors>>::at(w), make_const<Z>());
        }
    }
};
template <class Dict, class C, auto Z>
struct ev<Dict, greater_than<count_neighbors<C, von_neumann_4_neighbors>, constant<Z>>> {
    template <class Win>
    CELLATO_HD static word at(const Win& w) {
        if constexpr (Z == 0 && is_state_literal<C>::value && includes_cell<Win>() &&
                      kindex<index_of<Dict, C::value>, typename win_vn<Win>::type>::value >= 0) {
            // the own row's 3-cell OR (the cell included: never K where this is read) and the
            // masks of the cells above and below, all kept per row
            constexpr long long K = index_of<Dict, C::value>;
            word st, ct, sm, cm, sb, cb;
            w.template rsum<K, -1>(st, ct);
            w.template rsum<K, 0>(sm, cm);
            w.template rsum<K, 1>(sb, cb);
            return lop3<0xFE>(sm, ct, cb);
        } else if constexpr (Z == 0) {
            word m[4];
            vn_masks<Dict, C>::get(w, m);
            return m[0] | m[1] | m[2] | m[3];
        } else {
            return gt(ev<Dict, count_neighbors<C, von_neumann_4_neighbors>>::at(w), make_const<Z>());
        }
    }
};

// ---- Margolus neighbourhood ---------------------------------------------------------------
// In global coordinates: on even steps the 2x2 blocks start at even (x, y), on odd steps at
// odd (x, y) (standard.hpp:185-248). The vertical partner row and the horizontal partner
// column are chosen per cell from the row parity mask and the cell's x parity.
constexpr word even_cols = 0x5555555555555555ULL;   // cells with even x
constexpr word odd_cols = 0xAAAAAAAAAAAAAAAAULL;    // cells with odd x

// value of the vertical block partner (dy = +-1), for any per-cell function F
template <int dx, class Win, class F>
CELLATO_HD word vpartner_f(const Win& w, F f) {
    const word up = w.template nbf<dx, -1>(f), down = w.template nbf<dx, 1>(f);
    // even step: even rows pair with the row below; odd step: with the row above
    return (w.t & 1) == 0 ? select(w.yodd(), up, down) : select(w.yodd(), down, up);
}

template <class Dict, auto V>
struct ev<Dict, count_neighbors<state_constant<V>, margolus_alternating_neighborhood>> {
    template <class Win>
    CELLATO_HD static vint<3> at(const Win& w) {
        const is_index<index_of<Dict, V>> f{};
        const word own = w.template nbf<0, 0>(f);
        const word ver = vpartner_f<0>(w, f);
        word own_h, ver_h;   // the same two rows, at the horizontal block partner
        if ((w.t & 1) == 0) {
            own_h = swap_pairs(own);
            ver_h = swap_pairs(ver);
        } else {
            own_h = (w.template nbf<-1, 0>(f) & even_cols) | (w.template nbf<1, 0>(f) & odd_cols);
            ver_h = (vpartner_f<-1>(w, f) & even_cols) | (vpartner_f<1>(w, f) & odd_cols);
        }
        const word m[4] = {own, ver, own_h, ver_h};
        return count4(m);
    }
};

template <class Dict>
struct ev<Dict, margolus_180_neighbor> {
    template <class Win>
    CELLATO_HD static auto at(const Win& w) {
        const word yo = w.yodd();
        if ((w.t & 1) == 0) {
            auto v = select(yo, w.template nb<0, -1>(), w.template nb<0, 1>());
            decltype(v) r;
#pragma unroll
            for (int i = 0; i < decltype(v)::bits; ++i) r.p[i] = swap_pairs(v.p[i]);
            return r;
        }
        auto left = select(yo, w.template nb<-1, 1>(), w.template nb<-1, -1>());
        auto right = select(yo, w.template nb<1, 1>(), w.template nb<1, -1>());
        return zip(left, right, [](word l, word r) { return (l & even_cols) | (r & odd_cols); });
    }
};

// ---- Margolus table path ----------------------------------------------------------------
// A one-plane rule made of the cell, one state's Margolus block count, the 180-degree partner,
// constants, comparisons, and / or / not and if-then-else (Critters) depends only on the four
// cells of the cell's 2x2 block: the cell, its vertical partner, its horizontal partner and
// the diagonal one (which is also the 180-degree partner). mg_val<..>::at(cell, count,
// partner) evaluates such a rule at compile time; the kernel then forms the four block words
// and each result plane is one 4-input function of them (three ops per 32 bits) instead of
// the 4-way count, its comparisons and the selects.
template <class Dict, class E>
struct mg_val { static constexpr bool ok = false; static constexpr long long K = -1;
                static constexpr long long at(long long, long long, long long) { return 0; } };
template <class Dict, auto V>
struct mg_val<Dict, state_constant<V>> {
    static constexpr bool ok = true; static constexpr long long K = -1;
    static constexpr long long at(long long, long long, long long) { return index_of<Dict, V>; }
};
template <class Dict, auto V>
struct mg_val<Dict, constant<V>> {
    static constexpr bool ok = true; static constexpr long long K = -1;
    static constexpr long long at(long long, long long, long long) { return static_cast<long long>(V); }
};
template <class Dict>
struct mg_val<Dict, neighbor_at<0, 0>> {
    static constexpr bool ok = true; static constexpr long long K = -1;
    static constexpr long long at(long long c, long long, long long) { return c; }
};
template <class Dict>
struct mg_val<Dict, margolus_180_neighbor> {
    static constexpr bool ok = true; static constexpr long long K = -1;
    static constexpr long long at(long long, long long, long long r) { return r; }
};
template <class Dict, auto V>
struct mg_val<Dict, count_neighbors<state_constant<V>, margolus_alternating_neighborhood>> {
    static constexpr bool ok = true; static constexpr long long K = index_of<Dict, V>;
    static constexpr long long at(long long, long long n, long long) { return n; }
};
// the counted state of two subtrees: -1 none, -2 two different ones
constexpr long long mg_k(long long a, long long b) { return a == -1 ? b : b == -1 ? a : a == b ? a : -2; }
#define CELLATO_MG_BIN(node, expr)                                                              \
    template <class Dict, class L, class R>                                                     \
    struct mg_val<Dict, node<L, R>> {                                                           \
        static constexpr bool ok = mg_val<Dict, L>::ok && mg_val<Dict, R>::ok;                  \
        static constexpr long long K = mg_k(mg_val<Dict, L>::K, mg_val<Dict, R>::K);            \
        static constexpr long long at(long long c, long long n, long long r) {                  \
            const long long a = mg_val<Dict, L>::at(c, n, r), b = mg_val<Dict, R>::at(c, n, r); \
            return expr;                                                                        \
        }                                                                                       \
    };
CELLATO_MG_BIN(equals, a == b ? 1 : 0)
CELLATO_MG_BIN(not_equals, a != b ? 1 : 0)
CELLATO_MG_BIN(greater_than, a > b ? 1 : 0)
CELLATO_MG_BIN(less_than, a < b ? 1 : 0)
CELLATO_MG_BIN(and_, (a && b) ? 1 : 0)
CELLATO_MG_BIN(or_, (a || b) ? 1 : 0)
#undef CELLATO_MG_BIN
template <class Dict, class V>
struct mg_val<Dict, not_<V>> {
    static constexpr bool ok = mg_val<Dict, V>::ok; static constexpr long long K = mg_val<Dict, V>::K;
    static constexpr long long at(long long c, long long n, long long r) { return mg_val<Dict, V>::at(c, n, r) ? 0 : 1; }
};
template <class Dict, class C, class A, class B>
struct mg_val<Dict, if_then_else<C, A, B>> {
    static constexpr bool ok = mg_val<Dict, C>::ok && mg_val<Dict, A>::ok && mg_val<Dict, B>::ok;
    static constexpr long long K = mg_k(mg_k(mg_val<Dict, C>::K, mg_val<Dict, A>::K), mg_val<Dict, B>::K);
    static constexpr long long at(long long c, long long n, long long r) {
        return mg_val<Dict, C>::at(c, n, r) ? mg_val<Dict, A>::at(c, n, r) : mg_val<Dict, B>::at(c, n, r);
    }
};

template <class Dict, class E>
struct mg_rule {
    using X = mg_val<Dict, E>;
    static constexpr bool ok = X::ok && X::K >= 0 && Dict::needed_bits == 1;
    // result bit over the block cells' bits, index own + 2 ver + 4 own_h + 8 ver_h
    static constexpr unsigned long long table() {
        unsigned long long t = 0;
        for (int i = 0; i < 16; ++i) {
            const long long own = i & 1, ver = (i >> 1) & 1, own_h = (i >> 2) & 1, ver_h = (i >> 3) & 1;
            const long long n = (own == X::K) + (ver == X::K) + (own_h == X::K) + (ver_h == X::K);
            if (X::at(own, n, ver_h) & 1) t |= 1ull << i;
        }
        return t;
    }
};

struct plane0 {
    template <int N> CELLATO_HD word operator()(const vint<N>& v) const { return v.p[0]; }
};

template <int N, unsigned long long F, unsigned long long D>
CELLATO_HD word pt_fn(const word* in);

template <class Dict, class E, class Win>
CELLATO_HD word mg_eval(const Win& w) {
    const plane0 f{};
    word in[4];
    in[0] = w.template nbf<0, 0>(f);
    in[1] = vpartner_f<0>(w, f);
    if ((w.t & 1) == 0) {
        in[2] = swap_pairs(in[0]);
        in[3] = swap_pairs(in[1]);
    } else {
        in[2] = lop3<0xCA>(even_cols, w.template nbf<-1, 0>(f), w.template nbf<1, 0>(f));
        in[3] = lop3<0xCA>(even_cols, vpartner_f<-1>(w, f), vpartner_f<1>(w, f));
    }
    return pt_fn<4, mg_rule<Dict, E>::table(), 0ull>(in);
}

// ---- rule traits: which kinds of nodes a rule uses (compile-time walk of the AST) -----------
enum rule_trait : unsigned {
    TRAIT_MOORE = 1u,          // count over the Moore neighbourhood
    TRAIT_VON_NEUMANN = 2u,    // count over the von Neumann neighbourhood
    TRAIT_MARGOLUS = 4u,       // Margolus block count or 180-degree neighbour
    TRAIT_NEIGHBOR = 8u,       // a single neighbour read directly
    TRAIT_DYNAMIC_TARGET = 16u,// a count whose target state is computed per cell
    TRAIT_ARITHMETIC = 32u,    // plus / modulo on states
    TRAIT_ALTERNATING = 64u,   // alternate_algorithms (parity of the time step)
};

template <class E>
struct rule_traits {
    static constexpr unsigned value = 0;
};
template <template <class...> class N, class... A>
struct rule_traits<N<A...>> {
    static constexpr unsigned value = (0u | ... | rule_traits<A>::value);
};
template <int dx, int dy>
struct rule_traits<neighbor_at<dx, dy>> {
    static constexpr unsigned value = (dx == 0 && dy == 0) ? 0u : TRAIT_NEIGHBOR;
};
template <class V, int b>
struct rule_traits<has_bit_set<V, b>> {
    static constexpr unsigned value = rule_traits<V>::value;
};
template <class C>
struct rule_traits<count_neighbors<C, moore_8_neighbors>> {
    static constexpr unsigned value = TRAIT_MOORE | (is_state_literal<C>::value ? 0u : TRAIT_DYNAMIC_TARGET) | rule_traits<C>::value;
};
template <class C>
struct rule_traits<count_neighbors<C, von_neumann_4_neighbors>> {
    static constexpr unsigned value = TRAIT_VON_NEUMANN | (is_state_literal<C>::value ? 0u : TRAIT_DYNAMIC_TARGET) | rule_traits<C>::value;
};
template <class C>
struct rule_traits<count_neighbors<C, margolus_alternating_neighborhood>> {
    static constexpr unsigned value = TRAIT_MARGOLUS;
};
template <>
struct rule_traits<margolus_180_neighbor> {
    static constexpr unsigned value = TRAIT_MARGOLUS;
};
template <class L, class R>
struct rule_traits<plus<L, R>> {
    static constexpr unsigned value = TRAIT_ARITHMETIC | rule_traits<L>::value | rule_traits<R>::value;
};
template <class L, class R>
struct rule_traits<modulo<L, R>> {
    static constexpr unsigned value = TRAIT_ARITHMETIC | rule_traits<L>::value | rule_traits<R>::value;
};
template <class... A>
struct rule_traits<alternate_algorithms<A...>> {
    static constexpr unsigned value = TRAIT_ALTERNATING | (0u | ... | rule_traits<A>::value);
};

// ---- rules of the cell and one Moore count (1 plane) -----------------------------------------
// A 1-plane rule built only from the cell itself, one state's Moore count, constants,
// comparisons, and / or / not and if-then-else (Game of Life, Maze) is a table T[cell][count].
// With the count in adder form (count = b0 + 2q, q = s1 + k + 2 m1 in 0..4) the rule becomes,
// for each 2-input function g of (cell, b0) that T takes at some values of q,
// [q in S_g] & g(cell, b0), OR-ed over the distinct g: one three-input op for the set test and
// one for the term. Game of Life: q == 1 & (cell | b0), 2 ops per half word against the 6 the
// node-by-node evaluation takes.
template <class Dict, class E>
struct cc_expr {   // not of that form
    static constexpr bool ok = false;
    static constexpr long long K = -1;
    static constexpr long long at(long long, int) { return 0; }
};
template <class Dict, auto V>
struct cc_expr<Dict, constant<V>> {
    static constexpr bool ok = true;
    static constexpr long long K = -1;
    static constexpr long long at(long long, int) { return static_cast<long long>(V); }
};
template <class Dict, auto V>
struct cc_expr<Dict, state_constant<V>> {
    static constexpr bool ok = true;
    static constexpr long long K = -1;
    static constexpr long long at(long long, int) { return index_of<Dict, V>; }
};
template <class Dict>
struct cc_expr<Dict, neighbor_at<0, 0>> {   // the cell's dictionary index
    static constexpr bool ok = true;
    static constexpr long long K = -1;
    static constexpr long long at(long long cell, int) { return cell; }
};
template <class Dict, auto V>
struct cc_expr<Dict, count_neighbors<state_constant<V>, moore_8_neighbors>> {
    static constexpr bool ok = true;
    static constexpr long long K = index_of<Dict, V>;
    static constexpr long long at(long long, int count) { return count; }
};
// the operands' count, if any, must be of one state
template <class Dict, class... A>
struct cc_nodes {
    static constexpr long long K = [] {
        long long k = -1;
        ((k = cc_expr<Dict, A>::K >= 0 ? cc_expr<Dict, A>::K : k), ...);
        return k;
    }();
    static constexpr bool ok = (cc_expr<Dict, A>::ok && ...) && ((cc_expr<Dict, A>::K < 0 || cc_expr<Dict, A>::K == K) && ...);
};
#define CELLATO_CC_BINARY(node, op)                                                             \
    template <class Dict, class L, class R>                                                     \
    struct cc_expr<Dict, node<L, R>> : cc_nodes<Dict, L, R> {                                   \
        static constexpr long long at(long long c, int n) {                                     \
            return (cc_expr<Dict, L>::at(c, n) op cc_expr<Dict, R>::at(c, n)) ? 1 : 0;          \
        }                                                                                       \
    };
CELLATO_CC_BINARY(equals, ==)
CELLATO_CC_BINARY(not_equals, !=)
CELLATO_CC_BINARY(less_than, <)
CELLATO_CC_BINARY(greater_than, >)
CELLATO_CC_BINARY(and_, &&)
CELLATO_CC_BINARY(or_, ||)
#undef CELLATO_CC_BINARY
template <class Dict, class V>
struct cc_expr<Dict, not_<V>> : cc_nodes<Dict, V> {
    static constexpr long long at(long long c, int n) { return cc_expr<Dict, V>::at(c, n) ? 0 : 1; }
};
template <class Dict, class C, class A, class B>
struct cc_expr<Dict, if_then_else<C, A, B>> : cc_nodes<Dict, C, A, B> {
    static constexpr long long at(long long c, int n) {
        return cc_expr<Dict, C>::at(c, n) ? cc_expr<Dict, A>::at(c, n) : cc_expr<Dict, B>::at(c, n);
    }
};

template <class Dict, class E>
struct cc_rule {
    using X = cc_expr<Dict, E>;
    static constexpr bool ok = X::ok && X::K >= 0;
    // g_q as a truth table over (cell, b0), bit 2 cell + b0 (the result's plane 0). q = 4
    // means count 8 (b0 is then 0): its b0 = 1 column repeats count 8.
    static constexpr unsigned g(int q) {
        unsigned r = 0;
        for (int c = 0; c < 2; ++c)
            for (int b = 0; b < 2; ++b) {
                const int n = 2 * q + b <= 8 ? 2 * q + b : 2 * q;
                if (X::at(c, n) & 1) r |= 1u << (2 * c + b);
            }
        return r;
    }
    // truth table over (s1, k, m1) (index 4 s1 + 2 k + m1) of "g_q == G"
    static constexpr unsigned qset(unsigned G) {
        unsigned lut = 0;
        for (int i = 0; i < 8; ++i)
            if (g(((i >> 2) & 1) + ((i >> 1) & 1) + 2 * (i & 1)) == G) lut |= 1u << i;
        return lut;
    }
};

template <class Dict, class E, unsigned G>
CELLATO_HD word cc_term(const moore_sum& ms, word cell) {
    using R = cc_rule<Dict, E>;
    if constexpr (G == 0 || R::qset(G) == 0) {
        return word(0);
    } else {
        // (q in S_G, cell, b0) -> q in S_G and G(cell, b0): rows 4 + 2 cell + b0
        return lop3<(G << 4)>(lop3<R::qset(G)>(ms.s1, ms.k, ms.m1), cell, ms.b0);
    }
}

template <class Dict, class E, class Win>
CELLATO_HD word cc_eval(const Win& w) {
    const moore_sum ms = moore_sum_of<cc_rule<Dict, E>::X::K>(w);
    const word cell = w.template nbf<0, 0>(is_index<1>{});
    return [&]<unsigned... G>(std::integer_sequence<unsigned, G...>) {
        return (cc_term<Dict, E, G>(ms, cell) | ...);
    }(std::make_integer_sequence<unsigned, 16>{});
}

// ---- rules of the cell and neighbourhood predicates (any plane count) ----------------------
// A rule whose neighbourhood enters only through boolean predicates that do not read the cell
// itself (comparisons of a state's Moore or von Neumann count with a constant, and and / or /
// not of those), around logic on the cell's own state (WireWorld, Brian's Brain, Greenberg-
// Hastings, forest fire: an if / elif chain over the cell's state, one predicate), is a table
// of the cell's dictionary index and the predicate bits. The predicates are evaluated as
// usual (from row sums where the window has them); the table is evaluated at compile time and
// each plane of the result becomes a boolean function of the P cell planes and the M predicate
// masks, built from three-input ops by Shannon expansion (3 inputs: one op; 4: three), cell
// indices past the dictionary's states left free. Greenberg-Hastings' eight-way chain becomes
// three ops per plane instead of a compare and a 3-plane select per branch.
template <class... T>
struct tlist {
    static constexpr int size = sizeof...(T);
};
template <class L, class X>
struct tl_add;
template <class... T, class X>
struct tl_add<tlist<T...>, X> {
    using type = std::conditional_t<(std::is_same_v<T, X> || ...), tlist<T...>, tlist<T..., X>>;
};
template <class L, class X>
struct tl_index;
template <class X>
struct tl_index<tlist<>, X> { static constexpr int value = -1; };
template <class H, class... T, class X>
struct tl_index<tlist<H, T...>, X> {
    static constexpr int value = std::is_same_v<H, X> ? 0 : (tl_index<tlist<T...>, X>::value < 0 ? -1 : 1 + tl_index<tlist<T...>, X>::value);
};
template <class L, int I>
struct tl_at;
template <class H, class... T>
struct tl_at<tlist<H, T...>, 0> { using type = H; };
template <class H, class... T, int I>
struct tl_at<tlist<H, T...>, I> { using type = typename tl_at<tlist<T...>, I - 1>::type; };

// pt_pred: a boolean node that reads the neighbourhood and not the cell
template <class Dict, class E>
struct pt_pred : std::false_type {};
template <class N>
struct pt_nbhd : std::false_type {};
template <> struct pt_nbhd<moore_8_neighbors> : std::true_type {};
template <> struct pt_nbhd<von_neumann_4_neighbors> : std::true_type {};
#define CELLATO_PT_CMP(node)                                                                    \
    template <class Dict, auto V, class N, auto K>                                              \
    struct pt_pred<Dict, node<count_neighbors<state_constant<V>, N>, constant<K>>> : pt_nbhd<N> {};
CELLATO_PT_CMP(equals)
CELLATO_PT_CMP(not_equals)
CELLATO_PT_CMP(greater_than)
CELLATO_PT_CMP(less_than)
#undef CELLATO_PT_CMP
template <class Dict, class L, class R>
struct pt_pred<Dict, and_<L, R>> : std::bool_constant<pt_pred<Dict, L>::value && pt_pred<Dict, R>::value> {};
template <class Dict, class L, class R>
struct pt_pred<Dict, or_<L, R>> : std::bool_constant<pt_pred<Dict, L>::value && pt_pred<Dict, R>::value> {};
template <class Dict, class V>
struct pt_pred<Dict, not_<V>> : pt_pred<Dict, V> {};

// pt_expr<Dict, E, Preds>: whether E is of the form above (with its maximal predicates in
// Preds), and its value for a cell index and predicate bits (bit i = predicate i of Preds)
template <class Dict, class E, class Preds, bool Pred = pt_pred<Dict, E>::value>
struct pt_expr {   // not of that form
    static constexpr bool ok = false;
    using preds = Preds;
    static constexpr long long at(long long, unsigned) { return 0; }
};
template <class Dict, class E, class Preds>
struct pt_expr<Dict, E, Preds, true> {   // a predicate: a leaf
    static constexpr bool ok = true;
    using preds = typename tl_add<Preds, E>::type;
    static constexpr long long at(long long, unsigned bits) { return (bits >> tl_index<preds, E>::value) & 1u; }
};
template <class Dict, auto V, class Preds>
struct pt_expr<Dict, constant<V>, Preds, false> {
    static constexpr bool ok = true;
    using preds = Preds;
    static constexpr long long at(long long, unsigned) { return static_cast<long long>(V); }
};
template <class Dict, auto V, class Preds>
struct pt_expr<Dict, state_constant<V>, Preds, false> {
    static constexpr bool ok = true;
    using preds = Preds;
    static constexpr long long at(long long, unsigned) { return index_of<Dict, V>; }
};
template <class Dict, class Preds>
struct pt_expr<Dict, neighbor_at<0, 0>, Preds, false> {
    static constexpr bool ok = true;
    using preds = Preds;
    static constexpr long long at(long long cell, unsigned) { return cell; }
};
// interior nodes: the operands in order, each seeing the predicates collected so far; `at`
// evaluates against the final list
template <class Dict, class Preds, class... A>
struct pt_chain;
template <class Dict, class Preds>
struct pt_chain<Dict, Preds> {
    static constexpr bool ok = true;
    using preds = Preds;
};
template <class Dict, class Preds, class H, class... T>
struct pt_chain<Dict, Preds, H, T...> {
    using head = pt_expr<Dict, H, Preds>;
    using tail = pt_chain<Dict, typename head::preds, T...>;
    static constexpr bool ok = head::ok && tail::ok;
    using preds = typename tail::preds;
};
template <class Dict, class E, class Final>
struct pt_at {   // E evaluated with the predicate positions of the final list
    static constexpr long long at(long long c, unsigned b) { return pt_expr<Dict, E, Final>::at(c, b); }
};
#define CELLATO_PT_BINARY(node, op)                                                             \
    template <class Dict, class L, class R, class Preds>                                        \
    struct pt_expr<Dict, node<L, R>, Preds, false> : pt_chain<Dict, Preds, L, R> {              \
        using final_preds = typename pt_chain<Dict, Preds, L, R>::preds;                        \
        static constexpr long long at(long long c, unsigned b) {                                \
            return pt_at<Dict, L, final_preds>::at(c, b) op pt_at<Dict, R, final_preds>::at(c, b) ? 1 : 0; \
        }                                                                                       \
    };
CELLATO_PT_BINARY(equals, ==)
CELLATO_PT_BINARY(not_equals, !=)
CELLATO_PT_BINARY(less_than, <)
CELLATO_PT_BINARY(greater_than, >)
CELLATO_PT_BINARY(and_, &&)
CELLATO_PT_BINARY(or_, ||)
#undef CELLATO_PT_BINARY
template <class Dict, class V, class Preds>
struct pt_expr<Dict, not_<V>, Preds, false> : pt_chain<Dict, Preds, V> {
    using final_preds = typename pt_chain<Dict, Preds, V>::preds;
    static constexpr long long at(long long c, unsigned b) { return pt_at<Dict, V, final_preds>::at(c, b) ? 0 : 1; }
};
template <class Dict, class C, class A, class B, class Preds>
struct pt_expr<Dict, if_then_else<C, A, B>, Preds, false> : pt_chain<Dict, Preds, C, A, B> {
    using final_preds = typename pt_chain<Dict, Preds, C, A, B>::preds;
    static constexpr long long at(long long c, unsigned b) {
        return pt_at<Dict, C, final_preds>::at(c, b) ? pt_at<Dict, A, final_preds>::at(c, b) : pt_at<Dict, B, final_preds>::at(c, b);
    }
};

// comparisons of the cell with a state in E (`current_state == S`)
template <class E>
struct pt_cell_tests { static constexpr int value = 0; };
template <template <class...> class N, class... A>
struct pt_cell_tests<N<A...>> { static constexpr int value = (0 + ... + pt_cell_tests<A>::value); };
template <auto V>
struct pt_cell_tests<equals<neighbor_at<0, 0>, state_constant<V>>> { static constexpr int value = 1; };

// Taken for rules that compare the cell with two states or more (measured VALU per launch at
// 16384^2, table / node by node: Greenberg-Hastings 0.55-0.57, forest fire 0.68-0.69,
// WireWorld 0.95-0.96; Brian's Brain, with two comparisons, 1.12 until the table's counts could
// include the cell (self_free below), 0.97 since: loop VALU per word and stage 30.5 -> 29.7).
template <class Dict, class E, int P>
struct pt_rule {
    using X = pt_expr<Dict, E, tlist<>>;
    using preds = typename X::preds;
    static constexpr int M = preds::size;
    static constexpr int n = P + M;   // inputs: the cell's planes, then the predicates
    static constexpr bool ok = X::ok && M >= 1 && n <= 6 && pt_cell_tests<E>::value >= 2;
    // truth table of result plane `b` (bit i: inputs i, cell index in the low P bits) and the
    // inputs that do not occur (cell indices past the dictionary)
    static constexpr unsigned long long table(int b) {
        unsigned long long t = 0;
        for (int i = 0; i < (1 << n); ++i) {
            const long long cell = i & ((1 << P) - 1);
            if (cell < Dict::number_of_values && ((pt_at<Dict, E, preds>::at(cell, static_cast<unsigned>(i >> P)) >> b) & 1))
                t |= 1ull << i;
        }
        return t;
    }
    static constexpr unsigned long long free_inputs() {
        unsigned long long f = 0;
        for (int i = 0; i < (1 << n); ++i)
            if ((i & ((1 << P) - 1)) >= Dict::number_of_values) f |= 1ull << i;
        return f;
    }
    // whether predicate i changes the result at a cell of index c (for some value of the others)
    static constexpr bool pred_read_at(int i, long long c) {
        for (unsigned rest = 0; rest < (1u << M); ++rest)
            if (!((rest >> i) & 1u) && pt_at<Dict, E, preds>::at(c, rest) != pt_at<Dict, E, preds>::at(c, rest | (1u << i)))
                return true;
        return false;
    }
    // predicate i may count the cell itself: it is read only at cells that are none of the
    // states it counts
    template <int i, long long... K>
    static constexpr bool self_free(klist<K...>) { return !(pred_read_at(i, K) || ...); }
    // every predicate that counts state K may count the cell itself
    template <long long K>
    static constexpr bool self_free_counts();
};

// the states a predicate counts (dictionary indices)
template <class Dict, class E>
struct pt_counted { using type = klist<>; };
template <class Dict, template <class...> class N, class... A>
struct pt_counted<Dict, N<A...>> { using type = typename kjoin_all<typename pt_counted<Dict, A>::type...>::type; };
template <class Dict, auto V, class N>
struct pt_counted<Dict, count_neighbors<state_constant<V>, N>> { using type = klist<index_of<Dict, V>>; };

template <class Dict, class E, int P>
template <long long K>
constexpr bool pt_rule<Dict, E, P>::self_free_counts() {
    return []<int... I>(std::integer_sequence<int, I...>) {
        return ((kindex<K, typename pt_counted<Dict, typename tl_at<preds, I>::type>::type>::value < 0 ||
                 self_free<I>(typename pt_counted<Dict, typename tl_at<preds, I>::type>::type{})) && ...);
    }(std::make_integer_sequence<int, M>{});
}

// whether every Moore count of state K in E is the test "count > 0" (any neighbour)
template <class Dict, class E, long long K>
struct any_only : std::true_type {};
template <class Dict, template <class...> class N, class... A, long long K>
struct any_only<Dict, N<A...>, K> : std::bool_constant<(any_only<Dict, A, K>::value && ...)> {};
template <class Dict, auto V, long long K>
struct any_only<Dict, count_neighbors<state_constant<V>, moore_8_neighbors>, K>
    : std::bool_constant<index_of<Dict, V> != K> {};
template <class Dict, auto V, auto Z, long long K>
struct any_only<Dict, greater_than<count_neighbors<state_constant<V>, moore_8_neighbors>, constant<Z>>, K>
    : std::bool_constant<Z == 0 || index_of<Dict, V> != K> {};

// whether state K is counted only as "some von Neumann neighbour is K" (no other count of K)
template <class Dict, class E, long long K>
struct vn_any_only : std::true_type {};
template <class Dict, template <class...> class N, class... A, long long K>
struct vn_any_only<Dict, N<A...>, K> : std::bool_constant<(vn_any_only<Dict, A, K>::value && ...)> {};
template <class Dict, auto V, class Nb, long long K>
struct vn_any_only<Dict, count_neighbors<state_constant<V>, Nb>, K> : std::bool_constant<index_of<Dict, V> != K> {};
template <class Dict, auto V, auto Z, long long K>
struct vn_any_only<Dict, greater_than<count_neighbors<state_constant<V>, von_neumann_4_neighbors>, constant<Z>>, K>
    : std::bool_constant<Z == 0 || index_of<Dict, V> != K> {};
// the states a rule counts over the von Neumann neighbourhood
template <class Dict, class E>
struct vn_targets { using type = klist<>; };
template <class Dict, template <class...> class N, class... A>
struct vn_targets<Dict, N<A...>> { using type = typename kjoin_all<typename vn_targets<Dict, A>::type...>::type; };
template <class Dict, auto V>
struct vn_targets<Dict, count_neighbors<state_constant<V>, von_neumann_4_neighbors>> { using type = klist<index_of<Dict, V>>; };

// A von Neumann "any" state whose row sums a streaming caller may keep as (3-cell OR, mask):
// read only as "some neighbour is K", through the predicate table and counting the cell
// (the own row's OR then includes the cell, which is never K where it is read).
template <class Dict, class E, int P, long long K>
struct vn_sum_ok {
    static constexpr bool value = [] {
        if constexpr (P > 1 && pt_rule<Dict, E, P>::ok)
            return vn_any_only<Dict, E, K>::value && pt_rule<Dict, E, P>::template self_free_counts<K>();
        else
            return false;
    }();
};
template <class Dict, class E, int P, class L>
struct vn_sum_list;
template <class Dict, class E, int P, long long... K>
struct vn_sum_list<Dict, E, P, klist<K...>> {
    using type = typename kjoin_all<std::conditional_t<vn_sum_ok<Dict, E, P, K>::value, klist<K>, klist<>>...>::type;
};

// A window whose own cell matches nothing: a Moore count read through it includes the cell
// (the own row's 3-cell sum is not reduced by the cell). Exact for a predicate that the rule
// reads only at cells that are not the counted state, and two ops per 32 bits cheaper.
template <class Win>
struct self_incl_window {
    const Win& w;
    int t;
    static constexpr bool row_sums = has_row_sums<Win>();
    static constexpr bool includes_cell = true;
    using vn_targets = typename win_vn<Win>::type;
    CELLATO_HD word yodd() const { return w.yodd(); }
    template <long long K, int dy>
    CELLATO_HD void rsum(word& s, word& c) const { w.template rsum<K, dy>(s, c); }
    template <int dx, int dy>
    CELLATO_HD auto nb() const { return w.template nb<dx, dy>(); }
    template <int dx, int dy, class F>
    CELLATO_HD word nbf(F f) const {
        if constexpr (dx == 0 && dy == 0) return word(0);
        else return w.template nbf<dx, dy>(f);
    }
};

template <int N, unsigned long long F, unsigned long long D>
CELLATO_HD word pt_fn(const word* in);

// A compound predicate (and / or / not) of comparisons of one state's Moore count with
// constants: its value as a function of the count, cp_eval<..>::at(n).
template <class Dict, class E>
struct cp_eval { static constexpr bool ok = false; };
#define CELLATO_CP_CMP(node, op)                                                                \
    template <class Dict, auto V, auto C>                                                       \
    struct cp_eval<Dict, node<count_neighbors<state_constant<V>, moore_8_neighbors>, constant<C>>> { \
        static constexpr bool ok = true;                                                        \
        static constexpr bool at(long long n) { return n op static_cast<long long>(C); }        \
    };
CELLATO_CP_CMP(equals, ==)
CELLATO_CP_CMP(not_equals, !=)
CELLATO_CP_CMP(greater_than, >)
CELLATO_CP_CMP(less_than, <)
#undef CELLATO_CP_CMP
template <class Dict, class L, class R>
struct cp_eval<Dict, and_<L, R>> {
    static constexpr bool ok = cp_eval<Dict, L>::ok && cp_eval<Dict, R>::ok;
    static constexpr bool at(long long n) { return cp_eval<Dict, L>::at(n) && cp_eval<Dict, R>::at(n); }
};
template <class Dict, class L, class R>
struct cp_eval<Dict, or_<L, R>> {
    static constexpr bool ok = cp_eval<Dict, L>::ok && cp_eval<Dict, R>::ok;
    static constexpr bool at(long long n) { return cp_eval<Dict, L>::at(n) || cp_eval<Dict, R>::at(n); }
};
template <class Dict, class V>
struct cp_eval<Dict, not_<V>> {
    static constexpr bool ok = cp_eval<Dict, V>::ok;
    static constexpr bool at(long long n) { return !cp_eval<Dict, V>::at(n); }
};
template <class E> struct cp_compound : std::false_type {};
template <class L, class R> struct cp_compound<and_<L, R>> : std::true_type {};
template <class L, class R> struct cp_compound<or_<L, R>> : std::true_type {};
template <class V> struct cp_compound<not_<V>> : std::true_type {};
template <class L> struct kfirst;
template <long long H, long long... T> struct kfirst<klist<H, T...>> { static constexpr long long value = H; };

// Truth table of such a predicate over the adder form of the count, inputs (b0, s1, k, m1)
// (index b0 + 2 s1 + 4 k + 8 m1; count = b0 + 2 (s1 + k + 2 m1))
template <class Dict, class Pr>
constexpr unsigned long long cp_table() {
    unsigned long long t = 0;
    for (int i = 0; i < 16; ++i) {
        const long long n = (i & 1) + 2 * (((i >> 1) & 1) + ((i >> 2) & 1) + 2 * ((i >> 3) & 1));
        if (cp_eval<Dict, Pr>::at(n)) t |= 1ull << i;
    }
    return t;
}

template <bool SelfFree, class Dict, class Pr, class Win>
CELLATO_HD word pt_pred_value(const Win& w) {
    if constexpr (has_row_sums<Win>() && cp_compound<Pr>::value && cp_eval<Dict, Pr>::ok &&
                  pt_counted<Dict, Pr>::type::size == 1) {
        // one adder form of the count, then the predicate as one function of its four words
        // (WireWorld's "1 or 2 heads": three ops per 32 bits instead of five)
        constexpr long long K = kfirst<typename pt_counted<Dict, Pr>::type>::value;
        moore_sum ms;
        if constexpr (SelfFree) ms = moore_sum_of<K>(self_incl_window<Win>{w, w.t});
        else ms = moore_sum_of<K>(w);
        const word in[4] = {ms.b0, ms.s1, ms.k, ms.m1};
        return pt_fn<4, cp_table<Dict, Pr>(), 0ull>(in);
    } else if constexpr (SelfFree) {
        return ev<Dict, Pr>::at(self_incl_window<Win>{w, w.t});
    } else {
        return ev<Dict, Pr>::at(w);
    }
}

// A boolean function of the first N inputs in[0..N-1] (truth table F, free entries D) from
// three-input ops: one op for N <= 3, else a select on the last input between the two halves.
template <int N, unsigned long long F, unsigned long long D>
CELLATO_HD word pt_fn(const word* in) {
    constexpr unsigned long long all = N >= 6 ? ~0ull : ((1ull << (1 << N)) - 1);
    if constexpr (((F & ~D) & all) == 0) {
        return word(0);
    } else if constexpr (((F | D) & all) == all) {
        return ~word(0);
    } else if constexpr (N <= 3) {
        // free entries take 0; the LUT over (in[2], in[1], in[0]) = index 4 a + 2 b + c
        constexpr unsigned lut = [] {
            unsigned l = 0;
            for (int i = 0; i < 8; ++i) {
                const int j = i & ((1 << N) - 1);   // inputs past N do not matter
                if ((F >> j) & 1 & ~((D >> j) & 1)) l |= 1u << i;
            }
            return l;
        }();
        const word a = N >= 3 ? in[2] : word(0), b = N >= 2 ? in[1] : word(0), c = in[0];
        return lop3<lut>(a, b, c);
    } else {
        constexpr int h = 1 << (N - 1);
        constexpr unsigned long long lo_mask = (1ull << h) - 1;
        constexpr unsigned long long F0 = F & lo_mask, F1 = (F >> h) & lo_mask;
        constexpr unsigned long long D0 = D & lo_mask, D1 = (D >> h) & lo_mask;
        if constexpr (((F0 ^ F1) & ~(D0 | D1) & lo_mask) == 0) {
            // the last input does not matter (entries free in one half follow the other)
            return pt_fn<N - 1, (F0 & ~D0) | (F1 & ~D1), D0 & D1>(in);
        } else {
            const word f1 = pt_fn<N - 1, F1, D1>(in);
            const word f0 = pt_fn<N - 1, F0, D0>(in);
            return lop3<0xCA>(in[N - 1], f1, f0);   // in[N-1] ? f1 : f0
        }
    }
}

template <class Dict, class E, int P, class Win, class... Pr>
CELLATO_HD vint<P> pt_eval(const Win& w, tlist<Pr...>) {
    using R = pt_rule<Dict, E, P>;
    word in[P + sizeof...(Pr)];
    const vint<P> cell = w.template nb<0, 0>();
#pragma unroll
    for (int i = 0; i < P; ++i) in[i] = cell.p[i];
    [&]<int... I>(std::integer_sequence<int, I...>) {
        ((in[P + I] = pt_pred_value<R::template self_free<I>(typename pt_counted<Dict, Pr>::type{}), Dict, Pr>(w)), ...);
    }(std::make_integer_sequence<int, sizeof...(Pr)>{});
    vint<P> r;
    [&]<int... B>(std::integer_sequence<int, B...>) {
        ((r.p[B] = pt_fn<R::n, R::table(B), R::free_inputs()>(in)), ...);
    }(std::make_integer_sequence<int, P>{});
    return r;
}

// The reference's evaluator name: evaluator<word, dictionary, Expression>
// (evaluators/bit_planes.hpp:48-49). Only 64-bit words are implemented on the GPU.
template <typename cell_row_type, typename state_dictionary_type, typename Expression>
struct evaluator {
    static_assert(std::is_same_v<cell_row_type, word>, "the HIP traversers use 64-bit words");
    using dictionary = state_dictionary_type;
    using expression = Expression;
    static constexpr int planes = state_dictionary_type::needed_bits;
    // grid layout the evaluator reads (0 linear, 1 tiled 8x8): the HIP traversers check that
    // it matches the grid's
    static constexpr int layout = 0;
    // von Neumann "any" states whose row sums a streaming caller may keep as (3-cell OR, mask)
    using vn_sum_targets = typename vn_sum_list<state_dictionary_type, Expression, planes,
                                                typename vn_targets<state_dictionary_type, Expression>::type>::type;
    // states whose Moore-neighbourhood count a streaming caller keeps per-row sums of, then
    // the von Neumann "any" states above
    using row_sum_targets =
        typename kjoin<typename row_targets<state_dictionary_type, Expression>::type, vn_sum_targets>::type;
    // a counted state read only as "some neighbour is K", through the predicate table and
    // counting the cell itself: its row sum may be kept as the 3-cell OR alone (the sum word
    // holding the OR, the carry word 0), one word per ring slot instead of two
    template <long long K>
    static constexpr bool row_sum_vn_any = kindex<K, vn_sum_targets>::value >= 0;
    template <long long K>
    static constexpr bool row_sum_or_only = [] {
        if constexpr (planes > 1 && pt_rule<state_dictionary_type, Expression, planes>::ok)
            return any_only<state_dictionary_type, Expression, K>::value &&
                   pt_rule<state_dictionary_type, Expression, planes>::template self_free_counts<K>();
        else
            return false;
    }();
    // the rule reads the parity of the time step (alternating algorithms, Margolus blocks):
    // a streaming caller may then pass a window whose `t` is only that parity
    static constexpr bool reads_step_parity =
        (rule_traits<Expression>::value & (TRAIT_ALTERNATING | TRAIT_MARGOLUS)) != 0;

    template <class Win>
    CELLATO_HD static vint<planes> evaluate(const Win& w) {
        if constexpr (planes == 1 && mg_rule<state_dictionary_type, Expression>::ok) {
            vint<1> r;
            r.p[0] = mg_eval<state_dictionary_type, Expression>(w);
            return r;
        } else if constexpr (planes == 1 && has_row_sums<Win>() && cc_rule<state_dictionary_type, Expression>::ok) {
            vint<1> r;
            r.p[0] = cc_eval<state_dictionary_type, Expression>(w);
            return r;
        } else if constexpr (planes > 1 && pt_rule<state_dictionary_type, Expression, planes>::ok) {
            return pt_eval<state_dictionary_type, Expression, planes>(
                w, typename pt_rule<state_dictionary_type, Expression, planes>::preds{});
        } else {
            return resize<planes>(as_vint(eval<state_dictionary_type, Expression>(w)));
        }
    }
};

} // namespace cellato::evaluators::bit_planes

#endif // CELLATO_EVALUATORS_BIT_PLANES_HPP
