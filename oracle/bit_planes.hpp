// ORACLE — test infrastructure only (see oracle/standard.hpp for the rules of use).
//
// CPU restatement of the reference's bit_planes evaluator (include/evaluators/bit_planes.hpp)
// over the linear bit-plane layout (include/memory/bit_planes_grid.hpp): one word holds
// `bits(Wd)` horizontally adjacent cells of one plane, cell x -> bit x % bits of word x / bits.
// The bit-sliced integer follows include/core/vector_int.hpp: additions keep the wider
// operand's width (:227-286), constants take highest_bit + 2 planes (:752-776), bitwise
// and/or zero-extend (:687-720), and a result narrower than the grid stores only its own
// planes (:411-421).  Used for the reference's "CPU bit_planes" path (one of the two sides
// of the 1024^2 Game-of-Life equality check) and as a fast CPU cross-check.
#ifndef CELLATO_ORACLE_BIT_PLANES_HPP
#define CELLATO_ORACLE_BIT_PLANES_HPP

#include <array>
#include <cstdint>
#include <thread>
#include <vector>
#include <algorithm>

#include "cellato/core/ast.hpp"

namespace oracle::bit_planes {

using namespace cellato::ast;

template <class Wd>
constexpr int word_bits = static_cast<int>(sizeof(Wd) * 8);

template <class Wd, int N>
struct vint {
    static constexpr int bits = N;
    std::array<Wd, N> p{};
};

constexpr int highest_bit(long long v) { return v <= 0 ? -1 : 1 + highest_bit(v >> 1); }

// vector_int_factory::from_constant (vector_int.hpp:752-776)
template <class Wd, long long V>
constexpr auto from_constant() {
    constexpr int hb = highest_bit(V) + 1;
    if constexpr (hb == 0) {
        return vint<Wd, 1>{};
    } else {
        vint<Wd, hb + 1> r;
        for (int i = 0; i < hb + 1; ++i) r.p[i] = ((V >> i) & 1) ? static_cast<Wd>(~Wd(0)) : Wd(0);
        return r;
    }
}

template <class Wd, int A, int B>
vint<Wd, (A > B ? A : B)> add(const vint<Wd, A>& a, const vint<Wd, B>& b) {   // :227-286
    constexpr int R = A > B ? A : B;
    vint<Wd, R> r;
    Wd carry = 0;
    for (int i = 0; i < R; ++i) {
        const Wd x = i < A ? a.p[i] : Wd(0);
        const Wd y = i < B ? b.p[i] : Wd(0);
        r.p[i] = x ^ y ^ carry;
        carry = (x & y) | (carry & (x ^ y));
    }
    return r;
}

template <int K, class Wd, int N>
vint<Wd, K> resize(const vint<Wd, N>& a) {                                     // :323-334
    vint<Wd, K> r;
    for (int i = 0; i < K && i < N; ++i) r.p[i] = a.p[i];
    return r;
}

template <class Wd, int A, int B>
Wd equals_to(const vint<Wd, A>& a, const vint<Wd, B>& b) {                    // :423-458
    Wd r = static_cast<Wd>(~Wd(0));
    for (int i = 0; i < (A > B ? A : B); ++i) {
        const Wd x = i < A ? a.p[i] : Wd(0);
        const Wd y = i < B ? b.p[i] : Wd(0);
        r &= static_cast<Wd>(~(x ^ y));
    }
    return r;
}

template <int K, class Wd, int A>
Wd equals_const(const vint<Wd, A>& a) {                                      // :483-504
    Wd r = static_cast<Wd>(~Wd(0));
    for (int i = 0; i < A; ++i) r &= ((K >> i) & 1) ? a.p[i] : static_cast<Wd>(~a.p[i]);
    return r;
}

// a > b per column: scan from the most significant plane (:506-547)
template <class Wd, int A, int B>
Wd greater(const vint<Wd, A>& a, const vint<Wd, B>& b) {
    Wd res = 0, decided = 0;
    for (int i = (A > B ? A : B) - 1; i >= 0; --i) {
        const Wd x = i < A ? a.p[i] : Wd(0);
        const Wd y = i < B ? b.p[i] : Wd(0);
        res |= static_cast<Wd>(~decided) & (x ^ y) & x;
        decided |= x ^ y;
    }
    return res;
}

template <class Wd, int A, int B>
Wd less(const vint<Wd, A>& a, const vint<Wd, B>& b) { return greater(b, a); }   // :549-590

template <class Wd, int N>
vint<Wd, N> mask(const vint<Wd, N>& a, Wd m) {                                // :309-321
    vint<Wd, N> r;
    for (int i = 0; i < N; ++i) r.p[i] = a.p[i] & m;
    return r;
}

template <class Wd, int A, int B, class Op>
vint<Wd, (A > B ? A : B)> zip(const vint<Wd, A>& a, const vint<Wd, B>& b, Op op) {   // :687-720
    constexpr int R = A > B ? A : B;
    vint<Wd, R> r;
    for (int i = 0; i < R; ++i) r.p[i] = op(i < A ? a.p[i] : Wd(0), i < B ? b.p[i] : Wd(0));
    return r;
}

template <class Wd, int N>
vint<Wd, N> swap_pairs(const vint<Wd, N>& a) {                                // :592-607
    constexpr Wd m = static_cast<Wd>(0x5555555555555555ULL);
    vint<Wd, N> r;
    for (int i = 0; i < N; ++i) r.p[i] = static_cast<Wd>(((a.p[i] & m) << 1) | ((a.p[i] & static_cast<Wd>(~m)) >> 1));
    return r;
}

template <class Wd, int N>
vint<Wd, N> shl(const vint<Wd, N>& a, int s) {
    vint<Wd, N> r;
    for (int i = 0; i < N; ++i) r.p[i] = static_cast<Wd>(a.p[i] << s);
    return r;
}

template <class Wd, int P>
struct state {
    std::array<const Wd*, P> grid;
    int Wx, H;    // physical size: words per row, rows
    int x, y;     // word position
    int t;
    int idx(int xx, int yy) const { return ((yy + H) % H) * Wx + (xx + Wx) % Wx; }
    vint<Wd, P> load(int i) const {
        vint<Wd, P> r;
        for (int b = 0; b < P; ++b) r.p[b] = grid[b][i];
        return r;
    }
};

template <class Cfg, class Wd, class E>
struct eval;

template <class Cfg, class Wd>
using st = state<Wd, Cfg::state_dictionary::needed_bits>;

template <class Cfg, class Wd, auto V>
struct eval<Cfg, Wd, constant<V>> {                                           // bit_planes.hpp:58-63
    static auto at(const st<Cfg, Wd>&) { return from_constant<Wd, static_cast<long long>(V)>(); }
};

template <class Cfg, class Wd, auto V>
struct eval<Cfg, Wd, state_constant<V>> {                                     // :65-73
    static auto at(const st<Cfg, Wd>&) {
        constexpr int idx = Cfg::state_dictionary::state_to_index(V);
        return from_constant<Wd, idx>();
    }
};

template <class Cfg, class Wd, class Even, class Odd>
struct eval<Cfg, Wd, alternate_algorithms<Even, Odd>> {                       // :75-85
    static auto at(const st<Cfg, Wd>& s) {
        if (s.t % 2 == 0) return eval<Cfg, Wd, Even>::at(s);
        return eval<Cfg, Wd, Odd>::at(s);
    }
};

template <class T> struct is_vint : std::false_type {};
template <class Wd, int N> struct is_vint<vint<Wd, N>> : std::true_type {};

template <class Cfg, class Wd, class C, class A, class B>
struct eval<Cfg, Wd, if_then_else<C, A, B>> {                                 // :87-107
    static auto at(const st<Cfg, Wd>& s) {
        const Wd c = eval<Cfg, Wd, C>::at(s);
        auto a = eval<Cfg, Wd, A>::at(s);
        auto b = eval<Cfg, Wd, B>::at(s);
        if constexpr (is_vint<decltype(a)>::value) {
            return zip(mask(a, c), mask(b, static_cast<Wd>(~c)), [](Wd x, Wd y) { return static_cast<Wd>(x | y); });
        } else {
            return static_cast<Wd>((a & c) | (b & ~c));
        }
    }
};

template <class Cfg, class Wd, class L, class R>
struct eval<Cfg, Wd, bit_and_<L, R>> {                                        // :109-125
    static auto at(const st<Cfg, Wd>& s) {
        auto l = eval<Cfg, Wd, L>::at(s);
        auto r = eval<Cfg, Wd, R>::at(s);
        if constexpr (is_vint<decltype(l)>::value)
            return zip(l, r, [](Wd x, Wd y) { return static_cast<Wd>(x & y); });
        else
            return static_cast<Wd>(l & r);
    }
};

template <class Cfg, class Wd, class L, class R>
struct eval<Cfg, Wd, bit_or_<L, R>> {                                         // :127-143
    static auto at(const st<Cfg, Wd>& s) {
        auto l = eval<Cfg, Wd, L>::at(s);
        auto r = eval<Cfg, Wd, R>::at(s);
        if constexpr (is_vint<decltype(l)>::value)
            return zip(l, r, [](Wd x, Wd y) { return static_cast<Wd>(x | y); });
        else
            return static_cast<Wd>(l | r);
    }
};

template <class Cfg, class Wd, class L, class R>
struct eval<Cfg, Wd, plus<L, R>> {                                            // :145-161
    static auto at(const st<Cfg, Wd>& s) { return add(eval<Cfg, Wd, L>::at(s), eval<Cfg, Wd, R>::at(s)); }
};

template <class Cfg, class Wd, class L, int M>
struct eval<Cfg, Wd, modulo<L, constant<M>>> {                                // :163-182
    static_assert((M & (M - 1)) == 0, "only modulo by a power of two");
    static auto at(const st<Cfg, Wd>& s) { return resize<highest_bit(M)>(eval<Cfg, Wd, L>::at(s)); }
};

template <class Cfg, class Wd, class V>
struct eval<Cfg, Wd, not_<V>> {                                               // :184-194
    static Wd at(const st<Cfg, Wd>& s) { return static_cast<Wd>(~static_cast<Wd>(eval<Cfg, Wd, V>::at(s))); }
};
template <class Cfg, class Wd, class L, class R>
struct eval<Cfg, Wd, and_<L, R>> {                                            // :196-208
    static Wd at(const st<Cfg, Wd>& s) { return static_cast<Wd>(eval<Cfg, Wd, L>::at(s) & eval<Cfg, Wd, R>::at(s)); }
};
template <class Cfg, class Wd, class L, class R>
struct eval<Cfg, Wd, or_<L, R>> {                                             // :210-222
    static Wd at(const st<Cfg, Wd>& s) { return static_cast<Wd>(eval<Cfg, Wd, L>::at(s) | eval<Cfg, Wd, R>::at(s)); }
};

template <class Cfg, class Wd, class L, class R>
struct eval<Cfg, Wd, equals<L, R>> {                                          // :224-236
    static Wd at(const st<Cfg, Wd>& s) { return equals_to(eval<Cfg, Wd, L>::at(s), eval<Cfg, Wd, R>::at(s)); }
};
template <class Cfg, class Wd, class L, class R>
struct eval<Cfg, Wd, not_equals<L, R>> {       // a full mask word (the reference narrows it to bool, :272)
    static Wd at(const st<Cfg, Wd>& s) { return static_cast<Wd>(~equals_to(eval<Cfg, Wd, L>::at(s), eval<Cfg, Wd, R>::at(s))); }
};
template <class Cfg, class Wd, class L, class R>
struct eval<Cfg, Wd, greater_than<L, R>> {                                    // :238-250
    static Wd at(const st<Cfg, Wd>& s) { return greater(eval<Cfg, Wd, L>::at(s), eval<Cfg, Wd, R>::at(s)); }
};
template <class Cfg, class Wd, class L, class R>
struct eval<Cfg, Wd, less_than<L, R>> {                                       // :252-264
    static Wd at(const st<Cfg, Wd>& s) { return less(eval<Cfg, Wd, L>::at(s), eval<Cfg, Wd, R>::at(s)); }
};

template <class Cfg, class Wd, class V, int b>
struct eval<Cfg, Wd, has_bit_set<V, b>> {                                     // :280-290
    static Wd at(const st<Cfg, Wd>& s) { return eval<Cfg, Wd, V>::at(s).p[b]; }
};

// neighbor_at: centre word of row y+dy, funnel-shifted with the word at x+dx (:293-365)
template <class Cfg, class Wd, int dx, int dy>
struct eval<Cfg, Wd, neighbor_at<dx, dy>> {
    static constexpr int P = Cfg::state_dictionary::needed_bits;
    static vint<Wd, P> at(const st<Cfg, Wd>& s) {
        auto c = s.load(s.idx(s.x, s.y + dy));
        if constexpr (dx == 0) {
            return c;
        } else {
            auto n = s.load(s.idx(s.x + dx, s.y + dy));
            constexpr int w = word_bits<Wd>;
            vint<Wd, P> r;
            for (int b = 0; b < P; ++b) {
                if constexpr (dx > 0) r.p[b] = static_cast<Wd>((c.p[b] >> dx) | (n.p[b] << (w - dx)));
                else                  r.p[b] = static_cast<Wd>((c.p[b] << -dx) | (n.p[b] >> (w + dx)));
            }
            return r;
        }
    }
};

template <class Wd>
vint<Wd, 1> cond(Wd m) { vint<Wd, 1> r; r.p[0] = m; return r; }

// Moore count with a state-constant target (:508-551) or an expression target (:553-596).
template <class Cfg, class Wd, class C>
struct eval<Cfg, Wd, count_neighbors<C, moore_8_neighbors>> {
    template <int dx, int dy>
    static Wd m(const st<Cfg, Wd>& s, const auto& target) {
        return equals_to(eval<Cfg, Wd, neighbor_at<dx, dy>>::at(s), target);
    }
    static vint<Wd, 4> at(const st<Cfg, Wd>& s) {
        const auto target = eval<Cfg, Wd, C>::at(s);
        auto a = resize<2>(cond(m<-1, -1>(s, target)));
        a = add(a, cond(m<0, -1>(s, target)));
        a = add(a, cond(m<1, -1>(s, target)));
        auto b = resize<3>(a);
        b = add(b, cond(m<-1, 0>(s, target)));
        b = add(b, cond(m<1, 0>(s, target)));
        b = add(b, cond(m<-1, 1>(s, target)));
        b = add(b, cond(m<0, 1>(s, target)));
        auto c = resize<4>(b);
        return add(c, cond(m<1, 1>(s, target)));
    }
};

// greater_than<moore count, 0> as an OR of the eight equality masks (:598-628).
template <class Cfg, class Wd, class C>
struct eval<Cfg, Wd, greater_than<count_neighbors<C, moore_8_neighbors>, constant<0>>> {
    static Wd at(const st<Cfg, Wd>& s) {
        const auto target = eval<Cfg, Wd, C>::at(s);
        Wd r = 0;
        r |= equals_to(eval<Cfg, Wd, neighbor_at<-1, -1>>::at(s), target);
        r |= equals_to(eval<Cfg, Wd, neighbor_at<0, -1>>::at(s), target);
        r |= equals_to(eval<Cfg, Wd, neighbor_at<1, -1>>::at(s), target);
        r |= equals_to(eval<Cfg, Wd, neighbor_at<-1, 0>>::at(s), target);
        r |= equals_to(eval<Cfg, Wd, neighbor_at<1, 0>>::at(s), target);
        r |= equals_to(eval<Cfg, Wd, neighbor_at<-1, 1>>::at(s), target);
        r |= equals_to(eval<Cfg, Wd, neighbor_at<0, 1>>::at(s), target);
        r |= equals_to(eval<Cfg, Wd, neighbor_at<1, 1>>::at(s), target);
        return r;
    }
};

// von Neumann count, top/right/bottom/left (:630-663)
template <class Cfg, class Wd, class C>
struct eval<Cfg, Wd, count_neighbors<C, von_neumann_4_neighbors>> {
    static vint<Wd, 3> at(const st<Cfg, Wd>& s) {
        const auto target = eval<Cfg, Wd, C>::at(s);
        auto a = resize<2>(cond(equals_to(eval<Cfg, Wd, neighbor_at<0, -1>>::at(s), target)));
        a = add(a, cond(equals_to(eval<Cfg, Wd, neighbor_at<1, 0>>::at(s), target)));
        a = add(a, cond(equals_to(eval<Cfg, Wd, neighbor_at<0, 1>>::at(s), target)));
        auto b = resize<3>(a);
        return add(b, cond(equals_to(eval<Cfg, Wd, neighbor_at<-1, 0>>::at(s), target)));
    }
};

// Margolus block count (:367-455). Even step: pair swap of the column sums. Odd step: the
// same sums taken one column to the right, shifted left by one, and bit 0 patched from
// the block that straddles the word's left edge.
template <class Cfg, class Wd, auto V>
struct eval<Cfg, Wd, count_neighbors<state_constant<V>, margolus_alternating_neighborhood>> {
    static constexpr int k = Cfg::state_dictionary::state_to_index(V);
    template <int dx, int dy>
    static Wd m(const st<Cfg, Wd>& s) { return equals_const<k>(eval<Cfg, Wd, neighbor_at<dx, dy>>::at(s)); }

    static vint<Wd, 3> at(const st<Cfg, Wd>& s) {
        const bool yodd = (s.y % 2) != 0;
        if (s.t % 2 == 0) {
            const Wd cur = m<0, 0>(s);
            const Wd ver = yodd ? m<0, -1>(s) : m<0, 1>(s);
            auto col = add(resize<2>(cond(cur)), cond(ver));
            return add(resize<3>(col), swap_pairs(col));
        }
        const Wd cur = m<1, 0>(s);
        const Wd ver = yodd ? m<1, 1>(s) : m<1, -1>(s);
        auto col = add(resize<2>(cond(cur)), cond(ver));
        auto r = shl(add(resize<3>(col), swap_pairs(col)), 1);
        // bit 0: block made of columns x-1 and x (:438-451)
        const Wd lc = m<-1, 0>(s);
        const Wd lv = yodd ? m<-1, 1>(s) : m<-1, -1>(s);
        static constexpr int ones[4] = {0, 1, 1, 2};
        const int last = ones[lv & 3] + ones[lc & 3];
        for (int b = 0; b < 3; ++b) r.p[b] = static_cast<Wd>((r.p[b] & ~Wd(1)) | static_cast<Wd>((last >> b) & 1));
        return r;
    }
};

// Diagonally opposite cell of the Margolus block (:457-506).
template <class Cfg, class Wd>
struct eval<Cfg, Wd, margolus_180_neighbor> {
    static auto at(const st<Cfg, Wd>& s) {
        const bool yodd = (s.y % 2) != 0;
        if (s.t % 2 == 0) {
            auto v = yodd ? eval<Cfg, Wd, neighbor_at<0, -1>>::at(s) : eval<Cfg, Wd, neighbor_at<0, 1>>::at(s);
            return swap_pairs(v);
        }
        auto ev = yodd ? eval<Cfg, Wd, neighbor_at<1, 1>>::at(s) : eval<Cfg, Wd, neighbor_at<1, -1>>::at(s);
        auto od = yodd ? eval<Cfg, Wd, neighbor_at<-1, 1>>::at(s) : eval<Cfg, Wd, neighbor_at<-1, -1>>::at(s);
        constexpr Wd odd_bits = static_cast<Wd>(0xAAAAAAAAAAAAAAAAULL);
        constexpr Wd even_bits = static_cast<Wd>(0x5555555555555555ULL);
        return zip(mask(ev, odd_bits), mask(od, even_bits), [](Wd x, Wd y) { return static_cast<Wd>(x | y); });
    }
};

// CPU simple traverser over words (cpu/simple.hpp:60-85 with the bit_planes evaluator).
// planes: P consecutive arrays of Wx*H words. Writes min(P, result bits) planes.
template <class Cfg, class Wd>
void step_rows(const Wd* in, Wd* out, int Wx, int H, int t, int y0, int y1) {
    constexpr int P = Cfg::state_dictionary::needed_bits;
    const size_t plane = static_cast<size_t>(Wx) * H;
    st<Cfg, Wd> s;
    for (int b = 0; b < P; ++b) s.grid[b] = in + b * plane;
    s.Wx = Wx; s.H = H; s.t = t;
    for (int y = y0; y < y1; ++y) {
        s.y = y;
        for (int x = 0; x < Wx; ++x) {
            s.x = x;
            auto r = eval<Cfg, Wd, typename Cfg::algorithm>::at(s);
            constexpr int R = decltype(r)::bits;
            for (int b = 0; b < (R < P ? R : P); ++b) out[b * plane + static_cast<size_t>(y) * Wx + x] = r.p[b];
        }
    }
}

template <class Cfg, class Wd>
void run(Wd* planes, int Wx, int H, int t0, int steps, int threads) {
    constexpr int P = Cfg::state_dictionary::needed_bits;
    const size_t n = static_cast<size_t>(P) * Wx * H;
    // the second buffer starts as a copy, as the reference's traverser copies the input
    std::vector<Wd> tmp(planes, planes + n);
    Wd* cur = planes;
    Wd* nxt = tmp.data();
    threads = std::max(1, std::min(threads, H));
    for (int k = 0; k < steps; ++k) {
        if (threads == 1) {
            step_rows<Cfg, Wd>(cur, nxt, Wx, H, t0 + k, 0, H);
        } else {
            std::vector<std::thread> pool;
            for (int i = 0; i < threads; ++i) {
                const int y0 = static_cast<int>(static_cast<long long>(H) * i / threads);
                const int y1 = static_cast<int>(static_cast<long long>(H) * (i + 1) / threads);
                pool.emplace_back(step_rows<Cfg, Wd>, cur, nxt, Wx, H, t0 + k, y0, y1);
            }
            for (auto& th : pool) th.join();
        }
        std::swap(cur, nxt);
    }
    if (cur != planes) std::copy(cur, cur + n, planes);
}

} // namespace oracle::bit_planes

#endif
