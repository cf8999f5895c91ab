// State dictionaries: map a rule's cell-state enum onto the dense index that is stored in
// bit planes. Plane b holds bit b of the *index*, not of the enum value (for forest fire the
// two differ: enum ash=2, fire=3 but the dictionary lists fire before ash).
//
// Interface of the reference's include/memory/state_dictionary.hpp:10-70 (same names and
// meaning: number_of_values, needed_bits = floor(log2(n-1)) + 1, state_to_index,
// index_to_state, which throw std::out_of_range for unknown values).
#ifndef CELLATO_MEMORY_STATE_DICTIONARY_HPP
#define CELLATO_MEMORY_STATE_DICTIONARY_HPP

#include <stdexcept>

namespace cellato::memory::grids {

namespace detail {
constexpr int floor_log2(int n) { return n < 2 ? 0 : 1 + floor_log2(n / 2); }
}

template <auto... states>
class state_dictionary {
public:
    using index_t = int;
    using state_t = decltype((states, ...));

    static constexpr index_t number_of_values = sizeof...(states);
    static constexpr index_t needed_bits = detail::floor_log2(number_of_values - 1) + 1;

    static constexpr index_t state_to_index(state_t s) {
        constexpr state_t table[] = {states...};
        for (index_t i = 0; i < number_of_values; ++i)
            if (table[i] == s) return i;
        throw std::out_of_range("State not found in dictionary");
    }

    static constexpr state_t index_to_state(index_t i) {
        constexpr state_t table[] = {states...};
        if (i < 0 || i >= number_of_values) throw std::out_of_range("Index out of range");
        return table[i];
    }
};

// Integer states 0 .. 2^bits - 1, stored as themselves.
template <int bits>
class int_based_state_dictionary {
public:
    using index_t = int;
    using state_t = int;

    static constexpr index_t number_of_values = 1 << bits;
    static constexpr index_t needed_bits = bits;

    static constexpr index_t state_to_index(state_t s) {
        if (s < 0 || s >= number_of_values) throw std::out_of_range("State not found in dictionary");
        return s;
    }

    static constexpr state_t index_to_state(index_t i) {
        if (i < 0 || i >= number_of_values) throw std::out_of_range("Index out of range");
        return i;
    }
};

} // namespace cellato::memory::grids

#endif // CELLATO_MEMORY_STATE_DICTIONARY_HPP
