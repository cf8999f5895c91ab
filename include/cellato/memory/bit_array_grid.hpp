// The bit_array layout: the reference's memory::grids::bit_array::grid<dictionary, word>
// (include/memory/bit_array_grid.hpp:85-321 there), with its members (bits_per_cell,
// cells_per_word, cell_mask, states_dict_t, store_type, x/y_size_logical(), x/y_size_physical(),
// get_cell(), to_original_representation(), to_standard(), to_cuda(), to_cpu(),
// free_cuda_memory(), raw_data()) and its bit order: cells are numbered row-major, cell i lives
// in word i / cells_per_word at bits [(i % cells_per_word) * P, + P) as its dictionary index,
// and the width must be a multiple of cells_per_word (so that every row starts a word).
// Words of 32 or 64 bits (the reference's default is 32). Host code; the device copy is one
// hipMalloc'd array.
#ifndef CELLATO_MEMORY_BIT_ARRAY_GRID_HPP
#define CELLATO_MEMORY_BIT_ARRAY_GRID_HPP

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "bit_planes_grid.hpp"   // device, hip_check
#include "standard_grid.hpp"

namespace cellato::memory::grids::bit_array {

template <typename states_dictionary_t, typename store_word_type = std::uint32_t, device dev = device::CPU>
class grid {
public:
    static_assert(sizeof(store_word_type) == 4 || sizeof(store_word_type) == 8, "32- or 64-bit words");
    using states_dict_t = states_dictionary_t;
    static constexpr int bits_per_cell = states_dict_t::needed_bits;
    static constexpr int needed_bits = bits_per_cell;
    static constexpr int cells_per_word = static_cast<int>(sizeof(store_word_type) * 8) / bits_per_cell;
    static constexpr store_word_type cell_mask = (store_word_type(1) << bits_per_cell) - 1;
    static constexpr int layout = 2;   // the HIP traversers' id of this layout
    static constexpr bool is_bit_array = true;
    using original_state_t = typename states_dict_t::state_t;
    using cell_t = original_state_t;
    using store_type = store_word_type;

    grid() = default;

    grid(std::size_t y_size, std::size_t x_size) : _x(x_size), _y(y_size) {
        check_width(x_size);
        if constexpr (dev == device::CPU) _host.assign(words(), 0);
    }

    grid(std::size_t y_size, std::size_t x_size, const original_state_t* cells) requires (dev == device::CPU)
        : grid(y_size, x_size) {
        for (std::size_t i = 0; i < _x * _y; ++i) set_index(i, cells[i]);
    }

    // from a standard grid of state values; implicit, as the reference's
    grid(const standard::grid<original_state_t>& s) requires (dev == device::CPU)
        : grid(s.y_size_physical(), s.x_size_physical(), s.data()) {}

    std::size_t x_size_logical() const { return _x; }
    std::size_t y_size_logical() const { return _y; }
    std::size_t x_size_physical() const { return _x / cells_per_word; }
    std::size_t y_size_physical() const { return _y; }
    std::size_t x_size_original() const { return _x; }
    std::size_t y_size_original() const { return _y; }
    std::size_t words() const { return x_size_physical() * _y; }

    store_word_type* raw_data() const {
        if constexpr (dev == device::HIP) return _dev;
        else return const_cast<store_word_type*>(_host.data());
    }

    original_state_t get_cell(std::size_t x, std::size_t y) const requires (dev == device::CPU) {
        if (x >= _x || y >= _y) throw std::out_of_range("Cell coordinates out of range");
        return get_index(y * _x + x);
    }

    std::vector<original_state_t> to_original_representation() const requires (dev == device::CPU) {
        std::vector<original_state_t> r(_x * _y);
        for (std::size_t i = 0; i < r.size(); ++i) r[i] = get_index(i);
        return r;
    }

    standard::grid<original_state_t> to_standard() const requires (dev == device::CPU) {
        return standard::grid<original_state_t>(to_original_representation(), static_cast<int>(_x), static_cast<int>(_y));
    }

    grid<states_dict_t, store_word_type, device::HIP> to_cuda() const requires (dev == device::CPU) {
        grid<states_dict_t, store_word_type, device::HIP> g;
        g._x = _x;
        g._y = _y;
        hip_check(hipMalloc(reinterpret_cast<void**>(&g._dev), _host.size() * sizeof(store_word_type)), "hipMalloc");
        hip_check(hipMemcpy(g._dev, _host.data(), _host.size() * sizeof(store_word_type), hipMemcpyHostToDevice),
                  "hipMemcpy H2D");
        return g;
    }
    grid<states_dict_t, store_word_type, device::HIP> to_hip() const requires (dev == device::CPU) { return to_cuda(); }

    grid<states_dict_t, store_word_type, device::CPU> to_cpu() const requires (dev == device::HIP) {
        grid<states_dict_t, store_word_type, device::CPU> g(_y, _x);
        hip_check(hipMemcpy(g._host.data(), _dev, g._host.size() * sizeof(store_word_type), hipMemcpyDeviceToHost),
                  "hipMemcpy D2H");
        return g;
    }

    void free_cuda_memory() {
        if constexpr (dev == device::HIP) {
            if (_dev) hip_check(hipFree(_dev), "hipFree");
            _dev = nullptr;
        }
    }
    void free_device_memory() { free_cuda_memory(); }

    std::string get_checksum() const requires (dev == device::CPU) { return to_standard().get_checksum(); }

private:
    template <typename, typename, device> friend class grid;

    void check_width(std::size_t x_size) const {
        if (x_size % cells_per_word != 0)
            throw std::invalid_argument("Grid x_size must be divisible by cells_per_word. x_size=" + std::to_string(x_size) +
                                        ", cells_per_word=" + std::to_string(cells_per_word) + " (bits [" +
                                        std::to_string(sizeof(store_word_type) * 8) + "] / bits_per_cell [" +
                                        std::to_string(bits_per_cell) + "])");
    }
    void set_index(std::size_t i, original_state_t s) {
        const std::size_t w = i / cells_per_word, sh = (i % cells_per_word) * bits_per_cell;
        const auto idx = static_cast<store_word_type>(states_dict_t::state_to_index(s));
        _host[w] = (_host[w] & ~(cell_mask << sh)) | ((idx & cell_mask) << sh);
    }
    original_state_t get_index(std::size_t i) const {
        const std::size_t w = i / cells_per_word, sh = (i % cells_per_word) * bits_per_cell;
        return states_dict_t::index_to_state(static_cast<int>((_host[w] >> sh) & cell_mask));
    }

    std::size_t _x = 0, _y = 0;
    std::vector<store_word_type> _host;
    store_word_type* _dev = nullptr;
};

} // namespace cellato::memory::grids::bit_array

#endif
