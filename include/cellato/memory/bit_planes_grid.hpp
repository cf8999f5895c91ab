// Bit-plane grids for the HIP traversers: the reference's bit_planes::grid and
// tiled_bit_planes::grid (include/memory/bit_planes_grid.hpp:30-271,
// tiled_bit_planes_grid.hpp:30-289) with the same members (needed_bits, word_store_bits,
// x/y_word_tile_size, store_type, data(), x_size_physical(), y_size_physical(), to_standard(),
// get_cell(), to_cuda(), to_cpu(), free_cuda_memory(), and the implicit conversion from a
// standard grid that the reference's experiment manager relies on). The device copy lives in
// one hipMalloc'd block (planes back to back); to_hip() / free_device_memory() are the same
// calls under this library's names. Host code only; needs the HIP runtime library.
#ifndef CELLATO_MEMORY_BIT_PLANES_GRID_HPP
#define CELLATO_MEMORY_BIT_PLANES_GRID_HPP

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "standard_grid.hpp"

namespace cellato::memory::grids {

enum class device { CPU, HIP };

struct hip_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw hip_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Shared implementation; Tiled selects the 8x8 block layout.
template <typename word_t, typename dict_t, bool Tiled, device dev = device::CPU>
class plane_grid {
public:
    static_assert(sizeof(word_t) == 8, "64-bit words");
    static constexpr int needed_bits = dict_t::needed_bits;
    static constexpr int word_store_bits = sizeof(word_t) * 8;
    static constexpr int x_word_tile_size = Tiled ? 8 : 64;
    static constexpr int y_word_tile_size = Tiled ? 8 : 1;
    static constexpr int layout = Tiled ? 1 : 0;
    using store_type = word_t;
    using cell_t = word_t;
    using original_state_t = typename dict_t::state_t;
    using dictionary = dict_t;

    plane_grid() = default;

    // from a standard grid of state values (bit_planes_grid.hpp:80-84); implicit, as there
    plane_grid(const standard::grid<original_state_t>& s) requires (dev == device::CPU)
        : plane_grid(s.y_size_physical(), s.x_size_physical(), s.data()) {}

    plane_grid(std::size_t y_size, std::size_t x_size, const original_state_t* cells) requires (dev == device::CPU) {
        if (x_size % x_word_tile_size || y_size % y_word_tile_size)
            throw std::invalid_argument("grid size must be a multiple of the word tile (" +
                                        std::to_string(x_word_tile_size) + "x" + std::to_string(y_word_tile_size) + ")");
        _wx = x_size / x_word_tile_size;
        _h = y_size / y_word_tile_size;
        _host.assign(needed_bits * _wx * _h, 0);
        for (std::size_t y = 0; y < y_size; ++y)
            for (std::size_t x = 0; x < x_size; ++x) {
                const int idx = dict_t::state_to_index(cells[y * x_size + x]);
                const auto [w, bit] = locate(x, y);
                for (int b = 0; b < needed_bits; ++b)
                    if ((idx >> b) & 1) _host[b * plane_words() + w] |= word_t(1) << bit;
            }
    }

    std::size_t x_size_physical() const { return _wx; }
    std::size_t y_size_physical() const { return _h; }
    std::size_t x_size_original() const { return _wx * x_word_tile_size; }
    std::size_t y_size_original() const { return _h * y_word_tile_size; }
    std::size_t plane_words() const { return _wx * _h; }

    std::array<word_t*, needed_bits> data() const {
        std::array<word_t*, needed_bits> r{};
        word_t* base = dev == device::HIP ? _dev : const_cast<word_t*>(_host.data());
        for (int b = 0; b < needed_bits; ++b) r[b] = base + b * plane_words();
        return r;
    }

    original_state_t get_cell(std::size_t x, std::size_t y) const requires (dev == device::CPU) {
        if (x >= x_size_original() || y >= y_size_original()) throw std::out_of_range("Cell coordinates out of range");
        const auto [w, bit] = locate(x, y);
        int idx = 0;
        for (int b = 0; b < needed_bits; ++b) idx |= static_cast<int>((_host[b * plane_words() + w] >> bit) & 1) << b;
        return dict_t::index_to_state(idx);
    }

    standard::grid<original_state_t> to_standard() const requires (dev == device::CPU) {
        std::vector<original_state_t> cells(x_size_original() * y_size_original());
        for (std::size_t y = 0; y < y_size_original(); ++y)
            for (std::size_t x = 0; x < x_size_original(); ++x) cells[y * x_size_original() + x] = get_cell(x, y);
        return standard::grid<original_state_t>(std::move(cells), static_cast<int>(x_size_original()),
                                                static_cast<int>(y_size_original()));
    }

    plane_grid<word_t, dict_t, Tiled, device::HIP> to_hip() const requires (dev == device::CPU) {
        plane_grid<word_t, dict_t, Tiled, device::HIP> g;
        g._wx = _wx;
        g._h = _h;
        hip_check(hipMalloc(reinterpret_cast<void**>(&g._dev), _host.size() * sizeof(word_t)), "hipMalloc");
        hip_check(hipMemcpy(g._dev, _host.data(), _host.size() * sizeof(word_t), hipMemcpyHostToDevice), "hipMemcpy H2D");
        return g;
    }

    plane_grid<word_t, dict_t, Tiled, device::CPU> to_cpu() const requires (dev == device::HIP) {
        plane_grid<word_t, dict_t, Tiled, device::CPU> g;
        g._wx = _wx;
        g._h = _h;
        g._host.resize(needed_bits * plane_words());
        hip_check(hipMemcpy(g._host.data(), _dev, g._host.size() * sizeof(word_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        return g;
    }

    // the reference's name for the device copy (bit_planes_grid.hpp:132)
    plane_grid<word_t, dict_t, Tiled, device::HIP> to_cuda() const requires (dev == device::CPU) { return to_hip(); }

    void free_device_memory() {
        if constexpr (dev == device::HIP) {
            if (_dev) hip_check(hipFree(_dev), "hipFree");
            _dev = nullptr;
        }
    }
    // the reference's name (bit_planes_grid.hpp:86-95)
    void free_cuda_memory() { free_device_memory(); }

    std::string get_checksum() const requires (dev == device::CPU) { return to_standard().get_checksum(); }

private:
    template <typename, typename, bool, device> friend class plane_grid;

    // (word index within a plane, bit) of cell (x, y)
    std::pair<std::size_t, int> locate(std::size_t x, std::size_t y) const {
        if constexpr (Tiled) return {(y / 8) * _wx + x / 8, static_cast<int>((y % 8) * 8 + x % 8)};
        else return {y * _wx + x / 64, static_cast<int>(x % 64)};
    }

    std::size_t _wx = 0, _h = 0;
    std::vector<word_t> _host;
    word_t* _dev = nullptr;
};

namespace bit_planes {
template <typename word_t, typename dict_t, device dev = device::CPU>
using grid = plane_grid<word_t, dict_t, false, dev>;
}
namespace tiled_bit_planes {
template <typename word_t, typename dict_t, device dev = device::CPU>
using grid = plane_grid<word_t, dict_t, true, dev>;
}

} // namespace cellato::memory::grids

#endif
