// One state value per cell and the 16-tile checksum; the reference's standard::grid
// (include/memory/standard_grid.hpp:32-221, checksum :179-213), host side only.
#ifndef CELLATO_MEMORY_STANDARD_GRID_HPP
#define CELLATO_MEMORY_STANDARD_GRID_HPP

#include <algorithm>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace cellato::memory::grids::standard {

template <typename cell_type>
class grid {
public:
    using store_type = cell_type;

    grid() = default;
    grid(int x_size, int y_size) : _x(x_size), _y(y_size), _data(static_cast<std::size_t>(x_size) * y_size) {}
    grid(std::vector<cell_type>&& data, int x_size, int y_size) : _x(x_size), _y(y_size), _data(std::move(data)) {
        if (static_cast<long long>(_data.size()) != static_cast<long long>(x_size) * y_size)
            throw std::invalid_argument("Data size does not match grid dimensions");
    }

    cell_type* data() { return _data.data(); }
    const cell_type* data() const { return _data.data(); }
    int x_size_physical() const { return _x; }
    int y_size_physical() const { return _y; }

    // 4x4 tiles of ceil(W/4) x ceil(H/4) cells, sums of state values joined by '-'
    std::string get_checksum() const {
        const int tw = (_x + 3) / 4, th = (_y + 3) / 4;
        std::ostringstream out;
        for (int ty = 0; ty < 4; ++ty)
            for (int tx = 0; tx < 4; ++tx) {
                std::uint64_t sum = 0;
                for (int y = ty * th; y < std::min(ty * th + th, _y); ++y)
                    for (int x = tx * tw; x < std::min(tx * tw + tw, _x); ++x)
                        sum += static_cast<std::uint64_t>(_data[static_cast<std::size_t>(y) * _x + x]);
                if (ty || tx) out << '-';
                out << sum;
            }
        return out.str();
    }

private:
    int _x = 0, _y = 0;
    std::vector<cell_type> _data;
};

} // namespace cellato::memory::grids::standard

#endif
