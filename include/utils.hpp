// Drop-in counterpart of the parts of the reference's include/utils.hpp that
// the time step's callers use: the D2Q9 lattice (utils.hpp:50-70) and the
// index maps (utils.hpp:365-408).
#pragma once

#include <cstdint>
#include <iostream>
#include <vector>

#include "io.hpp"

struct lattice {
    //                   0  1  2   3  4  5   6   7  8
    std::vector<int> e_x{0, 1, 0, -1, 0, 1, -1, -1, 1};
    std::vector<int> e_y{0, 0, 1, 0, -1, 1, 1, -1, -1};
    std::vector<int> icOpp{0, 3, 4, 1, 2, 7, 8, 5, 6};
};

struct Index {
    int row;
    int col;
};

namespace utils {

// First subscript varies fastest; out-of-range subscripts print and give -1.
template <typename T>
inline int64_t sub2ind(const std::vector<int>& sizes, const std::vector<T>& subs) {
    if (subs.size() != sizes.size()) {
        std::cerr << "Error: Number of subscripts and sizes do not match." << std::endl;
        return -1;
    }
    int64_t index = 0, mul = 1;
    for (size_t i = 0; i < subs.size(); ++i) {
        if (subs[i] < 0 || subs[i] >= sizes[i]) {
            std::cerr << "Error: Subscript out of bounds." << std::endl;
            return -1;
        }
        index += static_cast<int64_t>(subs[i]) * mul;
        mul *= sizes[i];
    }
    return index;
}

inline std::vector<int> ind2sub(const std::vector<int>& size, int64_t index) {
    std::vector<int> s(size.size());
    for (size_t i = 0; i < size.size(); i++) {
        s[i] = static_cast<int>(index % size[i]);
        index /= size[i];
    }
    return s;
}

}  // namespace utils
