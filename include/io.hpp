// Drop-in counterpart of the reference's include/io.hpp (io.hpp:53-92):
// whitespace-separated text in, one value per line at 15 significant digits out.
#pragma once

#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace IO {

template <typename T>
void loadDataFromFile(const std::string& fileName, std::vector<T>& data) {
    std::ifstream file(fileName);
    if (!file.is_open()) {
        std::cout << "Tried loading " << fileName << std::endl;
        throw std::runtime_error("[ERROR]: File not found.");
    }
    for (size_t i = 0; i < data.size(); i++)
        if (!(file >> data[i])) throw std::runtime_error("[ERROR]: Insufficient data in the file.");
}

template <typename T>
void writeVectorToFile(const std::string& fileName, const std::vector<T>& data) {
    std::ofstream file(fileName);
    if (!file.is_open()) {
        std::cout << "Tried writing " << fileName << std::endl;
        throw std::runtime_error("[ERROR]: Unable to open file for writing.");
    }
    for (const auto& v : data) file << std::setprecision(15) << v << '\n';
}

}  // namespace IO
