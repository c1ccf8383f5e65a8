// Drop-in counterpart of the reference's include/io.hpp (io.hpp:53-92):
// whitespace-separated text in, one value per line at 15 significant digits out.
// The binary pair (no reference counterpart) stores the raw values: a
// lossless, fast form of the same vectors, e.g. f for an exact resume.
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

// Raw values in the host's byte order, no header; the reader checks the size.
template <typename T>
void writeBinary(const std::string& fileName, const std::vector<T>& data) {
    std::ofstream file(fileName, std::ios::binary);
    if (!file.is_open()) throw std::runtime_error("[ERROR]: Unable to open file for writing.");
    file.write(reinterpret_cast<const char*>(data.data()), (std::streamsize)(data.size() * sizeof(T)));
    if (!file) throw std::runtime_error("[ERROR]: Short write to " + fileName);
}

template <typename T>
void loadBinary(const std::string& fileName, std::vector<T>& data) {
    std::ifstream file(fileName, std::ios::binary | std::ios::ate);
    if (!file.is_open()) throw std::runtime_error("[ERROR]: File not found: " + fileName);
    if ((size_t)file.tellg() != data.size() * sizeof(T))
        throw std::runtime_error("[ERROR]: " + fileName + " does not hold " + std::to_string(data.size()) + " values.");
    file.seekg(0);
    file.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(data.size() * sizeof(T)));
}

}  // namespace IO
