// Drop-in counterpart of the reference's include/shape.hpp: the binary
// domain with per-pixel Kn and local pore size (shape.hpp:48-77), row-major
// [i = y][j = x].
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "globalParameters.hpp"
#include "io.hpp"

class Shape {
public:
    int Nx;
    int Ny;
    double resolution;  // meters
    std::vector<std::vector<int>> domain;
    std::vector<std::vector<double>> Kn;
    std::vector<std::vector<double>> LocPore;

    Shape(int x, int y, double res) : Nx(x), Ny(y), resolution(res) {
        domain.assign(Ny, std::vector<int>(Nx, 1));
        LocPore.assign(Ny, std::vector<double>(Nx, 0.0));
        Kn.assign(Ny, std::vector<double>(Nx, 0.0));
    }

    // shape.hpp:91-114
    void addHorizontalBoundary(int h) {
        for (int j = 0; j < Nx; j++) domain[h][j] = 0;
    }
    void addCircle(int R, int cx, int cy) {
        for (int i = 0; i < Ny; i++)
            for (int j = 0; j < Nx; j++)
                if (std::sqrt(std::pow((double)(i - cy), 2) + std::pow((double)(j - cx), 2)) <= (double)R) domain[i][j] = 0;
    }
    void addRectangle(int xs, int xe, int ys, int ye) {
        for (int i = 0; i < Ny; i++)
            for (int j = 0; j < Nx; j++)
                if (j >= xs && j <= xe && i >= ys && i <= ye) domain[i][j] = 0;
    }

    // shape.hpp:253-285: pore.dat, Kn.dat, localporesize.dat, column-major.
    void loadExistingModel(std::string& f) {
        std::vector<int32_t> pore((size_t)Nx * Ny);
        std::vector<double> kn((size_t)Nx * Ny), lp((size_t)Nx * Ny);
        if (Nx != Ny) throw std::invalid_argument("loadExistingModel expects a square domain");
        if (mts_load_existing_model(f.c_str(), Nx, pore.data(), kn.data(), lp.data()) != MTS_OK)
            throw std::runtime_error(mts_last_error());
        for (int i = 0; i < Ny; i++)
            for (int j = 0; j < Nx; j++) {
                domain[i][j] = pore[(size_t)i * Nx + j];
                Kn[i][j] = kn[(size_t)i * Nx + j];
                LocPore[i][j] = lp[(size_t)i * Nx + j];
            }
    }

    // shape.hpp:287-319
    void writeToText(std::string& folder) {
        IO::writeVectorToFile(folder + "dimensions.dat", std::vector<int>{Ny, Nx});
        std::vector<int> d;
        std::vector<double> l, k;
        for (int i = 0; i < Ny; i++)
            for (int j = 0; j < Nx; j++) {
                d.push_back(domain[i][j]);
                l.push_back(LocPore[i][j]);
                k.push_back(Kn[i][j]);
            }
        IO::writeVectorToFile(folder + "pore.dat", d);
        IO::writeVectorToFile(folder + "localporesize.dat", l);
        IO::writeVectorToFile(folder + "Kn.dat", k);
    }

    // One process holds the whole geometry.
    void MPI_broadcast(int = 0) {}

    // Skeleton / distance-transform pore-size estimate (shape.hpp:127-251),
    // computed on the GPU: sets resolution, LocPore and Kn of every fluid
    // pixel (wall pixels keep their values, as in the reference).  The
    // skeleton window of the reference (dispSkeleton) is not opened.
    void calculateProperties(double res, double mfp, bool = false) {
        resolution = res;
        const std::vector<int32_t> d = flatDomain();
        std::vector<double> lp((size_t)Nx * Ny), kn((size_t)Nx * Ny);
        if (mts_shape_properties(d.data(), Nx, Ny, res, mfp, lp.data(), kn.data(), nullptr) != MTS_OK)
            throw std::runtime_error(mts_last_error());
        for (int i = 0; i < Ny; i++)
            for (int j = 0; j < Nx; j++)
                if (domain[i][j] == 1) {
                    LocPore[i][j] = lp[(size_t)i * Nx + j];
                    Kn[i][j] = kn[(size_t)i * Nx + j];
                }
    }

    // row-major flat copies for the C ABI
    std::vector<int32_t> flatDomain() const {
        std::vector<int32_t> v;
        for (const auto& r : domain) v.insert(v.end(), r.begin(), r.end());
        return v;
    }
    std::vector<double> flat(const std::vector<std::vector<double>>& a) const {
        std::vector<double> v;
        for (const auto& r : a) v.insert(v.end(), r.begin(), r.end());
        return v;
    }
};
