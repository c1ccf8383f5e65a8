// Drop-in counterpart of the reference's include/grid.hpp (class Grid2D,
// grid.hpp:167-264).  The cell list, connectivity and boundary types are
// built by the library when LB2D is constructed (mts_create); Grid2D holds
// the geometry until then and mirrors the arrays afterwards.
#pragma once

#include <algorithm>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "mantiis_hip.h"
#include "shape.hpp"
#include "utils.hpp"

class Grid2D {
public:
    std::vector<int64_t> gridID;
    std::vector<int> gridType, gridLevel, gridIsBuffer;
    std::vector<std::vector<int>> gridIJ;
    std::vector<std::vector<int64_t>> gridConnect;
    std::vector<int64_t> solidID;
    std::vector<double> locpore, kn;
    std::vector<int> bndTypes;

    int ny, nx, maxLevel = 0;
    lattice latt;
    bool isMultigrid;
    int64_t globalGridSize, localGridSize = 0, startID = 0, endID = 0;

    const Shape* shape = nullptr;

    Grid2D(const Shape& s, const lattice& l, bool isMulti = false)
        : ny(s.Ny), nx(s.Nx), latt(l), isMultigrid(isMulti), globalGridSize((int64_t)s.Ny * s.Nx), shape(&s) {}

    // grid.hpp:271-311 (single-block): cells are the fluid pixels in
    // row-major order; solidID lists the solid pixels.
    void initialize(const Shape& s) {
        if (isMultigrid)
            throw std::invalid_argument("multi-block grids are not provided by the MI355X build of this step");
        shape = &s;
        std::cout << "Processing Grid..." << std::endl;
        solidID.clear();
        int64_t n = 0;
        for (int i = 0; i < s.Ny; i++)
            for (int j = 0; j < s.Nx; j++) {
                if (s.domain[i][j] == 0) solidID.push_back((int64_t)i * s.Nx + j);
                else n++;
            }
        globalGridSize = localGridSize = endID = n;
        startID = 0;
        maxLevel = 1;
    }

    // Filled from the library once the lattice exists (LB2D constructor).
    void mirror(mts_sim h) {
        int64_t n = 0;
        mts_cell_count(h, &n);
        std::vector<int32_t> ij(2 * n), lvl(n), bnd(n);
        std::vector<int64_t> con(9 * n);
        gridID.resize(n);
        mts_grid_info(h, gridID.data(), ij.data(), lvl.data(), con.data(), bnd.data());
        gridIJ.assign(n, std::vector<int>(2));
        gridConnect.assign(n, std::vector<int64_t>(9));
        gridLevel.assign(lvl.begin(), lvl.end());
        bndTypes.assign(bnd.begin(), bnd.end());
        gridType.assign(n, 0);
        gridIsBuffer.assign(n, 0);
        for (int64_t c = 0; c < n; c++) {
            gridIJ[c] = {ij[2 * c], ij[2 * c + 1]};
            for (int k = 0; k < 9; k++) gridConnect[c][k] = con[9 * c + k];
        }
        kn.resize(n);
        locpore.resize(n);
        for (int64_t c = 0; c < n; c++) {
            kn[c] = shape->Kn[gridIJ[c][1]][gridIJ[c][0]];
            locpore[c] = shape->LocPore[gridIJ[c][1]][gridIJ[c][0]];
        }
    }

    // One process holds the whole lattice.
    void MPI_broadcast(int = 0) {}
    void MPI_distributeGridData() {}
};
