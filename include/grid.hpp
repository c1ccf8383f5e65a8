// Drop-in counterpart of the reference's include/grid.hpp (class Grid2D,
// grid.hpp:167-264).  A single-block cell list, connectivity and boundary
// types are built by the library when LB2D is constructed (mts_create);
// Grid2D holds the geometry until then and mirrors the arrays afterwards.  A
// multi-block grid (isMultigrid) is built by the library's host builder in
// initialize (mts_grid_create: the quadtree, buffer cells, connectivity,
// boundary types), as the reference builds it there.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mantiis_hip.h"
#include "shape.hpp"
#include "utils.hpp"

// Switches read when a Grid2D is initialised or an LB2D constructed:
// MANTIIS_TORUS=1 selects the torus seam rule; MANTIIS_LEVELS=3 lets a
// multi-block grid have three quadtree levels (maxAllowedLevel = 3; the
// reference's two otherwise).
inline int mantiis_torus_from_env() {
    const char* e = std::getenv("MANTIIS_TORUS");
    return (e && std::string(e) == "1") ? 1 : 0;
}
inline int mantiis_levels_from_env() {
    const char* e = std::getenv("MANTIIS_LEVELS");
    return (e && std::string(e) == "3") ? 3 : 1;
}

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
        shape = &s;
        std::cout << "Processing Grid..." << std::endl;
        if (isMultigrid) {
            initializeMultigrid(s);
            return;
        }
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

    // grid.hpp:313-1157 (the multi-block branch of initialize): quadtree,
    // cells in the reference's order, buffer cells, connectivity and
    // boundary types, from the library's host builder.
    void initializeMultigrid(const Shape& s) {
        mts_params p;
        mts_default_params(&p);
        p.multiblock = mantiis_levels_from_env();
        p.periodic_y = mantiis_torus_from_env();
        const std::vector<int32_t> pore = s.flatDomain();
        const std::vector<double> knf = s.flat(s.Kn), lpf = s.flat(s.LocPore);
        mts_grid g = nullptr;
        if (mts_grid_create(s.Nx, s.Ny, pore.data(), knf.data(), lpf.data(), &p, &g) != MTS_OK)
            throw std::runtime_error(mts_last_error());
        int64_t n = 0;
        int32_t L = 0;
        mts_grid_size(g, &n, &L);
        std::vector<int32_t> ij(2 * n), lvl(n), bnd(n), typ(n), isb(n);
        std::vector<int64_t> con(9 * n);
        gridID.resize(n);
        const int rc = mts_grid_arrays(g, gridID.data(), ij.data(), lvl.data(), con.data(), bnd.data(), typ.data(),
                                       isb.data(), nullptr);
        mts_grid_destroy(g);
        if (rc != MTS_OK) throw std::runtime_error(mts_last_error());
        solidID.clear();
        for (int i = 0; i < s.Ny; i++)
            for (int j = 0; j < s.Nx; j++)
                if (s.domain[i][j] == 0) solidID.push_back((int64_t)i * s.Nx + j);
        gridIJ.assign(n, std::vector<int>(2));
        gridConnect.assign(n, std::vector<int64_t>(9));
        gridLevel.assign(lvl.begin(), lvl.end());
        bndTypes.assign(bnd.begin(), bnd.end());
        gridType.assign(typ.begin(), typ.end());
        gridIsBuffer.assign(isb.begin(), isb.end());
        kn.resize(n);
        locpore.resize(n);
        for (int64_t c = 0; c < n; c++) {
            gridIJ[c] = {ij[2 * c], ij[2 * c + 1]};
            for (int k = 0; k < 9; k++) gridConnect[c][k] = con[9 * c + k];
            kn[c] = s.Kn[gridIJ[c][1]][gridIJ[c][0]];
            locpore[c] = s.LocPore[gridIJ[c][1]][gridIJ[c][0]];
        }
        globalGridSize = localGridSize = endID = n;
        startID = 0;
        maxLevel = L;
    }

    // Filled from the library once the lattice exists (LB2D constructor); a
    // multi-block grid keeps the arrays of initialize (the device builder
    // behind mts_create forms the same tables).
    void mirror(mts_sim h) {
        if (isMultigrid) return;
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
