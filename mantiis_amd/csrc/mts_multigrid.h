// Two-level quadtree lattice description built on the host
// (mts_multigrid.cpp) and turned into per-level device images (mts_api.hip).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "mts_common.h"

namespace mts {

struct MultiGeometry {
    int nx = 0, ny = 0, maxLevel = 1;
    bool torus = false;
    int64_t ncell = 0;
    // cells in the reference's order: fluid cells (row-major by top-left
    // pixel), then buffer cells in creation order
    std::vector<int32_t> cx, cy, level, type, isBuffer, bnd;
    std::vector<int64_t> gridID, connect;   // connect[c*9+k]: cell, -1 wall, -2 missing
    std::vector<double> taus, kn;
    std::vector<int64_t> parentChild;       // rows {cell, tl, tr, bl, br}
    std::vector<int64_t> coarseRows;        // rows {coarse buffer, 4 children} (type 21)
    std::vector<int64_t> fineRows;          // rows {parent, 4 fine buffer cells} (type 12)
};

int build_multi_grid(int nx, int ny, const int32_t* pore, const double* kn, const double* locpore, const Physics& ph,
                     bool torus, MultiGeometry& G, std::string& err);

}  // namespace mts
