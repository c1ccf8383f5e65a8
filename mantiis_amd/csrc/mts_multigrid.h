// Quadtree lattice description built on the host (mts_multigrid.cpp) and
// turned into per-level device images (mts_api.hip).
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
    std::vector<double> taus, kn;           // taus: empty until multi_grid_taus
    std::vector<int32_t> pore;              // the pore image, row-major
    std::vector<uint8_t> lvl;               // level of each pixel's quadtree leaf, row-major
    std::vector<int64_t> parentChild;       // rows {cell, tl, tr, bl, br}
    // interface rows of the pairs of levels (l, l + 1), l = 1 .. maxLevel - 1,
    // one pair after the other; *Count[l - 1] rows belong to pair l
    std::vector<int64_t> coarseRows;        // rows {coarse buffer, 4 children} (type 10 (l + 1) + l)
    std::vector<int64_t> fineRows;          // rows {parent, 4 fine buffer cells} (type 10 l + l + 1)
    std::vector<int64_t> coarseCount, fineCount;
};

int build_multi_grid(int nx, int ny, const int32_t* pore, const double* kn, const double* locpore, const Physics& ph,
                     bool torus, int K, MultiGeometry& G, std::string& err);
// The stages of build_multi_grid: argument check (K: most levels), level
// decision (the level of each pixel's leaf, nx * ny), the cells (original
// and buffer cells, their positions, levels, types, gridID, Kn, the pore
// and level images) and their neighbour and interface tables.
int multi_check(int nx, int ny, const int32_t* pore, int K, std::string& err);
// Most levels of the quadtree named by mts_params.multiblock (1: the
// reference's two).
inline int quadtree_levels(int multiblock) { return multiblock >= 2 ? multiblock : 2; }
int split_distance(int K);
void multi_levels_host(int nx, int ny, const int32_t* pore, bool torus, int K, std::vector<uint8_t>& lvl);
int multi_cells(int nx, int ny, const int32_t* pore, const double* kn, bool torus, const std::vector<uint8_t>& lvl,
                MultiGeometry& G, std::string& err);
int multi_tables_host(MultiGeometry& G, std::string& err);
// Per-cell relaxation times of a built grid (needs the current device).
int multi_grid_taus(const Physics& ph, MultiGeometry& G, std::string& err);

}  // namespace mts

namespace mts {

// One level of a multi-block lattice as a dense site image.
struct LevelImage {
    int level = 1;           // 1 finest
    int nx = 0, ny = 0;      // image size (pixels >> (level-1))
    int64_t stride = 0;
    std::vector<uint32_t> meta;   // link words (mts_common.h)
    std::vector<uint16_t> pidx;   // parameter index per site
    std::vector<int64_t> cells;   // global cells of this level, in global order
    std::vector<int32_t> sites;   // their sites
};

// Link words and parameter indices of every level image (the classification
// of coreLB.hpp:292-356 applied to the multi-block connectivity).
// meta = false leaves the link words at zero (the device builder forms them).
int build_level_images(const MultiGeometry& G, std::vector<LevelImage>& L, std::vector<CellParam>& params,
                       const Physics& ph, std::string& err, bool meta = true);

}  // namespace mts

// the HIP stream type (this header is also compiled by the host compiler)
typedef struct ihipStream_t* hipStream_t;

namespace mts {

// The device builder (mts_quadtree.hip): the level decision, and, from the
// host's cells, the neighbour and interface tables of G and (levels !=
// nullptr, images sized by build_level_images) the link words of the level
// images -- entry for entry those of the host builder.
int multi_levels_device(int nx, int ny, const int32_t* pore, bool torus, int K, hipStream_t st,
                        std::vector<uint8_t>& lvl, std::string& err);
int multi_tables_device(MultiGeometry& G, std::vector<LevelImage>* levels, hipStream_t st, std::string& err);

}  // namespace mts
