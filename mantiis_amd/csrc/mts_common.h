// Definitions shared by the host builder (mts_grid.cpp), the kernels
// (mts_kernels.hip) and the C ABI (mts_api.hip).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mts {

// ---------------------------------------------------------------- D2Q9
// Direction order, opposites and weights of utils.hpp:50-70, coreLB.hpp:83.
constexpr int Q = 9;
constexpr int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

// ---------------------------------------------------------------- link word
// One 32-bit word per lattice site of the dense nx*ny image (0 = solid).
// It encodes the reference's per-cell link tables (coreLB.hpp:272-366) in the
// form a per-site gather ("pull") kernel needs:
//   bits  0.. 7  IN:   slot j (=bit j-1, j = 1..8) receives population j from
//                      the site at x - c_j (x/y wrap) -- a stream link or an
//                      x-seam periodic link of that source cell;
//   bits  8..15  OUT:  the outgoing link (x, ic) is in the Shan-Chen
//                      neighbour sum (stream or x-periodic link);
//   bits 16..23  WALL: the outgoing link (x, ic) is a boundary link;
//   bits 24..29  TYPE: boundary type 0..51 (grid.hpp:1111-1157), 63 = none;
//   bit  30      FLUID: the site is a cell;
//   bit  31      SEAM: the site is in column 0 or nx-1, where an IN link whose
//                source lies across the x-seam is a periodic link that
//                overwrites the slot after the wall rule (coreLB.hpp:969-977).
//                Links crossing only the y-seam are dropped (no IN bit).
// On a torus lattice (mts_params.periodic_y = 1) no site carries SEAM: links
// across both seams are classified like interior links.
constexpr uint32_t IN_SHIFT = 0, OUT_SHIFT = 8, WALL_SHIFT = 16, TYPE_SHIFT = 24;
constexpr uint32_t TYPE_NONE = 63;
constexpr uint32_t FLUID_BIT = 1u << 30;
constexpr uint32_t SEAM_BIT = 1u << 31;

// Per distinct (tau_s, Kn) pair: the relaxation rates and wall coefficients
// the reference recomputes per cell and step (coreLB.hpp:599-609, 705-715,
// 757-770, 925, 961).
struct CellParam {
    double taus;    // tau_s (coreLB.hpp:482-502)
    double kn;      // Kn
    double s_s;     // 1/tau_s
    double s_q;     // 1/tau_q, tau_q = 0.5 + (3 + pi (2 tau_s - 1)^2 A2) / (8 (2 tau_s - 1))
    double s_d;     // 1/tau_d of the non-regularised MRT and the Guo source
    double r_srbb;  // 0.4987 Kn^-0.131 - 0.25
    double r_mdbb;  // 0.3222 Kn^-0.217 - 0.25
    double pad;
};

// PR-EOS constants and unit conversions (globalParameters.hpp:45-222).
struct Physics {
    double Cl = 1.0e-9, Fbody = 1e-8, T_phy = 300.0, P_phy = 2.0e6;
    double rhoin = 0, mfp = 0, Cf = 0, T_lu = 0, a_alpha_lu = 0, Z = 0, Crho = 0, Cu = 0;
    bool cubic_ok = false;
};
void compute_physics(Physics& ph);
CellParam make_param(double taus, double kn, const Physics& ph);

// Host-side description of one single-block lattice.
struct Geometry {
    int nx = 0, ny = 0;
    bool torus = false;                  // seams classified like interior links
    int64_t stride = 0;                  // elements per population plane (>= nx*ny, multiple of 256)
    int64_t ncell = 0;                   // fluid cells
    std::vector<int32_t> cellPos;        // compact cell -> dense site y*nx + x
    std::vector<int64_t> gridConnect;    // [c*9+ic] (grid.hpp:916-972)
    std::vector<int32_t> bndTypes;       // [c]
    std::vector<double> taus, kn;        // [c]
    std::vector<uint32_t> meta;          // [nx*ny] link words
    std::vector<uint16_t> pidx;          // [nx*ny] index into params (0 for solids)
    std::vector<CellParam> params;
    int64_t gpIn = -1, gpOut = -1;       // dense sites of the generalized-periodic cells
};

// Builds the single-block grid exactly as Grid2D::initialize + LB2D::LB2D at
// one rank.  Returns 0 or a negative MTS_ERR_* code with a message in err.
int build_single_grid(int nx, int ny, const int32_t* pore, const double* kn, const double* locpore,
                      const Physics& ph, bool torus, Geometry& g, std::string& err);

// ---------------------------------------------------------------- cell order
// The FP32 fused passes carry the populations in cell order (the fluid sites
// in row-major order) and run one thread per cell.  Per cell:
//   cw (u32): bits 0..7 IN, 8..15 WALL (as in the link word), 16..21 TYPE,
//             bit 22 COMPLEX, bit 23 (x, y-1) is a cell, bit 24 (x, y+1) is a
//             cell, bits 25..31 x mod 128;
//   rw (u32): ru - c in bits 0..15 and c - rd in bits 16..31 (signed 16-bit),
//             where ru, rd are the numbers of cells before the site (x, y+1) /
//             (x, y-1) (y wraps), so that the sources of slots 4, 8, 7 are ru,
//             ru-1, ru+[(x,y+1) is a cell] and those of slots 2, 5, 6 are rd,
//             rd-1, rd+[...]; a cell whose offsets do not fit (the wrapped
//             rows of a torus) is COMPLEX;
//             the x neighbours of a cell are the cells c-1 and c+1;
//   outm:     the OUT bits (Shan-Chen neighbour sum).
// A COMPLEX cell (lattice columns 0 and nx-1, x-seam sites, the generalized-
// periodic cells and cells whose wall links share a specular slot) keeps its
// sources, position and full link word in an aux record; its rw is the record.
constexpr uint32_t CW_WALL_SHIFT = 8, CW_TYPE_SHIFT = 16;
constexpr uint32_t CW_COMPLEX = 1u << 22, CW_ISC_D = 1u << 23, CW_ISC_U = 1u << 24;
constexpr uint32_t CW_X_SHIFT = 25;
constexpr int CTILE_W = 128, CTILE_H = 8;  // Shan-Chen tile (sites)
constexpr int AUXW = 12;  // aux record: src[8], x, y, link word, override slot (-1: none)

struct CompactImage {
    int64_t ncell = 0, cstride = 0, naux = 0;
    int ntx = 0;                   // tile columns of CTILE_W sites
    std::vector<uint32_t> cw;      // [ncell]
    std::vector<uint32_t> rw;      // [ncell]: packed rank offsets (COMPLEX: aux record)
    std::vector<uint16_t> pidx;    // [ncell]
    std::vector<uint8_t> outm;     // [ncell]
    std::vector<int32_t> aux;      // [naux * AUXW]
    std::vector<int32_t> cidx;     // [nx*ny] site -> cell or -1
    std::vector<int32_t> cellPos;  // [ncell] cell -> site
    std::vector<int32_t> segs;     // [ny * (ntx+1)]: cells before the site (tx*CTILE_W, y)
    uint32_t spec[64] = {};        // per TYPE: slot j's specular source ic in bits 4(j-1)..4j-1
};

// The specular-source table of the cell-order kernels (the same for every
// lattice): per boundary type (52 = typeless, full bounce-back), slot j's
// source link ic, taken when ic is a wall link of the cell.
void canonical_spec(uint32_t spec[64]);

// A site image to put in cell order: link words and parameter indices per
// site, the generalized-periodic sites and, per site, the slot of an
// interface override (multi-block buffer cells; nullptr or -1: none).  A cell
// with an override slot is COMPLEX; its aux record keeps the slot.
struct LatticeView {
    int nx, ny;
    bool torus;
    const uint32_t* meta;
    const uint16_t* pidx;
    int64_t gpIn, gpOut;
    const int32_t* override;
};

// Builds the cell-order image of a single-block lattice or of one level.
void build_compact(const Geometry& g, CompactImage& ci);
void build_compact(const LatticeView& v, CompactImage& ci);

// Decodes the link lists of coreLB.hpp:272-366 from the link words.
void decode_links(const Geometry& g, int which, std::vector<int64_t>& id, std::vector<int64_t>& loc,
                  std::vector<int32_t>& ic);

}  // namespace mts
