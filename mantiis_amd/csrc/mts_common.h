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

// Decodes the link lists of coreLB.hpp:272-366 from the link words.
void decode_links(const Geometry& g, int which, std::vector<int64_t>& id, std::vector<int64_t>& loc,
                  std::vector<int32_t>& ic);

}  // namespace mts
