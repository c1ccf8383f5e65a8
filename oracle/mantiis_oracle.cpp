// ============================================================================
// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A plain CPU restatement (C++17, FP64, AoS populations f[cell*9+dir],
// explicit link lists, dense 9x9 matrix-vector MRT products) of the mantiis
// D2Q9 time step.  It exists to check the HIP library, never to be shipped or
// measured as the product: only tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg load it.  The HIP library never links or calls it.
//
// Every function cites the reference file:line it restates.  Parity pins:
//   * whole single-block transport / adsorption trajectories: the reference's
//     own gold files test/test_input_output/gold/{transport,adsorption}
//     (64x64, 100 steps) -- tests/test_oracle_gold.py; they also pin the
//     link tables (stream, wall, periodic) of that geometry, whose every
//     wall pattern enters the trajectory;
//   * constants (PR-EOS root, rho_in, mfp, Cf): the reference's printed
//     values and the cubic's residual -- tests/test_oracle_physics.py;
//   * equilibrium / regularisation / MRT / SRT: closed-form moment identities
//     and MRT(all rates 1) == BGK(tau 1) -- tests/test_oracle_physics.py;
//   * walls and seams: mass invariants of BB / SRBB / torus and the loss at a
//     fluid y-seam under the reference's seam rule -- test_oracle_physics.py;
//     the 52-type boundary tables against the neighbourhood rule --
//     tests/test_tables.py;
//   * quadtree multi-block grid: brute-force invariants on small images and
//     an independent second build (the library's) bit for bit --
//     tests/test_grid.py;
//   * fine-buffer interpolation (ALGLIB spline2dbuildbicubicv on the 3 x 2
//     nodes the reference passes, coreLB.hpp:1217-1224): a product of a
//     quadratic in x and a linear function in y sampled at the parent's
//     neighbours is reproduced exactly at every child's centre (a bicubic
//     Hermite patch with spline-derived slopes reproduces such a product;
//     a transposed Z order, swapped top/bottom rows or a wrong centre shift
//     does not) -- tests/test_oracle_multiblock_pins.py;
//   * per-level downsampling (cv::resize + threshold 128, utils.hpp:539-568):
//     every 2 x 2 pattern against OpenCV's documented area (factor 2) and
//     bilinear (factor 4) sampling with its 8-bit rounding --
//     tests/test_oracle_multiblock_pins.py.
// Build: oracle/Makefile (strict IEEE: -O2 -ffp-contract=off, the checker)
// and an -Ofast -fopenmp copy used only as the timed CPU baseline.
// ============================================================================

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------- lattice
// utils.hpp:50-70 (e_x, e_y, opposite) and coreLB.hpp:83 (weights).
static const int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const double W[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                            1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
static const int NC = 9;
// coreLB.hpp:60-64
static const double F1 = 3.0, F2 = 4.5, F3 = 1.5;
static const double CS2 = 1. / 3.;

// globalParameters.hpp:142-177 (I, M and the 15-decimal truncated Minv).
static const double Mm[81] = {
    1., 1., 1., 1., 1., 1., 1., 1., 1.,
    -4., -1., -1., -1., -1., 2., 2., 2., 2.,
    4., -2., -2., -2., -2., 1., 1., 1., 1.,
    0., 1., 0., -1., 0., 1., -1., -1., 1.,
    0., -2., 0., 2., 0., 1., -1., -1., 1.,
    0., 0., 1., 0., -1., 1., 1., -1., -1.,
    0., 0., -2., 0., 2., 1., 1., -1., -1.,
    0., 1., -1., 1., -1., 0., 0., 0., 0.,
    0., 0., 0., 0., 0., 1., -1., 1., -1.};
static const double a9 = 0.111111111111111, b36 = 0.027777777777778, c18 = 0.055555555555556,
                    d6 = 0.166666666666667, e12 = 0.083333333333333, q4 = 0.25;
static const double Minv[81] = {
    a9, -a9, a9, 0, -0., -0., -0., 0, 0,
    a9, -b36, -c18, d6, -d6, 0, 0, q4, -0.,
    a9, -b36, -c18, -0., -0., d6, -d6, -q4, 0,
    a9, -b36, -c18, -d6, d6, -0., -0., q4, -0.,
    a9, -b36, -c18, 0, 0, -d6, d6, -q4, -0.,
    a9, c18, b36, d6, e12, d6, e12, 0, q4,
    a9, c18, b36, -d6, -e12, d6, e12, 0, -q4,
    a9, c18, b36, -d6, -e12, -d6, -e12, 0, q4,
    a9, c18, b36, d6, e12, -d6, -e12, 0, -q4};
// The same matrices as the reference's global std::vectors (_GLOBAL_::M,
// _GLOBAL_::Minv), which the collisions read through matMulLin: a product
// whose operands may alias the output keeps its accumulator in memory, the
// cost the reference's loops carry.
static const std::vector<double> MmV(Mm, Mm + 81), MinvV(Minv, Minv + 81);

// ---------------------------------------------------------------- globals
// globalParameters.hpp:45-222 (_GLOBAL_ namespace + recompute()).
struct Globals {
    double Cl = 1.0e-9, Fbody = 1e-8, T_phy = 300.0, P_phy = 2.0e6;
    double Tc_phy = 190.564, Pc_phy = 4.5992 * 1.0e6, d_mol = 3.800e-10, omega = 0.011,
           Mw = 16.043 * 1.0e-3;
    const double NA = 6.02214076e23, R_EOS = 8.31446, a_lu = 2.0 / 49.0, b_lu = 2.0 / 21.0,
                 R_lu = 1.0, Mw_lu = 1.0, Rs_lu = 1.0;
    double k, alpha, a_alpha_lu, Rs, a, aG, bG, AG, BG, ZG2, ZG1, ZG0, ZG, vG, rho_phy_mol,
        rho_phy, C_Rs, b_Rs, a_Rs, C_as, C_bs, C_aob, Ct, CT, T_lu, C_Mw, Cm, Crho, Cn, Cu, Cf,
        Crho_mol, rhoin, mfp;
    double Gfs = -12.7, Gff = -1.0;  // 120-121 (compile-time in the reference)
    const double A2 = 0.9;           // 130
    double tau_SRT = 1.0;            // 141
    bool cubic_ok = true;

    // utils.hpp:571-600 -- Cardano, delta<0 branch only, pow(u,1/3) of both
    // terms exactly as written (NaN for a negative term, as in C).
    static std::vector<double> solveCubic(double a, double b, double c, double d) {
        std::vector<double> roots;
        double p = (3 * a * c - b * b) / (3 * a * a);
        double q = (2 * b * b * b - 9 * a * b * c + 27 * a * a * d) / (27 * a * a * a);
        double delta = -4 * p * p * p - 27 * q * q;
        if (delta < 0) {
            double u1 = -q / 2 + std::sqrt(q * q / 4 + p * p * p / 27);
            double u2 = -q / 2 - std::sqrt(q * q / 4 + p * p * p / 27);
            double t = std::pow(u1, 1.0 / 3.0) + std::pow(u2, 1.0 / 3.0);
            roots.push_back(t - b / (3 * a));
        }
        return roots;
    }
    void recompute() {
        k = 0.37464 + 1.54226 * omega - 0.26992 * std::pow(omega, 2);
        alpha = std::pow((1 + k * (1 - std::sqrt(T_phy / Tc_phy))), 2);
        a_alpha_lu = a_lu * alpha;
        Rs = R_EOS / Mw;
        a = 0.45724 * (R_EOS * Tc_phy) * (R_EOS * Tc_phy) / Pc_phy;
        aG = a * alpha;
        bG = 0.07780 * R_EOS * Tc_phy / Pc_phy;
        AG = aG * P_phy / (R_EOS * R_EOS) / (T_phy * T_phy);
        BG = bG * P_phy / R_EOS / T_phy;
        ZG2 = BG - 1;
        ZG1 = AG - 3 * (BG * BG) - 2 * BG;
        ZG0 = BG * BG * BG + (BG * BG) - AG * BG;
        std::vector<double> sol = solveCubic(1.0, ZG2, ZG1, ZG0);
        // utils.hpp:225-237 findMaxValue starts from numeric_limits<double>::min().
        ZG = std::numeric_limits<double>::min();
        for (double v : sol)
            if (v > ZG) ZG = v;
        cubic_ok = sol.size() == 1 && std::isfinite(ZG);
        vG = ZG * R_EOS * T_phy / P_phy;
        rho_phy_mol = 1 / vG;
        rho_phy = rho_phy_mol * Mw;
        C_Rs = Rs / Rs_lu;
        b_Rs = 0.07780 * Rs * Tc_phy / Pc_phy;
        a_Rs = 0.45724 * std::pow((Rs * Tc_phy), 2) / Pc_phy;
        C_as = a_Rs / a_lu;
        C_bs = b_Rs / b_lu;
        C_aob = C_as / C_bs;
        Ct = Cl / std::sqrt(C_aob);
        CT = 1 / (C_Rs / std::pow(Cl, 2) * std::pow(Ct, 2));
        T_lu = T_phy / CT;
        C_Mw = Mw / Mw_lu;
        Cm = std::pow(Cl, 3) / C_bs;
        Crho = Cm / (Cl * Cl * Cl);
        Cn = Cm / C_Mw;
        Cu = Cl / Ct;
        Cf = Cm / (std::pow(Ct, 2) * std::pow(Cl, 2));
        Crho_mol = Cn / std::pow(Cl, 3);
        rhoin = rho_phy / Crho;
        mfp = 1.0 / (std::sqrt(2.0) * Crho_mol * rhoin / Mw_lu * NA * M_PI *
                     (2 * d_mol / 2.0 + 2 * d_mol / 2.0) * (2 * d_mol / 2.0 + 2 * d_mol / 2.0));
    }
};
// globalParameters.hpp:122
static const double WF[9] = {0.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0,
                             1.0 / 12.0, 1.0 / 12.0, 1.0 / 12.0, 1.0 / 12.0};

// ---------------------------------------------------------------- tables
// utils.hpp:728-793: direction-major 9x52 text tables, icsr 1-based.
struct Tables {
    int n = 52;
    int bnd_types[52];
    bool ifStream[52][9];
    int icsr[52][9];
    double xy_norm[52][2];
};
static void loadNumbers(const std::string& fn, std::vector<double>& data) {
    std::ifstream f(fn);
    if (!f.is_open()) throw std::runtime_error("[ERROR]: File not found.");  // io.hpp:58-62
    for (size_t i = 0; i < data.size(); i++)
        if (!(f >> data[i])) throw std::runtime_error("[ERROR]: Insufficient data in the file.");
}
static void loadTables(const std::string& dir, Tables& T) {
    std::vector<double> d(52 * 9);
    loadNumbers(dir + "/ifstream.txt", d);
    for (int i = 0; i < 9; i++)
        for (int j = 0; j < 52; j++) T.ifStream[j][i] = static_cast<bool>(d[i * 52 + j]);
    loadNumbers(dir + "/icsr.txt", d);
    for (int i = 0; i < 9; i++)
        for (int j = 0; j < 52; j++) T.icsr[j][i] = static_cast<int>(d[i * 52 + j]) - 1;
    std::vector<double> d2(52 * 2);
    loadNumbers(dir + "/xy_norm.txt", d2);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 52; j++) T.xy_norm[j][i] = d2[i * 52 + j];
    std::vector<double> d3(52);
    loadNumbers(dir + "/bnd_types.txt", d3);
    for (int j = 0; j < 52; j++) T.bnd_types[j] = static_cast<int>(d3[j]);
}

// ---------------------------------------------------------------- index helpers
// utils.hpp:379-408 sub2ind (first subscript fastest; -1 when out of range)
static int64_t sub2ind(const std::vector<int>& sizes, const std::vector<int>& subs) {
    if (subs.size() != sizes.size()) return -1;
    int64_t idx = 0, mul = 1;
    for (size_t i = 0; i < subs.size(); ++i) {
        if (subs[i] < 0 || subs[i] >= sizes[i]) return -1;
        idx += static_cast<int64_t>(subs[i]) * mul;
        mul *= sizes[i];
    }
    return idx;
}
// utils.hpp:365-376
static std::vector<int> ind2sub(const std::vector<int>& size, int64_t index) {
    std::vector<int> s(size.size());
    for (size_t i = 0; i < size.size(); i++) {
        s[i] = static_cast<int>(index % size[i]);
        index /= size[i];
    }
    return s;
}

// ---------------------------------------------------------------- grid
struct Grid {
    int nx = 0, ny = 0, maxLevel = 0;
    int maxAllowedLevel = 2;                   // grid.hpp:1180 (a parameter here)
    bool isMultigrid = false;
    int64_t size = 0;                          // globalGridSize == localGridSize (np = 1)
    std::vector<int64_t> gridID;               // grid.hpp:167
    std::vector<int> gridType, gridLevel, gridIsBuffer;
    std::vector<int> gridIJ;                   // [c*2+0] = x (col), [c*2+1] = y (row)
    std::vector<int64_t> gridConnect;          // [c*9+ic]
    std::vector<int64_t> solidID;
    std::vector<double> locpore, kn;
    std::vector<int> bndTypes;
    std::vector<int64_t> parentChild;          // rows of {cell, tl, tr, bl, br}
    std::map<int, std::vector<std::vector<int64_t>>> coarseBuffer, fineBuffer;
    std::vector<int64_t> reconstructed;        // ny*nx, reconstructed.dat
};

struct Shape {  // shape.hpp:48-77 (row-major [i=y][j=x])
    int Nx = 0, Ny = 0;
    double resolution = 0;
    std::vector<int> domain;
    std::vector<double> Kn, LocPore;
};

// grid.hpp:401-424 processSinglegrid
static void processSinglegrid(const Shape& S, Grid& G, std::vector<int64_t>& img) {
    img.assign((size_t)S.Ny * S.Nx, 0);
    int64_t count = 1;
    for (int i = 0; i < S.Ny; i++)
        for (int j = 0; j < S.Nx; j++) {
            if (S.domain[(size_t)i * S.Nx + j] == 0) {
                img[(size_t)i * S.Nx + j] = -count;
                G.solidID.push_back(sub2ind({S.Ny, S.Nx}, {j, i}));
            } else
                img[(size_t)i * S.Nx + j] = count;
            count++;
        }
}

// Declared here, defined in oracle/multiblock_oracle.cpp (grid.hpp:313-399,
// 427-889, 974-1108, 1160-1269).
void processMultigrid(const Shape& S, Grid& G, std::vector<int64_t>& img, bool torus);
void rebuildGridMulti(Grid& G, const std::vector<int64_t>& img, std::vector<int64_t>& linInd, bool torus);
void determineParentChild(Grid& G);
void orderBufferIndex(Grid& G);
struct Sim;
void multigridTaus(Sim& s);

// grid.hpp:427-889 for maxLevel == 1 (every unique image value is one pixel;
// no interfaces so no buffers).  Cells are the pore pixels in row-major order.
static void rebuildGridSingle(Grid& G, const std::vector<int64_t>& img, std::vector<int64_t>& linInd) {
    const int ny = G.ny, nx = G.nx;
    G.maxLevel = 1;
    linInd.clear();
    std::vector<char> pore;
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            linInd.push_back(sub2ind({ny, nx, 1}, {j, i, 0}));
            pore.push_back(img[(size_t)i * nx + j] < 0 ? 0 : 1);
        }
    int64_t n = 0;
    for (char p : pore) n += p;
    G.size = n;
    G.gridID.assign(n, 0);
    G.gridIJ.assign(2 * n, 0);
    G.gridLevel.assign(n, 1);
    G.gridType.assign(n, 0);
    G.gridIsBuffer.assign(n, 0);
    G.gridConnect.assign(9 * n, 0);
    G.locpore.assign(n, 0);
    G.kn.assign(n, 0);
    int64_t c = 0;
    for (size_t i = 0; i < linInd.size(); i++)
        if (pore[i]) {
            G.gridID[c] = linInd[i];
            std::vector<int> s = ind2sub({ny, nx, 1}, linInd[i]);
            G.gridIJ[2 * c] = s[0];
            G.gridIJ[2 * c + 1] = s[1];
            c++;
        }
}

// grid.hpp:891-899
static void getProperties(const Shape& S, Grid& G) {
    for (int64_t i = 0; i < G.size; i++) {
        size_t p = (size_t)G.gridIJ[2 * i + 1] * S.Nx + G.gridIJ[2 * i];
        G.locpore[i] = S.LocPore[p];
        G.kn[i] = S.Kn[p];
    }
}

// grid.hpp:916-972: same-level neighbour with x/y wrap; -1 = a pixel of the
// image at that level that is not a cell (solid), -2 = nothing there.
static void buildConnections(Grid& G, const std::vector<int64_t>& linInd) {
    std::vector<int> vs = {G.ny, G.nx, G.maxLevel};
    std::unordered_map<int64_t, int64_t> idx;
    idx.reserve(G.size * 2);
    for (int64_t c = 0; c < G.size; c++) idx[G.gridID[c]] = c;
    std::unordered_set<int64_t> lin(linInd.begin(), linInd.end());
    for (int64_t c = 0; c < G.size; c++) {
        int i = G.gridIJ[2 * c], j = G.gridIJ[2 * c + 1], dim = G.gridLevel[c];
        int jump = (int)std::pow(2, dim - 1);
        for (int k = 0; k < 9; k++) {
            int ip = i + EX[k] * jump, jp = j + EY[k] * jump;
            jp = (jp <= -1) ? jp + G.ny : jp;
            jp = (jp >= G.ny) ? jp - G.ny : jp;
            ip = (ip <= -1) ? ip + G.nx : ip;
            ip = (ip >= G.nx) ? ip - G.nx : ip;
            int64_t ng = sub2ind(vs, {ip, jp, dim - 1});
            auto it = idx.find(ng);
            if (it != idx.end())
                G.gridConnect[9 * c + k] = it->second;
            else
                G.gridConnect[9 * c + k] = lin.count(ng) ? -1 : -2;
        }
    }
}

// grid.hpp:1111-1157: 9-character code over neighbours {7,3,6,4,0,2,8,1,5},
// MSB first, '0' where gridConnect == -1; 511 -> -1; unknown -> throw.
static void getBoundaryTypes(Grid& G, const Tables& T) {
    static const int order[9] = {7, 3, 6, 4, 0, 2, 8, 1, 5};
    G.bndTypes.assign(G.size, -1);
    for (int64_t i = 0; i < G.size; i++) {
        if (G.gridIsBuffer[i] != 0) continue;
        int code = 0;
        for (int b = 0; b < 9; b++) code = (code << 1) | (G.gridConnect[9 * i + order[b]] == -1 ? 0 : 1);
        if (code != 511) {
            int t = -1;
            for (int k = 0; k < 52; k++)
                if (T.bnd_types[k] == code) { t = k; break; }
            if (t < 0) throw std::runtime_error("[ERROR]: Boundary type is not contained.");
            G.bndTypes[i] = t;
        }
    }
}

// ---------------------------------------------------------------- solver
struct Sim {
    Globals g;
    Tables T;
    orc::Shape S;
    Grid G;
    int NX = 0, NY = 0;
    int t = 0;
    double diff = 1.0;
    std::vector<double> diff_over_time;
    int collision = 2, wall = 2;
    // false: the reference's seam rule (coreLB.hpp:301-331); true: torus, the
    // links across both seams are classified like interior links.
    bool periodicY = false;

    std::vector<double> rho, rho_old, ux, uy, Fx, Fy, uxsq, uysq, usq, usq_old, usqMinusOld;
    std::vector<double> f, fb, ftemp, feq, fneq, psi, taus;
    std::vector<int64_t> streamID, streamLOC, boundaryID, periodicID, periodicI2;
    std::vector<int> streamIC, boundaryIC, periodicIC;
    std::vector<int64_t> gpIn, gpIn2, gpOut, gpOut2;
    std::vector<std::vector<int>> sfIC, bfIC;
    std::vector<std::vector<int64_t>> sfI2;

    // ---- coreLB.hpp:482-502 calculateSingleGridTaus (fluid pixels row-major)
    void singleGridTaus() {
        taus.assign(G.size, 0.0);
        int64_t c = 0;
        for (int i = 0; i < S.Ny; i++)
            for (int j = 0; j < S.Nx; j++) {
                size_t p = (size_t)i * S.Nx + j;
                if (S.domain[p] == 1) {
                    taus[c] = 0.5 + std::sqrt(6.0 / M_PI) * S.LocPore[p] * S.Kn[p] / (1 + 2 * S.Kn[p]);
                    c++;
                }
            }
    }

    // ---- coreLB.hpp:231-378 link tables (np = 1: every neighbour in-rank)
    void buildLinks() {
        sfIC.assign(G.size, {});
        bfIC.assign(G.size, {});
        sfI2.assign(G.size, {});
        int ny_c = (NY % 2 == 1) ? (NY + 1) / 2 - 1 : NY / 2 - 1;
        const int i_in = ny_c, j_in = 0, i_out = ny_c, j_out = NX - 1;
        const int seam = (int)std::pow(2, G.maxLevel - 1);
        for (int64_t i = 0; i < G.size; i++) {
            for (int ic = 0; ic < NC; ic++) {
                int64_t c = G.gridConnect[9 * i + ic];
                int bt = G.bndTypes[i];
                if (c >= 0 && c < G.size) {
                    bool xp = std::abs(G.gridIJ[2 * i] - G.gridIJ[2 * c]) > seam;
                    bool yp = std::abs(G.gridIJ[2 * i + 1] - G.gridIJ[2 * c + 1]) > seam;
                    if (periodicY) xp = yp = false;  // torus: seams like interior links
                    if (!xp && !yp) {
                        bool isStream = (bt == -1) || T.ifStream[bt][ic];
                        if (isStream) {
                            streamID.push_back(i); streamIC.push_back(ic); streamLOC.push_back(c);
                            sfIC[i].push_back(ic); sfI2[i].push_back(c);
                        } else {
                            boundaryID.push_back(i); boundaryIC.push_back(ic);
                            bfIC[i].push_back(ic);
                        }
                    } else if (xp) {
                        periodicID.push_back(i); periodicI2.push_back(c); periodicIC.push_back(ic);
                        sfIC[i].push_back(ic); sfI2[i].push_back(c);
                    }
                } else if (c != -1 && (c < 0 || c >= G.size)) {
                    bool isStream = (bt == -1) || T.ifStream[bt][ic];
                    if (!isStream) {
                        boundaryID.push_back(i); boundaryIC.push_back(ic);
                        bfIC[i].push_back(ic);
                    }
                } else if (c == -1 && (bt < 0 || !T.ifStream[bt][ic])) {
                    // bt < 0 only for a buffer cell next to a wall across a
                    // periodic seam, where the reference reads ifStream[-1]:
                    // defined here as a wall link
                    boundaryID.push_back(i); boundaryIC.push_back(ic);
                    bfIC[i].push_back(ic);
                }
            }
            int gx = G.gridIJ[2 * i], gy = G.gridIJ[2 * i + 1];
            if (gy == i_in && gx == j_in) gpIn.push_back(i);
            else if (gy == i_out && gx == j_out) gpOut.push_back(i);
            else if (gy == i_in && gx == j_in + 1) gpIn2.push_back(i);
            else if (gy == i_out && gx == j_out - 1) gpOut2.push_back(i);
        }
    }

    // ---- coreLB.hpp:381-417
    void initialize() {
        const int64_t n = G.size;
        rho.assign(n, g.rhoin); rho_old.assign(n, 0); ux.assign(n, 0); uy.assign(n, 0);
        uxsq.assign(n, 0); uysq.assign(n, 0); usq.assign(n, 0); usq_old.assign(n, 0);
        usqMinusOld.assign(n, 0);
        f.assign(n * NC, 0); feq.assign(n * NC, 0); fb.assign(n * NC, 0);
        ftemp.assign(n * NC, 0); fneq.assign(n * NC, 0);
        psi.assign(n, 0); Fx.assign(n, 0); Fy.assign(n, 0);
        for (int64_t k = 0; k < n; k++)
            for (int ic = 0; ic < NC; ic++) {
                double cu = EX[ic] * ux[k] + EY[ic] * uy[k];
                f[k * NC + ic] = W[ic] * rho[k] * (1.0 + F1 * cu + F2 * (cu * cu) - F3 * usq[k]);
            }
        macroscopic();
    }

    // ---- coreLB.hpp:420-430
    // (std::pow as the reference writes it: pow(x, 2) is x * x exactly)
    void equilibrium() {
#pragma omp parallel for
        for (int64_t k = 0; k < G.size; k++)
            for (int ic = 0; ic < NC; ic++) {
                double cu = EX[ic] * ux[k] + EY[ic] * uy[k];
                feq[k * NC + ic] = W[ic] * rho[k] * (1.0 + F1 * cu + F2 * std::pow(cu, 2) - F3 * usq[k]);
            }
    }

    // ---- coreLB.hpp:634-643 (accumulates into mat3).  The collisions keep
    // the reference's containers: heap vectors for the rate matrix and the
    // four work vectors of each thread, refilled per cell -- the same
    // arithmetic as plain arrays, and the cost the timed CPU baseline should
    // carry (the reference's collisions are 65-82% of its transport step).
    template <class A, class B, class Out>
    static void matMulLin(const A& m1, const B& m2, Out& m3) {
        for (int u = 0; u < NC; u++)
            for (int v = 0; v < NC; v++) m3[u] += m1[NC * u + v] * m2[v];
    }
    // MRT rates diag(1, 1/1.1, 1/1.2, 1/td, 1/tq, 1/td, 1/tq, 1/ts, 1/ts) as a
    // dense matrix (coreLB.hpp:705-715, 760-770, 599-609).
    template <class Mat>
    static void fillR(Mat& R, double tau_d, double tau_q, double tau_s) {
        for (int u = 0; u < 81; u++) R[u] = 0.0;
        const double d[9] = {1.0 / 1.0, 1.0 / 1.1, 1.0 / 1.2, 1.0 / tau_d, 1.0 / tau_q,
                             1.0 / tau_d, 1.0 / tau_q, 1.0 / tau_s, 1.0 / tau_s};
        for (int u = 0; u < 9; u++) R[u * 10] = d[u];
    }
    double tauQ(int64_t k) const {
        return 0.5 + (3.0 + M_PI * (2.0 * taus[k] - 1.0) * (2.0 * taus[k] - 1.0) * g.A2) / (8.0 * (2.0 * taus[k] - 1.0));
    }
    double tauD(int64_t k) const {  // coreLB.hpp:604, 710
        return 0.5 + (3.0 / 2.0) * std::sqrt((double)3.0) * 1.0 /
                         std::pow((double)(1.0 / std::sqrt((double)(g.mfp / g.Cl * 1.0 / (1.0 + 2.0 * G.kn[k]))) * 2.0), (double)2.0);
    }

    // ---- coreLB.hpp:646-675 (all cells, every call)
    void regularize() {
        std::vector<double> P(4, 0.0), H(4, 0.0), ccr(4, 0.0);
#pragma omp parallel for firstprivate(P, H, ccr)
        for (int64_t k = 0; k < G.size; k++) {
            std::fill(P.begin(), P.end(), 0.0);
            for (int ic = 0; ic < NC; ic++) {
                ccr = {(double)(EX[ic] * EX[ic]), (double)(EX[ic] * EY[ic]), (double)(EX[ic] * EY[ic]),
                       (double)(EY[ic] * EY[ic])};
                double d = f[k * NC + ic] - feq[k * NC + ic];
                for (int u = 0; u < 4; u++) P[u] += d * ccr[u];
            }
            for (int ic = 0; ic < NC; ic++) {
                H = {EX[ic] * EX[ic] / CS2 - 1, EX[ic] * EY[ic] / CS2, EX[ic] * EY[ic] / CS2, EY[ic] * EY[ic] / CS2 - 1};
                fneq[k * NC + ic] = W[ic] * (1. / (2. * CS2)) * (P[0] * H[0] + P[1] * H[1] + P[2] * H[2] + P[3] * H[3]);
            }
        }
    }

    // ---- coreLB.hpp:678-691
    void SRT(int lvl) {
#pragma omp parallel for
        for (int64_t i = 0; i < G.size; i++)
            if (G.gridLevel[i] == lvl)
                for (int ic = 0; ic < NC; ic++)
                    f[i * NC + ic] = f[i * NC + ic] - 1.0 / g.tau_SRT * (f[i * NC + ic] - feq[i * NC + ic]);
    }
    // ---- coreLB.hpp:694-742
    void MRT(int lvl) {
        std::vector<double> R(81), t1(NC), t2(NC), t3(NC), t4(NC);
#pragma omp parallel for firstprivate(R, t1, t2, t3, t4)
        for (int64_t k = 0; k < G.size; k++) {
            if (G.gridLevel[k] != lvl) continue;
            std::fill(t2.begin(), t2.end(), 0.0);
            std::fill(t3.begin(), t3.end(), 0.0);
            std::fill(t4.begin(), t4.end(), 0.0);
            fillR(R, tauD(k), tauQ(k), taus[k]);
            for (int ic = 0; ic < NC; ic++) t1[ic] = f[k * NC + ic] - feq[k * NC + ic];
            matMulLin(MmV, t1, t2);
            matMulLin(R, t2, t3);
            matMulLin(MinvV, t3, t4);
            for (int ic = 0; ic < NC; ic++) f[k * NC + ic] = f[k * NC + ic] - t4[ic];
        }
    }
    // ---- coreLB.hpp:744-792
    void RMRT(int lvl) {
        regularize();
        std::vector<double> R(81), t1(NC), t2(NC), t3(NC), t4(NC);
#pragma omp parallel for firstprivate(R, t1, t2, t3, t4)
        for (int64_t k = 0; k < G.size; k++) {
            if (G.gridLevel[k] != lvl) continue;
            std::fill(t2.begin(), t2.end(), 0.0);
            std::fill(t3.begin(), t3.end(), 0.0);
            std::fill(t4.begin(), t4.end(), 0.0);
            fillR(R, 1.0, tauQ(k), taus[k]);
            for (int ic = 0; ic < NC; ic++) t1[ic] = fneq[k * NC + ic];
            matMulLin(MmV, t1, t2);
            matMulLin(R, t2, t3);
            matMulLin(MinvV, t3, t4);
            for (int ic = 0; ic < NC; ic++) f[k * NC + ic] = feq[k * NC + ic] + fneq[k * NC + ic] - t4[ic];
        }
    }
    void collide(int lvl) {  // coreLB.hpp:809-812 + mantiisApp.hpp:180-203
        if (collision == 2) RMRT(lvl);
        else if (collision == 1) MRT(lvl);
        else SRT(lvl);
    }

    // ---- coreLB.hpp:1233-1246
    void zeroDown(int lvl) {
        for (int64_t i = 0; i < G.size; i++)
            if (G.gridLevel[i] == lvl)
                for (int ic = 0; ic < NC; ic++) f[i * NC + ic] = 0;
    }
    // ---- coreLB.hpp:834-856 (push; the MPI exchange is empty at np = 1)
    void stream(int lvl) {
#pragma omp parallel for
        for (size_t i = 0; i < streamID.size(); i++)
            if (G.gridLevel[streamID[i]] == lvl)
                f[streamLOC[i] * NC + streamIC[i]] = ftemp[streamID[i] * NC + streamIC[i]];
    }
    // ---- coreLB.hpp:859-865
    void BBWall(int lvl) {
        for (size_t i = 0; i < boundaryID.size(); i++)
            if (G.gridLevel[boundaryID[i]] == lvl)
                f[boundaryID[i] * NC + OPP[boundaryIC[i]]] = ftemp[boundaryID[i] * NC + boundaryIC[i]];
    }
    // ---- coreLB.hpp:918-933 (sequential order: r part, then 1-r part, per link)
    void SRBBWall(int lvl) {
        for (size_t i = 0; i < boundaryID.size(); i++)
            if (G.gridLevel[boundaryID[i]] == lvl) {
                double r = 0.4987 * std::pow(G.kn[boundaryID[i]], -0.131) - 0.25;
                const int bt = G.bndTypes[boundaryID[i]];
                // typeless wall link of a buffer cell: full bounce-back target
                int srOpp = bt < 0 ? OPP[boundaryIC[i]] : T.icsr[bt][boundaryIC[i]];
                f[boundaryID[i] * NC + OPP[boundaryIC[i]]] += r * ftemp[boundaryID[i] * NC + boundaryIC[i]];
                f[boundaryID[i] * NC + srOpp] += (1 - r) * ftemp[boundaryID[i] * NC + boundaryIC[i]];
            }
    }
    // ---- coreLB.hpp:936-967
    void MDBBWall(int lvl) {
        for (size_t i = 0; i < boundaryID.size(); i++) {
            if (G.gridLevel[boundaryID[i]] != lvl) continue;
            int64_t k1 = boundaryID[i];
            int k2 = boundaryIC[i];
            if (G.bndTypes[k1] < 0) {  // typeless wall link: plain bounce-back
                f[k1 * NC + OPP[k2]] = ftemp[k1 * NC + k2];
                continue;
            }
            double Kdeno = 0.0, Kno = 0.0;
            double nx = T.xy_norm[G.bndTypes[k1]][0], ny = T.xy_norm[G.bndTypes[k1]][1];
            for (int ic = 0; ic < NC; ic++) {
                if ((EY[ic] * ny + EX[ic] * nx) > 0) Kdeno += feq[k1 * NC + ic];
                else if ((EY[ic] * ny + EX[ic] * nx) < 0) Kno += fb[k1 * NC + ic];
            }
            double r = 0.3222 * std::pow(G.kn[k1], -0.217) - 0.25;
            f[k1 * NC + OPP[k2]] = r * ftemp[k1 * NC + k2] + (1 - r) * Kno / Kdeno * feq[k1 * NC + OPP[k2]];
        }
    }
    void wallBC(int lvl) {  // mantiisApp.hpp:206-234
        if (wall == 2) SRBBWall(lvl);
        else if (wall == 1) MDBBWall(lvl);
        else BBWall(lvl);
    }
    // ---- coreLB.hpp:969-977
    void periodic(int lvl) {
        for (size_t i = 0; i < periodicID.size(); i++)
            if (G.gridLevel[periodicID[i]] == lvl)
                f[periodicI2[i] * NC + periodicIC[i]] = ftemp[periodicID[i] * NC + periodicIC[i]];
    }
    // ---- coreLB.hpp:980-1009 (note W[2] on the diagonals)
    void generalizedPeriodicAdsorption() {
        for (size_t i = 0; i < gpIn.size(); i++) {
            int64_t a = gpIn[i] * NC, b = gpOut[i] * NC;
            f[a + 1] = W[1] * g.rhoin;
            f[a + 5] = W[2] * g.rhoin;
            f[a + 8] = W[2] * g.rhoin;
            f[b + 3] = W[1] * g.rhoin;
            f[b + 6] = W[2] * g.rhoin;
            f[b + 7] = W[2] * g.rhoin;
        }
    }

    // ---- coreLB.hpp:1018-1054
    void macroscopic() {
#pragma omp parallel for
        for (int64_t i = 0; i < G.size; i++) {
            double jx = 0, jy = 0, r = 0;
            for (int ic = 0; ic < NC; ic++) {
                r += f[i * NC + ic];
                jx += f[i * NC + ic] * EX[ic];
                jy += f[i * NC + ic] * EY[ic];
            }
            rho[i] = r;
            ux[i] = jx / rho[i] + g.Fbody / (2.0 * rho[i]) * std::pow(2.0, G.gridLevel[i] - 1);
            uy[i] = jy / rho[i];
            uxsq[i] = ux[i] * ux[i];
            uysq[i] = uy[i] * uy[i];
            usq[i] = uxsq[i] + uysq[i];
        }
    }
    // ---- coreLB.hpp:1056-1067
    void density() {
        for (int64_t i = 0; i < G.size; i++) {
            double r = 0;
            for (int ic = 0; ic < NC; ic++) r += f[i * NC + ic];
            rho[i] = r;
        }
    }
    // ---- coreLB.hpp:1069-1099
    void velocity() {
        for (int64_t i = 0; i < G.size; i++) {
            double jx = 0, jy = 0;
            for (int ic = 0; ic < NC; ic++) {
                jx += f[i * NC + ic] * EX[ic];
                jy += f[i * NC + ic] * EY[ic];
            }
            double sc = std::pow(2.0, G.gridLevel[i] - 1);
            ux[i] = jx / rho[i] + Fx[i] / (2.0 * rho[i]) * sc;
            uy[i] = jy / rho[i] + Fy[i] / (2.0 * rho[i]) * sc;
            uxsq[i] = ux[i] * ux[i];
            uysq[i] = uy[i] * uy[i];
            usq[i] = uxsq[i] + uysq[i];
        }
    }
    // ---- coreLB.hpp:515-531 (negative argument -> psi = argument)
    int psiNegative = 0;
    void pseudoPotential() {
        psiNegative = 0;
        for (int64_t i = 0; i < G.size; i++) {
            double m = rho[i] / g.Mw_lu;
            double acc = (m * g.R_lu * g.T_lu / (1 - g.b_lu * m) -
                          g.a_alpha_lu * m * m / (1 + 2 * g.b_lu * m - g.b_lu * g.b_lu * m * m) - (1.0 / 3.0) * rho[i]) /
                         (3.0 * g.Gff);
            if (acc < 0) { psi[i] = acc; psiNegative++; }
            else psi[i] = std::sqrt(acc);
        }
    }
    // ---- coreLB.hpp:533-566
    void forces() {
        for (int64_t i = 0; i < G.size; i++) {
            double fxff = 0, fyff = 0, fxfw = 0, fyfw = 0;
            for (size_t k = 0; k < sfIC[i].size(); k++) {
                int ic = sfIC[i][k];
                int64_t nb = sfI2[i][k];
                fxff += (-6.0) * g.Gff * psi[i] * WF[ic] * psi[nb] * EX[ic];
                fyff += (-6.0) * g.Gff * psi[i] * WF[ic] * psi[nb] * EY[ic];
            }
            for (size_t k = 0; k < bfIC[i].size(); k++) {
                int ic = bfIC[i][k];
                fxfw += (-1.0) * g.Gfs * psi[i] * psi[i] * WF[ic] * EX[ic];
                fyfw += (-1.0) * g.Gfs * psi[i] * psi[i] * WF[ic] * EY[ic];
            }
            Fx[i] = fxff + fxfw + g.Fbody;
            Fy[i] = fyff + fyfw;
        }
    }
    // ---- coreLB.hpp:568-630 Guo source in MRT space, f += Minv (I - R/2) M S
    void sourceTerm() {
        static const double I81[81] = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1,
                                       0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0,
                                       0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0,
                                       0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1};
        for (int64_t k = 0; k < G.size; k++) {
            double uF[4] = {ux[k] * Fx[k], ux[k] * Fy[k], uy[k] * Fx[k], uy[k] * Fy[k]};
            double S9[9];
            for (int ic = 0; ic < NC; ic++) {
                double cDotF = EX[ic] * Fx[k] + EY[ic] * Fy[k];
                double ccs[4] = {(double)EX[ic] * EX[ic] - CS2, (double)EX[ic] * EY[ic],
                                 (double)EY[ic] * EX[ic], (double)EY[ic] * EY[ic] - CS2};
                S9[ic] = W[ic] * (cDotF / CS2 + (uF[0] * ccs[0] + uF[2] * ccs[1] + uF[1] * ccs[2] + uF[3] * ccs[3]) / (CS2 * CS2));
            }
            double R[81], t1[81], t2[9] = {0}, t3[9] = {0}, t4[9] = {0};
            fillR(R, tauD(k), tauQ(k), taus[k]);
            for (int u = 0; u < 81; u++) t1[u] = I81[u] - 0.5 * R[u];
            matMulLin(Mm, S9, t2);
            matMulLin(t1, t2, t3);
            matMulLin(Minv, t3, t4);
            for (int ic = 0; ic < NC; ic++) f[k * NC + ic] = f[k * NC + ic] + t4[ic];
        }
    }

    // ---- coreLB.hpp:1248-1267
    void velocityDifference() {
        double s1 = 0, s2 = 0;
        for (int64_t i = 0; i < G.size; i++) {
            usqMinusOld[i] = std::abs(std::sqrt(usq[i]) - std::sqrt(usq_old[i]));
            s1 += usqMinusOld[i];
            s2 += std::sqrt(usq[i]);
        }
        diff = s1 / s2;
        diff_over_time.push_back(diff);
    }
    // ---- coreLB.hpp:1269-1283
    void densityDifference() {
        double m = 0.0;
        for (int64_t i = 0; i < G.size; i++) {
            double d = std::abs(rho[i] - rho_old[i]) / rho_old[i];
            if (d > m) m = d;
        }
        if (m < 1.0e-15) m = 1.0e-5;
        diff = m;
        diff_over_time.push_back(diff);
    }

    // ---- coreLB.hpp:1101-1231 (defined in multiblock_oracle.cpp)
    void interpolateBlockInterface(int l1, int l2);

    // ---- coreLB.hpp:1310-1331
    void evolutionStepOfMultiBlock(int lvl) {
        int rep = (int)std::pow(2, G.maxLevel - lvl);
        for (int i = 0; i < rep; i++) {
            fb = f;
            collide(lvl);
            ftemp = f;
            zeroDown(lvl);
            stream(lvl);
            wallBC(lvl);
            periodic(lvl);
        }
        if (lvl > 1) evolutionStepOfMultiBlock(lvl - 1);
        if (lvl != G.maxLevel) interpolateBlockInterface(lvl, lvl + 1);
    }

    // ---- mantiisApp.hpp:239-284
    void transportStep() {
        usq_old = usq;
        equilibrium();
        evolutionStepOfMultiBlock(G.maxLevel);
        macroscopic();
        velocityDifference();
    }
    void adsorptionStep() {
        rho_old = rho;
        density();
        pseudoPotential();
        forces();
        velocity();
        equilibrium();
        MRT(1);
        sourceTerm();
        ftemp = f;
        zeroDown(1);
        stream(1);
        BBWall(1);
        periodic(1);
        generalizedPeriodicAdsorption();
        densityDifference();
    }
    void coupledStep() {
        rho_old = rho;
        density();
        pseudoPotential();
        forces();
        velocity();
        equilibrium();
        RMRT(1);
        sourceTerm();
        ftemp = f;
        zeroDown(1);
        stream(1);
        SRBBWall(1);
        periodic(1);
        velocityDifference();
    }
    // ---- mantiisApp.hpp:286-403 run loop: step, break on diff < tol, t++.
    int runLoop(int kind, int iters, double tol) {
        int done = 0;
        for (int i = 0; i < iters; i++) {
            if (kind == 0) transportStep();
            else if (kind == 1) adsorptionStep();
            else coupledStep();
            done++;
            if (diff < tol) break;
            t++;
        }
        return done;
    }
    int run(int kind, int iters, double tol) {
        if (kind != 2) return runLoop(kind, iters, tol);
        // mantiisApp.hpp:376-393 coupled: 100 adsorption steps with Fbody = 0,
        // then reset (ux filled twice, uy untouched; rho, feq reset).
        double fb0 = g.Fbody;
        g.Fbody = 0;
        int n = runLoop(1, 100, tol);
        g.Fbody = fb0;
        std::fill(ux.begin(), ux.end(), 0.0);
        std::fill(rho.begin(), rho.end(), g.rhoin);
        std::fill(feq.begin(), feq.end(), 0.0);
        return n + runLoop(2, iters, tol);
    }

    // ---- coreLB.hpp:1344-1411 completeSingleGridDomain (row-major, solids 0)
    void completeDomain(double* oux, double* ouy, double* orho) const {
        std::vector<int64_t> sid = G.solidID;
        std::sort(sid.begin(), sid.end());
        int64_t c = 0;
        for (int64_t i = 0; i < (int64_t)NX * NY; i++) {
            if (std::binary_search(sid.begin(), sid.end(), i)) {
                oux[i] = 0; ouy[i] = 0; orho[i] = 0;
            } else {
                oux[i] = ux[c]; ouy[i] = uy[c]; orho[i] = rho[c];
                c++;
            }
        }
    }
};

}  // namespace orc

#include "shape_oracle.inc"
#include "multiblock_oracle.inc"

// ============================================================================
// C ABI used by the Python tests (ctypes) and by bench.py's CPU leg.
// ============================================================================
using orc::Sim;

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Inputs are row-major [y][x] (the in-memory Shape::domain layout after the
// transposed load of shape.hpp:253-285).  multiblock=1 runs the quadtree path
// with the reference's maxAllowedLevel = 2 (grid.hpp:1180); multiblock = k >= 2
// runs it with maxAllowedLevel = k.
void* orc_create(int N, const int* pore, const double* kn, const double* locpore, double dx,
                 double Fbody, double T, double P, const char* tables_dir, int multiblock, int periodic_y) {
    try {
        Sim* s = new Sim();
        s->periodicY = periodic_y != 0;
        s->g.Cl = dx; s->g.Fbody = Fbody; s->g.T_phy = T; s->g.P_phy = P;
        s->g.recompute();
        orc::loadTables(tables_dir, s->T);
        s->NX = N; s->NY = N;
        s->S.Nx = N; s->S.Ny = N; s->S.resolution = dx;
        s->S.domain.assign(pore, pore + (size_t)N * N);
        s->S.Kn.assign(kn, kn + (size_t)N * N);
        s->S.LocPore.assign(locpore, locpore + (size_t)N * N);
        s->G.nx = N; s->G.ny = N; s->G.isMultigrid = multiblock != 0;
        s->G.maxAllowedLevel = multiblock >= 2 ? multiblock : 2;
        std::vector<int64_t> img, linInd;
        if (multiblock) {
            orc::processMultigrid(s->S, s->G, img, s->periodicY);
            orc::rebuildGridMulti(s->G, img, linInd, s->periodicY);
        } else {
            orc::processSinglegrid(s->S, s->G, img);
            orc::rebuildGridSingle(s->G, img, linInd);
        }
        s->G.reconstructed = img;
        orc::getProperties(s->S, s->G);
        orc::buildConnections(s->G, linInd);
        if (multiblock) {
            orc::determineParentChild(s->G);
            orc::orderBufferIndex(s->G);
        }
        orc::getBoundaryTypes(s->G, s->T);
        if (!multiblock) s->singleGridTaus();
        else orc::multigridTaus(*s);
        s->buildLinks();
        s->initialize();
        return s;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void orc_destroy(void* h) { delete static_cast<Sim*>(h); }

// globals: rhoin, mfp, Cf, T_lu, a_alpha_lu, ZG, Crho, Cu, Cl, cubic_ok
void orc_globals(void* h, double* out) {
    Sim* s = static_cast<Sim*>(h);
    const orc::Globals& g = s->g;
    double v[10] = {g.rhoin, g.mfp, g.Cf, g.T_lu, g.a_alpha_lu, g.ZG, g.Crho, g.Cu, g.Cl, g.cubic_ok ? 1.0 : 0.0};
    std::memcpy(out, v, sizeof(v));
}
// Standalone recompute for arbitrary (dx, T, P): same 10 values.
void orc_globals_for(double dx, double T, double P, double* out) {
    orc::Globals g;
    g.Cl = dx; g.T_phy = T; g.P_phy = P;
    g.recompute();
    double v[10] = {g.rhoin, g.mfp, g.Cf, g.T_lu, g.a_alpha_lu, g.ZG, g.Crho, g.Cu, g.Cl, g.cubic_ok ? 1.0 : 0.0};
    std::memcpy(out, v, sizeof(v));
}

int64_t orc_cell_count(void* h) { return static_cast<Sim*>(h)->G.size; }
int orc_max_level(void* h) { return static_cast<Sim*>(h)->G.maxLevel; }
void orc_set_operators(void* h, int coll, int wall) {
    Sim* s = static_cast<Sim*>(h);
    s->collision = coll;
    s->wall = wall;
}
void orc_set_interaction(void* h, double Gff, double Gfs) {
    Sim* s = static_cast<Sim*>(h);
    s->g.Gff = Gff;
    s->g.Gfs = Gfs;
}

// grid arrays (compact cell order)
void orc_grid(void* h, int64_t* gridID, int* gridIJ, int* level, int64_t* connect, int* bnd, int* isBuffer, int* type) {
    Sim* s = static_cast<Sim*>(h);
    const orc::Grid& G = s->G;
    for (int64_t c = 0; c < G.size; c++) {
        if (gridID) gridID[c] = G.gridID[c];
        if (gridIJ) { gridIJ[2 * c] = G.gridIJ[2 * c]; gridIJ[2 * c + 1] = G.gridIJ[2 * c + 1]; }
        if (level) level[c] = G.gridLevel[c];
        if (bnd) bnd[c] = G.bndTypes[c];
        if (isBuffer) isBuffer[c] = G.gridIsBuffer[c];
        if (type) type[c] = G.gridType[c];
        if (connect)
            for (int k = 0; k < 9; k++) connect[9 * c + k] = G.gridConnect[9 * c + k];
    }
}
void orc_reconstructed(void* h, int64_t* img) {
    Sim* s = static_cast<Sim*>(h);
    std::copy(s->G.reconstructed.begin(), s->G.reconstructed.end(), img);
}
// buffers: which = 0 coarse, 1 fine; returns number of rows for key, fills rows*5
int64_t orc_buffer(void* h, int which, int key, int64_t* rows) {
    Sim* s = static_cast<Sim*>(h);
    auto& m = which == 0 ? s->G.coarseBuffer : s->G.fineBuffer;
    auto it = m.find(key);
    if (it == m.end()) return 0;
    if (rows)
        for (size_t r = 0; r < it->second.size(); r++)
            for (int k = 0; k < 5; k++) rows[5 * r + k] = it->second[r][k];
    return (int64_t)it->second.size();
}

// link lists: which = 0 stream (id, loc, ic), 1 boundary (id, -, ic), 2 periodic (id, i2, ic)
int64_t orc_links(void* h, int which, int64_t* id, int64_t* loc, int* ic) {
    Sim* s = static_cast<Sim*>(h);
    const std::vector<int64_t>* A = which == 0 ? &s->streamID : which == 1 ? &s->boundaryID : &s->periodicID;
    const std::vector<int64_t>* B = which == 0 ? &s->streamLOC : which == 2 ? &s->periodicI2 : nullptr;
    const std::vector<int>* C = which == 0 ? &s->streamIC : which == 1 ? &s->boundaryIC : &s->periodicIC;
    int64_t n = (int64_t)A->size();
    for (int64_t i = 0; i < n; i++) {
        if (id) id[i] = (*A)[i];
        if (loc && B) loc[i] = (*B)[i];
        if (ic) ic[i] = (*C)[i];
    }
    return n;
}
int64_t orc_gp_cells(void* h, int64_t* out4) {
    Sim* s = static_cast<Sim*>(h);
    int64_t n = (int64_t)s->gpIn.size();
    if (out4 && n > 0 && s->gpOut.size() > 0) {
        out4[0] = s->gpIn[0]; out4[1] = s->gpOut[0];
        out4[2] = s->gpIn2.empty() ? -1 : s->gpIn2[0];
        out4[3] = s->gpOut2.empty() ? -1 : s->gpOut2[0];
    }
    return n;
}

static std::vector<double>* field(Sim* s, const char* name) {
    std::string n(name);
    if (n == "f") return &s->f;
    if (n == "ftemp") return &s->ftemp;
    if (n == "feq") return &s->feq;
    if (n == "fneq") return &s->fneq;
    if (n == "fb") return &s->fb;
    if (n == "rho") return &s->rho;
    if (n == "rho_old") return &s->rho_old;
    if (n == "ux") return &s->ux;
    if (n == "uy") return &s->uy;
    if (n == "usq") return &s->usq;
    if (n == "usq_old") return &s->usq_old;
    if (n == "psi") return &s->psi;
    if (n == "Fx") return &s->Fx;
    if (n == "Fy") return &s->Fy;
    if (n == "taus") return &s->taus;
    if (n == "kn") return &s->G.kn;
    if (n == "locpore") return &s->G.locpore;
    return nullptr;
}
int64_t orc_get(void* h, const char* name, double* out) {
    auto* v = field(static_cast<Sim*>(h), name);
    if (!v) return -1;
    if (out) std::copy(v->begin(), v->end(), out);
    return (int64_t)v->size();
}
int64_t orc_set(void* h, const char* name, const double* in) {
    auto* v = field(static_cast<Sim*>(h), name);
    if (!v) return -1;
    std::copy(in, in + v->size(), v->begin());
    return (int64_t)v->size();
}

// Individual LB2D methods for phase-level checks.
int orc_method(void* h, const char* name, int lvl) {
    Sim* s = static_cast<Sim*>(h);
    std::string n(name);
    if (n == "initialize") s->initialize();
    else if (n == "equilibrium") s->equilibrium();
    else if (n == "regularize") s->regularize();
    else if (n == "SRTCollision") s->SRT(lvl);
    else if (n == "MRTCollisionNonRegularized") s->MRT(lvl);
    else if (n == "MRTCollisionRegularized") s->RMRT(lvl);
    else if (n == "collide") s->collide(lvl);
    else if (n == "zeroDown") s->zeroDown(lvl);
    else if (n == "stream") s->stream(lvl);
    else if (n == "BBWall") s->BBWall(lvl);
    else if (n == "SRBBWall") s->SRBBWall(lvl);
    else if (n == "MDBBWall") s->MDBBWall(lvl);
    else if (n == "periodicBoundary") s->periodic(lvl);
    else if (n == "generalizedPeriodicBoundaryAdsorption") s->generalizedPeriodicAdsorption();
    else if (n == "calculateMacroscopicPorperties") s->macroscopic();
    else if (n == "calculateDensity") s->density();
    else if (n == "calculateVelocity") s->velocity();
    else if (n == "getPseudoPotential") s->pseudoPotential();
    else if (n == "calculateForces") s->forces();
    else if (n == "applySourceTerm") s->sourceTerm();
    else if (n == "calculateVelocityDifference") s->velocityDifference();
    else if (n == "calculateDensityDifference") s->densityDifference();
    else if (n == "evolutionStepOfMultiBlock") s->evolutionStepOfMultiBlock(lvl);
    else if (n == "interpolateBlockInterface") s->interpolateBlockInterface(lvl, lvl + 1);
    else if (n == "copy_f_to_ftemp") s->ftemp = s->f;
    else if (n == "copy_f_to_fb") s->fb = s->f;
    else if (n == "copy_usq_to_usq_old") s->usq_old = s->usq;
    else if (n == "copy_rho_to_rho_old") s->rho_old = s->rho;
    else return -1;
    return 0;
}

int orc_run(void* h, int kind, int iters, double tol) {
    try {
        return static_cast<Sim*>(h)->run(kind, iters, tol);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}
int orc_steps_taken(void* h) { return static_cast<Sim*>(h)->t; }
double orc_diff(void* h) { return static_cast<Sim*>(h)->diff; }
int64_t orc_diff_history(void* h, double* out) {
    Sim* s = static_cast<Sim*>(h);
    if (out) std::copy(s->diff_over_time.begin(), s->diff_over_time.end(), out);
    return (int64_t)s->diff_over_time.size();
}
void orc_complete_domain(void* h, double* ux, double* uy, double* rho) {
    static_cast<Sim*>(h)->completeDomain(ux, uy, rho);
}
int orc_psi_negative(void* h) { return static_cast<Sim*>(h)->psiNegative; }

// Multi-block field on the pixel image: each pixel takes the value of the
// non-buffer cell covering it (the assignment of LB2D::reconstructOriginalGrid,
// coreLB.hpp:1413-1552, whose interpolation is commented out); solids 0.
void orc_pixel_domain(void* h, double* ux, double* uy, double* rho) {
    Sim* s = static_cast<Sim*>(h);
    const int N = s->NX;
    std::fill(ux, ux + (size_t)N * N, 0.0);
    std::fill(uy, uy + (size_t)N * N, 0.0);
    std::fill(rho, rho + (size_t)N * N, 0.0);
    for (int64_t c = 0; c < s->G.size; c++) {
        if (s->G.gridIsBuffer[c]) continue;
        const int w = 1 << (s->G.gridLevel[c] - 1);
        for (int dy = 0; dy < w; dy++)
            for (int dx = 0; dx < w; dx++) {
                const size_t p = (size_t)(s->G.gridIJ[2 * c + 1] + dy) * N + s->G.gridIJ[2 * c] + dx;
                ux[p] = s->ux[c];
                uy[p] = s->uy[c];
                rho[p] = s->rho[c];
            }
    }
}
int64_t orc_parent_child(void* h, int64_t* rows) {
    Sim* s = static_cast<Sim*>(h);
    if (rows) std::copy(s->G.parentChild.begin(), s->G.parentChild.end(), rows);
    return (int64_t)s->G.parentChild.size() / 5;
}
// Shape::calculateProperties (shape.hpp:127-251).  domain ny x nx row-major;
// locpore, kn: ny x nx; skeleton (may be null): the cropped skeleton.
int orc_shape_properties(const int* domain, int nx, int ny, double resolution, double mfp, double* locpore,
                         double* kn, uint8_t* skeleton, int* iterations) {
    try {
        shp::Properties P;
        shp::calculate(domain, nx, ny, resolution, mfp, P);
        std::copy(P.locpore.begin(), P.locpore.end(), locpore);
        std::copy(P.kn.begin(), P.kn.end(), kn);
        if (skeleton) std::copy(P.skeleton.begin(), P.skeleton.end(), skeleton);
        if (iterations) *iterations = P.iterations;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}
// utils.hpp:518-536 in place on a 0/1 image; returns the iterations run.
int orc_thinning(uint8_t* img, int rows, int cols) {
    std::vector<uint8_t> v(img, img + (size_t)rows * cols);
    const int it = shp::thin(v, rows, cols);
    std::copy(v.begin(), v.end(), img);
    return it;
}
// Distance of every pixel to the nearest zero pixel of fg: the float value
// and the exact integer square (-1: no zero pixel).
void orc_distance(const uint8_t* fg, int rows, int cols, float* value, int64_t* square) {
    const std::vector<int64_t> k2 = shp::squaredDistance(fg, rows, cols);
    for (size_t p = 0; p < k2.size(); p++) {
        if (value) value[p] = shp::distanceValue(k2[p]);
        if (square) square[p] = k2[p];
    }
}
int64_t orc_nearest_set(const uint8_t* img, int rows, int cols, int i, int j, double D) {
    return shp::nearestSetPixel(img, rows, cols, i, j, D);
}
// utils.hpp:539-568 at factor f (1, 2, 4, ...): the n/f x n/f image that
// cv::resize + threshold(128) gives for the n x n 0/1 image dom (row-major);
// the per-level relaxation time uses its interior rows.
void orc_downsample(const int* dom, int n, int f, int* out) {
    orc::Shape S;
    S.Nx = S.Ny = n;
    S.domain.assign(dom, dom + (size_t)n * n);
    const int m = n / f;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) out[(size_t)i * m + j] = orc::resizedPixel(S, f, i, j);
}
int orc_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
// OpenMP threads of the -Ofast build (the reference's OMP_NUM_THREADS)
void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}
}  // extern "C"
