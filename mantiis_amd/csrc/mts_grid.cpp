// Host-side set-up of a single-block lattice: PR-EOS state, cell list,
// connectivity, boundary classification and the per-site link words that the
// HIP kernels decode.  Plain C++ (no HIP calls).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <utility>

#include "mantiis_hip.h"
#include "mts_common.h"
#include "mts_tables.inc"

namespace mts {

// PR-EOS gas root and lattice-unit conversions of globalParameters.hpp:180-221.
// The reference solves the cubic with Cardano's formula taking
// pow(u, 1/3) of both terms (utils.hpp:571-600); both terms are positive for
// supercritical methane, where that equals the real cube root used here.
// A (T, P) for which the cubic has three real roots or a negative Cardano
// term is flagged (cubic_ok = false) instead of producing a NaN density.
void compute_physics(Physics& ph) {
    // literals written as in globalParameters.hpp:57-61 so that the products round identically
    const double Tc = 190.564, Pc = 4.5992 * 1.0e6, dmol = 3.800e-10, omega = 0.011, Mw = 16.043 * 1.0e-3;
    const double NA = 6.02214076e23, R = 8.31446, a_lu = 2.0 / 49.0, b_lu = 2.0 / 21.0;
    const double kk = 0.37464 + 1.54226 * omega - 0.26992 * std::pow(omega, 2);
    const double alpha = std::pow((1 + kk * (1 - std::sqrt(ph.T_phy / Tc))), 2);
    ph.a_alpha_lu = a_lu * alpha;
    const double Rs = R / Mw;
    const double a = 0.45724 * (R * Tc) * (R * Tc) / Pc;
    const double aG = a * alpha;
    const double bG = 0.07780 * R * Tc / Pc;
    const double AG = aG * ph.P_phy / (R * R) / (ph.T_phy * ph.T_phy);
    const double BG = bG * ph.P_phy / R / ph.T_phy;
    // Z^3 + c2 Z^2 + c1 Z + c0 = 0
    const double c2 = BG - 1, c1 = AG - 3 * (BG * BG) - 2 * BG, c0 = BG * BG * BG + (BG * BG) - AG * BG;
    const double p = (3 * c1 - c2 * c2) / 3;
    const double q = (2 * c2 * c2 * c2 - 9 * c2 * c1 + 27 * c0) / 27;
    const double disc = -4 * p * p * p - 27 * q * q;
    ph.cubic_ok = false;
    ph.Z = std::numeric_limits<double>::quiet_NaN();
    if (disc < 0) {
        const double s = std::sqrt(q * q / 4 + p * p * p / 27);
        const double u1 = -q / 2 + s, u2 = -q / 2 - s;
        if (u1 > 0 && u2 > 0) {
            ph.Z = std::pow(u1, 1.0 / 3.0) + std::pow(u2, 1.0 / 3.0) - c2 / 3;
            ph.cubic_ok = std::isfinite(ph.Z);
        }
    }
    const double rho_phy = (1 / (ph.Z * R * ph.T_phy / ph.P_phy)) * Mw;
    const double C_Rs = Rs;
    const double b_Rs = 0.07780 * Rs * Tc / Pc;
    const double a_Rs = 0.45724 * std::pow((Rs * Tc), 2) / Pc;
    const double C_as = a_Rs / a_lu, C_bs = b_Rs / b_lu, C_aob = C_as / C_bs;
    const double Cl = ph.Cl;
    const double Ct = Cl / std::sqrt(C_aob);
    const double CT = 1 / (C_Rs / std::pow(Cl, 2) * std::pow(Ct, 2));
    ph.T_lu = ph.T_phy / CT;
    const double Cm = std::pow(Cl, 3) / C_bs;
    ph.Crho = Cm / (Cl * Cl * Cl);
    const double Cn = Cm / Mw;
    ph.Cu = Cl / Ct;
    ph.Cf = Cm / (std::pow(Ct, 2) * std::pow(Cl, 2));
    const double Crho_mol = Cn / std::pow(Cl, 3);
    ph.rhoin = rho_phy / ph.Crho;
    ph.mfp = 1.0 / (std::sqrt(2.0) * Crho_mol * ph.rhoin * NA * M_PI * (2 * dmol / 2.0 + 2 * dmol / 2.0) *
                    (2 * dmol / 2.0 + 2 * dmol / 2.0));
}

static int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Relaxation rates and wall coefficients of one (tau_s, Kn) pair, with the
// reference's formulas (coreLB.hpp:603-604, 757-759, 925, 961).
CellParam make_param(double ts, double K, const Physics& ph) {
    CellParam cp{};
    const double A2 = 0.9;
    const double tq = 0.5 + (3.0 + M_PI * (2.0 * ts - 1.0) * (2.0 * ts - 1.0) * A2) / (8.0 * (2.0 * ts - 1.0));
    const double base = 1.0 / std::sqrt((double)(ph.mfp / ph.Cl * 1.0 / (1.0 + 2.0 * K))) * 2.0;
    const double td = 0.5 + (3.0 / 2.0) * std::sqrt((double)3.0) * 1.0 / (base * base);
    cp.taus = ts;
    cp.kn = K;
    cp.s_s = 1.0 / ts;
    cp.s_q = 1.0 / tq;
    cp.s_d = 1.0 / td;
    cp.r_srbb = 0.4987 * std::pow(K, -0.131) - 0.25;
    cp.r_mdbb = 0.3222 * std::pow(K, -0.217) - 0.25;
    return cp;
}

int build_single_grid(int nx, int ny, const int32_t* pore, const double* kn, const double* locpore,
                      const Physics& ph, bool torus, Geometry& g, std::string& err) {
    if (nx < 4 || ny < 4) {
        err = "lattice must be at least 4x4";
        return MTS_ERR_ARG;
    }
    g = Geometry();
    g.nx = nx;
    g.ny = ny;
    g.torus = torus;
    const int64_t P = (int64_t)nx * ny;
    g.stride = (P + 255) / 256 * 256;

    // Cell list: fluid pixels in row-major order (grid.hpp:401-424 and the
    // single-level branch of rebuildGrid, grid.hpp:537-553, 861-873).
    std::vector<int64_t> cellOf(P, -1);
    for (int64_t p = 0; p < P; p++)
        if (pore[p] == 1) {
            cellOf[p] = (int64_t)g.cellPos.size();
            g.cellPos.push_back((int32_t)p);
        } else if (pore[p] != 0) {
            err = "pore image must hold only 0 (solid) and 1 (fluid)";
            return MTS_ERR_ARG;
        }
    const int64_t n = (int64_t)g.cellPos.size();
    g.ncell = n;
    if (n == 0) {
        err = "geometry has no fluid cell";
        return MTS_ERR_ARG;
    }

    // Same-level neighbours with x/y wrap (grid.hpp:916-972): a solid pixel
    // gives -1; at one level every pixel exists, so -2 never occurs.
    g.gridConnect.assign(n * 9, -1);
    for (int64_t c = 0; c < n; c++) {
        const int x = g.cellPos[c] % nx, y = g.cellPos[c] / nx;
        for (int k = 0; k < 9; k++) {
            const int64_t q = (int64_t)wrap(y + EY[k], ny) * nx + wrap(x + EX[k], nx);
            g.gridConnect[c * 9 + k] = cellOf[q];
        }
    }

    // Boundary type (grid.hpp:1111-1157): 9-bit code over neighbours
    // {7,3,6,4,0,2,8,1,5}, MSB first, 0 where the neighbour is a wall.
    static const int order[9] = {7, 3, 6, 4, 0, 2, 8, 1, 5};
    int codeToType[512];
    for (int i = 0; i < 512; i++) codeToType[i] = -2;
    for (int t = 0; t < MTS_NTYPES; t++) codeToType[MTS_BND_CODE[t]] = t;
    codeToType[511] = -1;
    g.bndTypes.assign(n, -1);
    for (int64_t c = 0; c < n; c++) {
        int code = 0;
        for (int b = 0; b < 9; b++) code = (code << 1) | (g.gridConnect[c * 9 + order[b]] == -1 ? 0 : 1);
        const int t = codeToType[code];
        if (t == -2) {
            const int x = g.cellPos[c] % nx, y = g.cellPos[c] / nx;
            err = "[ERROR]: Boundary type is not contained. (cell x=" + std::to_string(x) + " y=" +
                  std::to_string(y) + ", code " + std::to_string(code) + ")";
            return MTS_ERR_GEOMETRY;
        }
        g.bndTypes[c] = t;
    }

    // Relaxation times (coreLB.hpp:482-502) and per-cell Kn (grid.hpp:891-899),
    // folded into a table of distinct (tau_s, Kn) pairs.
    g.taus.assign(n, 0);
    g.kn.assign(n, 0);
    std::map<std::pair<double, double>, int> uniq;
    g.pidx.assign(P, 0);
    for (int64_t c = 0; c < n; c++) {
        const int64_t p = g.cellPos[c];
        const double K = kn[p], L = locpore[p];
        g.taus[c] = 0.5 + std::sqrt(6.0 / M_PI) * L * K / (1 + 2 * K);
        g.kn[c] = K;
        auto key = std::make_pair(g.taus[c], K);
        auto it = uniq.find(key);
        int id;
        if (it == uniq.end()) {
            id = (int)g.params.size();
            if (id > 65535) {
                err = "more than 65536 distinct (tau_s, Kn) pairs";
                return MTS_ERR_UNSUPPORTED;
            }
            uniq.emplace(key, id);
            const CellParam cp = make_param(g.taus[c], K, ph);
            g.params.push_back(cp);
        } else
            id = it->second;
        g.pidx[p] = (uint16_t)id;
    }

    // Link words from the classification of coreLB.hpp:292-356 at one rank.
    g.meta.assign(P, 0);
    for (int64_t c = 0; c < n; c++) {
        const int64_t p = g.cellPos[c];
        const int x = (int)(p % nx);
        const int t = g.bndTypes[c];
        uint32_t w = FLUID_BIT | ((uint32_t)(t < 0 ? TYPE_NONE : t) << TYPE_SHIFT);
        if (!torus && (x == 0 || x == nx - 1)) w |= SEAM_BIT;
        g.meta[p] |= w;
    }
    for (int64_t c = 0; c < n; c++) {
        const int64_t p = g.cellPos[c];
        const int x = (int)(p % nx), y = (int)(p / nx);
        const int t = g.bndTypes[c];
        for (int ic = 1; ic < 9; ic++) {
            const int64_t nb = g.gridConnect[c * 9 + ic];
            const bool ifs = (t == -1) || MTS_IFSTREAM[t][ic];
            if (nb >= 0) {
                const int64_t q = g.cellPos[nb];
                const int nxq = (int)(q % nx), nyq = (int)(q / nx);
                const bool xp = std::abs(x - nxq) > 1, yp = std::abs(y - nyq) > 1;
                if (torus || (!xp && !yp)) {
                    if (ifs) {  // stream link c -> nb
                        g.meta[q] |= 1u << (IN_SHIFT + ic - 1);
                        g.meta[p] |= 1u << (OUT_SHIFT + ic - 1);
                    } else
                        g.meta[p] |= 1u << (WALL_SHIFT + ic - 1);
                } else if (xp) {  // periodic link c -> nb (ifStream ignored)
                    g.meta[q] |= 1u << (IN_SHIFT + ic - 1);
                    g.meta[p] |= 1u << (OUT_SHIFT + ic - 1);
                }
                // otherwise a y-seam-only link is dropped (coreLB.hpp:301-331)
            } else if (!MTS_IFSTREAM[t < 0 ? 0 : t][ic] && t >= 0) {
                g.meta[p] |= 1u << (WALL_SHIFT + ic - 1);
            }
        }
        // ic = 0 is always a stream link of the cell to itself
        // (gridConnect[c][0] == c, ifStream[t][0] == 1 for all 52 types).
    }

    // Generalized-periodic inlet/outlet (coreLB.hpp:282-291, 357-365).
    const int ny_c = (ny % 2 == 1) ? (ny + 1) / 2 - 1 : ny / 2 - 1;
    const int64_t pin = (int64_t)ny_c * nx + 0, pout = (int64_t)ny_c * nx + nx - 1;
    if (cellOf[pin] >= 0 && cellOf[pout] >= 0) {
        g.gpIn = pin;
        g.gpOut = pout;
    }
    return MTS_OK;
}

void decode_links(const Geometry& g, int which, std::vector<int64_t>& id, std::vector<int64_t>& loc,
                  std::vector<int32_t>& icv) {
    id.clear();
    loc.clear();
    icv.clear();
    const int nx = g.nx, ny = g.ny;
    std::vector<int64_t> cellOf((size_t)nx * ny, -1);
    for (int64_t c = 0; c < g.ncell; c++) cellOf[g.cellPos[c]] = c;
    for (int64_t c = 0; c < g.ncell; c++) {
        const int64_t p = g.cellPos[c];
        const int x = (int)(p % nx), y = (int)(p / nx);
        const uint32_t w = g.meta[p];
        for (int ic = 0; ic < 9; ic++) {
            if (ic == 0) {
                if (which == 0) { id.push_back(c); loc.push_back(c); icv.push_back(0); }
                continue;
            }
            const int xq = wrap(x + EX[ic], nx), yq = wrap(y + EY[ic], ny);
            const int64_t q = (int64_t)yq * nx + xq;
            const bool xseam = !g.torus && ((x + EX[ic] < 0) || (x + EX[ic] >= nx));
            const bool linked = cellOf[q] >= 0 && ((g.meta[q] >> (IN_SHIFT + ic - 1)) & 1u);
            if (which == 0 && linked && !xseam) { id.push_back(c); loc.push_back(cellOf[q]); icv.push_back(ic); }
            if (which == 2 && linked && xseam) { id.push_back(c); loc.push_back(cellOf[q]); icv.push_back(ic); }
            if (which == 1 && ((w >> (WALL_SHIFT + ic - 1)) & 1u)) { id.push_back(c); loc.push_back(-1); icv.push_back(ic); }
        }
    }
}

void canonical_spec(uint32_t spec[64]) {
    // slot j <- the link ic that is not streamed for the type and whose
    // specular direction is j (the first such link when two share a slot)
    for (int t = 0; t < 64; t++) spec[t] = 0;
    for (int t = 0; t <= MTS_NTYPES; t++)
        for (int ic = 8; ic >= 1; ic--) {
            const bool wl = t == MTS_NTYPES ? true : !MTS_IFSTREAM[t][ic];
            const int sr = t == MTS_NTYPES ? OPP[ic] : MTS_ICSR[t][ic];
            if (!wl || sr <= 0) continue;
            spec[t] = (spec[t] & ~(15u << (4 * (sr - 1)))) | ((uint32_t)ic << (4 * (sr - 1)));
        }
}

void build_compact(const Geometry& g, CompactImage& ci) {
    LatticeView v{g.nx, g.ny, g.torus, g.meta.data(), g.pidx.data(), g.gpIn, g.gpOut, nullptr};
    build_compact(v, ci);
}

void build_compact(const LatticeView& g, CompactImage& ci) {
    const int nx = g.nx, ny = g.ny;
    const int64_t P = (int64_t)nx * ny;
    // cells = the sites carrying FLUID_BIT, in row-major order
    ci.cidx.assign(P, -1);
    ci.cellPos.clear();
    for (int64_t p = 0; p < P; p++)
        if (g.meta[p] & FLUID_BIT) {
            ci.cidx[p] = (int32_t)ci.cellPos.size();
            ci.cellPos.push_back((int32_t)p);
        }
    const int64_t n = (int64_t)ci.cellPos.size();
    ci.ncell = n;
    ci.cstride = (n + 255) / 256 * 256;
    // rank[p] = cells before site p (row-major)
    std::vector<int32_t> rank(P + 1);
    int32_t acc = 0;
    for (int64_t p = 0; p < P; p++) {
        rank[p] = acc;
        acc += ci.cidx[p] >= 0;
    }
    rank[P] = acc;
    ci.ntx = (nx + CTILE_W - 1) / CTILE_W;
    ci.segs.assign((size_t)ny * (ci.ntx + 1), 0);
    for (int y = 0; y < ny; y++)
        for (int tx = 0; tx <= ci.ntx; tx++) {
            const int x = std::min(tx * CTILE_W, nx);
            ci.segs[(size_t)y * (ci.ntx + 1) + tx] = rank[(int64_t)y * nx + x];
        }
    ci.cw.assign(n, 0);
    ci.rw.assign(n, 0u);
    ci.pidx.assign(n, 0);
    ci.outm.assign(n, 0);
    ci.aux.clear();
    canonical_spec(ci.spec);
    for (int64_t c = 0; c < n; c++) {
        const int64_t p = ci.cellPos[c];
        const int x = (int)(p % nx), y = (int)(p / nx);
        const uint32_t w = g.meta[p];
        const uint32_t in = (w >> IN_SHIFT) & 0xffu, out = (w >> OUT_SHIFT) & 0xffu, wall = (w >> WALL_SHIFT) & 0xffu;
        const int32_t ovr = g.override ? g.override[p] : -1;
        const uint32_t t = (w >> TYPE_SHIFT) & 63u;
        const int yu = wrap(y + 1, ny), yd = wrap(y - 1, ny);
        const int64_t pu = (int64_t)yu * nx + x, pd = (int64_t)yd * nx + x;
        uint32_t cw = in | (wall << CW_WALL_SHIFT) | (t << CW_TYPE_SHIFT) | ((uint32_t)(x & 127) << CW_X_SHIFT);
        if (ci.cidx[pd] >= 0) cw |= CW_ISC_D;
        if (ci.cidx[pu] >= 0) cw |= CW_ISC_U;
        // specular source of each slot (SRBB, coreLB.hpp:918-933)
        uint32_t spec = 0;
        bool ambiguous = false;
        for (int ic = 1; ic < 9; ic++) {
            if (!((wall >> (ic - 1)) & 1u)) continue;
            const int sr = t < (uint32_t)MTS_NTYPES ? MTS_ICSR[t][ic] : OPP[ic];
            if (sr <= 0) continue;
            if ((spec >> (4 * (sr - 1))) & 15u) ambiguous = true;
            spec |= (uint32_t)ic << (4 * (sr - 1));
        }
        // rows 0 and ny-1 of a non-torus lattice lose their y-seam links
        // the kernel takes slot j's specular source from the type table,
        // kept when that source is one of the cell's wall links
        uint32_t specT = 0;
        for (int j = 1; j < 9; j++) {
            const uint32_t sj = (ci.spec[t] >> (4 * (j - 1))) & 15u;
            if (sj && ((wall >> (sj - 1)) & 1u)) specT |= sj << (4 * (j - 1));
        }
        const bool complex = (w & SEAM_BIT) || x == 0 || x == nx - 1 || (!g.torus && (y == 0 || y == ny - 1)) ||
                             ambiguous || specT != spec || p == g.gpIn || p == g.gpOut || ny < 3 || ovr >= 0;
        auto push_aux = [&]() {  // explicit record of a COMPLEX cell
            const int32_t k = (int32_t)(ci.aux.size() / AUXW);
            for (int j = 1; j < 9; j++) {
                const int xs = wrap(x - EX[j], nx), ys = wrap(y - EY[j], ny);
                const int32_t q = ci.cidx[(int64_t)ys * nx + xs];
                ci.aux.push_back(((in >> (j - 1)) & 1u) ? q : -1);
            }
            ci.aux.push_back(x);
            ci.aux.push_back(y);
            ci.aux.push_back((int32_t)w);
            ci.aux.push_back(ovr);
            return k;
        };
        ci.pidx[c] = g.pidx[p];
        ci.outm[c] = (uint8_t)out;
        // rank offsets of the rows above and below, packed as two signed
        // 16-bit values (the wrapped rows of a torus do not fit)
        const int64_t du = (int64_t)rank[pu] - c, dd = c - (int64_t)rank[pd];
        const bool fits = du >= -32768 && du <= 32767 && dd >= -32768 && dd <= 32767;
        if (complex || !fits) {
            cw |= CW_COMPLEX;
            ci.rw[c] = (uint32_t)push_aux();
        } else {
            ci.rw[c] = ((uint32_t)(uint16_t)(int16_t)du) | ((uint32_t)(uint16_t)(int16_t)dd << 16);
            // every streamed slot must find its source by rank arithmetic
            const int64_t ru = rank[pu], rd = rank[pd];
            const int64_t src[9] = {c, c - 1, rd, c + 1, ru, rd - 1, rd + ((cw & CW_ISC_D) ? 1 : 0),
                                    ru + ((cw & CW_ISC_U) ? 1 : 0), ru - 1};
            bool ok = true;
            for (int j = 1; j < 9; j++) {
                if (!((in >> (j - 1)) & 1u)) continue;
                const int xs = wrap(x - EX[j], nx), ys = wrap(y - EY[j], ny);
                if (ci.cidx[(int64_t)ys * nx + xs] != src[j]) ok = false;
            }
            if (!ok) {
                cw |= CW_COMPLEX;
                ci.rw[c] = (uint32_t)push_aux();
            }
        }
        ci.cw[c] = cw;
    }
    ci.naux = (int64_t)ci.aux.size() / AUXW;
}

}  // namespace mts
