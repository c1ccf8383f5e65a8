// Host-side set-up of a quadtree lattice (the multi-block grid of
// grid.hpp:313-1157): wall distance, level decision, buffer cells,
// connectivity, parent/child and interface tables, and the per-level site
// images with link words that the HIP kernels decode.  Plain C++.
//
// With at most K levels, a cell of level l >= 2 is an aligned block of side
// h = 2^(l-1) whose pixels are fluid at a Euclidean distance of at least
// ceil(2^(K-1) sqrt 2) h from every wall pixel, and no larger block holding
// it is one (the quadtree split rule of grid.hpp:1180-1214 with
// maxAllowedLevel = K; the reference fixes K = 2: blocks of 2 x 2 at
// distance >= 6).  Cells are numbered in row-major order of their top-left
// pixel, followed by the buffer cells in creation order (grid.hpp:808-888).
// On a torus (periodic_y = 1) the distance transform and the buffer rule wrap
// around both seams; otherwise they stop at the image edge as the
// reference's do.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <unordered_map>

#include "mantiis_hip.h"
#include "mts_common.h"
#include "mts_multigrid.h"
#include "mts_tables.inc"

namespace mts {

namespace {

// 1-D squared distance transform (lower envelope of parabolas).
void dt1d(const std::vector<double>& f, std::vector<double>& d) {
    const int n = (int)f.size();
    std::vector<int> v(n);
    std::vector<double> z(n + 1);
    int k = 0;
    v[0] = 0;
    z[0] = -1e300;
    z[1] = 1e300;
    for (int q = 1; q < n; q++) {
        if (f[q] >= 1e299) continue;
        if (f[v[k]] >= 1e299) {
            v[k] = q;
            continue;
        }
        double s;
        while (true) {
            s = ((f[q] + (double)q * q) - (f[v[k]] + (double)v[k] * v[k])) / (2.0 * q - 2.0 * v[k]);
            if (s <= z[k] && k > 0) k--;
            else break;
        }
        if (s <= z[k]) {  // k == 0
            v[k] = q;
            z[k + 1] = 1e300;
            continue;
        }
        k++;
        v[k] = q;
        z[k] = s;
        z[k + 1] = 1e300;
    }
    k = 0;
    for (int q = 0; q < n; q++) {
        while (z[k + 1] < q) k++;
        const double dq = (double)(q - v[k]);
        d[q] = f[v[k]] >= 1e299 ? 1e300 : dq * dq + f[v[k]];
    }
}

// Exact squared Euclidean distance of every pixel to the nearest wall pixel
// (Felzenszwalb-Huttenlocher separable transform), optionally periodic.
std::vector<double> edt2(int nx, int ny, const int32_t* pore, bool wrap) {
    const double INF = 1e300;
    std::vector<double> g((size_t)nx * ny);
    const int rep = wrap ? 3 : 1;
    {
        std::vector<double> f((size_t)ny * rep), d(f.size());
        for (int x = 0; x < nx; x++) {
            for (int r = 0; r < rep; r++)
                for (int y = 0; y < ny; y++) f[(size_t)r * ny + y] = pore[(size_t)y * nx + x] == 0 ? 0.0 : INF;
            dt1d(f, d);
            for (int y = 0; y < ny; y++) g[(size_t)y * nx + x] = d[(size_t)(wrap ? ny : 0) + y];
        }
    }
    std::vector<double> out((size_t)nx * ny);
    std::vector<double> f((size_t)nx * rep), d(f.size());
    for (int y = 0; y < ny; y++) {
        for (int r = 0; r < rep; r++)
            for (int x = 0; x < nx; x++) f[(size_t)r * nx + x] = g[(size_t)y * nx + x];
        dt1d(f, d);
        for (int x = 0; x < nx; x++) out[(size_t)y * nx + x] = d[(size_t)(wrap ? nx : 0) + x];
    }
    return out;
}

int wrapc(int v, int n) { return ((v % n) + n) % n; }

}  // namespace

int multi_check(int nx, int ny, const int32_t* pore, int K, std::string& err) {
    if (K < 2 || K > MTS_MAX_LEVELS) {
        err = "a quadtree has 2 to " + std::to_string(MTS_MAX_LEVELS) + " levels";
        return MTS_ERR_UNSUPPORTED;
    }
    if (nx != ny || nx < 8 || (nx & (nx - 1)) != 0) {
        err = "Matrix dimensions must be divisible by 2 in each direction. (multi-block lattices are square, "
              "power-of-two sided)";
        return MTS_ERR_ARG;
    }
    for (int64_t p = 0; p < (int64_t)nx * ny; p++)
        if (pore[p] != 0 && pore[p] != 1) {
            err = "pore image must hold only 0 (solid) and 1 (fluid)";
            return MTS_ERR_ARG;
        }
    return MTS_OK;
}

// Split distance of the quadtree with at most K levels (grid.hpp:1180-1214):
// a block of side h <= 2^(K-1) is a leaf when it is fluid and no wall pixel
// lies closer than ceil(2^(K-1) sqrt 2) h to any of its pixels.
int split_distance(int K) { return (int)std::ceil(std::pow(2.0, K - 1) * std::sqrt(2.0)); }

// Level decision (the leaves of grid.hpp:1160-1269): each pixel's level is
// 1 + log2 of the side of the largest aligned block holding it that is a
// leaf; larger blocks are decided first, so a block inside a leaf is not
// looked at again.
void multi_levels_host(int nx, int ny, const int32_t* pore, bool torus, int K, std::vector<uint8_t>& lvl) {
    const std::vector<double> d2 = edt2(nx, ny, pore, torus);
    auto dist = [&](int x, int y) -> double {
        const double v = d2[(size_t)y * nx + x];
        return v >= 1e299 ? 1e30 : (double)(float)std::sqrt(v);
    };
    const int md = split_distance(K);
    lvl.assign((size_t)nx * ny, 1);
    for (int l = K; l >= 2; l--) {
        const int h = 1 << (l - 1);
        const double T = (double)md * h;
        for (int by = 0; by + h <= ny; by += h)
            for (int bx = 0; bx + h <= nx; bx += h) {
                if (lvl[(size_t)by * nx + bx] != 1) continue;
                bool ok = true;
                for (int y = by; y < by + h && ok; y++)
                    for (int x = bx; x < bx + h && ok; x++) ok = pore[(size_t)y * nx + x] == 1 && dist(x, y) >= T;
                if (!ok) continue;
                for (int y = by; y < by + h; y++)
                    for (int x = bx; x < bx + h; x++) lvl[(size_t)y * nx + x] = (uint8_t)l;
            }
    }
}

int build_multi_grid(int nx, int ny, const int32_t* pore, const double* kn, const double* locpore, const Physics& ph,
                     bool torus, int K, MultiGeometry& G, std::string& err) {
    (void)ph;
    (void)locpore;
    G = MultiGeometry();
    if (int rc = multi_check(nx, ny, pore, K, err)) return rc;
    std::vector<uint8_t> lvl;
    multi_levels_host(nx, ny, pore, torus, K, lvl);
    if (int rc = multi_cells(nx, ny, pore, kn, torus, lvl, G, err)) return rc;
    return multi_tables_host(G, err);
}

// The cells of the quadtree: original cells (fluid pixels of level 1 and
// the blocks of higher levels) in row-major order of their top-left pixel,
// then the buffer cells in creation order (grid.hpp:427-889; sequential: a
// buffer's placement depends on the buffers made before it).
int multi_cells(int nx, int ny, const int32_t* pore, const double* kn, bool torus, const std::vector<uint8_t>& lvl,
                MultiGeometry& G, std::string& err) {
    (void)err;
    G = MultiGeometry();
    G.nx = nx;
    G.ny = ny;
    G.torus = torus;
    int L = 1;
    for (uint8_t v : lvl) L = std::max<int>(L, v);
    G.maxLevel = L;
    G.lvl = lvl;
    auto gid = [&](int x, int y, int l) -> int64_t { return (int64_t)x + (int64_t)y * ny + (int64_t)(l - 1) * ny * nx; };
    auto lvlAt = [&](int x, int y) -> int { return lvl[(size_t)y * nx + x]; };

    // ---- original cells (incl. wall pixels) in row-major order of top-left pixel
    struct Orig {
        int x, y, lvl;
        bool fluid;
    };
    std::vector<Orig> orig;
    std::vector<int32_t> origAt((size_t)nx * ny, -1);  // top-left pixel -> orig index
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const int v = lvlAt(x, y), h = 1 << (v - 1);
            if (v > 1 && (x % h || y % h)) continue;
            origAt[(size_t)y * nx + x] = (int32_t)orig.size();
            orig.push_back({x, y, v, v > 1 || pore[(size_t)y * nx + x] == 1});
        }
    // existence of a cell (original or buffer) per level position
    std::vector<std::vector<uint8_t>> exist(L);
    for (int l = 0; l < L; l++) exist[l].assign((size_t)(nx >> l) * (ny >> l), 0);
    auto ex = [&](int l, int x, int y) -> uint8_t& { return exist[l - 1][(size_t)(y >> (l - 1)) * (nx >> (l - 1)) + (x >> (l - 1))]; };
    for (const Orig& o : orig) ex(o.lvl, o.x, o.y) = 1;
    // ---- buffer cells (grid.hpp:598-806): a cell of level g >= 2 with a
    // position of its level missing in its window (side 2 R, R the side of
    // the largest cells) makes its children at level g - 1 buffers when a
    // cell of its own level lies in range (it does itself), and for each
    // other cell in range, when no larger cell does, a buffer of its level
    // at the missing position nearest to that cell
    struct Buf {
        int x, y, lvl, type;
    };
    std::vector<Buf> bufs;
    const int R = 1 << (L - 1);
    for (const Orig& o : orig) {
        const int g = o.lvl;
        if (g < 2) continue;
        const int jump = 1 << (g - 1);
        // missing positions of the window, column-major (x outer, y inner),
        // unwrapped (off-image positions of a torus wrap)
        std::vector<std::pair<int, int>> mk;
        for (int ox = -R; ox <= R; ox += jump)
            for (int oy = -R; oy <= R; oy += jump) {
                const int x = o.x + ox, y = o.y + oy;
                if (!torus && (x < 0 || y < 0 || x >= nx || y >= ny)) continue;
                if (lvlAt(wrapc(x, nx), wrapc(y, ny)) != g) mk.push_back({x, y});
            }
        if (mk.empty()) continue;
        // cells whose top-left pixel lies within +-R of this cell's, in cell order
        std::vector<std::pair<int, std::pair<int, int>>> inr;  // (orig index, unwrapped top-left)
        for (int oy = -R; oy <= R; oy++)
            for (int ox = -R; ox <= R; ox++) {
                const int x = o.x + ox, y = o.y + oy;
                if (!torus && (x < 0 || y < 0 || x >= nx || y >= ny)) continue;
                const int k = origAt[(size_t)wrapc(y, ny) * nx + wrapc(x, nx)];
                if (k >= 0) inr.push_back({k, {x, y}});
            }
        std::sort(inr.begin(), inr.end());
        if (torus)  // a cell may be reached twice through the wrap on tiny lattices; keep the first
            inr.erase(std::unique(inr.begin(), inr.end(), [](const auto& a, const auto& b) { return a.first == b.first; }),
                      inr.end());
        int maxL = 0;
        for (const auto& e : inr) maxL = std::max(maxL, orig[e.first].lvl);
        std::vector<bool> used(mk.size(), false);
        for (const auto& e : inr) {
            const Orig& c = orig[e.first];
            if (c.lvl == g) {
                // the children of this cell (type 10 (g - 1) + g)
                const int hj = jump / 2;
                const int cand[4][2] = {{o.x, o.y}, {o.x + hj, o.y}, {o.x, o.y + hj}, {o.x + hj, o.y + hj}};
                for (auto& cd : cand)
                    if (!ex(g - 1, cd[0], cd[1])) {
                        ex(g - 1, cd[0], cd[1]) = 1;
                        bufs.push_back({cd[0], cd[1], g - 1, 10 * (g - 1) + g});
                    }
            } else if (maxL <= g) {
                // a buffer of this level at the missing position closest to
                // the cell in range (type 10 g + g - 1)
                const int rx = e.second.first, ry = e.second.second;
                const float shift = (float)(std::pow(2.0, g - 2) - 0.5);
                int best = 0;
                double minD = std::numeric_limits<double>::max();
                for (size_t m = 0; m < mk.size(); m++) {
                    float mx, my;
                    if (used[m]) {
                        // the reference marks a consumed key -1, whose position
                        // ind2sub(-1) reads as (-1, 0); on a torus it is skipped
                        if (torus) continue;
                        mx = -1.0f;
                        my = 0.0f;
                    } else {
                        mx = (float)mk[m].first;
                        my = (float)mk[m].second;
                    }
                    const float ddx = ((float)rx + shift) - (mx + shift), ddy = ((float)ry + shift) - (my + shift);
                    const double dd = std::sqrt((double)(std::pow(ddx, 2) + std::pow(ddy, 2)));
                    if (dd < minD) {
                        minD = dd;
                        best = (int)m;
                    }
                }
                if (!used[best]) {
                    const int mx = wrapc(mk[best].first, nx), my = wrapc(mk[best].second, ny);
                    if (!ex(g, mx, my)) {
                        ex(g, mx, my) = 1;
                        bufs.push_back({mx, my, g, 10 * g + g - 1});
                        used[best] = true;
                    }
                }
            }
        }
    }
    // ---- grid arrays (fluid original cells, then buffers)
    for (const Orig& o : orig)
        if (o.fluid) {
            G.cx.push_back(o.x);
            G.cy.push_back(o.y);
            G.level.push_back(o.lvl);
            G.type.push_back(0);
            G.isBuffer.push_back(0);
        }
    for (const Buf& b : bufs) {
        G.cx.push_back(b.x);
        G.cy.push_back(b.y);
        G.level.push_back(b.lvl);
        G.type.push_back(b.type);
        G.isBuffer.push_back(1);
    }
    const int64_t n = (int64_t)G.cx.size();
    G.ncell = n;
    G.gridID.resize(n);
    for (int64_t c = 0; c < n; c++) G.gridID[c] = gid(G.cx[c], G.cy[c], G.level[c]);
    // per-cell Kn (the loaded map at the cell's first pixel); the
    // relaxation times need the device (multi_grid_taus)
    G.kn.resize(n);
    for (int64_t c = 0; c < n; c++) G.kn[c] = kn[(size_t)G.cy[c] * nx + G.cx[c]];
    G.taus.clear();
    G.pore.assign(pore, pore + (size_t)nx * ny);
    return MTS_OK;
}

// Neighbour and interface tables of the cells (grid.hpp:916-1157):
// connectivity, boundary types, parent/child rows and the coarse and fine
// buffer rows of each pair of levels.  The device builder (mts_quadtree.hip)
// forms the same tables.
int multi_tables_host(MultiGeometry& G, std::string& err) {
    const int nx = G.nx, ny = G.ny, L = G.maxLevel;
    const int64_t n = G.ncell;
    // per-level index images: the lowest cell at a level position
    std::vector<std::vector<int64_t>> at(L);
    for (int l = 0; l < L; l++) at[l].assign((size_t)(nx >> l) * (ny >> l), -1);
    auto atp = [&](int l, int x, int y) -> int64_t& { return at[l - 1][(size_t)(y >> (l - 1)) * (nx >> (l - 1)) + (x >> (l - 1))]; };
    for (int64_t c = n - 1; c >= 0; c--) atp(G.level[c], G.cx[c], G.cy[c]) = c;
    // ---- connectivity (grid.hpp:916-972): -1 where an original cell (a
    // wall pixel) of the level sits, -2 where none does
    G.connect.assign(n * 9, -2);
    for (int64_t c = 0; c < n; c++) {
        const int l = G.level[c], jump = 1 << (l - 1);
        for (int k = 0; k < 9; k++) {
            const int x = wrapc(G.cx[c] + EX[k] * jump, nx), y = wrapc(G.cy[c] + EY[k] * jump, ny);
            const int64_t q = atp(l, x, y);
            G.connect[c * 9 + k] = q >= 0 ? q : (G.lvl[(size_t)y * nx + x] == l ? -1 : -2);
        }
    }
    // ---- boundary types of non-buffer cells (grid.hpp:1111-1157)
    static const int order[9] = {7, 3, 6, 4, 0, 2, 8, 1, 5};
    int codeToType[512];
    for (int i = 0; i < 512; i++) codeToType[i] = -2;
    for (int t = 0; t < MTS_NTYPES; t++) codeToType[MTS_BND_CODE[t]] = t;
    codeToType[511] = -1;
    G.bnd.assign(n, -1);
    for (int64_t c = 0; c < n; c++) {
        if (G.isBuffer[c]) continue;
        int code = 0;
        for (int b = 0; b < 9; b++) code = (code << 1) | (G.connect[c * 9 + order[b]] == -1 ? 0 : 1);
        const int t = codeToType[code];
        if (t == -2) {
            err = "[ERROR]: Boundary type is not contained. (cell x=" + std::to_string(G.cx[c]) + " y=" +
                  std::to_string(G.cy[c]) + ", code " + std::to_string(code) + ")";
            return MTS_ERR_GEOMETRY;
        }
        G.bnd[c] = t;
    }
    // ---- parent/child (grid.hpp:974-1038): the four cells of the level
    // below in the cell's quadrants
    std::vector<int64_t> rowOf(n, -1);
    G.parentChild.clear();
    for (int64_t c = 0; c < n; c++) {
        if (G.level[c] <= 1) continue;
        const int hj = 1 << (G.level[c] - 2);
        rowOf[c] = (int64_t)G.parentChild.size() / 5;
        G.parentChild.push_back(c);
        const int sub[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
        for (auto& s : sub) G.parentChild.push_back(atp(G.level[c] - 1, G.cx[c] + s[0] * hj, G.cy[c] + s[1] * hj));
    }
    // ---- buffer rows (grid.hpp:1041-1108) per pair of levels (l, l + 1):
    // coarse buffers (type 10 (l + 1) + l) with their children, fine
    // buffers (type 10 l + l + 1) four at a time under the parent of the
    // first row naming any of them
    G.coarseRows.clear();
    G.fineRows.clear();
    G.coarseCount.assign(std::max(L - 1, 0), 0);
    G.fineCount.assign(std::max(L - 1, 0), 0);
    std::unordered_map<int64_t, int64_t> firstRow;
    const int64_t rows = (int64_t)G.parentChild.size() / 5;
    for (int64_t r = rows - 1; r >= 0; r--)
        for (int k = 4; k >= 1; k--)
            if (G.parentChild[5 * r + k] >= 0) firstRow[G.parentChild[5 * r + k]] = r;
    for (int l = 1; l < L; l++) {
        const int ctype = 10 * (l + 1) + l, ftype = 10 * l + l + 1;
        for (int64_t c = 0; c < n; c++)
            if (G.type[c] == ctype) {
                const int64_t r = rowOf[c];
                G.coarseRows.insert(G.coarseRows.end(), G.parentChild.begin() + 5 * r, G.parentChild.begin() + 5 * r + 5);
                G.coarseCount[l - 1]++;
            }
        std::vector<int64_t> ids;
        for (int64_t c = 0; c < n; c++)
            if (G.type[c] == ftype) ids.push_back(c);
        for (size_t g4 = 0; g4 + 3 < ids.size(); g4 += 4) {
            int64_t best = std::numeric_limits<int64_t>::max();
            for (int k = 0; k < 4; k++) {
                auto it = firstRow.find(ids[g4 + k]);
                if (it != firstRow.end()) best = std::min(best, it->second);
            }
            if (best == std::numeric_limits<int64_t>::max()) {
                err = "fine buffer cells without a parent";
                return MTS_ERR_GEOMETRY;
            }
            const int64_t row[5] = {G.parentChild[5 * best], ids[g4], ids[g4 + 1], ids[g4 + 2], ids[g4 + 3]};
            G.fineRows.insert(G.fineRows.end(), row, row + 5);
            G.fineCount[l - 1]++;
        }
    }
    return MTS_OK;
}

// utils.hpp:539-568 at the sizes the quadtree uses: cv::resize (bilinear)
// of the 0/255 image to 1/f of its side, then threshold at 128.  Output
// pixel (i, j) samples source position f (i, j) + (f - 1) / 2, half way
// between the pixels f/2 - 1 and f/2 of its block in each direction, so its
// value is the mean of those 2 x 2 pixels rounded to 8 bits (0, 64, 128,
// 191, 255): fluid when at least three of the four are fluid.  f = 1 is
// the image itself.
static int32_t resized_pixel(const int32_t* pore, int nx, int f, int i, int j) {
    if (f == 1) return pore[(size_t)i * nx + j] != 0;
    const int r0 = f * i + f / 2 - 1, c0 = f * j + f / 2 - 1;
    const int n = (pore[(size_t)r0 * nx + c0] != 0) + (pore[(size_t)r0 * nx + c0 + 1] != 0) +
                  (pore[(size_t)(r0 + 1) * nx + c0] != 0) + (pore[(size_t)(r0 + 1) * nx + c0 + 1] != 0);
    return n >= 3;
}

// LB2D::calculateMultigridTauS (coreLB.hpp:432-480): for each level
// lvl = 0..maxLevel-1 the domain resized by 2^-lvl, with wall rows first and
// last (Shape::addHorizontalBoundary), goes through calculateProperties at
// resolution Cl 2^lvl (on the device, mts_shape_properties); the level's
// map is tau = 0.5 + sqrt(6/pi) LocPore Kn / (1 + 2 Kn) (0.5 at walls), and
// a cell of level L takes its map's value at (y, x) / 2^(L-1).
int multi_grid_taus(const Physics& ph, MultiGeometry& G, std::string& err) {
    const int nx = G.nx, ny = G.ny;
    if (G.pore.size() != (size_t)nx * ny) {
        err = "multi-block grid without its pore image";
        return MTS_ERR_STATE;
    }
    std::vector<std::vector<double>> tauMap(G.maxLevel);
    for (int lvl = 0; lvl < G.maxLevel; lvl++) {
        const int f = 1 << lvl, lx = nx / f, ly = ny / f;
        std::vector<int32_t> dom((size_t)lx * ly, 0);
        for (int i = 1; i < ly - 1; i++)
            for (int j = 0; j < lx; j++) dom[(size_t)i * lx + j] = resized_pixel(G.pore.data(), nx, f, i, j);
        std::vector<double> lp((size_t)lx * ly), kn((size_t)lx * ly);
        const int rc = mts_shape_properties(dom.data(), lx, ly, ph.Cl * f, ph.mfp, lp.data(), kn.data(), nullptr);
        if (rc != MTS_OK) {
            err = std::string("per-level pore size (level ") + std::to_string(lvl + 1) + "): " + mts_last_error();
            return rc;
        }
        std::vector<double>& T = tauMap[lvl];
        T.resize((size_t)lx * ly);
        for (size_t p = 0; p < T.size(); p++) T[p] = 0.5 + std::sqrt(6.0 / M_PI) * lp[p] * kn[p] / (1 + 2 * kn[p]);
    }
    G.taus.resize(G.ncell);
    for (int64_t c = 0; c < G.ncell; c++) {
        const int L = G.level[c], f = 1 << (L - 1), lx = nx / f;
        G.taus[c] = tauMap[L - 1][(size_t)(G.cy[c] / f) * lx + G.cx[c] / f];
    }
    return MTS_OK;
}

}  // namespace mts

namespace mts {

int build_level_images(const MultiGeometry& G, std::vector<LevelImage>& Ls, std::vector<CellParam>& params,
                       const Physics& ph, std::string& err, bool meta) {
    Ls.assign(G.maxLevel, LevelImage());
    for (int l = 0; l < G.maxLevel; l++) {
        LevelImage& L = Ls[l];
        L.level = l + 1;
        L.nx = G.nx >> l;
        L.ny = G.ny >> l;
        const int64_t P = (int64_t)L.nx * L.ny;
        L.stride = (P + 255) / 256 * 256;
        L.meta.assign(P, 0);
        L.pidx.assign(P, 0);
    }
    auto site = [&](int64_t c) -> int64_t {
        const int l = G.level[c] - 1;
        return (int64_t)(G.cy[c] >> l) * Ls[l].nx + (G.cx[c] >> l);
    };
    // parameter table over distinct (tau_s, Kn)
    std::map<std::pair<double, double>, int> uniq;
    params.clear();
    for (int64_t c = 0; c < G.ncell; c++) {
        LevelImage& L = Ls[G.level[c] - 1];
        const int64_t s = site(c);
        L.cells.push_back(c);
        L.sites.push_back((int32_t)s);
        auto key = std::make_pair(G.taus[c], G.kn[c]);
        auto it = uniq.find(key);
        int id;
        if (it == uniq.end()) {
            id = (int)params.size();
            if (id > 65535) {
                err = "more than 65536 distinct (tau_s, Kn) pairs";
                return MTS_ERR_UNSUPPORTED;
            }
            uniq.emplace(key, id);
            params.push_back(make_param(G.taus[c], G.kn[c], ph));
        } else
            id = it->second;
        L.pidx[s] = (uint16_t)id;
    }
    if (!meta) return MTS_OK;  // the link words come from the device builder
    const int half = 1 << (G.maxLevel - 1);
    for (int64_t c = 0; c < G.ncell; c++) {
        LevelImage& L = Ls[G.level[c] - 1];
        const int64_t s = site(c);
        const int bt = G.bnd[c];
        uint32_t w = FLUID_BIT;
        uint32_t wall = 0;
        for (int ic = 1; ic < 9; ic++) {
            const int64_t nb = G.connect[c * 9 + ic];
            if (nb >= 0) {
                bool xp = std::abs(G.cx[c] - G.cx[nb]) > half, yp = std::abs(G.cy[c] - G.cy[nb]) > half;
                if (G.torus) xp = yp = false;
                if (!xp && !yp) {
                    if (bt == -1 || MTS_IFSTREAM[bt][ic]) {
                        Ls[G.level[nb] - 1].meta[site(nb)] |= 1u << (IN_SHIFT + ic - 1);
                        w |= 1u << (OUT_SHIFT + ic - 1);
                    } else
                        wall |= 1u << (ic - 1);
                } else if (xp) {
                    Ls[G.level[nb] - 1].meta[site(nb)] |= 1u << (IN_SHIFT + ic - 1);
                    w |= 1u << (OUT_SHIFT + ic - 1);
                }
            } else if (nb == -2) {
                if (!(bt == -1 || MTS_IFSTREAM[bt][ic])) wall |= 1u << (ic - 1);
            } else {  // wall pixel; a typeless (buffer) cell makes it a wall link
                if (bt < 0 || !MTS_IFSTREAM[bt][ic]) wall |= 1u << (ic - 1);
            }
        }
        const uint32_t type = bt >= 0 ? (uint32_t)bt : (wall ? (uint32_t)MTS_NTYPES : TYPE_NONE);
        w |= wall << WALL_SHIFT;
        w |= type << TYPE_SHIFT;
        const int xl = G.cx[c] >> (G.level[c] - 1);
        if (!G.torus && (xl == 0 || xl == L.nx - 1)) w |= SEAM_BIT;
        L.meta[s] |= w;
    }
    return MTS_OK;
}

}  // namespace mts
