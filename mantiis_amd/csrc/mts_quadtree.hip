// Quadtree tables built on the device: the level decision, the neighbour
// table (connectivity), boundary types, parent/child rows, the coarse and
// fine interface rows, and the link words of the level images.  The host
// keeps the one sequential stage -- the buffer cells, whose placement
// depends on the buffers made before them (grid.hpp:665-806) -- and the
// parameter table.  Every table equals the host builder's
// (mts_multigrid.cpp) entry for entry: the rules are integer lookups, and
// the scattered link bits are ORs, which commute.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <hipcub/hipcub.hpp>
#include <string>
#include <vector>

#include "mantiis_hip.h"
#include "mts_common.h"
#include "mts_multigrid.h"
#include "mts_tables.inc"

namespace mts {
int set_error(int code, const std::string& msg);  // mts_api.hip

namespace qt {

__constant__ int c_bndcode[MTS_NTYPES];
__constant__ unsigned char c_ifstream[MTS_NTYPES][9];

__device__ __forceinline__ int wrapd(int v, int n) { return ((v % n) + n) % n; }
__device__ __forceinline__ int exq(int k) { return (k == 1 || k == 5 || k == 8) ? 1 : ((k == 3 || k == 6 || k == 7) ? -1 : 0); }
__device__ __forceinline__ int eyq(int k) { return (k == 2 || k == 5 || k == 6) ? 1 : ((k == 4 || k == 7 || k == 8) ? -1 : 0); }

// Level decision (grid.hpp:1160-1269): launched for l = K .. 2 in turn; an
// aligned block of side h = 2^(l-1) not inside a larger leaf becomes a leaf
// of level l when its pixels are fluid and no wall pixel lies at a squared
// distance below T^2 (T = ceil(2^(K-1) sqrt 2) h) from any of them -- the
// exact-distance test d >= T of the host's Euclidean transform.  Wall
// pixels beyond the image edge do not exist unless the lattice is a torus.
__global__ void k_levels(int nx, int ny, int torus, int l, int T, const int32_t* __restrict__ pore,
                         uint8_t* __restrict__ lvl) {
    const int h = 1 << (l - 1), bnx = nx / h, bny = ny / h;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bnx * bny) return;
    const int x0 = (b % bnx) * h, y0 = (b / bnx) * h;
    if (lvl[(size_t)y0 * nx + x0] != 1) return;  // inside a larger leaf
    bool ok = true;
    const int T2 = T * T;
    for (int oy = -(T - 1); oy <= h - 1 + (T - 1) && ok; oy++) {
        const int ddy = oy < 0 ? -oy : (oy >= h ? oy - h + 1 : 0);
        int yy = y0 + oy;
        if (torus) yy = wrapd(yy, ny);
        else if (yy < 0 || yy >= ny) continue;
        for (int ox = -(T - 1); ox <= h - 1 + (T - 1) && ok; ox++) {
            const int ddx = ox < 0 ? -ox : (ox >= h ? ox - h + 1 : 0);
            if (ddx * ddx + ddy * ddy >= T2) continue;
            int xx = x0 + ox;
            if (torus) xx = wrapd(xx, nx);
            else if (xx < 0 || xx >= nx) continue;
            if (pore[(size_t)yy * nx + xx] != 1) ok = false;
        }
    }
    if (!ok) return;
    for (int y = y0; y < y0 + h; y++)
        for (int x = x0; x < x0 + h; x++) lvl[(size_t)y * nx + x] = (uint8_t)l;
}

struct Cells {
    int nx, ny, L;
    int64_t n;
    const int32_t *cx, *cy, *level, *type, *isBuffer;
    const uint8_t* lvl;             // level of each pixel's leaf
    int32_t* at[MTS_MAX_LEVELS];    // per level position: the lowest cell there (INT_MAX: none)
    int lnx[MTS_MAX_LEVELS];        // level image widths
};

__device__ __forceinline__ int32_t& at_of(const Cells& C, int l, int x, int y) {
    return C.at[l - 1][(size_t)(y >> (l - 1)) * C.lnx[l - 1] + (x >> (l - 1))];
}

__global__ void k_fill(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_index(Cells C) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.n) return;
    atomicMin(&at_of(C, C.level[c], C.cx[c], C.cy[c]), (int32_t)c);
}

// grid.hpp:916-972: neighbour k of cell c at its level's spacing (wrapped),
// -1 when an original cell of that level (a wall pixel) sits there, -2 when
// nothing does.
__global__ void k_connect(Cells C, int64_t* __restrict__ connect) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.n * 9) return;
    const int64_t c = i / 9;
    const int k = (int)(i - c * 9);
    const int l = C.level[c], jump = 1 << (l - 1);
    const int x = wrapd(C.cx[c] + exq(k) * jump, C.nx), y = wrapd(C.cy[c] + eyq(k) * jump, C.ny);
    const int32_t q = at_of(C, l, x, y);
    connect[i] = q != INT_MAX ? (int64_t)q : (C.lvl[(size_t)y * C.nx + x] == l ? -1 : -2);
}

// grid.hpp:1111-1157: the boundary type of a non-buffer cell from the
// wall/fluid code of its neighbourhood; an unknown code is reported through
// *bad (the lowest such cell).
__global__ void k_bnd(Cells C, const int64_t* __restrict__ connect, int32_t* __restrict__ bnd,
                      unsigned long long* bad) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.n) return;
    if (C.isBuffer[c]) {
        bnd[c] = -1;
        return;
    }
    const int order[9] = {7, 3, 6, 4, 0, 2, 8, 1, 5};
    int code = 0;
    for (int b = 0; b < 9; b++) code = (code << 1) | (connect[c * 9 + order[b]] == -1 ? 0 : 1);
    int t = code == 511 ? -1 : -2;
    for (int k = 0; k < MTS_NTYPES && t == -2; k++)
        if (c_bndcode[k] == code) t = k;
    bnd[c] = t;
    if (t == -2) atomicMin(bad, (unsigned long long)c);
}

// grid.hpp:974-1038: row rowOf[c] of a cell of level >= 2 {c, children
// tl, tr, bl, br} (-1 where no cell of the level below is).
__global__ void k_parent(Cells C, const int32_t* __restrict__ rowOf, int64_t* __restrict__ pc) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.n || C.level[c] <= 1) return;
    const int64_t r = rowOf[c];
    const int l = C.level[c], hj = 1 << (l - 2);
    pc[5 * r] = c;
    for (int s = 0; s < 4; s++) {
        const int32_t q = at_of(C, l - 1, C.cx[c] + (s & 1) * hj, C.cy[c] + (s >> 1) * hj);
        pc[5 * r + 1 + s] = q != INT_MAX ? (int64_t)q : -1;
    }
}

// Coarse buffer rows of one type: the parent/child row of each, in cell order.
__global__ void k_coarse_rows(Cells C, int ctype, const int32_t* __restrict__ rowOf, const int32_t* __restrict__ pos,
                              const int64_t* __restrict__ pc, int64_t* __restrict__ rows) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.n || C.type[c] != ctype) return;
    const int64_t r = rowOf[c], o = pos[c];
    for (int k = 0; k < 5; k++) rows[5 * o + k] = pc[5 * r + k];
}

// The first parent/child row that names each cell as a child.
__global__ void k_first_row(int64_t nrows, const int64_t* __restrict__ pc, int32_t* __restrict__ first) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    for (int k = 1; k <= 4; k++) {
        const int64_t q = pc[5 * r + k];
        if (q >= 0) atomicMin(&first[q], (int32_t)r);
    }
}

// grid.hpp:1041-1108: the fine buffer cells of one type in cell order ...
__global__ void k_fine_ids(Cells C, int ftype, const int32_t* __restrict__ pos, int64_t* __restrict__ ids) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < C.n && C.type[c] == ftype) ids[pos[c]] = c;
}

// ... four at a time, under the parent of the first row naming any of the four.
__global__ void k_fine_groups(const int32_t* __restrict__ first, const int64_t* __restrict__ pc, int64_t ngroups,
                              const int64_t* __restrict__ ids, int64_t* __restrict__ rows, unsigned long long* bad) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    int32_t best = INT_MAX;
    for (int k = 0; k < 4; k++) best = min(best, first[ids[4 * g + k]]);
    if (best == INT_MAX) {
        atomicMin(bad, (unsigned long long)g);
        return;
    }
    rows[5 * g] = pc[5 * (int64_t)best];
    for (int k = 0; k < 4; k++) rows[5 * g + 1 + k] = ids[4 * g + k];
}

// Link words of the level images (the classification of coreLB.hpp:292-356
// on the multi-block connectivity, as build_level_images forms it): each
// cell ORs its own word into its site and its neighbours' "in" bits into
// theirs.
struct Meta {
    uint32_t* meta[MTS_MAX_LEVELS];
    int lnx[MTS_MAX_LEVELS];
};
__global__ void k_meta(Cells C, const int64_t* __restrict__ connect, const int32_t* __restrict__ bnd, int torus,
                       int half, Meta M) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.n) return;
    auto site = [&](int64_t q) -> int64_t {
        const int l = C.level[q] - 1;
        return (int64_t)(C.cy[q] >> l) * M.lnx[l] + (C.cx[q] >> l);
    };
    const int bt = bnd[c];
    uint32_t w = FLUID_BIT, wall = 0;
    for (int ic = 1; ic < 9; ic++) {
        const int64_t nb = connect[c * 9 + ic];
        const bool streams = bt == -1 || c_ifstream[bt < 0 ? 0 : bt][ic];
        if (nb >= 0) {
            bool xp = abs(C.cx[c] - C.cx[nb]) > half, yp = abs(C.cy[c] - C.cy[nb]) > half;
            if (torus) xp = yp = false;
            if (!xp && !yp) {
                if (streams) {
                    atomicOr(&M.meta[C.level[nb] - 1][site(nb)], 1u << (IN_SHIFT + ic - 1));
                    w |= 1u << (OUT_SHIFT + ic - 1);
                } else {
                    wall |= 1u << (ic - 1);
                }
            } else if (xp) {
                atomicOr(&M.meta[C.level[nb] - 1][site(nb)], 1u << (IN_SHIFT + ic - 1));
                w |= 1u << (OUT_SHIFT + ic - 1);
            }
        } else if (nb == -2) {
            if (!streams) wall |= 1u << (ic - 1);
        } else {  // wall pixel; a typeless (buffer) cell makes it a wall link
            if (bt < 0 || !c_ifstream[bt][ic]) wall |= 1u << (ic - 1);
        }
    }
    const uint32_t type = bt >= 0 ? (uint32_t)bt : (wall ? (uint32_t)MTS_NTYPES : TYPE_NONE);
    w |= wall << WALL_SHIFT;
    w |= type << TYPE_SHIFT;
    const int l = C.level[c] - 1;
    const int xl = C.cx[c] >> l;
    if (!torus && (xl == 0 || xl == M.lnx[l] - 1)) w |= SEAM_BIT;
    atomicOr(&M.meta[l][site(c)], w);
}

}  // namespace qt

namespace {

struct DevBuf {
    std::vector<void*> ptrs;
    ~DevBuf() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T* get(size_t n, hipError_t& e) {
        void* p = nullptr;
        if (e == hipSuccess) e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(p);
        return (T*)p;
    }
};

unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>((n + 255) / 256, 1); }

int hip_fail(hipError_t e, const char* what, std::string& err) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return MTS_ERR_HIP;
}

// Exclusive prefix sum of the 0/1 flags (int32) into out.
hipError_t exclusive_sum(const int32_t* flags, int32_t* out, int64_t n, hipStream_t st, DevBuf& B) {
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, flags, out, (int)n, st);
    void* tmp = B.get<char>(bytes, e);
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, flags, out, (int)n, st);
    return e;
}

__global__ void k_flag(const int32_t* __restrict__ a, int64_t n, int mode, int32_t* __restrict__ f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f[i] = mode == 0 ? (a[i] > 1 ? 1 : 0) : (a[i] == mode ? 1 : 0);  // 0: level > 1; else: type == mode
}

}  // namespace

int multi_levels_device(int nx, int ny, const int32_t* pore, bool torus, int K, hipStream_t st,
                        std::vector<uint8_t>& lvl, std::string& err) {
    DevBuf B;
    hipError_t e = hipSuccess;
    const int64_t P = (int64_t)nx * ny;
    int32_t* d_pore = B.get<int32_t>(P, e);
    uint8_t* d_lvl = B.get<uint8_t>(P, e);
    if (e == hipSuccess) e = hipMemcpyAsync(d_pore, pore, P * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_lvl, 1, P, st);
    const int md = split_distance(K);
    for (int l = K; l >= 2 && e == hipSuccess; l--) {
        const int h = 1 << (l - 1);
        const int64_t nb = (int64_t)(nx / h) * (ny / h);
        qt::k_levels<<<grid_of(nb), 256, 0, st>>>(nx, ny, torus ? 1 : 0, l, md * h, d_pore, d_lvl);
        e = hipGetLastError();
    }
    lvl.assign(P, 1);
    if (e == hipSuccess) e = hipMemcpyAsync(lvl.data(), d_lvl, P, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? MTS_OK : hip_fail(e, "quadtree level decision", err);
}

int multi_tables_device(MultiGeometry& G, std::vector<LevelImage>* levels, hipStream_t st, std::string& err) {
    const int nx = G.nx, ny = G.ny, L = G.maxLevel;
    const int64_t n = G.ncell;
    if (n == 0) {  // no fluid cell: no tables (as the host builder leaves them)
        return multi_tables_host(G, err);
    }
    if (L > MTS_MAX_LEVELS) {
        err = "quadtree tables: too many levels";
        return MTS_ERR_UNSUPPORTED;
    }
    DevBuf B;
    hipError_t e = hipSuccess;
    qt::Cells C;
    C.nx = nx;
    C.ny = ny;
    C.L = L;
    C.n = n;
    int32_t* cx = B.get<int32_t>(n, e);
    int32_t* cy = B.get<int32_t>(n, e);
    int32_t* level = B.get<int32_t>(n, e);
    int32_t* type = B.get<int32_t>(n, e);
    int32_t* isBuf = B.get<int32_t>(n, e);
    uint8_t* d_lvl = B.get<uint8_t>((size_t)nx * ny, e);
    for (int l = 0; l < MTS_MAX_LEVELS; l++) {
        C.lnx[l] = nx >> l;
        C.at[l] = l < L ? B.get<int32_t>((size_t)(nx >> l) * (ny >> l), e) : nullptr;
    }
    int64_t* connect = B.get<int64_t>((size_t)n * 9, e);
    int32_t* bnd = B.get<int32_t>(n, e);
    int32_t* flag = B.get<int32_t>(n, e);
    int32_t* rowOf = B.get<int32_t>(n, e);
    int32_t* pos = B.get<int32_t>(n, e);
    int32_t* first = B.get<int32_t>(n, e);
    unsigned long long* bad = B.get<unsigned long long>(2, e);
    if (e != hipSuccess) return hip_fail(e, "quadtree tables: device memory", err);
    auto up = [&](int32_t* d, const std::vector<int32_t>& h) {
        if (e == hipSuccess && n > 0) e = hipMemcpyAsync(d, h.data(), n * 4, hipMemcpyHostToDevice, st);
    };
    up(cx, G.cx);
    up(cy, G.cy);
    up(level, G.level);
    up(type, G.type);
    up(isBuf, G.isBuffer);
    if (e == hipSuccess) e = hipMemcpyAsync(d_lvl, G.lvl.data(), G.lvl.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(qt::c_bndcode), MTS_BND_CODE, sizeof(MTS_BND_CODE), 0,
                                                    hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyToSymbolAsync(HIP_SYMBOL(qt::c_ifstream), MTS_IFSTREAM, sizeof(MTS_IFSTREAM), 0,
                                                    hipMemcpyHostToDevice, st);
    const unsigned long long none[2] = {ULLONG_MAX, ULLONG_MAX};
    if (e == hipSuccess) e = hipMemcpyAsync(bad, none, sizeof(none), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "quadtree tables: upload", err);
    C.cx = cx;
    C.cy = cy;
    C.level = level;
    C.type = type;
    C.isBuffer = isBuf;
    C.lvl = d_lvl;
    for (int l = 0; l < L; l++) {
        const int64_t P = (int64_t)(nx >> l) * (ny >> l);
        qt::k_fill<<<grid_of(P), 256, 0, st>>>(C.at[l], P, INT_MAX);
    }
    qt::k_fill<<<grid_of(n), 256, 0, st>>>(first, n, INT_MAX);
    qt::k_index<<<grid_of(n), 256, 0, st>>>(C);
    qt::k_connect<<<grid_of(n * 9), 256, 0, st>>>(C, connect);
    qt::k_bnd<<<grid_of(n), 256, 0, st>>>(C, connect, bnd, bad);
    // parent/child rows: one per cell of level >= 2, in cell order
    k_flag<<<grid_of(n), 256, 0, st>>>(level, n, 0, flag);
    e = hipGetLastError();
    if (e == hipSuccess) e = exclusive_sum(flag, rowOf, n, st, B);
    if (e != hipSuccess) return hip_fail(e, "quadtree tables: scans", err);
    int64_t nrows = 0;
    std::vector<int64_t> nct(std::max(L - 1, 0), 0), nft(std::max(L - 1, 0), 0);
    for (int64_t c = 0; c < n; c++) {
        nrows += G.level[c] > 1;
        for (int l = 1; l < L; l++) {
            nct[l - 1] += G.type[c] == 10 * (l + 1) + l;
            nft[l - 1] += G.type[c] == 10 * l + l + 1;
        }
    }
    int64_t ncrow = 0, nfrow = 0, nfid = 0;
    for (int l = 1; l < L; l++) {
        ncrow += nct[l - 1];
        nfrow += nft[l - 1] / 4;
        nfid = std::max(nfid, nft[l - 1]);
    }
    int64_t* pc = B.get<int64_t>((size_t)nrows * 5, e);
    int64_t* crow = B.get<int64_t>((size_t)ncrow * 5, e);
    int64_t* ids = B.get<int64_t>((size_t)nfid, e);
    int64_t* frow = B.get<int64_t>((size_t)nfrow * 5, e);
    if (e != hipSuccess) return hip_fail(e, "quadtree tables: device memory", err);
    qt::k_parent<<<grid_of(n), 256, 0, st>>>(C, rowOf, pc);
    qt::k_first_row<<<grid_of(nrows), 256, 0, st>>>(nrows, pc, first);
    // interface rows per pair of levels, one pair after the other
    int64_t co = 0, fo = 0;
    for (int l = 1; l < L && e == hipSuccess; l++) {
        const int ctype = 10 * (l + 1) + l, ftype = 10 * l + l + 1;
        k_flag<<<grid_of(n), 256, 0, st>>>(type, n, ctype, flag);
        e = exclusive_sum(flag, pos, n, st, B);
        if (e != hipSuccess) break;
        qt::k_coarse_rows<<<grid_of(n), 256, 0, st>>>(C, ctype, rowOf, pos, pc, crow + 5 * co);
        k_flag<<<grid_of(n), 256, 0, st>>>(type, n, ftype, flag);
        e = exclusive_sum(flag, pos, n, st, B);
        if (e != hipSuccess) break;
        const int64_t ngroups = nft[l - 1] / 4;
        qt::k_fine_ids<<<grid_of(n), 256, 0, st>>>(C, ftype, pos, ids);
        qt::k_fine_groups<<<grid_of(ngroups), 256, 0, st>>>(first, pc, ngroups, ids, frow + 5 * fo, bad + 1);
        e = hipGetLastError();
        co += nct[l - 1];
        fo += ngroups;
    }
    // link words of the level images
    uint32_t* meta[MTS_MAX_LEVELS] = {};
    if (levels && e == hipSuccess) {
        qt::Meta M;
        for (int l = 0; l < MTS_MAX_LEVELS; l++) {
            M.meta[l] = nullptr;
            M.lnx[l] = nx >> l;
        }
        for (int l = 0; l < L; l++) {
            const int64_t P = (int64_t)(nx >> l) * (ny >> l);
            meta[l] = B.get<uint32_t>(P, e);
            if (e == hipSuccess) qt::k_fill<<<grid_of(P), 256, 0, st>>>((int32_t*)meta[l], P, 0);
            M.meta[l] = meta[l];
        }
        if (e == hipSuccess) {
            qt::k_meta<<<grid_of(n), 256, 0, st>>>(C, connect, bnd, G.torus ? 1 : 0, 1 << (L - 1), M);
            e = hipGetLastError();
        }
    }
    if (e != hipSuccess) return hip_fail(e, "quadtree tables: kernels", err);
    // back to the host grid
    unsigned long long badh[2];
    G.connect.resize((size_t)n * 9);
    G.bnd.resize(n);
    G.parentChild.resize((size_t)nrows * 5);
    G.coarseRows.resize((size_t)ncrow * 5);
    G.fineRows.resize((size_t)nfrow * 5);
    G.coarseCount = nct;
    G.fineCount.assign(std::max(L - 1, 0), 0);
    for (int l = 1; l < L; l++) G.fineCount[l - 1] = nft[l - 1] / 4;
    auto down = [&](void* h, const void* d, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st);
    };
    down(badh, bad, sizeof(badh));
    down(G.connect.data(), connect, G.connect.size() * 8);
    down(G.bnd.data(), bnd, G.bnd.size() * 4);
    down(G.parentChild.data(), pc, G.parentChild.size() * 8);
    down(G.coarseRows.data(), crow, G.coarseRows.size() * 8);
    down(G.fineRows.data(), frow, G.fineRows.size() * 8);
    if (levels)
        for (int l = 0; l < L; l++) {
            LevelImage& I = (*levels)[l];
            down(I.meta.data(), meta[l], I.meta.size() * 4);
        }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "quadtree tables: download", err);
    if (badh[0] != ULLONG_MAX) {
        const int64_t c = (int64_t)badh[0];
        err = "[ERROR]: Boundary type is not contained. (cell x=" + std::to_string(G.cx[c]) + " y=" +
              std::to_string(G.cy[c]) + ")";
        return MTS_ERR_GEOMETRY;
    }
    if (badh[1] != ULLONG_MAX) {
        err = "fine buffer cells without a parent";
        return MTS_ERR_GEOMETRY;
    }
    return MTS_OK;
}

}  // namespace mts

extern "C" {

int mts_quadtree_check(int nx, int ny, const int32_t* pore, const double* kn, const mts_params* p, int64_t* ndiff) {
    using namespace mts;
    if (!pore || !kn || !ndiff) return set_error(MTS_ERR_ARG, "null argument");
    *ndiff = -1;
    mts_params prm;
    if (p) prm = *p;
    else mts_default_params(&prm);
    const bool torus = prm.periodic_y != 0;
    const int K = quadtree_levels(prm.multiblock);
    std::string err;
    int rc = multi_check(nx, ny, pore, K, err);
    if (rc) return set_error(rc, err);
    hipError_t e = hipSetDevice(prm.device);
    if (e != hipSuccess) return set_error(MTS_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    hipStream_t st = nullptr;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) return set_error(MTS_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
    // host: transform-based level decision and the host tables
    std::vector<uint8_t> ch, cd;
    multi_levels_host(nx, ny, pore, torus, K, ch);
    MultiGeometry H, D;
    rc = multi_cells(nx, ny, pore, kn, torus, ch, H, err);
    if (!rc) rc = multi_tables_host(H, err);
    // device: window-based level decision and the device tables
    if (!rc) rc = multi_levels_device(nx, ny, pore, torus, K, st, cd, err);
    if (!rc) rc = multi_cells(nx, ny, pore, kn, torus, cd, D, err);
    std::vector<LevelImage> LH, LD;
    std::vector<CellParam> par;
    Physics ph;
    ph.Cl = prm.dx;
    if (!rc) {
        H.taus.assign(H.ncell, 1.0);
        D.taus.assign(D.ncell, 1.0);
        rc = build_level_images(H, LH, par, ph, err, true);
    }
    if (!rc) rc = build_level_images(D, LD, par, ph, err, false);
    if (!rc) rc = multi_tables_device(D, &LD, st, err);
    (void)hipStreamDestroy(st);
    if (rc) return set_error(rc, err);
    int64_t d = 0;
    auto cmp = [&](const auto& a, const auto& b) {
        if (a.size() != b.size()) {
            d += (int64_t)std::max(a.size(), b.size());
            return;
        }
        for (size_t i = 0; i < a.size(); i++) d += a[i] != b[i];
    };
    cmp(ch, cd);
    cmp(H.cx, D.cx);
    cmp(H.cy, D.cy);
    cmp(H.level, D.level);
    cmp(H.type, D.type);
    cmp(H.connect, D.connect);
    cmp(H.bnd, D.bnd);
    cmp(H.parentChild, D.parentChild);
    cmp(H.coarseRows, D.coarseRows);
    cmp(H.fineRows, D.fineRows);
    cmp(H.coarseCount, D.coarseCount);
    cmp(H.fineCount, D.fineCount);
    if (LH.size() != LD.size()) d += 1;
    for (size_t l = 0; l < LH.size() && l < LD.size(); l++) cmp(LH[l].meta, LD[l].meta);
    *ndiff = d;
    return MTS_OK;
}

}  // extern "C"
