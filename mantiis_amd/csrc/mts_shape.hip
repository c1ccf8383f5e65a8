// Geometry preprocessing (C ABI mts_shape_*): the local pore size and
// Knudsen number of every fluid pixel of a binary pore image, as
// Shape::calculateProperties computes them (shape.hpp:127-251):
//   1. a canvas with max(nx, ny) pixels of margin on every side, the domain
//      in the middle and its interior rows run on as fluid across the side
//      margins (shape.hpp:131-150);
//   2. Zhang-Suen thinning of the canvas with the reference's bound of 2..5
//      set neighbours (utils.hpp:444-536), cropped back to the domain;
//   3. exact Euclidean distance of every pixel to the nearest wall and to the
//      nearest skeleton pixel, as OpenCV's float values (shape.hpp:184-210);
//   4. per fluid pixel, the first skeleton pixel in row-major order at the
//      skeleton distance (utils.hpp:116-134); LocPore = 2 x its wall
//      distance, Kn = mfp / LocPore / resolution (shape.hpp:212-250).
//
// Layout on the device: the canvas is bit-packed, one uint64 per 64 pixels
// of a row (bit b = column 64 w + b), so that a thread evaluates the
// deletion rule for 64 pixels with bitwise logic (bit-sliced counters for
// the neighbour count and the 0->1 steps around the ring).  One launch runs
// THIN_S subiterations of a 128 x 64-pixel tile in LDS (temporal blocking:
// a halo of one word left and right and THIN_S rows above and below keeps
// the inner tile exact).  Distances are exact integer squares: per-column
// nearest zero (chunked scans), then the row minimum with early exit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mantiis_hip.h"

namespace mts {
int set_error(int code, const std::string& msg);  // mts_api.hip
}
using mts::set_error;

#define SH_TRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(e_ == hipErrorOutOfMemory ? MTS_ERR_OOM : MTS_ERR_HIP,        \
                             std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

namespace {

typedef unsigned long long u64;

constexpr int THIN_S = 32;                     // subiterations per launch (even)
constexpr int THIN_TW = 2;                     // inner tile width, words
constexpr int THIN_TH = 64;                    // inner tile height, rows
constexpr int THIN_LW = THIN_TW + 2;           // LDS words per row (one halo word each side)
constexpr int THIN_LR = THIN_TH + 2 * THIN_S;  // LDS rows
constexpr int THIN_THREADS = THIN_LW * THIN_LR;  // one word per thread
static_assert(THIN_S % 2 == 0 && THIN_THREADS <= 1024, "tile shape");
static_assert(THIN_S < THIN_TH && THIN_S < 64 * THIN_TW, "a launch's reach must stay inside the neighbour tiles");
constexpr int SEG = 32;                      // rows per column chunk
constexpr int MAX_SIDE = 2048;               // exact float distance sums
constexpr int CHECK_EVERY = 8;               // launches between convergence reads

// error bits raised by the kernels
constexpr int ERR_VALUE = 1;      // a domain byte other than 0 / 1
constexpr int ERR_NO_SKEL = 2;    // no skeleton pixel inside the domain
constexpr int ERR_NOT_FOUND = 4;  // no skeleton pixel at the skeleton distance

// Canvas word (r, w): bit b is canvas pixel (r, 64 w + b).
__global__ void k_canvas(const uint8_t* __restrict__ dom, int nx, int ny, int m, int H, int Wc, int WW,
                         u64* __restrict__ cv, int* __restrict__ err) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)H * WW) return;
    const int r = (int)(t / WW), w = (int)(t % WW);
    const int i = r - m;
    u64 bits = 0;
    if (i >= 0 && i < ny) {
        const bool band = i >= 1 && i < ny - 1;
        int bad = 0;
        for (int b = 0; b < 64; b++) {
            const int col = 64 * w + b;
            if (col >= Wc) break;
            const int j = col - m;
            bool v;
            if (j >= 0 && j < nx) {
                const uint8_t d = dom[(size_t)i * nx + j];
                bad |= d > 1;
                v = d == 1;
            } else {
                v = band;
            }
            bits |= (u64)v << b;
        }
        if (bad) atomicOr(err, ERR_VALUE);
    }
    cv[t] = bits;
}

// Truth table of a 3-input bitwise function f(a, b, c): bit (a << 2 | b << 1 | c).
template <class F>
constexpr unsigned lut3(F f) {
    unsigned r = 0;
    for (int i = 0; i < 8; i++)
        if (f((i >> 2) & 1, (i >> 1) & 1, i & 1)) r |= 1u << i;
    return r;
}
constexpr unsigned L_STEP = lut3([](int a, int b, int) { return !a && b; });         // a 0 -> 1 step
constexpr unsigned L_SEEN = lut3([](int s, int a, int b) { return s || (!a && b); });  // seen | step
constexpr unsigned L_TWICE = lut3([](int s, int a, int b) { return s && !a && b; });  // seen & step
constexpr unsigned L_XOR3 = lut3([](int a, int b, int c) { return (a ^ b ^ c) != 0; });
constexpr unsigned L_MAJ = lut3([](int a, int b, int c) { return a + b + c >= 2; });
constexpr unsigned L_OR3 = lut3([](int a, int b, int c) { return a || b || c; });
constexpr unsigned L_K3 = lut3([](int k, int s, int c) { return k && (s || c); });
constexpr unsigned L_ANDN = lut3([](int a, int b, int) { return a && !b; });
constexpr unsigned L_ONE = lut3([](int o, int sm, int th) { return o && sm && !th; });
constexpr unsigned L_DEL = lut3([](int a, int t, int m) { return a && !t && m; });
constexpr unsigned L_AND3 = lut3([](int a, int b, int c) { return a && b && c; });

// One v_bitop3_b32: f applied bitwise to a, b, c.
template <unsigned LUT>
__device__ __forceinline__ uint32_t op3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, LUT);
}

// The pixels of a 32-pixel half word that survive subiteration SUB, given
// the half words p0..p7 of their neighbours N, NE, E, SE, S, SW, W, NW.  A
// set pixel in mask is deleted when (utils.hpp:495-503): exactly one 0 -> 1
// step goes round the ring; B, the number of set neighbours, lies in 2..5;
// and N*E*S = E*S*W = 0 (SUB 0) or N*E*W = N*S*W = 0 (SUB 1).
template <int SUB>
__device__ __forceinline__ uint32_t zs_keep(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3, uint32_t p4,
                                            uint32_t p5, uint32_t p6, uint32_t p7, uint32_t me, uint32_t mask) {
    // steps around the ring: seen = at least one, twice = at least two
    uint32_t seen = op3<L_STEP>(p0, p1, 0u);
    uint32_t twice = op3<L_TWICE>(seen, p1, p2);
    seen = op3<L_SEEN>(seen, p1, p2);
    twice |= op3<L_TWICE>(seen, p2, p3);
    seen = op3<L_SEEN>(seen, p2, p3);
    twice |= op3<L_TWICE>(seen, p3, p4);
    seen = op3<L_SEEN>(seen, p3, p4);
    twice |= op3<L_TWICE>(seen, p4, p5);
    seen = op3<L_SEEN>(seen, p4, p5);
    twice |= op3<L_TWICE>(seen, p5, p6);
    seen = op3<L_SEEN>(seen, p5, p6);
    twice |= op3<L_TWICE>(seen, p6, p7);
    seen = op3<L_SEEN>(seen, p6, p7);
    twice |= op3<L_TWICE>(seen, p7, p0);
    seen = op3<L_SEEN>(seen, p7, p0);
    const uint32_t oneStep = op3<L_ANDN>(seen, twice, 0u);
    // B = s4 + 2 (c1 + c2 + c3 + c4) by carry-save adders; with
    // c1 + c2 + c3 = s + 2 k: B >= 2 iff s | k | c4, B >= 6 iff k & (s | c4)
    const uint32_t c1 = op3<L_MAJ>(p0, p1, p2), s1 = op3<L_XOR3>(p0, p1, p2);
    const uint32_t c2 = op3<L_MAJ>(p3, p4, p5), s2 = op3<L_XOR3>(p3, p4, p5);
    const uint32_t c3 = p6 & p7, s3 = p6 ^ p7;
    const uint32_t c4 = op3<L_MAJ>(s1, s2, s3);
    const uint32_t k = op3<L_MAJ>(c1, c2, c3), s = op3<L_XOR3>(c1, c2, c3);
    const uint32_t atLeast6 = op3<L_K3>(k, s, c4);
    const uint32_t atLeast2 = op3<L_OR3>(s, k, c4);
    const uint32_t tri = SUB == 0 ? op3<L_AND3>(p2, p4, p0 | p6) : op3<L_AND3>(p0, p6, p2 | p4);
    const uint32_t del = op3<L_DEL>(op3<L_ONE>(oneStep, atLeast2, atLeast6), tri, mask);
    return op3<L_ANDN>(me, del, 0u);
}

template <int SUB>
__device__ __forceinline__ u64 zs_keep64(u64 p0, u64 p1, u64 p2, u64 p3, u64 p4, u64 p5, u64 p6, u64 p7, u64 me,
                                         u64 mask) {
    const uint32_t lo = zs_keep<SUB>((uint32_t)p0, (uint32_t)p1, (uint32_t)p2, (uint32_t)p3, (uint32_t)p4,
                                     (uint32_t)p5, (uint32_t)p6, (uint32_t)p7, (uint32_t)me, (uint32_t)mask);
    const uint32_t hi = zs_keep<SUB>((uint32_t)(p0 >> 32), (uint32_t)(p1 >> 32), (uint32_t)(p2 >> 32),
                                     (uint32_t)(p3 >> 32), (uint32_t)(p4 >> 32), (uint32_t)(p5 >> 32),
                                     (uint32_t)(p6 >> 32), (uint32_t)(p7 >> 32), (uint32_t)(me >> 32),
                                     (uint32_t)(mask >> 32));
    return (u64)hi << 32 | lo;
}

// THIN_S Zhang-Suen subiterations (starting with subiteration 0) of one
// tile; out receives the inner tile.  Thinning only deletes, and one launch
// moves information at most THIN_S pixels, so a tile whose 3 x 3 tile
// neighbourhood did not change in the previous launch cannot change in this
// one: it is skipped, and both ping-pong buffers already hold it (it was
// unchanged, or skipped, the launch before).  chgIn / chgOut: per tile,
// changed in the previous / this launch; cnt counts changed tiles.
__global__ void __launch_bounds__(THIN_THREADS)
    k_thin(const u64* __restrict__ in, u64* __restrict__ out, int H, int Wc, int WW, int tilesX, int tilesY,
           const uint8_t* __restrict__ chgIn, uint8_t* __restrict__ chgOut, unsigned* __restrict__ cnt) {
    const int bx = blockIdx.x, by = blockIdx.y;
    bool live = false;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            const int tx = bx + dx, ty = by + dy;
            if (tx >= 0 && tx < tilesX && ty >= 0 && ty < tilesY) live |= chgIn[ty * tilesX + tx] != 0;
        }
    if (!live) {
        if (threadIdx.x == 0) chgOut[by * tilesX + bx] = 0;
        return;
    }
    // LDS rows 1..THIN_LR and words 1..THIN_LW hold the tile and its halo;
    // the border around them is never written, so every neighbour load is
    // plain.  Whatever it holds reaches at most THIN_S pixels in, inside the
    // halo; it holds set pixels, so that a fluid halo next to it is a full
    // neighbourhood the early-out below skips rather than a false edge that
    // erodes on every subiteration.
    __shared__ u64 L[2][THIN_LR + 2][THIN_LW + 2];
    const int t = threadIdx.x;
    for (int e = t; e < 2 * (THIN_LR + 2); e += THIN_THREADS) {
        L[e & 1][e >> 1][0] = ~0ull;
        L[e & 1][e >> 1][THIN_LW + 1] = ~0ull;
    }
    for (int e = t; e < 2 * THIN_LW; e += THIN_THREADS) {
        const int b = e / THIN_LW, c = e % THIN_LW + 1;
        L[b][0][c] = ~0ull;
        L[b][THIN_LR + 1][c] = ~0ull;
    }
    const int lc = t % THIN_LW, lr = t / THIN_LW;
    const int gw = bx * THIN_TW - 1 + lc;
    const int gr = by * THIN_TH - THIN_S + lr;
    const bool inimg = gr >= 0 && gr < H && gw >= 0 && gw < WW;
    const u64 orig = inimg ? in[(size_t)gr * WW + gw] : 0ull;
    // pixels that may be deleted: image rows 1..H-2, columns 1..Wc-2
    u64 mask = 0;
    if (inimg && gr >= 1 && gr <= H - 2) {
        const int lo = max(1 - 64 * gw, 0), hi = min(Wc - 2 - 64 * gw, 63);
        if (lo <= hi) mask = (hi == 63 ? ~0ull : ((1ull << (hi + 1)) - 1)) & ~((1ull << lo) - 1);
    }
    u64 me = orig;
    L[0][lr + 1][lc + 1] = me;
    __syncthreads();
#pragma unroll 2
    for (int sub = 0; sub < THIN_S; sub++) {
        const int cur = sub & 1;
        const u64 up = L[cur][lr][lc + 1], dn = L[cur][lr + 2][lc + 1];
        const u64 upL = L[cur][lr][lc], upR = L[cur][lr][lc + 2];
        const u64 meL = L[cur][lr + 1][lc], meR = L[cur][lr + 1][lc + 2];
        const u64 dnL = L[cur][lr + 2][lc], dnR = L[cur][lr + 2][lc + 2];
        // nothing to delete in a word without deletable set pixels, or whose
        // whole neighbourhood is set (B = 8): whole waves skip the rule
        if ((me & mask) != 0 && (me & up & dn & upL & upR & meL & meR & dnL & dnR) != ~0ull) {
            // neighbour ring N, NE, E, SE, S, SW, W, NW (value of the pixel
            // one column right: word >> 1 with the next word's bit 0 on top)
            const u64 p0 = up;
            const u64 p1 = (up >> 1) | (upR << 63);
            const u64 p2 = (me >> 1) | (meR << 63);
            const u64 p3 = (dn >> 1) | (dnR << 63);
            const u64 p4 = dn;
            const u64 p5 = (dn << 1) | (dnL >> 63);
            const u64 p6 = (me << 1) | (meL >> 63);
            const u64 p7 = (up << 1) | (upL >> 63);
            me = cur == 0 ? zs_keep64<0>(p0, p1, p2, p3, p4, p5, p6, p7, me, mask)
                          : zs_keep64<1>(p0, p1, p2, p3, p4, p5, p6, p7, me, mask);
        }
        L[cur ^ 1][lr + 1][lc + 1] = me;
        __syncthreads();
    }
    const bool inner = inimg && lr >= THIN_S && lr < THIN_S + THIN_TH && lc >= 1 && lc <= THIN_TW;
    if (inner) out[(size_t)gr * WW + gw] = me;
    const bool changed = inner && me != orig;
    const bool any = __syncthreads_or(changed);
    if (t == 0) {
        chgOut[by * tilesX + bx] = any ? 1 : 0;
        if (any) atomicAdd(cnt, 1u);
    }
}

// Cropped skeleton byte (i, j) = canvas pixel (i + m, j + m).
__global__ void k_crop(const u64* __restrict__ cv, int WW, int m, int nx, int ny, uint8_t* __restrict__ skel) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)nx * ny) return;
    const int i = (int)(p / nx), j = (int)(p % nx);
    const int r = i + m, col = j + m;
    skel[p] = (uint8_t)((cv[(size_t)r * WW + (col >> 6)] >> (col & 63)) & 1ull);
}

// A "zero" pixel of image z (z = 0: domain, zero where the byte is 0;
// z = 1: skeleton, zero where the byte is set).
__device__ __forceinline__ bool zero_at(const uint8_t* img, int z, size_t p) { return z ? img[p] != 0 : img[p] == 0; }

// Per column chunk of SEG rows: its first and last zero row (-1: none).
__global__ void k_colseg(const uint8_t* __restrict__ dom, const uint8_t* __restrict__ skel, int nx, int ny,
                         int nseg, int* __restrict__ first, int* __restrict__ last) {
    const int z = blockIdx.y;
    const uint8_t* img = z ? skel : dom;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nseg * nx) return;
    const int sg = (int)(t / nx), j = (int)(t % nx);
    const int r0 = sg * SEG, r1 = min(r0 + SEG, ny);
    int f = -1, l = -1;
    for (int r = r0; r < r1; r++)
        if (zero_at(img, z, (size_t)r * nx + j)) {
            if (f < 0) f = r;
            l = r;
        }
    const size_t o = (size_t)z * nseg * nx + t;
    first[o] = f;
    last[o] = l;
}

// Distance in its column from every pixel to the nearest zero (-1: none),
// and that zero's row (the upper one of two at the same distance).
__global__ void k_colpass(const uint8_t* __restrict__ dom, const uint8_t* __restrict__ skel, int nx, int ny,
                          int nseg, const int* __restrict__ first, const int* __restrict__ last,
                          int* __restrict__ g, int* __restrict__ zrow) {
    const int z = blockIdx.y;
    const uint8_t* img = z ? skel : dom;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)nx * ny) return;
    const int i = (int)(p / nx), j = (int)(p % nx);
    const int sg = i / SEG;
    const int* F = first + (size_t)z * nseg * nx;
    const int* Ls = last + (size_t)z * nseg * nx;
    int up = -1, dn = -1;
    for (int r = i; r >= sg * SEG; r--)
        if (zero_at(img, z, (size_t)r * nx + j)) {
            up = r;
            break;
        }
    for (int s = sg - 1; up < 0 && s >= 0; s--) up = Ls[(size_t)s * nx + j];
    const int r1 = min((sg + 1) * SEG, ny);
    for (int r = i; r < r1; r++)
        if (zero_at(img, z, (size_t)r * nx + j)) {
            dn = r;
            break;
        }
    for (int s = sg + 1; dn < 0 && s < nseg; s++) dn = F[(size_t)s * nx + j];
    const int zr = up >= 0 && (dn < 0 || i - up <= dn - i) ? up : dn;
    g[(size_t)z * nx * ny + p] = zr < 0 ? -1 : (zr <= i ? i - zr : zr - i);
    zrow[(size_t)z * nx * ny + p] = zr;
}

// Squared distance to the nearest zero: min over the row of dx^2 + g^2
// (-1: the image has no zero pixel), and for the skeleton image the index of
// the first nearest zero in row-major order (ties resolved on (row, col):
// within a column only the column's nearest zero can reach the minimum).
__global__ void k_rowpass(const int* __restrict__ g, const int* __restrict__ zrow, int nx, int ny,
                          int* __restrict__ k2, int* __restrict__ nearest) {
    const int z = blockIdx.y;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)nx * ny) return;
    const int i = (int)(p / nx), j = (int)(p % nx);
    const int* row = g + (size_t)z * nx * ny + (size_t)i * nx;
    const int* zr = zrow + (size_t)z * nx * ny + (size_t)i * nx;
    const int64_t NONE = INT64_MAX;
    int64_t best = NONE, bq = INT64_MAX;
    auto consider = [&](int k, int64_t dx2) {
        if (row[k] < 0) return;
        const int64_t v = dx2 + (int64_t)row[k] * row[k], q = (int64_t)zr[k] * nx + k;
        if (v < best || (v == best && q < bq)) {
            best = v;
            bq = q;
        }
    };
    consider(j, 0);
    for (int r = 1; r < nx && (int64_t)r * r <= best; r++) {
        const int64_t r2 = (int64_t)r * r;
        if (j - r >= 0) consider(j - r, r2);
        if (j + r < nx) consider(j + r, r2);
    }
    k2[(size_t)z * nx * ny + p] = best == NONE ? -1 : (int)best;
    if (z == 1) nearest[p] = best == NONE ? -1 : (int)bq;
}

// OpenCV's float value of a squared distance: sqrtf of the exact integer;
// 1e15f (its infinity) under the root when the image has no zero pixel.
// The double root rounded to float is the correctly rounded float root.
__device__ __forceinline__ float dist_value(int k2) {
    return k2 < 0 ? (float)__builtin_sqrt((double)1e15f) : (float)__builtin_sqrt((double)k2);
}

// Integer square roots of t >= 0 (t < 2^31): floor and ceiling.
__device__ __forceinline__ int64_t isqrt_floor(int64_t t) {
    int64_t x = (int64_t)__builtin_sqrtf((float)t);
    while (x * x > t) x--;
    while ((x + 1) * (x + 1) <= t) x++;
    return x;
}
__device__ __forceinline__ int64_t isqrt_ceil(int64_t t) {
    const int64_t f = isqrt_floor(t);
    return f * f == t ? f : f + 1;
}

// LocPore and Kn of fluid pixel p.  The window of utils.hpp:118-131 is
// searched row by row in the reference's order.  A column can pass the
// reference's test |sqrt(dy^2 + dx^2) - D| < 1e-4 only if the integer
// s = dy^2 + dx^2 lies in [sLo, sHi] (the test's interval squared, widened
// by one): in a row only the columns with dx^2 in [sLo - dy^2, sHi - dy^2]
// are tested, in ascending order, with the reference's own double test.
__global__ void k_props(const uint8_t* __restrict__ dom, const uint8_t* __restrict__ skel,
                        const int* __restrict__ kw, const int* __restrict__ ks, const int* __restrict__ nearest, int nx,
                        int ny, double resolution, double mfp, double* __restrict__ locpore, double* __restrict__ kn,
                        int* __restrict__ err) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)nx * ny) return;
    if (dom[p] != 1) {
        locpore[p] = 0.0;
        kn[p] = 0.0;
        return;
    }
    const int i = (int)(p / nx), j = (int)(p % nx);
    if (ks[p] < 0) {
        atomicOr(err, ERR_NO_SKEL);
        locpore[p] = 0.0;
        kn[p] = 0.0;
        return;
    }
    const double D = (double)dist_value(ks[p]);
    // Below 1024 the float value of D is within 6.1e-5 of sqrt(ks) and the
    // next integer square is over 4.3e-4 away, so the reference's test passes
    // exactly the skeleton pixels at squared distance ks: the first of them
    // in row-major order is the distance pass's nearest.  It is checked with
    // the reference's own test; the window search covers the rest.
    if (D < 1024.0) {
        const int q = nearest[p];
        if (q >= 0) {
            const double dy = (double)(i - q / nx), dx = (double)(j - q % nx);
            if (fabs(__builtin_sqrt(dy * dy + dx * dx) - D) < 1.0e-4 && skel[q]) {
                const double lp = 2.0 * (double)dist_value(kw[q]);
                locpore[p] = lp;
                kn[p] = mfp / lp / resolution;
                return;
            }
        }
    }
    const int64_t r0 = (int64_t)(i - D - 2), c0 = (int64_t)(j - D - 2);
    const double rEnd = i + D + 2, cEnd = j + D + 2;
    const double lo = D - 1.0e-4, hi = D + 1.0e-4;
    const int64_t sLo = lo > 0 ? (int64_t)(lo * lo) - 1 : 0;
    const int64_t sHi = (int64_t)(hi * hi) + 2;
    int64_t q = -1;
    for (int64_t r = max(r0, (int64_t)0); r < ny && (double)r < rEnd && q < 0; r++) {
        const int64_t dy2 = (int64_t)(i - r) * (i - r);
        if (dy2 > sHi) continue;
        const int64_t dxLo = isqrt_ceil(max(sLo - dy2, (int64_t)0)), dxHi = isqrt_floor(sHi - dy2);
        if (dxLo > dxHi) continue;
        const double dy = (double)(i - r);
        for (int side = 0; side < 2 && q < 0; side++) {
            // side 0: columns j - dxHi .. j - dxLo; side 1: j + dx, dx > 0
            const int64_t a = side == 0 ? dxHi : max(dxLo, (int64_t)1);
            const int64_t n = side == 0 ? dxHi - dxLo + 1 : dxHi - a + 1;
            for (int64_t k = 0; k < n; k++) {
                const int64_t dx = side == 0 ? a - k : a + k;
                const int64_t c = side == 0 ? j - dx : j + dx;
                if (c < c0 || c < 0 || c >= nx || !((double)c < cEnd)) continue;
                const double ddx = (double)(j - c);
                const double dist = __builtin_sqrt(dy * dy + ddx * ddx);
                if (fabs(dist - D) < 1.0e-4 && skel[(size_t)r * nx + c]) {
                    q = r * nx + c;
                    break;
                }
            }
        }
    }
    if (q < 0) {
        atomicOr(err, ERR_NOT_FOUND);
        locpore[p] = 0.0;
        kn[p] = 0.0;
        return;
    }
    const double lp = 2.0 * (double)dist_value(kw[q]);
    locpore[p] = lp;
    kn[p] = mfp / lp / resolution;
}

// Fills n bytes with v (the flags and counters of a run are cleared by a
// kernel rather than a memset node).
__global__ void k_fill8(uint8_t* __restrict__ p, int64_t n, uint8_t v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

int grid_of(int64_t n, int b) { return (int)((n + b - 1) / b); }

// Work buffers of one run for sides nx x ny, as byte offsets into one
// region (each 256-byte aligned): the handle's buffers in the order
// mts_shape_create allocates them.
struct ShapeLayout {
    int m, H, Wc, WW, nseg, tilesX, tilesY;
    size_t cv0, cv1, skel, first, last, g, k2, zrow, nearest, cnt, chg, err, total;
};
constexpr size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
constexpr ShapeLayout shape_layout(int nx, int ny) {
    ShapeLayout L{};
    L.m = nx > ny ? nx : ny;
    L.H = ny + 2 * L.m;
    L.Wc = nx + 2 * L.m;
    L.WW = (L.Wc + 63) / 64;
    L.nseg = (ny + SEG - 1) / SEG;
    L.tilesX = (L.WW + THIN_TW - 1) / THIN_TW;
    L.tilesY = (L.H + THIN_TH - 1) / THIN_TH;
    const size_t np = (size_t)nx * ny, nw = (size_t)L.H * L.WW;
    size_t o = 0;
    L.cv0 = o, o = align256(o + nw * sizeof(u64));
    L.cv1 = o, o = align256(o + nw * sizeof(u64));
    L.skel = o, o = align256(o + np);
    L.first = o, o = align256(o + 2 * (size_t)L.nseg * nx * sizeof(int));
    L.last = o, o = align256(o + 2 * (size_t)L.nseg * nx * sizeof(int));
    L.g = o, o = align256(o + 2 * np * sizeof(int));
    L.k2 = o, o = align256(o + 2 * np * sizeof(int));
    L.zrow = o, o = align256(o + 2 * np * sizeof(int));
    L.nearest = o, o = align256(o + np * sizeof(int));
    L.cnt = o, o = align256(o + CHECK_EVERY * sizeof(unsigned));
    L.chg = o, o = align256(o + 2 * (size_t)L.tilesX * L.tilesY);
    L.err = o, o = align256(o + sizeof(int));
    L.total = o;
    return L;
}

}  // namespace

// ---------------------------------------------------------------------------
// Work memory of mts_shape_run_static: one file-scope device array per GPU,
// sized for the largest sides that entry point accepts.  Host code never
// uses the name `scratch` as an address: the device address of each GPU's
// copy comes from hipGetSymbolAddress, once per device, and is kept below.
constexpr int SHAPE_STATIC_MAX_SIDE = 512;
constexpr size_t SHAPE_SCRATCH_BYTES = shape_layout(SHAPE_STATIC_MAX_SIDE, SHAPE_STATIC_MAX_SIDE).total;
static __device__ __align__(256) unsigned char scratch[SHAPE_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of the current device's scratch array.
int scratch_of_current_device(unsigned char** out) {
    int dev = 0;
    SH_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(MTS_ERR_ARG, "mts_shape_run_static: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(MTS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return MTS_OK;
}
}  // namespace

struct mts_shape_s {
    int nx = 0, ny = 0, m = 0, H = 0, Wc = 0, WW = 0, nseg = 0;
    hipStream_t stream = nullptr;
    u64* cv[2] = {nullptr, nullptr};
    uint8_t* skel = nullptr;
    int* first = nullptr;
    int* last = nullptr;
    int* g = nullptr;
    int* zrow = nullptr;
    int* k2 = nullptr;
    int* nearest = nullptr;
    unsigned* cnt = nullptr;
    uint8_t* chg = nullptr;  // two arrays of per-tile change flags
    int tilesX = 0, tilesY = 0;
    int* err = nullptr;
    bool own_stream = true;  // false: the caller's stream (mts_shape_run_static)
    std::vector<void*> allocs;
    int64_t launches = 0;  // thinning launches of the last run

    template <class T>
    int alloc(T** p, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();  // a failed allocation must not surface at a later launch check
            return set_error(MTS_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        allocs.push_back(q);
        *p = (T*)q;
        return MTS_OK;
    }
    ~mts_shape_s() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* q : allocs) (void)hipFree(q);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

int mts_shape_create(int nx, int ny, mts_shape* out) {
    if (!out) return set_error(MTS_ERR_ARG, "mts_shape_create: null output");
    *out = nullptr;
    if (nx < 2 || ny < 2 || nx > MAX_SIDE || ny > MAX_SIDE)
        return set_error(MTS_ERR_ARG, "mts_shape_create: sides must lie in [2, " + std::to_string(MAX_SIDE) + "]");
    mts_shape_s* h = new mts_shape_s();
    h->nx = nx;
    h->ny = ny;
    h->m = std::max(nx, ny);
    h->H = ny + 2 * h->m;
    h->Wc = nx + 2 * h->m;
    h->WW = (h->Wc + 63) / 64;
    h->nseg = (ny + SEG - 1) / SEG;
    h->tilesX = (h->WW + THIN_TW - 1) / THIN_TW;
    h->tilesY = (h->H + THIN_TH - 1) / THIN_TH;
    const size_t np = (size_t)nx * ny, nw = (size_t)h->H * h->WW;
    int rc = MTS_OK;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = set_error(MTS_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    if (rc == MTS_OK) rc = h->alloc(&h->cv[0], nw);
    if (rc == MTS_OK) rc = h->alloc(&h->cv[1], nw);
    if (rc == MTS_OK) rc = h->alloc(&h->skel, np);
    if (rc == MTS_OK) rc = h->alloc(&h->first, 2 * (size_t)h->nseg * nx);
    if (rc == MTS_OK) rc = h->alloc(&h->last, 2 * (size_t)h->nseg * nx);
    if (rc == MTS_OK) rc = h->alloc(&h->g, 2 * np);
    if (rc == MTS_OK) rc = h->alloc(&h->k2, 2 * np);
    if (rc == MTS_OK) rc = h->alloc(&h->zrow, 2 * np);
    if (rc == MTS_OK) rc = h->alloc(&h->nearest, np);
    if (rc == MTS_OK) rc = h->alloc(&h->cnt, CHECK_EVERY);
    if (rc == MTS_OK) rc = h->alloc(&h->chg, 2 * (size_t)h->tilesX * h->tilesY);
    if (rc == MTS_OK) rc = h->alloc(&h->err, 1);
    if (rc != MTS_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return MTS_OK;
}

int mts_shape_destroy(mts_shape h) {
    delete h;
    return MTS_OK;
}

int mts_shape_stream(mts_shape h, void** stream) {
    if (!h || !stream) return set_error(MTS_ERR_ARG, "mts_shape_stream: null argument");
    *stream = h->stream;
    return MTS_OK;
}

int mts_shape_run(mts_shape h, const uint8_t* domain, double resolution, double mfp, double* locpore, double* kn,
                  uint8_t* skeleton, int* iterations) {
    if (!h || !domain || !locpore || !kn) return set_error(MTS_ERR_ARG, "mts_shape_run: null argument");
    if (!(resolution > 0) || !(mfp > 0)) return set_error(MTS_ERR_ARG, "mts_shape_run: resolution and mfp must be > 0");
    const int nx = h->nx, ny = h->ny;
    const int64_t np = (int64_t)nx * ny, nw = (int64_t)h->H * h->WW;
    hipStream_t s = h->stream;
    k_fill8<<<1, 64, 0, s>>>((uint8_t*)h->err, sizeof(int), 0);
    k_canvas<<<grid_of(nw, 256), 256, 0, s>>>(domain, nx, ny, h->m, h->H, h->Wc, h->WW, h->cv[0], h->err);
    SH_TRY(hipGetLastError());
    int herr = 0;
    SH_TRY(hipMemcpyAsync(&herr, h->err, sizeof(int), hipMemcpyDeviceToHost, s));
    SH_TRY(hipStreamSynchronize(s));
    if (herr & ERR_VALUE) return set_error(MTS_ERR_ARG, "mts_shape_run: domain values must be 0 or 1");
    // thinning: launches of THIN_S subiterations until one changes nothing;
    // every later launch would leave the image as it is
    const int ntiles = h->tilesX * h->tilesY;
    const dim3 tg(h->tilesX, h->tilesY);
    const int64_t maxLaunches = nw * 64 / (THIN_S / 2) + 2;  // each iteration but the last deletes a pixel
    k_fill8<<<grid_of(ntiles, 256), 256, 0, s>>>(h->chg, ntiles, 1);  // every tile is live in the first launch
    int cur = 0;
    bool done = false;
    h->launches = 0;
    while (!done) {
        if (h->launches >= maxLaunches) return set_error(MTS_ERR_STATE, "mts_shape_run: thinning did not settle");
        k_fill8<<<1, 64, 0, s>>>((uint8_t*)h->cnt, CHECK_EVERY * sizeof(unsigned), 0);
        for (int k = 0; k < CHECK_EVERY; k++) {
            const int64_t l = h->launches + k;
            k_thin<<<tg, THIN_THREADS, 0, s>>>(h->cv[cur], h->cv[cur ^ 1], h->H, h->Wc, h->WW, h->tilesX, h->tilesY,
                                               h->chg + (l & 1) * ntiles, h->chg + ((l & 1) ^ 1) * ntiles,
                                               h->cnt + k);
            cur ^= 1;
        }
        SH_TRY(hipGetLastError());
        h->launches += CHECK_EVERY;
        unsigned c[CHECK_EVERY];
        SH_TRY(hipMemcpyAsync(c, h->cnt, sizeof(c), hipMemcpyDeviceToHost, s));
        SH_TRY(hipStreamSynchronize(s));
        for (int k = 0; k < CHECK_EVERY; k++) done |= c[k] == 0;
    }
    if (iterations) *iterations = (int)std::min<int64_t>(h->launches * (THIN_S / 2), INT32_MAX);
    uint8_t* sk = skeleton ? skeleton : h->skel;
    k_crop<<<grid_of(np, 256), 256, 0, s>>>(h->cv[cur], h->WW, h->m, nx, ny, sk);
    k_colseg<<<dim3(grid_of((int64_t)h->nseg * nx, 256), 2), 256, 0, s>>>(domain, sk, nx, ny, h->nseg, h->first,
                                                                         h->last);
    k_colpass<<<dim3(grid_of(np, 256), 2), 256, 0, s>>>(domain, sk, nx, ny, h->nseg, h->first, h->last, h->g,
                                                        h->zrow);
    k_rowpass<<<dim3(grid_of(np, 256), 2), 256, 0, s>>>(h->g, h->zrow, nx, ny, h->k2, h->nearest);
    k_props<<<grid_of(np, 256), 256, 0, s>>>(domain, sk, h->k2, h->k2 + np, h->nearest, nx, ny, resolution, mfp,
                                            locpore, kn, h->err);
    SH_TRY(hipGetLastError());
    SH_TRY(hipMemcpyAsync(&herr, h->err, sizeof(int), hipMemcpyDeviceToHost, s));
    SH_TRY(hipStreamSynchronize(s));
    if (herr & ERR_NO_SKEL) return set_error(MTS_ERR_GEOMETRY, "mts_shape_run: the skeleton has no pixel inside the domain");
    if (herr & ERR_NOT_FOUND)
        return set_error(MTS_ERR_GEOMETRY, "mts_shape_run: no skeleton pixel at the skeleton distance");
    return MTS_OK;
}

int mts_shape_properties(const int32_t* domain, int nx, int ny, double resolution, double mfp, double* locpore,
                         double* kn, uint8_t* skeleton) {
    if (!domain || !locpore || !kn) return set_error(MTS_ERR_ARG, "mts_shape_properties: null argument");
    if (nx < 2 || ny < 2 || nx > MAX_SIDE || ny > MAX_SIDE)
        return set_error(MTS_ERR_ARG, "mts_shape_properties: sides must lie in [2, " + std::to_string(MAX_SIDE) + "]");
    const size_t np = (size_t)nx * ny;
    std::vector<uint8_t> d(np);
    for (size_t p = 0; p < np; p++) {
        if (domain[p] != 0 && domain[p] != 1) return set_error(MTS_ERR_ARG, "mts_shape_properties: domain values must be 0 or 1");
        d[p] = (uint8_t)domain[p];
    }
    mts_shape h = nullptr;
    int rc = mts_shape_create(nx, ny, &h);
    if (rc != MTS_OK) return rc;
    uint8_t* dd = nullptr;
    double* dl = nullptr;
    double* dk = nullptr;
    uint8_t* ds = nullptr;
    rc = h->alloc(&dd, np);
    if (rc == MTS_OK) rc = h->alloc(&dl, np);
    if (rc == MTS_OK) rc = h->alloc(&dk, np);
    if (rc == MTS_OK) rc = h->alloc(&ds, np);
    hipError_t e = hipSuccess;
    if (rc == MTS_OK) e = hipMemcpy(dd, d.data(), np, hipMemcpyHostToDevice);
    if (rc == MTS_OK && e != hipSuccess) rc = set_error(MTS_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    if (rc == MTS_OK) rc = mts_shape_run(h, dd, resolution, mfp, dl, dk, ds, nullptr);
    if (rc == MTS_OK) e = hipMemcpy(locpore, dl, np * sizeof(double), hipMemcpyDeviceToHost);
    if (rc == MTS_OK && e == hipSuccess) e = hipMemcpy(kn, dk, np * sizeof(double), hipMemcpyDeviceToHost);
    if (rc == MTS_OK && e == hipSuccess && skeleton) e = hipMemcpy(skeleton, ds, np, hipMemcpyDeviceToHost);
    if (rc == MTS_OK && e != hipSuccess) rc = set_error(MTS_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    mts_shape_destroy(h);
    return rc;
}

int mts_shape_run_static(int nx, int ny, const uint8_t* domain, double resolution, double mfp, double* locpore,
                         double* kn, uint8_t* skeleton, int* iterations, void* stream) {
    if (!domain || !locpore || !kn) return set_error(MTS_ERR_ARG, "mts_shape_run_static: null argument");
    if (nx < 2 || ny < 2 || nx > SHAPE_STATIC_MAX_SIDE || ny > SHAPE_STATIC_MAX_SIDE)
        return set_error(MTS_ERR_ARG, "mts_shape_run_static: sides must lie in [2, " +
                                          std::to_string(SHAPE_STATIC_MAX_SIDE) + "] (the static work memory holds " +
                                          std::to_string(SHAPE_SCRATCH_BYTES) + " bytes)");
    const ShapeLayout L = shape_layout(nx, ny);
    if (L.total > SHAPE_SCRATCH_BYTES) return set_error(MTS_ERR_ARG, "mts_shape_run_static: sides too large");
    unsigned char* base = nullptr;
    if (int rc = scratch_of_current_device(&base)) return rc;
    // a handle view over the scratch array: no allocation, the caller's stream
    mts_shape_s h;
    h.own_stream = false;
    h.stream = (hipStream_t)stream;
    h.nx = nx;
    h.ny = ny;
    h.m = L.m;
    h.H = L.H;
    h.Wc = L.Wc;
    h.WW = L.WW;
    h.nseg = L.nseg;
    h.tilesX = L.tilesX;
    h.tilesY = L.tilesY;
    h.cv[0] = (u64*)(base + L.cv0);
    h.cv[1] = (u64*)(base + L.cv1);
    h.skel = base + L.skel;
    h.first = (int*)(base + L.first);
    h.last = (int*)(base + L.last);
    h.g = (int*)(base + L.g);
    h.k2 = (int*)(base + L.k2);
    h.zrow = (int*)(base + L.zrow);
    h.nearest = (int*)(base + L.nearest);
    h.cnt = (unsigned*)(base + L.cnt);
    h.chg = base + L.chg;
    h.err = (int*)(base + L.err);
    return mts_shape_run(&h, domain, resolution, mfp, locpore, kn, skeleton, iterations);
}

}  // extern "C"
