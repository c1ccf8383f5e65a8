// Learned warm start: a four-level convolutional encoder-decoder (U-Net)
// that maps (fluid mask, log10 Kn, forcing) to the steady velocity field of a
// periodic porous lattice, with its own forward and backward kernels.
//
// Every convolution (3x3 and 1x1, periodic padding) is an implicit GEMM on
// BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation):
//   forward        Y[p, co]      = sum_{kk, ci} X[p + o(kk), ci] W[kk, co, ci]
//   backward data  dX[p, ci]     = sum_{kk, co} dY[p - o(kk), co] W[kk, co, ci]
//   backward wgt   dW[kk, co, ci] = sum_p dY[p, co] X[p + o(kk), ci]
// One kernel template runs all three: a 64x64 block tile, K in chunks of 32
// staged through LDS by per-element gather functors, four wave64s each on a
// 32x32 sub-tile (2x2 MFMA tiles).  The weight gradient splits K (pixels)
// over blocks and sums the partials in a fixed order.
// Activations are NHWC BF16; activation gradients FP32 (converted to BF16,
// with the ReLU mask applied, before they enter a GEMM); parameters FP32
// masters with BF16 copies; Adam in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mantiis_hip.h"

namespace mts {
int set_error(int code, const std::string& msg);  // mts_api.hip
}
using mts::set_error;

#define NET_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(e_ == hipErrorOutOfMemory ? MTS_ERR_OOM : MTS_ERR_HIP,        \
                             std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

namespace net {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 64, LDK = TK + 8;

__device__ __forceinline__ int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// ---------------------------------------------------------------- GEMM
// C[M x N] = A[M x K] B[K x N] over the K range of blockIdx.z (split-K).
// Lane layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l holds
// A[l % 16][8 (l / 16) .. +7], B[8 (l / 16) .. +7][l % 16] and
// D[4 (l / 16) .. +3][l % 16].
// A stage is TK = 64 deep: each thread holds two 8-element fragments of A and
// of B in registers.  The next stage's fragments are loaded from memory
// before the current stage's eight MFMAs per wave, and written to LDS after
// them, so the loads overlap the matrix work.
typedef uint4 v16;
struct Frag {
    v16 v;
};

template <class Op>
__global__ void __launch_bounds__(256) k_gemm(Op op, int M, int N, int K, int nsplit, int kchunk) {
    __shared__ bf16 sA[TM][LDK];
    __shared__ bf16 sB[TN][LDK];
    const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN, z = blockIdx.z;
    const int kb = (z % nsplit) * kchunk, ke = min(K, kb + kchunk);
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    Frag fa[2], fb[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        op.load_a(fa[h], m0, kb, ke, M, z, t, h);
        op.load_b(fb[h], n0, kb, ke, N, z, t, h);
    }
    for (int k0 = kb; k0 < ke; k0 += TK) {
        op.store_a(sA, fa, t);
        op.store_b(sB, fb, t);
        __syncthreads();
        if (k0 + TK < ke) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                op.load_a(fa[h], m0, k0 + TK, ke, M, z, t, h);
                op.load_b(fb[h], n0, k0 + TK, ke, N, z, t, h);
            }
        }
        const int i = l & 15, q = (l >> 4) * 8;
#pragma unroll
        for (int kh = 0; kh < 2; kh++) {
            bf16x8 af[2], bfm[2];
#pragma unroll
            for (int a = 0; a < 2; a++) af[a] = *(const bf16x8*)&sA[wm + a * 16 + i][32 * kh + q];
#pragma unroll
            for (int b = 0; b < 2; b++) bfm[b] = *(const bf16x8*)&sB[wn + b * 16 + i][32 * kh + q];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfm[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int m = m0 + wm + a * 16 + 4 * (l >> 4) + rr, n = n0 + wn + b * 16 + (l & 15);
                if (m < M && n < N) op.store(m, n, acc[a][b][rr], z);
            }
}

// Fragment loads and LDS stores.  Row layout: thread t takes tile row t/4
// and the 8 consecutive k from 8 (t%4) + 32 h; one 16-byte load when the
// operand keeps those 8 k contiguous (pa/pb), element loads otherwise.
// Transposed layout (operands whose 8 contiguous elements run along the
// tile's row index, e.g. 8 channels of one pixel): thread t takes rows
// 8 (t%8) .. +7 at k = 2 (t/8) + h, so its two fragments fill 4-byte pairs.
__device__ __forceinline__ void row_store(bf16 (*s)[LDK], const Frag* f, int t) {
#pragma unroll
    for (int h = 0; h < 2; h++) *(v16*)&s[t >> 2][(t & 3) * 8 + 32 * h] = f[h].v;
}
__device__ __forceinline__ void tr_store(bf16 (*s)[LDK], const Frag* f, int t) {
    const uint16_t* a = (const uint16_t*)&f[0].v;
    const uint16_t* b = (const uint16_t*)&f[1].v;
    const int gq = (t & 7) * 8, p = 2 * (t >> 3);
#pragma unroll
    for (int e = 0; e < 8; e++) *(uint32_t*)&s[gq + e][p] = (uint32_t)a[e] | ((uint32_t)b[e] << 16);
}
template <class Op>
__device__ __forceinline__ void row_load_a(const Op& op, Frag& f, int m0, int k0, int ke, int M, int z, int t, int h) {
    const int m = m0 + (t >> 2), k = k0 + (t & 3) * 8 + 32 * h;
    if (op.vec_a && m < M && k + 7 < ke) {
        f.v = *(const v16*)op.pa(m, k, z);
        return;
    }
    bf16* e8 = (bf16*)&f.v;
#pragma unroll
    for (int e = 0; e < 8; e++) e8[e] = (m < M && k + e < ke) ? op.a(m, k + e, z) : (bf16)0.0f;
}
template <class Op>
__device__ __forceinline__ void row_load_b(const Op& op, Frag& f, int n0, int k0, int ke, int N, int z, int t, int h) {
    const int n = n0 + (t >> 2), k = k0 + (t & 3) * 8 + 32 * h;
    if (op.vec_b && n < N && k + 7 < ke) {
        f.v = *(const v16*)op.pb(k, n, z);
        return;
    }
    bf16* e8 = (bf16*)&f.v;
#pragma unroll
    for (int e = 0; e < 8; e++) e8[e] = (n < N && k + e < ke) ? op.b(k + e, n, z) : (bf16)0.0f;
}

// row-layout operands (dense and convolution forward / backward-data)
#define MTS_ROW_FILLS                                                                                       \
    __device__ void load_a(Frag& f, int m0, int k0, int ke, int M, int z, int t, int h) const {             \
        row_load_a(*this, f, m0, k0, ke, M, z, t, h);                                                       \
    }                                                                                                       \
    __device__ void load_b(Frag& f, int n0, int k0, int ke, int N, int z, int t, int h) const {             \
        row_load_b(*this, f, n0, k0, ke, N, z, t, h);                                                       \
    }                                                                                                       \
    __device__ void store_a(bf16 (*s)[LDK], const Frag* f, int t) const { row_store(s, f, t); }             \
    __device__ void store_b(bf16 (*s)[LDK], const Frag* f, int t) const { row_store(s, f, t); }

// dense row-major operands (the GEMM check)
struct OpDense {
    const bf16 *A, *B;
    float* C;
    int K, N;
    static constexpr bool vec_a = false, vec_b = false;
    __device__ bf16 a(int m, int k, int) const { return A[(int64_t)m * K + k]; }
    __device__ bf16 b(int k, int n, int) const { return B[(int64_t)k * N + n]; }
    __device__ const bf16* pa(int, int, int) const { return nullptr; }
    __device__ const bf16* pb(int, int, int) const { return nullptr; }
    MTS_ROW_FILLS
    __device__ void store(int m, int n, float v, int) const { C[(int64_t)m * N + n] = v; }
};

struct Geom {
    int B, H, W;
    __device__ __forceinline__ void pix(int m, int& b, int& y, int& x) const {
        x = m % W;
        const int t = m / W;
        y = t % H;
        b = t / H;
    }
};

// kernel offsets: kk = 3 (dy + 1) + (dx + 1) for 3x3, kk = 0 for 1x1
__device__ __forceinline__ void koff(int kk, int ksz, int& dy, int& dx) {
    dy = ksz == 3 ? kk / 3 - 1 : 0;
    dx = ksz == 3 ? kk % 3 - 1 : 0;
}

struct OpConvFwd {
    Geom g;
    const bf16* X;   // [p][Cin]
    const bf16* Wt;  // [kk][Cout][Cin]
    const float* bias;
    bf16* Y;         // [p][Cout] (BF16, ReLU) or
    float* Yf;       // [p][Cout] (FP32, no ReLU)
    int Cin, Cout, ksz, relu;
    bool vec_a, vec_b;  // Cin % 8 == 0: 8 consecutive k are 8 channels of one kk
    __device__ const bf16* pa(int m, int k, int) const {
        const int kk = k / Cin, ci = k - kk * Cin;
        int b, y, x, dy, dx;
        g.pix(m, b, y, x);
        koff(kk, ksz, dy, dx);
        const int yy = wrapi(y + dy, g.H), xx = wrapi(x + dx, g.W);
        return X + (((int64_t)b * g.H + yy) * g.W + xx) * Cin + ci;
    }
    __device__ const bf16* pb(int k, int n, int) const {
        const int kk = k / Cin, ci = k - kk * Cin;
        return Wt + ((int64_t)kk * Cout + n) * Cin + ci;
    }
    __device__ bf16 a(int m, int k, int z) const { return *pa(m, k, z); }
    __device__ bf16 b(int k, int n, int z) const { return *pb(k, n, z); }
    MTS_ROW_FILLS
    __device__ void store(int m, int n, float v, int) const {
        v += bias[n];
        if (relu) v = fmaxf(v, 0.0f);
        if (Yf) Yf[(int64_t)m * Cout + n] = v;
        else Y[(int64_t)m * Cout + n] = (bf16)v;
    }
};

struct OpConvBwdData {
    Geom g;
    const bf16* dY;   // [p][Cout]
    const bf16* WtT;  // [kk][Cin][Cout] (transposed copy)
    float* dX;        // [p][Cin]
    int Cin, Cout, ksz, add;
    bool vec_a, vec_b;  // Cout % 8 == 0
    __device__ const bf16* pa(int m, int k, int) const {
        const int kk = k / Cout, co = k - kk * Cout;
        int b, y, x, dy, dx;
        g.pix(m, b, y, x);
        koff(kk, ksz, dy, dx);
        const int yy = wrapi(y - dy, g.H), xx = wrapi(x - dx, g.W);
        return dY + (((int64_t)b * g.H + yy) * g.W + xx) * Cout + co;
    }
    __device__ const bf16* pb(int k, int n, int) const {
        const int kk = k / Cout, co = k - kk * Cout;
        return WtT + ((int64_t)kk * Cin + n) * Cout + co;
    }
    __device__ bf16 a(int m, int k, int z) const { return *pa(m, k, z); }
    __device__ bf16 b(int k, int n, int z) const { return *pb(k, n, z); }
    MTS_ROW_FILLS
    __device__ void store(int m, int n, float v, int) const {
        float* o = dX + (int64_t)m * Cin + n;
        *o = add ? *o + v : v;
    }
};

// Weight gradient of a convolution as one GEMM per split of the pixels:
// dW[kk][co][ci] = sum_p dY[p][co] X[p shifted by kk][ci], with the kernel
// offsets kk along the columns (n = kk * Cin + ci), so a block's dY tile
// serves every offset of its columns.
struct OpConvBwdW {
    Geom g;
    const bf16* dY;  // [p][Cout]
    const bf16* X;   // [p][Cin]
    float* part;     // [z][Cout][nkk * Cin], z = split of the pixels
    int Cin, Cout, ksz, ncol;
    bool vec_a, vec_b;  // Cout % 8 == 0, Cin % 8 == 0
    __device__ int64_t xrow(int k, int kk) const {  // site of pixel k shifted by offset kk
        int b, y, x, dy, dx;
        g.pix(k, b, y, x);
        koff(kk, ksz, dy, dx);
        const int yy = wrapi(y + dy, g.H), xx = wrapi(x + dx, g.W);
        return ((int64_t)b * g.H + yy) * g.W + xx;
    }
    __device__ bf16 a(int m, int k, int) const { return dY[(int64_t)k * Cout + m]; }
    __device__ bf16 b(int k, int n, int) const {
        const int kk = n / Cin;
        return X[xrow(k, kk) * Cin + (n - kk * Cin)];
    }
    // transposed fragments: thread t loads the 8 channels 8 (t%8).. of pixel
    // k0 + 2 (t/8) + h
    __device__ void load_a(Frag& f, int m0, int k0, int ke, int M, int, int t, int h) const {
        const int m = m0 + (t & 7) * 8, k = k0 + 2 * (t >> 3) + h;
        if (vec_a && m + 7 < M && k < ke) {
            f.v = *(const v16*)(dY + (int64_t)k * Cout + m);
            return;
        }
        bf16* e8 = (bf16*)&f.v;
#pragma unroll
        for (int e = 0; e < 8; e++) e8[e] = (m + e < M && k < ke) ? dY[(int64_t)k * Cout + m + e] : (bf16)0.0f;
    }
    __device__ void load_b(Frag& f, int n0, int k0, int ke, int N, int z, int t, int h) const {
        const int n = n0 + (t & 7) * 8, k = k0 + 2 * (t >> 3) + h;
        bf16* e8 = (bf16*)&f.v;
        if (k >= ke) {
#pragma unroll
            for (int e = 0; e < 8; e++) e8[e] = (bf16)0.0f;
            return;
        }
        if (vec_b && n + 7 < N) {  // 8 channels of one offset (Cin % 8 == 0)
            const int kk = n / Cin;
            f.v = *(const v16*)(X + xrow(k, kk) * Cin + (n - kk * Cin));
            return;
        }
#pragma unroll
        for (int e = 0; e < 8; e++) e8[e] = n + e < N ? b(k, n + e, z) : (bf16)0.0f;
    }
    __device__ void store_a(bf16 (*s)[LDK], const Frag* f, int t) const { tr_store(s, f, t); }
    __device__ void store_b(bf16 (*s)[LDK], const Frag* f, int t) const { tr_store(s, f, t); }
    __device__ void store(int m, int n, float v, int z) const { part[((int64_t)z * Cout + m) * ncol + n] = v; }
};

// pixels per split of the weight-gradient GEMM (at most BW_SPLITS splits)
constexpr int BW_PIX = 2048, BW_SPLITS = 256;
inline int bw_nsplit(int64_t Pn) { return (int)std::max<int64_t>(1, std::min<int64_t>(BW_SPLITS, (Pn + BW_PIX - 1) / BW_PIX)); }

// ---------------------------------------------------------------- small kernels
// FP32 sum over the 64 lanes of a wave (butterfly: the same order on every
// lane and every run)
__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// dW[kk][co][ci] = the sum over the splits of part[z][co][kk Cin + ci]: one
// wave per weight, lane l taking the splits l, l + 64, ..., then the wave's
// butterfly
__global__ void __launch_bounds__(256) k_sum_parts(const float* part, float* dW, int nkk, int Cout, int Cin,
                                                   int nsplit) {
    const int64_t per = (int64_t)Cout * Cin;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int l = threadIdx.x & 63;
    if (i >= nkk * per) return;  // whole waves
    const int kk = (int)(i / per);
    const int64_t e = i - kk * per;
    const int co = (int)(e / Cin), ci = (int)(e - (int64_t)co * Cin);
    const int64_t ncol = (int64_t)nkk * Cin, off = (int64_t)co * ncol + (int64_t)kk * Cin + ci;
    float s = 0.0f;
    for (int q = l; q < nsplit; q += 64) s += part[(int64_t)q * Cout * ncol + off];
    s = wave_sumf(s);
    if (l == 0) dW[i] = s;
}

// bias gradient db[c] = sum_p g[p][c] in two passes with a fixed order:
// per chunk of BG_CHUNK pixels, CW channels x 256/CW pixel lanes per block
// (each pixel row of CW channels read as one run), the lanes combined in
// order; then one block per channel adds the chunks by a fixed tree.
constexpr int BG_CHUNK = 1024;
template <int CW>
__global__ void __launch_bounds__(256) k_bias_part(const bf16* g, int64_t P, int C, float* part) {
    constexpr int PL = 256 / CW;
    __shared__ float s[PL][CW];
    const int cl = threadIdx.x % CW, pl = threadIdx.x / CW;
    const int c = blockIdx.y * CW + cl;
    const int64_t p0 = (int64_t)blockIdx.x * BG_CHUNK, p1 = min(P, p0 + BG_CHUNK);
    float v = 0.0f;
    if (c < C)
        for (int64_t p = p0 + pl; p < p1; p += PL) v += (float)g[p * C + c];
    s[pl][cl] = v;
    __syncthreads();
    if (pl == 0 && c < C) {
        float a = 0.0f;
#pragma unroll
        for (int q = 0; q < PL; q++) a += s[q][cl];
        part[(int64_t)blockIdx.x * C + c] = a;
    }
}
__global__ void __launch_bounds__(256) k_bias_sum(const float* part, int nchunk, int C, float* db) {
    __shared__ float s[256];
    const int c = blockIdx.x, t = threadIdx.x;
    float v = 0.0f;
    for (int b = t; b < nchunk; b += 256) v += part[(int64_t)b * C + c];
    s[t] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    if (t == 0) db[c] = s[0];
}

// g (FP32) -> BF16 operand, times the ReLU mask of the layer output y
__global__ void k_grad_prep(const float* g, const bf16* y, bf16* out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (bf16)((y && (float)y[i] <= 0.0f) ? 0.0f : g[i]);
}

__global__ void k_f2bf(const float* in, bf16* out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (bf16)in[i];
}

// 2x2 average pool (NHWC): out [B][H/2][W/2][C]
__global__ void k_pool(const bf16* in, bf16* out, int B, int H, int W, int C) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = H / 2, w = W / 2;
    if (i >= (int64_t)B * h * w * C) return;
    const int c = (int)(i % C);
    int64_t t = i / C;
    const int x = (int)(t % w);
    t /= w;
    const int y = (int)(t % h);
    const int b = (int)(t / h);
    auto at = [&](int yy, int xx) { return (float)in[(((int64_t)b * H + yy) * W + xx) * C + c]; };
    out[i] = (bf16)(0.25f * (at(2 * y, 2 * x) + at(2 * y, 2 * x + 1) + at(2 * y + 1, 2 * x) + at(2 * y + 1, 2 * x + 1)));
}
// its gradient, added into gin [B][H][W][C]
__global__ void k_pool_bwd(const float* gout, float* gin, int B, int H, int W, int C) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * H * W * C) return;
    const int c = (int)(i % C);
    int64_t t = i / C;
    const int x = (int)(t % W);
    t /= W;
    const int y = (int)(t % H);
    const int b = (int)(t / H);
    gin[i] += 0.25f * gout[(((int64_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
}

// nearest 2x upsample of lo [B][H/2][W/2][C1] next to skip [B][H][W][C2]:
// out [B][H][W][C1 + C2]
__global__ void k_upcat(const bf16* lo, int C1, const bf16* skip, int C2, bf16* out, int B, int H, int W) {
    const int C = C1 + C2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * H * W * C) return;
    const int c = (int)(i % C);
    const int64_t p = i / C;
    const int x = (int)(p % W);
    const int64_t t = p / W;
    const int y = (int)(t % H);
    const int b = (int)(t / H);
    out[i] = c < C1 ? lo[(((int64_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C1 + c] : skip[p * C2 + (c - C1)];
}
// its gradient: glo [B][H/2][W/2][C1] = sum of the four children (overwritten),
// gskip [B][H][W][C2] = the skip part (overwritten)
__global__ void k_upcat_bwd_lo(const float* g, int C1, int C2, float* glo, int B, int H, int W) {
    const int C = C1 + C2, h = H / 2, w = W / 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)B * h * w * C1) return;
    const int c = (int)(i % C1);
    int64_t t = i / C1;
    const int x = (int)(t % w);
    t /= w;
    const int y = (int)(t % h);
    const int b = (int)(t / h);
    float s = 0.0f;
    for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) s += g[(((int64_t)b * H + 2 * y + dy) * W + 2 * x + dx) * C + c];
    glo[i] = s;
}
__global__ void k_upcat_bwd_skip(const float* g, int C1, int C2, float* gskip, int64_t P) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P * C2) return;
    const int64_t p = i / C2;
    const int c = (int)(i - p * C2);
    gskip[i] = g[p * (C1 + C2) + C1 + c];
}

// masked mean-square loss over B*H*W pixels of 2 channels:
// loss = sum m (o - t)^2 / (2 sum m); grad = m (o - t) / sum m.
// Pass 1: per chunk of LOSS_CHUNK pixels, (sum m (o - t)^2, sum m) by a
// fixed tree; pass 2 (one thread): the chunks in order; pass 3: the gradient.
constexpr int LOSS_CHUNK = 4096;
__global__ void __launch_bounds__(256) k_loss_part(const float* o, const float* t, const float* mask, int64_t P,
                                                   float* part) {
    __shared__ float s[2][256];
    const int tid = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * LOSS_CHUNK, p1 = min(P, p0 + LOSS_CHUNK);
    float se = 0.0f, sm = 0.0f;
    for (int64_t p = p0 + tid; p < p1; p += 256) {
        const float m = mask[p];
        const float d0 = o[2 * p] - t[2 * p], d1 = o[2 * p + 1] - t[2 * p + 1];
        se += m * (d0 * d0 + d1 * d1);
        sm += m;
    }
    s[0][tid] = se;
    s[1][tid] = sm;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            s[0][tid] += s[0][tid + h];
            s[1][tid] += s[1][tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        part[2 * blockIdx.x] = s[0][0];
        part[2 * blockIdx.x + 1] = s[1][0];
    }
}
__global__ void k_loss_sum(const float* part, int nchunk, float* out2) {
    float se = 0.0f, sm = 0.0f;
    for (int b = 0; b < nchunk; b++) {
        se += part[2 * b];
        sm += part[2 * b + 1];
    }
    out2[0] = se / (2.0f * fmaxf(sm, 1.0f));
    out2[1] = sm;
}
__global__ void k_loss_grad(const float* o, const float* t, const float* mask, int64_t P, const float* out2,
                            float* grad) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const float m = mask[p] / fmaxf(out2[1], 1.0f);
    grad[2 * p] = m * (o[2 * p] - t[2 * p]);
    grad[2 * p + 1] = m * (o[2 * p + 1] - t[2 * p + 1]);
}

__global__ void k_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copy of one layer's weights transposed to [kk][cin][cout]
__global__ void k_wtrans(const float* w, bf16* wt, int nkk, int cout, int cin) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)cout * cin;
    if (i >= nkk * per) return;
    const int kk = (int)(i / per);
    const int64_t e = i - kk * per;
    const int co = (int)(e / cin), ci = (int)(e - (int64_t)co * cin);
    wt[(int64_t)kk * per + (int64_t)ci * cout + co] = (bf16)w[i];
}

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

template <class Op>
void gemm(const Op& op, int M, int N, int K, int nz, int nsplit, int kchunk, hipStream_t st) {
    dim3 grid((M + TM - 1) / TM, (N + TN - 1) / TN, nz);
    hipLaunchKernelGGL(k_gemm<Op>, grid, dim3(256), 0, st, op, M, N, K, nsplit, kchunk);
}

}  // namespace net

using namespace net;

// ---------------------------------------------------------------- the network
struct ConvL {
    int cin, cout, ksz, relu;
    int64_t woff, boff;  // offsets in the parameter array
};
struct Tens {
    int C = 0, level = 0;  // level: resolution H >> level
    bf16* a = nullptr;     // activation
    float* g = nullptr;    // gradient
};

struct mts_net_s {
    hipStream_t stream = nullptr;
    int c[4];
    int maxB = 0, maxH = 0, maxW = 0;
    std::vector<ConvL> L;
    int64_t nparam = 0;
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    bf16* Pb = nullptr;   // BF16 copy, weights [kk][cout][cin]
    bf16* PbT = nullptr;  // BF16 copy, weights [kk][cin][cout]
    // tensors
    enum {
        X0, E0A, E0, P0, E1A, E1, P1, E2A, E2, P2, E3A, E3, C2, D2A, D2, C1, D1A, D1, C0, D0A, D0, NT
    };
    Tens T[NT];
    float* out = nullptr;   // [B][H][W][2]
    float* gout = nullptr;  // its gradient
    bf16* gb = nullptr;     // BF16 gradient operand scratch
    float* part = nullptr;  // split-K partials
    int64_t part_n = 0;
    float* loss = nullptr;  // 2 floats
    float* red = nullptr;   // partial sums of the bias gradients and the loss
    int64_t red_n = 0;
    int B = 0, H = 0, W = 0;
    std::vector<void*> allocs;
    template <class Tp>
    int alloc(Tp** p, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(Tp));
        if (e != hipSuccess) {
            (void)hipGetLastError();  // a failed allocation must not surface at a later launch check
            return set_error(MTS_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        allocs.push_back(q);
        *p = (Tp*)q;
        return MTS_OK;
    }
    ~mts_net_s() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* q : allocs) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int64_t npix(int level) const { return (int64_t)B * (H >> level) * (W >> level); }
    // BF16 copies of the FP32 parameters (both weight layouts)
    void refresh_bf16() {
        hipLaunchKernelGGL(k_f2bf, dim3(nblk(nparam)), dim3(256), 0, stream, P, Pb, nparam);
        for (const ConvL& l : L) {
            const int nkk = l.ksz * l.ksz;
            const int64_t nw = (int64_t)nkk * l.cout * l.cin;
            hipLaunchKernelGGL(k_wtrans, dim3(nblk(nw)), dim3(256), 0, stream, P + l.woff, PbT + l.woff, nkk, l.cout,
                               l.cin);
        }
    }
    Geom geom(int level) const { return Geom{B, H >> level, W >> level}; }

    void conv_fwd(int li, const Tens& in, Tens& out_t, float* outf) {
        const ConvL& l = L[li];
        OpConvFwd op{geom(in.level), in.a, Pb + l.woff, P + l.boff, out_t.a,  outf,
                     l.cin,          l.cout, l.ksz,       l.relu,      l.cin % 8 == 0, l.cin % 8 == 0};
        gemm(op, (int)npix(in.level), l.cout, l.cin * l.ksz * l.ksz, 1, 1, l.cin * l.ksz * l.ksz, stream);
    }
    // gradient of layer li given the FP32 gradient of its output (gy) and its
    // output activation (y, for the ReLU mask; nullptr for the head)
    void conv_bwd(int li, const Tens& in, const float* gy, const bf16* y, float* gin, int add) {
        const ConvL& l = L[li];
        const int64_t Pn = npix(in.level);
        const int64_t ny = Pn * l.cout;
        hipLaunchKernelGGL(k_grad_prep, dim3(nblk(ny)), dim3(256), 0, stream, gy, y, gb, ny);
        const int nbc = (int)((Pn + BG_CHUNK - 1) / BG_CHUNK);
        hipLaunchKernelGGL(k_bias_part<32>, dim3(nbc, (l.cout + 31) / 32), dim3(256), 0, stream, gb, Pn, l.cout, red);
        hipLaunchKernelGGL(k_bias_sum, dim3(l.cout), dim3(256), 0, stream, red, nbc, l.cout, G + l.boff);
        const int nkk = l.ksz * l.ksz;
        const int nsplit = bw_nsplit(Pn);
        const int kchunk = (int)(((Pn + nsplit - 1) / nsplit + TK - 1) / TK * TK);
        OpConvBwdW ow{geom(in.level), gb, in.a, part, l.cin, l.cout, l.ksz, nkk * l.cin, l.cout % 8 == 0, l.cin % 8 == 0};
        gemm(ow, l.cout, nkk * l.cin, (int)Pn, nsplit, nsplit, kchunk, stream);
        const int64_t per = (int64_t)l.cout * l.cin;
        hipLaunchKernelGGL(k_sum_parts, dim3(nblk(nkk * per * 64)), dim3(256), 0, stream, part, G + l.woff, nkk,
                           l.cout, l.cin, nsplit);
        if (gin) {
            OpConvBwdData od{geom(in.level), gb, PbT + l.woff, gin, l.cin, l.cout, l.ksz, add, l.cout % 8 == 0,
                             l.cout % 8 == 0};
            gemm(od, (int)Pn, l.cin, l.cout * nkk, 1, 1, l.cout * nkk, stream);
        }
    }
};

namespace {

int check_net(mts_net n) { return n ? MTS_OK : set_error(MTS_ERR_ARG, "null network handle"); }

// resolution level of each layer's input: two convolutions per encoder level
// 0..3, two per decoder level 2..0, then the 1x1 head at level 0
int layer_level(size_t i) {
    static const int level[15] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    return i < 15 ? level[i] : 0;
}

// layer table of the U-Net (widths c0..c3, input 3 channels, output 2)
void layers(mts_net_s* n) {
    const int* c = n->c;
    const int spec[15][4] = {{3, c[0], 3, 1},          {c[0], c[0], 3, 1},        {c[0], c[1], 3, 1},
                             {c[1], c[1], 3, 1},       {c[1], c[2], 3, 1},        {c[2], c[2], 3, 1},
                             {c[2], c[3], 3, 1},       {c[3], c[3], 3, 1},        {c[3] + c[2], c[2], 3, 1},
                             {c[2], c[2], 3, 1},       {c[2] + c[1], c[1], 3, 1}, {c[1], c[1], 3, 1},
                             {c[1] + c[0], c[0], 3, 1}, {c[0], c[0], 3, 1},       {c[0], 2, 1, 0}};
    int64_t off = 0;
    for (auto& s : spec) {
        ConvL l{s[0], s[1], s[2], s[3], 0, 0};
        l.woff = off;
        off += (int64_t)l.ksz * l.ksz * l.cout * l.cin;
        l.boff = off;
        off += l.cout;
        n->L.push_back(l);
    }
    n->nparam = off;
}

}  // namespace

extern "C" {

int mts_net_create(int c0, int max_batch, int max_h, int max_w, uint64_t seed, mts_net* out) {
    // c0 % 8 == 0 keeps every layer's weights 16-byte aligned for the vector loads
    if (!out || c0 <= 0 || c0 % 8 || max_batch <= 0 || max_h <= 0 || max_w <= 0 || max_h % 8 || max_w % 8)
        return set_error(MTS_ERR_ARG, "bad network shape (c0 and the sizes must be multiples of 8)");
    *out = nullptr;
    mts_net_s* n = new (std::nothrow) mts_net_s();
    if (!n) return set_error(MTS_ERR_OOM, "host allocation");
    auto bail = [&](int rc) {
        delete n;
        return rc;
    };
    n->c[0] = c0;
    n->c[1] = 2 * c0;
    n->c[2] = 4 * c0;
    n->c[3] = 8 * c0;
    n->maxB = max_batch;
    n->maxH = max_h;
    n->maxW = max_w;
    layers(n);
    hipError_t e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(set_error(MTS_ERR_HIP, std::string("stream: ") + hipGetErrorString(e)));
    int rc;
    const int64_t np = n->nparam;
    if ((rc = n->alloc(&n->P, np)) || (rc = n->alloc(&n->G, np)) || (rc = n->alloc(&n->Mo, np)) ||
        (rc = n->alloc(&n->Vo, np)) || (rc = n->alloc(&n->Pb, np)) || (rc = n->alloc(&n->PbT, np)))
        return bail(rc);
    // tensors: channels and level
    const int* c = n->c;
    const int tc[mts_net_s::NT][2] = {{3, 0},        {c[0], 0}, {c[0], 0},         {c[0], 1}, {c[1], 1},
                                      {c[1], 1},     {c[1], 2}, {c[2], 2},         {c[2], 2}, {c[2], 3},
                                      {c[3], 3},     {c[3], 3}, {c[3] + c[2], 2},  {c[2], 2}, {c[2], 2},
                                      {c[2] + c[1], 1}, {c[1], 1}, {c[1], 1},      {c[1] + c[0], 0}, {c[0], 0},
                                      {c[0], 0}};
    int64_t maxe = 0;
    for (int t = 0; t < mts_net_s::NT; t++) {
        Tens& T = n->T[t];
        T.C = tc[t][0];
        T.level = tc[t][1];
        const int64_t ne = (int64_t)max_batch * (max_h >> T.level) * (max_w >> T.level) * T.C;
        maxe = std::max(maxe, ne);
        if ((rc = n->alloc(&T.a, ne)) || (rc = n->alloc(&T.g, ne))) return bail(rc);
    }
    const int64_t P0 = (int64_t)max_batch * max_h * max_w;
    int64_t pmax = 0;
    for (size_t i = 0; i < n->L.size(); i++) {  // split-K partials at each layer's resolution
        const ConvL& l = n->L[i];
        const int nsplit = bw_nsplit(P0 >> (2 * layer_level(i)));
        pmax = std::max(pmax, (int64_t)l.ksz * l.ksz * nsplit * l.cout * l.cin);
    }
    n->part_n = pmax;
    // reduction partials: the bias gradient's chunks x channels, the loss's chunks x 2
    int cmax = 2;
    for (const ConvL& l : n->L) cmax = std::max(cmax, l.cout);
    n->red_n = std::max((P0 + BG_CHUNK - 1) / BG_CHUNK * cmax, (P0 + LOSS_CHUNK - 1) / LOSS_CHUNK * 2);
    if ((rc = n->alloc(&n->out, P0 * 2)) || (rc = n->alloc(&n->gout, P0 * 2)) || (rc = n->alloc(&n->gb, maxe)) ||
        (rc = n->alloc(&n->part, pmax)) || (rc = n->alloc(&n->loss, 2)) || (rc = n->alloc(&n->red, n->red_n)))
        return bail(rc);
    // He-normal weights, zero biases, zero Adam moments
    std::vector<float> h(np, 0.0f);
    std::mt19937_64 rng(seed);
    for (const ConvL& l : n->L) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)(l.cin * l.ksz * l.ksz)));
        const float sc = l.relu ? 1.0f : 0.1f;  // small head: the first predictions stay near zero
        for (int64_t i = 0; i < (int64_t)l.ksz * l.ksz * l.cout * l.cin; i++) h[l.woff + i] = sc * nd(rng);
    }
    if ((e = hipMemcpy(n->P, h.data(), np * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(n->Mo, 0, np * 4)) != hipSuccess || (e = hipMemset(n->Vo, 0, np * 4)) != hipSuccess ||
        (e = hipMemset(n->G, 0, np * 4)) != hipSuccess)
        return bail(set_error(MTS_ERR_HIP, std::string("init: ") + hipGetErrorString(e)));
    n->refresh_bf16();
    if ((e = hipStreamSynchronize(n->stream)) != hipSuccess)
        return bail(set_error(MTS_ERR_HIP, std::string("init: ") + hipGetErrorString(e)));
    *out = n;
    return MTS_OK;
}

int mts_net_destroy(mts_net n) {
    delete n;
    return MTS_OK;
}

int mts_net_flops(mts_net n, int B, int H, int W, double* fwd) {
    if (int rc = check_net(n)) return rc;
    if (!fwd || B <= 0 || H <= 0 || W <= 0) return set_error(MTS_ERR_ARG, "bad argument");
    double f = 0.0;
    for (size_t i = 0; i < n->L.size(); i++) {
        const ConvL& l = n->L[i];
        const int lv = layer_level(i);
        f += 2.0 * B * (double)(H >> lv) * (double)(W >> lv) * l.cin * l.cout * l.ksz * l.ksz;
    }
    *fwd = f;
    return MTS_OK;
}

int mts_net_param_count(mts_net n, int64_t* count) {
    if (int rc = check_net(n)) return rc;
    if (!count) return set_error(MTS_ERR_ARG, "null argument");
    *count = n->nparam;
    return MTS_OK;
}

int mts_net_get_params(mts_net n, float* host) {
    if (int rc = check_net(n)) return rc;
    if (!host) return set_error(MTS_ERR_ARG, "null argument");
    NET_TRY(hipMemcpyAsync(host, n->P, n->nparam * 4, hipMemcpyDeviceToHost, n->stream));
    NET_TRY(hipStreamSynchronize(n->stream));
    return MTS_OK;
}

int mts_net_set_params(mts_net n, const float* host) {
    if (int rc = check_net(n)) return rc;
    if (!host) return set_error(MTS_ERR_ARG, "null argument");
    NET_TRY(hipMemcpyAsync(n->P, host, n->nparam * 4, hipMemcpyHostToDevice, n->stream));
    n->refresh_bf16();
    NET_TRY(hipStreamSynchronize(n->stream));
    return MTS_OK;
}

int mts_net_get_grads(mts_net n, float* host) {
    if (int rc = check_net(n)) return rc;
    if (!host) return set_error(MTS_ERR_ARG, "null argument");
    NET_TRY(hipMemcpyAsync(host, n->G, n->nparam * 4, hipMemcpyDeviceToHost, n->stream));
    NET_TRY(hipStreamSynchronize(n->stream));
    return MTS_OK;
}

int mts_net_forward(mts_net n, const float* in, int B, int H, int W, float* out) {
    if (int rc = check_net(n)) return rc;
    if (!in || B <= 0 || B > n->maxB || H <= 0 || W <= 0 || H > n->maxH || W > n->maxW || H % 8 || W % 8 ||
        (int64_t)H * W * B > (int64_t)n->maxH * n->maxW * n->maxB)
        return set_error(MTS_ERR_ARG, "input shape exceeds the network's buffers or is not a multiple of 8");
    n->B = B;
    n->H = H;
    n->W = W;
    hipStream_t st = n->stream;
    auto& T = n->T;
    const int64_t p0 = n->npix(0);
    hipLaunchKernelGGL(k_f2bf, dim3(nblk(p0 * 3)), dim3(256), 0, st, in, T[mts_net_s::X0].a, p0 * 3);
    typedef mts_net_s S;
    auto pool = [&](int a, int b) {
        const Tens& x = T[a];
        hipLaunchKernelGGL(k_pool, dim3(nblk(n->npix(x.level + 1) * x.C)), dim3(256), 0, st, x.a, T[b].a, B,
                           H >> x.level, W >> x.level, x.C);
    };
    auto upcat = [&](int lo, int skip, int out_t) {
        const Tens& s = T[skip];
        hipLaunchKernelGGL(k_upcat, dim3(nblk(n->npix(s.level) * T[out_t].C)), dim3(256), 0, st, T[lo].a, T[lo].C, s.a,
                           s.C, T[out_t].a, B, H >> s.level, W >> s.level);
    };
    n->conv_fwd(0, T[S::X0], T[S::E0A], nullptr);
    n->conv_fwd(1, T[S::E0A], T[S::E0], nullptr);
    pool(S::E0, S::P0);
    n->conv_fwd(2, T[S::P0], T[S::E1A], nullptr);
    n->conv_fwd(3, T[S::E1A], T[S::E1], nullptr);
    pool(S::E1, S::P1);
    n->conv_fwd(4, T[S::P1], T[S::E2A], nullptr);
    n->conv_fwd(5, T[S::E2A], T[S::E2], nullptr);
    pool(S::E2, S::P2);
    n->conv_fwd(6, T[S::P2], T[S::E3A], nullptr);
    n->conv_fwd(7, T[S::E3A], T[S::E3], nullptr);
    upcat(S::E3, S::E2, S::C2);
    n->conv_fwd(8, T[S::C2], T[S::D2A], nullptr);
    n->conv_fwd(9, T[S::D2A], T[S::D2], nullptr);
    upcat(S::D2, S::E1, S::C1);
    n->conv_fwd(10, T[S::C1], T[S::D1A], nullptr);
    n->conv_fwd(11, T[S::D1A], T[S::D1], nullptr);
    upcat(S::D1, S::E0, S::C0);
    n->conv_fwd(12, T[S::C0], T[S::D0A], nullptr);
    n->conv_fwd(13, T[S::D0A], T[S::D0], nullptr);
    Tens dummy;
    n->conv_fwd(14, T[S::D0], dummy, n->out);
    if (out) NET_TRY(hipMemcpyAsync(out, n->out, p0 * 2 * 4, hipMemcpyDeviceToDevice, st));
    NET_TRY(hipGetLastError());
    NET_TRY(hipStreamSynchronize(st));
    return MTS_OK;
}

int mts_net_backward(mts_net n, const float* target, const float* mask, float* loss_host) {
    if (int rc = check_net(n)) return rc;
    if (!target || !mask || n->B == 0) return set_error(MTS_ERR_ARG, "run mts_net_forward first");
    hipStream_t st = n->stream;
    auto& T = n->T;
    typedef mts_net_s S;
    const int B = n->B, H = n->H, W = n->W;
    const int64_t p0 = n->npix(0);
    const int nlc = (int)((p0 + LOSS_CHUNK - 1) / LOSS_CHUNK);
    hipLaunchKernelGGL(k_loss_part, dim3(nlc), dim3(256), 0, st, n->out, target, mask, p0, n->red);
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(1), 0, st, n->red, nlc, n->loss);
    hipLaunchKernelGGL(k_loss_grad, dim3(nblk(p0)), dim3(256), 0, st, n->out, target, mask, p0, n->loss, n->gout);
    // head, then decoder, then encoder (reverse order of the forward pass)
    n->conv_bwd(14, T[S::D0], n->gout, nullptr, T[S::D0].g, 0);
    n->conv_bwd(13, T[S::D0A], T[S::D0].g, T[S::D0].a, T[S::D0A].g, 0);
    n->conv_bwd(12, T[S::C0], T[S::D0A].g, T[S::D0A].a, T[S::C0].g, 0);
    auto upcat_bwd = [&](int lo, int skip, int cat) {
        const Tens& s = T[skip];
        const int hh = H >> s.level, ww = W >> s.level;
        hipLaunchKernelGGL(k_upcat_bwd_lo, dim3(nblk(n->npix(s.level + 1) * T[lo].C)), dim3(256), 0, st, T[cat].g,
                           T[lo].C, s.C, T[lo].g, B, hh, ww);
        hipLaunchKernelGGL(k_upcat_bwd_skip, dim3(nblk(n->npix(s.level) * s.C)), dim3(256), 0, st, T[cat].g, T[lo].C,
                           s.C, T[skip].g, n->npix(s.level));
    };
    auto pool_bwd = [&](int pooled, int x) {
        const Tens& t = T[x];
        hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(n->npix(t.level) * t.C)), dim3(256), 0, st, T[pooled].g, t.g, B,
                           H >> t.level, W >> t.level, t.C);
    };
    upcat_bwd(S::D1, S::E0, S::C0);  // D1.g, E0.g (first part)
    n->conv_bwd(11, T[S::D1A], T[S::D1].g, T[S::D1].a, T[S::D1A].g, 0);
    n->conv_bwd(10, T[S::C1], T[S::D1A].g, T[S::D1A].a, T[S::C1].g, 0);
    upcat_bwd(S::D2, S::E1, S::C1);
    n->conv_bwd(9, T[S::D2A], T[S::D2].g, T[S::D2].a, T[S::D2A].g, 0);
    n->conv_bwd(8, T[S::C2], T[S::D2A].g, T[S::D2A].a, T[S::C2].g, 0);
    upcat_bwd(S::E3, S::E2, S::C2);
    n->conv_bwd(7, T[S::E3A], T[S::E3].g, T[S::E3].a, T[S::E3A].g, 0);
    n->conv_bwd(6, T[S::P2], T[S::E3A].g, T[S::E3A].a, T[S::P2].g, 0);
    pool_bwd(S::P2, S::E2);  // adds into E2.g
    n->conv_bwd(5, T[S::E2A], T[S::E2].g, T[S::E2].a, T[S::E2A].g, 0);
    n->conv_bwd(4, T[S::P1], T[S::E2A].g, T[S::E2A].a, T[S::P1].g, 0);
    pool_bwd(S::P1, S::E1);
    n->conv_bwd(3, T[S::E1A], T[S::E1].g, T[S::E1].a, T[S::E1A].g, 0);
    n->conv_bwd(2, T[S::P0], T[S::E1A].g, T[S::E1A].a, T[S::P0].g, 0);
    pool_bwd(S::P0, S::E0);
    n->conv_bwd(1, T[S::E0A], T[S::E0].g, T[S::E0].a, T[S::E0A].g, 0);
    n->conv_bwd(0, T[S::X0], T[S::E0A].g, T[S::E0A].a, nullptr, 0);
    NET_TRY(hipGetLastError());
    float l2[2];
    NET_TRY(hipMemcpyAsync(l2, n->loss, 8, hipMemcpyDeviceToHost, st));
    NET_TRY(hipStreamSynchronize(st));
    if (loss_host) *loss_host = l2[0];
    return MTS_OK;
}

int mts_net_adam(mts_net n, float lr, float beta1, float beta2, float eps, int step) {
    if (int rc = check_net(n)) return rc;
    if (step < 1) return set_error(MTS_ERR_ARG, "Adam step counts from 1");
    const float c1 = 1.0f - std::pow(beta1, (float)step), c2 = 1.0f - std::pow(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n->nparam)), dim3(256), 0, n->stream, n->P, n->G, n->Mo, n->Vo, n->nparam, lr,
                       beta1, beta2, eps, c1, c2);
    n->refresh_bf16();
    NET_TRY(hipGetLastError());
    NET_TRY(hipStreamSynchronize(n->stream));
    return MTS_OK;
}

// The GEMM kernel on dense host matrices (A: M x K, B: K x N, row-major,
// rounded to BF16 on the device): C = A B in FP32.
int mts_net_gemm_check(const float* A, const float* Bm, int M, int N, int K, float* C) {
    if (!A || !Bm || !C || M <= 0 || N <= 0 || K <= 0 || (int64_t)M * N > (1 << 24) || (int64_t)M * K > (1 << 24) ||
        (int64_t)K * N > (1 << 24))
        return set_error(MTS_ERR_ARG, "bad GEMM check shape");
    float *dA, *dB, *dC;
    bf16 *hA, *hB;
    NET_TRY(hipMalloc(&dA, (size_t)M * K * 4));
    NET_TRY(hipMalloc(&dB, (size_t)K * N * 4));
    NET_TRY(hipMalloc(&dC, (size_t)M * N * 4));
    NET_TRY(hipMalloc(&hA, (size_t)M * K * 2));
    NET_TRY(hipMalloc(&hB, (size_t)K * N * 2));
    NET_TRY(hipMemcpy(dA, A, (size_t)M * K * 4, hipMemcpyHostToDevice));
    NET_TRY(hipMemcpy(dB, Bm, (size_t)K * N * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_f2bf, dim3(nblk((int64_t)M * K)), dim3(256), 0, nullptr, dA, hA, (int64_t)M * K);
    hipLaunchKernelGGL(k_f2bf, dim3(nblk((int64_t)K * N)), dim3(256), 0, nullptr, dB, hB, (int64_t)K * N);
    OpDense op{hA, hB, dC, K, N};
    gemm(op, M, N, K, 1, 1, K, nullptr);
    NET_TRY(hipGetLastError());
    NET_TRY(hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost));
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dC);
    (void)hipFree(hA);
    (void)hipFree(hB);
    return MTS_OK;
}

int mts_net_stream(mts_net n, void** stream) {
    if (int rc = check_net(n)) return rc;
    if (!stream) return set_error(MTS_ERR_ARG, "null argument");
    *stream = (void*)n->stream;
    return MTS_OK;
}

}  // extern "C"
