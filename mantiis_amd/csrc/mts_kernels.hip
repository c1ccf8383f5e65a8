// HIP kernels of the D2Q9 step for gfx950 (MI355X).
//
// Layout in HBM (one lattice): a dense nx*ny site image, row-major
// (site p = y*nx + x), one SoA plane per direction:
//   pop[q*stride + p]   q = 0..8   (double, or float holding f - w_q*rho0)
//   meta[p]             uint32 link word (mts_common.h), 0 = solid
//   pidx[p]             uint16 index into the (tau_s, Kn) parameter table
// A wave64 touches 64 consecutive sites of one row per plane: every
// population load/store is a contiguous 512 B (FP64) or 256 B (FP32)
// segment, offset by one site for the x-neighbours.
//
// Time stepping uses the post-collision populations A as the carried state:
// the fused pass reads A_{n-1} by gathering population j of site x from
// x - c_j (streaming, bounce-back and periodic rules of the reference folded
// into that gather, no atomics), forms the post-stream f_n, its moments,
// and writes A_n = collide(f_n).  A call of K steps is one collide-only pass
// from the canonical FP64 f, K-1 fused passes and one stream-only pass back
// to the canonical f (so f equals the reference's vector between calls).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mantiis_hip.h"
#include <hip/hip_ext.h>

#include "mts_common.h"
#include "mts_kernels.h"
#include "mts_tables.inc"

namespace mts {
namespace dev {

// ---------------------------------------------------------------- constants
__device__ __forceinline__ constexpr int ex(int j) { return (j == 1 || j == 5 || j == 8) ? 1 : (j == 3 || j == 6 || j == 7) ? -1 : 0; }
__device__ __forceinline__ constexpr int ey(int j) { return (j == 2 || j == 5 || j == 6) ? 1 : (j == 4 || j == 7 || j == 8) ? -1 : 0; }
__device__ __forceinline__ constexpr int opp(int j) { return j == 0 ? 0 : (j <= 4 ? ((j + 1) % 4) + 1 : ((j - 5 + 2) % 4) + 5); }
__device__ __forceinline__ constexpr double wq(int j) { return j == 0 ? 4.0 / 9.0 : (j <= 4 ? 1.0 / 9.0 : 1.0 / 36.0); }
__device__ __forceinline__ constexpr double wf(int j) { return j == 0 ? 0.0 : (j <= 4 ? 1.0 / 3.0 : 1.0 / 12.0); }

// MRT basis (globalParameters.hpp:154-163) and the 15-decimal inverse the
// reference uses (globalParameters.hpp:166-175).
__device__ __forceinline__ constexpr double Mc(int k, int v) {
    constexpr double M[81] = {
        1., 1., 1., 1., 1., 1., 1., 1., 1.,
        -4., -1., -1., -1., -1., 2., 2., 2., 2.,
        4., -2., -2., -2., -2., 1., 1., 1., 1.,
        0., 1., 0., -1., 0., 1., -1., -1., 1.,
        0., -2., 0., 2., 0., 1., -1., -1., 1.,
        0., 0., 1., 0., -1., 1., 1., -1., -1.,
        0., 0., -2., 0., 2., 1., 1., -1., -1.,
        0., 1., -1., 1., -1., 0., 0., 0., 0.,
        0., 0., 0., 0., 0., 1., -1., 1., -1.};
    return M[k * 9 + v];
}
__device__ __forceinline__ constexpr double Mi(int k, int v) {
    constexpr double a9 = 0.111111111111111, b36 = 0.027777777777778, c18 = 0.055555555555556,
                     d6 = 0.166666666666667, e12 = 0.083333333333333, q4 = 0.25;
    constexpr double MI[81] = {
        a9, -a9, a9, 0, 0, 0, 0, 0, 0,
        a9, -b36, -c18, d6, -d6, 0, 0, q4, 0,
        a9, -b36, -c18, 0, 0, d6, -d6, -q4, 0,
        a9, -b36, -c18, -d6, d6, 0, 0, q4, 0,
        a9, -b36, -c18, 0, 0, -d6, d6, -q4, 0,
        a9, c18, b36, d6, e12, d6, e12, 0, q4,
        a9, c18, b36, -d6, -e12, d6, e12, 0, -q4,
        a9, c18, b36, -d6, -e12, -d6, -e12, 0, q4,
        a9, c18, b36, d6, e12, -d6, -e12, 0, -q4};
    return MI[k * 9 + v];
}

// icsr table (utils.hpp:749-767), 0-based; row MTS_NTYPES is the typeless
// wall link of a buffer cell (full bounce-back: target = opposite).
__constant__ signed char c_icsr[MTS_NTYPES + 1][9];

// Site helpers ---------------------------------------------------------------
struct Lat {
    int nx, ny;
    int64_t stride;
    __device__ __forceinline__ int64_t src(int x, int y, int j) const {
        int xs = x - ex(j), ys = y - ey(j);
        xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
        ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);
        return (int64_t)ys * nx + xs;
    }
    __device__ __forceinline__ int64_t nbr(int x, int y, int j) const {
        int xs = x + ex(j), ys = y + ey(j);
        xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
        ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);
        return (int64_t)ys * nx + xs;
    }
    // slot j of site (x, y) is fed across the x-seam (a periodic link when its
    // IN bit is set; only consulted for sites carrying SEAM_BIT)
    __device__ __forceinline__ bool seam_in(int x, int y, int j) const {
        return (x == 0 && ex(j) == 1) || (x == nx - 1 && ex(j) == -1);
    }
};

template <class S>
struct Store;
template <>
struct Store<double> {
    static constexpr bool dev = false;
    __device__ __forceinline__ static double ld(const double* a, int64_t i) { return __builtin_nontemporal_load(a + i); }
    __device__ __forceinline__ static void st(double* a, int64_t i, double v) { __builtin_nontemporal_store(v, a + i); }
};
template <>
struct Store<float> {
    static constexpr bool dev = true;
    __device__ __forceinline__ static float ld(const float* a, int64_t i) { return __builtin_nontemporal_load(a + i); }
    __device__ __forceinline__ static void st(float* a, int64_t i, float v) { __builtin_nontemporal_store(v, a + i); }
};

// ============================================================================
// FP64 reference-order arithmetic.  The reference evaluates dense 9x9
// matrix-vector products, summing over all nine columns in order; skipping
// the columns whose coefficient is zero leaves every partial sum unchanged,
// so these routines reproduce its rounding exactly (with FMA contraction
// disabled).
// ============================================================================
#pragma clang fp contract(off)

struct Macro64 {
    double rho, ux, uy, usq;
};

// coreLB.hpp:1018-1054 (force shift along x scaled by 2^(L-1): scale)
__device__ __forceinline__ Macro64 moments64(const double* f, double shiftx, double scale) {
    double r = 0.0, jx = 0.0, jy = 0.0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        r += f[j];
        if (ex(j) == 1) jx += f[j];
        if (ex(j) == -1) jx += -f[j];
        if (ey(j) == 1) jy += f[j];
        if (ey(j) == -1) jy += -f[j];
    }
    Macro64 m;
    m.rho = r;
    m.ux = jx / r + shiftx / (2.0 * r) * scale;
    m.uy = jy / r;
    m.usq = m.ux * m.ux + m.uy * m.uy;
    return m;
}

// coreLB.hpp:420-430
__device__ __forceinline__ double feq64(int j, const Macro64& m) {
    const double cu = (double)ex(j) * m.ux + (double)ey(j) * m.uy;
    return wq(j) * m.rho * (1.0 + 3.0 * cu + 4.5 * (cu * cu) - 1.5 * m.usq);
}

// out = Minv * diag(s) * M * in  (coreLB.hpp:634-643, three matMulLin calls)
__device__ __forceinline__ void mrt64(const double* in, const double* s, double* out) {
    double m[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        double acc = 0.0;
#pragma unroll
        for (int v = 0; v < 9; v++) {
            constexpr double z = 0.0;
            if (Mc(k, v) != z) acc += Mc(k, v) * in[v];
        }
        m[k] = s[k] * acc;
    }
#pragma unroll
    for (int j = 0; j < 9; j++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 9; k++)
            if (Mi(j, k) != 0.0) acc += Mi(j, k) * m[k];
        out[j] = acc;
    }
}

// coreLB.hpp:646-675: Hermite-regularised non-equilibrium part.
__device__ __forceinline__ void regularize64(const double* f, const double* fe, double* fneq) {
    double P0 = 0.0, P1 = 0.0, P2 = 0.0, P3 = 0.0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const double d = f[j] - fe[j];
        if (ex(j) != 0) P0 += d;
        if (ex(j) * ey(j) == 1) { P1 += d; P2 += d; }
        if (ex(j) * ey(j) == -1) { P1 += -d; P2 += -d; }
        if (ey(j) != 0) P3 += d;
    }
    const double cs2 = 1.0 / 3.0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const double H0 = ex(j) * ex(j) / cs2 - 1, H1 = ex(j) * ey(j) / cs2, H3 = ey(j) * ey(j) / cs2 - 1;
        fneq[j] = wq(j) * (1. / (2. * cs2)) * (P0 * H0 + P1 * H1 + P2 * H1 + P3 * H3);
    }
}

// Collision of one site in FP64 (coreLB.hpp:678-792); f in, post-collision
// out.  ep: parameter row.  feq from the moments m.
template <int COLL>
__device__ __forceinline__ void collide64(double* f, const Macro64& m, const CellParam& ep) {
    double fe[9];
#pragma unroll
    for (int j = 0; j < 9; j++) fe[j] = feq64(j, m);
    if (COLL == MTS_COLL_SRT) {
#pragma unroll
        for (int j = 0; j < 9; j++) f[j] = f[j] - 1.0 / 1.0 * (f[j] - fe[j]);
    } else if (COLL == MTS_COLL_MRT) {
        double d[9], t[9];
        const double s[9] = {1.0 / 1.0, 1.0 / 1.1, 1.0 / 1.2, ep.s_d, ep.s_q, ep.s_d, ep.s_q, ep.s_s, ep.s_s};
#pragma unroll
        for (int j = 0; j < 9; j++) d[j] = f[j] - fe[j];
        mrt64(d, s, t);
#pragma unroll
        for (int j = 0; j < 9; j++) f[j] = f[j] - t[j];
    } else {
        double fn[9], t[9];
        const double s[9] = {1.0 / 1.0, 1.0 / 1.1, 1.0 / 1.2, 1.0 / 1.0, ep.s_q, 1.0 / 1.0, ep.s_q, ep.s_s, ep.s_s};
        regularize64(f, fe, fn);
        mrt64(fn, s, t);
#pragma unroll
        for (int j = 0; j < 9; j++) f[j] = fe[j] + fn[j] - t[j];
    }
}

// Guo source in MRT space (coreLB.hpp:568-630): f += Minv (I - R/2) M S.
template <int COLL>
__device__ __forceinline__ void source64(double* f, double ux, double uy, double Fx, double Fy, const CellParam& ep) {
    const double cs2 = 1. / 3.;
    const double uF0 = ux * Fx, uF1 = ux * Fy, uF2 = uy * Fx, uF3 = uy * Fy;
    double S[9], t[9];
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const double cF = ex(j) * Fx + ey(j) * Fy;
        const double c0 = (double)ex(j) * ex(j) - cs2, c1 = (double)ex(j) * ey(j), c3 = (double)ey(j) * ey(j) - cs2;
        S[j] = wq(j) * (cF / cs2 + (uF0 * c0 + uF2 * c1 + uF1 * c1 + uF3 * c3) / (cs2 * cs2));
    }
    // the source always uses the non-regularised rates (tau_d from Kn)
    const double R[9] = {1.0 / 1.0, 1.0 / 1.1, 1.0 / 1.2, ep.s_d, ep.s_q, ep.s_d, ep.s_q, ep.s_s, ep.s_s};
    double s[9];
#pragma unroll
    for (int k = 0; k < 9; k++) s[k] = 1.0 - 0.5 * R[k];
    mrt64(S, s, t);
#pragma unroll
    for (int j = 0; j < 9; j++) f[j] = f[j] + t[j];
}

#pragma clang fp contract(on)

// ============================================================================
// Gather of the post-stream populations of one site (zeroDown -> stream ->
// wall rule -> periodic, coreLB.hpp:1319-1321, 814-818, in gather form).
// For FP32 deviation storage the bookkeeping of the w*rho0 offsets is done in
// FP64: base_j = (# of w-weighted contributions) - w_j.
// ============================================================================
template <class S, int WALL>
__device__ __forceinline__ void gather(const S* __restrict__ A, const Lat& L, int x, int y, int64_t p, uint32_t w,
                                       const CellParam& ep, double rho0, double* f) {
#pragma clang fp contract(off)
    // rest population and streamed/periodic slots
    f[0] = (double)Store<S>::ld(A, p);
    double base[9];
    base[0] = 0.0;
#pragma unroll
    for (int j = 1; j < 9; j++) {
        const bool in = (w >> (IN_SHIFT + j - 1)) & 1u;
        f[j] = in ? (double)Store<S>::ld(A, (int64_t)j * L.stride + L.src(x, y, j)) : 0.0;
        base[j] = in ? 0.0 : -wq(j);
    }
    const uint32_t wall = (w >> WALL_SHIFT) & 0xffu;
    if (wall) {
        double keep[9];
#pragma unroll
        for (int j = 1; j < 9; j++) keep[j] = f[j];
        const int t = (int)((w >> TYPE_SHIFT) & 63u);
        const double r = WALL == MTS_WALL_SRBB ? ep.r_srbb : 0.0;
#pragma unroll
        for (int ic = 1; ic < 9; ic++) {
            if (!((wall >> (ic - 1)) & 1u)) continue;
            const double a = (double)Store<S>::ld(A, (int64_t)ic * L.stride + p);
            if (WALL == MTS_WALL_BB) {
                f[opp(ic)] = a;  // coreLB.hpp:859-865
                base[opp(ic)] = wq(ic) - wq(opp(ic));
            } else {  // SRBB, coreLB.hpp:918-933
                f[opp(ic)] += r * a;
                base[opp(ic)] += r * wq(ic);
                const int sr = c_icsr[t][ic];
#pragma unroll
                for (int j = 1; j < 9; j++)
                    if (sr == j) {
                        f[j] += (1 - r) * a;
                        base[j] += (1 - r) * wq(ic);
                    }
            }
        }
        // periodic links overwrite after the wall rule (coreLB.hpp:969-977)
        if (w & SEAM_BIT) {
#pragma unroll
            for (int j = 1; j < 9; j++)
                if (L.seam_in(x, y, j) && ((w >> (IN_SHIFT + j - 1)) & 1u)) {
                    f[j] = keep[j];
                    base[j] = 0.0;
                }
        }
    }
    if (Store<S>::dev) {
        // f currently holds sums of deviations; add back the offsets.
#pragma unroll
        for (int j = 0; j < 9; j++) f[j] = f[j] + wq(j) * rho0 + base[j] * rho0;
    }
}

// ============================================================================
// Transport: MODE 0 collide-only (canonical f -> A), 1 fused (A -> A'),
// 2 stream-only (A -> canonical f, writes the moments).
// ============================================================================
struct TransportArgs {
    const void* Ain;
    void* Aout;
    double* fcan;
    const uint32_t* meta;
    const uint16_t* pidx;
    const CellParam* params;
    Lat L;
    double rho0, Fbody, fscale;   // fscale = 2^(L-1) of the level
    double *rho, *ux, *uy, *usq;  // moments output (mode 2) or nullptr
    double* usq_old;              // |u|^2 of the moments this pass computes (modes 0/1) or nullptr
    // mode 0: equilibrium from these moments instead of f's own (the feq of
    // coreLB.hpp:420-430 computed once per coarse step, stale on fine substeps)
    const double *min_rho, *min_ux, *min_uy, *min_usq;
};

template <class S, int MODE, int COLL, int WALL>
__global__ void __launch_bounds__(256) k_transport(TransportArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.L.nx) return;
    const int64_t p = (int64_t)y * a.L.nx + x;
    const uint32_t w = a.meta[p];
    if (!(w & FLUID_BIT)) return;
    const CellParam ep = a.params[a.pidx[p]];
    double f[9];
    if (MODE == 0) {
#pragma unroll
        for (int j = 0; j < 9; j++) f[j] = a.fcan[(int64_t)j * a.L.stride + p];
    } else {
        gather<S, WALL>((const S*)a.Ain, a.L, x, y, p, w, ep, a.rho0, f);
    }
    if (MODE == 2) {
#pragma unroll
        for (int j = 0; j < 9; j++) a.fcan[(int64_t)j * a.L.stride + p] = f[j];
        if (a.rho) {
            const Macro64 m = moments64(f, a.Fbody, a.fscale);
            a.rho[p] = m.rho;
            a.ux[p] = m.ux;
            a.uy[p] = m.uy;
            a.usq[p] = m.usq;
        }
        return;
    }
    Macro64 m;
    if (MODE == 0 && a.min_rho) m = Macro64{a.min_rho[p], a.min_ux[p], a.min_uy[p], a.min_usq[p]};
    else m = moments64(f, a.Fbody, a.fscale);
    if (a.usq_old) a.usq_old[p] = m.usq;
    collide64<COLL>(f, m, ep);
    S* out = (S*)a.Aout;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const double v = Store<S>::dev ? f[j] - wq(j) * a.rho0 : f[j];
        Store<S>::st(out, (int64_t)j * a.L.stride + p, (S)v);
    }
}

// ============================================================================
// Shan-Chen step (adsorption / coupled): LDS tile of TW x TH sites with a
// one-site halo of FP64 pseudopotential.  MODE 0 collide-only from canonical
// f, 1 fused.  The moments of the site (rho, u with the force shift), psi and
// F can be written for output.
// ============================================================================
constexpr int TW = 64, TH = 16, TROWS = 4;  // 256 threads = 64 x 4, 4 rows each

struct ScArgs {
    const void* Ain;
    void* Aout;
    const double* fcan;
    const uint32_t* meta;
    const uint16_t* pidx;
    const CellParam* params;
    Lat L;
    double rho0, Fbody, Gff, Gfs, T_lu, a_alpha_lu, b_lu;
    double *rho, *ux, *uy, *usq, *psi, *Fx, *Fy;  // outputs of this pass or nullptr
    int* psi_negative;                              // count of psi^2 < 0 sites (or nullptr)
};

// coreLB.hpp:515-531: psi = sqrt(acc), or psi = acc when acc < 0.
__device__ __forceinline__ double pseudo_potential(double rho, const ScArgs& a, bool& neg) {
#pragma clang fp contract(off)
    const double m = rho / 1.0;
    const double acc = (m * 1.0 * a.T_lu / (1 - a.b_lu * m) - a.a_alpha_lu * m * m / (1 + 2 * a.b_lu * m - a.b_lu * a.b_lu * m * m) -
                        (1.0 / 3.0) * rho) /
                       (3.0 * a.Gff);
    neg = acc < 0;
    return neg ? acc : sqrt(acc);
}

template <class S, int MODE, int COLL, int WALL, bool GENPER>
__device__ __forceinline__ void site_f(const ScArgs& a, int x, int y, int64_t p, uint32_t w, const CellParam& ep,
                                       int64_t gpIn, int64_t gpOut, double rhoin, double* f) {
#pragma clang fp contract(off)
    if (MODE == 0) {
#pragma unroll
        for (int j = 0; j < 9; j++) f[j] = a.fcan[(int64_t)j * a.L.stride + p];
    } else {
        gather<S, WALL>((const S*)a.Ain, a.L, x, y, p, w, ep, a.rho0, f);
        if (GENPER) {  // coreLB.hpp:980-1009 (W[2] on the diagonals)
            if (p == gpIn) {
                f[1] = (1.0 / 9.0) * rhoin;
                f[5] = (1.0 / 9.0) * rhoin;
                f[8] = (1.0 / 9.0) * rhoin;
            } else if (p == gpOut) {
                f[3] = (1.0 / 9.0) * rhoin;
                f[6] = (1.0 / 9.0) * rhoin;
                f[7] = (1.0 / 9.0) * rhoin;
            }
        }
    }
}

template <class S, int MODE, int COLL, int WALL, bool GENPER>
__global__ void __launch_bounds__(256) k_shanchen(ScArgs a, int64_t gpIn, int64_t gpOut, double rhoin) {
#pragma clang fp contract(off)
    __shared__ double s_psi[TH + 2][TW + 2];
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const int tid = threadIdx.y * TW + threadIdx.x;
    int neg = 0;
    // phase 1: psi of the tile and its halo
    for (int e = tid; e < (TH + 2) * (TW + 2); e += 256) {
        const int hy = e / (TW + 2), hx = e % (TW + 2);
        int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
        gx = gx < 0 ? gx + a.L.nx : (gx >= a.L.nx ? gx - a.L.nx : gx);
        gy = gy < 0 ? gy + a.L.ny : (gy >= a.L.ny ? gy - a.L.ny : gy);
        double ps = 0.0;
        const bool inside = (x0 - 1 + hx) < a.L.nx + 1 && (y0 - 1 + hy) < a.L.ny + 1;
        if (inside) {
            const int64_t p = (int64_t)gy * a.L.nx + gx;
            const uint32_t w = a.meta[p];
            if (w & FLUID_BIT) {
                const CellParam ep = a.params[a.pidx[p]];
                double f[9];
                site_f<S, MODE, COLL, WALL, GENPER>(a, gx, gy, p, w, ep, gpIn, gpOut, rhoin, f);
                double r = 0.0;
#pragma unroll
                for (int j = 0; j < 9; j++) r += f[j];
                bool ng;
                ps = pseudo_potential(r, a, ng);
                const bool own = hx >= 1 && hx <= TW && hy >= 1 && hy <= TH;
                if (own && ng) neg++;
            }
        }
        s_psi[hy][hx] = ps;
    }
    __syncthreads();
    // phase 2: forces, velocity, collision, source for the tile's own sites
    for (int k = 0; k < TROWS; k++) {
        const int lx = threadIdx.x, ly = threadIdx.y + 4 * k;
        const int x = x0 + lx, y = y0 + ly;
        if (x >= a.L.nx || y >= a.L.ny) continue;
        const int64_t p = (int64_t)y * a.L.nx + x;
        const uint32_t w = a.meta[p];
        if (!(w & FLUID_BIT)) continue;
        const CellParam ep = a.params[a.pidx[p]];
        double f[9];
        site_f<S, MODE, COLL, WALL, GENPER>(a, x, y, p, w, ep, gpIn, gpOut, rhoin, f);
        double r = 0.0, jx = 0.0, jy = 0.0;
#pragma unroll
        for (int j = 0; j < 9; j++) {
            r += f[j];
            if (ex(j) == 1) jx += f[j];
            if (ex(j) == -1) jx += -f[j];
            if (ey(j) == 1) jy += f[j];
            if (ey(j) == -1) jy += -f[j];
        }
        // forces, coreLB.hpp:533-566
        const double psi_i = s_psi[ly + 1][lx + 1];
        double fxff = 0.0, fyff = 0.0, fxfw = 0.0, fyfw = 0.0;
        const uint32_t outm = (w >> OUT_SHIFT) & 0xffu, wallm = (w >> WALL_SHIFT) & 0xffu;
#pragma unroll
        for (int ic = 1; ic < 9; ic++) {
            if ((outm >> (ic - 1)) & 1u) {
                const double pn = s_psi[ly + 1 + ey(ic)][lx + 1 + ex(ic)];
                const double c = (-6.0) * a.Gff * psi_i * wf(ic) * pn;
                if (ex(ic) != 0) fxff += c * ex(ic);
                if (ey(ic) != 0) fyff += c * ey(ic);
            }
            if ((wallm >> (ic - 1)) & 1u) {
                const double c = (-1.0) * a.Gfs * psi_i * psi_i * wf(ic);
                if (ex(ic) != 0) fxfw += c * ex(ic);
                if (ey(ic) != 0) fyfw += c * ey(ic);
            }
        }
        const double Fx = fxff + fxfw + a.Fbody, Fy = fyff + fyfw;
        // velocity with the half-force shift, coreLB.hpp:1069-1099
        Macro64 m;
        m.rho = r;
        m.ux = jx / r + Fx / (2.0 * r) * 1.0;
        m.uy = jy / r + Fy / (2.0 * r) * 1.0;
        m.usq = m.ux * m.ux + m.uy * m.uy;
        if (a.rho) {
            a.rho[p] = m.rho;
            a.ux[p] = m.ux;
            a.uy[p] = m.uy;
            a.usq[p] = m.usq;
        }
        if (a.psi) {
            a.psi[p] = psi_i;
            a.Fx[p] = Fx;
            a.Fy[p] = Fy;
        }
        collide64<COLL>(f, m, ep);
        source64<COLL>(f, m.ux, m.uy, Fx, Fy, ep);
        S* out = (S*)a.Aout;
#pragma unroll
        for (int j = 0; j < 9; j++) {
            const double v = Store<S>::dev ? f[j] - wq(j) * a.rho0 : f[j];
            Store<S>::st(out, (int64_t)j * a.L.stride + p, (S)v);
        }
    }
    if (a.psi_negative && neg) atomicAdd(a.psi_negative, neg);
}

// Stream-only pass of the Shan-Chen steps: A -> canonical f (wall rule of the
// step, periodic, generalized periodic).
template <class S, int WALL, bool GENPER>
__global__ void __launch_bounds__(256) k_sc_stream(ScArgs a, int64_t gpIn, int64_t gpOut, double rhoin) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.L.nx) return;
    const int64_t p = (int64_t)y * a.L.nx + x;
    const uint32_t w = a.meta[p];
    if (!(w & FLUID_BIT)) return;
    const CellParam ep = a.params[a.pidx[p]];
    double f[9];
    site_f<S, 1, MTS_COLL_MRT, WALL, GENPER>(a, x, y, p, w, ep, gpIn, gpOut, rhoin, f);
    double* fc = const_cast<double*>(a.fcan);
#pragma unroll
    for (int j = 0; j < 9; j++) fc[(int64_t)j * a.L.stride + p] = f[j];
}

// ============================================================================
// Reductions (FP64, fixed tree -> deterministic)
// ============================================================================
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// kind 0: {sum | |u| - |u_old| |, sum |u|} (coreLB.hpp:1248-1267)
// kind 1: {max |rho - rho_old| / rho_old, sum of the same}  (coreLB.hpp:1269-1283)
// kind 2: {sum rho, sum rho*ux... } diagnostics: {sum rho, sum jx, sum jy, count}
__global__ void __launch_bounds__(256) k_reduce(int kind, const uint32_t* meta, int64_t P, const double* a0,
                                                const double* a1, const double* a2, double* part) {
    __shared__ double sh[4][4];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < P; p += (int64_t)gridDim.x * 256) {
        if (!(meta[p] & FLUID_BIT)) continue;
        if (kind == 0) {
            const double un = sqrt(a0[p]), uo = sqrt(a1[p]);
            s0 += fabs(un - uo);
            s1 += un;
        } else if (kind == 1) {
            const double d = fabs(a0[p] - a1[p]) / a1[p];
            s0 = fmax(s0, d);
            s1 += d;  // fmax drops a NaN; the sum keeps it for the finiteness check
        } else {
            s0 += a0[p];
            s1 += a0[p] * a1[p];
            s2 += a0[p] * a2[p];
            s3 += 1.0;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (kind == 1) {
        s0 = wave_max(s0);
        s1 = wave_sum(s1);
    } else {
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        s2 = wave_sum(s2);
        s3 = wave_sum(s3);
    }
    if (lane == 0) {
        sh[wv][0] = s0;
        sh[wv][1] = s1;
        sh[wv][2] = s2;
        sh[wv][3] = s3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r0 = sh[0][0], r1 = sh[0][1], r2 = sh[0][2], r3 = sh[0][3];
        for (int i = 1; i < 4; i++) {
            r0 = kind == 1 ? fmax(r0, sh[i][0]) : r0 + sh[i][0];
            r1 += sh[i][1];
            r2 += sh[i][2];
            r3 += sh[i][3];
        }
        part[blockIdx.x * 4 + 0] = r0;
        part[blockIdx.x * 4 + 1] = r1;
        part[blockIdx.x * 4 + 2] = r2;
        part[blockIdx.x * 4 + 3] = r3;
    }
}

// 256 threads: strided partial sums, then a fixed-order tree in LDS
// (deterministic for a given partial count).
__global__ void __launch_bounds__(256) k_reduce_final(int kind, int nparts, const double* part, double* out) {
    __shared__ double s[4][256];
    const int t = threadIdx.x;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    for (int i = t; i < nparts; i += 256) {
        r0 = kind == 1 ? fmax(r0, part[4 * i]) : r0 + part[4 * i];
        r1 += part[4 * i + 1];
        r2 += part[4 * i + 2];
        r3 += part[4 * i + 3];
    }
    s[0][t] = r0;
    s[1][t] = r1;
    s[2][t] = r2;
    s[3][t] = r3;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            s[0][t] = kind == 1 ? fmax(s[0][t], s[0][t + h]) : s[0][t] + s[0][t + h];
            s[1][t] += s[1][t + h];
            s[2][t] += s[2][t + h];
            s[3][t] += s[3][t + h];
        }
        __syncthreads();
    }
    if (t < 4) out[t] = s[t][0];
}

// Gather/scatter between the dense site image and compact cell order.
__global__ void k_gather(const double* __restrict__ dense, int64_t stride, int q, const int32_t* cellPos, int64_t n,
                         double* __restrict__ compact) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int j = 0; j < q; j++) compact[c * q + j] = dense[(int64_t)j * stride + cellPos[c]];
}
__global__ void k_scatter(double* __restrict__ dense, int64_t stride, int q, const int32_t* cellPos, int64_t n,
                          const double* __restrict__ compact) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int j = 0; j < q; j++) dense[(int64_t)j * stride + cellPos[c]] = compact[c * q + j];
}

// LB2D::initialize (coreLB.hpp:381-417): f = W rho_in (u = 0).
__global__ void k_init(double* fcan, const uint32_t* meta, int64_t P, int64_t stride, double rhoin) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !(meta[p] & FLUID_BIT)) return;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const double cu = (double)ex(j) * 0.0 + (double)ey(j) * 0.0;
        fcan[(int64_t)j * stride + p] = wq(j) * rhoin * (1.0 + 3.0 * cu + 4.5 * (cu * cu) - 1.5 * 0.0);
    }
}

// Moments of the canonical f (calculateMacroscopicPorperties, coreLB.hpp:1018-1054).
__global__ void k_macro(const double* fcan, const uint32_t* meta, int64_t P, int64_t stride, double Fbody, double fscale,
                        double* rho, double* ux, double* uy, double* usq) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || !(meta[p] & FLUID_BIT)) return;
    double f[9];
#pragma unroll
    for (int j = 0; j < 9; j++) f[j] = fcan[(int64_t)j * stride + p];
    Macro64 m = moments64(f, Fbody, fscale);
    rho[p] = m.rho;
    ux[p] = m.ux;
    uy[p] = m.uy;
    usq[p] = m.usq;
}

}  // namespace dev

// ============================================================================
// Host launchers
// ============================================================================
using namespace dev;

// Launch on the context's stream; with timer events set, the events take the
// kernel's own start and end timestamps.
template <typename K, typename... A>
static void klaunch(K k, dim3 g, dim3 b, uint32_t sh, const LaunchCtx& c, A... a) {
    if (c.ev0) hipExtLaunchKernelGGL(k, g, b, sh, c.stream, c.ev0, c.ev1, 0, a...);
    else hipLaunchKernelGGL(k, g, b, sh, c.stream, a...);
}

void upload_icsr(hipStream_t st) {
    signed char t[MTS_NTYPES + 1][9];
    for (int i = 0; i < MTS_NTYPES; i++)
        for (int j = 0; j < 9; j++) t[i][j] = MTS_ICSR[i][j] < 0 ? 0 : MTS_ICSR[i][j];
    for (int j = 0; j < 9; j++) t[MTS_NTYPES][j] = (signed char)OPP[j];
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_icsr), t, sizeof(t), 0, hipMemcpyHostToDevice, st);
}

template <class S, int MODE, int COLL, int WALL>
static void launch_t(const LaunchCtx& c, const TransportArgs& a) {
    dim3 blk(256, 1), grd((c.nx + 255) / 256, c.ny);
    klaunch(k_transport<S, MODE, COLL, WALL>, grd, blk, 0, c, a);
}

template <class S, int MODE>
static int dispatch_t(const LaunchCtx& c, const TransportArgs& a, int coll, int wall) {
#define MTS_T(CO, WA)                                      \
    if (coll == CO && wall == WA) {                        \
        launch_t<S, MODE, CO, WA>(c, a);                   \
        return 0;                                          \
    }
    MTS_T(MTS_COLL_RMRT, MTS_WALL_SRBB)
    MTS_T(MTS_COLL_RMRT, MTS_WALL_BB)
    MTS_T(MTS_COLL_MRT, MTS_WALL_SRBB)
    MTS_T(MTS_COLL_MRT, MTS_WALL_BB)
    MTS_T(MTS_COLL_SRT, MTS_WALL_SRBB)
    MTS_T(MTS_COLL_SRT, MTS_WALL_BB)
#undef MTS_T
    return MTS_ERR_UNSUPPORTED;
}

int launch_transport(const LaunchCtx& c, int mode, const TransportSpec& s) {
    TransportArgs a;
    a.Ain = s.Ain;
    a.Aout = s.Aout;
    a.fcan = s.fcan;
    a.meta = c.meta;
    a.pidx = c.pidx;
    a.params = c.params;
    a.L = Lat{c.nx, c.ny, c.stride};
    a.rho0 = c.rho0;
    a.Fbody = c.Fbody;
    a.fscale = c.fscale;
    a.rho = s.rho;
    a.ux = s.ux;
    a.uy = s.uy;
    a.usq = s.usq;
    a.usq_old = s.usq_old;
    a.min_rho = s.min_rho;
    a.min_ux = s.min_ux;
    a.min_uy = s.min_uy;
    a.min_usq = s.min_usq;
    if (c.fp32) return c.cimg ? launch_ctransport(c, mode, s) : MTS_ERR_STATE;  // FP32 runs in cell order
    if (mode == 0) return dispatch_t<double, 0>(c, a, c.coll, c.wall);
    if (mode == 1) return dispatch_t<double, 1>(c, a, c.coll, c.wall);
    return dispatch_t<double, 2>(c, a, c.coll, c.wall);
}

template <class S, int MODE, int COLL, int WALL, bool GP>
static void launch_sc(const LaunchCtx& c, const ScArgs& a) {
    dim3 blk(TW, 4), grd((c.nx + TW - 1) / TW, (c.ny + TH - 1) / TH);
    hipLaunchKernelGGL((k_shanchen<S, MODE, COLL, WALL, GP>), grd, blk, 0, c.stream, a, c.gpIn, c.gpOut, c.rhoin);
}
template <class S, int WALL, bool GP>
static void launch_scs(const LaunchCtx& c, const ScArgs& a) {
    dim3 blk(256, 1), grd((c.nx + 255) / 256, c.ny);
    hipLaunchKernelGGL((k_sc_stream<S, WALL, GP>), grd, blk, 0, c.stream, a, c.gpIn, c.gpOut, c.rhoin);
}

// kind 0 adsorption (MRT, BB, generalized periodic); kind 1 coupled (RMRT, SRBB)
template <class S>
static void sc_dispatch(const LaunchCtx& c, int kind, int mode, const ScArgs& a) {
    if (kind == 0) {
        if (mode == 0) launch_sc<S, 0, MTS_COLL_MRT, MTS_WALL_BB, true>(c, a);
        else if (mode == 1) launch_sc<S, 1, MTS_COLL_MRT, MTS_WALL_BB, true>(c, a);
        else launch_scs<S, MTS_WALL_BB, true>(c, a);
    } else {
        if (mode == 0) launch_sc<S, 0, MTS_COLL_RMRT, MTS_WALL_SRBB, false>(c, a);
        else if (mode == 1) launch_sc<S, 1, MTS_COLL_RMRT, MTS_WALL_SRBB, false>(c, a);
        else launch_scs<S, MTS_WALL_SRBB, false>(c, a);
    }
}

int launch_shanchen(const LaunchCtx& c, int kind, int mode, const ScSpec& s) {
    ScArgs a;
    a.Ain = s.Ain;
    a.Aout = s.Aout;
    a.fcan = s.fcan;
    a.meta = c.meta;
    a.pidx = c.pidx;
    a.params = c.params;
    a.L = Lat{c.nx, c.ny, c.stride};
    a.rho0 = c.rho0;
    a.Fbody = c.Fbody;
    a.Gff = c.Gff;
    a.Gfs = c.Gfs;
    a.T_lu = c.T_lu;
    a.a_alpha_lu = c.a_alpha_lu;
    a.b_lu = 2.0 / 21.0;
    a.rho = s.rho;
    a.ux = s.ux;
    a.uy = s.uy;
    a.usq = s.usq;
    a.psi = s.psi;
    a.Fx = s.Fx;
    a.Fy = s.Fy;
    a.psi_negative = s.psi_negative;
    if (c.fp32) return c.cimg ? launch_cshanchen(c, kind, mode, s) : MTS_ERR_STATE;
    sc_dispatch<double>(c, kind, mode, a);
    return 0;
}

void launch_reduce(const LaunchCtx& c, int kind, const double* a0, const double* a1, const double* a2, double* part,
                   int nparts, double* out) {
    const int64_t P = (int64_t)c.nx * c.ny;
    hipLaunchKernelGGL(k_reduce, dim3(nparts), dim3(256), 0, c.stream, kind, c.meta, P, a0, a1, a2, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c.stream, kind, nparts, part, out);
}

namespace dev {
__global__ void __launch_bounds__(256) k_probe_copy(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}
}  // namespace dev

void launch_probe_copy(float* dst, const float* src, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(dev::k_probe_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, src, n);
}

namespace dev {
// Bandwidth copy: one 16-byte nontemporal load and store per lane (n4
// four-float elements); of the widths and per-thread counts tried (4, 8 and
// 16 bytes, 1 to 8 per thread) the fastest, at about 6.6 TB/s on MI355X.
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_probe_copy16(v4f* __restrict__ dst, const v4f* __restrict__ src, int64_t n4) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}
}  // namespace dev

void launch_probe_copy16(float* dst, const float* src, int64_t n, hipStream_t st) {
    const int64_t n4 = n / 4;
    hipLaunchKernelGGL(dev::k_probe_copy16, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, (dev::v4f*)dst,
                       (const dev::v4f*)src, n4);
}

void launch_gather(const LaunchCtx& c, const double* dense, int q, double* compact) {
    hipLaunchKernelGGL(k_gather, dim3((c.ncell + 255) / 256), dim3(256), 0, c.stream, dense, c.stride, q, c.cellPos,
                       c.ncell, compact);
}
void launch_scatter(const LaunchCtx& c, double* dense, int q, const double* compact) {
    hipLaunchKernelGGL(k_scatter, dim3((c.ncell + 255) / 256), dim3(256), 0, c.stream, dense, c.stride, q, c.cellPos,
                       c.ncell, compact);
}
void launch_init(const LaunchCtx& c, double* fcan) {
    const int64_t P = (int64_t)c.nx * c.ny;
    hipLaunchKernelGGL(k_init, dim3((P + 255) / 256), dim3(256), 0, c.stream, fcan, c.meta, P, c.stride, c.rhoin);
}
void launch_macro(const LaunchCtx& c, const double* fcan, double* rho, double* ux, double* uy, double* usq) {
    const int64_t P = (int64_t)c.nx * c.ny;
    hipLaunchKernelGGL(k_macro, dim3((P + 255) / 256), dim3(256), 0, c.stream, fcan, c.meta, P, c.stride, c.Fbody,
                       c.fscale, rho, ux, uy, usq);
}

}  // namespace mts
#include "mts_phase.inc"
#include "mts_fast.inc"
#include "mts_compact.inc"
#include "mts_tile.inc"
#include "mts_multi.inc"
