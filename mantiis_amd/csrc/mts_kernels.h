// Launch interface between the C ABI (mts_api.hip) and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mts_common.h"

namespace mts {

struct LaunchCtx {
    hipStream_t stream;
    int nx, ny;
    int64_t stride, ncell;
    const uint32_t* meta;
    const uint16_t* pidx;
    const CellParam* params;
    const int32_t* cellPos;
    bool fp32;
    int coll, wall;
    double rho0, rhoin, Fbody, Gff, Gfs, T_lu, a_alpha_lu;
    int64_t gpIn, gpOut;
    bool one_param;  // a single (tau_s, Kn) pair: the per-site index is not read
};

struct TransportSpec {
    const void* Ain = nullptr;
    void* Aout = nullptr;
    double* fcan = nullptr;
    double *rho = nullptr, *ux = nullptr, *uy = nullptr, *usq = nullptr;
    double* usq_old = nullptr;
};

struct ScSpec {
    const void* Ain = nullptr;
    void* Aout = nullptr;
    double* fcan = nullptr;
    double *rho = nullptr, *ux = nullptr, *uy = nullptr, *usq = nullptr;
    double *psi = nullptr, *Fx = nullptr, *Fy = nullptr;
    int* psi_negative = nullptr;
};

void upload_icsr(hipStream_t st);
// FP32 production kernels (mts_fast.inc)
int launch_transport32(const LaunchCtx& c, int mode, const TransportSpec& s);
int launch_shanchen32(const LaunchCtx& c, int kind, int mode, const ScSpec& s);
// mode 0 collide-only (fcan -> Aout), 1 fused (Ain -> Aout), 2 stream-only (Ain -> fcan)
int launch_transport(const LaunchCtx& c, int mode, const TransportSpec& s);
// kind 0 adsorption, 1 coupled; mode as above
int launch_shanchen(const LaunchCtx& c, int kind, int mode, const ScSpec& s);
// kind 0 velocity residual partials, 1 density residual, 2 diagnostics; out[4]
void launch_reduce(const LaunchCtx& c, int kind, const double* a0, const double* a1, const double* a2, double* part,
                   int nparts, double* out);
void launch_gather(const LaunchCtx& c, const double* dense, int q, double* compact);
void launch_scatter(const LaunchCtx& c, double* dense, int q, const double* compact);
void launch_init(const LaunchCtx& c, double* fcan);
void launch_macro(const LaunchCtx& c, const double* fcan, double* rho, double* ux, double* uy, double* usq);
// Unfused LB2D methods on the canonical FP64 arrays (mts_phase.hip).
struct PhaseArrays {
    double *f, *ftemp, *feq, *fneq, *fb;
    double *rho, *rho_old, *ux, *uy, *usq, *usq_old, *psi, *Fx, *Fy;
};
int launch_phase(const LaunchCtx& c, int method, int lvl, const PhaseArrays& a, int* psi_negative);

}  // namespace mts
