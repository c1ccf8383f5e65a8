// Launch interface between the C ABI (mts_api.hip) and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "mts_common.h"

namespace mts {

// Device copy of a CompactImage (cell-order FP32 passes).
struct CompactDev {
    const uint32_t* cw = nullptr;
    const uint32_t* rw = nullptr;  // packed rank offsets / aux record
    const uint16_t* pidx = nullptr;
    const uint8_t* outm = nullptr;
    const int32_t* aux = nullptr;
    const int32_t* cidx = nullptr;
    const int32_t* segs = nullptr;
    const void* cpar = nullptr;  // FP32 parameter records (32 B per (tau_s, Kn) pair)
    const void* book = nullptr;  // SRBB bookkeeping table (24 floats per pair)
    const int32_t* cellPos = nullptr;
    int64_t ncell = 0, cstride = 0;
    int ntx = 0;
    const float* O = nullptr;  // interface overrides (multi-block buffer cells)
    int64_t OS = 0;
    int ovr = 0;  // 0 gather, 1 take the override, 2 skip (see CImg)
    const float* M = nullptr;  // MDBB state read by the gather (set per launch)
};

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
    bool one_param;       // a single (tau_s, Kn) pair: the per-site index is not read
    double fscale = 1.0;  // force-shift scale 2^(L-1) of the lattice level
    const CompactDev* cimg = nullptr;  // FP32 passes in cell order when set
    // kernel timer: when set, the dominant launch records these events at
    // its own start and end (hipExtLaunchKernel)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct TransportSpec {
    const void* Ain = nullptr;
    void* Aout = nullptr;
    double* fcan = nullptr;
    double *rho = nullptr, *ux = nullptr, *uy = nullptr, *usq = nullptr;
    double* usq_old = nullptr;
    const double *min_rho = nullptr, *min_ux = nullptr, *min_uy = nullptr, *min_usq = nullptr;
    // MDBB (FP32 cell order): the previous collided state of the wall cells
    // read by the gather, and the one this pass writes (4 planes each)
    const float* Min = nullptr;
    float* Mout = nullptr;
};

struct ScSpec {
    const void* Ain = nullptr;
    void* Aout = nullptr;
    double* fcan = nullptr;
    double *rho = nullptr, *ux = nullptr, *uy = nullptr, *usq = nullptr;
    double *psi = nullptr, *Fx = nullptr, *Fy = nullptr;
    int* psi_negative = nullptr;
};

// One level substep of the multi-block FP32 path (cell order).
struct MbSpec {
    const float* Ain = nullptr;
    float* Aout = nullptr;
    const double* fcan = nullptr;  // mode 0
    float* m = nullptr;            // stored fields (written for mac 0, read for mac 1)
    double* usq_old = nullptr;
};
// Interface rows of the FP32 path in cell indices of the level images, with
// the weights formed on the host in FP64 (as the FP64 kernels form them).
struct CoarseRow32 {
    int32_t parent, child[4];  // level-2 / level-1 cells (-1 absent)
    float k;                   // 2 tau_c / mean tau_f
    int32_t nch;
    int32_t slot;              // override slot of the parent
};
struct FineRow32 {
    int32_t child[4];  // level-1 cells (fine buffers), override slots 4 r + k
    int32_t src[6];    // level-2 cells of the interpolation nodes
    float L[4][6];     // Lagrange weights of the 3 x 2 node grid per child
    float ratio[4];
};
// Interface of the multi-block FP32 path: a block gathers up to IF_TASKS
// post-stream states into LDS, forms up to IF_CO coarse-buffer values from
// them and up to IF_FO fine-buffer values.  A fine row's interpolation node
// that is a coarse-buffer cell takes the value formed in the same block from
// its children, so no block waits for another.  The values are either
// written as overrides, or are the buffer cells' post-stream states of the
// next step, updated in the same launch as that step's other cells.
constexpr int IF_TASKS = 256, IF_CO = 64, IF_FO = 256;
struct IfHead {
    int32_t ntask, nco, nfo, pad;
};
struct IfTask {    // one post-stream state to gather
    int32_t cell;  // cell of its level image
    uint32_t cw, rw;
    uint32_t pl;   // parameter index, bit 31: level 2
};
struct IfCoarse {     // a coarse-buffer value (coreLB.hpp:1101-1150)
    int16_t src[4];   // task slots of the children (-1 absent)
    int32_t parent;   // level-2 cell (stored fields)
    float k;          // 2 tau_c / mean tau_f
    int32_t nch;
    int32_t out;      // override slot to write, -1: used in this block only
    int32_t pidx;     // parameter index of the parent
    int32_t pad;
};
struct IfFine {       // one fine-buffer value (coreLB.hpp:1152-1231)
    int32_t child;    // level-1 cell (stored fields)
    int32_t out;      // override slot 4 r + k
    float ratio;
    int16_t node[6];  // LDS value slots of the interpolation nodes
    uint16_t pidx;    // parameter index of the child
    int16_t pad;
    float L[6];       // Lagrange weights of the 3 x 2 node grid
};
struct IfSpec {
    const float *A1 = nullptr, *A2 = nullptr;
    const float *m1 = nullptr, *m2 = nullptr;
    float *O1 = nullptr, *O2 = nullptr;
    const IfHead* head = nullptr;
    const IfTask* task = nullptr;   // [nblocks * IF_TASKS]
    const IfCoarse* co = nullptr;   // [nblocks * IF_CO]
    const IfFine* fo = nullptr;     // [nblocks * IF_FO]
    int nblocks = 0;
    int child_ovr = 0;              // 1: a child with an override takes it
};
// Interface of one coarse step fused with the next step's first substeps
// (launch_cmultistep2 with the overrides formed in the launch).
int launch_cinterface_step(const LaunchCtx& c1, const MbSpec& s1, const LaunchCtx& c2, const MbSpec& s2,
                           const IfSpec& f);

void upload_icsr(hipStream_t st);
// FP32 production kernels in cell order (mts_compact.inc)
void upload_spec(const uint32_t* spec64, hipStream_t st);
// The FP32 parameter records (8 floats per (tau_s, Kn) pair) and the SRBB
// bookkeeping table (24 floats per pair) of a parameter list.
void fp32_params(const std::vector<CellParam>& params, double rho0, std::vector<float>& cpar, std::vector<float>& book);
int launch_ctransport(const LaunchCtx& c, int mode, const TransportSpec& s);
int launch_cshanchen(const LaunchCtx& c, int kind, int mode, const ScSpec& s);
// mode 0 from the canonical state, 1 gather; mac 0 fresh fields, 1 stored
int launch_cmultistep(const LaunchCtx& c, int mode, int mac, const MbSpec& s);
int launch_cinterface(const LaunchCtx& c1, const LaunchCtx& c2, const IfSpec& s);
// Substeps of n = 2 or 3 level images in one launch (c[0] gives the
// operators), all with the same mode; bit i of mac: image i uses the step's
// stored fields.
int launch_cmultistep3(int n, const LaunchCtx* c, const MbSpec* s, int mode, int mac);
// The interface of levels (1, 2) of a three-level lattice in the launch of
// the next step's first level-1 substep: fine buffer values update their
// cells there, coarse buffer values are written as overrides.
int launch_cinterface_step_fine(const LaunchCtx& c1, const MbSpec& s1, const LaunchCtx& c2, const IfSpec& f);
// An interface of a pair of levels written as overrides, and in the same
// launch a substep of one level image that reads none of them.
int launch_cinterface_beside(const LaunchCtx& c1, const LaunchCtx& c2, const IfSpec& f, const LaunchCtx& ca,
                             const MbSpec& sa, int mode, int mac);
// Batched single-block lattices (cell order): one block of 256 cells of one
// lattice per queue entry.
// A 256-cell block of the batched image: cells [start, end) of lattice
// geom; id = its index among the batch's blocks; pair = the lattice's
// parameter pair when it has only one (its cells then skip their pair
// index), -1 otherwise.  An entry with start == end pads the work queue to
// the grid size.
struct BatBlock {
    int32_t start, end, geom, id, pair;
};
struct BatchLaunch {
    LaunchCtx c;  // cimg: the concatenated image
    const float* Ain = nullptr;
    float* Aout = nullptr;
    const BatBlock* queue = nullptr;  // maxblocks entries: the active blocks, then padding
    int maxblocks = 0;                // grid size
    float* mac = nullptr;
    double* part = nullptr;  // per-block (sum rho, sum j_x)
    int from_rest = 0;
    int threads = 256;       // threads per 256-cell block: 64, 128 or 256
    int xrun = 32;           // XCD run length of the queue map (0: none)
    const float* Min = nullptr;  // MDBB state (see TransportSpec)
    float* Mout = nullptr;
};
int launch_cbatch(const BatchLaunch& L);
// Batched tile kernels (mts_tile.inc): one block per TILE_X x TILE_Y site
// tile of a torus lattice in tile-major cell order, nsteps (1 or 2) steps per
// launch.
constexpr int TILE_X = 32, TILE_Y = 16;
struct TileLaunch {
    hipStream_t stream = nullptr;
    int coll = 0, wall = 0;
    int nsteps = 2;
    int grid = 0;                     // blocks (tiles past *nactive exit)
    const float* Ain = nullptr;
    float* Aout = nullptr;
    const uint32_t* tw = nullptr;     // per cell: link bits | wall bits | type | CW_COMPLEX
    const uint16_t* pidx = nullptr;
    const void* cpar = nullptr;       // FP32 parameter records
    const void* book = nullptr;       // SRBB bookkeeping table
    const uint32_t* tmask = nullptr;  // per tile row: fluid mask
    const int32_t* tpre = nullptr;    // per tile row: first cell
    const int32_t* tnb = nullptr;     // per tile: 3 x 3 tiles around it
    const int32_t* queue = nullptr;
    const int32_t* nactive = nullptr;
    float* mac = nullptr;
    double* part = nullptr;
    int64_t S = 0;
    double rho0 = 0;
    float Fbody = 0;
    int xrun = 0;
    bool one = false;
};
int launch_ctile(const TileLaunch& L);
// acc[2g..2g+1] = sum of part over the blocks [first[g], first[g+1])
void launch_bsum(const double* part, const int32_t* first, double* acc, int nb, hipStream_t st);
// first substeps (fresh fields) of two levels in one launch
int launch_cmultistep2(const LaunchCtx& ca, const MbSpec& sa, const LaunchCtx& cb, const MbSpec& sb, int mode);
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
void launch_probe_copy(float* dst, const float* src, int64_t n, hipStream_t st);
void launch_probe_copy16(float* dst, const float* src, int64_t n, hipStream_t st);
void launch_macro(const LaunchCtx& c, const double* fcan, double* rho, double* ux, double* uy, double* usq);
// Multi-block interface (coreLB.hpp:1101-1231) on the canonical FP64 arrays.
struct CoarseRow {  // coarse buffer site (level 2) and its 4 children (level-1 sites, -1 absent)
    int64_t parent, child[4];
    double tc;       // tau of the coarse buffer
    double tf;       // mean tau of the children as the reference forms it
    int nch;
};
struct FineRow {  // fine buffer children of one coarse cell
    int64_t child[4];  // level-1 sites
    int64_t src[6];    // level-2 sites of the 3x2 interpolation nodes Z[3*b + a]
    double xg[3], yg[2];
    double qx[4], qy[4];
    double ratio[4];  // tau_f / (2 tau_c)
};
void launch_coarse_interp(hipStream_t st, int64_t nrows, const CoarseRow* rows, double* f2, int64_t stride2,
                          const double* f1, int64_t stride1, const double* rho2, const double* ux2, const double* uy2,
                          const double* usq2);
void launch_fine_interp(hipStream_t st, int64_t nrows, const FineRow* rows, double* f1, int64_t stride1, const double* f2,
                        int64_t stride2, const double* rho1, const double* ux1, const double* uy1, const double* usq1);
void launch_pixel_out(hipStream_t st, int64_t npix, const int64_t* src, const double* const* level_arrays, double* out);
// Unfused LB2D methods on the canonical FP64 arrays (mts_phase.hip).
struct PhaseArrays {
    double *f, *ftemp, *feq, *fneq, *fb;
    double *rho, *rho_old, *ux, *uy, *usq, *usq_old, *psi, *Fx, *Fy;
};
int launch_phase(const LaunchCtx& c, int method, int lvl, const PhaseArrays& a, int* psi_negative);

}  // namespace mts
