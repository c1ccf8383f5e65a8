/*
 * mantiis_hip.h -- C ABI of the MI355X (gfx950) D2Q9 lattice-Boltzmann step.
 *
 * This is the drop-in boundary for the mantiis time step: the C++ classes
 * LB2D / Grid2D / Shape / MantiisApp (the .hpp headers beside this file,
 * named like include/{coreLB,grid,shape,mantiisApp}.hpp of the reference)
 * are thin wrappers over these functions.  Plain pointers and sizes only; no
 * exception crosses this boundary.  Every function returns MTS_OK (0) or a
 * negative status; mts_last_error() gives the message of the calling thread's
 * last failure.
 *
 * Threading: a handle is used from one host thread at a time; all work of a
 * handle is issued on one HIP stream owned by the handle.  Functions that
 * return host data synchronise that stream.
 *
 * Cell order: "compact" arrays follow the reference's cell list -- fluid
 * pixels in row-major order for the single-block grid (grid.hpp:401-424,
 * 427-889) -- so that a compact array here and an LB2D std::vector of the
 * reference hold the same cell at the same index.
 */
#ifndef MANTIIS_HIP_H
#define MANTIIS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference throws std::runtime_error / calls exit) */
#define MTS_OK 0
#define MTS_ERR_ARG (-1)         /* bad argument / size                                  */
#define MTS_ERR_GEOMETRY (-2)    /* "[ERROR]: Boundary type is not contained." grid.hpp:1143-1150 */
#define MTS_ERR_IO (-3)          /* missing / short input file, io.hpp:58-69             */
#define MTS_ERR_HIP (-4)         /* HIP runtime failure                                  */
#define MTS_ERR_OOM (-5)         /* device allocation failed                             */
#define MTS_ERR_STATE (-6)       /* call not valid in the handle's current state         */
#define MTS_ERR_UNSUPPORTED (-7) /* operator/precision combination not provided          */
#define MTS_ERR_NONFINITE (-8)   /* NaN/Inf detected in a residual reduction             */
#define MTS_ERR_PHYSICS (-9)     /* PR-EOS cubic has no single real root (utils.hpp:571-600) */

/* ---- operator codes: Settings.collision_method / wall_boundary
 * (mantiisApp.hpp:46-64, 180-234) */
#define MTS_COLL_SRT 0
#define MTS_COLL_MRT 1
#define MTS_COLL_RMRT 2
#define MTS_WALL_BB 0
#define MTS_WALL_MDBB 1
#define MTS_WALL_SRBB 2

/* ---- storage precision of the fused step */
#define MTS_FP64 0     /* populations and moments in FP64                         */
#define MTS_FP32DEV 1  /* populations stored in FP32 as f_i - w_i*rho0; FP64 sums  */

typedef struct mts_sim_s* mts_sim;

/* Physical and numerical settings.  dx/Fbody/T/P/collision/wall mirror
 * Settings (mantiisApp.hpp:46-64); Gff/Gfs are the Shan-Chen strengths that
 * the reference fixes at compile time (globalParameters.hpp:120-121). */
typedef struct mts_params {
    double dx;       /* lattice spacing Cl [m]                 default 1e-9  */
    double Fbody;    /* body force along +x, lattice units     default 1e-8  */
    double T;        /* temperature [K]                        default 300   */
    double P;        /* pressure [Pa]                          default 2e6   */
    double Gff;      /* fluid-fluid strength                   default -1    */
    double Gfs;      /* fluid-solid (adsorption) strength      default -12.7 */
    int collision;   /* MTS_COLL_*                             default RMRT  */
    int wall;        /* MTS_WALL_*                             default SRBB  */
    int precision;   /* MTS_FP64 / MTS_FP32DEV                 default FP64  */
    int multiblock;  /* 0: single block; 1: the reference's quadtree of
                      * at most two levels (maxAllowedLevel = 2, grid.hpp:
                      * 313-399, 1180); k = 2..MTS_MAX_LEVELS: at most k
                      * levels (the same recursion with maxAllowedLevel = k) */
    int device;      /* HIP device ordinal                     default 0     */
    int periodic_y;  /* 0: the reference's seam rule (coreLB.hpp:301-331):
                      *    x-seam links periodic (overwrite after the wall
                      *    rule), y-seam-only links dropped (default);
                      * 1: torus -- links across either seam are classified
                      *    like interior links (mass-conserving)            */
    int batch_steps; /* batched solver (mts_batch_*): time steps per launch --
                      * 1: one step per launch in cell order (default);
                      * 2: the tile kernels, two steps per launch staged in
                      *    LDS (every side must be a multiple of 32 x 16,
                      *    else one step per launch; not with MDBB walls)  */
} mts_params;

/* Fills the defaults above. */
void mts_default_params(mts_params* p);

/* Last error message of the calling thread ("" if none). */
const char* mts_last_error(void);

/* _GLOBAL_::recompute() (globalParameters.hpp:180-221): out[10] =
 * {rhoin, mfp, Cf, T_lu, a_alpha_lu, Z_gas, Crho, Cu, Cl, cubic_ok}. */
int mts_globals(const mts_params* p, double* out10);

/* Shape::loadExistingModel (shape.hpp:253-285): reads folder/pore.dat,
 * Kn.dat, localporesize.dat (column-major text, n*n values each) into
 * row-major [y][x] arrays supplied by the caller. */
int mts_load_existing_model(const char* folder, int n, int32_t* pore, double* kn, double* locpore);

/* Grid2D(shape) + Grid2D::initialize + LB2D::LB2D (grid.hpp:266-311,
 * coreLB.hpp:231-378): builds the cell list, neighbour connectivity,
 * boundary types and link tables of a row-major [y][x] geometry
 * (pore 1 = fluid), the per-cell relaxation times (coreLB.hpp:482-502) and
 * the device state; then LB2D::initialize (coreLB.hpp:381-417). */
int mts_create(int nx, int ny, const int32_t* pore, const double* kn, const double* locpore,
               const mts_params* p, mts_sim* out);
int mts_destroy(mts_sim s);

/* Largest level count of a quadtree lattice (mts_params.multiblock). */
#define MTS_MAX_LEVELS 3

/* ---- Grid2D on the host (grid.hpp:167-1157), no device needed.
 * mts_grid_create builds the cell list of a row-major [y][x] geometry exactly
 * as Grid2D(shape) + Grid2D::initialize (single-block, or the quadtree when
 * p->multiblock is set; p->periodic_y selects the seam rule).
 * Fails with MTS_ERR_GEOMETRY on an unsupported wall pattern. */
typedef struct mts_grid_s* mts_grid;
int mts_grid_create(int nx, int ny, const int32_t* pore, const double* kn, const double* locpore,
                    const mts_params* p, mts_grid* out);
int mts_grid_destroy(mts_grid g);
/* *n = Grid2D::globalGridSize, *max_level = Grid2D::maxLevel */
int mts_grid_size(mts_grid g, int64_t* n, int32_t* max_level);
/* Grid2D arrays (any pointer may be NULL): gridID [n], gridIJ [n*2] = {x, y},
 * gridLevel [n], gridConnect [n*9], bndTypes [n], gridType [n],
 * gridIsBuffer [n], taus [n] (relaxation time of each cell).  The taus of a
 * multi-block grid come from per-level pore sizes (LB2D::calculateMultigridTauS,
 * coreLB.hpp:432-480) computed on the current device: only that request
 * needs a GPU. */
int mts_grid_arrays(mts_grid g, int64_t* gridID, int32_t* gridIJ, int32_t* level, int64_t* connect, int32_t* bnd,
                    int32_t* type, int32_t* isBuffer, double* taus);
/* which = 0: Grid2D::gridParentChild rows {cell, tl, tr, bl, br};
 * 1: the coarseBuffer rows of every pair of levels (coarseBuffer[21], then
 * [32], ...); 2: the fineBuffer rows (fineBuffer[12], then [23], ...);
 * which = a buffer type (21, 32, 12, 23): that type's rows only.  Rows of 5
 * int64; call with rows = NULL to get *count (rows). */
int mts_grid_rows(mts_grid g, int which, int64_t* rows, int64_t* count);

/* ---- Quadtree tables on the device (no counterpart in the reference,
 * which builds them on the host in grid.hpp:313-1157).  mts_create of a
 * multi-block lattice makes the level decision and the neighbour and
 * interface tables (connectivity, boundary types, parent/child, coarse and
 * fine buffer rows) and the level images' link words on the device; the
 * buffer cells, whose placement is sequential, stay on the host.
 * mts_quadtree_check builds one geometry both ways (the host builder behind
 * mts_grid_create, and the device builder) and sets *ndiff to the number of
 * table entries that differ (0: identical).  Needs a GPU. */
int mts_quadtree_check(int nx, int ny, const int32_t* pore, const double* kn, const mts_params* p, int64_t* ndiff);

/* Grid2D::globalGridSize (grid.hpp:195). */
int mts_cell_count(mts_sim s, int64_t* n);

/* Grid2D arrays in compact order (grid.hpp:170-186): gridID [n],
 * gridIJ [n*2] = {x, y}, gridLevel [n], gridConnect [n*9], bndTypes [n].
 * Any pointer may be NULL. */
int mts_grid_info(mts_sim s, int64_t* gridID, int32_t* gridIJ, int32_t* level, int64_t* connect,
                  int32_t* bnd);

/* Link tables of LB2D::LB2D (coreLB.hpp:272-366) decoded from the device's
 * per-cell link words: which = 0 stream (id, loc, ic), 1 boundary (id, ic),
 * 2 periodic (id, i2, ic).  Call with NULL arrays to get *count first. */
int mts_links(mts_sim s, int which, int64_t* id, int64_t* loc, int32_t* ic, int64_t* count);

/* LB2D::setCollision / setWallBoundary (coreLB.hpp:794-807). */
int mts_set_operators(mts_sim s, int collision, int wall);

/* _GLOBAL_::Fbody (globalParameters.hpp:49), which MantiisApp::run sets to
 * 0 for the adsorption warm-up of the coupled mode (mantiisApp.hpp:379-384). */
int mts_set_body_force(mts_sim s, double Fbody);

/* LB2D::initialize (coreLB.hpp:381-417): f = W*rho_in, then moments. */
int mts_initialize(mts_sim s);

/* One call = nsteps of MantiisApp::transport_simulation_loop
 * (mantiisApp.hpp:239-246): equilibrium, evolutionStepOfMultiBlock,
 * calculateMacroscopicPorperties, calculateVelocityDifference -- fused into
 * one collide-stream pass per step.  The residual of the last step is
 * available from mts_residual; f, rho, ux, uy, usq afterwards equal the
 * reference's vectors after nsteps. */
int mts_step_transport(mts_sim s, int nsteps);

/* nsteps of MantiisApp::adsorption_simulation_loop (mantiisApp.hpp:248-265)
 * (Shan-Chen psi, forces, MRT + Guo source, BB, periodic, generalized
 * periodic, density residual), fused. */
int mts_step_adsorption(mts_sim s, int nsteps);

/* nsteps of MantiisApp::coupled_simulation_loop (mantiisApp.hpp:267-284). */
int mts_step_coupled(mts_sim s, int nsteps);

/* LB2D::diff (coreLB.hpp:154) after the last step call. */
int mts_residual(mts_sim s, double* diff);

/* Unfused LB2D methods, one HIP kernel each, for phase-level use by the
 * LB2D wrapper.  lvl is the level argument of the reference method. */
#define MTS_M_EQUILIBRIUM 1            /* coreLB.hpp:420-430   */
#define MTS_M_REGULARIZE 2             /* 646-675              */
#define MTS_M_SRT 3                    /* 678-691              */
#define MTS_M_MRT 4                    /* 694-742              */
#define MTS_M_RMRT 5                   /* 744-792              */
#define MTS_M_COLLIDE 6                /* 809-812              */
#define MTS_M_ZERODOWN 7               /* 1233-1246            */
#define MTS_M_STREAM 8                 /* 834-856              */
#define MTS_M_BBWALL 9                 /* 859-865              */
#define MTS_M_SRBBWALL 10              /* 918-933              */
#define MTS_M_MDBBWALL 11              /* 936-967              */
#define MTS_M_PERIODIC 12              /* 969-977              */
#define MTS_M_GENPERIODIC 13           /* 980-1009             */
#define MTS_M_MACROSCOPIC 14           /* 1018-1054            */
#define MTS_M_DENSITY 15               /* 1056-1067            */
#define MTS_M_VELOCITY 16              /* 1069-1099            */
#define MTS_M_PSEUDOPOTENTIAL 17       /* 515-531              */
#define MTS_M_FORCES 18                /* 533-566              */
#define MTS_M_SOURCE 19                /* 568-630              */
#define MTS_M_VELOCITY_DIFF 20         /* 1248-1267            */
#define MTS_M_DENSITY_DIFF 21          /* 1269-1283            */
#define MTS_M_APPLY_BC 22              /* 814-831              */
#define MTS_M_EVOLUTION 23             /* 1310-1331            */
#define MTS_M_COPY_F_FTEMP 24          /* ftemp = f            */
#define MTS_M_COPY_F_FB 25             /* fb = f               */
#define MTS_M_COPY_USQ_OLD 26          /* usq_old = usq        */
#define MTS_M_COPY_RHO_OLD 27          /* rho_old = rho        */
#define MTS_M_INTERPOLATE 28           /* 1101-1231: interpolateBlockInterface(lvl, lvl + 1) */
int mts_method(mts_sim s, int method, int lvl);

/* Public vectors of LB2D (coreLB.hpp:89-114) in compact order.
 * Population fields are AoS [cell*9 + dir] like the reference. */
#define MTS_F_F 0
#define MTS_F_FTEMP 1
#define MTS_F_FEQ 2
#define MTS_F_FNEQ 3
#define MTS_F_FB 4
#define MTS_F_RHO 5
#define MTS_F_RHO_OLD 6
#define MTS_F_UX 7
#define MTS_F_UY 8
#define MTS_F_USQ 9
#define MTS_F_USQ_OLD 10
#define MTS_F_PSI 11
#define MTS_F_FX 12
#define MTS_F_FY 13
#define MTS_F_TAUS 14
#define MTS_F_KN 15
int mts_field_size(mts_sim s, int field, int64_t* n);
int mts_get_field(mts_sim s, int field, double* out);
int mts_set_field(mts_sim s, int field, const double* in);

/* LB2D::completeSingleGridDomain (coreLB.hpp:1344-1411): ux, uy, rho as
 * nx*ny row-major arrays with solid pixels set to 0. */
int mts_complete_domain(mts_sim s, double* ux, double* uy, double* rho);

/* FP64 permeability report of the current state (no reference counterpart;
 * the flux integration a user of MantiisApp computes from ux.txt): over the
 * nx*ny pixels of completeSingleGridDomain (coreLB.hpp:1344-1411), out8 =
 * {U_D = sum u_x / (nx ny) (Darcy, superficial), mean u_x of the fluid
 * pixels, sum rho u_x, sum rho, porosity, tau = mean tau_s of the finest
 * level's cells, k = nu U_D / Fbody with nu = (tau - 0.5) / 3 (lattice
 * units), k dx^2 [m^2]}; sums in FP64.  k is 0 when Fbody is 0. */
int mts_permeability(mts_sim s, double* out8);

/* FP64 diagnostics of the current state (no reference counterpart):
 * out4 = {sum rho, sum j_x, sum j_y, fluid cell count}. */
int mts_diagnostics(mts_sim s, double* out4);

/* The HIP stream (hipStream_t) on which the handle launches its kernels. */
int mts_get_stream(mts_sim s, void** stream);
int mts_synchronize(mts_sim s);

/* Kernel timer (no reference counterpart): while enabled, every launch of
 * the step's dominant kernel -- the fused FP32/FP64 pass of a single-block
 * step, the first-substep pass of both levels of a quadtree step -- is
 * bracketed by HIP events on the handle's stream.  mts_kernel_timer_read
 * returns the summed event time (ms) and the launch count since the last
 * read and clears them (it synchronizes the stream). */
int mts_kernel_timer(mts_sim s, int enable);
int mts_kernel_timer_read(mts_sim s, double* ms, int64_t* launches);

/* ------------------------------------------------------------------------
 * Batched solver (no reference counterpart): nb independent single-block
 * lattices (torus seams, FP32 deviations, the operators of p) stepped
 * together, one launch per step over a device work queue of the blocks of
 * the lattices that have not converged.  Lattice g's inputs are the row-major
 * nx[g]*ny[g] arrays pore[g], kn[g], locpore[g].  Creation initialises every
 * lattice at rest (f = w rho_in).
 * ---------------------------------------------------------------------- */
typedef struct mts_batch_s* mts_batch;
int mts_batch_create(int nb, const int* nx, const int* ny, const int32_t* const* pore, const double* const* kn,
                     const double* const* locpore, const mts_params* p, mts_batch* out);
int mts_batch_destroy(mts_batch b);
/* fluid cells of every lattice (nb entries) */
int mts_batch_cells(mts_batch b, int64_t* cells);
/* Steps the active lattices in chunks of check_every (even) steps until each
 * one's mean velocity <u_x> = sum(rho u_x) / sum(rho) changes by less than
 * tol per step between two checks, or max_steps have run.  Per lattice:
 * steps to convergence (negative: -steps run, not converged), seconds since
 * the call began, and the last <u_x>; cell_updates = cell updates done.
 * use_graph captures a chunk in a hipGraph and replays it. */
int mts_batch_run(mts_batch b, int64_t max_steps, int check_every, double tol, int use_graph, int64_t* steps,
                  double* seconds, double* ux_mean, int64_t* cell_updates);
/* nsteps steps of the active lattices, one launch each, with no statistics
 * and no convergence test (the fixed-length step of a timing window; it
 * tunes first if the batch was never tuned).  Asynchronous on the batch's
 * stream; mts_batch_fields reports the state of the last statistics step
 * (creation or the last check of mts_batch_run). */
int mts_batch_step(mts_batch b, int nsteps);
/* the batch's step kernel: *steps_per_launch = 2 for the tile kernels, 1 for
 * the one-step cell-order kernel; *units = its tiles (32 x 16 sites) or
 * 256-cell blocks (either pointer may be NULL) */
int mts_batch_kernel(mts_batch b, int* steps_per_launch, int64_t* units);
/* marks lattices active (1) or not (0) for the next runs */
int mts_batch_activate(mts_batch b, const int32_t* active);
/* rho, u_x, u_y (nx*ny row-major, solids 0) of lattice g's post-stream
 * state at the last check */
int mts_batch_fields(mts_batch b, int g, double* rho, double* ux, double* uy);
/* sets lattice g's post-stream populations (9 planes of nx*ny doubles,
 * f[j*nx*ny + site]; solids ignored): a warm start */
int mts_batch_set_state(mts_batch b, int g, const double* f);
/* the same for n lattices g[0..n-1] with populations f[0..n-1], collided in
 * one launch */
int mts_batch_set_states(mts_batch b, int n, const int32_t* g, const double* const* f);
/* warm start from velocity fields (no counterpart in the reference):
 * lattices g[0..n-1] get, on the device, the populations of the equilibrium
 * at rho plus the regularised non-equilibrium part -3 tau[i] w_j rho
 * (Q_j : grad u) of the given u (ux[i], uy[i]: nx*ny row-major doubles,
 * zero in the solids; central periodic differences), collided in one
 * launch like mts_batch_set_states */
int mts_batch_seed(mts_batch b, int n, const int32_t* g, const double* const* ux, const double* const* uy,
                   const double* tau, double rho);
int mts_batch_stream(mts_batch b, void** stream);
/* Autotuner (no reference counterpart): steps_each > 0 times every launch
 * configuration of the batched step -- threads per 256-cell block {64, 128,
 * 256} (each thread then takes 4, 2 or 1 cells) x XCD run length of the
 * queue map {0 (queue order), 8, 32} -- for steps_each steps from the current
 * state, keeps the fastest (of the fastest width's queue maps within 1% of
 * it, the longest XCD run) and puts the state back (the configuration
 * changes no result).  steps_each == 0 takes the default (256, 32) untimed;
 * steps_each < 0 only reports.  mts_batch_run tunes with 8 steps on its first
 * call if this was never called.  Outputs (may be null): the configuration
 * and ms_per_step[9] per candidate (row-major threads x run length, -1: not
 * timed). */
int mts_batch_tune(mts_batch b, int steps_each, int* threads, int* xcd_run, double* ms_per_step);
/* Sets the launch configuration by hand (threads 64 / 128 / 256, XCD run
 * length >= 0) and marks the batch tuned. */
int mts_batch_set_config(mts_batch b, int threads, int xcd_run);

/* ------------------------------------------------------------------------
 * Learned warm start (no reference counterpart): a four-level convolutional
 * encoder-decoder (widths c0, 2c0, 4c0, 8c0; 3x3 periodic convolutions, ReLU,
 * 2x2 average pooling, nearest upsampling with skip concatenation, a 1x1
 * head) from 3 input channels (fluid mask, log10 Kn, forcing) to 2 output
 * channels (scaled velocity), trained with its own BF16 matrix-core forward
 * and backward kernels.  Inputs, targets, masks and outputs are device
 * pointers, NHWC FP32: in [B][H][W][3], out/target [B][H][W][2], mask
 * [B][H][W]; c0, H and W multiples of 8, within the sizes given at creation.
 * ---------------------------------------------------------------------- */
typedef struct mts_net_s* mts_net;
int mts_net_create(int c0, int max_batch, int max_h, int max_w, uint64_t seed, mts_net* out);
int mts_net_destroy(mts_net n);
int mts_net_param_count(mts_net n, int64_t* count);
/* floating-point operations of one forward pass over B images of H x W
 * (2 per multiply-add of every convolution; a training step's backward
 * products are twice that) */
int mts_net_flops(mts_net n, int B, int H, int W, double* fwd);
/* parameters and gradients as one FP32 host array: per layer the weights
 * [3*3 or 1][cout][cin] then the biases [cout] */
int mts_net_get_params(mts_net n, float* host);
int mts_net_set_params(mts_net n, const float* host);
int mts_net_get_grads(mts_net n, float* host);
int mts_net_forward(mts_net n, const float* in, int B, int H, int W, float* out);
/* after mts_net_forward: masked mean-square loss sum m |o - t|^2 / (2 sum m)
 * and the gradient of every parameter */
int mts_net_backward(mts_net n, const float* target, const float* mask, float* loss);
int mts_net_adam(mts_net n, float lr, float beta1, float beta2, float eps, int step);
/* the convolutions' GEMM kernel on dense host matrices (check of the matrix
 * core lane layout): C[M x N] = A[M x K] B[K x N] with BF16 operands */
int mts_net_gemm_check(const float* A, const float* B, int M, int N, int K, float* C);
int mts_net_stream(mts_net n, void** stream);

/* Device memory for the network's tensors and the calibration copy (no
 * reference counterpart): plain hipMalloc / hipFree / synchronous copies. */
int mts_dev_alloc(size_t bytes, void** ptr);
int mts_dev_free(void* ptr);
int mts_dev_upload(void* dst, const void* src, size_t bytes);
int mts_dev_download(void* dst, const void* src, size_t bytes);

/* Calibration copy (no reference counterpart): copies n FP32 values from
 * src to dst (device pointers) with one 4-byte nontemporal load and store
 * per lane, the access width of the lattice kernels, `launches` times on the
 * default stream, then synchronizes.  Used to scale the rocprofv3
 * FETCH_SIZE / WRITE_SIZE counters. */
int mts_probe_copy(float* dst, const float* src, int64_t n, int launches);

/* Bandwidth copy (no reference counterpart): as mts_probe_copy with one
 * 16-byte nontemporal load and store per lane; n a multiple of 4 and 16-byte
 * aligned pointers (MTS_ERR_ARG otherwise).  Its rate is the measured copy
 * ceiling beside the nominal HBM peak. */
int mts_probe_copy16(float* dst, const float* src, int64_t n, int launches);

/* ------------------------------------------------------------------------
 * Geometry preprocessing: replaces Shape::calculateProperties
 * (include/shape.hpp:127-251 of the reference) -- Zhang-Suen skeleton of the
 * domain on a canvas with max(nx, ny) pixels of margin (interior rows run on
 * as fluid), exact distances to the walls and to the skeleton, and per
 * fluid pixel LocPore = 2 x (wall distance at the first skeleton pixel, in
 * row-major order, at the skeleton distance), Kn = mfp / LocPore /
 * resolution.  Wall pixels get 0.  Sides 2..2048 (float distances stay
 * exact).  Errors: MTS_ERR_ARG for values other than 0 / 1,
 * MTS_ERR_GEOMETRY when the skeleton misses the domain (the reference then
 * reads out of bounds, shape.hpp:232-245).
 * ---------------------------------------------------------------------- */
typedef struct mts_shape_s* mts_shape;
/* one-shot form of the reference call: host domain ny x nx row-major int32
 * (Shape::domain), host outputs ny x nx (Shape::LocPore, Shape::Kn) and,
 * if non-null, the cropped skeleton (ny x nx bytes) */
int mts_shape_properties(const int32_t* domain, int nx, int ny, double resolution, double mfp, double* locpore,
                         double* kn, uint8_t* skeleton);
/* device-resident form: a handle holds the work buffers of one image size;
 * mts_shape_run reads a device domain (ny x nx bytes, 0 / 1) and writes
 * device locpore / kn (and skeleton when non-null) on the handle's stream,
 * which it synchronizes.  iterations (may be null): full thinning
 * iterations run on the device, a multiple of 16 no smaller than the
 * reference's count. */
int mts_shape_create(int nx, int ny, mts_shape* out);
int mts_shape_destroy(mts_shape h);
int mts_shape_run(mts_shape h, const uint8_t* domain, double resolution, double mfp, double* locpore, double* kn,
                  uint8_t* skeleton, int* iterations);
int mts_shape_stream(mts_shape h, void** stream);

/* Allocation-free form of mts_shape_run (no reference counterpart: for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture elsewhere in their pipeline, tight serving loops, harnesses that
 * forbid hipMalloc).  Same inputs and outputs as mts_shape_run, given
 * directly (device domain, device outputs, sides nx, ny in [2, 512]); the
 * work buffers live in one static device array of the library (8.3 MB per
 * GPU, held until the process ends) instead of a handle, and the run goes on
 * the caller's stream (NULL: the null stream), which it synchronizes.  It
 * calls no hipMalloc / hipFree.  Sides beyond 512 return MTS_ERR_ARG.  Every
 * call on one GPU uses that GPU's one array: calls that overlap on one GPU
 * (from several host threads or streams) race and must be serialized by the
 * caller. */
int mts_shape_run_static(int nx, int ny, const uint8_t* domain, double resolution, double mfp, double* locpore,
                         double* kn, uint8_t* skeleton, int* iterations, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MANTIIS_HIP_H */
