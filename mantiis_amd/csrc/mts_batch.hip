// Batched solver (C ABI mts_batch_*): many independent single-block
// lattices in one cell space, stepped by one launch per step over a device
// work queue of the blocks of the unconverged lattices, with per-lattice
// convergence on the mean velocity and hipGraph-captured step chunks.
// The per-lattice model is the single-block FP32 transport step of
// mts_api.hip (torus seams).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "mantiis_hip.h"
#include "mts_common.h"
#include "mts_kernels.h"

using namespace mts;

namespace mts {
int set_error(int code, const std::string& msg);  // mts_api.hip
}

#define BT_TRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(e_ == hipErrorOutOfMemory ? MTS_ERR_OOM : MTS_ERR_HIP,        \
                             std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

struct mts_batch_s {
    mts_params prm{};
    Physics ph;
    hipStream_t stream = nullptr;
    int nb = 0;
    std::vector<int> nx, ny;
    std::vector<int64_t> off, ncell;           // cell range of each lattice
    std::vector<std::vector<int32_t>> cellPos;  // per lattice: cell -> site
    std::vector<CellParam> params;
    CompactImage cat;                          // concatenated image (host)
    std::vector<float> cpar_h;
    float* book = nullptr;  // SRBB bookkeeping table of the FP32 parameter records
    CompactDev dev;
    std::vector<BatBlock> hblocks;
    std::vector<std::vector<int32_t>> blocks_of;  // per lattice: its block ids
    float* A[2] = {nullptr, nullptr};
    // MDBB: the collided state of the wall cells beside A[cur] (4 planes)
    float* M[2] = {nullptr, nullptr};
    float* msnap = nullptr;
    int cur = 0;
    float* mac = nullptr;
    double* acc = nullptr;
    double* part = nullptr;      // per block
    int32_t* d_first = nullptr;  // per lattice: first block id (nb + 1)
    BatBlock* d_queue = nullptr;  // the active blocks' records, padded to the grid
    int nactive = 0;
    // tile kernels (two steps per launch): tile-major cell order, per tile
    // row the fluid mask and first cell, per tile its 3 x 3 neighbourhood
    bool tiled = false;
    int ntiles = 0;
    std::vector<std::vector<int32_t>> tiles_of;  // per lattice: tiles holding a cell
    std::vector<int32_t> tfirst;                 // per lattice: first tile (nb + 1)
    int32_t* d_tnb = nullptr;
    uint32_t* d_tmask = nullptr;
    int32_t* d_tpre = nullptr;
    int32_t* d_tqueue = nullptr;
    int32_t* d_tfirst = nullptr;
    int32_t* d_ntactive = nullptr;
    double* tpart = nullptr;  // per tile sums (zero for the tiles without a cell)
    int ntactive = 0;
    std::vector<char> active;
    std::vector<double> u_prev;
    std::vector<void*> allocs;
    // launch configuration of k_cbat (chosen by tune()): threads per
    // 256-cell block and XCD run length of the queue map
    int threads = 256, xrun = 32;
    bool tuned = false;
    std::vector<double> tune_ms;  // per candidate: ms per step (-1: not timed)
    float* snap = nullptr;        // carried state saved while tuning
    // graph of one chunk of steps (replayed while the chunk length and the
    // grid hold)
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int gchunk = 0;
    int gcur = -1;
    int ggrid = -1;

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
    ~mts_batch_s() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* q : allocs) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
    LaunchCtx ctx() const {
        LaunchCtx c{};
        c.stream = stream;
        c.nx = 0;
        c.ny = 0;
        c.stride = 0;
        c.ncell = cat.ncell;
        c.params = nullptr;
        c.fp32 = true;
        c.coll = prm.collision;
        c.wall = prm.wall;
        c.rho0 = ph.rhoin;
        c.rhoin = ph.rhoin;
        c.Fbody = prm.Fbody;
        c.one_param = params.size() == 1;
        c.cimg = &dev;
        return c;
    }
    int upload_queue() {
        // the active blocks' records in queue order, padded with empty
        // entries to the grid of a step
        std::vector<BatBlock> q;
        std::vector<int32_t> tq;
        for (int g = 0; g < nb; g++)
            if (active[g]) {
                for (int32_t bid : blocks_of[g]) q.push_back(hblocks[bid]);
                if (tiled) tq.insert(tq.end(), tiles_of[g].begin(), tiles_of[g].end());
            }
        nactive = (int)q.size();
        q.resize((size_t)grid_for(nactive), BatBlock{0, 0, -1, -1, -1});
        if (!q.empty())
            BT_TRY(hipMemcpyAsync(d_queue, q.data(), q.size() * sizeof(BatBlock), hipMemcpyHostToDevice, stream));
        if (tiled) {
            ntactive = (int)tq.size();
            if (!tq.empty()) BT_TRY(hipMemcpyAsync(d_tqueue, tq.data(), tq.size() * 4, hipMemcpyHostToDevice, stream));
            BT_TRY(hipMemcpyAsync(d_ntactive, &ntactive, 4, hipMemcpyHostToDevice, stream));
        }
        return hipStreamSynchronize(stream) == hipSuccess ? MTS_OK : set_error(MTS_ERR_HIP, "queue upload");
    }
    // Grid of a step: the active blocks rounded up to a power of two (at
    // least 1024), so that a shrinking queue is not launched over every
    // block of the batch (the queue's padding entries exit at once, but
    // hundreds of thousands of them still cost their dispatch), while the
    // captured chunk is replaced only when the bucket changes.
    int grid_for(int n) const {
        if (n <= 0) return 0;
        int g = 1024;
        while (g < n) g *= 2;
        return std::min<int>(g, (int)hblocks.size());
    }
    int tgrid_for(int n) const {
        if (n <= 0) return 0;
        int g = 1024;
        while (g < n) g *= 2;
        return std::min<int>(g, ntiles);
    }
    int grid_key() const { return tiled ? tgrid_for(ntactive) : grid_for(nactive); }
    // tile kernels: one launch of ns (1 or 2) steps over the tile queue,
    // A[cur] -> A[cur^1]; with stats the last post-stream state's fields and
    // per-lattice sums
    int advance(int ns, bool stats) {
        TileLaunch T;
        T.stream = stream;
        T.coll = prm.collision;
        T.wall = prm.wall;
        T.nsteps = ns;
        T.grid = tgrid_for(ntactive);
        T.Ain = A[cur];
        T.Aout = A[cur ^ 1];
        T.tw = dev.cw;
        T.pidx = dev.pidx;
        T.cpar = dev.cpar;
        T.book = dev.book;
        T.tmask = d_tmask;
        T.tpre = d_tpre;
        T.tnb = d_tnb;
        T.queue = d_tqueue;
        T.nactive = d_ntactive;
        T.mac = stats ? mac : nullptr;
        T.part = stats ? tpart : nullptr;
        T.S = dev.cstride;
        T.rho0 = ph.rhoin;
        T.Fbody = (float)prm.Fbody;
        T.xrun = xrun;
        T.one = params.size() == 1;
        const int rc = launch_ctile(T);
        if (stats) launch_bsum(tpart, d_tfirst, acc, nb, stream);
        cur ^= 1;
        return rc;
    }
    // nsteps plain steps (no statistics)
    int steps(int nsteps) {
        int rc;
        if (tiled) {
            for (int i = 0; i < nsteps / 2; i++)
                if ((rc = advance(2, false))) return rc;
            if (nsteps & 1) return advance(1, false);
            return MTS_OK;
        }
        for (int i = 0; i < nsteps; i++)
            if ((rc = step(false, 0))) return rc;
        return MTS_OK;
    }
    // one step over the queue: A[cur] -> A[cur^1]
    int step(bool stats, int from) {
        BatchLaunch L;
        L.c = ctx();
        L.Ain = A[cur];
        L.Aout = A[cur ^ 1];
        L.queue = d_queue;
        L.maxblocks = grid_for(nactive);
        L.mac = stats ? mac : nullptr;
        L.part = stats ? part : nullptr;
        L.from_rest = from;
        L.threads = threads;
        L.xrun = xrun;
        L.Min = M[cur];
        L.Mout = M[cur ^ 1];
        if (from == 2) {  // the given state sits in A[cur^1]: read it, write A[cur]
            L.Ain = A[cur ^ 1];
            L.Aout = A[cur];
            L.Mout = M[cur];
        }
        const int rc = launch_cbatch(L);
        if (stats) launch_bsum(part, d_first, acc, nb, stream);
        if (from != 2) cur ^= 1;
        return rc;
    }
    // n steps (n even), the last with statistics
    int chunk_eager(int n) {
        int rc;
        if (tiled) {
            for (int i = 0; i < n / 2; i++)
                if ((rc = advance(2, i == n / 2 - 1))) return rc;
            return MTS_OK;
        }
        for (int i = 0; i < n; i++)
            if ((rc = step(i == n - 1, 0))) return rc;
        return MTS_OK;
    }
    // Autotuner: times every launch configuration (threads per block x XCD
    // run length) for `each` steps from the current state and keeps the
    // fastest.  The configuration changes no result, and the state is saved
    // before and put back after, so tuning leaves the simulation where it
    // was.  each == 0 takes the default configuration without timing.
    static constexpr int NCAND = 9;
    static int cand_threads(int k) { return k / 3 == 0 ? 64 : (k / 3 == 1 ? 128 : 256); }
    static int cand_xrun(int k) { return k % 3 == 0 ? 0 : (k % 3 == 1 ? 8 : 32); }
    int tune(int each) {
        tune_ms.assign(NCAND, -1.0);
        tuned = true;
        if (each <= 0 || nactive == 0) {
            threads = 256;
            xrun = 32;
            return MTS_OK;
        }
        each += each & 1;
        const size_t elems = (size_t)9 * dev.cstride;
        if (!snap) {
            if (int rc = alloc(&snap, elems)) return rc;
        }
        const int cur0 = cur;
        const size_t melems = (size_t)4 * dev.cstride;
        if (M[0] && !msnap) {
            if (int rc = alloc(&msnap, melems)) return rc;
        }
        BT_TRY(hipMemcpyAsync(snap, A[cur0], elems * sizeof(float), hipMemcpyDeviceToDevice, stream));
        if (M[0]) BT_TRY(hipMemcpyAsync(msnap, M[cur0], melems * sizeof(float), hipMemcpyDeviceToDevice, stream));
        // puts the saved state back (also after a failed step: tuning never
        // leaves the batch partly advanced)
        auto restore = [&]() -> int {
            cur = cur0;
            if (hipMemcpyAsync(A[cur0], snap, elems * sizeof(float), hipMemcpyDeviceToDevice, stream) != hipSuccess ||
                (M[0] && hipMemcpyAsync(M[cur0], msnap, melems * sizeof(float), hipMemcpyDeviceToDevice, stream) !=
                             hipSuccess) ||
                hipStreamSynchronize(stream) != hipSuccess)
                return set_error(MTS_ERR_HIP, "tune: restoring the state");
            return MTS_OK;
        };
        hipEvent_t e0 = nullptr, e1 = nullptr;
        int best = -1;
        int rc = MTS_OK;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
            rc = set_error(MTS_ERR_HIP, "tune: event creation");
        for (int k = 0; k < NCAND && rc == MTS_OK; k++) {
            threads = cand_threads(k);
            xrun = cand_xrun(k);
            if (tiled && threads != 256) continue;  // the tile kernels run 256 threads per tile
            rc = steps(2);                          // first use of the variant
            if (rc) break;
            if (hipEventRecord(e0, stream) != hipSuccess) rc = set_error(MTS_ERR_HIP, "tune: event");
            rc = steps(each);
            if (rc) break;
            float ms = 0.0f;
            if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = set_error(MTS_ERR_HIP, "tune: event timing");
                break;
            }
            tune_ms[k] = (double)ms / each;
            if (best < 0 || tune_ms[k] < tune_ms[best]) best = k;
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        const int rrc = restore();
        if (rc || rrc || best < 0) {
            threads = 256;
            xrun = 32;
            return rc ? rc : rrc;
        }
        // the short timings do not separate the queue maps of one launch width
        // within about 1%: of those that close to the fastest, take the
        // longest XCD run (the fewest bytes fetched from memory)
        const int b0 = best;
        for (int k = 0; k < NCAND; k++)
            if (tune_ms[k] > 0.0 && cand_threads(k) == cand_threads(b0) && cand_xrun(k) > cand_xrun(best) &&
                tune_ms[k] <= 1.01 * tune_ms[b0])
                best = k;
        threads = cand_threads(best);
        xrun = cand_xrun(best);
        // a captured chunk holds the old configuration
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        return MTS_OK;
    }
    int chunk_graph(int n) {
        if (!gexec || gchunk != n || gcur != cur || ggrid != grid_key()) {
            if (gexec) (void)hipGraphExecDestroy(gexec);
            if (graph) (void)hipGraphDestroy(graph);
            gexec = nullptr;
            graph = nullptr;
            const int c0 = cur;
            BT_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            const int rc = chunk_eager(n);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(stream, &g);
            cur = c0;  // capture ran nothing, whatever chunk_eager flipped
            if (rc || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc ? rc : set_error(MTS_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            }
            graph = g;
            if (hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(graph);
                graph = nullptr;
                gexec = nullptr;
                return set_error(MTS_ERR_HIP, "graph instantiation");
            }
            gchunk = n;
            gcur = c0;
            ggrid = grid_key();
        }
        BT_TRY(hipGraphLaunch(gexec, stream));
        const int flips = tiled ? n / 2 : n;  // launches in the chunk
        if (flips & 1) cur ^= 1;
        return MTS_OK;
    }
};

namespace {

int check(mts_batch b) { return b ? MTS_OK : set_error(MTS_ERR_ARG, "null batch handle"); }

}  // namespace

extern "C" {

int mts_batch_create(int nb, const int* nx, const int* ny, const int32_t* const* pore, const double* const* kn,
                     const double* const* locpore, const mts_params* p, mts_batch* out) {
    if (!out || nb <= 0 || !nx || !ny || !pore || !kn || !locpore) return set_error(MTS_ERR_ARG, "bad argument");
    *out = nullptr;
    mts_params prm;
    if (p) prm = *p;
    else mts_default_params(&prm);
    if (prm.precision != MTS_FP32DEV) return set_error(MTS_ERR_UNSUPPORTED, "the batched solver runs FP32 deviations");
    if (prm.multiblock) return set_error(MTS_ERR_UNSUPPORTED, "the batched solver runs single-block lattices");
    if (prm.collision < 0 || prm.collision > 2 || prm.wall < 0 || prm.wall > 2)
        return set_error(MTS_ERR_ARG, "operator combination has no batched kernel");
    mts_batch_s* b = new (std::nothrow) mts_batch_s();
    if (!b) return set_error(MTS_ERR_OOM, "host allocation");
    auto bail = [&](int rc) {
        delete b;
        return rc;
    };
    b->prm = prm;
    b->prm.periodic_y = 1;
    b->ph.Cl = prm.dx;
    b->ph.Fbody = prm.Fbody;
    b->ph.T_phy = prm.T;
    b->ph.P_phy = prm.P;
    compute_physics(b->ph);
    if (!b->ph.cubic_ok) return bail(set_error(MTS_ERR_PHYSICS, "PR-EOS: no single positive real root at this (T, P)"));
    b->nb = nb;
    b->nx.assign(nx, nx + nb);
    b->ny.assign(ny, ny + nb);
    b->off.resize(nb);
    b->ncell.resize(nb);
    b->cellPos.resize(nb);
    b->blocks_of.resize(nb);
    b->tiles_of.resize(nb);
    b->tfirst.assign(nb + 1, 0);
    // tile kernels when asked for and every side is a multiple of the tile
    b->tiled = prm.batch_steps == 2 && prm.wall != MTS_WALL_MDBB;  // the tile kernels have no MDBB
    for (int g = 0; g < nb && b->tiled; g++)
        if (nx[g] % TILE_X || ny[g] % TILE_Y) b->tiled = false;
    std::vector<int32_t> htnb, htpre;
    std::vector<uint32_t> htmask;
    CompactImage& C = b->cat;
    int64_t base = 0;
    // The lattices' grids and cell images are independent: host threads
    // build them a chunk at a time, then they are appended in lattice order.
    struct Built {
        Geometry geo;
        CompactImage ci;
        int rc = MTS_OK;
        std::string err;
    };
    const int nthr = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    const int chunk = 4 * nthr;
    std::vector<Built> built;
    for (int g = 0; g < nb; g++) {
        if (g % chunk == 0) {
            const int g0 = g, g1 = std::min(nb, g + chunk);
            built.clear();
            built.resize(g1 - g0);
            std::atomic<int> next{g0};
            auto work = [&]() {
                for (int h; (h = next.fetch_add(1)) < g1;) {
                    Built& B = built[h - g0];
                    if (nx[h] <= 0 || ny[h] <= 0 || !pore[h] || !kn[h] || !locpore[h]) {
                        B.rc = MTS_ERR_ARG;
                        B.err = "bad geometry";
                        continue;
                    }
                    B.rc = build_single_grid(nx[h], ny[h], pore[h], kn[h], locpore[h], b->ph, true, B.geo, B.err);
                    if (B.rc == MTS_OK) build_compact(B.geo, B.ci);
                }
            };
            std::vector<std::thread> pool;
            for (int i = 1; i < std::min(nthr, g1 - g0); i++) pool.emplace_back(work);
            work();
            for (auto& t : pool) t.join();
        }
        Built& B = built[g % chunk];
        if (B.rc) return bail(set_error(B.rc, "lattice " + std::to_string(g) + ": " + B.err));
        const Geometry& geo = B.geo;
        const CompactImage& ci = B.ci;
        const int32_t poff = (int32_t)b->params.size();
        if (b->params.size() + geo.params.size() > 65536)
            return bail(set_error(MTS_ERR_UNSUPPORTED, "more than 65536 distinct (tau_s, Kn) pairs in the batch"));
        b->params.insert(b->params.end(), geo.params.begin(), geo.params.end());
        const int64_t n = ci.ncell, naux0 = (int64_t)C.aux.size() / AUXW;
        if (base + ((n + 255) / 256) * 256 > ((int64_t)1 << 31) - 256)
            return bail(set_error(MTS_ERR_UNSUPPORTED, "more than 2^31 cells in the batch"));
        b->off[g] = base;
        b->ncell[g] = n;
        b->cellPos[g] = ci.cellPos;
        C.cw.resize(base + n);
        C.rw.resize(base + n);
        C.pidx.resize(base + n);
        C.outm.resize(base + n);
        if (b->tiled) {
            // tile-major order: tiles in row-major tile order, in a tile its
            // fluid sites in row-major order; per cell the link bits, the
            // boundary type and whether the one-step pass applies the wall
            // rule in link order
            constexpr int TX = TILE_X, TY = TILE_Y;
            const int ntx = nx[g] / TX, nty = ny[g] / TY, t0 = b->ntiles;
            b->tfirst[g] = t0;
            std::vector<int32_t> pos(n);
            int64_t k = 0;
            constexpr uint32_t keep = 0xffu | (0xffu << CW_WALL_SHIFT) | (63u << CW_TYPE_SHIFT) | CW_COMPLEX;
            for (int ty = 0; ty < nty; ty++)
                for (int tx = 0; tx < ntx; tx++) {
                    const int t = t0 + ty * ntx + tx;
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++)
                            htnb.push_back(t0 + ((ty + dy + nty) % nty) * ntx + (tx + dx + ntx) % ntx);
                    const int64_t k0 = k;
                    for (int ly = 0; ly < TY; ly++) {
                        uint32_t m = 0;
                        htpre.push_back((int32_t)(base + k));
                        for (int lx = 0; lx < TX; lx++) {
                            const int64_t p = (int64_t)(ty * TY + ly) * nx[g] + tx * TX + lx;
                            const int32_t r = ci.cidx[p];
                            if (r < 0) continue;
                            m |= 1u << lx;
                            C.cw[base + k] = ci.cw[r] & keep;
                            C.pidx[base + k] = (uint16_t)(ci.pidx[r] + poff);
                            C.rw[base + k] = 0u;
                            C.outm[base + k] = 0;
                            pos[k] = (int32_t)p;
                            k++;
                        }
                        htmask.push_back(m);
                    }
                    if (k > k0) b->tiles_of[g].push_back(t);
                }
            if (k != n) return bail(set_error(MTS_ERR_STATE, "lattice " + std::to_string(g) + ": tile order lost cells"));
            b->ntiles = t0 + ntx * nty;
            b->cellPos[g] = pos;
        }
        for (int64_t c = 0; c < n && !b->tiled; c++) {
            const uint32_t cw = ci.cw[c];
            C.cw[base + c] = cw;
            C.pidx[base + c] = (uint16_t)(ci.pidx[c] + poff);
            C.outm[base + c] = ci.outm[c];
            // rank offsets are relative to the cell; aux records move with the
            // lattice's first record
            C.rw[base + c] = (cw & CW_COMPLEX) ? (uint32_t)(ci.rw[c] + naux0) : ci.rw[c];
        }
        for (int64_t k = 0; k < ci.naux && !b->tiled; k++) {
            const int32_t* r = &ci.aux[k * AUXW];
            for (int j = 0; j < 8; j++) C.aux.push_back(r[j] >= 0 ? (int32_t)(r[j] + base) : -1);
            for (int j = 8; j < AUXW; j++) C.aux.push_back(r[j]);
        }
        for (int64_t s0 = 0; s0 < n; s0 += 256) {
            const int64_t s1 = std::min<int64_t>(s0 + 256, n);
            const BatBlock B{(int32_t)(base + s0), (int32_t)(base + s1), g, (int32_t)b->hblocks.size(),
                             geo.params.size() == 1 ? poff : -1};
            b->blocks_of[g].push_back(B.id);
            b->hblocks.push_back(B);
        }
        base += ((n + 255) / 256) * 256;
        // padding cells between lattices are never in a block
        C.cw.resize(base, 0);
        C.rw.resize(base, 0u);
        C.pidx.resize(base, 0);
        C.outm.resize(base, 0);
        if (g == 0) canonical_spec(C.spec);
        B = Built();  // release the lattice's host tables
    }
    b->tfirst[nb] = b->ntiles;
    C.ncell = base;
    C.cstride = base;
    C.naux = (int64_t)C.aux.size() / AUXW;
    C.cellPos.assign(1, 0);
    C.cidx.assign(1, -1);
    C.segs.assign(1, 0);
    C.ntx = 0;
    hipError_t e = hipSetDevice(prm.device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(set_error(MTS_ERR_HIP, std::string("stream: ") + hipGetErrorString(e)));
    int rc;
    uint32_t* cw;
    uint32_t* rw;
    int32_t* aux;
    uint16_t* pidx;
    uint8_t* outm;
    float* cpar;
    std::vector<float> book;
    fp32_params(b->params, b->ph.rhoin, b->cpar_h, book);
    const size_t nA = (size_t)base * 9;
    if ((rc = b->alloc(&cw, base)) || (rc = b->alloc(&rw, base)) || (rc = b->alloc(&pidx, base)) ||
        (rc = b->alloc(&outm, base)) || (rc = b->alloc(&aux, std::max<size_t>(C.aux.size(), 1))) ||
        (rc = b->alloc(&cpar, b->cpar_h.size())) || (rc = b->alloc(&b->book, book.size())) ||
        (rc = b->alloc(&b->A[0], nA)) || (rc = b->alloc(&b->A[1], nA)) ||
        (rc = b->alloc(&b->mac, (size_t)base * 3)) || (rc = b->alloc(&b->acc, 2 * (size_t)nb)) ||
        (prm.wall == MTS_WALL_MDBB && ((rc = b->alloc(&b->M[0], (size_t)base * 4)) ||
                                       (rc = b->alloc(&b->M[1], (size_t)base * 4)))) ||
        (rc = b->alloc(&b->part, 2 * b->hblocks.size())) ||
        (rc = b->alloc(&b->d_first, (size_t)nb + 1)) ||
        (rc = b->alloc(&b->d_queue, b->hblocks.size())))
        return bail(rc);
    e = hipMemcpyAsync(cw, C.cw.data(), base * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rw, C.rw.data(), base * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pidx, C.pidx.data(), base * 2, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(outm, C.outm.data(), base, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess && !C.aux.empty())
        e = hipMemcpyAsync(aux, C.aux.data(), C.aux.size() * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(cpar, b->cpar_h.data(), b->cpar_h.size() * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->book, book.data(), book.size() * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->mac, 0, (size_t)base * 3 * 4, b->stream);
    std::vector<int32_t> first(nb + 1);
    for (int g = 0; g < nb; g++) first[g] = b->blocks_of[g].empty() ? 0 : b->blocks_of[g][0];
    first[nb] = (int32_t)b->hblocks.size();
    for (int g = nb - 1; g >= 0; g--)
        if (b->blocks_of[g].empty()) first[g] = first[g + 1];
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_first, first.data(), (nb + 1) * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->part, 0, 2 * b->hblocks.size() * 8, b->stream);
    if (e == hipSuccess && b->tiled) {
        if ((rc = b->alloc(&b->d_tnb, htnb.size())) || (rc = b->alloc(&b->d_tmask, htmask.size())) ||
            (rc = b->alloc(&b->d_tpre, htpre.size())) || (rc = b->alloc(&b->d_tqueue, (size_t)b->ntiles)) ||
            (rc = b->alloc(&b->d_tfirst, (size_t)nb + 1)) || (rc = b->alloc(&b->d_ntactive, 1)) ||
            (rc = b->alloc(&b->tpart, 2 * (size_t)b->ntiles)))
            return bail(rc);
        e = hipMemsetAsync(b->tpart, 0, 2 * (size_t)b->ntiles * 8, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(b->d_tnb, htnb.data(), htnb.size() * 4, hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(b->d_tmask, htmask.data(), htmask.size() * 4, hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(b->d_tpre, htpre.data(), htpre.size() * 4, hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->d_tfirst, b->tfirst.data(), (nb + 1) * 4, hipMemcpyHostToDevice, b->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bail(set_error(MTS_ERR_HIP, std::string("upload: ") + hipGetErrorString(e)));
    upload_spec(C.spec, b->stream);
    upload_icsr(b->stream);
    b->dev.cw = cw;
    b->dev.rw = rw;
    b->dev.pidx = pidx;
    b->dev.outm = outm;
    b->dev.aux = aux;
    b->dev.cpar = cpar;
    b->dev.book = b->book;
    b->dev.ncell = base;
    b->dev.cstride = base;
    // drop the host copies of what now lives on the device
    C.cw = std::vector<uint32_t>();
    C.rw = std::vector<uint32_t>();
    C.pidx = std::vector<uint16_t>();
    C.outm = std::vector<uint8_t>();
    C.aux = std::vector<int32_t>();
    b->active.assign(nb, 1);
    b->u_prev.assign(nb, 0.0);
    if ((rc = b->upload_queue())) return bail(rc);
    if ((rc = b->step(true, 1))) return bail(rc);  // rest equilibrium -> first post-collision state
    e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return bail(set_error(MTS_ERR_HIP, std::string("init: ") + hipGetErrorString(e)));
    *out = b;
    return MTS_OK;
}

int mts_batch_destroy(mts_batch b) {
    delete b;
    return MTS_OK;
}

int mts_batch_cells(mts_batch b, int64_t* cells) {
    if (int rc = check(b)) return rc;
    if (!cells) return set_error(MTS_ERR_ARG, "null argument");
    std::copy(b->ncell.begin(), b->ncell.end(), cells);
    return MTS_OK;
}

int mts_batch_run(mts_batch b, int64_t max_steps, int check_every, double tol, int use_graph, int64_t* steps,
                  double* seconds, double* ux_mean, int64_t* cell_updates) {
    if (int rc = check(b)) return rc;
    if (max_steps <= 0 || check_every < 2 || check_every % 2 || !(tol > 0))
        return set_error(MTS_ERR_ARG, "max_steps > 0, even check_every >= 2 and tol > 0 are required");
    const int nb = b->nb;
    std::vector<int64_t> st(nb, -1);
    std::vector<double> sec(nb, -1.0), um(nb, 0.0), hacc(2 * nb);
    int64_t upd = 0, done = 0;
    int rc;
    // baseline mean velocity of the current state
    for (int g = 0; g < nb; g++)
        if (!b->active[g]) st[g] = 0;
    if (!b->tuned && (rc = b->tune(8))) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    bool first = true;
    while (done < max_steps && b->nactive > 0) {
        const int n = check_every;
        if ((rc = use_graph ? b->chunk_graph(n) : b->chunk_eager(n))) return rc;
        BT_TRY(hipMemcpyAsync(hacc.data(), b->acc, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        BT_TRY(hipStreamSynchronize(b->stream));
        done += n;
        int64_t act_cells = 0;
        for (int g = 0; g < nb; g++)
            if (b->active[g]) act_cells += b->ncell[g];
        upd += act_cells * n;
        const double tnow = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        bool changed = false;
        for (int g = 0; g < nb; g++) {
            if (!b->active[g]) continue;
            const double u = hacc[2 * g + 1] / hacc[2 * g];
            um[g] = u;
            if (!std::isfinite(u)) return set_error(MTS_ERR_NONFINITE, "lattice " + std::to_string(g) + ": non-finite");
            const double crit = u != 0.0 ? std::fabs(u - b->u_prev[g]) / (n * std::fabs(u)) : 1.0;
            b->u_prev[g] = u;
            if (!first && crit < tol) {
                b->active[g] = 0;
                st[g] = done;
                sec[g] = tnow;
                changed = true;
            }
        }
        first = false;
        if (changed && (rc = b->upload_queue())) return rc;
    }
    for (int g = 0; g < nb; g++) {
        if (st[g] < 0) st[g] = -done;  // not converged within max_steps
        if (steps) steps[g] = st[g];
        if (seconds) seconds[g] = sec[g];
        if (ux_mean) ux_mean[g] = b->active[g] || st[g] > 0 ? um[g] : b->u_prev[g];
    }
    if (cell_updates) *cell_updates = upd;
    return MTS_OK;
}

int mts_batch_kernel(mts_batch b, int* steps_per_launch, int64_t* units) {
    if (int rc = check(b)) return rc;
    if (steps_per_launch) *steps_per_launch = b->tiled ? 2 : 1;
    if (units) *units = b->tiled ? (int64_t)b->ntiles : (int64_t)b->hblocks.size();
    return MTS_OK;
}

int mts_batch_step(mts_batch b, int nsteps) {
    if (int rc = check(b)) return rc;
    if (nsteps <= 0) return set_error(MTS_ERR_ARG, "nsteps must be positive");
    if (!b->tuned) {
        if (int rc = b->tune(8)) return rc;
    }
    if (int rc = b->steps(nsteps)) return rc;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(MTS_ERR_HIP, std::string("step: ") + hipGetErrorString(e));
    return MTS_OK;
}

int mts_batch_tune(mts_batch b, int steps_each, int* threads, int* xcd_run, double* ms_per_step) {
    if (int rc = check(b)) return rc;
    if (steps_each >= 0) {
        if (int rc = b->tune(steps_each)) return rc;
    }
    if (threads) *threads = b->threads;
    if (xcd_run) *xcd_run = b->xrun;
    if (ms_per_step)
        for (int k = 0; k < mts_batch_s::NCAND; k++)
            ms_per_step[k] = k < (int)b->tune_ms.size() ? b->tune_ms[k] : -1.0;
    return MTS_OK;
}

int mts_batch_set_config(mts_batch b, int threads, int xcd_run) {
    if (int rc = check(b)) return rc;
    if ((threads != 64 && threads != 128 && threads != 256) || xcd_run < 0)
        return set_error(MTS_ERR_ARG, "threads must be 64, 128 or 256 and xcd_run >= 0");
    b->threads = threads;
    b->xrun = xcd_run;
    b->tuned = true;
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    return MTS_OK;
}

int mts_batch_activate(mts_batch b, const int32_t* active) {
    if (int rc = check(b)) return rc;
    if (!active) return set_error(MTS_ERR_ARG, "null argument");
    for (int g = 0; g < b->nb; g++) {
        b->active[g] = active[g] ? 1 : 0;
        b->u_prev[g] = 0.0;
    }
    return b->upload_queue();
}

int mts_batch_fields(mts_batch b, int g, double* rho, double* ux, double* uy) {
    if (int rc = check(b)) return rc;
    if (g < 0 || g >= b->nb || !rho || !ux || !uy) return set_error(MTS_ERR_ARG, "bad argument");
    const int64_t n = b->ncell[g], o = b->off[g], S = b->dev.cstride;
    std::vector<float> h(3 * n);
    for (int k = 0; k < 3; k++)
        BT_TRY(hipMemcpyAsync(&h[k * n], b->mac + k * S + o, n * 4, hipMemcpyDeviceToHost, b->stream));
    BT_TRY(hipStreamSynchronize(b->stream));
    const int64_t P = (int64_t)b->nx[g] * b->ny[g];
    std::fill(rho, rho + P, 0.0);
    std::fill(ux, ux + P, 0.0);
    std::fill(uy, uy + P, 0.0);
    const auto& cp = b->cellPos[g];
    for (int64_t c = 0; c < n; c++) {
        rho[cp[c]] = b->ph.rhoin + (double)h[c];
        ux[cp[c]] = h[n + c];
        uy[cp[c]] = h[2 * n + c];
    }
    return MTS_OK;
}

int mts_batch_set_states(mts_batch b, int n, const int32_t* gs, const double* const* f) {
    if (int rc = check(b)) return rc;
    if (n <= 0 || !gs || !f) return set_error(MTS_ERR_ARG, "bad argument");
    for (int i = 0; i < n; i++)
        if (gs[i] < 0 || gs[i] >= b->nb || !f[i]) return set_error(MTS_ERR_ARG, "bad argument");
    static const double w[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
    const int64_t S = b->dev.cstride;
    // the given post-stream states go to the other buffer (deviations from
    // w rho0 in cell order, converted on host threads); the lattices' blocks
    // alone then collide them into the carried state in one launch
    std::vector<std::vector<float>> h(n);
    {
        std::atomic<int> next{0};
        auto work = [&]() {
            for (int i; (i = next.fetch_add(1)) < n;) {
                const int g = gs[i];
                const int64_t nc = b->ncell[g], P = (int64_t)b->nx[g] * b->ny[g];
                const auto& cp = b->cellPos[g];
                h[i].resize(9 * nc);
                for (int64_t c = 0; c < nc; c++)
                    for (int j = 0; j < 9; j++) h[i][j * nc + c] = (float)(f[i][(int64_t)j * P + cp[c]] - w[j] * b->ph.rhoin);
            }
        };
        const int nthr = std::max(1, std::min({16, n, (int)std::thread::hardware_concurrency()}));
        std::vector<std::thread> pool;
        for (int t = 1; t < nthr; t++) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
    }
    for (int i = 0; i < n; i++) {
        const int g = gs[i];
        const int64_t nc = b->ncell[g], o = b->off[g];
        for (int j = 0; j < 9; j++)
            BT_TRY(hipMemcpyAsync(b->A[b->cur ^ 1] + j * S + o, &h[i][j * nc], nc * 4, hipMemcpyHostToDevice, b->stream));
    }
    std::vector<char> keep = b->active;
    std::fill(b->active.begin(), b->active.end(), 0);
    for (int i = 0; i < n; i++) b->active[gs[i]] = 1;
    int rc;
    if ((rc = b->upload_queue())) return rc;
    if ((rc = b->step(true, 2))) return rc;
    b->active = keep;
    for (int i = 0; i < n; i++) b->u_prev[gs[i]] = 0.0;
    if ((rc = b->upload_queue())) return rc;
    BT_TRY(hipStreamSynchronize(b->stream));  // the host copies go out of scope
    return MTS_OK;
}

namespace {

// Warm-start populations of one lattice from a velocity field (row-major
// site images, zero in the solids): equilibrium at rho plus the regularised
// non-equilibrium part -3 tau w_i rho (Q_i : grad u) with central periodic
// differences, written as deviations from w_i rho0 into the lattice's cells
// of the 9 planes at A (plane stride S).
__global__ void k_seed(int nx, int ny, int64_t ncell, const int32_t* __restrict__ cellPos, const double* __restrict__ ux,
                       const double* __restrict__ uy, double tau, double rho, double rho0, float* __restrict__ A,
                       int64_t S) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int p = cellPos[c], x = p % nx, y = p / nx;
    const int xp = x + 1 == nx ? 0 : x + 1, xm = x == 0 ? nx - 1 : x - 1;
    const int yp = y + 1 == ny ? 0 : y + 1, ym = y == 0 ? ny - 1 : y - 1;
    const double u = ux[p], v = uy[p];
    const double dxux = (ux[(int64_t)y * nx + xp] - ux[(int64_t)y * nx + xm]) / 2;
    const double dyuy = (uy[(int64_t)yp * nx + x] - uy[(int64_t)ym * nx + x]) / 2;
    const double dxuy = (uy[(int64_t)y * nx + xp] - uy[(int64_t)y * nx + xm]) / 2;
    const double dyux = (ux[(int64_t)yp * nx + x] - ux[(int64_t)ym * nx + x]) / 2;
    const double usq = u * u + v * v;
    const int ex[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1}, ey[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    const double w[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
    for (int j = 0; j < 9; j++) {
        const double cu = ex[j] * u + ey[j] * v;
        const double feq = w[j] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq);
        const double q = (ex[j] * ex[j] - 1.0 / 3) * dxux + (ey[j] * ey[j] - 1.0 / 3) * dyuy + ex[j] * ey[j] * (dxuy + dyux);
        const double f = feq - 3 * tau * w[j] * rho * q;
        A[j * S + c] = (float)(f - w[j] * rho0);
    }
}

}  // namespace

int mts_batch_seed(mts_batch b, int n, const int32_t* gs, const double* const* ux, const double* const* uy,
                   const double* tau, double rho) {
    if (int rc = check(b)) return rc;
    if (n <= 0 || !gs || !ux || !uy || !tau) return set_error(MTS_ERR_ARG, "bad argument");
    for (int i = 0; i < n; i++)
        if (gs[i] < 0 || gs[i] >= b->nb || !ux[i] || !uy[i]) return set_error(MTS_ERR_ARG, "bad argument");
    // one staging buffer for the largest lattice: velocities and cell sites
    int64_t Pmax = 1, nmax = 1;
    for (int i = 0; i < n; i++) {
        Pmax = std::max<int64_t>(Pmax, (int64_t)b->nx[gs[i]] * b->ny[gs[i]]);
        nmax = std::max<int64_t>(nmax, b->ncell[gs[i]]);
    }
    double *d_u = nullptr, *d_v = nullptr;
    int32_t* d_pos = nullptr;
    hipError_t e = hipMalloc(&d_u, Pmax * 8);
    if (e == hipSuccess) e = hipMalloc(&d_v, Pmax * 8);
    if (e == hipSuccess) e = hipMalloc(&d_pos, nmax * 4);
    float* dst = b->A[b->cur ^ 1];
    for (int i = 0; i < n && e == hipSuccess; i++) {
        const int g = gs[i];
        const int64_t P = (int64_t)b->nx[g] * b->ny[g], nc = b->ncell[g];
        e = hipMemcpyAsync(d_u, ux[i], P * 8, hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_v, uy[i], P * 8, hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_pos, b->cellPos[g].data(), nc * 4, hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess) {
            k_seed<<<(unsigned)((nc + 255) / 256), 256, 0, b->stream>>>(b->nx[g], b->ny[g], nc, d_pos, d_u, d_v, tau[i],
                                                                          rho, b->ph.rhoin, dst + b->off[g],
                                                                          b->dev.cstride);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (d_u) (void)hipFree(d_u);
    if (d_v) (void)hipFree(d_v);
    if (d_pos) (void)hipFree(d_pos);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(e == hipErrorOutOfMemory ? MTS_ERR_OOM : MTS_ERR_HIP,
                         std::string("warm-start seeding: ") + hipGetErrorString(e));
    }
    // the seeded post-stream states are collided in one launch
    std::vector<char> keep = b->active;
    std::fill(b->active.begin(), b->active.end(), 0);
    for (int i = 0; i < n; i++) b->active[gs[i]] = 1;
    int rc;
    if ((rc = b->upload_queue())) return rc;
    if ((rc = b->step(true, 2))) return rc;
    b->active = keep;
    for (int i = 0; i < n; i++) b->u_prev[gs[i]] = 0.0;
    if ((rc = b->upload_queue())) return rc;
    BT_TRY(hipStreamSynchronize(b->stream));
    return MTS_OK;
}

int mts_batch_set_state(mts_batch b, int g, const double* f) {
    const int32_t gs[1] = {g};
    const double* fs[1] = {f};
    return mts_batch_set_states(b, 1, gs, fs);
}

int mts_batch_stream(mts_batch b, void** stream) {
    if (int rc = check(b)) return rc;
    if (!stream) return set_error(MTS_ERR_ARG, "null argument");
    *stream = (void*)b->stream;
    return MTS_OK;
}

}  // extern "C"
