// C ABI of the library (include/mantiis_hip.h): handle life cycle, device
// memory, pass sequencing of the fused steps, per-method entry points and
// host/device transfers.  No exception crosses the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "mantiis_hip.h"
#include "mts_common.h"
#include "mts_kernels.h"
#include "mts_multigrid.h"

namespace mts {
void upload_xynorm(hipStream_t st);
}

using namespace mts;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// Stores the residual of a fused step call (LB2D::diff) and reports a NaN or
// Inf in it: a diverged run stops with MTS_ERR_NONFINITE instead of running
// on (its test diff < tol is never true).
static int set_residual(double* diff, double v) {
    *diff = v;
    return std::isfinite(v) ? MTS_OK : fail(MTS_ERR_NONFINITE, "residual is not finite (the run diverged)");
}

namespace mts {
// error reporting shared with the batched solver (mts_batch.hip)
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace mts

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? MTS_ERR_OOM : MTS_ERR_HIP,               \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

struct mts_sim_s {
    Geometry geo;
    Physics ph;
    mts_params prm{};
    hipStream_t stream = nullptr;
    int64_t P = 0;
    uint32_t* d_meta = nullptr;
    uint16_t* d_pidx = nullptr;
    CellParam* d_params = nullptr;
    int32_t* d_cellPos = nullptr;
    void* d_A = nullptr;
    void* d_B = nullptr;
    // MDBB with FP32 populations: the collided state of the wall cells, 4
    // planes per buffer, read by a pass's gather and written by its collision
    float* d_M[2] = {nullptr, nullptr};
    // FP32 carried populations live in cell order (single block)
    CompactImage cimg_h;
    std::vector<float> cpar_h;
    const float* d_book = nullptr;  // SRBB bookkeeping table of the FP32 parameter records
    CompactDev cimg;
    bool use_cimg = false;
    double* d_f = nullptr;
    double *d_rho = nullptr, *d_rho_old = nullptr, *d_ux = nullptr, *d_uy = nullptr, *d_usq = nullptr,
           *d_usq_old = nullptr;
    double *d_ftemp = nullptr, *d_feq = nullptr, *d_fneq = nullptr, *d_fb = nullptr, *d_psi = nullptr,
           *d_Fx = nullptr, *d_Fy = nullptr;
    double* d_part = nullptr;
    double* d_red = nullptr;
    double* d_compact = nullptr;
    int* d_neg = nullptr;
    int nparts = 0;
    double diff = 1.0;
    double Fbody = 1e-8;
    std::vector<void*> allocs;

    // ---- multi-block lattice: one site image per level
    struct LevelDev {
        int nx = 0, ny = 0, level = 1;
        int64_t stride = 0, P = 0, ncell = 0;
        uint32_t* d_meta = nullptr;
        uint16_t* d_pidx = nullptr;
        int32_t* d_sites = nullptr;
        void* d_A = nullptr;
        double *d_f = nullptr, *d_rho = nullptr, *d_ux = nullptr, *d_uy = nullptr, *d_usq = nullptr, *d_usq_old = nullptr;
        // arrays of the per-method calls (allocated by the first one)
        double *d_ftemp = nullptr, *d_feq = nullptr, *d_fneq = nullptr, *d_fb = nullptr, *d_rho_old = nullptr;
    };
    bool multi = false;
    MultiGeometry mg;
    std::vector<LevelImage> limg;
    std::vector<LevelDev> lv;
    std::vector<CellParam> mparams;
    // interface rows of every pair of levels (l, l + 1), one pair after the
    // other: pair p = l - 1 owns rows [off[p], off[p + 1])
    CoarseRow* d_crows = nullptr;
    FineRow* d_frows = nullptr;
    int64_t ncrows = 0, nfrows = 0;
    std::vector<int64_t> crow_off, frow_off;
    // FP32 interface plans (one launch per pair of levels and coarse step),
    // blocks of pair p: [ifb_off[p], ifb_off[p + 1])
    IfHead* d_ifhead = nullptr;
    IfTask* d_iftask = nullptr;
    IfCoarse* d_ifco = nullptr;
    IfFine* d_iffo = nullptr;
    std::vector<int> ifb_off;
    // FP32 multi-block state in cell order, per level: image (host and device,
    // with and without interface overrides), carried states, stored fields
    // (two sets: a step's interface reads the previous step's while the
    // step's first substeps write their own)
    struct LevelC {
        CompactImage h;
        CompactDev plain, ovr;
        float* A[2] = {nullptr, nullptr};
        float* m[2] = {nullptr, nullptr};
        float* O = nullptr;
    };
    std::vector<LevelC> lc;

    LaunchCtx lctx(int l) const {
        LaunchCtx c = ctx();
        const LevelDev& L = lv[l];
        c.nx = L.nx;
        c.ny = L.ny;
        c.stride = L.stride;
        c.ncell = L.ncell;
        c.meta = L.d_meta;
        c.pidx = L.d_pidx;
        c.cellPos = L.d_sites;
        c.one_param = mparams.size() == 1;
        c.fscale = (double)(1 << l);
        c.gpIn = c.gpOut = -1;
        c.cimg = nullptr;
        return c;
    }

    LaunchCtx ctx() const {
        LaunchCtx c;
        c.stream = stream;
        c.nx = geo.nx;
        c.ny = geo.ny;
        c.stride = geo.stride;
        c.ncell = geo.ncell;
        c.meta = d_meta;
        c.pidx = d_pidx;
        c.params = d_params;
        c.cellPos = d_cellPos;
        c.fp32 = prm.precision == MTS_FP32DEV;
        c.coll = prm.collision;
        c.wall = prm.wall;
        c.rho0 = ph.rhoin;
        c.rhoin = ph.rhoin;
        c.Fbody = Fbody;
        c.Gff = prm.Gff;
        c.Gfs = prm.Gfs;
        c.T_lu = ph.T_lu;
        c.a_alpha_lu = ph.a_alpha_lu;
        c.gpIn = geo.gpIn;
        c.gpOut = geo.gpOut;
        c.one_param = geo.params.size() == 1;
        c.cimg = use_cimg ? &cimg : nullptr;
        return c;
    }
    template <class T>
    int alloc(T** p, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();  // a failed allocation must not surface at a later launch check
            return fail(MTS_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        e = hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), stream);
        if (e != hipSuccess) return fail(MTS_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
        allocs.push_back(q);
        *p = (T*)q;
        return MTS_OK;
    }
    // kernel timer: event pairs around the dominant kernel's launches
    bool ktimer = false;
    std::vector<hipEvent_t> kev;
    size_t kev_used = 0;
    double ktime_ms = 0.0;
    int64_t klaunches = 0;
    // events for the next timed launch (nullptr when the timer is off)
    void tick(LaunchCtx& c) {
        c.ev0 = c.ev1 = nullptr;
        if (!ktimer) return;
        if (kev_used + 2 > kev.size()) {
            if (kev.size() >= 8192) {  // pool full: fold the pending pairs first
                fold_timer();
            } else {
                for (int i = 0; i < 256; i++) {
                    hipEvent_t e;
                    if (hipEventCreate(&e) != hipSuccess) return;
                    kev.push_back(e);
                }
            }
        }
        if (kev_used + 2 > kev.size()) return;
        c.ev0 = kev[kev_used];
        c.ev1 = kev[kev_used + 1];
    }
    void tock(LaunchCtx& c) {  // after the timed launch
        if (c.ev0) kev_used += 2;
        c.ev0 = c.ev1 = nullptr;
    }
    void fold_timer() {
        (void)hipStreamSynchronize(stream);
        for (size_t i = 0; i + 1 < kev_used; i += 2) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, kev[i], kev[i + 1]) == hipSuccess) {
                ktime_ms += ms;
                klaunches++;
            }
        }
        kev_used = 0;
    }
    ~mts_sim_s() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : kev) (void)hipEventDestroy(e);
        for (void* q : allocs) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int ensure_phase() {
        if (d_ftemp) return MTS_OK;
        const size_t np = (size_t)geo.stride * 9;
        int rc;
        if ((rc = alloc(&d_ftemp, np)) || (rc = alloc(&d_feq, np)) || (rc = alloc(&d_fneq, np)) ||
            (rc = alloc(&d_fb, np)))
            return rc;
        return MTS_OK;
    }
    int ensure_sc() {
        if (d_psi) return MTS_OK;
        int rc;
        if ((rc = alloc(&d_psi, P)) || (rc = alloc(&d_Fx, P)) || (rc = alloc(&d_Fy, P))) return rc;
        return MTS_OK;
    }
    int launched() {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(MTS_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
        return MTS_OK;
    }
    int reduce(int kind, const double* a0, const double* a1, const double* a2, double* out4) {
        launch_reduce(ctx(), kind, a0, a1, a2, d_part, nparts, d_red);
        int rc = launched();
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(out4, d_red, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return MTS_OK;
    }
};

// ---------------------------------------------------------------- multi-block
static int upload(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return MTS_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return fail(MTS_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    return MTS_OK;
}

// FP32 parameter records (s_s, s_q, s_d, r as hi + lo) of a parameter table.
static int upload_cpar(mts_sim_s* s, const std::vector<CellParam>& params, const float** out) {
    std::vector<float> book;
    fp32_params(params, s->ph.rhoin, s->cpar_h, book);
    float* cpar;
    float* dbook;
    int rc;
    if ((rc = s->alloc(&cpar, s->cpar_h.size())) ||
        (rc = upload(cpar, s->cpar_h.data(), s->cpar_h.size() * 4, s->stream)) ||
        (rc = s->alloc(&dbook, book.size())) || (rc = upload(dbook, book.data(), book.size() * 4, s->stream)))
        return rc;
    s->d_book = dbook;
    *out = cpar;
    return MTS_OK;
}

// Device copy of a cell-order image (the host image must outlive the copies).
static int upload_compact(mts_sim_s* s, const CompactImage& h, const float* cpar, CompactDev& d) {
    uint32_t* cw;
    uint32_t* rw;
    int32_t *aux, *cidx, *segs, *cpos;
    uint16_t* pidx;
    uint8_t* outm;
    int rc;
    if ((rc = s->alloc(&cw, h.ncell)) || (rc = s->alloc(&rw, h.ncell)) ||
        (rc = s->alloc(&pidx, h.ncell)) || (rc = s->alloc(&outm, h.ncell)) ||
        (rc = s->alloc(&aux, std::max<size_t>(h.aux.size(), 1))) || (rc = s->alloc(&cidx, h.cidx.size())) ||
        (rc = s->alloc(&segs, h.segs.size())) || (rc = s->alloc(&cpos, h.ncell)))
        return rc;
    if ((rc = upload(cw, h.cw.data(), h.ncell * 4, s->stream)) || (rc = upload(rw, h.rw.data(), h.ncell * 4, s->stream)) ||
        (rc = upload(pidx, h.pidx.data(), h.ncell * 2, s->stream)) ||
        (rc = upload(outm, h.outm.data(), h.ncell, s->stream)) ||
        (rc = upload(aux, h.aux.data(), h.aux.size() * 4, s->stream)) ||
        (rc = upload(cidx, h.cidx.data(), h.cidx.size() * 4, s->stream)) ||
        (rc = upload(segs, h.segs.data(), h.segs.size() * 4, s->stream)) ||
        (rc = upload(cpos, h.cellPos.data(), h.ncell * 4, s->stream)))
        return rc;
    d.cw = cw;
    d.rw = rw;
    d.pidx = pidx;
    d.outm = outm;
    d.aux = aux;
    d.cidx = cidx;
    d.segs = segs;
    d.cellPos = cpos;
    d.cpar = cpar;
    d.book = s->d_book;
    d.ncell = h.ncell;
    d.cstride = h.cstride;
    d.ntx = h.ntx;
    return MTS_OK;
}

// Interface rows of coreLB.hpp:1101-1231 with every coordinate and weight
// input formed on the host as the reference forms them, for each pair of
// levels (l1, l2 = l1 + 1) in turn; off[p] are the first rows of pair p.
static void interface_rows(const MultiGeometry& G, const std::vector<LevelImage>& L, std::vector<CoarseRow>& cr,
                           std::vector<FineRow>& fr, std::vector<int64_t>& coff, std::vector<int64_t>& foff) {
    auto site = [&](int64_t c) -> int64_t {
        const int l = G.level[c] - 1;
        return (int64_t)(G.cy[c] >> l) * L[l].nx + (G.cx[c] >> l);
    };
    const int np = std::max(G.maxLevel - 1, 0);
    coff.assign(np + 1, 0);
    foff.assign(np + 1, 0);
    for (int p = 0; p < np; p++) {
        coff[p + 1] = coff[p] + G.coarseCount[p];
        foff[p + 1] = foff[p] + G.fineCount[p];
    }
    for (size_t r = 0; r + 4 < G.coarseRows.size() + 1; r += 5) {
        CoarseRow R{};
        const int64_t* row = &G.coarseRows[r];
        R.parent = site(row[0]);
        R.tc = G.taus[row[0]];
        double tf = 0;
        int nch = 0;
        for (int k = 0; k < 4; k++) {
            R.child[k] = row[1 + k] >= 0 ? site(row[1 + k]) : -1;
            if (row[1 + k] >= 0) {
                tf += G.taus[row[1 + k]];
                nch++;
            }
        }
        R.nch = nch;
        R.tf = nch ? tf / (nch == 4 ? 4 : nch) : 1.0;
        cr.push_back(R);
    }
    static const int top[9] = {7, 4, 8, 3, 0, 1, 6, 2, 5};
    static const int bottom[9] = {6, 2, 5, 3, 0, 1, 7, 4, 8};
    int p = 0;
    for (size_t r = 0; r + 4 < G.fineRows.size() + 1; r += 5) {
        while (p + 1 < np && (int64_t)(r / 5) >= foff[p + 1]) p++;
        const int l1 = p + 1, l2 = p + 2;
        // centre shifts of the parent's and the children's level (coreLB.hpp:1134-1135)
        const double pshift = std::pow(2.0, l2 - 2) - 0.5, cshift = l1 == 1 ? 0.0 : std::pow(2.0, l1 - 2) - 0.5;
        FineRow R{};
        const int64_t* row = &G.fineRows[r];
        const int64_t par = row[0];
        const int64_t* pc = &G.connect[9 * par];
        const int py = G.cy[par], px = G.cx[par];
        const double half = std::pow(2, G.level[par] - 1);
        std::vector<double> X(9), Y(9);
        for (int ic = 0; ic < 9; ic++) {
            const int ny_ = pc[ic] >= 0 ? G.cy[pc[ic]] : py + 2 * EY[ic];
            const int nx_ = pc[ic] >= 0 ? G.cx[pc[ic]] : px + 2 * EX[ic];
            if (std::abs(py - ny_) > half && py - ny_ < 0) X[ic] = py + pshift + 1 - 2;
            else if (std::abs(py - ny_) > half && py - ny_ > 0) X[ic] = py + pshift + 1 + 2;
            else X[ic] = ny_ + pshift + 1;
            if (std::abs(px - nx_) > half && px - nx_ < 0) Y[ic] = px + pshift + 1 - 2;
            else if (std::abs(px - nx_) > half && px - nx_ > 0) Y[ic] = px + pshift + 1 + 2;
            else Y[ic] = nx_ + pshift + 1;
        }
        std::sort(X.begin(), X.end());
        X.erase(std::unique(X.begin(), X.end()), X.end());
        std::sort(Y.begin(), Y.end());
        Y.erase(std::unique(Y.begin(), Y.end()), Y.end());
        const bool topHalf = py < G.nx / 2;
        for (int k = 0; k < 3; k++) R.xg[k] = Y[std::min<size_t>(k, Y.size() - 1)];
        for (int k = 0; k < 2; k++) R.yg[k] = topHalf ? X[std::min<size_t>(k, X.size() - 1)] : X[X.size() - k - 1];
        const int* ord = topHalf ? top : bottom;
        for (int k = 0; k < 6; k++) R.src[k] = site(pc[ord[k]] >= 0 ? pc[ord[k]] : par);
        for (int ch = 0; ch < 4; ch++) {
            const int64_t cc = row[1 + ch];
            R.child[ch] = site(cc);
            R.qx[ch] = G.cx[cc] + cshift + 1;
            R.qy[ch] = G.cy[cc] + cshift + 1;
            R.ratio[ch] = G.taus[cc] / (2 * G.taus[par]);
        }
        fr.push_back(R);
    }
}

// Interface plan of the FP32 path (IfHead / IfTask / IfCoarse / IfFine):
// fine rows are packed in order into blocks while the block's gathers,
// coarse-buffer values and fine outputs fit; a node that is a coarse-buffer
// cell adds that coarse row (its children's gathers) to the block, and the
// first block to add a coarse row writes its override.  The coarse rows no
// fine row uses fill further blocks.  Every array is padded to the block's
// capacity, so a block's records start at fixed offsets.
struct IfPlan {
    std::vector<IfHead> head;
    std::vector<IfTask> task;
    std::vector<IfCoarse> co;
    std::vector<IfFine> fo;
};

// fbase: the first override slot of the fine rows on the fine level (its
// coarse-buffer slots, as the coarse level of the pair below, come first).
static void build_if_plan(const std::vector<CoarseRow32>& cr, const std::vector<FineRow32>& fr, const CompactImage& h1,
                          const CompactImage& h2, int64_t fbase, IfPlan& P) {
    std::unordered_map<int32_t, int32_t> crowOf;  // coarse-buffer cell -> coarse row
    for (size_t r = 0; r < cr.size(); r++)
        if (cr[r].nch > 0) crowOf[cr[r].parent] = (int32_t)r;
    std::vector<char> owned(cr.size(), 0);
    std::unordered_map<int64_t, int> tslot;  // (level, cell) -> task slot of the open block
    std::unordered_map<int32_t, int> cslot;  // coarse row -> value slot of the open block
    std::vector<IfTask> T;
    std::vector<IfCoarse> C;
    std::vector<IfFine> F;
    auto key = [](int lvl, int32_t cell) { return ((int64_t)lvl << 32) | (uint32_t)cell; };
    auto flush = [&]() {
        if (T.empty() && C.empty() && F.empty()) return;
        // gathers in (level, cell) order, so that neighbouring lanes read
        // neighbouring elements of each plane
        std::vector<int> ord(T.size()), at(T.size());
        for (size_t i = 0; i < T.size(); i++) ord[i] = (int)i;
        std::sort(ord.begin(), ord.end(), [&](int x, int y) {
            return std::make_pair(T[x].pl >> 31, T[x].cell) < std::make_pair(T[y].pl >> 31, T[y].cell);
        });
        std::vector<IfTask> Ts(T.size());
        for (size_t i = 0; i < T.size(); i++) {
            Ts[i] = T[ord[i]];
            at[ord[i]] = (int)i;
        }
        T.swap(Ts);
        for (IfCoarse& o : C)
            for (int k = 0; k < 4; k++)
                if (o.src[k] >= 0) o.src[k] = (int16_t)at[o.src[k]];
        for (IfFine& o : F)
            for (int k = 0; k < 6; k++)
                if (o.node[k] < IF_TASKS) o.node[k] = (int16_t)at[o.node[k]];
        P.head.push_back(IfHead{(int32_t)T.size(), (int32_t)C.size(), (int32_t)F.size(), 0});
        T.resize(IF_TASKS, IfTask{-1, 0u, 0u, 0u});
        C.resize(IF_CO, IfCoarse{{-1, -1, -1, -1}, 0, 0.0f, 0, -1, 0, 0});
        F.resize(IF_FO, IfFine{0, 0, 0.0f, {0, 0, 0, 0, 0, 0}, 0, 0, {0, 0, 0, 0, 0, 0}});
        P.task.insert(P.task.end(), T.begin(), T.end());
        P.co.insert(P.co.end(), C.begin(), C.end());
        P.fo.insert(P.fo.end(), F.begin(), F.end());
        T.clear();
        C.clear();
        F.clear();
        tslot.clear();
        cslot.clear();
    };
    auto task = [&](int lvl, int32_t cell) -> int {
        const auto it = tslot.find(key(lvl, cell));
        if (it != tslot.end()) return it->second;
        const CompactImage& h = lvl == 2 ? h2 : h1;
        T.push_back(IfTask{cell, h.cw[cell], h.rw[cell], (uint32_t)h.pidx[cell] | (lvl == 2 ? 0x80000000u : 0u)});
        return tslot[key(lvl, cell)] = (int)T.size() - 1;
    };
    auto coarse = [&](int32_t r) -> int {
        const auto it = cslot.find(r);
        if (it != cslot.end()) return IF_TASKS + it->second;
        const CoarseRow32& R = cr[r];
        IfCoarse o{{-1, -1, -1, -1}, R.parent, R.k, R.nch, owned[r] ? -1 : R.slot, (int32_t)h2.pidx[R.parent], 0};
        owned[r] = 1;
        for (int k = 0; k < 4; k++)
            if (R.child[k] >= 0) o.src[k] = (int16_t)task(1, R.child[k]);
        C.push_back(o);
        return IF_TASKS + (cslot[r] = (int)C.size() - 1);
    };
    // gathers and coarse values a fine row would add to the open block
    auto need = [&](const FineRow32& R, int& nt, int& nc) {
        nt = nc = 0;
        std::vector<int64_t> seen;
        for (int k = 0; k < 6; k++) {
            const auto it = crowOf.find(R.src[k]);
            if (it == crowOf.end()) {
                const int64_t kk = key(2, R.src[k]);
                if (!tslot.count(kk) && std::find(seen.begin(), seen.end(), kk) == seen.end()) {
                    seen.push_back(kk);
                    nt++;
                }
            } else if (!cslot.count(it->second) &&
                       std::find(seen.begin(), seen.end(), -1 - (int64_t)it->second) == seen.end()) {
                seen.push_back(-1 - (int64_t)it->second);
                nc++;
                nt += cr[it->second].nch;
            }
        }
    };
    for (size_t r = 0; r < fr.size(); r++) {
        const FineRow32& R = fr[r];
        int nt, nc;
        need(R, nt, nc);
        if ((int)T.size() + nt > IF_TASKS || (int)C.size() + nc > IF_CO || (int)F.size() + 4 > IF_FO) flush();
        int16_t node[6];
        for (int k = 0; k < 6; k++) {
            const auto it = crowOf.find(R.src[k]);
            node[k] = (int16_t)(it == crowOf.end() ? task(2, R.src[k]) : coarse(it->second));
        }
        for (int ch = 0; ch < 4; ch++) {
            IfFine o{R.child[ch], (int32_t)(fbase + 4 * r + ch), R.ratio[ch], {0, 0, 0, 0, 0, 0}, h1.pidx[R.child[ch]], 0,
                     {0, 0, 0, 0, 0, 0}};
            for (int k = 0; k < 6; k++) {
                o.node[k] = node[k];
                o.L[k] = R.L[ch][k];
            }
            F.push_back(o);
        }
    }
    flush();
    for (size_t r = 0; r < cr.size(); r++) {
        if (cr[r].nch <= 0 || owned[r]) continue;
        if ((int)T.size() + cr[r].nch > IF_TASKS || (int)C.size() + 1 > IF_CO) flush();
        coarse((int32_t)r);
    }
    flush();
}

// Bounds of every index the interface kernel takes from the plan, and each
// override slot written exactly once; false names the first violation.
static bool check_if_plan(const IfPlan& P, size_t ncr, const std::vector<CoarseRow32>& cr, size_t nfr, int64_t fbase,
                          int64_t ncell1, int64_t ncell2, size_t nparam, std::string& err) {
    std::vector<int> wrote2(ncr, 0), wrote1(4 * nfr, 0);
    for (size_t b = 0; b < P.head.size(); b++) {
        const IfHead& H = P.head[b];
        if (H.ntask < 0 || H.ntask > IF_TASKS || H.nco < 0 || H.nco > IF_CO || H.nfo < 0 || H.nfo > IF_FO) {
            err = "interface plan: block capacity";
            return false;
        }
        for (int t = 0; t < H.ntask; t++) {
            const IfTask& T = P.task[b * IF_TASKS + t];
            if (T.cell < 0 || T.cell >= ((T.pl >> 31) ? ncell2 : ncell1)) {
                err = "interface plan: task cell";
                return false;
            }
        }
        for (int i = 0; i < H.nco; i++) {
            const IfCoarse& C = P.co[b * IF_CO + i];
            for (int k = 0; k < 4; k++)
                if (C.src[k] < -1 || C.src[k] >= H.ntask) {
                    err = "interface plan: child slot";
                    return false;
                }
            if (C.parent < 0 || C.parent >= ncell2 || C.nch <= 0 || C.out < -1 || C.out >= (int64_t)ncr ||
                C.pidx < 0 || (size_t)C.pidx >= nparam) {
                err = "interface plan: coarse record";
                return false;
            }
            if (C.out >= 0) wrote2[C.out]++;
        }
        for (int i = 0; i < H.nfo; i++) {
            const IfFine& F = P.fo[b * IF_FO + i];
            for (int k = 0; k < 6; k++)
                if (F.node[k] < 0 || (F.node[k] >= H.ntask && (F.node[k] < IF_TASKS || F.node[k] >= IF_TASKS + H.nco))) {
                    err = "interface plan: node slot";
                    return false;
                }
            if (F.child < 0 || F.child >= ncell1 || F.out < fbase || F.out >= fbase + (int64_t)(4 * nfr) ||
                F.pidx >= nparam) {
                err = "interface plan: fine record";
                return false;
            }
            wrote1[F.out - fbase]++;
        }
    }
    for (size_t r = 0; r < ncr; r++)
        if (wrote2[r] != (cr[r].nch > 0 ? 1 : 0)) {
            err = "interface plan: coarse row not written once";
            return false;
        }
    for (size_t i = 0; i < wrote1.size(); i++)
        if (wrote1[i] != 1) {
            err = "interface plan: fine buffer not written once";
            return false;
        }
    return true;
}

// Cell-order images of the levels for the FP32 path.  Buffer cells written
// by the interface get an override slot on their level: the coarse buffer
// of row r of the pair below -> r, then fine buffer k of row r of the pair
// above -> (coarse rows of the pair below) + 4 r + k.
static int create_multi32(mts_sim_s* s, const std::vector<CoarseRow>& cr, const std::vector<FineRow>& fr,
                          std::string& err) {
    const int Lmax = (int)s->limg.size(), np = Lmax - 1;
    const std::vector<int64_t>&co = s->crow_off, &fo = s->frow_off;
    // every interface site must be a cell of its level image
    auto is_cell = [&](int l, int64_t site) {
        const LevelImage& I = s->limg[l];
        return site >= 0 && site < (int64_t)I.nx * I.ny && (I.meta[site] & FLUID_BIT);
    };
    for (int p = 0; p < np; p++) {
        for (int64_t r = co[p]; r < co[p + 1]; r++) {
            const CoarseRow& R = cr[r];
            if (R.nch > 0 && !is_cell(p + 1, R.parent)) return fail(MTS_ERR_GEOMETRY, "interface row: parent is not a cell");
            for (int k = 0; k < 4; k++)
                if (R.child[k] >= 0 && !is_cell(p, R.child[k]))
                    return fail(MTS_ERR_GEOMETRY, "interface row: child is not a cell");
        }
        for (int64_t r = fo[p]; r < fo[p + 1]; r++) {
            const FineRow& R = fr[r];
            for (int k = 0; k < 6; k++)
                if (!is_cell(p + 1, R.src[k])) return fail(MTS_ERR_GEOMETRY, "interface row: node is not a cell");
            for (int k = 0; k < 4; k++)
                if (!is_cell(p, R.child[k])) return fail(MTS_ERR_GEOMETRY, "interface row: buffer is not a cell");
        }
    }
    // first fine-buffer slot of level l (after the coarse buffers of pair l - 1)
    auto fbase = [&](int l) -> int64_t { return l >= 1 ? co[l] - co[l - 1] : 0; };
    s->lc.resize(Lmax);
    const float* cpar;
    int rc;
    if ((rc = upload_cpar(s, s->mparams, &cpar))) return rc;
    for (int l = 0; l < Lmax; l++) {
        const LevelImage& I = s->limg[l];
        auto& C = s->lc[l];
        const int64_t P = (int64_t)I.nx * I.ny;
        std::vector<int32_t> ovr(P, -1);
        int64_t nslot = fbase(l);
        if (l >= 1)
            for (int64_t r = co[l - 1]; r < co[l]; r++)
                if (cr[r].nch > 0) {
                    if (ovr[cr[r].parent] >= 0) {
                        err = "coarse buffer cell in two interface rows";
                        return MTS_ERR_GEOMETRY;
                    }
                    ovr[cr[r].parent] = (int32_t)(r - co[l - 1]);
                }
        if (l < np) {
            for (int64_t r = fo[l]; r < fo[l + 1]; r++)
                for (int k = 0; k < 4; k++) {
                    if (ovr[fr[r].child[k]] >= 0) {
                        err = "fine buffer cell in two interface rows";
                        return MTS_ERR_GEOMETRY;
                    }
                    ovr[fr[r].child[k]] = (int32_t)(fbase(l) + 4 * (r - fo[l]) + k);
                }
            nslot += 4 * (fo[l + 1] - fo[l]);
        }
        LatticeView v{I.nx, I.ny, s->mg.torus, I.meta.data(), I.pidx.data(), -1, -1, ovr.data()};
        build_compact(v, C.h);
        if ((rc = upload_compact(s, C.h, cpar, C.plain))) return rc;
        const int64_t OS = std::max<int64_t>((nslot + 63) / 64 * 64, 64);
        if ((rc = s->alloc(&C.A[0], (size_t)C.h.cstride * 9)) || (rc = s->alloc(&C.A[1], (size_t)C.h.cstride * 9)) ||
            (rc = s->alloc(&C.m[0], (size_t)C.h.cstride * 3)) || (rc = s->alloc(&C.m[1], (size_t)C.h.cstride * 3)) ||
            (rc = s->alloc(&C.O, (size_t)OS * 9)))
            return rc;
        C.plain.O = C.O;
        C.plain.OS = OS;
        C.plain.ovr = 0;
        C.ovr = C.plain;
        C.ovr.ovr = 1;
    }
    upload_spec(s->lc[0].h.spec, s->stream);
    // interface plans of the pairs, in cell indices with host-formed weights
    IfPlan all;
    s->ifb_off.assign(1, 0);
    for (int p = 0; p < np; p++) {
        const auto& ci1 = s->lc[p].h.cidx;
        const auto& ci2 = s->lc[p + 1].h.cidx;
        std::vector<CoarseRow32> c32(co[p + 1] - co[p]);
        for (size_t r = 0; r < c32.size(); r++) {
            const CoarseRow& R = cr[co[p] + r];
            CoarseRow32& o = c32[r];
            o.parent = R.nch > 0 ? ci2[R.parent] : 0;
            for (int k = 0; k < 4; k++) o.child[k] = R.child[k] >= 0 ? ci1[R.child[k]] : -1;
            o.k = R.nch > 0 ? (float)(2 * R.tc / R.tf) : 0.0f;
            o.nch = R.nch;
            o.slot = (int32_t)r;
        }
        std::vector<FineRow32> f32(fo[p + 1] - fo[p]);
        for (size_t r = 0; r < f32.size(); r++) {
            const FineRow& R = fr[fo[p] + r];
            FineRow32& o = f32[r];
            for (int k = 0; k < 4; k++) o.child[k] = ci1[R.child[k]];
            for (int k = 0; k < 6; k++) o.src[k] = ci2[R.src[k]];
            for (int ch = 0; ch < 4; ch++) {
                const double qx = R.qx[ch], qy = R.qy[ch];
                for (int b = 0; b < 2; b++) {
                    const double Ly = (qy - R.yg[1 - b]) / (R.yg[b] - R.yg[1 - b]);
                    for (int a = 0; a < 3; a++) {
                        double Lx = 1;
                        for (int mm = 0; mm < 3; mm++)
                            if (mm != a) Lx *= (qx - R.xg[mm]) / (R.xg[a] - R.xg[mm]);
                        o.L[ch][3 * b + a] = (float)(Lx * Ly);
                    }
                }
                o.ratio[ch] = (float)R.ratio[ch];
            }
        }
        IfPlan P;
        build_if_plan(c32, f32, s->lc[p].h, s->lc[p + 1].h, fbase(p), P);
        if (!check_if_plan(P, c32.size(), c32, f32.size(), fbase(p), s->lc[p].h.ncell, s->lc[p + 1].h.ncell,
                           s->mparams.size(), err))
            return fail(MTS_ERR_GEOMETRY, err);
        all.head.insert(all.head.end(), P.head.begin(), P.head.end());
        all.task.insert(all.task.end(), P.task.begin(), P.task.end());
        all.co.insert(all.co.end(), P.co.begin(), P.co.end());
        all.fo.insert(all.fo.end(), P.fo.begin(), P.fo.end());
        s->ifb_off.push_back((int)all.head.size());
    }
    if (!all.head.empty()) {
        if ((rc = s->alloc((char**)&s->d_ifhead, all.head.size() * sizeof(IfHead))) ||
            (rc = s->alloc((char**)&s->d_iftask, all.task.size() * sizeof(IfTask))) ||
            (rc = s->alloc((char**)&s->d_ifco, all.co.size() * sizeof(IfCoarse))) ||
            (rc = s->alloc((char**)&s->d_iffo, all.fo.size() * sizeof(IfFine))) ||
            (rc = upload(s->d_ifhead, all.head.data(), all.head.size() * sizeof(IfHead), s->stream)) ||
            (rc = upload(s->d_iftask, all.task.data(), all.task.size() * sizeof(IfTask), s->stream)) ||
            (rc = upload(s->d_ifco, all.co.data(), all.co.size() * sizeof(IfCoarse), s->stream)) ||
            (rc = upload(s->d_iffo, all.fo.data(), all.fo.size() * sizeof(IfFine), s->stream)))
            return rc;
        HIP_TRY(hipStreamSynchronize(s->stream));  // the host plan goes out of scope
    }
    return MTS_OK;
}

static int create_multi(mts_sim_s* s, int nx, int ny, const int32_t* pore, const double* kn, const double* locpore,
                        std::string& err) {
    s->multi = true;
    hipError_t e = hipSetDevice(s->prm.device);
    if (e != hipSuccess) return fail(MTS_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(MTS_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
    // the quadtree: level decision and tables on the device, the buffer
    // cells (a sequential rule) and the parameter table on the host
    (void)locpore;
    const bool torus = s->prm.periodic_y != 0;
    const int K = quadtree_levels(s->prm.multiblock);
    std::vector<uint8_t> lvl;
    int rc = multi_check(nx, ny, pore, K, err);
    if (rc) return rc;
    if ((rc = multi_levels_device(nx, ny, pore, torus, K, s->stream, lvl, err))) return rc;
    if ((rc = multi_cells(nx, ny, pore, kn, torus, lvl, s->mg, err))) return rc;
    {  // a domain with no wall near some level (a wall-free torus) leaves the finer levels empty
        std::vector<int64_t> count(s->mg.maxLevel + 1, 0);
        for (int64_t c = 0; c < s->mg.ncell; c++) count[s->mg.level[c]]++;
        for (int l = 1; l <= s->mg.maxLevel; l++)
            if (!count[l])
                return fail(MTS_ERR_UNSUPPORTED, "multi-block lattice whose level " + std::to_string(l) +
                                                     " holds no cells (no wall near it); use multiblock = 0");
    }
    if ((rc = multi_grid_taus(s->ph, s->mg, err))) return rc;
    if ((rc = build_level_images(s->mg, s->limg, s->mparams, s->ph, err, false))) return rc;
    if ((rc = multi_tables_device(s->mg, &s->limg, s->stream, err))) return rc;
    const size_t esz = s->prm.precision == MTS_FP32DEV ? 4 : 8;
    int64_t Pmax = 0;
    s->geo.nx = nx;
    s->geo.ny = ny;
    s->geo.ncell = s->mg.ncell;
    s->lv.resize(s->limg.size());
    for (size_t l = 0; l < s->limg.size(); l++) {
        const LevelImage& I = s->limg[l];
        mts_sim_s::LevelDev& L = s->lv[l];
        L.nx = I.nx;
        L.ny = I.ny;
        L.level = I.level;
        L.stride = I.stride;
        L.P = (int64_t)I.nx * I.ny;
        L.ncell = (int64_t)I.cells.size();
        Pmax = std::max(Pmax, L.P);
        const size_t np = (size_t)I.stride * 9;
        if ((rc = s->alloc(&L.d_meta, L.P)) || (rc = s->alloc(&L.d_pidx, L.P)) || (rc = s->alloc(&L.d_sites, L.ncell)) ||
            (rc = s->alloc((char**)&L.d_A, np * esz)) || (rc = s->alloc(&L.d_f, np)) || (rc = s->alloc(&L.d_rho, L.P)) ||
            (rc = s->alloc(&L.d_ux, L.P)) || (rc = s->alloc(&L.d_uy, L.P)) || (rc = s->alloc(&L.d_usq, L.P)) ||
            (rc = s->alloc(&L.d_usq_old, L.P)))
            return rc;
        if ((rc = upload(L.d_meta, I.meta.data(), L.P * 4, s->stream)) ||
            (rc = upload(L.d_pidx, I.pidx.data(), L.P * 2, s->stream)) ||
            (rc = upload(L.d_sites, I.sites.data(), L.ncell * 4, s->stream)))
            return rc;
    }
    s->P = Pmax;
    s->nparts = (int)std::min<int64_t>(1024, (Pmax + 255) / 256);
    if ((rc = s->alloc(&s->d_params, s->mparams.size())) || (rc = s->alloc(&s->d_part, (size_t)4 * s->nparts)) ||
        (rc = s->alloc(&s->d_red, 4)) || (rc = s->alloc(&s->d_neg, 1)))
        return rc;
    if ((rc = upload(s->d_params, s->mparams.data(), s->mparams.size() * sizeof(CellParam), s->stream))) return rc;
    std::vector<CoarseRow> cr;
    std::vector<FineRow> fr;
    interface_rows(s->mg, s->limg, cr, fr, s->crow_off, s->frow_off);
    s->ncrows = (int64_t)cr.size();
    s->nfrows = (int64_t)fr.size();
    if ((rc = s->alloc((char**)&s->d_crows, cr.size() * sizeof(CoarseRow))) ||
        (rc = s->alloc((char**)&s->d_frows, fr.size() * sizeof(FineRow))))
        return rc;
    if ((rc = upload(s->d_crows, cr.data(), cr.size() * sizeof(CoarseRow), s->stream)) ||
        (rc = upload(s->d_frows, fr.data(), fr.size() * sizeof(FineRow), s->stream)))
        return rc;
    if (s->prm.precision == MTS_FP32DEV && (rc = create_multi32(s, cr, fr, err))) return rc;
    upload_icsr(s->stream);
    upload_xynorm(s->stream);
    for (size_t l = 0; l < s->lv.size(); l++) {
        LaunchCtx c = s->lctx((int)l);
        launch_init(c, s->lv[l].d_f);
        launch_macro(c, s->lv[l].d_f, s->lv[l].d_rho, s->lv[l].d_ux, s->lv[l].d_uy, s->lv[l].d_usq);
    }
    return s->launched();
}

// One call = nsteps coarse steps of the transport loop on the quadtree
// (mantiisApp.hpp:239-246 with evolutionStepOfMultiBlock, coreLB.hpp:1310-1331).
// FP32 quadtree step in cell order: per coarse step level l does 2^(L-l)
// fused substeps (all but the first with the stored fields of the first),
// then the interface forms the buffer cells' post-stream states.  On two
// levels, from the second step of a call on, the interface of the previous
// step runs in the launch of the step's first substeps and updates the
// buffer cells itself, and the last step's interface writes the overrides;
// on three, both interfaces of a step run in the next step's first two
// launches.  A call starts from and ends in the canonical FP64 level images.
static int multi_transport32(mts_sim_s* s, int nsteps) {
    const int Lmax = (int)s->lv.size();
    int rc;
    std::vector<int> cur(Lmax, 0);
    int mi = 0;  // stored-field set of the current step
    // the interface of the pair of levels (p + 1, p + 2)
    auto iface = [&](int p, int mprev) {
        IfSpec f;
        f.A1 = s->lc[p].A[cur[p]];
        f.A2 = s->lc[p + 1].A[cur[p + 1]];
        f.m1 = s->lc[p].m[mprev];
        f.m2 = s->lc[p + 1].m[mprev];
        f.O1 = s->lc[p].O;
        f.O2 = s->lc[p + 1].O;
        const int b0 = s->ifb_off[p];
        f.head = s->d_ifhead + b0;
        f.task = s->d_iftask + (int64_t)b0 * IF_TASKS;
        f.co = s->d_ifco + (int64_t)b0 * IF_CO;
        f.fo = s->d_iffo + (int64_t)b0 * IF_FO;
        f.nblocks = s->ifb_off[p + 1] - b0;
        // above the finest pair, children may be coarse buffers of the pair
        // below, whose values that pair's interface has just written
        f.child_ovr = p > 0 ? 1 : 0;
        return f;
    };
    for (int step = 0; step < nsteps; step++) {
        const bool first = step == 0, last = step == nsteps - 1;
        if (Lmax == 2) {
            // both levels' first substeps in one launch (with the previous
            // step's interface after the first step), then the second
            // level-1 substep with the stored fields
            const IfSpec f = iface(0, mi ^ 1);
            MbSpec a[2];
            LaunchCtx c[2];
            for (int l = 0; l < 2; l++) {
                auto& C = s->lc[l];
                c[l] = s->lctx(l);
                c[l].cimg = &C.ovr;
                a[l].Ain = C.A[cur[l]];
                a[l].Aout = C.A[cur[l] ^ 1];
                a[l].fcan = s->lv[l].d_f;
                a[l].m = C.m[mi];
                a[l].usq_old = last ? s->lv[l].d_usq_old : nullptr;
                cur[l] ^= 1;
            }
            s->tick(c[0]);
            if ((rc = first ? launch_cmultistep2(c[0], a[0], c[1], a[1], 0)
                            : launch_cinterface_step(c[0], a[0], c[1], a[1], f)))
                return fail(rc, "operator combination has no kernel");
            s->tock(c[0]);
            auto& C = s->lc[0];
            c[0].cimg = &C.plain;
            MbSpec b;
            b.Ain = C.A[cur[0]];
            b.Aout = C.A[cur[0] ^ 1];
            b.m = C.m[mi];
            launch_cmultistep(c[0], 1, 1, b);
            cur[0] ^= 1;
            mi ^= 1;
            if ((rc = s->launched())) return rc;
            continue;
        }
        // three levels: level l does 2^(L-1-l) substeps (coreLB.hpp:1310-1331),
        // then the interfaces of levels (1, 2) and (2, 3).  The levels do not
        // meet between interfaces; level 1 reads only the (1, 2) interface and
        // levels 2 and 3 read both.  After the first step of a call, launch g
        // runs level 1's substep g and:
        //   g = 0: the previous step's (1, 2) interface, whose fine values
        //          update their cells (their first substep) and whose coarse
        //          values are written as overrides;
        //   g = 1: the previous step's (2, 3) interface, as overrides;
        //   g = 2: the first substeps of levels 2 and 3;
        //   g = 3: level 2's second substep --
        // four launches per coarse step.  The interfaces read the previous
        // step's stored fields while the substeps write the step's own (two
        // sets, mi).  The first step of a call runs substep g of every level
        // that has one in launch g.
        const bool pending = !first && Lmax == 3;
        for (int g = 0; g < (1 << (Lmax - 1)); g++) {
            // the interface specs read the carried states before this launch
            const IfSpec f12 = pending && g == 0 ? iface(0, mi ^ 1) : IfSpec{};
            const IfSpec f23 = pending && g == 1 ? iface(1, mi ^ 1) : IfSpec{};
            LaunchCtx c[MTS_MAX_LEVELS];
            MbSpec a[MTS_MAX_LEVELS];
            int n = 0, mac = 0;
            for (int l = 0; l < Lmax; l++) {
                const int r = pending && l > 0 ? g - 2 : g;  // substep of level l in launch g
                if (r < 0 || r >= (1 << (Lmax - 1 - l))) continue;
                auto& C = s->lc[l];
                c[n] = s->lctx(l);
                c[n].cimg = r == 0 ? &C.ovr : &C.plain;
                a[n].Ain = C.A[cur[l]];
                a[n].Aout = C.A[cur[l] ^ 1];
                a[n].m = C.m[mi];
                if (r == 0) {
                    a[n].fcan = s->lv[l].d_f;
                    a[n].usq_old = last ? s->lv[l].d_usq_old : nullptr;
                } else {
                    mac |= 1 << n;
                }
                cur[l] ^= 1;
                n++;
            }
            const int mode = g == 0 && first ? 0 : 1;
            if (pending && g == 0) {
                LaunchCtx c2 = s->lctx(1);
                c2.cimg = &s->lc[1].ovr;
                rc = launch_cinterface_step_fine(c[0], a[0], c2, f12);
            } else if (pending && g == 1) {
                LaunchCtx c2 = s->lctx(1), c3 = s->lctx(2);
                c2.cimg = &s->lc[1].ovr;
                c3.cimg = &s->lc[2].ovr;
                rc = launch_cinterface_beside(c2, c3, f23, c[0], a[0], mode, mac & 1);
            } else {
                rc = n == 1 ? launch_cmultistep(c[0], mode, mac & 1, a[0]) : launch_cmultistep3(n, c, a, mode, mac);
            }
            if (rc) return fail(rc, "operator combination has no kernel");
        }
        mi ^= 1;
        if ((rc = s->launched())) return rc;
    }
    if (Lmax == 3) {  // the last step's interfaces, as overrides for the export
        LaunchCtx c1 = s->lctx(0), c2 = s->lctx(1), c3 = s->lctx(2);
        c1.cimg = &s->lc[0].ovr;
        c2.cimg = &s->lc[1].ovr;
        c3.cimg = &s->lc[2].ovr;
        if ((rc = launch_cinterface(c1, c2, iface(0, mi ^ 1))) || (rc = launch_cinterface(c2, c3, iface(1, mi ^ 1))))
            return fail(rc, "interface has no kernel");
    }
    if (Lmax == 2) {  // the last step's interface, as overrides for the export
        LaunchCtx c1 = s->lctx(0), c2 = s->lctx(1);
        c1.cimg = &s->lc[0].ovr;
        c2.cimg = &s->lc[1].ovr;
        if ((rc = launch_cinterface(c1, c2, iface(0, mi ^ 1)))) return fail(rc, "interface has no kernel");
    }
    // export: post-stream state with the interface overrides -> canonical
    for (int l = 0; l < Lmax; l++) {
        auto& C = s->lc[l];
        auto& L = s->lv[l];
        LaunchCtx c = s->lctx(l);
        c.cimg = &C.ovr;
        TransportSpec z;
        z.Ain = C.A[cur[l]];
        z.fcan = L.d_f;
        z.rho = L.d_rho;
        z.ux = L.d_ux;
        z.uy = L.d_uy;
        z.usq = L.d_usq;
        launch_ctransport(c, 2, z);
    }
    if ((rc = s->launched())) return rc;
    double s1 = 0, s2 = 0;
    for (int l = 0; l < Lmax; l++) {
        double r[4];
        const LaunchCtx c = s->lctx(l);
        launch_reduce(c, 0, s->lv[l].d_usq, s->lv[l].d_usq_old, nullptr, s->d_part, s->nparts, s->d_red);
        if ((rc = s->launched())) return rc;
        HIP_TRY(hipMemcpyAsync(r, s->d_red, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        s1 += r[0];
        s2 += r[1];
    }
    return set_residual(&s->diff, s1 / s2);
}

static int multi_transport_phases(mts_sim s, int nsteps);

static int multi_transport(mts_sim_s* s, int nsteps) {
    if (!s->lc.empty()) return multi_transport32(s, nsteps);
    if (s->prm.wall == MTS_WALL_MDBB) return multi_transport_phases(s, nsteps);
    const int Lmax = (int)s->lv.size();
    int rc;
    for (int step = 0; step < nsteps; step++) {
        const bool last = step == nsteps - 1;
        if (last)
            for (auto& L : s->lv)
                HIP_TRY(hipMemcpyAsync(L.d_usq_old, L.d_usq, L.P * 8, hipMemcpyDeviceToDevice, s->stream));
        for (int l = Lmax - 1; l >= 0; l--) {
            const LaunchCtx c = s->lctx(l);
            auto& L = s->lv[l];
            for (int rep = 0; rep < (1 << (Lmax - 1 - l)); rep++) {
                TransportSpec a;
                a.fcan = L.d_f;
                a.Aout = L.d_A;
                a.min_rho = L.d_rho;
                a.min_ux = L.d_ux;
                a.min_uy = L.d_uy;
                a.min_usq = L.d_usq;
                if ((rc = launch_transport(c, 0, a))) return fail(rc, "operator combination has no kernel");
                TransportSpec b;
                b.Ain = L.d_A;
                b.fcan = L.d_f;
                launch_transport(c, 2, b);
            }
        }
        // interpolateBlockInterface(l, l + 1) for l = 1 .. Lmax - 1 in turn
        for (int p = 0; p + 1 < Lmax; p++) {
            auto &L1 = s->lv[p], &L2 = s->lv[p + 1];
            launch_coarse_interp(s->stream, s->crow_off[p + 1] - s->crow_off[p], s->d_crows + s->crow_off[p], L2.d_f,
                                 L2.stride, L1.d_f, L1.stride, L2.d_rho, L2.d_ux, L2.d_uy, L2.d_usq);
            launch_fine_interp(s->stream, s->frow_off[p + 1] - s->frow_off[p], s->d_frows + s->frow_off[p], L1.d_f,
                               L1.stride, L2.d_f, L2.stride, L1.d_rho, L1.d_ux, L1.d_uy, L1.d_usq);
        }
        for (int l = 0; l < Lmax; l++) {
            auto& L = s->lv[l];
            launch_macro(s->lctx(l), L.d_f, L.d_rho, L.d_ux, L.d_uy, L.d_usq);
        }
        if ((rc = s->launched())) return rc;
    }
    double s1 = 0, s2 = 0;
    for (int l = 0; l < Lmax; l++) {
        double r[4];
        const LaunchCtx c = s->lctx(l);
        launch_reduce(c, 0, s->lv[l].d_usq, s->lv[l].d_usq_old, nullptr, s->d_part, s->nparts, s->d_red);
        if ((rc = s->launched())) return rc;
        HIP_TRY(hipMemcpyAsync(r, s->d_red, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        s1 += r[0];
        s2 += r[1];
    }
    return set_residual(&s->diff, s1 / s2);
}

// Downloads one per-cell field of a multi-block lattice into compact order.
static int multi_ensure_phase(mts_sim s);

static int multi_get(mts_sim_s* s, int field, double* out) {
    const bool pop = field == MTS_F_F || field == MTS_F_FTEMP || field == MTS_F_FEQ || field == MTS_F_FNEQ ||
                     field == MTS_F_FB;
    const int q = pop ? 9 : 1;
    int rc;
    if ((field == MTS_F_FTEMP || field == MTS_F_FEQ || field == MTS_F_FNEQ || field == MTS_F_FB ||
         field == MTS_F_RHO_OLD) && (rc = multi_ensure_phase(s)))
        return rc;
    for (size_t l = 0; l < s->lv.size(); l++) {
        auto& L = s->lv[l];
        const double* src = field == MTS_F_F ? L.d_f : field == MTS_F_RHO ? L.d_rho : field == MTS_F_UX ? L.d_ux
                          : field == MTS_F_UY ? L.d_uy : field == MTS_F_USQ ? L.d_usq : field == MTS_F_USQ_OLD ? L.d_usq_old
                          : field == MTS_F_FTEMP ? L.d_ftemp : field == MTS_F_FEQ ? L.d_feq : field == MTS_F_FNEQ ? L.d_fneq
                          : field == MTS_F_FB ? L.d_fb : field == MTS_F_RHO_OLD ? L.d_rho_old : nullptr;
        if (!src) return fail(MTS_ERR_UNSUPPORTED, "field not held by multi-block lattices");
        std::vector<double> h((size_t)(q == 9 ? L.stride * 9 : L.P));
        HIP_TRY(hipMemcpyAsync(h.data(), src, h.size() * 8, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        const LevelImage& I = s->limg[l];
        for (size_t k = 0; k < I.cells.size(); k++)
            for (int j = 0; j < q; j++)
                out[I.cells[k] * q + j] = h[(size_t)(q == 9 ? j * L.stride : 0) + I.sites[k]];
    }
    return MTS_OK;
}

extern "C" {

void mts_default_params(mts_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->dx = 1.0e-9;
    p->Fbody = 1e-8;
    p->T = 300.0;
    p->P = 2.0e6;
    p->Gff = -1.0;
    p->Gfs = -12.7;
    p->collision = MTS_COLL_RMRT;
    p->wall = MTS_WALL_SRBB;
    p->precision = MTS_FP64;
    p->multiblock = 0;
    p->device = 0;
    p->periodic_y = 0;
    p->batch_steps = 1;
}

const char* mts_last_error(void) { return g_err.c_str(); }

int mts_globals(const mts_params* p, double* out) {
    if (!p || !out) return fail(MTS_ERR_ARG, "null argument");
    Physics ph;
    ph.Cl = p->dx;
    ph.Fbody = p->Fbody;
    ph.T_phy = p->T;
    ph.P_phy = p->P;
    compute_physics(ph);
    const double v[10] = {ph.rhoin, ph.mfp, ph.Cf, ph.T_lu, ph.a_alpha_lu, ph.Z, ph.Crho, ph.Cu, ph.Cl, ph.cubic_ok ? 1.0 : 0.0};
    std::memcpy(out, v, sizeof(v));
    return ph.cubic_ok ? MTS_OK : fail(MTS_ERR_PHYSICS, "PR-EOS: no single positive real root at this (T, P)");
}

int mts_load_existing_model(const char* folder, int n, int32_t* pore, double* kn, double* locpore) {
    if (!folder || n <= 0 || !pore || !kn || !locpore) return fail(MTS_ERR_ARG, "null argument");
    const char* names[3] = {"pore.dat", "Kn.dat", "localporesize.dat"};
    std::vector<double> data((size_t)n * n);
    for (int k = 0; k < 3; k++) {
        std::string fn = std::string(folder) + names[k];
        std::ifstream in(fn);
        if (!in.is_open()) return fail(MTS_ERR_IO, "[ERROR]: File not found. (" + fn + ")");
        for (size_t i = 0; i < data.size(); i++)
            if (!(in >> data[i])) return fail(MTS_ERR_IO, "[ERROR]: Insufficient data in the file. (" + fn + ")");
        // column-major on disk: value (row i, col j) = data[j*n + i]
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                const double v = data[(size_t)j * n + i];
                if (k == 0) pore[(size_t)i * n + j] = (int32_t)v;
                else if (k == 1) kn[(size_t)i * n + j] = v;
                else locpore[(size_t)i * n + j] = v;
            }
    }
    return MTS_OK;
}

int mts_create(int nx, int ny, const int32_t* pore, const double* kn, const double* locpore, const mts_params* p,
               mts_sim* out) {
    if (!out) return fail(MTS_ERR_ARG, "null handle pointer");
    *out = nullptr;
    if (!pore || !kn || !locpore) return fail(MTS_ERR_ARG, "null geometry array");
    mts_params prm;
    if (p) prm = *p;
    else mts_default_params(&prm);
    if (prm.precision != MTS_FP64 && prm.precision != MTS_FP32DEV) return fail(MTS_ERR_ARG, "bad precision");
    if (prm.collision < 0 || prm.collision > 2) return fail(MTS_ERR_ARG, "Incorrect keyword for collision");
    if (prm.wall < 0 || prm.wall > 2) return fail(MTS_ERR_ARG, "Incorrect keyword for wall boundary method");
    if ((int64_t)nx * ny > (int64_t)1 << 31) return fail(MTS_ERR_ARG, "lattice larger than 2^31 sites");
    mts_sim_s* s = new (std::nothrow) mts_sim_s();
    if (!s) return fail(MTS_ERR_OOM, "host allocation");
    s->prm = prm;
    s->Fbody = prm.Fbody;
    s->ph.Cl = prm.dx;
    s->ph.Fbody = prm.Fbody;
    s->ph.T_phy = prm.T;
    s->ph.P_phy = prm.P;
    compute_physics(s->ph);
    if (!s->ph.cubic_ok) {
        delete s;
        return fail(MTS_ERR_PHYSICS, "PR-EOS: no single positive real root at this (T, P)");
    }
    std::string err;
    if (prm.multiblock) {
        if (prm.wall == MTS_WALL_MDBB && prm.precision == MTS_FP32DEV) {
            delete s;
            return fail(MTS_ERR_UNSUPPORTED, "MDBB walls on multi-block lattices run in FP64 (precision MTS_FP64)");
        }
        int rc = create_multi(s, nx, ny, pore, kn, locpore, err);
        if (rc) {
            delete s;
            return err.empty() ? rc : fail(rc, err);  // empty: the inner fail() holds the message
        }
        *out = s;
        return MTS_OK;
    }
    int rc = build_single_grid(nx, ny, pore, kn, locpore, s->ph, prm.periodic_y != 0, s->geo, err);
    if (rc) {
        delete s;
        return err.empty() ? rc : fail(rc, err);
    }
    auto bail = [&](int code) {
        delete s;
        return code;
    };
    hipError_t e = hipSetDevice(prm.device);
    if (e != hipSuccess) return bail(fail(MTS_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)));
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(fail(MTS_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
    const Geometry& g = s->geo;
    s->P = (int64_t)g.nx * g.ny;
    const size_t np = (size_t)g.stride * 9;
    const bool fp32 = prm.precision == MTS_FP32DEV;
    if (fp32) build_compact(g, s->cimg_h);
    const size_t nA = fp32 ? (size_t)s->cimg_h.cstride * 9 * 4 : np * 8;
    s->nparts = (int)std::min<int64_t>(1024, (s->P + 255) / 256);
    if ((rc = s->alloc(&s->d_meta, s->P)) || (rc = s->alloc(&s->d_pidx, s->P)) ||
        (rc = s->alloc(&s->d_params, g.params.size())) || (rc = s->alloc(&s->d_cellPos, g.ncell)) ||
        (rc = s->alloc((char**)&s->d_A, nA)) || (rc = s->alloc((char**)&s->d_B, nA)) ||
        (rc = s->alloc(&s->d_f, np)) || (rc = s->alloc(&s->d_rho, s->P)) || (rc = s->alloc(&s->d_rho_old, s->P)) ||
        (rc = s->alloc(&s->d_ux, s->P)) || (rc = s->alloc(&s->d_uy, s->P)) || (rc = s->alloc(&s->d_usq, s->P)) ||
        (rc = s->alloc(&s->d_usq_old, s->P)) || (rc = s->alloc(&s->d_part, (size_t)4 * s->nparts)) ||
        (rc = s->alloc(&s->d_red, 4)) || (rc = s->alloc(&s->d_neg, 1)))
        return bail(rc);
    e = hipMemcpyAsync(s->d_meta, g.meta.data(), s->P * 4, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->d_pidx, g.pidx.data(), s->P * 2, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_params, g.params.data(), g.params.size() * sizeof(CellParam), hipMemcpyHostToDevice,
                           s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_cellPos, g.cellPos.data(), g.ncell * 4, hipMemcpyHostToDevice, s->stream);
    if (e != hipSuccess) return bail(fail(MTS_ERR_HIP, std::string("upload: ") + hipGetErrorString(e)));
    if (fp32) {
        const float* cpar;
        if ((rc = upload_cpar(s, g.params, &cpar)) || (rc = upload_compact(s, s->cimg_h, cpar, s->cimg))) return bail(rc);
        upload_spec(s->cimg_h.spec, s->stream);
        s->use_cimg = true;
    }
    upload_icsr(s->stream);
    upload_xynorm(s->stream);
    *out = s;
    rc = mts_initialize(s);
    if (rc) {
        *out = nullptr;
        return bail(rc);
    }
    return MTS_OK;
}

int mts_destroy(mts_sim s) {
    delete s;
    return MTS_OK;
}

}  // extern "C"

// Host grid handle: the single-block Geometry or the quadtree MultiGeometry
// presented as one cell list.
struct mts_grid_s {
    bool multi = false;
    Geometry single;
    MultiGeometry mg;
    Physics ph;
    int64_t n() const { return multi ? mg.ncell : single.ncell; }
};

static int build_grid(int nx, int ny, const int32_t* pore, const double* kn, const double* locpore, const mts_params& prm,
                      const Physics& ph, mts_grid_s& g, std::string& err) {
    g.multi = prm.multiblock != 0;
    if (g.multi)
        return build_multi_grid(nx, ny, pore, kn, locpore, ph, prm.periodic_y != 0, quadtree_levels(prm.multiblock), g.mg,
                                err);
    return build_single_grid(nx, ny, pore, kn, locpore, ph, prm.periodic_y != 0, g.single, err);
}

extern "C" {

int mts_grid_create(int nx, int ny, const int32_t* pore, const double* kn, const double* locpore, const mts_params* p,
                    mts_grid* out) {
    if (!out || !pore || !kn || !locpore) return fail(MTS_ERR_ARG, "null argument");
    *out = nullptr;
    mts_params prm;
    if (p) prm = *p;
    else mts_default_params(&prm);
    Physics ph;
    ph.Cl = prm.dx;
    ph.Fbody = prm.Fbody;
    ph.T_phy = prm.T;
    ph.P_phy = prm.P;
    compute_physics(ph);
    if (!ph.cubic_ok) return fail(MTS_ERR_PHYSICS, "PR-EOS: no single positive real root at this (T, P)");
    mts_grid_s* g = new (std::nothrow) mts_grid_s();
    if (!g) return fail(MTS_ERR_OOM, "host allocation");
    std::string err;
    g->ph = ph;
    const int rc = build_grid(nx, ny, pore, kn, locpore, prm, ph, *g, err);
    if (rc) {
        delete g;
        return err.empty() ? rc : fail(rc, err);
    }
    *out = g;
    return MTS_OK;
}

int mts_grid_destroy(mts_grid g) {
    delete g;
    return MTS_OK;
}

int mts_grid_size(mts_grid g, int64_t* n, int32_t* max_level) {
    if (!g) return fail(MTS_ERR_ARG, "null grid");
    if (n) *n = g->n();
    if (max_level) *max_level = g->multi ? g->mg.maxLevel : 1;
    return MTS_OK;
}

int mts_grid_arrays(mts_grid g, int64_t* gridID, int32_t* gridIJ, int32_t* level, int64_t* connect, int32_t* bnd,
                    int32_t* type, int32_t* isBuffer, double* taus) {
    if (!g) return fail(MTS_ERR_ARG, "null grid");
    if (taus && g->multi && g->mg.taus.empty()) {
        std::string err;
        const int rc = multi_grid_taus(g->ph, g->mg, err);
        if (rc) return err.empty() ? rc : fail(rc, err);
    }
    const int64_t n = g->n();
    for (int64_t c = 0; c < n; c++) {
        if (g->multi) {
            const MultiGeometry& m = g->mg;
            if (gridID) gridID[c] = m.gridID[c];
            if (gridIJ) { gridIJ[2 * c] = m.cx[c]; gridIJ[2 * c + 1] = m.cy[c]; }
            if (level) level[c] = m.level[c];
            if (bnd) bnd[c] = m.bnd[c];
            if (type) type[c] = m.type[c];
            if (isBuffer) isBuffer[c] = m.isBuffer[c];
            if (taus) taus[c] = m.taus[c];
            if (connect)
                for (int k = 0; k < 9; k++) connect[9 * c + k] = m.connect[9 * c + k];
        } else {
            const Geometry& s = g->single;
            const int64_t p = s.cellPos[c];
            if (gridID) gridID[c] = (p % s.nx) + (p / s.nx) * (int64_t)s.ny;
            if (gridIJ) { gridIJ[2 * c] = (int32_t)(p % s.nx); gridIJ[2 * c + 1] = (int32_t)(p / s.nx); }
            if (level) level[c] = 1;
            if (bnd) bnd[c] = s.bndTypes[c];
            if (type) type[c] = 0;
            if (isBuffer) isBuffer[c] = 0;
            if (taus) taus[c] = s.taus[c];
            if (connect)
                for (int k = 0; k < 9; k++) connect[9 * c + k] = s.gridConnect[9 * c + k];
        }
    }
    return MTS_OK;
}

int mts_grid_rows(mts_grid g, int which, int64_t* rows, int64_t* count) {
    if (!g || !count) return fail(MTS_ERR_ARG, "null argument");
    static const std::vector<int64_t> none;
    const MultiGeometry& G = g->mg;
    const std::vector<int64_t>* v = &none;
    int64_t r0 = 0, r1 = 0;
    if (which >= 0 && which <= 2) {
        v = !g->multi ? &none : which == 0 ? &G.parentChild : which == 1 ? &G.coarseRows : &G.fineRows;
        r1 = (int64_t)v->size() / 5;
    } else {
        // one buffer type: coarse 10 (l + 1) + l or fine 10 l + l + 1
        const int hi = which / 10, lo = which % 10;
        const bool coarse = hi == lo + 1, fine = lo == hi + 1;
        if (!(coarse || fine) || std::min(hi, lo) < 1 || std::max(hi, lo) > MTS_MAX_LEVELS)
            return fail(MTS_ERR_ARG, "which: 0, 1, 2 or a buffer type (21, 32, 12, 23)");
        const int p = std::min(hi, lo) - 1;
        if (g->multi && p < G.maxLevel - 1) {
            const std::vector<int64_t>& cnt = coarse ? G.coarseCount : G.fineCount;
            v = coarse ? &G.coarseRows : &G.fineRows;
            for (int q = 0; q < p; q++) r0 += cnt[q];
            r1 = r0 + cnt[p];
        }
    }
    *count = r1 - r0;
    if (rows) std::copy(v->begin() + 5 * r0, v->begin() + 5 * r1, rows);
    return MTS_OK;
}

int mts_cell_count(mts_sim s, int64_t* n) {
    if (!s || !n) return fail(MTS_ERR_ARG, "null argument");
    *n = s->geo.ncell;
    return MTS_OK;
}

int mts_grid_info(mts_sim s, int64_t* gridID, int32_t* gridIJ, int32_t* level, int64_t* connect, int32_t* bnd) {
    if (!s) return fail(MTS_ERR_ARG, "null handle");
    if (s->multi) {
        const MultiGeometry& m = s->mg;
        for (int64_t c = 0; c < m.ncell; c++) {
            if (gridID) gridID[c] = m.gridID[c];
            if (gridIJ) { gridIJ[2 * c] = m.cx[c]; gridIJ[2 * c + 1] = m.cy[c]; }
            if (level) level[c] = m.level[c];
            if (bnd) bnd[c] = m.bnd[c];
            if (connect)
                for (int k = 0; k < 9; k++) connect[9 * c + k] = m.connect[9 * c + k];
        }
        return MTS_OK;
    }
    const Geometry& g = s->geo;
    for (int64_t c = 0; c < g.ncell; c++) {
        const int64_t p = g.cellPos[c];
        if (gridID) gridID[c] = p;  // sub2ind({ny,nx,1},{x,y,0}) == y*nx + x for square lattices
        if (gridIJ) {
            gridIJ[2 * c] = (int32_t)(p % g.nx);
            gridIJ[2 * c + 1] = (int32_t)(p / g.nx);
        }
        if (level) level[c] = 1;
        if (bnd) bnd[c] = g.bndTypes[c];
        if (connect)
            for (int k = 0; k < 9; k++) connect[9 * c + k] = g.gridConnect[9 * c + k];
    }
    return MTS_OK;
}

int mts_links(mts_sim s, int which, int64_t* id, int64_t* loc, int32_t* ic, int64_t* count) {
    if (!s || !count || which < 0 || which > 2) return fail(MTS_ERR_ARG, "bad argument");
    if (s->multi) {
        // decode the level images' link words as held on the device
        std::vector<std::vector<uint32_t>> meta(s->lv.size());
        std::vector<std::vector<int64_t>> cellAt(s->lv.size());
        for (size_t l = 0; l < s->lv.size(); l++) {
            meta[l].resize(s->lv[l].P);
            HIP_TRY(hipMemcpyAsync(meta[l].data(), s->lv[l].d_meta, s->lv[l].P * 4, hipMemcpyDeviceToHost, s->stream));
            cellAt[l].assign(s->lv[l].P, -1);
            for (size_t k = 0; k < s->limg[l].cells.size(); k++) cellAt[l][s->limg[l].sites[k]] = s->limg[l].cells[k];
        }
        HIP_TRY(hipStreamSynchronize(s->stream));
        std::vector<int64_t> a, b;
        std::vector<int32_t> c;
        const MultiGeometry& G = s->mg;
        for (int64_t cc = 0; cc < G.ncell; cc++) {
            const int l = G.level[cc] - 1;
            const int nxl = s->lv[l].nx, nyl = s->lv[l].ny;
            const int x = G.cx[cc] >> l, y = G.cy[cc] >> l;
            const uint32_t w = meta[l][(size_t)y * nxl + x];
            for (int k = 0; k < 9; k++) {
                if (k == 0) {
                    if (which == 0) { a.push_back(cc); b.push_back(cc); c.push_back(0); }
                    continue;
                }
                const int xq = ((x + EX[k]) % nxl + nxl) % nxl, yq = ((y + EY[k]) % nyl + nyl) % nyl;
                const int64_t q = (int64_t)yq * nxl + xq;
                const bool xseam = !G.torus && (x + EX[k] < 0 || x + EX[k] >= nxl);
                const bool linked = cellAt[l][q] >= 0 && ((meta[l][q] >> (IN_SHIFT + k - 1)) & 1u);
                if ((which == 0 && linked && !xseam) || (which == 2 && linked && xseam)) {
                    a.push_back(cc);
                    b.push_back(cellAt[l][q]);
                    c.push_back(k);
                }
                if (which == 1 && ((w >> (WALL_SHIFT + k - 1)) & 1u)) { a.push_back(cc); b.push_back(-1); c.push_back(k); }
            }
        }
        *count = (int64_t)a.size();
        if (id) std::copy(a.begin(), a.end(), id);
        if (loc) std::copy(b.begin(), b.end(), loc);
        if (ic) std::copy(c.begin(), c.end(), ic);
        return MTS_OK;
    }
    Geometry g = s->geo;  // decode from the words held on the device
    HIP_TRY(hipMemcpyAsync(g.meta.data(), s->d_meta, s->P * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<int64_t> a, b;
    std::vector<int32_t> c;
    decode_links(g, which, a, b, c);
    *count = (int64_t)a.size();
    if (id) std::copy(a.begin(), a.end(), id);
    if (loc) std::copy(b.begin(), b.end(), loc);
    if (ic) std::copy(c.begin(), c.end(), ic);
    return MTS_OK;
}

int mts_set_operators(mts_sim s, int collision, int wall) {
    if (!s) return fail(MTS_ERR_ARG, "null handle");
    if (collision < 0 || collision > 2)
        return fail(MTS_ERR_ARG,
                    "Incorrect keyword for collision: available options are\n\t 2 for regularized MRT\n\t 1 for simple MRT\n\t 0 for simple BGK\n");
    if (wall < 0 || wall > 2)
        return fail(MTS_ERR_ARG,
                    "Incorrect keyword for wall boundary method: available options are\n\t 2 for SRBBWall\n\t 1 for MDBBWall\n\t 0 for BBWall\n");
    s->prm.collision = collision;
    s->prm.wall = wall;
    return MTS_OK;
}

int mts_set_body_force(mts_sim s, double Fbody) {
    if (!s) return fail(MTS_ERR_ARG, "null handle");
    s->Fbody = Fbody;
    return MTS_OK;
}

int mts_initialize(mts_sim s) {
    if (!s) return fail(MTS_ERR_ARG, "null handle");
    if (s->multi) {
        for (size_t l = 0; l < s->lv.size(); l++) {
            LaunchCtx c = s->lctx((int)l);
            launch_init(c, s->lv[l].d_f);
            launch_macro(c, s->lv[l].d_f, s->lv[l].d_rho, s->lv[l].d_ux, s->lv[l].d_uy, s->lv[l].d_usq);
        }
        s->diff = 1.0;
        return s->launched();
    }
    LaunchCtx c = s->ctx();
    launch_init(c, s->d_f);
    launch_macro(c, s->d_f, s->d_rho, s->d_ux, s->d_uy, s->d_usq);
    int rc = s->launched();
    if (rc) return rc;
    s->diff = 1.0;
    return MTS_OK;
}

// ---- unfused per-step fallback (operators without a fused kernel: MDBB)
static int transport_step_phases(mts_sim s) {
    int rc;
    if ((rc = mts_method(s, MTS_M_COPY_USQ_OLD, 1)) || (rc = mts_method(s, MTS_M_EQUILIBRIUM, 1)) ||
        (rc = mts_method(s, MTS_M_EVOLUTION, 1)) || (rc = mts_method(s, MTS_M_MACROSCOPIC, 1)) ||
        (rc = mts_method(s, MTS_M_VELOCITY_DIFF, 1)))
        return rc;
    return MTS_OK;
}

int mts_step_transport(mts_sim s, int nsteps) {
    if (!s) return fail(MTS_ERR_ARG, "null handle");
    if (nsteps <= 0) return fail(MTS_ERR_ARG, "nsteps must be positive");
    if (s->multi) return multi_transport(s, nsteps);
    const bool mdbb = s->prm.wall == MTS_WALL_MDBB;
    if (mdbb && s->prm.precision == MTS_FP64) {  // FP64: the per-method kernels
        for (int i = 0; i < nsteps; i++) {
            int rc = transport_step_phases(s);
            if (rc) return rc;
        }
        return MTS_OK;
    }
    int rc;
    if (mdbb && !s->d_M[0]) {
        const size_t nm = (size_t)4 * s->cimg.cstride;
        if ((rc = s->alloc(&s->d_M[0], nm)) || (rc = s->alloc(&s->d_M[1], nm))) return rc;
    }
    LaunchCtx c = s->ctx();
    TransportSpec t;
    int mc = 0;  // MDBB state: written by a pass's collision, read by the next pass's gather
    // pass 0: collide the canonical state
    t.fcan = s->d_f;
    t.Aout = s->d_A;
    t.usq_old = nsteps == 1 ? s->d_usq_old : nullptr;
    t.Mout = s->d_M[0];
    if ((rc = launch_transport(c, 0, t))) return fail(rc, "operator combination has no fused kernel");
    for (int i = 1; i < nsteps; i++) {
        TransportSpec u;
        u.Ain = s->d_A;
        u.Aout = s->d_B;
        u.usq_old = (i == nsteps - 1) ? s->d_usq_old : nullptr;
        u.Min = s->d_M[mc];
        u.Mout = s->d_M[mc ^ 1];
        s->tick(c);
        launch_transport(c, 1, u);
        s->tock(c);
        std::swap(s->d_A, s->d_B);
        mc ^= 1;
    }
    TransportSpec z;
    z.Min = s->d_M[mc];
    z.Ain = s->d_A;
    z.fcan = s->d_f;
    z.rho = s->d_rho;
    z.ux = s->d_ux;
    z.uy = s->d_uy;
    z.usq = s->d_usq;
    launch_transport(c, 2, z);
    if ((rc = s->launched())) return rc;
    double r[4];
    if ((rc = s->reduce(0, s->d_usq, s->d_usq_old, nullptr, r))) return rc;
    return set_residual(&s->diff, r[0] / r[1]);
}

static int sc_steps(mts_sim s, int kind, int nsteps) {
    if (!s) return fail(MTS_ERR_ARG, "null handle");
    if (s->multi) return fail(MTS_ERR_UNSUPPORTED, "multi-block lattices provide the fused transport step only");
    if (nsteps <= 0) return fail(MTS_ERR_ARG, "nsteps must be positive");
    int rc;
    if ((rc = s->ensure_sc())) return rc;
    LaunchCtx c = s->ctx();
    HIP_TRY(hipMemsetAsync(s->d_neg, 0, sizeof(int), s->stream));
    if (kind == 0 && nsteps == 1)  // rho_old = rho
        HIP_TRY(hipMemcpyAsync(s->d_rho_old, s->d_rho, s->P * 8, hipMemcpyDeviceToDevice, s->stream));
    for (int i = 0; i < nsteps; i++) {
        ScSpec sp;
        sp.fcan = s->d_f;
        sp.Ain = s->d_A;
        sp.Aout = i == 0 ? s->d_A : s->d_B;
        if (i == nsteps - 1) {
            sp.rho = s->d_rho;
            sp.ux = s->d_ux;
            sp.uy = s->d_uy;
            sp.usq = s->d_usq;
            sp.psi = s->d_psi;
            sp.Fx = s->d_Fx;
            sp.Fy = s->d_Fy;
            sp.psi_negative = s->d_neg;
        } else if (i == nsteps - 2 && kind == 0) {
            // the density of the step before the last is the residual's rho_old
            sp.rho = s->d_rho_old;
            sp.ux = s->d_ux;
            sp.uy = s->d_uy;
            sp.usq = s->d_usq;
        }
        if (i > 0) s->tick(c);
        launch_shanchen(c, kind, i == 0 ? 0 : 1, sp);
        s->tock(c);
        if (i > 0) std::swap(s->d_A, s->d_B);
    }
    ScSpec z;
    z.Ain = s->d_A;
    z.fcan = s->d_f;
    launch_shanchen(c, kind, 2, z);
    if ((rc = s->launched())) return rc;
    double r[4];
    if (kind == 0) {
        if ((rc = s->reduce(1, s->d_rho, s->d_rho_old, nullptr, r))) return rc;
        double d = r[0];
        if (d < 1.0e-15) d = 1.0e-5;  // coreLB.hpp:1280
        s->diff = std::isfinite(r[1]) ? d : r[1];  // the max drops NaN, the sum does not
    } else {
        // coreLB.hpp:1248-1267 with usq_old never refreshed by this loop
        if ((rc = s->reduce(0, s->d_usq, s->d_usq_old, nullptr, r))) return rc;
        s->diff = r[0] / r[1];
    }
    if (!std::isfinite(s->diff)) return fail(MTS_ERR_NONFINITE, "residual is not finite");
    return MTS_OK;
}

int mts_step_adsorption(mts_sim s, int nsteps) { return sc_steps(s, 0, nsteps); }
int mts_step_coupled(mts_sim s, int nsteps) { return sc_steps(s, 1, nsteps); }

int mts_residual(mts_sim s, double* diff) {
    if (!s || !diff) return fail(MTS_ERR_ARG, "null argument");
    *diff = s->diff;
    return MTS_OK;
}

// ---- per-method calls on a multi-block lattice: each level's dense site
// image takes the single-block kernels.  A method with a level
// (coreLB.hpp:678-977) runs on that level's image; a method over every cell
// runs on every image; the regularised collision of one level first forms
// fneq of the others (regularize() runs over all cells, coreLB.hpp:744-792).
static int multi_ensure_phase(mts_sim s) {
    int rc;
    for (auto& V : s->lv) {
        if (V.d_ftemp) continue;
        const size_t np = (size_t)V.stride * 9;
        if ((rc = s->alloc(&V.d_ftemp, np)) || (rc = s->alloc(&V.d_feq, np)) || (rc = s->alloc(&V.d_fneq, np)) ||
            (rc = s->alloc(&V.d_fb, np)) || (rc = s->alloc(&V.d_rho_old, (size_t)V.P)))
            return rc;
        HIP_TRY(hipMemsetAsync(V.d_ftemp, 0, np * 8, s->stream));
        HIP_TRY(hipMemsetAsync(V.d_feq, 0, np * 8, s->stream));
        HIP_TRY(hipMemsetAsync(V.d_fneq, 0, np * 8, s->stream));
        HIP_TRY(hipMemsetAsync(V.d_fb, 0, np * 8, s->stream));
        HIP_TRY(hipMemsetAsync(V.d_rho_old, 0, (size_t)V.P * 8, s->stream));
    }
    return MTS_OK;
}

static int multi_method(mts_sim s, int method, int lvl) {
    const int L = (int)s->lv.size();
    int rc;
    if ((rc = multi_ensure_phase(s))) return rc;
    auto arrays = [&](int l) {
        auto& V = s->lv[l];
        return PhaseArrays{V.d_f, V.d_ftemp, V.d_feq, V.d_fneq, V.d_fb, V.d_rho, V.d_rho_old, V.d_ux, V.d_uy,
                           V.d_usq, V.d_usq_old, nullptr, nullptr, nullptr};
    };
    auto on = [&](int l, int m) -> int {
        const int r = launch_phase(s->lctx(l), m, 1, arrays(l), s->d_neg);
        return r ? fail(r, "unknown method id") : s->launched();
    };
    using LevelDev = mts_sim_s::LevelDev;
    auto copy = [&](double* LevelDev::*dst, double* LevelDev::*src, bool pop) -> int {
        for (auto& V : s->lv)
            HIP_TRY(hipMemcpyAsync(V.*dst, V.*src, (pop ? (size_t)V.stride * 9 : (size_t)V.P) * 8,
                                   hipMemcpyDeviceToDevice, s->stream));
        return MTS_OK;
    };
    switch (method) {
        case MTS_M_COPY_F_FTEMP: return copy(&LevelDev::d_ftemp, &LevelDev::d_f, true);
        case MTS_M_COPY_F_FB: return copy(&LevelDev::d_fb, &LevelDev::d_f, true);
        case MTS_M_COPY_USQ_OLD: return copy(&LevelDev::d_usq_old, &LevelDev::d_usq, false);
        case MTS_M_COPY_RHO_OLD: return copy(&LevelDev::d_rho_old, &LevelDev::d_rho, false);
        case MTS_M_VELOCITY_DIFF: {  // sums over the cells of every level
            double s1 = 0, s2 = 0;
            for (int l = 0; l < L; l++) {
                double r[4];
                launch_reduce(s->lctx(l), 0, s->lv[l].d_usq, s->lv[l].d_usq_old, nullptr, s->d_part, s->nparts, s->d_red);
                if ((rc = s->launched())) return rc;
                HIP_TRY(hipMemcpyAsync(r, s->d_red, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
                HIP_TRY(hipStreamSynchronize(s->stream));
                s1 += r[0];
                s2 += r[1];
            }
            s->diff = s1 / s2;
            return MTS_OK;
        }
        case MTS_M_DENSITY_DIFF: {  // the largest over every level
            double m = 0;
            for (int l = 0; l < L; l++) {
                double r[4];
                launch_reduce(s->lctx(l), 1, s->lv[l].d_rho, s->lv[l].d_rho_old, nullptr, s->d_part, s->nparts, s->d_red);
                if ((rc = s->launched())) return rc;
                HIP_TRY(hipMemcpyAsync(r, s->d_red, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
                HIP_TRY(hipStreamSynchronize(s->stream));
                m = std::max(m, r[0]);
            }
            s->diff = m < 1.0e-15 ? 1.0e-5 : m;
            return MTS_OK;
        }
        case MTS_M_INTERPOLATE: {  // interpolateBlockInterface(lvl, lvl + 1)
            if (lvl < 1 || lvl >= L) return fail(MTS_ERR_ARG, "interpolateBlockInterface: no level pair (lvl, lvl + 1)");
            const int p = lvl - 1;
            auto &L1 = s->lv[p], &L2 = s->lv[p + 1];
            launch_coarse_interp(s->stream, s->crow_off[p + 1] - s->crow_off[p], s->d_crows + s->crow_off[p], L2.d_f,
                                 L2.stride, L1.d_f, L1.stride, L2.d_rho, L2.d_ux, L2.d_uy, L2.d_usq);
            launch_fine_interp(s->stream, s->frow_off[p + 1] - s->frow_off[p], s->d_frows + s->frow_off[p], L1.d_f,
                               L1.stride, L2.d_f, L2.stride, L1.d_rho, L1.d_ux, L1.d_uy, L1.d_usq);
            return s->launched();
        }
        case MTS_M_EVOLUTION: {  // coreLB.hpp:1310-1331, with the recursion and the interfaces
            if (lvl < 1 || lvl > L) return MTS_OK;
            for (int i = 0; i < (1 << (L - lvl)); i++)
                if ((rc = multi_method(s, MTS_M_COPY_F_FB, lvl)) || (rc = multi_method(s, MTS_M_COLLIDE, lvl)) ||
                    (rc = multi_method(s, MTS_M_COPY_F_FTEMP, lvl)) || (rc = multi_method(s, MTS_M_ZERODOWN, lvl)) ||
                    (rc = multi_method(s, MTS_M_STREAM, lvl)) || (rc = multi_method(s, MTS_M_APPLY_BC, lvl)))
                    return rc;
            if (lvl > 1 && (rc = multi_method(s, MTS_M_EVOLUTION, lvl - 1))) return rc;
            if (lvl != L && (rc = multi_method(s, MTS_M_INTERPOLATE, lvl))) return rc;
            return MTS_OK;
        }
        case MTS_M_EQUILIBRIUM:
        case MTS_M_REGULARIZE:
        case MTS_M_MACROSCOPIC:
            for (int l = 0; l < L; l++)
                if ((rc = on(l, method))) return rc;
            return MTS_OK;
        case MTS_M_COLLIDE:
            return multi_method(s, s->prm.collision == MTS_COLL_RMRT ? MTS_M_RMRT
                                   : s->prm.collision == MTS_COLL_MRT ? MTS_M_MRT : MTS_M_SRT, lvl);
        case MTS_M_APPLY_BC:
            if ((rc = multi_method(s, s->prm.wall == MTS_WALL_SRBB ? MTS_M_SRBBWALL
                                      : s->prm.wall == MTS_WALL_MDBB ? MTS_M_MDBBWALL : MTS_M_BBWALL, lvl)))
                return rc;
            return multi_method(s, MTS_M_PERIODIC, lvl);
        case MTS_M_SRT:
        case MTS_M_MRT:
        case MTS_M_RMRT:
        case MTS_M_ZERODOWN:
        case MTS_M_STREAM:
        case MTS_M_BBWALL:
        case MTS_M_SRBBWALL:
        case MTS_M_MDBBWALL:
        case MTS_M_PERIODIC:
            if (lvl < 1 || lvl > L) return MTS_OK;  // no cell has that level
            if (method == MTS_M_RMRT)
                for (int l = 0; l < L; l++)
                    if (l != lvl - 1 && (rc = on(l, MTS_M_REGULARIZE))) return rc;
            return on(lvl - 1, method);
        default:
            return fail(MTS_ERR_UNSUPPORTED, "multi-block lattices: the adsorption methods are single-block only");
    }
}

// The transport loop of mantiisApp.hpp:239-246 method by method (operators
// without a fused multi-block kernel: MDBB walls).
static int multi_transport_phases(mts_sim s, int nsteps) {
    const int L = (int)s->lv.size();
    int rc;
    for (int i = 0; i < nsteps; i++)
        if ((rc = multi_method(s, MTS_M_COPY_USQ_OLD, 1)) || (rc = multi_method(s, MTS_M_EQUILIBRIUM, 1)) ||
            (rc = multi_method(s, MTS_M_EVOLUTION, L)) || (rc = multi_method(s, MTS_M_MACROSCOPIC, 1)) ||
            (rc = multi_method(s, MTS_M_VELOCITY_DIFF, 1)))
            return rc;
    return set_residual(&s->diff, s->diff);
}

int mts_method(mts_sim s, int method, int lvl) {
    if (!s) return fail(MTS_ERR_ARG, "null handle");
    if (s->multi) return multi_method(s, method, lvl);
    int rc;
    if ((rc = s->ensure_phase()) || (rc = s->ensure_sc())) return rc;
    const size_t np = (size_t)s->geo.stride * 9 * 8;
    switch (method) {
        case MTS_M_COPY_F_FTEMP:
            HIP_TRY(hipMemcpyAsync(s->d_ftemp, s->d_f, np, hipMemcpyDeviceToDevice, s->stream));
            return MTS_OK;
        case MTS_M_COPY_F_FB:
            HIP_TRY(hipMemcpyAsync(s->d_fb, s->d_f, np, hipMemcpyDeviceToDevice, s->stream));
            return MTS_OK;
        case MTS_M_COPY_USQ_OLD:
            HIP_TRY(hipMemcpyAsync(s->d_usq_old, s->d_usq, s->P * 8, hipMemcpyDeviceToDevice, s->stream));
            return MTS_OK;
        case MTS_M_COPY_RHO_OLD:
            HIP_TRY(hipMemcpyAsync(s->d_rho_old, s->d_rho, s->P * 8, hipMemcpyDeviceToDevice, s->stream));
            return MTS_OK;
        case MTS_M_VELOCITY_DIFF: {
            double r[4];
            if ((rc = s->reduce(0, s->d_usq, s->d_usq_old, nullptr, r))) return rc;
            s->diff = r[0] / r[1];
            return MTS_OK;
        }
        case MTS_M_DENSITY_DIFF: {
            double r[4];
            if ((rc = s->reduce(1, s->d_rho, s->d_rho_old, nullptr, r))) return rc;
            s->diff = r[0] < 1.0e-15 ? 1.0e-5 : r[0];
            return MTS_OK;
        }
        case MTS_M_EVOLUTION: {  // coreLB.hpp:1310-1331 at one level
            if (lvl != 1) return MTS_OK;
            if ((rc = mts_method(s, MTS_M_COPY_F_FB, 1)) || (rc = mts_method(s, MTS_M_COLLIDE, 1)) ||
                (rc = mts_method(s, MTS_M_COPY_F_FTEMP, 1)) || (rc = mts_method(s, MTS_M_ZERODOWN, 1)) ||
                (rc = mts_method(s, MTS_M_STREAM, 1)) || (rc = mts_method(s, MTS_M_APPLY_BC, 1)))
                return rc;
            return MTS_OK;
        }
        default: break;
    }
    PhaseArrays a{s->d_f, s->d_ftemp, s->d_feq, s->d_fneq, s->d_fb, s->d_rho, s->d_rho_old, s->d_ux, s->d_uy,
                  s->d_usq, s->d_usq_old, s->d_psi, s->d_Fx, s->d_Fy};
    rc = launch_phase(s->ctx(), method, lvl, a, s->d_neg);
    if (rc) return fail(rc, "unknown method id");
    return s->launched();
}

static double* field_ptr(mts_sim s, int field, int* q) {
    *q = 1;
    switch (field) {
        case MTS_F_F: *q = 9; return s->d_f;
        case MTS_F_FTEMP: *q = 9; return s->d_ftemp;
        case MTS_F_FEQ: *q = 9; return s->d_feq;
        case MTS_F_FNEQ: *q = 9; return s->d_fneq;
        case MTS_F_FB: *q = 9; return s->d_fb;
        case MTS_F_RHO: return s->d_rho;
        case MTS_F_RHO_OLD: return s->d_rho_old;
        case MTS_F_UX: return s->d_ux;
        case MTS_F_UY: return s->d_uy;
        case MTS_F_USQ: return s->d_usq;
        case MTS_F_USQ_OLD: return s->d_usq_old;
        case MTS_F_PSI: return s->d_psi;
        case MTS_F_FX: return s->d_Fx;
        case MTS_F_FY: return s->d_Fy;
        default: return nullptr;
    }
}

int mts_field_size(mts_sim s, int field, int64_t* n) {
    if (!s || !n) return fail(MTS_ERR_ARG, "null argument");
    if (field < 0 || field > MTS_F_KN) return fail(MTS_ERR_ARG, "unknown field");
    const bool pop = field <= MTS_F_FB;
    *n = s->geo.ncell * (pop ? 9 : 1);
    return MTS_OK;
}

int mts_get_field(mts_sim s, int field, double* out) {
    if (!s || !out) return fail(MTS_ERR_ARG, "null argument");
    if (field == MTS_F_TAUS || field == MTS_F_KN) {
        const auto& v = s->multi ? (field == MTS_F_TAUS ? s->mg.taus : s->mg.kn) : (field == MTS_F_TAUS ? s->geo.taus : s->geo.kn);
        std::copy(v.begin(), v.end(), out);
        return MTS_OK;
    }
    if (s->multi) return multi_get(s, field, out);
    int rc;
    if (field == MTS_F_FTEMP || field == MTS_F_FEQ || field == MTS_F_FNEQ || field == MTS_F_FB)
        if ((rc = s->ensure_phase())) return rc;
    if (field == MTS_F_PSI || field == MTS_F_FX || field == MTS_F_FY)
        if ((rc = s->ensure_sc())) return rc;
    int q;
    double* d = field_ptr(s, field, &q);
    if (!d) return fail(MTS_ERR_ARG, "unknown field");
    const int64_t n = s->geo.ncell * q;
    double* tmp = nullptr;
    HIP_TRY(hipMallocAsync((void**)&tmp, n * 8, s->stream));
    LaunchCtx c = s->ctx();
    const int64_t stride = q == 9 ? s->geo.stride : 0;
    LaunchCtx c2 = c;
    c2.stride = stride;
    launch_gather(c2, d, q, tmp);
    if ((rc = s->launched())) return rc;
    HIP_TRY(hipMemcpyAsync(out, tmp, n * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipFreeAsync(tmp, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MTS_OK;
}

int mts_set_field(mts_sim s, int field, const double* in) {
    if (!s || !in) return fail(MTS_ERR_ARG, "null argument");
    if (s->multi) return fail(MTS_ERR_UNSUPPORTED, "fields of multi-block lattices are read-only");
    if (field == MTS_F_TAUS || field == MTS_F_KN) return fail(MTS_ERR_ARG, "read-only field");
    int rc;
    if (field == MTS_F_FTEMP || field == MTS_F_FEQ || field == MTS_F_FNEQ || field == MTS_F_FB)
        if ((rc = s->ensure_phase())) return rc;
    if (field == MTS_F_PSI || field == MTS_F_FX || field == MTS_F_FY)
        if ((rc = s->ensure_sc())) return rc;
    int q;
    double* d = field_ptr(s, field, &q);
    if (!d) return fail(MTS_ERR_ARG, "unknown field");
    const int64_t n = s->geo.ncell * q;
    double* tmp = nullptr;
    HIP_TRY(hipMallocAsync((void**)&tmp, n * 8, s->stream));
    HIP_TRY(hipMemcpyAsync(tmp, in, n * 8, hipMemcpyHostToDevice, s->stream));
    LaunchCtx c2 = s->ctx();
    c2.stride = q == 9 ? s->geo.stride : 0;
    launch_scatter(c2, d, q, tmp);
    if ((rc = s->launched())) return rc;
    HIP_TRY(hipFreeAsync(tmp, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MTS_OK;
}

int mts_complete_domain(mts_sim s, double* ux, double* uy, double* rho) {
    if (!s) return fail(MTS_ERR_ARG, "null handle");
    if (s->multi) {
        // each pixel takes the value of the non-buffer cell covering it (the
        // assignment of coreLB.hpp:1413-1552); wall pixels 0
        const int nx = s->mg.nx, ny = s->mg.ny;
        const int64_t n = s->mg.ncell;
        std::vector<double> a(n), b(n), c(n);
        int rc;
        if ((rc = multi_get(s, MTS_F_UX, a.data())) || (rc = multi_get(s, MTS_F_UY, b.data())) ||
            (rc = multi_get(s, MTS_F_RHO, c.data())))
            return rc;
        double* outs[3] = {ux, uy, rho};
        const std::vector<double>* vals[3] = {&a, &b, &c};
        for (int k = 0; k < 3; k++)
            if (outs[k]) std::fill(outs[k], outs[k] + (size_t)nx * ny, 0.0);
        for (int64_t cc = 0; cc < n; cc++) {
            if (s->mg.isBuffer[cc]) continue;
            const int w = 1 << (s->mg.level[cc] - 1);
            for (int dy = 0; dy < w; dy++)
                for (int dx = 0; dx < w; dx++) {
                    const size_t p = (size_t)(s->mg.cy[cc] + dy) * nx + s->mg.cx[cc] + dx;
                    for (int k = 0; k < 3; k++)
                        if (outs[k]) outs[k][p] = (*vals[k])[cc];
                }
        }
        return MTS_OK;
    }
    double* outs[3] = {ux, uy, rho};
    double* src[3] = {s->d_ux, s->d_uy, s->d_rho};
    for (int k = 0; k < 3; k++) {
        if (!outs[k]) continue;
        // dense site image is row-major and holds 0 at solid sites
        HIP_TRY(hipMemcpyAsync(outs[k], src[k], s->P * 8, hipMemcpyDeviceToHost, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MTS_OK;
}

int mts_permeability(mts_sim s, double* out8) {
    if (!s || !out8) return fail(MTS_ERR_ARG, "null argument");
    const int nx = s->multi ? s->mg.nx : s->geo.nx, ny = s->multi ? s->mg.ny : s->geo.ny;
    const size_t P = (size_t)nx * ny;
    std::vector<double> ux(P), rho(P);
    int rc;
    if ((rc = mts_complete_domain(s, ux.data(), nullptr, rho.data()))) return rc;
    std::vector<uint8_t> fluid(P, 0);
    if (s->multi)
        for (size_t p = 0; p < P; p++) fluid[p] = s->mg.pore[p] == 1;
    else
        for (int32_t p : s->geo.cellPos) fluid[p] = 1;
    double su = 0, sj = 0, sr = 0;
    int64_t nf = 0;
    for (size_t p = 0; p < P; p++) {
        if (!fluid[p]) continue;
        su += ux[p];
        sj += rho[p] * ux[p];
        sr += rho[p];
        nf++;
    }
    // mean relaxation time of the finest level's cells (the report's units)
    const std::vector<double>& taus = s->multi ? s->mg.taus : s->geo.taus;
    double st = 0;
    int64_t nt = 0;
    for (size_t c = 0; c < taus.size(); c++) {
        if (s->multi && (s->mg.level[c] != 1 || s->mg.isBuffer[c])) continue;
        st += taus[c];
        nt++;
    }
    const double UD = su / (double)P, tau = nt ? st / (double)nt : 0.5, nu = (tau - 0.5) / 3.0;
    const double F = s->Fbody;
    const double k = F != 0.0 ? nu * UD / F : 0.0;
    const double out[8] = {UD, nf ? su / (double)nf : 0.0, sj, sr, (double)nf / (double)P, tau, k,
                           k * s->ph.Cl * s->ph.Cl};
    std::copy(out, out + 8, out8);
    return std::isfinite(UD) ? MTS_OK : fail(MTS_ERR_NONFINITE, "velocity field is not finite");
}

int mts_diagnostics(mts_sim s, double* out4) {
    if (!s || !out4) return fail(MTS_ERR_ARG, "null argument");
    if (s->multi) {  // sums over the cells of every level, weighted by the pixels each covers
        double acc[4] = {0, 0, 0, 0};
        for (size_t l = 0; l < s->lv.size(); l++) {
            double r[4];
            launch_reduce(s->lctx((int)l), 2, s->lv[l].d_rho, s->lv[l].d_ux, s->lv[l].d_uy, s->d_part, s->nparts, s->d_red);
            int rc = s->launched();
            if (rc) return rc;
            HIP_TRY(hipMemcpyAsync(r, s->d_red, 32, hipMemcpyDeviceToHost, s->stream));
            HIP_TRY(hipStreamSynchronize(s->stream));
            const double wgt = (double)(1 << (2 * l));
            for (int k = 0; k < 4; k++) acc[k] += wgt * r[k];
        }
        std::copy(acc, acc + 4, out4);
        return MTS_OK;
    }
    return s->reduce(2, s->d_rho, s->d_ux, s->d_uy, out4);
}

int mts_get_stream(mts_sim s, void** stream) {
    if (!s || !stream) return fail(MTS_ERR_ARG, "null argument");
    *stream = (void*)s->stream;
    return MTS_OK;
}

int mts_synchronize(mts_sim s) {
    if (!s) return fail(MTS_ERR_ARG, "null handle");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MTS_OK;
}

int mts_dev_alloc(size_t bytes, void** ptr) {
    if (!ptr || bytes == 0) return fail(MTS_ERR_ARG, "bad argument");
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // a failed allocation must not surface at a later launch check
        *ptr = nullptr;
        return fail(MTS_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return MTS_OK;
}

int mts_dev_free(void* ptr) {
    if (ptr) HIP_TRY(hipFree(ptr));
    return MTS_OK;
}

int mts_dev_upload(void* dst, const void* src, size_t bytes) {
    if (!dst || !src) return fail(MTS_ERR_ARG, "null argument");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MTS_OK;
}

int mts_dev_download(void* dst, const void* src, size_t bytes) {
    if (!dst || !src) return fail(MTS_ERR_ARG, "null argument");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MTS_OK;
}

int mts_probe_copy(float* dst, const float* src, int64_t n, int launches) {
    if (!dst || !src || n <= 0 || n > ((int64_t)1 << 38) || launches <= 0) return fail(MTS_ERR_ARG, "bad argument");
    for (int i = 0; i < launches; i++) launch_probe_copy(dst, src, n, nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(MTS_ERR_HIP, std::string("probe copy: ") + hipGetErrorString(e));
    return MTS_OK;
}

int mts_probe_copy16(float* dst, const float* src, int64_t n, int launches) {
    if (!dst || !src || n <= 0 || n % 4 || n > ((int64_t)1 << 38) || launches <= 0 || ((uintptr_t)dst & 15) ||
        ((uintptr_t)src & 15))
        return fail(MTS_ERR_ARG, "bad argument (n a multiple of 4, 16-byte aligned pointers)");
    for (int i = 0; i < launches; i++) launch_probe_copy16(dst, src, n, nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(MTS_ERR_HIP, std::string("probe copy: ") + hipGetErrorString(e));
    return MTS_OK;
}

int mts_kernel_timer(mts_sim s, int enable) {
    if (!s) return fail(MTS_ERR_ARG, "null handle");
    s->fold_timer();
    s->ktimer = enable != 0;
    return MTS_OK;
}

int mts_kernel_timer_read(mts_sim s, double* ms, int64_t* launches) {
    if (!s || !ms || !launches) return fail(MTS_ERR_ARG, "null argument");
    s->fold_timer();
    *ms = s->ktime_ms;
    *launches = s->klaunches;
    s->ktime_ms = 0.0;
    s->klaunches = 0;
    return MTS_OK;
}

}  // extern "C"
