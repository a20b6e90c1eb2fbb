// windows.cpp -- the batched form over model windows of uneven size
// (aiso_batch_*).  Many shots, each on a window of a model with its own
// width, velocity, sources and receivers, advance together: one launch per
// fine step covers every tile of every window of a launch group (k_step_win),
// the time loop of each group is captured once in a HIP graph and replayed
// by later calls, and the tile shape, windows per launch and workgroups per
// CU are picked by timing candidates on the windows themselves
// (aiso_batch_tune).
//
// Each window's result is the one the nonlinear forward operator gives for
// that shot alone on that window (F/src/nonlinearPropShotsGpu.cpp:62-152,
// F/src/nonlinearShotsGpuFunctions.cu:227-309): the step kernel applies the
// same operations in the same order to every point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "engine_internal.hpp"

namespace aiso {
namespace {

// Tile lists of one geometry of one window for one tile shape.
struct WinGeom {
    DBuf tilePtr, ent;
    int nReg = 0;
};

struct BatchConfig {
    int tz = 32, tx = 16;
    int spl = 0;        // windows per launch (0: all)
    int persist = 0;    // workgroups per CU (0: one workgroup per tile)
    int lanes = 2;      // streams the launch groups alternate over
    // the device tables depend on the tile, the grouping and the grid only
    bool operator==(const BatchConfig& o) const
    {
        return tz == o.tz && tx == o.tx && spl == o.spl && persist == o.persist;
    }
};

}  // namespace
}  // namespace aiso

struct aiso_batch {
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    aiso_fd_params par{};
    struct Win {
        aiso::FdParams p;
        std::vector<float> vel;
        aiso::Geometry src, rec;
        aiso::DBuf v2dt2, cornerReg, weight;
        std::map<int, std::unique_ptr<aiso::WinGeom>> srcL, recL;   // by tz * 1000 + tx
        long interior() const { return p.interior(); }
    };
    std::vector<std::unique_ptr<Win>> win;
    aiso::DBuf cd;
    // one region per window, in window order
    aiso::DBuf pairs, dreg, sig, slots, desc, tileMap, outStage;
    std::vector<size_t> pairOff, dregOff, sigOff, outOff;
    int nSrcDev = 0;
    // the configuration the device tables hold, and the graphs captured with it
    aiso::BatchConfig cfg, prepared;
    bool isPrepared = false;
    std::vector<int> groupTiles;      // tiles of each launch group
    std::vector<long> groupPts;       // interior points of each launch group
    int useGraph = 1;
    std::vector<hipGraphExec_t> graphs;
    double lastCaptureMs = 0;
    hipEvent_t evIn = nullptr, e0 = nullptr, e1 = nullptr;
    // Lanes: launch group g runs on lane g % lanes; the lanes' kernels run
    // concurrently, so one group's launch tail is filled by another's work.
    std::vector<hipStream_t> lane;
    hipEvent_t forkEv = nullptr;
    std::vector<hipEvent_t> joinEv;
    aiso::PinnedBuf stage;
    float lastMs = 0;
};

namespace aiso {
namespace {

long window_tiles(const FdParams& p, int tz, int tx)
{
    return (long)((p.nz - 2 * kFat + tz - 1) / tz) * ((p.nx - 2 * kFat + tx - 1) / tx);
}

WinGeom* win_geom(aiso_batch* b, aiso_batch::Win& w, const Geometry& g, bool isSrc, int tz, int tx)
{
    auto& m = isSrc ? w.srcL : w.recL;
    const int key = tz * 1000 + tx;
    auto it = m.find(key);
    if (it != m.end()) return it->second.get();
    TileLists L = tile_lists(g, w.p.nz, w.p.nx, tz, tx);
    if (!L.ring.empty())
        throw Error(AISO_ERR_GEOM, "the batched form takes no source or receiver in the FAT ring of a window");
    auto wg = std::make_unique<WinGeom>();
    upload(wg->tilePtr, L.ptr, b->stream);
    upload(wg->ent, L.ent, b->stream);
    wg->nReg = (int)g.regPos.size();
    HIPC(hipStreamSynchronize(b->stream));
    WinGeom* out = wg.get();
    m.emplace(key, std::move(wg));
    return out;
}

void drop_graphs(aiso_batch* b)
{
    for (hipGraphExec_t g : b->graphs)
        if (g) (void)hipGraphExecDestroy(g);
    b->graphs.clear();
}

int n_groups(const aiso_batch* b, const BatchConfig& c)
{
    const int n = (int)b->win.size(), spl = c.spl > 0 ? std::min(c.spl, n) : n;
    return (n + spl - 1) / spl;
}

// Device tables of configuration c: slots, window descriptors, tile map.
void prepare(aiso_batch* b, const BatchConfig& c)
{
    if (b->isPrepared && b->prepared == c) return;
    if (c.tz != 32 && c.tz != 64) throw Error(AISO_ERR_PARAM, "tile tz must be 32 or 64");
    if (c.tx != 16 && c.tx != 32) throw Error(AISO_ERR_PARAM, "tile tx must be 16 or 32");
    if (c.spl < 0 || c.persist < 0) throw Error(AISO_ERR_PARAM, "windows per launch and workgroups per CU are >= 0");
    HIPC(hipStreamSynchronize(b->stream));   // tables are rewritten: no launch may still read them
    drop_graphs(b);
    const int n = (int)b->win.size(), spl = c.spl > 0 ? std::min(c.spl, n) : n;
    std::vector<ShotSlot> slots(n);
    std::vector<WinDesc> desc(n);
    std::vector<int2> map;
    b->groupTiles.clear();
    b->groupPts.clear();
    for (int g0 = 0; g0 < n; g0 += spl) {
        int tiles = 0;
        long pts = 0;
        for (int w = g0; w < std::min(n, g0 + spl); ++w) {
            aiso_batch::Win& W = *b->win[w];
            WinGeom* sg = win_geom(b, W, W.src, true, c.tz, c.tx);
            WinGeom* rg = win_geom(b, W, W.rec, false, c.tz, c.tx);
            ShotSlot& S = slots[w];
            std::memset(&S, 0, sizeof(S));
            S.p[0] = b->pairs.as<float>() + b->pairOff[w];
            S.p[1] = S.p[0] + W.p.nzx();
            S.dreg = b->dreg.as<float>() + b->dregOff[w];
            S.sig = b->sig.as<float>() + b->sigOff[w];
            S.srcTile = sg->tilePtr.as<int>();
            S.srcEnt = sg->ent.as<int2>();
            S.recTile = rg->tilePtr.as<int>();
            S.recEnt = rg->ent.as<int2>();
            WinDesc& D = desc[w];
            D.v2dt2 = W.v2dt2.as<float>();
            D.nz = W.p.nz;
            D.nx = W.p.nx;
            D.tilesZ = (W.p.nz - 2 * kFat + c.tz - 1) / c.tz;
            D.ntiles = (int)window_tiles(W.p, c.tz, c.tx);
            for (int t = 0; t < D.ntiles; ++t) map.push_back(make_int2(w - g0, t));
            tiles += D.ntiles;
            pts += W.interior();
        }
        b->groupTiles.push_back(tiles);
        b->groupPts.push_back(pts);
    }
    upload(b->slots, slots, b->stream);
    upload(b->desc, desc, b->stream);
    upload(b->tileMap, map, b->stream);
    HIPC(hipStreamSynchronize(b->stream));   // the host vectors go out of scope
    b->prepared = c;
    b->isPrepared = true;
    b->graphs.assign(b->groupTiles.size(), nullptr);
}

WinStepArgs group_args(const aiso_batch* b, int g)
{
    const BatchConfig& c = b->prepared;
    const int n = (int)b->win.size(), spl = c.spl > 0 ? std::min(c.spl, n) : n;
    const FdParams& p = b->win[0]->p;
    long tileOff = 0;
    for (int k = 0; k < g; ++k) tileOff += b->groupTiles[k];
    WinStepArgs w{};
    StepArgs& a = w.a;
    a.slots = b->slots.as<ShotSlot>() + (size_t)g * spl;
    a.cd = b->cd.as<float>();
    a.minPad = p.minPad;
    a.nts = p.nts;
    a.ntw = p.ntw;
    for (int k = 0; k < 6; ++k) { a.zc[k] = p.zc[k]; a.xc[k] = p.xc[k]; }
    a.nshots = 1;
    w.win = b->desc.as<WinDesc>() + (size_t)g * spl;
    w.tileMap = b->tileMap.as<int2>() + tileOff;
    w.ntotal = b->groupTiles[g];
    return w;
}

int group_grid(const aiso_batch* b, int g)
{
    const int tiles = b->groupTiles[g];
    long grid = (tiles + 7) / 8 * 8;
    if (b->prepared.persist > 0) grid = std::min<long>(grid, (long)b->cus * b->prepared.persist);
    return (int)std::max<long>(8, (grid + 7) / 8 * 8);
}

std::pair<int, int> group_range(const aiso_batch* b, int g)
{
    const int n = (int)b->win.size(), spl = b->prepared.spl > 0 ? std::min(b->prepared.spl, n) : n;
    return {g * spl, std::min(n, (g + 1) * spl)};
}

void ensure_lanes(aiso_batch* b, int n)
{
    while ((int)b->lane.size() < n) {
        hipStream_t st = nullptr;
        HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));   // default priority
        b->lane.push_back(st);
        hipEvent_t e = nullptr;
        HIPC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->joinEv.push_back(e);
    }
    if (!b->forkEv) HIPC(hipEventCreateWithFlags(&b->forkEv, hipEventDisableTiming));
}

// Lanes used for G launch groups, and the stream of group g.
int lanes_for(const aiso_batch* b, const BatchConfig& c, int G) { return std::max(1, std::min(c.lanes, G)); }
hipStream_t group_stream(const aiso_batch* b, int g, int L) { return L > 1 ? b->lane[g % L] : b->stream; }

// The lanes start after the work queued on the batch stream...
void fork_lanes(aiso_batch* b, int L)
{
    if (L <= 1) return;
    ensure_lanes(b, L);
    HIPC(hipEventRecord(b->forkEv, b->stream));
    for (int l = 0; l < L; ++l) HIPC(hipStreamWaitEvent(b->lane[l], b->forkEv, 0));
}

// ...and the batch stream continues after all of them.
void join_lanes(aiso_batch* b, int L)
{
    if (L <= 1) return;
    for (int l = 0; l < L; ++l) {
        HIPC(hipEventRecord(b->joinEv[l], b->lane[l]));
        HIPC(hipStreamWaitEvent(b->stream, b->joinEv[l], 0));
    }
}

// Fine steps [t0, t1) of group g on stream st, t counted from 0 (its = t / sub).
void launch_steps(aiso_batch* b, int g, int t0, int t1, hipStream_t st, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr)
{
    const FdParams& p = b->win[0]->p;
    const int h = p.h();
    const float* f = p.filt.data();
    WinStepArgs w = group_args(b, g);
    const int grid = group_grid(b, g);
    for (int t = t0; t < t1; ++t) {
        const int its = t / p.sub, it2 = t % p.sub + 1;
        w.a.cur = (t & 1) ? 0 : 1;
        w.a.its = its;
        w.a.itwInj = its * p.sub + it2 - 1;
        w.a.fe = f[it2];
        w.a.ff = f[h + it2];
        hipEvent_t a0 = t == t0 ? e0 : nullptr, a1 = t == t1 - 1 ? e1 : nullptr;
        if (a0 && !a1) HIPC(hipEventRecord(a0, st));
        HIPC(launch_step_windows(MODE_NL_FWD, b->prepared.tz, b->prepared.tx, grid, w, st,
                                 a0 && a1 ? a0 : nullptr, a0 && a1 ? a1 : nullptr));
        if (a1 && !a0) HIPC(hipEventRecord(a1, st));
    }
}

// Zero the wavefield pairs and receiver traces of group g (on stream st).
void clear_group(aiso_batch* b, int g, hipStream_t st)
{
    auto r = group_range(b, g);
    const int last = r.second - 1;
    const size_t p0 = b->pairOff[r.first], p1 = b->pairOff[last] + 2 * b->win[last]->p.nzx();
    const size_t d0 = b->dregOff[r.first], d1 = b->dregOff[last] + (size_t)b->win[last]->rec.regPos.size() *
                                                                      b->win[last]->p.nts;
    HIPC(launch_fill(b->pairs.as<float>() + p0, (long)(p1 - p0), st));
    HIPC(launch_fill(b->dreg.as<float>() + d0, (long)(d1 - d0), st));
}

// The whole time loop of group g on stream st: replayed from its graph,
// captured on the first call (a graph runs on whichever stream launches it).
void run_group(aiso_batch* b, int g, hipStream_t st)
{
    const FdParams& p = b->win[0]->p;
    const int T = p.ntw - 1;
    if (!b->useGraph) {
        clear_group(b, g, st);
        launch_steps(b, g, 0, T, st);
        return;
    }
    if (!b->graphs[g]) {
        const auto h0 = std::chrono::steady_clock::now();
        hipGraph_t graph = nullptr;
        HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            clear_group(b, g, st);
            launch_steps(b, g, 0, T, st);
        } catch (...) {
            hipGraph_t dummy = nullptr;
            (void)hipStreamEndCapture(st, &dummy);
            if (dummy) (void)hipGraphDestroy(dummy);
            throw;
        }
        HIPC(hipStreamEndCapture(st, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPC(e);
        b->graphs[g] = exec;
        b->lastCaptureMs +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    }
    HIPC(hipGraphLaunch(b->graphs[g], st));
}

// Source chain of a nonlinear shot on window W (F/src/nonlinearPropGpu.cpp:132-139).
std::vector<float> window_source(const aiso_batch::Win& W, const float* m)
{
    const FdParams& p = W.p;
    const int nReg = (int)W.src.regPos.size();
    std::vector<float> reg((size_t)nReg * p.nts), fine((size_t)nReg * p.ntw);
    W.src.adjoint(false, reg.data(), m, p.nts);
    scale_source(p, W.vel.data(), W.src.regPos, p.nts, reg.data());
    interp_time_fwd(p, nReg, reg.data(), fine.data());
    return fine;
}

void batch_nl_forward(aiso_batch* b, int add, const float* const* model, int nModel, float* const* data, int mem)
{
    const int n = (int)b->win.size();
    if (n == 0) throw Error(AISO_ERR_STATE, "no windows set");
    if (nModel != 1 && nModel != n) throw Error(AISO_ERR_PARAM, "model count must be 1 or the number of windows");
    for (int k = 0; k < nModel; ++k)
        if (!model[k]) throw Error(AISO_ERR_PARAM, "null model");
    for (int w = 0; w < n; ++w)
        if (!data[w]) throw Error(AISO_ERR_PARAM, "null data");
    hipStream_t s = b->stream;
    if (mem == AISO_MEM_DEVICE) {
        if (!b->evIn) HIPC(hipEventCreateWithFlags(&b->evIn, hipEventDisableTiming));
        HIPC(hipEventRecord(b->evIn, nullptr));
        HIPC(hipStreamWaitEvent(s, b->evIn, 0));
    }
    prepare(b, b->cfg);
    const FdParams& p0 = b->win[0]->p;
    const size_t modelN = (size_t)b->nSrcDev * p0.nts;
    // source traces of every window, staged through pinned memory
    std::vector<float> hm(mem == AISO_MEM_DEVICE ? modelN * nModel : 0);
    for (int k = 0; mem == AISO_MEM_DEVICE && k < nModel; ++k)
        HIPC(hipMemcpy(hm.data() + k * modelN, model[k], modelN * 4, hipMemcpyDeviceToHost));
    const size_t sigN = b->sigOff.back();
    float* st = (float*)b->stage.need(sigN * 4);
    HIPC(hipStreamSynchronize(s));             // the previous call's copies from the stage have run
    for (int w = 0; w < n; ++w) {
        const int k = nModel == 1 ? 0 : w;
        const float* m = mem == AISO_MEM_DEVICE ? hm.data() + k * modelN : model[k];
        std::vector<float> fine = window_source(*b->win[w], m);
        std::memcpy(st + b->sigOff[w], fine.data(), fine.size() * 4);
    }
    HIPC(hipMemcpyAsync(b->sig.p, st, sigN * 4, hipMemcpyHostToDevice, s));
    if (!b->e0) {
        HIPC(hipEventCreate(&b->e0));
        HIPC(hipEventCreate(&b->e1));
    }
    HIPC(hipEventRecord(b->e0, s));
    const int G = (int)b->groupTiles.size();
    const int L = lanes_for(b, b->cfg, G);
    fork_lanes(b, L);
    for (int g = 0; g < G; ++g) {
        hipStream_t gs = group_stream(b, g, L);
        run_group(b, g, gs);
        auto r = group_range(b, g);
        for (int w = r.first; w < r.second; ++w) {
            aiso_batch::Win& W = *b->win[w];
            float* dst = mem == AISO_MEM_DEVICE ? data[w] : b->outStage.as<float>() + b->outOff[w];
            const int addHere = mem == AISO_MEM_DEVICE ? add : 0;
            HIPC(launch_dev_forward(b->dreg.as<float>() + b->dregOff[w], dst, W.cornerReg.as<int>(),
                                    W.weight.as<float>(), W.rec.nDev, W.rec.nb, W.p.nts, addHere, gs));
        }
    }
    join_lanes(b, L);
    HIPC(hipEventRecord(b->e1, s));
    if (mem != AISO_MEM_DEVICE) {
        HIPC(hipStreamSynchronize(s));
        for (int w = 0; w < n; ++w) {
            const aiso_batch::Win& W = *b->win[w];
            const size_t cnt = (size_t)W.rec.nDev * W.p.nts;
            std::vector<float> tmp(cnt);
            HIPC(hipMemcpy(tmp.data(), b->outStage.as<float>() + b->outOff[w], cnt * 4, hipMemcpyDeviceToHost));
            if (add)
                for (size_t i = 0; i < cnt; ++i) data[w][i] += tmp[i];
            else
                std::memcpy(data[w], tmp.data(), cnt * 4);
        }
    }
    HIPC(hipStreamSynchronize(s));
    HIPC(hipEventElapsedTime(&b->lastMs, b->e0, b->e1));
}

// Time `steps` fine steps of every launch group under configuration c, the
// groups on their lanes as a call runs them; returns ms per interior
// point-update.
double time_config(aiso_batch* b, const BatchConfig& c, int steps)
{
    prepare(b, c);
    const int G = (int)b->groupTiles.size();
    const int L = lanes_for(b, c, G);
    hipStream_t s = b->stream;
    for (int g = 0; g < G; ++g) clear_group(b, g, s);
    HIPC(hipEventRecord(b->e0, s));      // after the clears, before the lanes fork
    fork_lanes(b, L);
    double pts = 0;
    for (int g = 0; g < G; ++g) {
        launch_steps(b, g, 0, steps, group_stream(b, g, L));
        pts += (double)b->groupPts[g] * steps;
    }
    join_lanes(b, L);
    HIPC(hipEventRecord(b->e1, s));
    HIPC(hipStreamSynchronize(s));
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, b->e0, b->e1));
    return ms / pts;
}

}  // namespace
}  // namespace aiso

using namespace aiso;

extern "C" {

aiso_status aiso_batch_create(const aiso_fd_params* par, int device, aiso_batch** out)
{
    return guard_call([&] {
        if (!par || !out) throw Error(AISO_ERR_PARAM, "null argument");
        auto b = std::make_unique<aiso_batch>();
        b->device = device;
        b->par = *par;
        HIPC(hipSetDevice(device));
        hipDeviceProp_t prop{};
        HIPC(hipGetDeviceProperties(&prop, device));
        b->cus = std::max(1, prop.multiProcessorCount);
        HIPC(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));   // default priority
        *out = b.release();
    });
}

aiso_status aiso_batch_destroy(aiso_batch* b)
{
    if (!b) return AISO_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (hipEvent_t e : {b->evIn, b->e0, b->e1, b->forkEv})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->joinEv) (void)hipEventDestroy(e);
    for (hipStream_t st : b->lane) (void)hipStreamDestroy(st);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return AISO_OK;
}

aiso_status aiso_batch_set_windows(aiso_batch* b, int nWin, const int* nz, const int* nx, const float* const* vel,
                                   aiso_device* const* sources, aiso_device* const* receivers, int mem)
{
    return guard_call([&] {
        if (!b || !nz || !nx || !vel || !sources || !receivers) throw Error(AISO_ERR_PARAM, "null argument");
        if (nWin < 1) throw Error(AISO_ERR_PARAM, "nWin must be >= 1");
        HIPC(hipSetDevice(b->device));
        HIPC(hipStreamSynchronize(b->stream));
        drop_graphs(b);
        b->isPrepared = false;
        std::vector<std::unique_ptr<aiso_batch::Win>> win;
        int nSrcDev = -1;
        for (int w = 0; w < nWin; ++w) {
            if (!vel[w] || !sources[w] || !receivers[w]) throw Error(AISO_ERR_PARAM, "null window argument");
            aiso_fd_params q = b->par;
            q.nz = nz[w];
            q.nx = nx[w];
            auto W = std::make_unique<aiso_batch::Win>();
            W->p = make_params(q);
            if (W->p.freeSurface) throw Error(AISO_ERR_PARAM, "the batched form takes no free surface");
            if (W->p.sub >= kSubMax) throw Error(AISO_ERR_PARAM, "**** ERROR: Subsampling parameter too high ****");
            if (W->p.minPad >= kPadMax) throw Error(AISO_ERR_PARAM, "**** ERROR: Padding value is too high ****");
            W->vel.resize(W->p.nzx());
            if (mem == AISO_MEM_DEVICE)
                HIPC(hipMemcpy(W->vel.data(), vel[w], W->vel.size() * 4, hipMemcpyDeviceToHost));
            else
                std::memcpy(W->vel.data(), vel[w], W->vel.size() * 4);
            check_velocity(W->p, W->vel.data());
            W->src = sources[w]->g;
            W->rec = receivers[w]->g;
            for (int v : W->src.gridIdx)
                if (v >= W->p.nzx()) throw Error(AISO_ERR_GEOM, "source outside its window");
            for (int v : W->rec.gridIdx)
                if (v >= W->p.nzx()) throw Error(AISO_ERR_GEOM, "receiver outside its window");
            if (W->src.nt != W->p.nts || W->rec.nt != W->p.nts)
                throw Error(AISO_ERR_GEOM, "device trace length differs from nts");
            if (nSrcDev < 0) nSrcDev = W->src.nDev;
            if (W->src.nDev != nSrcDev) throw Error(AISO_ERR_GEOM, "sources per window must be constant");
            std::vector<float> v2(W->p.nzx()), refl(W->p.nzx());
            velocity_scales(W->p, W->vel.data(), v2.data(), refl.data());
            upload(W->v2dt2, v2, b->stream);
            upload(W->cornerReg, W->rec.cornerReg, b->stream);
            upload(W->weight, W->rec.weight, b->stream);
            HIPC(hipStreamSynchronize(b->stream));
            win.push_back(std::move(W));
        }
        const FdParams& p0 = win[0]->p;
        upload(b->cd, p0.cd, b->stream);
        b->pairOff.assign(nWin + 1, 0);
        b->dregOff.assign(nWin + 1, 0);
        b->sigOff.assign(nWin + 1, 0);
        b->outOff.assign(nWin + 1, 0);
        for (int w = 0; w < nWin; ++w) {
            const aiso_batch::Win& W = *win[w];
            auto al = [](size_t n) { return (n + 63) / 64 * 64; };   // 256-byte regions
            b->pairOff[w + 1] = b->pairOff[w] + al(2 * W.p.nzx());
            b->dregOff[w + 1] = b->dregOff[w] + al(std::max<size_t>(1, W.rec.regPos.size()) * W.p.nts);
            b->sigOff[w + 1] = b->sigOff[w] + al(std::max<size_t>(1, W.src.regPos.size()) * W.p.ntw);
            b->outOff[w + 1] = b->outOff[w] + al((size_t)W.rec.nDev * W.p.nts);
        }
        size_t freeB = 0, totalB = 0;
        HIPC(hipMemGetInfo(&freeB, &totalB));
        const size_t need = (b->pairOff[nWin] + b->dregOff[nWin] + b->sigOff[nWin] + b->outOff[nWin]) * 4;
        const size_t have = b->pairs.n + b->dreg.n + b->sig.n + b->outStage.n;
        if (need > (freeB + have) * 0.9)
            throw Error(AISO_ERR_MEMORY, "the windows need more device memory than is free");
        b->pairs.need(b->pairOff[nWin] * 4);
        b->dreg.need(b->dregOff[nWin] * 4);
        b->sig.need(b->sigOff[nWin] * 4);
        b->outStage.need(b->outOff[nWin] * 4);
        // source traces start at zero (aiso_batch_tune may run before a call)
        HIPC(launch_fill(b->sig.as<float>(), (long)b->sigOff[nWin], b->stream));
        HIPC(hipStreamSynchronize(b->stream));
        b->win = std::move(win);
        b->nSrcDev = nSrcDev;
    });
}

aiso_status aiso_batch_set_config(aiso_batch* b, int tz, int tx, int windowsPerLaunch, int blocksPerCU, int lanes,
                                  int graph)
{
    return guard_call([&] {
        if (!b) throw Error(AISO_ERR_PARAM, "null argument");
        if (tz != 32 && tz != 64) throw Error(AISO_ERR_PARAM, "tile tz must be 32 or 64");
        if (tx != 16 && tx != 32) throw Error(AISO_ERR_PARAM, "tile tx must be 16 or 32");
        if (windowsPerLaunch < 0 || blocksPerCU < 0) throw Error(AISO_ERR_PARAM, "negative configuration value");
        if (lanes < 1 || lanes > 4) throw Error(AISO_ERR_PARAM, "lanes must be 1..4");
        b->cfg = BatchConfig{tz, tx, windowsPerLaunch, blocksPerCU, lanes};
        if (graph != b->useGraph) {
            HIPC(hipSetDevice(b->device));
            HIPC(hipStreamSynchronize(b->stream));
            drop_graphs(b);
            if (b->isPrepared) b->graphs.assign(b->groupTiles.size(), nullptr);
        }
        b->useGraph = graph ? 1 : 0;
    });
}

aiso_status aiso_batch_get_config(const aiso_batch* b, int* tz, int* tx, int* windowsPerLaunch, int* blocksPerCU,
                                  int* lanes, int* graph)
{
    if (!b) return AISO_ERR_PARAM;
    if (tz) *tz = b->cfg.tz;
    if (tx) *tx = b->cfg.tx;
    if (windowsPerLaunch) *windowsPerLaunch = b->cfg.spl;
    if (blocksPerCU) *blocksPerCU = b->cfg.persist;
    if (lanes) *lanes = b->cfg.lanes;
    if (graph) *graph = b->useGraph;
    return AISO_OK;
}

aiso_status aiso_batch_tune(aiso_batch* b, int steps, double* msPerPointUpdate)
{
    return guard_call([&] {
        if (!b) throw Error(AISO_ERR_PARAM, "null argument");
        if (b->win.empty()) throw Error(AISO_ERR_STATE, "no windows set");
        HIPC(hipSetDevice(b->device));
        if (!b->e0) {
            HIPC(hipEventCreate(&b->e0));
            HIPC(hipEventCreate(&b->e1));
        }
        const int n = (int)b->win.size();
        const int T = b->win[0]->p.ntw - 1;
        steps = std::max(1, std::min(steps > 0 ? steps : 16, T));
        std::vector<int> spls;
        for (int s : {8, 16, 32, 64})
            if (s < n) spls.push_back(s);
        spls.push_back(0);
        // Screen every candidate, then re-time the four fastest three times
        // over 3 x steps and keep the best of those times.
        std::vector<std::pair<double, BatchConfig>> cand;
        for (int tz : {32, 64})
            for (int tx : {16, 32})
                for (int spl : spls)
                    for (int persist : {0, 4, 8})
                        for (int lanes : {1, 2}) {
                            const BatchConfig c{tz, tx, spl, persist, lanes};
                            cand.emplace_back(time_config(b, c, steps), c);
                        }
        std::stable_sort(cand.begin(), cand.end(),
                         [](const auto& x, const auto& y) { return x.first < y.first; });
        cand.resize(std::min<size_t>(cand.size(), 4));
        const int longSteps = std::min(3 * steps, T);
        for (auto& cd : cand) {
            double t = 1e300;
            for (int r = 0; r < 3; ++r) t = std::min(t, time_config(b, cd.second, longSteps));
            cd.first = t;
        }
        const auto best = *std::min_element(cand.begin(), cand.end(),
                                            [](const auto& x, const auto& y) { return x.first < y.first; });
        b->cfg = best.second;
        if (msPerPointUpdate) *msPerPointUpdate = best.first;
    });
}

aiso_status aiso_batch_nl_forward(aiso_batch* b, int add, const float* const* model, int nModel, float* const* data,
                                  int mem)
{
    return guard_call([&] {
        if (!b || !model || !data) throw Error(AISO_ERR_PARAM, "null argument");
        HIPC(hipSetDevice(b->device));
        batch_nl_forward(b, add, model, nModel, data, mem);
    });
}

aiso_status aiso_batch_info(const aiso_batch* b, int* nWin, int* launchGroups, long long* pointUpdates,
                            float* lastMs, double* captureMs)
{
    if (!b) return AISO_ERR_PARAM;
    long long pts = 0;
    for (auto& W : b->win) pts += (long long)W->interior() * (W->p.ntw - 1);
    if (nWin) *nWin = (int)b->win.size();
    if (launchGroups) *launchGroups = b->win.empty() ? 0 : n_groups(b, b->cfg);
    if (pointUpdates) *pointUpdates = pts;
    if (lastMs) *lastMs = b->lastMs;
    if (captureMs) *captureMs = b->lastCaptureMs;
    return AISO_OK;
}

aiso_status aiso_batch_stream(const aiso_batch* b, void** stream)
{
    if (!b || !stream) return AISO_ERR_PARAM;
    *stream = (void*)b->stream;
    return AISO_OK;
}

}  // extern "C"
