// engine.cpp -- the C-ABI of libaiso.so: contexts, geometry, the shared
// helpers of the operator passes, and the adjoint-state velocity gradient.
//
// The reference runs one shot per GPU at a time, re-creating its operator
// objects and device buffers on every call (F/src/BornShotsGpu.cpp:82-143).
// Here a context keeps the velocity tables, geometry lists and work buffers
// resident on the device, and every fine time step advances a whole batch of
// shots with one launch of the fused step kernel (kernels.hip); the passes of
// each operator call are in passes.cpp.  Per-shot results are reduced in
// shot order, so they do not depend on the batch size.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "engine_internal.hpp"

using namespace aiso;

namespace {
thread_local std::string g_err;
std::atomic<long> g_geom_id{1};
}  // namespace

namespace aiso {

long next_geom_id() { return g_geom_id++; }

void set_last_error(const char* msg) { g_err = msg; }

void hip_check(hipError_t e, const char* what)
{
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        throw Error(AISO_ERR_MEMORY, std::string("out of device memory in ") + what);
    if (e != hipSuccess)
        throw Error(AISO_ERR_HIP, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

int tiles_z(const aiso_ctx* c) { return (c->p.nz - 2 * kFat + c->tz - 1) / c->tz; }
int tiles_x(const aiso_ctx* c) { return (c->p.nx - 2 * kFat + c->tx - 1) / c->tx; }

// Bucket the unique points of a geometry by tile of a tz x tx tiling of
// the nz x nx grid's interior: per-tile CSR offsets and (local index, trace)
// entries; points in the FAT ring are listed apart with their grid index.
TileLists tile_lists(const Geometry& g, int nz, int nx, int tz, int tx)
{
    TileLists L;
    const int tZ = (nz - 2 * kFat + tz - 1) / tz, nt = tZ * ((nx - 2 * kFat + tx - 1) / tx);
    std::vector<std::vector<int2>> buckets(nt);
    for (int r = 0; r < (int)g.regPos.size(); ++r) {
        int pos = g.regPos[r], iz = pos % nz, ix = pos / nz;
        if (iz >= kFat && iz < nz - kFat && ix >= kFat && ix < nx - kFat) {
            int t = (iz - kFat) / tz + tZ * ((ix - kFat) / tx);
            int local = ((ix - kFat) % tx) * tz + (iz - kFat) % tz;
            buckets[t].push_back(make_int2(local, r));
        } else {
            L.ring.push_back(make_int2(pos, r));
        }
    }
    L.ptr.assign(nt + 1, 0);
    for (int t = 0; t < nt; ++t) {
        L.ptr[t + 1] = L.ptr[t] + (int)buckets[t].size();
        L.ent.insert(L.ent.end(), buckets[t].begin(), buckets[t].end());
    }
    return L;
}

// The tile lists of a geometry for the context's current tile, resident on
// the device (cached by geometry and tile).
DevGeom* dev_geom(aiso_ctx* c, const aiso_device* d)
{
    auto key = std::make_tuple(d->id, c->tz, c->tx);
    auto it = c->geoms.find(key);
    if (it != c->geoms.end()) return it->second.get();
    auto dg = std::make_unique<DevGeom>();
    const Geometry& g = d->g;
    TileLists L = tile_lists(g, c->p.nz, c->p.nx, c->tz, c->tx);
    upload(dg->tilePtr, L.ptr, c->stream);
    upload(dg->ent, L.ent, c->stream);
    upload(dg->ring, L.ring, c->stream);
    upload(dg->cornerReg, g.cornerReg, c->stream);
    upload(dg->weight, g.weight, c->stream);
    upload(dg->adjPtr, g.adjPtr, c->stream);
    upload(dg->adjEnt, g.adjEnt, c->stream);
    dg->nRing = (int)L.ring.size();
    dg->nReg = (int)g.regPos.size();
    HIPC(hipStreamSynchronize(c->stream));
    DevGeom* out = dg.get();
    c->geoms.emplace(key, std::move(dg));
    return out;
}

// Tiling of the two-step kernel: output tiles of OZ x OX points, each
// computing step A on the EZ x EX region around it.
int tiles_z2(const aiso_ctx* c) { return (c->p.nz - 2 * kFat + kOZ2 - 1) / kOZ2; }
int tiles_x2(const aiso_ctx* c) { return (c->p.nx - 2 * kFat + kOX2 - 1) / kOX2; }

// Sparse lists of a geometry for the two-step kernel.  With `region` every
// tile whose EZ x EX region holds a point lists it (injections feed step A on
// the whole region); otherwise only the owning tile does (records).  Local
// indices are ex * EZ + ez in the region frame.  Points in the FAT ring are
// not listed: the caller checks there are none.
DevGeom* dev_geom_pair2(aiso_ctx* c, const aiso_device* d, bool region)
{
    auto key = std::make_tuple(d->id, region ? -1 : -2, kPair2EZ * 1000 + kPair2EX);
    auto it = c->geoms.find(key);
    if (it != c->geoms.end()) return it->second.get();
    auto dg = std::make_unique<DevGeom>();
    const Geometry& g = d->g;
    const int nz = c->p.nz, nx = c->p.nx, tZ = tiles_z2(c), tX = tiles_x2(c), nt = tZ * tX;
    std::vector<std::vector<int2>> buckets(nt);
    for (int r = 0; r < (int)g.regPos.size(); ++r) {
        const int pos = g.regPos[r], iz = pos % nz, ix = pos / nz;
        if (!(iz >= kFat && iz < nz - kFat && ix >= kFat && ix < nx - kFat)) continue;
        int zlo = (iz - kFat) / kOZ2, zhi = zlo, xlo = (ix - kFat) / kOX2, xhi = xlo;
        if (region) {   // tiles t with t*O <= i < t*O + E (region origin t*O)
            zlo = std::max(0, (iz - kPair2EZ + kOZ2) / kOZ2);
            zhi = std::min(tZ - 1, iz / kOZ2);
            xlo = std::max(0, (ix - kPair2EX + kOX2) / kOX2);
            xhi = std::min(tX - 1, ix / kOX2);
        }
        for (int tx = xlo; tx <= xhi; ++tx)
            for (int tz = zlo; tz <= zhi; ++tz) {
                const int lz = iz - tz * kOZ2, lx = ix - tx * kOX2;
                if (lz < 0 || lz >= kPair2EZ || lx < 0 || lx >= kPair2EX) continue;
                buckets[tz + tZ * tx].push_back(make_int2(lx * kPair2EZ + lz, r));
            }
    }
    std::vector<int> ptr(nt + 1, 0);
    std::vector<int2> ent;
    for (int t = 0; t < nt; ++t) {
        ptr[t + 1] = ptr[t] + (int)buckets[t].size();
        ent.insert(ent.end(), buckets[t].begin(), buckets[t].end());
    }
    upload(dg->tilePtr, ptr, c->stream);
    upload(dg->ent, ent, c->stream);
    dg->nReg = (int)g.regPos.size();
    HIPC(hipStreamSynchronize(c->stream));
    DevGeom* out = dg.get();
    c->geoms.emplace(key, std::move(dg));
    return out;
}

hipEvent_t pooled_event(aiso_ctx* c)
{
    if (!c->evPool.empty()) {
        hipEvent_t e = c->evPool.back();
        c->evPool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    HIPC(hipEventCreate(&e));
    return e;
}

// Device-memory operands may have been written by the caller on the legacy
// default stream (PyTorch's default): the context's non-blocking stream waits
// for that work before it reads them.
void wait_default_stream(aiso_ctx* c)
{
    if (!c->evIn) HIPC(hipEventCreateWithFlags(&c->evIn, hipEventDisableTiming));
    HIPC(hipEventRecord(c->evIn, nullptr));
    HIPC(hipStreamWaitEvent(c->stream, c->evIn, 0));
}

// Shots [k0, k0 + cnt) of a launch over `nshots` that lane l advances.
void lane_range(int nshots, int lanes, int l, int& k0, int& cnt)
{
    const int per = (nshots + lanes - 1) / lanes;
    k0 = l * per;
    cnt = std::max(0, std::min(per, nshots - k0));
}

StepArgs lane_args(const StepArgs& a, int k0, int cnt)
{
    StepArgs y = a;
    y.slots = a.slots + k0;
    y.nshots = cnt;
    y.gridLogical = a.ntiles * cnt;
    return y;
}

hipStream_t lane_stream(const aiso_ctx* c, int l) { return c->curLanes > 1 ? c->lane[l] : c->stream; }

void ensure_lanes(aiso_ctx* c)
{
    while ((int)c->lane.size() < c->lanes) {
        hipStream_t st = nullptr;
        HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));   // default priority
        c->lane.push_back(st);
        hipEvent_t e = nullptr;
        HIPC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->joinEv.push_back(e);
    }
    if (!c->forkEv) HIPC(hipEventCreateWithFlags(&c->forkEv, hipEventDisableTiming));
}

// Lane streams start after the work queued on the context's stream...
void fork_lanes(aiso_ctx* c)
{
    if (c->curLanes <= 1) return;
    HIPC(hipEventRecord(c->forkEv, c->stream));
    for (int l = 0; l < c->curLanes; ++l) HIPC(hipStreamWaitEvent(c->lane[l], c->forkEv, 0));
}

// ...and the context's stream continues after all of them.
void join_lanes(aiso_ctx* c)
{
    if (c->curLanes <= 1) return;
    for (int l = 0; l < c->curLanes; ++l) {
        HIPC(hipEventRecord(c->joinEv[l], c->lane[l]));
        HIPC(hipStreamWaitEvent(c->stream, c->joinEv[l], 0));
    }
}

// Whether this dispatch of `kind` carries profiling events.  A sampled
// dispatch (profEvery > 1) that follows other kernels back to back reads
// several microseconds longer than the kernel runs (MI355X, against
// rocprofv3), so sampled figures other than the pair kernel's windows are
// upper bounds.
bool sample_event(aiso_ctx* c, int kind)
{
    // With lanes the dispatches of different lanes overlap, so per-dispatch
    // durations are not the kernel's; such passes are timed whole instead.
    return c->profile && c->curLanes == 1 && c->profSeq[kind]++ % c->profEvery == 0;
}

hipEvent_t pass_timer_begin(aiso_ctx* c)
{
    if (!c->profile || c->curLanes <= 1) return nullptr;
    hipEvent_t e0 = pooled_event(c);
    HIPC(hipEventRecord(e0, c->stream));
    return e0;
}

void pass_timer_end(aiso_ctx* c, hipEvent_t e0, int kind, double pts, int steps)
{
    if (!e0) return;
    hipEvent_t e1 = pooled_event(c);
    HIPC(hipEventRecord(e1, c->stream));
    c->ev.push_back({e0, e1, kind, pts, steps});
}

void launch_timed(aiso_ctx* c, unsigned mode, int kind, const StepArgs& a, bool generic)
{
    for (int l = 0; l < c->curLanes; ++l) {
        int k0, cnt;
        lane_range(a.nshots, c->curLanes, l, k0, cnt);
        if (cnt == 0) continue;
        const StepArgs y = c->curLanes > 1 ? lane_args(a, k0, cnt) : a;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (sample_event(c, kind)) {
            e0 = pooled_event(c);
            e1 = pooled_event(c);
            c->ev.push_back({e0, e1, kind, double(c->p.interior()) * cnt});
        }
        if (generic) HIPC(launch_step_generic(mode, c->tz, c->tx, y, lane_stream(c, l), e0, e1));
        else HIPC(launch_step(mode, c->tz, c->tx, c->spb, y, lane_stream(c, l), e0, e1, c->spb == 1 ? c->occ : 0));
    }
}

void drain_profile(aiso_ctx* c)
{
    if (c->ev.empty() && c->batchEv.empty()) return;
    HIPC(hipStreamSynchronize(c->stream));
    for (auto& t : c->ev) {
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, t.e0, t.e1));
        c->ms[t.kind] += ms;
        c->pts[t.kind] += t.pts;
        c->launches[t.kind] += t.n;
        c->evPool.push_back(t.e0);
        c->evPool.push_back(t.e1);
    }
    c->ev.clear();
    for (auto& b : c->batchEv) {
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, b.e0, b.e1));
        c->batchDone.emplace_back(ms, b.shots);
        c->evPool.push_back(b.e0);
        c->evPool.push_back(b.e1);
    }
    c->batchEv.clear();
}

// Host signal chain for one Born shot: S = I_t(D_v(D2_t(R_src^T w)))
// (F/include/seismicOperator2D.cpp:151-180).
std::vector<float> born_source(const aiso_ctx* c, const Geometry& src, const float* wav)
{
    const FdParams& p = c->p;
    const int nReg = (int)src.regPos.size();
    std::vector<float> reg((size_t)nReg * p.nts), d2((size_t)nReg * p.nts), fine((size_t)nReg * p.ntw);
    src.adjoint(false, reg.data(), wav, p.nts);
    second_time_derivative(p, nReg, reg.data(), d2.data());
    scale_source(p, c->vel.data(), src.regPos, p.nts, d2.data());
    interp_time_fwd(p, nReg, d2.data(), fine.data());
    return fine;
}

// Nonlinear source chain (F/src/nonlinearPropGpu.cpp:132-139).
std::vector<float> nl_source(const aiso_ctx* c, const Geometry& src, const float* m)
{
    const FdParams& p = c->p;
    const int nReg = (int)src.regPos.size();
    std::vector<float> reg((size_t)nReg * p.nts), fine((size_t)nReg * p.ntw);
    src.adjoint(false, reg.data(), m, p.nts);
    scale_source(p, c->vel.data(), src.regPos, p.nts, reg.data());
    interp_time_fwd(p, nReg, reg.data(), fine.data());
    return fine;
}

// Pinned staging of one batch's host-to-device copies.
void stage_begin(aiso_ctx* c, size_t bytes)
{
    if (!c->stageEv) HIPC(hipEventCreateWithFlags(&c->stageEv, hipEventDisableTiming));
    if (c->stagePending) {   // the previous batch's copies read the buffer until they have run
        HIPC(hipEventSynchronize(c->stageEv));
        c->stagePending = false;
    }
    c->stage.need(bytes);
    c->stageOff = 0;
}

void stage_copy(aiso_ctx* c, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return;
    if (c->stageOff + stage_size(bytes) > c->stage.n) throw Error(AISO_ERR_STATE, "internal: staging buffer overrun");
    char* h = static_cast<char*>(c->stage.p) + c->stageOff;
    std::memcpy(h, src, bytes);
    HIPC(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, c->stream));
    c->stageOff += stage_size(bytes);
}

void stage_end(aiso_ctx* c)
{
    HIPC(hipEventRecord(c->stageEv, c->stream));
    c->stagePending = true;
}

}  // namespace aiso

namespace {

template <class F>
aiso_status guard(F&& f)
{
    try {
        f();
        return AISO_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return AISO_ERR_PARAM;
    }
}

// Velocity gradient of the nonlinear operator by the adjoint-state method
// (no counterpart in the reference, whose FWI takes the Born adjoint as its
// gradient, PY/fwiMain.py:123-142).  The forward recursion of one fine step,
//   p_t = D (v2dt2 * L p_{t-1} + 2 p_{t-1} - D p_{t-2} + s_t),
// has as transpose the nonlinear adjoint propagation; driven by dJ/dd its
// damped slices nu_t give dJ/dv2dt2 = sum_t (L p_{t-1}) * nu_t, plus at the
// source points sum_t nu_t * u_t for the v2dt2 factor of the injected source
// s = v2dt2 * u.  The forward slices are recomputed backward in segments of
// K steps from checkpoints (the pair of slices every K steps).  Per batch of
// shots: one forward pass writing checkpoints, then per segment, last first,
// a forward recompute into W (K slices) and the K adjoint steps, each
// followed by the imaging kernel.
void run_nl_grad(aiso_ctx* c, const float* model, int nShotSig, const float* dataGrad, float* velGrad,
                 aiso_device* const* srcs, int nShot, aiso_device* const* recs, int nRec, int every, int mem)
{
    const FdParams& p = c->p;
    c->tz = c->tzAdj;
    c->tx = c->txAdj;
    if (!c->haveVel) throw Error(AISO_ERR_STATE, "velocity not set");
    if (nShot < 1) throw Error(AISO_ERR_PARAM, "nShot must be >= 1");
    if (nRec != 1 && nRec != nShot) throw Error(AISO_ERR_PARAM, "receiver geometry count must be 1 or nShot");
    if (nShotSig != 1 && nShotSig != nShot) throw Error(AISO_ERR_PARAM, "model shot axis must be 1 or nShot");
    if (p.sub >= kSubMax) throw Error(AISO_ERR_PARAM, "**** ERROR: Subsampling parameter too high ****");
    if (every < 0) throw Error(AISO_ERR_PARAM, "checkpoint interval must be >= 0");
    if (p.freeSurface) throw Error(AISO_ERR_PARAM, "the velocity gradient does not support a free surface");
    const int nRecDev = recs[0]->g.nDev, nSrcDev = srcs[0]->g.nDev;
    for (int i = 0; i < nRec; ++i)
        if (recs[i]->g.nDev != nRecDev) throw Error(AISO_ERR_GEOM, "receivers per shot must be constant");
    for (int i = 0; i < nShot; ++i)
        if (srcs[i]->g.nDev != nSrcDev) throw Error(AISO_ERR_GEOM, "sources per shot must be constant");

    hipStream_t s = c->stream;
    if (mem == AISO_MEM_DEVICE) wait_default_stream(c);
    const long nzx = p.nzx();
    const int h = p.h(), T = p.ntw - 1;
    const float* f = p.filt.data();
    const int K = every > 0 ? std::min(every, T) : std::max(1, (int)std::ceil(std::sqrt((double)T)));
    const int nCk = (T - 1) / K + 1;                   // checkpoints at t = 0, K, 2K, ... < T

    std::vector<float> hModel((size_t)nShotSig * nSrcDev * p.nts);
    if (mem == AISO_MEM_DEVICE) HIPC(hipMemcpy(hModel.data(), model, hModel.size() * 4, hipMemcpyDeviceToHost));
    else std::memcpy(hModel.data(), model, hModel.size() * 4);
    const long dataN = (long)nShot * nRecDev * p.nts;
    const float* dGrad = dataGrad;
    if (mem != AISO_MEM_DEVICE) {
        float* d = (float*)c->dataDev.need(dataN * 4);
        HIPC(hipMemcpyAsync(d, dataGrad, dataN * 4, hipMemcpyHostToDevice, s));
        dGrad = d;
    }
    std::vector<DevGeom*> srcG(nShot), recG(nRec);
    int nSrcRegMax = 1, nRecRegMax = 1;
    for (int i = 0; i < nShot; ++i) {
        srcG[i] = dev_geom(c, srcs[i]);
        nSrcRegMax = std::max(nSrcRegMax, srcG[i]->nReg);
    }
    for (int i = 0; i < nRec; ++i) {
        recG[i] = dev_geom(c, recs[i]);
        nRecRegMax = std::max(nRecRegMax, recG[i]->nReg);
    }
    int maxSrcRing = 0, maxRecRing = 0;
    for (auto* g : srcG) maxSrcRing = std::max(maxSrcRing, g->nRing);
    for (auto* g : recG) maxRecRing = std::max(maxRecRing, g->nRing);

    // per shot: forward pair, adjoint pair, K segment slices, gradient, checkpoints
    const size_t perShot = (size_t)(2 + 2 + K + 1 + 2 * nCk) * nzx * 4 + (size_t)nRecRegMax * p.nts * 4 +
                           2 * (size_t)nSrcRegMax * p.ntw * 4;
    // 16 shots per batch by default, 8 per lane: at 400 x 1000 with 8 per
    // batch a lane's launch covers the GPU about once (MI355X: 230 against
    // 254 Gpts/s for 16; 12 and 32 gave 249 and 251).
    int B = c->batchAdj > 0 ? c->batchAdj : (c->batchReq > 0 ? c->batchReq : 16);
    {
        size_t freeB = 0, totalB = 0;
        HIPC(hipMemGetInfo(&freeB, &totalB));
        size_t have = c->pairs.n + c->pairsQ.n + c->W.n + c->img.n + c->dreg.n + c->sig.n + c->ss.n + c->ck.n;
        B = std::min(B, (int)std::max<size_t>(1, (size_t)((freeB + have) * 0.85) / perShot));
    }
    B = std::min(B, nShot);
    c->batchUsed = B;
    float* pairs = (float*)c->pairs.need((size_t)B * 2 * nzx * 4);
    float* pairsA = (float*)c->pairsQ.need((size_t)B * 2 * nzx * 4);
    float* W = (float*)c->W.need((size_t)B * K * nzx * 4);
    float* img = (float*)c->img.need((size_t)B * nzx * 4);
    float* ck = (float*)c->ck.need((size_t)B * 2 * nCk * nzx * 4);
    float* dreg = (float*)c->dreg.need((size_t)B * nRecRegMax * p.nts * 4);
    float* sig = (float*)c->sig.need((size_t)B * nSrcRegMax * p.ntw * 4);
    float* sigA = (float*)c->ss.need((size_t)B * nSrcRegMax * p.ntw * 4);
    ShotSlot* dF = (ShotSlot*)c->slots.need(sizeof(ShotSlot) * B);
    ShotSlot* dA = (ShotSlot*)c->slotsQ.need(sizeof(ShotSlot) * B);
    float* acc = (float*)c->acc.need(nzx * 4);
    HIPC(launch_fill(acc, nzx, s));
    c->curLanes = 1;

    StepArgs base{};
    base.v2dt2 = c->v2dt2.as<float>();
    base.cd = c->cd.as<float>();
    base.nz = p.nz; base.nx = p.nx; base.minPad = p.minPad; base.nts = p.nts; base.ntw = p.ntw;
    base.tilesZ = tiles_z(c);
    base.ntiles = base.tilesZ * tiles_x(c);
    for (int k = 0; k < 6; ++k) { base.zc[k] = p.zc[k]; base.xc[k] = p.xc[k]; }
    base.ntW = c->streamNT;
    GradArgs ga{};
    ga.nz = p.nz; ga.nx = p.nx;
    ga.tilesZ = base.tilesZ;
    ga.ntiles = base.ntiles;
    for (int k = 0; k < 6; ++k) { ga.zc[k] = p.zc[k]; ga.xc[k] = p.xc[k]; }

    // source-scaling terms: per shot, per regular source point
    std::vector<std::vector<std::pair<int, double>>> srcTerms(nShot);
    for (int b0 = 0; b0 < nShot; b0 += B) {
        const int nb = std::min(B, nShot - b0);
        std::vector<ShotSlot> fs(nb), as(nb);
        std::vector<float> sigHost((size_t)nb * nSrcRegMax * p.ntw, 0.f);
        std::vector<std::vector<float>> uHost(nb);
        for (int k = 0; k < nb; ++k) {
            const int shot = b0 + k;
            DevGeom* sg = srcG[shot];
            DevGeom* rg = recG[nRec == 1 ? 0 : shot];
            ShotSlot& F = fs[k];
            std::memset(&F, 0, sizeof(F));
            F.p[0] = pairs + (size_t)k * 2 * nzx;
            F.p[1] = F.p[0] + nzx;
            F.W = W + (size_t)k * K * nzx;
            F.img = img + (size_t)k * nzx;
            F.dreg = dreg + (size_t)k * nRecRegMax * p.nts;
            F.sig = sig + (size_t)k * nSrcRegMax * p.ntw;
            F.srcTile = sg->tilePtr.as<int>();
            F.srcEnt = sg->ent.as<int2>();
            F.srcRing = sg->ring.as<int2>();
            F.nSrcRing = sg->nRing;
            F.recTile = rg->tilePtr.as<int>();
            F.recEnt = rg->ent.as<int2>();
            F.recRing = rg->ring.as<int2>();
            F.nRecRing = rg->nRing;
            ShotSlot& A = as[k];
            A = F;
            A.p[0] = pairsA + (size_t)k * 2 * nzx;
            A.p[1] = A.p[0] + nzx;
            A.sig = sigA + (size_t)k * nSrcRegMax * p.ntw;
            const Geometry& sgm = srcs[shot]->g;
            const float* m = hModel.data() + (size_t)(nShotSig == 1 ? 0 : shot) * nSrcDev * p.nts;
            std::vector<float> fine = nl_source(c, sgm, m);
            std::memcpy(sigHost.data() + (size_t)k * nSrcRegMax * p.ntw, fine.data(), fine.size() * 4);
            const int nReg = (int)sgm.regPos.size();
            std::vector<float> reg((size_t)nReg * p.nts);
            uHost[k].resize((size_t)nReg * p.ntw);
            sgm.adjoint(false, reg.data(), m, p.nts);
            interp_time_fwd(p, nReg, reg.data(), uHost[k].data());   // unscaled source
            HIPC(launch_dev_adjoint(dreg + (size_t)k * nRecRegMax * p.nts, dGrad + (size_t)shot * nRecDev * p.nts,
                                    rg->adjPtr.as<int>(), rg->adjEnt.as<int>(), rg->weight.as<float>(), rg->nReg,
                                    recs[nRec == 1 ? 0 : shot]->g.nb, p.nts, s));
        }
        HIPC(hipMemcpyAsync(dF, fs.data(), sizeof(ShotSlot) * nb, hipMemcpyHostToDevice, s));
        HIPC(hipMemcpyAsync(dA, as.data(), sizeof(ShotSlot) * nb, hipMemcpyHostToDevice, s));
        HIPC(hipMemcpyAsync(sig, sigHost.data(), sigHost.size() * 4, hipMemcpyHostToDevice, s));
        HIPC(launch_fill(sigA, (long)nb * nSrcRegMax * p.ntw, s));
        HIPC(launch_fill(img, (long)nb * nzx, s));
        HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
        HIPC(launch_fill(pairsA, (long)nb * 2 * nzx, s));
        // The recompute writes interior points only; the imaging kernel's
        // Laplacian reads the FAT ring of these slices, which must hold 0.
        HIPC(launch_fill(W, (long)nb * K * nzx, s));

        StepArgs a = base;
        a.nshots = nb;
        a.gridLogical = a.ntiles * nb;
        ga.gridLogical = a.gridLogical;
        auto ckSlice = [&](int k, int j, int which) { return ck + (((size_t)k * nCk + j) * 2 + which) * nzx; };
        // Lanes: the batch's shots split over the context's streams; every
        // per-shot copy and every launch goes to the lane of its shots.
        auto lane_of = [&](int k) {
            for (int l = 0; l < c->curLanes; ++l) {
                int k0, cnt;
                lane_range(nb, c->curLanes, l, k0, cnt);
                if (k >= k0 && k < k0 + cnt) return lane_stream(c, l);
            }
            return s;
        };
        auto save = [&](int j, int cur) {
            for (int k = 0; k < nb; ++k) {
                hipStream_t ls = lane_of(k);
                HIPC(hipMemcpyAsync(ckSlice(k, j, 0), fs[k].p[cur], nzx * 4, hipMemcpyDeviceToDevice, ls));
                HIPC(hipMemcpyAsync(ckSlice(k, j, 1), fs[k].p[cur ^ 1], nzx * 4, hipMemcpyDeviceToDevice, ls));
            }
        };
        auto ring_lanes = [&](unsigned mode, const StepArgs& x, int mr) {
            for (int l = 0; l < c->curLanes; ++l) {
                int k0, cnt;
                lane_range(nb, c->curLanes, l, k0, cnt);
                if (cnt > 0) HIPC(launch_ring(mode, c->curLanes > 1 ? lane_args(x, k0, cnt) : x, mr, lane_stream(c, l)));
            }
        };
        // forward: one fine step = MODE_SRC without capture (step, injection, sponge)
        auto fwd_step = [&](int t, int cur, long capture) {
            StepArgs x = a;
            x.slots = dF;
            x.cur = cur;
            x.itwInj = t - 1;
            x.fe = 1.f;
            x.ff = 0.f;
            x.assignA = capture >= 0;
            x.skipA = capture < 0;
            x.wA = capture >= 0 ? capture : 0;
            x.assignB = 0;
            x.skipB = 1;
            launch_timed(c, MODE_SRC, AISO_K_SRC, x);
            if (maxSrcRing > 0) ring_lanes(M_INJ_SIG, x, maxSrcRing);
        };
        c->curLanes = std::max(1, std::min(c->lanes, nb));
        if (c->curLanes > 1) ensure_lanes(c);
        fork_lanes(c);
        int cur = 1;
        save(0, cur);
        for (int t = 1; t < T; ++t) {
            fwd_step(t, cur, -1);
            cur ^= 1;
            if (t % K == 0) save(t / K, cur);
        }
        // backward, segment by segment
        StepArgs r = a;
        r.slots = dA;
        r.cur = 1;
        ga.slots = dA;
        for (int j = nCk - 1; j >= 0; --j) {
            const int t0 = j * K, t1 = std::min(t0 + K, T);
            for (int k = 0; k < nb; ++k) {
                hipStream_t ls = lane_of(k);
                HIPC(hipMemcpyAsync(fs[k].p[1], ckSlice(k, j, 0), nzx * 4, hipMemcpyDeviceToDevice, ls));
                HIPC(hipMemcpyAsync(fs[k].p[0], ckSlice(k, j, 1), nzx * 4, hipMemcpyDeviceToDevice, ls));
                HIPC(hipMemcpyAsync(fs[k].W, ckSlice(k, j, 0), nzx * 4, hipMemcpyDeviceToDevice, ls));
            }
            int fc = 1;
            for (int t = t0 + 1; t < t1; ++t) {          // W[t - t0] = p_t
                fwd_step(t, fc, (long)(t - t0) * nzx);
                fc ^= 1;
            }
            for (int t = t1; t > t0; --t) {              // nu_t, then img += L(p_{t-1}) nu_t
                const int itw = t - 1, its = itw / p.sub, it2 = itw % p.sub;
                r.its = its;
                r.itwRec = itw;
                r.fc = f[it2 + 1];
                r.fd = f[h + it2 + 1];
                launch_timed(c, MODE_NL_ADJ, AISO_K_NL_ADJ, r);
                if (maxSrcRing > 0 || maxRecRing > 0)
                    ring_lanes(M_INJ_DATA | M_REC_SIG, r, std::max(maxSrcRing, maxRecRing));
                r.cur ^= 1;
                ga.wF = (long)(t - 1 - t0) * nzx;
                ga.nu = r.cur;
                for (int l = 0; l < c->curLanes; ++l) {
                    int k0, cnt;
                    lane_range(nb, c->curLanes, l, k0, cnt);
                    if (cnt == 0) continue;
                    GradArgs gl = ga;
                    gl.slots = dA + k0;
                    gl.gridLogical = ga.ntiles * cnt;
                    HIPC(launch_grad(c->tz, c->tx, gl, lane_stream(c, l)));
                }
            }
        }
        join_lanes(c);
        c->curLanes = 1;
        for (int k = 0; k < nb; ++k) HIPC(launch_accumulate(acc, img + (size_t)k * nzx, nzx, s));
        std::vector<float> sa((size_t)nb * nSrcRegMax * p.ntw);
        HIPC(hipMemcpyAsync(sa.data(), sigA, sa.size() * 4, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        for (int k = 0; k < nb; ++k) {
            const Geometry& sgm = srcs[b0 + k]->g;
            for (int rr = 0; rr < (int)sgm.regPos.size(); ++rr) {
                double sum = 0.0;
                const float* x = sa.data() + (size_t)k * nSrcRegMax * p.ntw + (size_t)rr * p.ntw;
                const float* u = uHost[k].data() + (size_t)rr * p.ntw;
                for (int i = 0; i < p.ntw; ++i) sum += (double)x[i] * (double)u[i];
                srcTerms[b0 + k].push_back({sgm.regPos[rr], sum});
            }
        }
        drain_profile(c);
    }
    std::vector<float> G2(nzx);
    HIPC(hipMemcpyAsync(G2.data(), acc, nzx * 4, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    for (auto& shot : srcTerms)
        for (auto& t : shot) G2[t.first] += (float)t.second;
    const float dtw = p.dtw;
    for (long i = 0; i < nzx; ++i) G2[i] = G2[i] * (((2.f * c->vel[i]) * dtw) * dtw);
    if (mem == AISO_MEM_DEVICE) HIPC(hipMemcpyAsync(velGrad, G2.data(), nzx * 4, hipMemcpyHostToDevice, s));
    else std::memcpy(velGrad, G2.data(), nzx * 4);
    HIPC(hipStreamSynchronize(s));
}

aiso_status op_call(aiso_ctx* c, Op op, int add, const float* mIn, float* mOut, const float* dIn, float* dOut,
                    int nShotSig, aiso_device* const* src, int nShot, const float* wav, int nWav,
                    aiso_device* const* rec, int nRec, int mem)
{
    return guard([&] {
        if (!c) throw Error(AISO_ERR_STATE, "null context");
        if (!src || !rec) throw Error(AISO_ERR_GEOM, "null geometry list");
        HIPC(hipSetDevice(c->device));
        Request q{op, add, mIn, mOut, dIn, dOut, nShotSig, src, nShot, wav, nWav, rec, nRec, mem};
        run_call(c, q);
    });
}

}  // namespace

// ================================================================= C-ABI ==
extern "C" {

const char* aiso_last_error(void) { return g_err.c_str(); }

aiso_status aiso_ctx_create(const aiso_fd_params* par, int device, aiso_ctx** out)
{
    return guard([&] {
        if (!par || !out) throw Error(AISO_ERR_PARAM, "null argument");
        auto c = std::make_unique<aiso_ctx>();
        c->p = make_params(*par);
        int n = 0;
        HIPC(hipGetDeviceCount(&n));
        if (device < 0 || device >= n)
            throw Error(AISO_ERR_PARAM, "**** ERROR [getGpuInfo]: One of the element of the GPU Id list is not a valid GPU Id number ****");
        c->device = device;
        HIPC(hipSetDevice(device));
        HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        *out = c.release();
    });
}

aiso_status aiso_ctx_destroy(aiso_ctx* c)
{
    return guard([&] {
        if (!c) return;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (auto& e : c->ev) { (void)hipEventDestroy(e.e0); (void)hipEventDestroy(e.e1); }
        for (auto e : c->evPool) (void)hipEventDestroy(e);
        if (c->evIn) (void)hipEventDestroy(c->evIn);
        if (c->forkEv) (void)hipEventDestroy(c->forkEv);
        if (c->stageEv) (void)hipEventDestroy(c->stageEv);
        for (auto e : c->joinEv) (void)hipEventDestroy(e);
        for (auto l : c->lane) {
            (void)hipStreamSynchronize(l);
            (void)hipStreamDestroy(l);
        }
        c->geoms.clear();
        hipStream_t s = c->stream;
        delete c;
        (void)hipStreamDestroy(s);
    });
}

aiso_status aiso_ctx_set_velocity(aiso_ctx* c, const float* vel, int mem)
{
    return guard([&] {
        if (!c || !vel) throw Error(AISO_ERR_PARAM, "null argument");
        HIPC(hipSetDevice(c->device));
        const FdParams& p = c->p;
        std::vector<float> v(p.nzx());
        if (mem == AISO_MEM_DEVICE) HIPC(hipMemcpy(v.data(), vel, v.size() * 4, hipMemcpyDeviceToHost));
        else std::memcpy(v.data(), vel, v.size() * 4);
        check_velocity(p, v.data());
        c->vel = std::move(v);
        c->v2dt2h.assign(p.nzx(), 0.f);
        c->reflh.assign(p.nzx(), 0.f);
        velocity_scales(p, c->vel.data(), c->v2dt2h.data(), c->reflh.data());
        upload(c->v2dt2, c->v2dt2h, c->stream);
        upload(c->refl, c->reflh, c->stream);
        upload(c->cd, p.cd, c->stream);
        HIPC(hipStreamSynchronize(c->stream));
        c->haveVel = true;
    });
}

aiso_status aiso_ctx_set_batch(aiso_ctx* c, int n)
{
    return guard([&] {
        if (!c || n < 0) throw Error(AISO_ERR_PARAM, "bad batch size");
        c->batchReq = n;
    });
}

aiso_status aiso_ctx_set_tile(aiso_ctx* c, int tz, int tx)
{
    return guard([&] {
        if (!c) throw Error(AISO_ERR_PARAM, "null context");
        if ((tz != 32 && tz != 64) || (tx != 16 && tx != 32))
            throw Error(AISO_ERR_PARAM, "tile must be tz in {32,64}, tx in {16,32}");
        c->tzFwd = c->tzAdj = tz;
        c->txFwd = c->txAdj = tx;
        c->tileFwdSet = true;
    });
}

aiso_status aiso_ctx_set_option(aiso_ctx* c, const char* key, int value)
{
    return guard([&] {
        if (!c || !key) throw Error(AISO_ERR_PARAM, "null argument");
        std::string k(key);
        if (k == "streamNT") c->streamNT = value != 0;
        else if (k == "lockstep") c->lockstep = value != 0;
        else if (k == "pair2") c->pair2 = value != 0;
        else if (k == "leg1") c->leg1 = value != 0;
        else if (k == "leg2") c->leg2 = value != 0;
        else if (k == "sloth") c->sloth = value != 0;
        else if (k == "imagePerShot") c->imagePerShot = value != 0;
        else if (k == "saveWavefield") c->saveWavefield = value != 0;
        else if (k == "wavefieldShot") {
            if (value < 0) throw Error(AISO_ERR_PARAM, "wavefieldShot must be >= 0");
            c->wavefieldShot = value;
        }
        else if (k == "extBatch") {
            if (value < 1 || value > kExtBatchMax) throw Error(AISO_ERR_PARAM, "extBatch must be 1..16");
            c->extBatch = value;
        }
        else if (k == "lanes") {
            if (value < 1 || value > 4) throw Error(AISO_ERR_PARAM, "lanes must be 1..4");
            c->lanes = value;
        }
        else if (k == "tileFwd" || k == "tileAdj") {   // tz * 100 + tx
            const int tz = value / 100, tx = value % 100;
            if ((tz != 32 && tz != 64) || (tx != 16 && tx != 32))
                throw Error(AISO_ERR_PARAM, "tile must be tz in {32,64}, tx in {16,32} (value tz*100+tx)");
            if (k == "tileFwd") { c->tzFwd = tz; c->txFwd = tx; c->tileFwdSet = true; }
            else { c->tzAdj = tz; c->txAdj = tx; }
        }
        else if (k == "batchFwd") c->batchFwd = std::max(0, value);
        else if (k == "batchAdj") c->batchAdj = std::max(0, value);
        else if (k == "batchRec") c->batchRec = std::max(0, value);
        else if (k == "shotsPerBlock") {
            if (value != 1 && value != 2) throw Error(AISO_ERR_PARAM, "shotsPerBlock must be 1 or 2");
            c->spb = value;
        }
        else if (k == "occupancy") {
            if (value != 0 && value != 6 && value != 8) throw Error(AISO_ERR_PARAM, "occupancy must be 0, 6 or 8");
            c->occ = value;
        }
        else if (k == "cacheBudgetMB") c->cacheBudget = (long)std::max(1, value) << 20;
        else if (k == "profileEvery") {
            if (value < 1) throw Error(AISO_ERR_PARAM, "profileEvery must be >= 1");
            c->profEvery = value;
        }
        else throw Error(AISO_ERR_PARAM, "unknown option " + k);
    });
}

aiso_status aiso_ctx_get_option(const aiso_ctx* c, const char* key, int* value)
{
    return guard([&] {
        if (!c || !key || !value) throw Error(AISO_ERR_PARAM, "null argument");
        const std::string k(key);
        if (k == "streamNT") *value = c->streamNT;
        else if (k == "lockstep") *value = c->lockstep;
        else if (k == "pair2") *value = c->pair2;
        else if (k == "leg1") *value = c->leg1;
        else if (k == "leg2") *value = c->leg2;
        else if (k == "sloth") *value = c->sloth;
        else if (k == "imagePerShot") *value = c->imagePerShot;
        else if (k == "saveWavefield") *value = c->saveWavefield;
        else if (k == "wavefieldShot") *value = c->wavefieldShot;
        else if (k == "extBatch") *value = c->extBatch;
        else if (k == "lanes") *value = c->lanes;
        else if (k == "tileFwd") *value = c->tzFwd * 100 + c->txFwd;
        else if (k == "tileAdj") *value = c->tzAdj * 100 + c->txAdj;
        else if (k == "batchFwd") *value = c->batchFwd;
        else if (k == "batchAdj") *value = c->batchAdj;
        else if (k == "batchRec") *value = c->batchRec;
        else if (k == "shotsPerBlock") *value = c->spb;
        else if (k == "occupancy") *value = c->occ;
        else if (k == "cacheBudgetMB") *value = (int)(c->cacheBudget >> 20);
        else if (k == "profileEvery") *value = c->profEvery;
        else throw Error(AISO_ERR_PARAM, "unknown option " + k);
    });
}

aiso_status aiso_ctx_info(const aiso_ctx* c, int* ntw, int* batchUsed, size_t* bytesPerShot)
{
    return guard([&] {
        if (!c) throw Error(AISO_ERR_PARAM, "null context");
        if (ntw) *ntw = c->p.ntw;
        if (batchUsed) *batchUsed = c->batchUsed;
        if (bytesPerShot) *bytesPerShot = c->bytesPerShot;
    });
}

aiso_status aiso_ctx_stream(const aiso_ctx* c, void** stream)
{
    return guard([&] {
        if (!c || !stream) throw Error(AISO_ERR_PARAM, "null argument");
        *stream = (void*)c->stream;
    });
}

aiso_status aiso_ctx_profile(aiso_ctx* c, int enable)
{
    return guard([&] {
        if (!c) throw Error(AISO_ERR_PARAM, "null context");
        drain_profile(c);
        c->profile = enable != 0;
        for (int k = 0; k < AISO_K_COUNT; ++k) {
            c->launches[k] = 0;
            c->ms[k] = 0;
            c->pts[k] = 0;
            c->profSeq[k] = 0;
        }
        c->batchDone.clear();
    });
}

aiso_status aiso_ctx_batch_read(aiso_ctx* c, float* ms, int* shots, int maxBatches, int* count)
{
    return guard([&] {
        if (!c) throw Error(AISO_ERR_PARAM, "null context");
        drain_profile(c);
        const int n = (int)c->batchDone.size();
        for (int i = 0; i < std::min(n, maxBatches); ++i) {
            if (ms) ms[i] = c->batchDone[i].first;
            if (shots) shots[i] = c->batchDone[i].second;
        }
        if (count) *count = n;
    });
}

aiso_status aiso_ctx_profile_read(aiso_ctx* c, int which, long long* launches, double* ms, double* pts)
{
    return guard([&] {
        if (!c) throw Error(AISO_ERR_PARAM, "null context");
        if (which < AISO_K_ALL || which >= AISO_K_COUNT) throw Error(AISO_ERR_PARAM, "bad kernel variant");
        drain_profile(c);
        long long l = 0;
        double m = 0, p = 0;
        for (int k = 0; k < AISO_K_COUNT; ++k)
            if (which == AISO_K_ALL || which == k) {
                l += c->launches[k];
                m += c->ms[k];
                p += c->pts[k];
            }
        if (launches) *launches = l;
        if (ms) *ms = m;
        if (pts) *pts = p;
    });
}

static aiso_status wrap_dev(Geometry&& g, aiso_device** out)
{
    auto d = std::make_unique<aiso_device>();
    d->g = std::move(g);
    *out = d.release();
    return AISO_OK;
}

aiso_status aiso_device_create_coords(int nDev, const float* z, const float* x, int nzVel, int nxVel, float oz,
                                      float dz, float ox, float dx, int nt, int dipole, float zShift, float xShift,
                                      aiso_device** out)
{
    return guard([&] {
        if (!z || !x || !out) throw Error(AISO_ERR_PARAM, "null argument");
        wrap_dev(geom_coords(nDev, z, x, nzVel, nxVel, oz, dz, ox, dx, nt, dipole, zShift, xShift), out);
    });
}

aiso_status aiso_device_create_ints(int nDev, const int* z, const int* x, int nzVel, int nxVel, int nt, int dipole,
                                    int zShift, int xShift, aiso_device** out)
{
    return guard([&] {
        if (!z || !x || !out) throw Error(AISO_ERR_PARAM, "null argument");
        wrap_dev(geom_ints(nDev, z, x, nzVel, nxVel, nt, dipole, zShift, xShift), out);
    });
}

aiso_status aiso_device_create_grid(int nzDev, int ozDev, int dzDev, int nxDev, int oxDev, int dxDev, int nzVel,
                                    int nxVel, int nt, int dipole, int zShift, int xShift, aiso_device** out)
{
    return guard([&] {
        if (!out) throw Error(AISO_ERR_PARAM, "null argument");
        wrap_dev(geom_grid(nzDev, ozDev, dzDev, nxDev, oxDev, dxDev, nzVel, nxVel, nt, dipole, zShift, xShift), out);
    });
}

aiso_status aiso_device_destroy(aiso_device* d)
{
    delete d;
    return AISO_OK;
}

aiso_status aiso_device_sizes(const aiso_device* d, int* nIrreg, int* nReg, int* nb)
{
    return guard([&] {
        if (!d) throw Error(AISO_ERR_PARAM, "null device");
        if (nIrreg) *nIrreg = d->g.nDev;
        if (nReg) *nReg = (int)d->g.regPos.size();
        if (nb) *nb = d->g.nb;
    });
}

aiso_status aiso_device_reg_pos(const aiso_device* d, int* regPos)
{
    return guard([&] {
        if (!d || !regPos) throw Error(AISO_ERR_PARAM, "null argument");
        std::copy(d->g.regPos.begin(), d->g.regPos.end(), regPos);
    });
}

aiso_status aiso_device_corners(const aiso_device* d, int* gridIdx, float* weight)
{
    return guard([&] {
        if (!d) throw Error(AISO_ERR_PARAM, "null device");
        if (gridIdx) std::copy(d->g.gridIdx.begin(), d->g.gridIdx.end(), gridIdx);
        if (weight) std::copy(d->g.weight.begin(), d->g.weight.end(), weight);
    });
}

aiso_status aiso_device_forward(const aiso_device* d, int add, const float* reg, float* irreg, int nt)
{
    return guard([&] {
        if (!d || !reg || !irreg) throw Error(AISO_ERR_PARAM, "null argument");
        d->g.forward(add != 0, reg, irreg, nt);
    });
}

aiso_status aiso_device_adjoint(const aiso_device* d, int add, float* reg, const float* irreg, int nt)
{
    return guard([&] {
        if (!d || !reg || !irreg) throw Error(AISO_ERR_PARAM, "null argument");
        d->g.adjoint(add != 0, reg, irreg, nt);
    });
}

aiso_status aiso_nl_forward(aiso_ctx* c, int add, const float* model, int nShotSig, float* data,
                            aiso_device* const* src, int nShot, aiso_device* const* rec, int nRec, int mem)
{
    return op_call(c, Op::NlFwd, add, model, nullptr, nullptr, data, nShotSig, src, nShot, nullptr, 1, rec, nRec, mem);
}

aiso_status aiso_nl_adjoint(aiso_ctx* c, int add, float* model, int nShotSig, const float* data,
                            aiso_device* const* src, int nShot, aiso_device* const* rec, int nRec, int mem)
{
    return op_call(c, Op::NlAdj, add, nullptr, model, data, nullptr, nShotSig, src, nShot, nullptr, 1, rec, nRec, mem);
}

aiso_status aiso_nl_velocity_gradient(aiso_ctx* c, const float* model, int nShotSig, const float* dataGrad,
                                      float* velGrad, aiso_device* const* src, int nShot, aiso_device* const* rec,
                                      int nRec, int checkpointEvery, int mem)
{
    return guard([&] {
        if (!c) throw Error(AISO_ERR_STATE, "null context");
        if (!src || !rec || !model || !dataGrad || !velGrad) throw Error(AISO_ERR_PARAM, "null argument");
        HIPC(hipSetDevice(c->device));
        run_nl_grad(c, model, nShotSig, dataGrad, velGrad, src, nShot, rec, nRec, checkpointEvery, mem);
    });
}

aiso_status aiso_born_forward(aiso_ctx* c, int add, const float* model, float* data, aiso_device* const* src,
                              int nShot, const float* wav, int nWav, aiso_device* const* rec, int nRec, int mem)
{
    return op_call(c, Op::BornFwd, add, model, nullptr, nullptr, data, 1, src, nShot, wav, nWav, rec, nRec, mem);
}

aiso_status aiso_born_adjoint(aiso_ctx* c, int add, float* model, const float* data, aiso_device* const* src,
                              int nShot, const float* wav, int nWav, aiso_device* const* rec, int nRec, int mem)
{
    return op_call(c, Op::BornAdj, add, nullptr, model, data, nullptr, 1, src, nShot, wav, nWav, rec, nRec, mem);
}

aiso_status aiso_bornext_forward(aiso_ctx* c, int add, const float* model, float* data, aiso_device* const* src,
                                 int nShot, const float* wav, int nWav, aiso_device* const* rec, int nRec, int mem)
{
    return op_call(c, Op::ExtFwd, add, model, nullptr, nullptr, data, 1, src, nShot, wav, nWav, rec, nRec, mem);
}

aiso_status aiso_bornext_adjoint(aiso_ctx* c, int add, float* model, const float* data, aiso_device* const* src,
                                 int nShot, const float* wav, int nWav, aiso_device* const* rec, int nRec, int mem)
{
    return op_call(c, Op::ExtAdj, add, nullptr, model, data, nullptr, 1, src, nShot, wav, nWav, rec, nRec, mem);
}

aiso_status aiso_ctx_wavefield(aiso_ctx* c, int which, float* out, int mem)
{
    return guard([&] {
        if (!c || !out) throw Error(AISO_ERR_PARAM, "null argument");
        if (which != AISO_WAVEFIELD_SOURCE && which != AISO_WAVEFIELD_SECONDARY)
            throw Error(AISO_ERR_PARAM, "which must be AISO_WAVEFIELD_SOURCE or AISO_WAVEFIELD_SECONDARY");
        const bool src = which == AISO_WAVEFIELD_SOURCE;
        if (!(src ? c->wfSrcValid : c->wfSecValid))
            throw Error(AISO_ERR_STATE, src ? "no wavefield saved: call an operator with saveWavefield set first"
                                            : "no secondary wavefield saved: only the Born forward / adjoint save one");
        HIPC(hipSetDevice(c->device));
        const size_t bytes = (size_t)c->p.nts * c->p.nzx() * 4;
        const aiso::DBuf& b = src ? c->wfSrc : c->wfSec;
        HIPC(hipMemcpyAsync(out, b.p, bytes, mem == AISO_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                            c->stream));
        HIPC(hipStreamSynchronize(c->stream));
    });
}

aiso_status aiso_ctx_set_ext_reflectivity(aiso_ctx* c, const float* refl, int mem)
{
    return guard([&] {
        if (!c || !refl) throw Error(AISO_ERR_PARAM, "null argument");
        HIPC(hipSetDevice(c->device));
        const size_t n = (size_t)c->p.nzx() * c->p.nExt;
        float* d = (float*)c->extRefl.need(n * 4);
        if (mem == AISO_MEM_DEVICE) {
            wait_default_stream(c);
            HIPC(hipMemcpyAsync(d, refl, n * 4, hipMemcpyDeviceToDevice, c->stream));
        } else {
            HIPC(hipMemcpyAsync(d, refl, n * 4, hipMemcpyHostToDevice, c->stream));
        }
        HIPC(hipStreamSynchronize(c->stream));
        c->haveExtRefl = true;
    });
}

aiso_status aiso_tomo_forward(aiso_ctx* c, int add, const float* model, float* data, aiso_device* const* src,
                              int nShot, const float* wav, int nWav, aiso_device* const* rec, int nRec, int mem)
{
    return op_call(c, Op::TomoFwd, add, model, nullptr, nullptr, data, 1, src, nShot, wav, nWav, rec, nRec, mem);
}

aiso_status aiso_tomo_adjoint(aiso_ctx* c, int add, float* model, const float* data, aiso_device* const* src,
                              int nShot, const float* wav, int nWav, aiso_device* const* rec, int nRec, int mem)
{
    return op_call(c, Op::TomoAdj, add, nullptr, model, data, nullptr, 1, src, nShot, wav, nWav, rec, nRec, mem);
}

aiso_status aiso_wemva_forward(aiso_ctx* c, int add, const float* model, float* image, const float* data,
                               aiso_device* const* src, int nShot, const float* wav, int nWav,
                               aiso_device* const* rec, int nRec, int mem)
{
    return op_call(c, Op::WemvaFwd, add, model, image, data, nullptr, 1, src, nShot, wav, nWav, rec, nRec, mem);
}

aiso_status aiso_wemva_adjoint(aiso_ctx* c, int add, float* model, const float* image, const float* data,
                               aiso_device* const* src, int nShot, const float* wav, int nWav,
                               aiso_device* const* rec, int nRec, int mem)
{
    return op_call(c, Op::WemvaAdj, add, image, model, data, nullptr, 1, src, nShot, wav, nWav, rec, nRec, mem);
}

int aiso_pad_plus(int n, int pad, int blockSize) { return pad_plus(n, pad, blockSize); }
int aiso_pad_plus_free_surface(int nz, int zPad, int blockSize) { return pad_plus_free_surface(nz, zPad, blockSize); }

}  // extern "C"
