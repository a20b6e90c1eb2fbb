// engine.cpp -- the C-ABI of libaiso.so: contexts, geometry, and the
// multi-shot operators (nonlinear, Born, offset-extended Born).
//
// The reference runs one shot per GPU at a time, re-creating its operator
// objects and device buffers on every call (F/src/BornShotsGpu.cpp:82-143).
// Here a context keeps the velocity tables, geometry lists and work buffers
// resident on the device, and every fine time step advances a whole batch of
// shots with one launch of the fused step kernel (kernels.hip).  Per-shot
// results are reduced in shot order, so they do not depend on the batch size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <functional>
#include <vector>

#include "aiso.h"
#include "host.hpp"
#include "kernels.hpp"

using namespace aiso;

namespace {

thread_local std::string g_err;

void hip_check(hipError_t e, const char* what)
{
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        throw Error(AISO_ERR_MEMORY, std::string("out of device memory in ") + what);
    if (e != hipSuccess)
        throw Error(AISO_ERR_HIP, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

#define HIPC(x) hip_check((x), #x)

// Grow-only device buffer.
struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void* need(size_t bytes)
    {
        if (bytes <= n && p) return p;
        release();
        if (bytes == 0) bytes = 4;
        HIPC(hipMalloc(&p, bytes));
        n = bytes;
        return p;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

template <class T>
void upload(DBuf& b, const std::vector<T>& v, hipStream_t s)
{
    b.need(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (!v.empty()) HIPC(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

std::atomic<long> g_geom_id{1};

// Per-tile sparse lists of one geometry for one tile shape, plus the arrays
// of the device-side spatial interpolation.
struct DevGeom {
    DBuf tilePtr, ent, ring, cornerReg, weight, adjPtr, adjEnt;
    int nRing = 0, nReg = 0;
};

}  // namespace

struct aiso_device {
    Geometry g;
    long id = g_geom_id++;
};

struct aiso_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    FdParams p;
    bool haveVel = false;
    std::vector<float> vel, v2dt2h, reflh;
    DBuf v2dt2, refl, cd;
    int batchReq = 0, batchUsed = 0;
    int tz = 32, tx = 32;               // tile of the call in progress
    // Step-kernel tile per direction: the forward's pair kernel runs best on
    // 32 x 16 tiles, the adjoint passes on 32 x 32 (MI355X measurements).
    int tzFwd = 32, txFwd = 16, tzAdj = 32, txAdj = 32;
    bool tileFwdSet = false;            // the caller chose the forward tile
    int streamNT = 1;
    size_t bytesPerShot = 0;
    // work buffers
    DBuf pairs, pairsQ, W, ss, img, dreg, sig, slots, slotsQ, mtil, acc, modelDev, dataDev;
    DBuf ck;                            // checkpoints of the velocity-gradient pass
    // tomographic operator: the extended reflectivity it linearizes about
    // (set once, [nExt][nx][nz]), its per-call 2/v^3 scaling, and the
    // scattered / receiver wavefield and capture buffers
    DBuf extRefl, rExt, P, cap, slotsC, slotsI, slotsJ, imgDev;
    bool haveExtRefl = false;
    int leg1 = 1, leg2 = 1;
    // time-lag extension: the per-point model scale of the call and the
    // sloth key (scale by -v^2 dtw^2 instead of 2/v^3 v^2 dtw^2)
    DBuf tscale;
    int sloth = 0;
    // Born and extended-Born adjoints: one image per shot ([nShot][nExt][nx][nz])
    // instead of their sum, for callers that batch independent requests.
    int imagePerShot = 0;
    hipEvent_t evIn = nullptr;          // orders device inputs after the default stream's work
    // Lanes: the shots of a batch split over streams whose kernels run
    // concurrently, so one lane's launch gaps and tails are filled by the
    // other's.  On MI355X two lanes gain about 1 % on the Born benchmark, so
    // the option is off by default.
    int lanes = 1;
    int extBatch = 16;                  // coarse samples per offset-imaging / secondary-source launch
    int curLanes = 1;                   // lanes of the call in progress
    std::vector<hipStream_t> lane;
    hipEvent_t forkEv = nullptr;
    std::vector<hipEvent_t> joinEv;
    int lockstep = 1;
    // Two fine steps per launch in the Born forward and adjoint (sub == 2).
    // They move half the HBM bytes of the one-step kernels but run more
    // instructions per point (step A covers the tile plus halo), which on
    // MI355X costs more than it saves, so they are opt-in.
    int pair2 = 0;
    int batchFwd = 0, batchAdj = 0;     // per-direction overrides of batchReq
    int batchRec = 0;                   // receiver-pass group size within an adjoint batch
    int spb = 1;                        // shots per workgroup of the step kernel
    long cacheBudget = 192l << 20;      // bytes of hot slices per launch (Infinity Cache is 256 MB)
    std::map<std::tuple<long, int, int>, std::unique_ptr<DevGeom>> geoms;
    // profiling of the step kernels
    bool profile = false;
    // With lanes > 1 the dispatches of different lanes overlap in time, so
    // their durations are those of concurrent kernels.
    struct Timed { hipEvent_t e0, e1; int kind; double pts; int n = 1; };   // n launches between e0 and e1
    std::vector<Timed> ev;
    struct Batch { hipEvent_t e0, e1; int shots; };
    std::vector<Batch> batchEv;
    std::vector<std::pair<float, int>> batchDone;
    std::vector<hipEvent_t> evPool;
    long long launches[AISO_K_COUNT] = {0};
    // Events go on every profEvery-th dispatch of each kind (sampling keeps
    // the timestamps' cost off the other dispatches).
    int profEvery = 1;
    long long profSeq[AISO_K_COUNT] = {0};
    double ms[AISO_K_COUNT] = {0}, pts[AISO_K_COUNT] = {0};   // pts: point-updates (both fields of a pair)
};

namespace {

int tilesZ(const aiso_ctx* c) { return (c->p.nz - 2 * kFat + c->tz - 1) / c->tz; }
int tilesX(const aiso_ctx* c) { return (c->p.nx - 2 * kFat + c->tx - 1) / c->tx; }

// Bucket the unique points of a geometry by step-kernel tile.
DevGeom* dev_geom(aiso_ctx* c, const aiso_device* d)
{
    auto key = std::make_tuple(d->id, c->tz, c->tx);
    auto it = c->geoms.find(key);
    if (it != c->geoms.end()) return it->second.get();
    auto dg = std::make_unique<DevGeom>();
    const Geometry& g = d->g;
    const int nz = c->p.nz, nx = c->p.nx, tZ = tilesZ(c), nt = tZ * tilesX(c);
    std::vector<std::vector<int2>> buckets(nt);
    std::vector<int2> ring;
    for (int r = 0; r < (int)g.regPos.size(); ++r) {
        int pos = g.regPos[r], iz = pos % nz, ix = pos / nz;
        if (iz >= kFat && iz < nz - kFat && ix >= kFat && ix < nx - kFat) {
            int t = (iz - kFat) / c->tz + tZ * ((ix - kFat) / c->tx);
            int local = ((ix - kFat) % c->tx) * c->tz + (iz - kFat) % c->tz;
            buckets[t].push_back(make_int2(local, r));
        } else {
            ring.push_back(make_int2(pos, r));
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
    upload(dg->ring, ring, c->stream);
    upload(dg->cornerReg, g.cornerReg, c->stream);
    upload(dg->weight, g.weight, c->stream);
    upload(dg->adjPtr, g.adjPtr, c->stream);
    upload(dg->adjEnt, g.adjEnt, c->stream);
    dg->nRing = (int)ring.size();
    dg->nReg = (int)g.regPos.size();
    HIPC(hipStreamSynchronize(c->stream));
    DevGeom* out = dg.get();
    c->geoms.emplace(key, std::move(dg));
    return out;
}

// Tiling of the two-step kernel: output tiles of OZ x OX points, each
// computing step A on the EZ x EX region around it.
constexpr int kOZ2 = kPair2EZ - 2 * kFat, kOX2 = kPair2EX - 2 * kFat;
int tilesZ2(const aiso_ctx* c) { return (c->p.nz - 2 * kFat + kOZ2 - 1) / kOZ2; }
int tilesX2(const aiso_ctx* c) { return (c->p.nx - 2 * kFat + kOX2 - 1) / kOX2; }

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
    const int nz = c->p.nz, nx = c->p.nx, tZ = tilesZ2(c), tX = tilesX2(c), nt = tZ * tX;
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

enum class Op { NlFwd, NlAdj, BornFwd, BornAdj, ExtFwd, ExtAdj, TomoFwd, TomoAdj, WemvaFwd, WemvaAdj };

// The adjoint of each operator writes the model side (for WEMVA, the
// velocity perturbation); both WEMVA directions read observed data.
bool is_adj(Op op)
{
    return op == Op::NlAdj || op == Op::BornAdj || op == Op::ExtAdj || op == Op::TomoAdj || op == Op::WemvaAdj;
}
bool is_born(Op op) { return op != Op::NlFwd && op != Op::NlAdj; }
bool is_ext(Op op) { return op == Op::ExtFwd || op == Op::ExtAdj; }
bool is_tomo(Op op) { return op == Op::TomoFwd || op == Op::TomoAdj; }
bool is_wemva(Op op) { return op == Op::WemvaFwd || op == Op::WemvaAdj; }

struct Request {
    Op op;
    int add;
    const float* modelIn;   // forward input
    float* modelOut;        // adjoint output
    const float* dataIn;
    float* dataOut;
    int nShotSig;           // NL model shots
    aiso_device* const* src;
    int nShot;
    const float* wavelets;
    int nWav;
    aiso_device* const* rec;
    int nRec;
    int mem;
};

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
    return c->profile && c->profSeq[kind]++ % c->profEvery == 0;
}

void launch_timed(aiso_ctx* c, unsigned mode, int kind, const StepArgs& a, int nshots, bool generic = false)
{
    (void)nshots;
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
        else HIPC(launch_step(mode, c->tz, c->tx, c->spb, y, lane_stream(c, l), e0, e1));
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

void run(aiso_ctx* c, const Request& q)
{
    const FdParams& p = c->p;
    c->tz = is_adj(q.op) ? c->tzAdj : c->tzFwd;
    c->tx = is_adj(q.op) ? c->txAdj : c->txFwd;
    // The forward passes of the extended, tomographic and WEMVA operators,
    // and every pass with a free surface, run on the generic step kernel,
    // which is faster on 32 x 32 tiles than on the specialized forward
    // kernels' 32 x 16 (2-6 % on MI355X).
    const bool onGeneric = is_ext(q.op) || is_tomo(q.op) || is_wemva(q.op) || c->p.freeSurface != 0;
    if (!is_adj(q.op) && !c->tileFwdSet && onGeneric) {
        c->tz = 32;
        c->tx = 32;
    }
    if (!c->haveVel) throw Error(AISO_ERR_STATE, "velocity not set");
    if (q.nShot < 1) throw Error(AISO_ERR_PARAM, "nShot must be >= 1");
    if (q.nRec != 1 && q.nRec != q.nShot) throw Error(AISO_ERR_PARAM, "receiver geometry count must be 1 or nShot");
    if ((!is_born(q.op) || is_ext(q.op)) && p.sub >= kSubMax) throw Error(AISO_ERR_PARAM, "**** ERROR: Subsampling parameter too high ****");
    if ((!is_born(q.op) || is_ext(q.op)) && p.minPad >= kPadMax) throw Error(AISO_ERR_PARAM, "**** ERROR: Padding value is too high ****");
    if (is_ext(q.op) && p.extension == AISO_EXT_NONE && p.nExt > 1)
        throw Error(AISO_ERR_PARAM, "extended Born needs extension=offset or extension=time");
    if (is_tomo(q.op) || is_wemva(q.op)) {
        if (p.extension == AISO_EXT_NONE)
            throw Error(AISO_ERR_PARAM, is_tomo(q.op) ? "the tomographic operator needs extension=offset or extension=time"
                                                      : "the WEMVA operator needs extension=offset or extension=time");
        if (is_tomo(q.op) && !c->haveExtRefl) throw Error(AISO_ERR_STATE, "extended reflectivity not set");
        if (p.sub >= kSubMax) throw Error(AISO_ERR_PARAM, "**** ERROR: Subsampling parameter too high ****");
        if (p.minPad >= kPadMax) throw Error(AISO_ERR_PARAM, "**** ERROR: Padding value is too high ****");
    }
    if (is_born(q.op) && q.nWav != 1 && q.nWav != q.nShot) throw Error(AISO_ERR_PARAM, "wavelet count must be 1 or nShot");
    if (!is_born(q.op) && q.nShotSig != 1 && q.nShotSig != q.nShot)
        throw Error(AISO_ERR_PARAM, "model shot axis must be 1 or nShot");
    const int nRecDev = q.rec[0]->g.nDev;
    for (int i = 0; i < q.nRec; ++i)
        if (q.rec[i]->g.nDev != nRecDev) throw Error(AISO_ERR_GEOM, "receivers per shot must be constant");
    const int nSrcDev = q.src[0]->g.nDev;
    for (int i = 0; i < q.nShot; ++i)
        if (q.src[i]->g.nDev != nSrcDev) throw Error(AISO_ERR_GEOM, "sources per shot must be constant");
    for (int i = 0; i < q.nShot; ++i)
        for (int v : q.src[i]->g.gridIdx)
            if (v >= p.nzx()) throw Error(AISO_ERR_GEOM, "source outside the velocity grid");
    for (int i = 0; i < q.nRec; ++i)
        for (int v : q.rec[i]->g.gridIdx)
            if (v >= p.nzx()) throw Error(AISO_ERR_GEOM, "receiver outside the velocity grid");
    // The reference's WEMVA has no free-surface routines: it propagates with
    // the absorbing top whatever the key (F/src/wemvaExtGpu.cpp:131-135).
    if (p.freeSurface && !is_wemva(q.op)) {   // the surface condition rewrites the rows above it every step
        auto above = [&](const aiso_device* d) {
            for (size_t k = 0; k < d->g.gridIdx.size(); ++k)
                if (d->g.weight[k] != 0.f && d->g.gridIdx[k] % p.nz < kFat) return true;
            return false;
        };
        for (int i = 0; i < q.nShot; ++i)
            if (above(q.src[i])) throw Error(AISO_ERR_GEOM, "with a free surface, sources must lie at or below the surface row");
        for (int i = 0; i < q.nRec; ++i)
            if (above(q.rec[i])) throw Error(AISO_ERR_GEOM, "with a free surface, receivers must lie at or below the surface row");
    }

    hipStream_t s = c->stream;
    if (q.mem == AISO_MEM_DEVICE) wait_default_stream(c);
    const long nzx = p.nzx();
    const int nExt = is_ext(q.op) ? p.nExt : 1;
    const long modelN = is_born(q.op) ? nzx * nExt : (long)q.nShotSig * nSrcDev * p.nts;
    // WEMVA: the velocity perturbation and the extended image are its two
    // sides (model in / out differ in size); both directions read data.
    const bool wemva = is_wemva(q.op);
    const int nImg = is_ext(q.op) || q.op == Op::WemvaFwd ? p.nExt : 1;   // slices of the per-shot image
    const bool perShotImg = c->imagePerShot && (q.op == Op::BornAdj || q.op == Op::ExtAdj);
    const long mInN = q.op == Op::WemvaAdj ? nzx * p.nExt : modelN;
    const long mOutN = q.op == Op::WemvaFwd ? nzx * p.nExt : perShotImg ? (long)q.nShot * nImg * nzx : modelN;
    const bool dataInput = is_adj(q.op) || wemva;
    const long dataN = (long)q.nShot * nRecDev * p.nts;
    const int h = p.h();
    const float* f = p.filt.data();

    // ---- operands on the device
    const float* dModelIn = nullptr;
    float* dModelOut = nullptr;
    const float* dDataIn = nullptr;
    float* dDataOut = nullptr;
    std::vector<float> hModel;   // NL models are assembled on the host
    if (!is_born(q.op)) {
        hModel.resize(modelN);
        if (q.op == Op::NlFwd) {
            if (q.mem == AISO_MEM_DEVICE) HIPC(hipMemcpy(hModel.data(), q.modelIn, modelN * 4, hipMemcpyDeviceToHost));
            else std::memcpy(hModel.data(), q.modelIn, modelN * 4);
        } else if (q.add) {
            if (q.mem == AISO_MEM_DEVICE) HIPC(hipMemcpy(hModel.data(), q.modelOut, modelN * 4, hipMemcpyDeviceToHost));
            else std::memcpy(hModel.data(), q.modelOut, modelN * 4);
        } else {
            std::fill(hModel.begin(), hModel.end(), 0.f);
        }
    } else if (q.mem == AISO_MEM_DEVICE) {
        dModelIn = q.modelIn;
        dModelOut = q.modelOut;
    } else if (wemva) {
        float* mi = (float*)c->modelDev.need(mInN * 4);
        float* mo = (float*)c->imgDev.need(mOutN * 4);
        HIPC(hipMemcpyAsync(mi, q.modelIn, mInN * 4, hipMemcpyHostToDevice, s));
        if (q.add) HIPC(hipMemcpyAsync(mo, q.modelOut, mOutN * 4, hipMemcpyHostToDevice, s));
        dModelIn = mi;
        dModelOut = mo;
    } else {
        float* m = c->modelDev.need(std::max(modelN, mOutN) * 4) ? c->modelDev.as<float>() : nullptr;
        if (!is_adj(q.op) || q.add)
            HIPC(hipMemcpyAsync(m, is_adj(q.op) ? q.modelOut : q.modelIn, (is_adj(q.op) ? mOutN : modelN) * 4,
                                hipMemcpyHostToDevice, s));
        dModelIn = m;
        dModelOut = m;
    }
    if (q.mem == AISO_MEM_DEVICE) {
        dDataIn = q.dataIn;
        dDataOut = q.dataOut;
    } else {
        float* d = c->dataDev.need(dataN * 4) ? c->dataDev.as<float>() : nullptr;
        if (dataInput || q.add) HIPC(hipMemcpyAsync(d, dataInput ? q.dataIn : q.dataOut, dataN * 4, hipMemcpyHostToDevice, s));
        dDataIn = d;
        dDataOut = d;
    }

    // ---- geometry lists
    std::vector<DevGeom*> srcG(q.nShot), recG(q.nRec);
    int nSrcRegMax = 1, nRecRegMax = 1;
    for (int i = 0; i < q.nShot; ++i) {
        srcG[i] = dev_geom(c, q.src[i]);
        nSrcRegMax = std::max(nSrcRegMax, srcG[i]->nReg);
    }
    for (int i = 0; i < q.nRec; ++i) {
        recG[i] = dev_geom(c, q.rec[i]);
        nRecRegMax = std::max(nRecRegMax, recG[i]->nReg);
    }

    // ---- batch size from the HBM budget
    // The Born forward with sub == 2 runs its two propagations in lockstep:
    // the secondary source then lives in a 3-slice ring instead of nts slices.
    // A free surface runs on the generic step kernel, without the lockstep
    // pair or two steps per launch.
    const bool fsurf = p.freeSurface != 0 && !wemva;
    const bool lockstep = q.op == Op::BornFwd && p.sub == 2 && c->lockstep && !fsurf;
    // Two fine steps per launch (k_fwd_pair2) when no source or receiver
    // point lies in the FAT ring, whose values that kernel takes as 0.
    // The Born adjoint with sub == 2 likewise runs both of its passes two
    // fine steps per launch (k_fwd_pair2 as source pass, k_rec2).
    bool noRing = true;
    for (int i = 0; noRing && i < q.nShot; ++i) noRing = srcG[i]->nRing == 0;
    for (int i = 0; noRing && i < q.nRec; ++i) noRing = recG[i]->nRing == 0;
    const bool pair2 = lockstep && c->pair2 && noRing;
    const bool adj2 = q.op == Op::BornAdj && p.sub == 2 && c->pair2 && noRing && !fsurf;
    const int pairSlices = pair2 || adj2 ? 4 : 2;
    const long wSlices = lockstep ? 3 : p.nts;
    const bool needW = is_born(q.op), needSS = is_ext(q.op) || is_tomo(q.op) || wemva || q.op == Op::BornAdj;
    // The offset-extended passes keep a ring of coarse-rate slices per shot:
    // extK samples batched per imaging / secondary-source launch, plus the
    // one the time steps are writing or reading.
    const int extK = std::min(c->extBatch, p.nts);
    const int ssSlices = is_ext(q.op) || is_tomo(q.op) || wemva ? extK + 1 : 2;
    auto ssSlot = [&](int k) { return (long)(k % ssSlices) * nzx; };
    // Time lags of a batch: sample k uses slices e in [lo, hi), those whose
    // W sample k - dir * 2 (e - hExt) lies in [0, nts)
    // (F/src/BornExtShotsGpuFunctions.cu:362-366, 374-378 for dir = +1;
    // F/src/tomoExtShotsGpuFunctions.cu:2630-2633 for dir = -1).
    auto time_lags = [&](ExtBatchArgs& e, int dir) {
        e.nts = p.nts;
        e.lagDir = dir;
        e.eLo = 2 * p.hExt + 1;
        e.eHi = 0;
        for (int j = 0; j < e.nk; ++j) {
            const int k = (int)(e.wOff[j] / nzx);
            if (dir > 0) {
                e.lo[j] = std::max((k + 1 - p.nts) / 2, -p.hExt) + p.hExt;
                e.hi[j] = std::min(k / 2, p.hExt) + p.hExt + 1;
            } else {
                e.lo[j] = std::max((-k) / 2, -p.hExt) + p.hExt;
                e.hi[j] = std::min((p.nts - 1 - k) / 2, p.hExt) + p.hExt + 1;
            }
            e.eLo = std::min(e.eLo, e.lo[j]);
            e.eHi = std::max(e.eHi, e.hi[j]);
        }
        if (e.eLo > e.eHi) e.eLo = e.eHi;
    };
    const bool needImg = q.op == Op::BornAdj || q.op == Op::ExtAdj || q.op == Op::TomoAdj || wemva;
    const bool tomo = is_tomo(q.op) || wemva;   // the tomographic and WEMVA passes share the machinery below
    const int capSlices = wemva ? extK + 1 : 2;     // captured coarse samples kept per shot
    size_t perShot = pairSlices * nzx * 4 + (size_t)nRecRegMax * p.nts * 4 + (size_t)nSrcRegMax * p.ntw * 4;
    if (needW) perShot += (size_t)wSlices * nzx * 4;
    if (lockstep) perShot += pairSlices * nzx * 4;
    if (needSS) perShot += (size_t)ssSlices * nzx * 4;
    if (needImg) perShot += (size_t)nImg * nzx * 4;
    if (tomo) perShot += (size_t)(p.nts + capSlices) * nzx * 4;   // scattered / receiver wavefield + captures
    c->bytesPerShot = perShot;
    // Shots per launch: an explicit request, else as many as keep the
    // slices every fine step touches (per shot) within the on-die cache
    // budget; never more than device memory holds.
    int B = is_adj(q.op) ? c->batchAdj : c->batchFwd;
    if (B <= 0) B = c->batchReq;
    if (B <= 0) {
        int hot = 2;                                  // wavefield pair
        if (lockstep) hot = 7;                        // two pairs + secondary-source ring
        else if (q.op == Op::BornAdj) hot = 3;        // + image
        else if (q.op == Op::ExtFwd) hot = 4;         // + two secondary-source slices
        else if (q.op == Op::ExtAdj) hot = 3;         // + receiver slice (the image is touched once per extK samples)
        else if (tomo) hot = 4;                       // + secondary source and captured slices
        B = (int)std::max<long>(4, c->cacheBudget / ((long)hot * nzx * 4));
        B = std::min(B, 32);
    }
    {
        size_t freeB = 0, totalB = 0;
        HIPC(hipMemGetInfo(&freeB, &totalB));
        size_t have = c->pairs.n + c->pairsQ.n + c->W.n + c->ss.n + c->img.n + c->dreg.n + c->sig.n + c->P.n +
                      c->cap.n + c->imgDev.n;
        size_t budget = (size_t)((freeB + have) * 0.85);
        B = std::min(B, (int)std::max<size_t>(1, budget / perShot));
    }
    B = std::min(B, q.nShot);
    B = (q.nShot + (q.nShot + B - 1) / B - 1) / ((q.nShot + B - 1) / B);   // same batch count, balanced sizes
    c->batchUsed = B;

    float* pairs = c->pairs.need((size_t)B * pairSlices * nzx * 4) ? c->pairs.as<float>() : nullptr;
    float* W = needW ? (float*)c->W.need((size_t)B * wSlices * nzx * 4) : nullptr;
    float* pairsQ = lockstep ? (float*)c->pairsQ.need((size_t)B * pairSlices * nzx * 4) : nullptr;
    ShotSlot* dSlotsQ = lockstep || adj2 || tomo ? (ShotSlot*)c->slotsQ.need(sizeof(ShotSlot) * B) : nullptr;
    // Tomographic operator: P holds the scattered wavefield of the forward
    // legs or the receiver wavefield of the adjoint (slots dSlotsQ see it as
    // W); cap holds two captured coarse samples (slots dSlotsC).
    float* Pw = tomo ? (float*)c->P.need((size_t)B * p.nts * nzx * 4) : nullptr;
    float* cap = tomo ? (float*)c->cap.need((size_t)B * capSlices * nzx * 4) : nullptr;
    ShotSlot* dSlotsC = tomo ? (ShotSlot*)c->slotsC.need(sizeof(ShotSlot) * B) : nullptr;
    // WEMVA forward's batched offset imaging: slots that see (ss: captures,
    // W: W'') and (ss: the receiver wavefield, W: captures)
    ShotSlot* dSlotsI = q.op == Op::WemvaFwd ? (ShotSlot*)c->slotsI.need(sizeof(ShotSlot) * B) : nullptr;
    ShotSlot* dSlotsJ = q.op == Op::WemvaFwd ? (ShotSlot*)c->slotsJ.need(sizeof(ShotSlot) * B) : nullptr;
    float* ss = needSS ? (float*)c->ss.need((size_t)B * ssSlices * nzx * 4) : nullptr;
    float* img = needImg ? (float*)c->img.need((size_t)B * nImg * nzx * 4) : nullptr;
    float* dreg = (float*)c->dreg.need((size_t)B * nRecRegMax * p.nts * 4);
    float* sig = (float*)c->sig.need((size_t)B * nSrcRegMax * p.ntw * 4);
    ShotSlot* dSlots = (ShotSlot*)c->slots.need(sizeof(ShotSlot) * B);

    // ---- per-call model preparation
    const float* mtil = nullptr;
    float* acc = nullptr;
    if (q.op == Op::BornFwd) {
        float* m = (float*)c->mtil.need(nzx * 4);
        HIPC(launch_scale_model(dModelIn, m, c->v2dt2.as<float>(), c->refl.as<float>(), p.nz, p.nx, s));
        mtil = m;
    }
    // Per-point scale of the extended model: 2/v^3 for offsets; for time
    // lags also v^2 dtw^2, or -v^2 dtw^2 with sloth (which neither the
    // reference's free-surface routines nor its WEMVA take).
    const bool tlag = (is_ext(q.op) || tomo) && p.extension == AISO_EXT_TIME;
    const float* extScale = c->refl.as<float>();
    if (tlag) {
        float* t = (float*)c->tscale.need(nzx * 4);
        HIPC(launch_time_scale(c->v2dt2.as<float>(), c->refl.as<float>(), c->sloth && !fsurf && !wemva, t, p.nz,
                               p.nx, s));
        extScale = t;
    }
    if (q.op == Op::ExtFwd) {
        float* m = (float*)c->mtil.need(nzx * nExt * 4);
        HIPC(launch_scale_ext(dModelIn, m, extScale, p.nz, p.nx, nExt, s));
        mtil = m;
    }
    const float* rExt = nullptr;   // extended reflectivity (tomographic) or image (WEMVA adjoint) scaled by 2/v^3
    if (is_tomo(q.op)) {
        if (q.op == Op::TomoFwd) {
            float* m = (float*)c->mtil.need(nzx * 4);
            HIPC(launch_scale_model(dModelIn, m, c->v2dt2.as<float>(), c->refl.as<float>(), p.nz, p.nx, s));
            mtil = m;
        }
        float* r = (float*)c->rExt.need(nzx * p.nExt * 4);
        HIPC(launch_scale_ext(c->extRefl.as<float>(), r, extScale, p.nz, p.nx, p.nExt, s));
        rExt = r;
    } else if (q.op == Op::WemvaFwd) {
        float* m = (float*)c->mtil.need(nzx * 4);
        HIPC(launch_scale_model(dModelIn, m, c->v2dt2.as<float>(), c->refl.as<float>(), p.nz, p.nx, s));
        mtil = m;
    } else if (q.op == Op::WemvaAdj) {   // the extended image in, scaled like a model
        float* r = (float*)c->rExt.need(nzx * p.nExt * 4);
        HIPC(launch_scale_ext(dModelIn, r, extScale, p.nz, p.nx, p.nExt, s));
        rExt = r;
    }
    if (needImg && !perShotImg) {
        acc = (float*)c->acc.need(nzx * nImg * 4);
        HIPC(launch_fill(acc, nzx * nImg, s));
    }
    std::vector<float> nlAcc;
    if (q.op == Op::NlAdj && q.nShotSig == 1) nlAcc.assign((size_t)nSrcDev * p.nts, 0.f);

    StepArgs base{};
    base.slots = dSlots;
    base.v2dt2 = c->v2dt2.as<float>();
    base.cd = c->cd.as<float>();
    base.mtil = mtil;
    base.nz = p.nz; base.nx = p.nx; base.minPad = p.minPad; base.nts = p.nts; base.ntw = p.ntw;
    base.tilesZ = tilesZ(c);
    base.ntiles = base.tilesZ * tilesX(c);
    for (int k = 0; k < 6; ++k) { base.zc[k] = p.zc[k]; base.xc[k] = p.xc[k]; }
    base.wA = base.wB = base.wImg = 0;
    base.ntW = c->streamNT;
    base.fs = fsurf;

    std::vector<float> sigHost;
    for (int b0 = 0; b0 < q.nShot; b0 += B) {
        const int nb = std::min(B, q.nShot - b0);
        hipEvent_t batch0 = nullptr;
        if (c->profile) {
            batch0 = pooled_event(c);
            HIPC(hipEventRecord(batch0, s));
        }
        // -- slot descriptors and source traces
        std::vector<ShotSlot> slots(nb);
        int maxSrcRing = 0, maxRecRing = 0;
        sigHost.assign((size_t)nb * nSrcRegMax * p.ntw, 0.f);
        for (int k = 0; k < nb; ++k) {
            const int shot = b0 + k;
            const aiso_device* sd = q.src[shot];
            DevGeom* sg = srcG[shot];
            DevGeom* rg = recG[q.nRec == 1 ? 0 : shot];
            ShotSlot& S = slots[k];
            std::memset(&S, 0, sizeof(S));
            for (int i = 0; i < pairSlices; ++i) S.p[i] = pairs + ((size_t)k * pairSlices + i) * nzx;
            if (W) S.W = W + (size_t)k * wSlices * nzx;
            if (pairsQ)
                for (int i = 0; i < pairSlices; ++i) S.q[i] = pairsQ + ((size_t)k * pairSlices + i) * nzx;
            if (ss) { S.ss[0] = ss + (size_t)k * ssSlices * nzx; S.ss[1] = S.ss[0] + nzx; }
            if (img) S.img = img + (size_t)k * nImg * nzx;
            S.dreg = dreg + (size_t)k * nRecRegMax * p.nts;
            S.sig = sig + (size_t)k * nSrcRegMax * p.ntw;
            S.srcTile = sg->tilePtr.as<int>();
            S.srcEnt = sg->ent.as<int2>();
            S.srcRing = sg->ring.as<int2>();
            S.nSrcRing = sg->nRing;
            S.recTile = rg->tilePtr.as<int>();
            S.recEnt = rg->ent.as<int2>();
            S.recRing = rg->ring.as<int2>();
            S.nRecRing = rg->nRing;
            maxSrcRing = std::max(maxSrcRing, sg->nRing);
            maxRecRing = std::max(maxRecRing, rg->nRing);
            std::vector<float> fine;
            if (is_born(q.op)) {
                const float* wav = q.wavelets + (size_t)(q.nWav == 1 ? 0 : shot) * nSrcDev * p.nts;
                fine = born_source(c, sd->g, wav);
            } else if (q.op == Op::NlFwd) {
                fine = nl_source(c, sd->g, hModel.data() + (size_t)(q.nShotSig == 1 ? 0 : shot) * nSrcDev * p.nts);
            }
            if (!fine.empty()) std::memcpy(sigHost.data() + (size_t)k * nSrcRegMax * p.ntw, fine.data(), fine.size() * 4);
        }
        HIPC(hipMemcpyAsync(dSlots, slots.data(), sizeof(ShotSlot) * nb, hipMemcpyHostToDevice, s));
        std::vector<ShotSlot> slotsP, slotsCv;
        std::vector<ShotSlot> slotsIv, slotsJv;
        if (tomo) {   // the same shots seeing P, or the captured samples, as W
            slotsP = slots;
            slotsCv = slots;
            for (int k = 0; k < nb; ++k) {
                slotsP[k].W = Pw + (size_t)k * p.nts * nzx;
                slotsCv[k].W = cap + (size_t)k * capSlices * nzx;
            }
            HIPC(hipMemcpyAsync(dSlotsQ, slotsP.data(), sizeof(ShotSlot) * nb, hipMemcpyHostToDevice, s));
            HIPC(hipMemcpyAsync(dSlotsC, slotsCv.data(), sizeof(ShotSlot) * nb, hipMemcpyHostToDevice, s));
        }
        if (q.op == Op::WemvaFwd) {
            // offsets: I = (captures, W''), J = (R, captures); time lags:
            // J = (captures, R), the lag-shifted field being the W side
            slotsIv = slots;
            slotsJv = tlag ? slotsP : slotsCv;
            for (int k = 0; k < nb; ++k) {
                slotsIv[k].ss[0] = cap + (size_t)k * capSlices * nzx;
                slotsJv[k].ss[0] = tlag ? cap + (size_t)k * capSlices * nzx : Pw + (size_t)k * p.nts * nzx;
            }
            HIPC(hipMemcpyAsync(dSlotsI, slotsIv.data(), sizeof(ShotSlot) * nb, hipMemcpyHostToDevice, s));
            HIPC(hipMemcpyAsync(dSlotsJ, slotsJv.data(), sizeof(ShotSlot) * nb, hipMemcpyHostToDevice, s));
        }
        std::vector<ShotSlot> slotsQ;
        if (lockstep) {   // the scattered field seen as "the" pair, for single-field launches
            slotsQ = slots;
            for (auto& S : slotsQ) { S.p[0] = S.q[0]; S.p[1] = S.q[1]; }
            HIPC(hipMemcpyAsync(dSlotsQ, slotsQ.data(), sizeof(ShotSlot) * nb, hipMemcpyHostToDevice, s));
        }
        if (q.op != Op::NlAdj) HIPC(hipMemcpyAsync(sig, sigHost.data(), sigHost.size() * 4, hipMemcpyHostToDevice, s));
        else HIPC(launch_fill(sig, (long)nb * nSrcRegMax * p.ntw, s));

        StepArgs a = base;
        a.nshots = nb;
        a.gridLogical = a.ntiles * nb;

        // -- adjoint (and WEMVA): data traces to the regular receiver grid
        if (dataInput) {
            for (int k = 0; k < nb; ++k) {
                const int shot = b0 + k;
                const aiso_device* rd = q.rec[q.nRec == 1 ? 0 : shot];
                DevGeom* rg = recG[q.nRec == 1 ? 0 : shot];
                HIPC(launch_dev_adjoint(dreg + (size_t)k * nRecRegMax * p.nts, dDataIn + (size_t)shot * nRecDev * p.nts,
                                        rg->adjPtr.as<int>(), rg->adjEnt.as<int>(), rg->weight.as<float>(),
                                        rg->nReg, rd->g.nb, p.nts, s));
            }
        } else {
            HIPC(launch_fill(dreg, (long)nb * nRecRegMax * p.nts, s));
        }

        auto ring = [&](unsigned mode, const StepArgs& sa) {   // true if it launched anything
            int mr = 0;
            if (mode & (M_INJ_SIG | M_REC_SIG)) mr = std::max(mr, maxSrcRing);
            if (mode & (M_INJ_DATA | M_REC_DATA)) mr = std::max(mr, maxRecRing);
            if (mr <= 0) return false;
            for (int l = 0; l < c->curLanes; ++l) {
                int k0, cnt;
                lane_range(sa.nshots, c->curLanes, l, k0, cnt);
                if (cnt > 0) HIPC(launch_ring(mode, c->curLanes > 1 ? lane_args(sa, k0, cnt) : sa, mr, lane_stream(c, l)));
            }
            return true;
        };
        c->curLanes = is_ext(q.op) || tomo || pair2 || adj2 ? 1 : std::max(1, std::min(c->lanes, nb));
        if (c->curLanes > 1) ensure_lanes(c);

        // Three consecutive fine slices hold everything a coarse sample needs
        // when sub == 2, so captures and imaging are written once per coarse
        // step instead of accumulated at every fine step.
        const bool win = p.sub == 2;
        // The extended-Born operators step on the unspecialized kernel: the
        // specialized one ended a GPU run in a memory fault on this path with
        // no cause found in its source (it is exact and in bounds when the
        // same source runs on the host).
        const bool generic = is_ext(q.op) || tomo || fsurf;

        if (pair2) {
            // The scattered field lags the source field by one launch (two
            // fine steps); SS[k] lives in a 3-slice ring of W.
            std::vector<ShotSlot> s2 = slots;
            for (int k = 0; k < nb; ++k) {
                const int shot = b0 + k;
                DevGeom* sgE = dev_geom_pair2(c, q.src[shot], true);
                DevGeom* rgT = dev_geom_pair2(c, q.rec[q.nRec == 1 ? 0 : shot], false);
                s2[k].srcTile = sgE->tilePtr.as<int>();
                s2[k].srcEnt = sgE->ent.as<int2>();
                s2[k].recTile = rgT->tilePtr.as<int>();
                s2[k].recEnt = rgT->ent.as<int2>();
            }
            HIPC(hipMemcpyAsync(dSlotsQ, s2.data(), sizeof(ShotSlot) * nb, hipMemcpyHostToDevice, s));
            HIPC(launch_fill(pairs, (long)nb * 4 * nzx, s));
            HIPC(launch_fill(pairsQ, (long)nb * 4 * nzx, s));
            auto ring3 = [&](long k) { return (k % 3) * nzx; };
            Pair2Args a2{};
            a2.slots = dSlotsQ;
            a2.v2dt2 = base.v2dt2;
            a2.cd = base.cd;
            a2.mtil = mtil;
            a2.nz = p.nz; a2.nx = p.nx; a2.minPad = p.minPad; a2.nts = p.nts; a2.ntw = p.ntw;
            a2.tilesZ = tilesZ2(c);
            a2.ntiles = a2.tilesZ * tilesX2(c);
            a2.gridLogical = a2.ntiles * nb;
            for (int k = 0; k < 6; ++k) { a2.zc[k] = p.zc[k]; a2.xc[k] = p.xc[k]; }
            a2.wa = f[h + 1]; a2.wb = f[h + 2]; a2.wc = f[1];
            a2.ta = f[h + 1]; a2.tb = f[h + 2];
            a2.faA = f[0]; a2.faB = f[1]; a2.fbB = f[h + 1];
            a2.feA = f[1]; a2.ffA = f[h + 1]; a2.feB = f[2]; a2.ffB = f[h + 2];
            a2.ntW = c->streamNT;
            a2.ssTail = ring3(p.nts - 1);
            int rs[4] = {0, 1, 2, 3}, rq[4] = {0, 1, 2, 3};   // current, old, new A, new B
            const int M = (p.ntw - 1) / 2;
            for (int m = 0; m <= M; ++m) {
                a2.doS = m < M;
                a2.doQ = m >= 1;
                a2.tailS = m == M - 1;
                a2.sC = rs[0]; a2.sO = rs[1]; a2.sA = rs[2]; a2.sB = rs[3];
                a2.qC = rq[0]; a2.qO = rq[1]; a2.qA = rq[2]; a2.qB = rq[3];
                a2.injA = 2 * m;
                a2.injB = 2 * m + 1;
                a2.ssNew = ring3(m);
                a2.ssPrev = m >= 1 ? ring3(m - 1) : 0;
                a2.its = m - 1;
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (sample_event(c, AISO_K_BORN_PAIR)) {
                    e0 = pooled_event(c);
                    e1 = pooled_event(c);
                    c->ev.push_back({e0, e1, AISO_K_BORN_PAIR, 2.0 * (a2.doS + a2.doQ) * double(p.interior()) * nb});
                }
                HIPC(launch_fwd_pair2(a2, s, e0, e1));
                if (a2.doS) { int t[4] = {rs[3], rs[2], rs[0], rs[1]}; std::copy(t, t + 4, rs); }
                if (a2.doQ) { int t[4] = {rq[3], rq[2], rq[0], rq[1]}; std::copy(t, t + 4, rq); }
            }
            for (int k = 0; k < nb; ++k) {
                const int shot = b0 + k;
                const aiso_device* rd = q.rec[q.nRec == 1 ? 0 : shot];
                DevGeom* rg = recG[q.nRec == 1 ? 0 : shot];
                HIPC(launch_dev_forward(dreg + (size_t)k * nRecRegMax * p.nts, dDataOut + (size_t)shot * nRecDev * p.nts,
                                        rg->cornerReg.as<int>(), rg->weight.as<float>(), nRecDev, rd->g.nb, p.nts,
                                        q.add, s));
            }
        } else if (lockstep) {
            HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
            HIPC(launch_fill(pairsQ, (long)nb * 2 * nzx, s));
            const int T = p.ntw - 1;   // fine steps of each propagation
            auto ring3 = [&](long k) { return (k % 3) * nzx; };
            int curS = 1, curQ = 1;
            a.winOut = WIN_SS;
            a.wa = f[h + 1];
            a.wb = f[h + 2];
            a.wc = f[1];
            fork_lanes(c);
            // With sampled profiling on one lane, the pair kernel is timed
            // over windows of profEvery consecutive launches between two
            // plain events (the average time per launch, back to back); a
            // window that another kernel entered is dropped.
            struct { hipEvent_t e0 = nullptr; int n = 0; bool spoiled = false; } pw;
            const bool windows = c->profile && c->profEvery > 1 && c->curLanes == 1;
            auto close_window = [&] {
                if (!pw.e0) return;
                hipEvent_t e1 = pooled_event(c);
                HIPC(hipEventRecord(e1, s));
                if (!pw.spoiled && pw.n > 0) {
                    c->ev.push_back({pw.e0, e1, AISO_K_BORN_PAIR, 2.0 * double(p.interior()) * nb * pw.n, pw.n});
                } else {
                    c->evPool.push_back(pw.e0);
                    c->evPool.push_back(e1);
                }
                pw.e0 = nullptr;
                pw.n = 0;
                pw.spoiled = false;
            };
            for (int t = 1; t <= T + 2; ++t) {
                const bool doS = t <= T, doQ = t - 2 >= 1;
                const int itsS = (t - 1) / p.sub, it2S = t - itsS * p.sub;
                const int tq = t - 2, itsQ = doQ ? (tq - 1) / p.sub : 0, it2Q = doQ ? tq - itsQ * p.sub : 1;
                StepArgs x = a;
                x.cur = curS;
                x.curQ = curQ;
                x.itwInj = t - 1;
                x.win = doS && it2S == 1;
                x.wA = ring3(itsS);
                x.its = itsQ;
                x.fa = f[it2Q - 1];
                x.fb = f[h + it2Q - 1];
                x.wQa = ring3(itsQ);
                x.wQb = ring3(itsQ + 1);
                x.fe = f[it2Q];
                x.ff = f[h + it2Q];
                if (doS && doQ) {
                    if (windows) {
                        if (!pw.e0) {
                            pw.e0 = pooled_event(c);
                            HIPC(hipEventRecord(pw.e0, s));
                        }
                        HIPC(launch_fwd_pair(c->tz, c->tx, x, s, nullptr, nullptr));
                        ++pw.n;
                    } else {
                        for (int l = 0; l < c->curLanes; ++l) {
                            int k0, cnt;
                            lane_range(nb, c->curLanes, l, k0, cnt);
                            if (cnt == 0) continue;
                            hipEvent_t e0 = nullptr, e1 = nullptr;
                            if (sample_event(c, AISO_K_BORN_PAIR)) {
                                e0 = pooled_event(c);
                                e1 = pooled_event(c);
                                c->ev.push_back({e0, e1, AISO_K_BORN_PAIR, 2.0 * double(p.interior()) * cnt});
                            }
                            HIPC(launch_fwd_pair(c->tz, c->tx, c->curLanes > 1 ? lane_args(x, k0, cnt) : x,
                                                 lane_stream(c, l), e0, e1));
                        }
                    }
                    bool other = ring(MODE_SRC_WIN, x);
                    StepArgs xq = x;
                    xq.slots = dSlotsQ;
                    xq.cur = curQ;
                    other = ring(MODE_BORN_SCAT_PRE, xq) || other;
                    if (windows) {
                        pw.spoiled = pw.spoiled || other;
                        if (pw.n == c->profEvery) close_window();
                    }
                } else if (doS) {
                    if (windows) close_window();
                    launch_timed(c, MODE_SRC_WIN, AISO_K_SRC, x, nb);
                    ring(MODE_SRC_WIN, x);
                } else {
                    if (windows) close_window();
                    StepArgs xq = x;
                    xq.slots = dSlotsQ;
                    xq.cur = curQ;
                    xq.wA = x.wQa;
                    xq.wB = x.wQb;
                    launch_timed(c, MODE_BORN_SCAT_PRE, AISO_K_BORN_SCAT, xq, nb);
                    ring(MODE_BORN_SCAT_PRE, xq);
                }
                if (doS) curS ^= 1;
                if (doQ) curQ ^= 1;
                if (t == T) {   // last coarse sample of the source wavefield
                    if (windows) close_window();
                    for (int l = 0; l < c->curLanes; ++l) {
                        int k0, cnt;
                        lane_range(nb, c->curLanes, l, k0, cnt);
                        if (cnt > 0)
                            HIPC(launch_win_tail(dSlots + k0, cnt, curS, WIN_SS, f[h + 1], f[h + 2], ring3(p.nts - 1),
                                                 0, 0, mtil, p.nz, p.nx, lane_stream(c, l)));
                    }
                }
            }
            if (windows) close_window();
            join_lanes(c);
            for (int k = 0; k < nb; ++k) {
                const int shot = b0 + k;
                const aiso_device* rd = q.rec[q.nRec == 1 ? 0 : shot];
                DevGeom* rg = recG[q.nRec == 1 ? 0 : shot];
                HIPC(launch_dev_forward(dreg + (size_t)k * nRecRegMax * p.nts, dDataOut + (size_t)shot * nRecDev * p.nts,
                                        rg->cornerReg.as<int>(), rg->weight.as<float>(), nRecDev, rd->g.nb, p.nts,
                                        q.add, s));
            }
        }

        if (adj2) {
            // Source pass (W[k] for every coarse k) then receiver pass with
            // imaging, two fine steps per launch each.
            std::vector<ShotSlot> s2 = slots;
            for (int k = 0; k < nb; ++k) {
                const int shot = b0 + k;
                DevGeom* sgE = dev_geom_pair2(c, q.src[shot], true);
                DevGeom* rgE = dev_geom_pair2(c, q.rec[q.nRec == 1 ? 0 : shot], true);
                s2[k].srcTile = sgE->tilePtr.as<int>();
                s2[k].srcEnt = sgE->ent.as<int2>();
                s2[k].recTile = rgE->tilePtr.as<int>();
                s2[k].recEnt = rgE->ent.as<int2>();
            }
            HIPC(hipMemcpyAsync(dSlotsQ, s2.data(), sizeof(ShotSlot) * nb, hipMemcpyHostToDevice, s));
            HIPC(launch_fill(pairs, (long)nb * 4 * nzx, s));
            Pair2Args a2{};
            a2.slots = dSlotsQ;
            a2.v2dt2 = base.v2dt2;
            a2.cd = base.cd;
            a2.mtil = nullptr;                            // raw W
            a2.nz = p.nz; a2.nx = p.nx; a2.minPad = p.minPad; a2.nts = p.nts; a2.ntw = p.ntw;
            a2.tilesZ = tilesZ2(c);
            a2.ntiles = a2.tilesZ * tilesX2(c);
            a2.gridLogical = a2.ntiles * nb;
            for (int k = 0; k < 6; ++k) { a2.zc[k] = p.zc[k]; a2.xc[k] = p.xc[k]; }
            a2.wa = f[h + 1]; a2.wb = f[h + 2]; a2.wc = f[1];
            a2.ta = f[h + 1]; a2.tb = f[h + 2];
            a2.ntW = c->streamNT;
            a2.ssTail = (long)(p.nts - 1) * nzx;
            a2.doS = 1;
            int rs[4] = {0, 1, 2, 3};
            const int M = (p.ntw - 1) / 2;
            for (int m = 0; m < M; ++m) {
                a2.tailS = m == M - 1;
                a2.sC = rs[0]; a2.sO = rs[1]; a2.sA = rs[2]; a2.sB = rs[3];
                a2.injA = 2 * m;
                a2.injB = 2 * m + 1;
                a2.ssNew = (long)m * nzx;
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (sample_event(c, AISO_K_SRC)) {
                    e0 = pooled_event(c);
                    e1 = pooled_event(c);
                    c->ev.push_back({e0, e1, AISO_K_SRC, 2.0 * double(p.interior()) * nb});
                }
                HIPC(launch_fwd_pair2(a2, s, e0, e1));
                int t[4] = {rs[3], rs[2], rs[0], rs[1]};
                std::copy(t, t + 4, rs);
            }
            HIPC(launch_fill(pairs, (long)nb * 4 * nzx, s));
            HIPC(launch_fill(img, (long)nb * nExt * nzx, s));
            Rec2Args r2{};
            r2.slots = dSlotsQ;
            r2.v2dt2 = base.v2dt2;
            r2.cd = base.cd;
            r2.nz = p.nz; r2.nx = p.nx; r2.minPad = p.minPad; r2.nts = p.nts;
            r2.tilesZ = a2.tilesZ;
            r2.ntiles = a2.ntiles;
            r2.gridLogical = a2.gridLogical;
            for (int k = 0; k < 6; ++k) { r2.zc[k] = p.zc[k]; r2.xc[k] = p.xc[k]; }
            r2.fcA = f[2]; r2.fdA = f[h + 2];               // it2 = 1
            r2.fcB = f[1]; r2.fdB = f[h + 1];               // it2 = 0
            r2.wa = f[1]; r2.wb = f[0]; r2.wc = f[h + 1];
            r2.ta = f[1]; r2.tb = f[0];
            r2.ntW = c->streamNT;
            int rr[4] = {0, 1, 2, 3};
            for (int its = p.nts - 2; its >= 0; --its) {
                r2.sC = rr[0]; r2.sO = rr[1]; r2.sA = rr[2]; r2.sB = rr[3];
                r2.its = its;
                r2.wImg = (long)(its + 1) * nzx;
                r2.tail = its == 0;
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (sample_event(c, AISO_K_BORN_REC)) {
                    e0 = pooled_event(c);
                    e1 = pooled_event(c);
                    c->ev.push_back({e0, e1, AISO_K_BORN_REC, 2.0 * double(p.interior()) * nb});
                }
                HIPC(launch_rec2(r2, s, e0, e1));
                int t[4] = {rr[3], rr[2], rr[0], rr[1]};
                std::copy(t, t + 4, rr);
            }
            HIPC(launch_born_adj_finish(dSlots, nb, -1, base.v2dt2, c->refl.as<float>(), p.nz, p.nx, s));
        }

        // -- source wavefield pass (all Born-type operators)
        if (is_born(q.op) && !lockstep && !adj2) {
            HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
            a.cur = 1;
            const unsigned mode = win ? MODE_SRC_WIN : MODE_SRC;
            const int winOut = q.op == Op::BornFwd ? WIN_SS : WIN_W;
            a.winOut = winOut;
            a.wa = f[h + 1];
            a.wb = f[h + 2 < 2 * h ? h + 2 : h + 1];
            a.wc = f[1];
            fork_lanes(c);
            for (int its = 0; its < p.nts - 1; ++its)
                for (int it2 = 1; it2 <= p.sub; ++it2) {
                    a.its = its;
                    a.itwInj = its * p.sub + it2 - 1;
                    a.wA = (long)its * nzx;
                    a.wB = (long)(its + 1) * nzx;
                    a.fe = f[it2];
                    a.ff = f[h + it2];
                    a.skipA = a.fe == 0.f;
                    a.skipB = a.ff == 0.f;
                    a.assignA = its == 0 && it2 == 1;
                    a.assignB = it2 == 1;
                    a.win = it2 == 1;
                    launch_timed(c, mode, AISO_K_SRC, a, nb, generic);
                    ring(mode, a);
                    a.cur ^= 1;
                }
            join_lanes(c);
            if (win)   // W[nts-1] = p(last-1)*f[h+1] + p(last)*f[h+2]
                HIPC(launch_win_tail(dSlots, nb, a.cur, winOut, f[h + 1], f[h + 2], (long)(p.nts - 1) * nzx, 0, 0,
                                     mtil, p.nz, p.nx, s));
            a.win = 0;
        }

        // -- tomographic and WEMVA operators (extension=offset): both legs,
        // forward or adjoint, after the source pass above
        // (F/src/tomoExtShotsGpuFunctions.cu:467-572, 912-1020;
        // F/src/wemvaExtShotsGpuFunctions.cu:458-574, 689-808)
        if (tomo) {
            const bool haveAll = ss && Pw && cap && dSlotsQ && dSlotsC && (!needImg || img) &&
                                 (q.op != Op::WemvaFwd || (dSlotsI && dSlotsJ)) &&
                                 ((q.op != Op::TomoFwd && q.op != Op::WemvaFwd) || mtil) &&
                                 ((!is_tomo(q.op) && q.op != Op::WemvaAdj) || rExt);
            if (!haveAll) throw Error(AISO_ERR_STATE, "internal: a work buffer of the tomographic / WEMVA passes is missing");
            // W <- its coarse second time derivative, in place
            const float d2 = p.dts * p.dts;
            HIPC(launch_dt2_coarse(dSlots, nb, p.nts, (float)(1.0 / (double)d2), (float)(-2.0 / (double)d2), p.nz,
                                   p.nx, s));
            enum { SS_EXT, SS_MUL, SS_ADJ };   // ring producers
            auto produce = [&](int kind, const ShotSlot* from, const float* scale, int k0, int n) {
                ExtBatchArgs e{};
                e.slots = from;
                e.v2dt2 = base.v2dt2;
                e.mext = scale;
                e.nz = p.nz; e.nx = p.nx; e.hExt = p.hExt;
                e.nk = n;
                for (int j = 0; j < n; ++j) {
                    e.ssOff[j] = ssSlot(k0 + j);
                    e.wOff[j] = (long)(k0 + j) * nzx;
                }
                if (kind == SS_MUL) {
                    HIPC(launch_mul_batch(e, nb, s));
                } else if (tlag) {   // imagingTimeFwdGpu / imagingTimeTomoAdjGpu
                    time_lags(e, kind == SS_EXT ? 1 : -1);
                    HIPC(launch_time_ss_batch(e, nb, s));
                } else if (kind == SS_EXT) {
                    HIPC(launch_ext_ss_batch(e, nb, s));
                } else {
                    HIPC(launch_tomo_adj_ss(e, nb, s));
                }
            };
            auto capOff = [&](int k) { return (long)(k % capSlices) * nzx; };
            // Forward in time.  ssKind >= 0: inject ring samples made from
            // `from`'s W (scaled by `scale`); capture: 1 = into the launch
            // slots' W at the sample index (nts slices), 2 = into the capture
            // ring; done(k) runs once coarse sample k is complete.
            auto forward_pass = [&](unsigned mode, const ShotSlot* on, int ssKind, const ShotSlot* from,
                                    const float* scale, int capture, const std::function<void(int)>& done) {
                HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
                StepArgs x = a;
                x.slots = on;
                x.cur = 1;
                x.win = 0;
                int next = 0;
                auto off = [&](int k) { return capture == 1 ? (long)k * nzx : capOff(k); };
                for (int its = 0; its < p.nts - 1; ++its) {
                    while (ssKind >= 0 && next <= its + 1) {
                        const int n = std::min(extK, p.nts - next);
                        produce(ssKind, from, scale, next, n);
                        next += n;
                    }
                    for (int it2 = 1; it2 <= p.sub; ++it2) {
                        x.its = its;
                        x.fa = f[it2 - 1];
                        x.fb = f[h + it2 - 1];
                        x.fe = f[it2];
                        x.ff = f[h + it2];
                        if (ssKind >= 0) {
                            x.ssA = x.ssB = 0;
                            x.ssOffA = ssSlot(its);
                            x.ssOffB = ssSlot(its + 1);
                        }
                        if (capture) {
                            x.wA = off(its);
                            x.wB = off(its + 1);
                            x.skipA = x.fe == 0.f;
                            x.skipB = x.ff == 0.f;
                            x.assignA = its == 0 && it2 == 1;
                            x.assignB = it2 == 1;
                        } else {   // MODE_BORN_SCAT reads W[its], W[its + 1]
                            x.wA = (long)its * nzx;
                            x.wB = (long)(its + 1) * nzx;
                            x.skipA = x.skipB = x.assignA = x.assignB = 0;
                        }
                        launch_timed(c, mode, AISO_K_TOMO, x, nb, true);
                        ring(mode, x);
                        x.cur ^= 1;
                    }
                    if (done) done(its);
                }
                if (done) done(p.nts - 1);
            };
            // Backward in time on the adjoint stepper from ring samples made
            // from `from`'s W, injected with the weights of fine index it2 + 1,
            // captured into the ring; done(k) once coarse sample k is complete.
            auto backward_pass = [&](int ssKind, const ShotSlot* from, const float* scale,
                                     const std::function<void(int)>& done) {
                HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
                HIPC(launch_fill(cap, (long)nb * capSlices * nzx, s));
                StepArgs y = a;
                y.slots = dSlotsC;
                y.cur = 1;
                y.win = 0;
                int next = p.nts - 1;   // highest sample not yet in the ring
                for (int its = p.nts - 2; its >= 0; --its) {
                    while (next >= its) {
                        const int n = std::min(extK, next + 1);
                        produce(ssKind, from, scale, next - n + 1, n);
                        next -= n;
                    }
                    for (int it2 = p.sub - 1; it2 >= 0; --it2) {
                        y.its = its;
                        y.fa = f[it2 + 1];
                        y.fb = f[h + it2 + 1];
                        y.ssA = y.ssB = 0;
                        y.ssOffA = ssSlot(its);
                        y.ssOffB = ssSlot(its + 1);
                        y.wA = capOff(its);
                        y.wB = capOff(its + 1);
                        y.fe = f[it2];
                        y.ff = f[h + it2];
                        y.skipA = y.fe == 0.f;
                        y.skipB = y.ff == 0.f;
                        y.assignA = it2 == p.sub - 1;
                        y.assignB = 0;
                        launch_timed(c, MODE_TOMO_ADJ_CAP, AISO_K_TOMO, y, nb, true);
                        y.cur ^= 1;
                    }
                    done(its + 1);
                }
                done(0);
            };
            // img += X[k] * captured sample k, one coarse sample at a time
            auto image_with = [&](const ShotSlot* X) {
                return [&, X](int k) { HIPC(launch_img_acc(X, (long)k * nzx, dSlotsC, capOff(k), nb, p.nz, p.nx, s)); };
            };
            if (dataInput) {   // receiver wavefield R into P (computeTomoRecWfld, :1481-1515)
                HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
                StepArgs x = a;
                x.slots = dSlotsQ;
                x.cur = 1;
                x.win = 0;
                for (int its = p.nts - 2; its >= 0; --its)
                    for (int it2 = p.sub - 1; it2 >= 0; --it2) {
                        x.its = its;
                        x.itwRec = its * p.sub + it2;
                        x.fc = f[it2 + 1];
                        x.fd = f[h + it2 + 1];
                        x.wA = (long)its * nzx;
                        x.wB = (long)(its + 1) * nzx;
                        x.fe = f[it2];
                        x.ff = f[h + it2];
                        x.skipA = x.fe == 0.f;
                        x.skipB = x.ff == 0.f;
                        x.assignA = it2 == p.sub - 1;
                        x.assignB = its == p.nts - 2 && it2 == p.sub - 1;
                        launch_timed(c, MODE_TOMO_REC, AISO_K_TOMO, x, nb, true);
                        ring(MODE_TOMO_REC, x);
                        x.cur ^= 1;
                    }
            }
            if (img) HIPC(launch_fill(img, (long)nb * nImg * nzx, s));
            if (q.op == Op::TomoFwd) {
                if (c->leg1) {   // source -> extended reflectivity -> P -> model -> data
                    forward_pass(MODE_TOMO_CAP, dSlotsQ, SS_EXT, dSlots, rExt, 1, nullptr);
                    forward_pass(MODE_BORN_SCAT, dSlotsQ, -1, nullptr, nullptr, 0, nullptr);
                }
                if (c->leg2) {   // source -> model -> P -> extended reflectivity -> data
                    forward_pass(MODE_TOMO_CAP, dSlotsQ, SS_MUL, dSlots, mtil, 1, nullptr);
                    forward_pass(MODE_EXT_SCAT, dSlotsQ, SS_EXT, dSlotsQ, rExt, 0, nullptr);
                }
                for (int k = 0; k < nb; ++k) {
                    const int shot = b0 + k;
                    const aiso_device* rd = q.rec[q.nRec == 1 ? 0 : shot];
                    DevGeom* rg = recG[q.nRec == 1 ? 0 : shot];
                    HIPC(launch_dev_forward(dreg + (size_t)k * nRecRegMax * p.nts, dDataOut + (size_t)shot * nRecDev * p.nts,
                                            rg->cornerReg.as<int>(), rg->weight.as<float>(), nRecDev, rd->g.nb, p.nts,
                                            q.add, s));
                }
            } else if (q.op == Op::TomoAdj || q.op == Op::WemvaAdj) {
                // The same two passes: forward from the extended scattering of
                // W'' imaged against R, and backward from the adjoint extended
                // scattering of R imaged against W''; the tomographic adjoint
                // runs them as legs 1, 2, WEMVA's as legs 2, 1.
                auto fwdLeg = [&] { forward_pass(MODE_TOMO_CAP, dSlotsC, SS_EXT, dSlots, rExt, 2, image_with(dSlotsQ)); };
                auto bwdLeg = [&] { backward_pass(SS_ADJ, dSlotsQ, rExt, image_with(dSlots)); };
                if (q.op == Op::TomoAdj) {
                    if (c->leg1) fwdLeg();
                    if (c->leg2) bwdLeg();
                } else {
                    if (c->leg1) bwdLeg();
                    if (c->leg2) fwdLeg();
                }
                HIPC(launch_born_adj_finish(dSlots, nb, -1, base.v2dt2, c->refl.as<float>(), p.nz, p.nx, s));
            } else {   // WEMVA forward: velocity perturbation -> extended image
                std::vector<int> pend;
                auto flush = [&](bool wemvaFwd) {
                    if (pend.empty()) return;
                    ExtBatchArgs e{};
                    e.slots = wemvaFwd ? dSlotsJ : dSlotsI;
                    e.v2dt2 = base.v2dt2;
                    e.nz = p.nz; e.nx = p.nx; e.hExt = p.hExt;
                    e.nk = (int)pend.size();
                    for (int j = 0; j < e.nk; ++j) {
                        const int k = pend[j];
                        const bool rSide = wemvaFwd && !tlag;
                        e.ssOff[j] = rSide ? (long)k * nzx : capOff(k);   // receiver side: R[k] / sample k
                        e.wOff[j] = rSide ? capOff(k) : (long)k * nzx;    // source side: sample k / W''[k] (R[k])
                    }
                    if (tlag) {   // imagingWemvaTimeAdjGpu / imagingTimeAdjGpu
                        time_lags(e, wemvaFwd ? -1 : 1);
                        HIPC(launch_time_img_batch(e, nb, s));
                    } else {
                        HIPC(launch_ext_img_batch(e, nb, s, wemvaFwd));
                    }
                    pend.clear();
                };
                if (c->leg1) {   // forward from m * W'', imaged against R (batches of extK samples)
                    forward_pass(MODE_TOMO_CAP, dSlotsC, SS_MUL, dSlots, mtil, 2, [&](int k) {
                        pend.push_back(k);
                        if ((int)pend.size() == extK || k == p.nts - 1) flush(true);
                    });
                }
                if (c->leg2) {   // backward from m * R, imaged against W''
                    backward_pass(SS_MUL, dSlotsQ, mtil, [&](int k) {
                        pend.push_back(k);
                        if ((int)pend.size() == extK || k == 0) flush(false);
                    });
                }
                HIPC(launch_ext_adj_finish(dSlots, nb, extScale, p.nz, p.nx, p.nExt, s));
            }
        }

        // -- second pass
        if (!lockstep && !adj2 && !tomo) {
            HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
            a.cur = 1;
            a.assignA = a.assignB = a.skipA = a.skipB = 0;
            if (q.op == Op::NlFwd || q.op == Op::BornFwd || q.op == Op::ExtFwd) {
                unsigned mode = MODE_NL_FWD;
                int kind = AISO_K_NL_FWD;
                if (q.op == Op::BornFwd) { mode = win ? MODE_BORN_SCAT_PRE : MODE_BORN_SCAT; kind = AISO_K_BORN_SCAT; }
                if (q.op == Op::ExtFwd) { mode = MODE_EXT_SCAT; kind = AISO_K_EXT_SCAT; }
                int L = 0, R = 1;
                // Secondary sources of the offset-extended forward, extK
                // samples per launch into the ring, ahead of their use.
                int nextSS = 0;
                auto ext_ss_upto = [&](int upto) {
                    while (nextSS <= upto) {
                        ExtBatchArgs e{};
                        e.slots = dSlots;
                        e.v2dt2 = base.v2dt2;
                        e.mext = mtil;
                        e.nz = p.nz; e.nx = p.nx; e.hExt = p.hExt;
                        e.nk = std::min(extK, p.nts - nextSS);
                        for (int j = 0; j < e.nk; ++j) {
                            e.ssOff[j] = ssSlot(nextSS + j);
                            e.wOff[j] = (long)(nextSS + j) * nzx;
                        }
                        if (tlag) {
                            time_lags(e, 1);
                            HIPC(launch_time_ss_batch(e, nb, s));
                        } else {
                            HIPC(launch_ext_ss_batch(e, nb, s));
                        }
                        nextSS += e.nk;
                    }
                };
                fork_lanes(c);
                for (int its = 0; its < p.nts - 1; ++its) {
                    if (q.op == Op::ExtFwd) {
                        ext_ss_upto(its + 1);
                        L = R = 0;
                        a.ssOffA = ssSlot(its);
                        a.ssOffB = ssSlot(its + 1);
                    }
                    for (int it2 = 1; it2 <= p.sub; ++it2) {
                        a.its = its;
                        a.itwInj = its * p.sub + it2 - 1;
                        a.fa = f[it2 - 1];
                        a.fb = f[h + it2 - 1];
                        a.wA = (long)its * nzx;
                        a.wB = (long)(its + 1) * nzx;
                        a.ssA = L;
                        a.ssB = R;
                        a.fe = f[it2];
                        a.ff = f[h + it2];
                        launch_timed(c, mode, kind, a, nb, generic);
                        ring(mode, a);
                        a.cur ^= 1;
                    }
                    std::swap(L, R);
                }
                join_lanes(c);
                for (int k = 0; k < nb; ++k) {
                    const int shot = b0 + k;
                    const aiso_device* rd = q.rec[q.nRec == 1 ? 0 : shot];
                    DevGeom* rg = recG[q.nRec == 1 ? 0 : shot];
                    HIPC(launch_dev_forward(dreg + (size_t)k * nRecRegMax * p.nts, dDataOut + (size_t)shot * nRecDev * p.nts,
                                            rg->cornerReg.as<int>(), rg->weight.as<float>(), nRecDev, rd->g.nb, p.nts,
                                            q.add, s));
                }
            } else {
                unsigned mode = MODE_NL_ADJ;
                int kind = AISO_K_NL_ADJ;
                if (q.op == Op::BornAdj) { mode = win ? MODE_REC_WIN : MODE_BORN_REC; kind = AISO_K_BORN_REC; }
                if (q.op == Op::ExtAdj) { mode = win ? MODE_REC_WIN : MODE_EXT_REC; kind = AISO_K_EXT_REC; }
                if (ss) HIPC(launch_fill(ss, (long)nb * ssSlices * nzx, s));
                if (img) HIPC(launch_fill(img, (long)nb * nExt * nzx, s));
                a.winOut = q.op == Op::BornAdj ? WIN_IMG : WIN_SLICE;
                a.wa = f[1];
                a.wb = f[0];
                a.wc = f[h + 1];
                // The stored source wavefield lets the receiver pass run the batch
                // in smaller groups, whose per-shot slices (pair + image) stay in
                // the on-die cache.
                const int subB = c->batchRec > 0 ? std::min(c->batchRec, nb) : nb;
                for (int k0 = 0; k0 < nb; k0 += subB) {
                    const int nbr = std::min(subB, nb - k0);
                    ShotSlot* sl = dSlots + k0;
                    StepArgs r = a;
                    r.slots = sl;
                    r.nshots = nbr;
                    r.gridLogical = r.ntiles * nbr;
                    r.cur = 1;
                    int L = 0, R = 1;
                    // Completed receiver-wavefield samples awaiting the offset
                    // imaging, newest last; imaged extK at a time.
                    std::vector<int> pend;
                    auto ext_img_flush = [&] {
                        if (pend.empty()) return;
                        ExtBatchArgs e{};
                        e.slots = sl;
                        e.v2dt2 = base.v2dt2;
                        e.nz = p.nz; e.nx = p.nx; e.hExt = p.hExt;
                        e.nk = (int)pend.size();
                        for (int j = 0; j < e.nk; ++j) {
                            e.ssOff[j] = ssSlot(pend[j]);
                            e.wOff[j] = (long)pend[j] * nzx;
                        }
                        if (tlag) {
                            time_lags(e, 1);
                            HIPC(launch_time_img_batch(e, nbr, s));
                        } else {
                            HIPC(launch_ext_img_batch(e, nbr, s));
                        }
                        pend.clear();
                    };
                    fork_lanes(c);
                    for (int its = p.nts - 2; its >= 0; --its) {
                        for (int it2 = p.sub - 1; it2 >= 0; --it2) {
                            r.its = its;
                            r.itwRec = its * p.sub + it2;
                            r.fc = f[it2 + 1];
                            r.fd = f[h + it2 + 1];
                            r.ssA = win ? 0 : L;
                            r.ssB = R;
                            if (q.op == Op::ExtAdj) {   // A: sample its (window: its + 1), B: sample its + 1
                                r.ssA = r.ssB = 0;
                                r.ssOffA = ssSlot(win ? its + 1 : its);
                                r.ssOffB = ssSlot(its + 1);
                            }
                            r.fe = f[it2];
                            r.ff = f[h + it2];
                            r.skipA = r.fe == 0.f;
                            r.skipB = r.ff == 0.f;
                            r.assignA = it2 == p.sub - 1;
                            r.assignB = 0;
                            r.image = (q.op == Op::BornAdj && it2 == 0) ? 1 : 0;
                            r.win = it2 == 1;
                            r.wImg = (long)(its + 1) * nzx;
                            launch_timed(c, mode, kind, r, nbr, generic);
                            ring(mode, r);
                            r.cur ^= 1;
                        }
                        if (q.op == Op::ExtAdj) {
                            pend.push_back(its + 1);
                            if ((int)pend.size() == extK) ext_img_flush();
                        }
                        std::swap(L, R);
                    }
                    join_lanes(c);
                    if (q.op == Op::BornAdj) {
                        if (win) {   // img += W[0] * (q(1)*f[1] + q(0)*f[0]), then the scaling
                            HIPC(launch_win_tail(sl, nbr, r.cur, WIN_IMG, f[1], f[0], 0, 0, 0, nullptr, p.nz, p.nx, s));
                            HIPC(launch_born_adj_finish(sl, nbr, -1, base.v2dt2, c->refl.as<float>(), p.nz, p.nx, s));
                        } else {
                            HIPC(launch_born_adj_finish(sl, nbr, R, base.v2dt2, c->refl.as<float>(), p.nz, p.nx, s));
                        }
                    } else if (q.op == Op::ExtAdj) {
                        if (win)
                            HIPC(launch_win_tail(sl, nbr, r.cur, WIN_SLICE, f[1], f[0], ssSlot(0), 0, 0, nullptr, p.nz,
                                                 p.nx, s));
                        pend.push_back(0);
                        ext_img_flush();
                        HIPC(launch_ext_adj_finish(sl, nbr, extScale, p.nz, p.nx, nExt, s));
                    }
                }
            }
        }
        if (needImg && perShotImg) {
            for (int k = 0; k < nb; ++k)
                HIPC(launch_add_to(dModelOut + (size_t)(b0 + k) * nImg * nzx, img + (size_t)k * nImg * nzx, nzx * nImg,
                                   q.add, s));
        } else if (needImg) {
            for (int k = 0; k < nb; ++k) HIPC(launch_accumulate(acc, img + (size_t)k * nImg * nzx, nzx * nImg, s));
        }
        if (q.op == Op::NlAdj) {
            std::vector<float> fineAll((size_t)nb * nSrcRegMax * p.ntw);
            HIPC(hipMemcpyAsync(fineAll.data(), sig, fineAll.size() * 4, hipMemcpyDeviceToHost, s));
            HIPC(hipStreamSynchronize(s));
            for (int k = 0; k < nb; ++k) {
                const int shot = b0 + k;
                const Geometry& sgm = q.src[shot]->g;
                const int nReg = (int)sgm.regPos.size();
                std::vector<float> reg((size_t)nReg * p.nts);
                interp_time_adj(p, nReg, reg.data(), fineAll.data() + (size_t)k * nSrcRegMax * p.ntw);
                scale_source(p, c->vel.data(), sgm.regPos, p.nts, reg.data());
                if (q.nShotSig == 1) {
                    sgm.forward(true, reg.data(), nlAcc.data(), p.nts);
                } else {
                    std::vector<float> sl((size_t)nSrcDev * p.nts);
                    sgm.forward(false, reg.data(), sl.data(), p.nts);
                    float* dst = hModel.data() + (size_t)shot * nSrcDev * p.nts;
                    for (size_t i = 0; i < sl.size(); ++i) dst[i] += sl[i];
                }
            }
        }
        if (c->profile) {
            hipEvent_t batch1 = pooled_event(c);
            HIPC(hipEventRecord(batch1, s));
            c->batchEv.push_back({batch0, batch1, nb});
        }
        drain_profile(c);
    }

    // ---- outputs
    if (needImg && !perShotImg) HIPC(launch_add_to(dModelOut, acc, nzx * nImg, q.add, s));
    if (q.op == Op::NlAdj) {
        if (q.nShotSig == 1)
            for (size_t i = 0; i < nlAcc.size(); ++i) hModel[i] += nlAcc[i];
        if (q.mem == AISO_MEM_DEVICE) HIPC(hipMemcpyAsync(q.modelOut, hModel.data(), modelN * 4, hipMemcpyHostToDevice, s));
        else std::memcpy(q.modelOut, hModel.data(), modelN * 4);
    }
    if (q.mem == AISO_MEM_HOST) {
        if ((is_adj(q.op) && is_born(q.op)) || wemva)
            HIPC(hipMemcpyAsync(q.modelOut, dModelOut, mOutN * 4, hipMemcpyDeviceToHost, s));
        if (!is_adj(q.op) && !wemva) HIPC(hipMemcpyAsync(q.dataOut, dDataOut, dataN * 4, hipMemcpyDeviceToHost, s));
    }
    HIPC(hipStreamSynchronize(s));
    HIPC(hipGetLastError());
}

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
    int B = c->batchAdj > 0 ? c->batchAdj : (c->batchReq > 0 ? c->batchReq : 8);
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
    base.tilesZ = tilesZ(c);
    base.ntiles = base.tilesZ * tilesX(c);
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
        auto save = [&](int j, int cur) {
            for (int k = 0; k < nb; ++k) {
                HIPC(hipMemcpyAsync(ckSlice(k, j, 0), fs[k].p[cur], nzx * 4, hipMemcpyDeviceToDevice, s));
                HIPC(hipMemcpyAsync(ckSlice(k, j, 1), fs[k].p[cur ^ 1], nzx * 4, hipMemcpyDeviceToDevice, s));
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
            launch_timed(c, MODE_SRC, AISO_K_SRC, x, nb);
            if (maxSrcRing > 0) HIPC(launch_ring(M_INJ_SIG, x, maxSrcRing, s));
        };
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
                HIPC(hipMemcpyAsync(fs[k].p[1], ckSlice(k, j, 0), nzx * 4, hipMemcpyDeviceToDevice, s));
                HIPC(hipMemcpyAsync(fs[k].p[0], ckSlice(k, j, 1), nzx * 4, hipMemcpyDeviceToDevice, s));
                HIPC(hipMemcpyAsync(fs[k].W, ckSlice(k, j, 0), nzx * 4, hipMemcpyDeviceToDevice, s));
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
                launch_timed(c, MODE_NL_ADJ, AISO_K_NL_ADJ, r, nb);
                if (maxSrcRing > 0 || maxRecRing > 0)
                    HIPC(launch_ring(M_INJ_DATA | M_REC_SIG, r, std::max(maxSrcRing, maxRecRing), s));
                r.cur ^= 1;
                ga.wF = (long)(t - 1 - t0) * nzx;
                ga.nu = r.cur;
                HIPC(launch_grad(c->tz, c->tx, ga, s));
            }
        }
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
        run(c, q);
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
        else if (k == "cacheBudgetMB") c->cacheBudget = (long)std::max(1, value) << 20;
        else if (k == "profileEvery") {
            if (value < 1) throw Error(AISO_ERR_PARAM, "profileEvery must be >= 1");
            c->profEvery = value;
        }
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
