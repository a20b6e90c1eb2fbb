// engine_internal.hpp -- state shared by the engine's translation units: the
// context and geometry handles behind the C-ABI, device and pinned host
// buffers, and the helpers the operator passes use.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "aiso.h"
#include "host.hpp"
#include "kernels.hpp"

namespace aiso {

void hip_check(hipError_t e, const char* what);
#define HIPC(x) ::aiso::hip_check((x), #x)

// Grow-only device buffer.  Growing frees the old allocation first; hipFree
// waits for the device work that may still read it.
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

// Grow-only pinned host buffer (hipHostMalloc): copies from it are truly
// asynchronous, so its contents must not change before they have run.
struct PinnedBuf {
    void* p = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    void* need(size_t bytes)
    {
        if (bytes <= n && p) return p;
        release();
        HIPC(hipHostMalloc(&p, std::max<size_t>(bytes, 256), hipHostMallocDefault));
        n = std::max<size_t>(bytes, 256);
        return p;
    }
};

template <class T>
void upload(DBuf& b, const std::vector<T>& v, hipStream_t s)
{
    b.need(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (!v.empty()) HIPC(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// Per-tile sparse lists of one geometry for one tile shape, plus the arrays
// of the device-side spatial interpolation.
struct DevGeom {
    DBuf tilePtr, ent, ring, cornerReg, weight, adjPtr, adjEnt;
    int nRing = 0, nReg = 0;
};

long next_geom_id();

}  // namespace aiso

struct aiso_device {
    aiso::Geometry g;
    long id = aiso::next_geom_id();
};

struct aiso_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    aiso::FdParams p;
    bool haveVel = false;
    std::vector<float> vel, v2dt2h, reflh;
    aiso::DBuf v2dt2, refl, cd;
    int batchReq = 0, batchUsed = 0;
    int tz = 32, tx = 32;               // tile of the call in progress
    // Step-kernel tile per direction: the forward's pair kernel runs best on
    // 32 x 16 tiles, the adjoint passes on 32 x 32 (MI355X measurements).
    int tzFwd = 32, txFwd = 16, tzAdj = 32, txAdj = 32;
    bool tileFwdSet = false;            // the caller chose the forward tile
    int streamNT = 1;
    size_t bytesPerShot = 0;
    // work buffers
    aiso::DBuf pairs, pairsQ, W, ss, img, dreg, sig, slots, slotsQ, mtil, acc, modelDev, dataDev;
    aiso::DBuf ck;                      // checkpoints of the velocity-gradient pass
    // tomographic operator: the extended reflectivity it linearizes about
    // (set once, [nExt][nx][nz]), its per-call 2/v^3 scaling, and the
    // scattered / receiver wavefield and capture buffers
    aiso::DBuf extRefl, rExt, P, cap, slotsC, slotsI, slotsJ, imgDev;
    bool haveExtRefl = false;
    int leg1 = 1, leg2 = 1;
    // time-lag extension: the per-point model scale of the call and the
    // sloth key (scale by -v^2 dtw^2 instead of 2/v^3 v^2 dtw^2)
    aiso::DBuf tscale;
    int sloth = 0;
    // Born and extended-Born adjoints: one image per shot ([nShot][nExt][nx][nz])
    // instead of their sum, for callers that batch independent requests.
    int imagePerShot = 0;
    // Wavefield saving (the reference's saveWavefield): the coarse source
    // wavefield [nts][nx][nz] of shot `wavefieldShot` of the last Born-type
    // call, and its scattered (forward) or receiver (adjoint) wavefield; the
    // nonlinear operators save their propagated wavefield.
    int saveWavefield = 0, wavefieldShot = 0;
    aiso::DBuf wfSrc, wfSec, wfScratch;
    int wfSrcValid = 0, wfSecValid = 0;
    hipEvent_t evIn = nullptr;          // orders device inputs after the default stream's work
    // Host staging of the per-batch tables and source traces: pinned, reused
    // batch after batch once the copies of the previous batch have run.
    aiso::PinnedBuf stage;
    size_t stageOff = 0;
    hipEvent_t stageEv = nullptr;
    bool stagePending = false;
    // Lanes: the shots of a batch split over streams whose kernels run
    // concurrently, so one lane's launch tail (CUs idle while its last
    // workgroups finish) and launch gap are filled by the other's.  On
    // MI355X two lanes raise the Born benchmark from 279 to 296 Gpts/s;
    // three and four lanes measured slower than one.
    int lanes = 2;
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
    // 6 / 8: the specialized step kernels compiled to the register budget
    // of that many waves per SIMD (k_step_w); 0: the compiler's own budget.
    int occ = 6;
    long cacheBudget = 192l << 20;      // bytes of hot slices per launch (Infinity Cache is 256 MB)
    std::map<std::tuple<long, int, int>, std::unique_ptr<aiso::DevGeom>> geoms;
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

namespace aiso {

enum class Op { NlFwd, NlAdj, BornFwd, BornAdj, ExtFwd, ExtAdj, TomoFwd, TomoAdj, WemvaFwd, WemvaAdj };

// The adjoint of each operator writes the model side (for WEMVA, the
// velocity perturbation); both WEMVA directions read observed data.
inline bool is_adj(Op op)
{
    return op == Op::NlAdj || op == Op::BornAdj || op == Op::ExtAdj || op == Op::TomoAdj || op == Op::WemvaAdj;
}
inline bool is_born(Op op) { return op != Op::NlFwd && op != Op::NlAdj; }
inline bool is_ext(Op op) { return op == Op::ExtFwd || op == Op::ExtAdj; }
inline bool is_tomo(Op op) { return op == Op::TomoFwd || op == Op::TomoAdj; }
inline bool is_wemva(Op op) { return op == Op::WemvaFwd || op == Op::WemvaAdj; }

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

int tiles_z(const aiso_ctx* c);
int tiles_x(const aiso_ctx* c);
struct TileLists {
    std::vector<int> ptr;        // tiles + 1 CSR offsets
    std::vector<int2> ent;       // (local index in tile, trace)
    std::vector<int2> ring;      // (grid index, trace) of points in the FAT ring
};
TileLists tile_lists(const Geometry& g, int nz, int nx, int tz, int tx);
DevGeom* dev_geom(aiso_ctx* c, const aiso_device* d);
// Tiling of the two-step kernel: output tiles of OZ x OX points.
constexpr int kOZ2 = kPair2EZ - 2 * kFat, kOX2 = kPair2EX - 2 * kFat;
int tiles_z2(const aiso_ctx* c);
int tiles_x2(const aiso_ctx* c);
DevGeom* dev_geom_pair2(aiso_ctx* c, const aiso_device* d, bool region);

hipEvent_t pooled_event(aiso_ctx* c);
void wait_default_stream(aiso_ctx* c);
void lane_range(int nshots, int lanes, int l, int& k0, int& cnt);
StepArgs lane_args(const StepArgs& a, int k0, int cnt);
hipStream_t lane_stream(const aiso_ctx* c, int l);
void ensure_lanes(aiso_ctx* c);
void fork_lanes(aiso_ctx* c);
void join_lanes(aiso_ctx* c);
bool sample_event(aiso_ctx* c, int kind);
// Pass-level timing of a pass whose shots run on several lanes: an event on
// the context's stream before the lanes fork and one after they join, the
// interval credited to `kind` with the point-updates of every lane and
// `steps` launches (one per fine step, counting the lanes' concurrent
// dispatches as one).  No-op unless profiling with more than one lane.
hipEvent_t pass_timer_begin(aiso_ctx* c);
void pass_timer_end(aiso_ctx* c, hipEvent_t e0, int kind, double pts, int steps);
void launch_timed(aiso_ctx* c, unsigned mode, int kind, const StepArgs& a, bool generic = false);
void drain_profile(aiso_ctx* c);

// Pinned staging of one batch's host-to-device copies: stage_begin waits for
// the copies of the previous batch and sizes the buffer; stage_copy appends
// and queues one copy on the context's stream; stage_end marks the batch's
// copies so the next stage_begin can reuse the buffer.
void stage_begin(aiso_ctx* c, size_t bytes);
void stage_copy(aiso_ctx* c, void* dst, const void* src, size_t bytes);
void stage_end(aiso_ctx* c);
inline size_t stage_size(size_t bytes) { return (bytes + 255) / 256 * 256; }

std::vector<float> born_source(const aiso_ctx* c, const Geometry& src, const float* wav);
std::vector<float> nl_source(const aiso_ctx* c, const Geometry& src, const float* m);

// One operator call (passes.cpp).
void run_call(aiso_ctx* c, const Request& q);

// The C-ABI's error convention: run f, turn an exception into a status code
// and the message aiso_last_error() returns.
void set_last_error(const char* msg);
template <class F>
aiso_status guard_call(F&& f)
{
    try {
        f();
        return AISO_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return AISO_ERR_PARAM;
    }
}

}  // namespace aiso
