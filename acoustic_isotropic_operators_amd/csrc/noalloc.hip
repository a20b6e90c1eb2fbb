// noalloc.hip -- aiso_nl_forward_noalloc: the nonlinear forward of
// aiso_nl_forward with no device allocation, for callers that cannot
// allocate (code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// Every buffer the call needs beyond its inputs and outputs (wavefield
// pairs, regular-grid traces and source signals, slot table, per-tile device
// lists) is carved from `scratch`, one file-scope device array of
// kScratchBytes per GPU, held until the process ends.  The host side never
// uses the array's name as an address: it asks hipGetSymbolAddress for the
// array's address on each device once, with that device current, and passes
// that pointer to the kernels and copies.  Calls that overlap on one GPU
// share the array and race; each call clears what it needs zeroed.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine_internal.hpp"

namespace aiso {
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_scratchMu;
void* g_scratchPtr[kMaxDevices] = {};

// The scratch array's address on `device` (current when this is called):
// looked up once per device, then reused.
char* scratch_on(int device)
{
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (device < 0 || device >= kMaxDevices) throw Error(AISO_ERR_PARAM, "device id out of range for the scratch table");
    if (!g_scratchPtr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error(AISO_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchPtr[device] = ptr;
    }
    return static_cast<char*>(g_scratchPtr[device]);
}

// Bump allocation inside the scratch array, 256-byte aligned.
struct Carve {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count)
    {
        const size_t bytes = (count * sizeof(T) + 255) / 256 * 256;
        if (off + bytes > kScratchBytes)
            throw Error(AISO_ERR_MEMORY, "the call needs more than the 64 MiB scratch of aiso_nl_forward_noalloc "
                                         "(use fewer receivers, a smaller grid, or aiso_nl_forward)");
        T* p = reinterpret_cast<T*>(base + off);
        off += bytes;
        return p;
    }
};

template <class T>
void put(T* dst, const std::vector<T>& v, hipStream_t s)
{
    if (!v.empty()) HIPC(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// F/src/nonlinearPropShotsGpu.cpp:62-152 (the shot loop) and
// F/src/nonlinearShotsGpuFunctions.cu:227-309 (per shot), as aiso_nl_forward
// runs them: the same kernels and launch sequence, on scratch buffers.
void nl_forward_noalloc(aiso_ctx* c, int add, const float* model, int nShotSig, float* data,
                        aiso_device* const* src, int nShot, aiso_device* const* rec, int nRec)
{
    const FdParams& p = c->p;
    if (!c->haveVel) throw Error(AISO_ERR_STATE, "velocity not set");
    if (nShot < 1) throw Error(AISO_ERR_PARAM, "nShot must be >= 1");
    if (nRec != 1 && nRec != nShot) throw Error(AISO_ERR_PARAM, "receiver geometry count must be 1 or nShot");
    if (nShotSig != 1 && nShotSig != nShot) throw Error(AISO_ERR_PARAM, "model shot axis must be 1 or nShot");
    if (p.sub >= kSubMax) throw Error(AISO_ERR_PARAM, "**** ERROR: Subsampling parameter too high ****");
    if (p.minPad >= kPadMax) throw Error(AISO_ERR_PARAM, "**** ERROR: Padding value is too high ****");
    if (p.freeSurface) throw Error(AISO_ERR_PARAM, "aiso_nl_forward_noalloc does not take a free surface");
    const int nRecDev = rec[0]->g.nDev, nSrcDev = src[0]->g.nDev;
    for (int i = 0; i < nRec; ++i)
        if (rec[i]->g.nDev != nRecDev) throw Error(AISO_ERR_GEOM, "receivers per shot must be constant");
    for (int i = 0; i < nShot; ++i)
        if (src[i]->g.nDev != nSrcDev) throw Error(AISO_ERR_GEOM, "sources per shot must be constant");
    for (int i = 0; i < nShot; ++i)
        for (int v : src[i]->g.gridIdx)
            if (v >= p.nzx()) throw Error(AISO_ERR_GEOM, "source outside the velocity grid");
    for (int i = 0; i < nRec; ++i)
        for (int v : rec[i]->g.gridIdx)
            if (v >= p.nzx()) throw Error(AISO_ERR_GEOM, "receiver outside the velocity grid");

    hipStream_t s = c->stream;
    wait_default_stream(c);
    Carve cv{scratch_on(c->device)};
    const long nzx = p.nzx();
    const int h = p.h();
    const float* f = p.filt.data();
    const int tz = c->tzFwd, tx = c->txFwd;
    const int tZ = (p.nz - 2 * kFat + tz - 1) / tz, nTiles = tZ * ((p.nx - 2 * kFat + tx - 1) / tx);

    // the wavelets, through the host like aiso_nl_forward
    std::vector<float> hModel((size_t)nShotSig * nSrcDev * p.nts);
    HIPC(hipMemcpy(hModel.data(), model, hModel.size() * 4, hipMemcpyDeviceToHost));

    // per-tile lists of every source and receiver geometry, in scratch
    struct Dev {
        int* tilePtr;
        int2* ent;
        int2* ring;
        int nRing, nReg;
        int* cornerReg;
        float* weight;
    };
    std::vector<TileLists> hostLists;
    hostLists.reserve(nShot + nRec);
    auto place = [&](const aiso_device* d) {
        hostLists.push_back(tile_lists(d->g, p.nz, p.nx, tz, tx));
        const TileLists& L = hostLists.back();
        Dev out{};
        out.tilePtr = cv.take<int>(L.ptr.size());
        out.ent = cv.take<int2>(std::max<size_t>(L.ent.size(), 1));
        out.ring = cv.take<int2>(std::max<size_t>(L.ring.size(), 1));
        out.cornerReg = cv.take<int>(d->g.cornerReg.size());
        out.weight = cv.take<float>(d->g.weight.size());
        put(out.tilePtr, L.ptr, s);
        put(out.ent, L.ent, s);
        put(out.ring, L.ring, s);
        put(out.cornerReg, d->g.cornerReg, s);
        put(out.weight, d->g.weight, s);
        out.nRing = (int)L.ring.size();
        out.nReg = (int)d->g.regPos.size();
        return out;
    };
    std::vector<Dev> srcL, recL;
    int nSrcRegMax = 1, nRecRegMax = 1;
    for (int i = 0; i < nShot; ++i) {
        srcL.push_back(place(src[i]));
        nSrcRegMax = std::max(nSrcRegMax, srcL.back().nReg);
    }
    for (int i = 0; i < nRec; ++i) {
        recL.push_back(place(rec[i]));
        nRecRegMax = std::max(nRecRegMax, recL.back().nReg);
    }

    // shots per launch: as many as the rest of the scratch holds
    const size_t perShot = ((2 * nzx * 4 + 255) / 256 + ((size_t)nRecRegMax * p.nts * 4 + 255) / 256 +
                            ((size_t)nSrcRegMax * p.ntw * 4 + 255) / 256) * 256 + sizeof(ShotSlot);
    const size_t left = kScratchBytes - cv.off;
    int B = (int)std::min<size_t>((size_t)nShot, left / (perShot + 256));
    if (B < 1)
        throw Error(AISO_ERR_MEMORY, "one shot needs more than the 64 MiB scratch of aiso_nl_forward_noalloc "
                                     "(use fewer receivers, a smaller grid, or aiso_nl_forward)");
    ShotSlot* dSlots = cv.take<ShotSlot>(B);
    float* pairs = cv.take<float>((size_t)B * 2 * nzx);
    float* dreg = cv.take<float>((size_t)B * nRecRegMax * p.nts);
    float* sig = cv.take<float>((size_t)B * nSrcRegMax * p.ntw);

    StepArgs base{};
    base.slots = dSlots;
    base.v2dt2 = c->v2dt2.as<float>();
    base.cd = c->cd.as<float>();
    base.nz = p.nz; base.nx = p.nx; base.minPad = p.minPad; base.nts = p.nts; base.ntw = p.ntw;
    base.tilesZ = tZ;
    base.ntiles = nTiles;
    for (int k = 0; k < 6; ++k) { base.zc[k] = p.zc[k]; base.xc[k] = p.xc[k]; }
    base.ntW = c->streamNT;

    for (int b0 = 0; b0 < nShot; b0 += B) {
        const int nb = std::min(B, nShot - b0);
        std::vector<ShotSlot> slots(nb);
        std::vector<float> sigHost((size_t)nb * nSrcRegMax * p.ntw, 0.f);
        int maxSrcRing = 0, maxRecRing = 0;
        for (int k = 0; k < nb; ++k) {
            const int shot = b0 + k;
            const Dev& sg = srcL[shot];
            const Dev& rg = recL[nRec == 1 ? 0 : shot];
            ShotSlot& S = slots[k];
            std::memset(&S, 0, sizeof(S));
            S.p[0] = pairs + (size_t)k * 2 * nzx;
            S.p[1] = S.p[0] + nzx;
            S.dreg = dreg + (size_t)k * nRecRegMax * p.nts;
            S.sig = sig + (size_t)k * nSrcRegMax * p.ntw;
            S.srcTile = sg.tilePtr;
            S.srcEnt = sg.ent;
            S.srcRing = sg.ring;
            S.nSrcRing = sg.nRing;
            S.recTile = rg.tilePtr;
            S.recEnt = rg.ent;
            S.recRing = rg.ring;
            S.nRecRing = rg.nRing;
            maxSrcRing = std::max(maxSrcRing, sg.nRing);
            maxRecRing = std::max(maxRecRing, rg.nRing);
            std::vector<float> fine =
                nl_source(c, src[shot]->g, hModel.data() + (size_t)(nShotSig == 1 ? 0 : shot) * nSrcDev * p.nts);
            std::memcpy(sigHost.data() + (size_t)k * nSrcRegMax * p.ntw, fine.data(), fine.size() * 4);
        }
        put(dSlots, slots, s);
        put(sig, sigHost, s);
        HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
        HIPC(launch_fill(dreg, (long)nb * nRecRegMax * p.nts, s));
        StepArgs a = base;
        a.nshots = nb;
        a.gridLogical = a.ntiles * nb;
        a.cur = 1;
        const int mr = std::max(maxSrcRing, maxRecRing);
        for (int its = 0; its < p.nts - 1; ++its)
            for (int it2 = 1; it2 <= p.sub; ++it2) {
                a.its = its;
                a.itwInj = its * p.sub + it2 - 1;
                a.fe = f[it2];
                a.ff = f[h + it2];
                HIPC(launch_step(MODE_NL_FWD, tz, tx, 1, a, s));
                if (mr > 0) HIPC(launch_ring(MODE_NL_FWD, a, mr, s));
                a.cur ^= 1;
            }
        for (int k = 0; k < nb; ++k) {
            const int shot = b0 + k;
            const Dev& rg = recL[nRec == 1 ? 0 : shot];
            HIPC(launch_dev_forward(dreg + (size_t)k * nRecRegMax * p.nts, data + (size_t)shot * nRecDev * p.nts,
                                    rg.cornerReg, rg.weight, nRecDev, rec[nRec == 1 ? 0 : shot]->g.nb, p.nts, add, s));
        }
        // the host vectors behind this batch's copies live until here
        HIPC(hipStreamSynchronize(s));
    }
    HIPC(hipGetLastError());
}

}  // namespace
}  // namespace aiso

extern "C" aiso_status aiso_nl_forward_noalloc(aiso_ctx* c, int add, const float* model, int nShotSig, float* data,
                                               aiso_device* const* sources, int nShot, aiso_device* const* receivers,
                                               int nRecGeom)
{
    return aiso::guard_call([&] {
        if (!c) throw aiso::Error(AISO_ERR_STATE, "null context");
        if (!model || !data || !sources || !receivers) throw aiso::Error(AISO_ERR_PARAM, "null argument");
        HIPC(hipSetDevice(c->device));
        aiso::nl_forward_noalloc(c, add, model, nShotSig, data, sources, nShot, receivers, nRecGeom);
    });
}
