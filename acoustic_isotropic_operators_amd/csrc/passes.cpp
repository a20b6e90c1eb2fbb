// passes.cpp -- one call of a multi-shot operator, as a sequence of stages.
//
// A call checks its request, stages its operands on the device, sizes the
// batches of shots that one launch advances, and for every batch runs the
// propagation passes of its operator:
//
//   Born forward, sub == 2     lockstep pair (or two fine steps per launch)
//   Born adjoint, sub == 2     two fine steps per launch (opt-in), else:
//   every Born-type operator   source pass (coarse source wavefield W)
//   tomographic / WEMVA        their two legs
//   NL / Born / ext forward    scattered (or nonlinear) pass, data recorded
//   NL / Born / ext adjoint    receiver pass with imaging
//
// The reference runs these per shot with 4-5 launches per fine step
// (F/src/BornShotsGpuFunctions.cu:260-388, 829-957;
// F/src/nonlinearShotsGpuFunctions.cu:227-309, 579-652;
// F/src/BornExtShotsGpuFunctions.cu:582-724, 1848-1989); here each fine step
// of a pass is one launch of the fused step kernel over the batch.
#include <cmath>
#include <functional>
#include <memory>

#include "engine_internal.hpp"

namespace aiso {

namespace {

class Call {
public:
    Call(aiso_ctx* ctx, const Request& req) : c(ctx), q(req), p(ctx->p) {}
    void run();

private:
    // ---- setup stages
    void select_tile();
    void validate() const;
    void derive_sizes();
    void stage_operands();
    void build_geometry();
    void plan_batches();
    void allocate();
    void prepare_model();
    void init_step_args();
    // ---- per batch
    void begin_batch(int first, int count);
    void upload_tables();
    void load_receiver_traces();
    void pass_pair2();
    void pass_lockstep();
    void pass_adjoint2();
    void pass_source();
    void passes_tomo();
    void pass_forward();
    void pass_adjoint();
    void record_data();
    void reduce_images();
    void finish_nl_adjoint();
    void save_wavefields();
    void end_batch();
    // ---- end of the call
    void outputs();

    // helpers
    bool ring(unsigned mode, const StepArgs& sa);
    long ss_slot(int k) const { return (long)(k % ssSlices) * nzx; }
    long cap_off(int k) const { return (long)(k % capSlices) * nzx; }
    void time_lags(ExtBatchArgs& e, int dir) const;
    const aiso_device* rec_dev(int shot) const { return q.rec[q.nRec == 1 ? 0 : shot]; }
    DevGeom* rec_geom(int shot) const { return recG[q.nRec == 1 ? 0 : shot]; }

    aiso_ctx* c;
    const Request& q;
    const FdParams& p;
    hipStream_t s = nullptr;

    // sizes and modes of the call
    long nzx = 0, modelN = 0, mInN = 0, mOutN = 0, dataN = 0, wSlices = 0;
    int nExt = 1, nImg = 1, nRecDev = 0, nSrcDev = 0, h = 0;
    const float* f = nullptr;
    bool wemva = false, tomo = false, perShotImg = false, dataInput = false, fsurf = false, tlag = false;
    bool lockstep = false, pair2 = false, adj2 = false, win = false, generic = false;
    bool needW = false, needSS = false, needImg = false;
    int pairSlices = 2, ssSlices = 2, capSlices = 2, extK = 1, B = 1;

    // operands on the device
    const float* dModelIn = nullptr;
    float* dModelOut = nullptr;
    const float* dDataIn = nullptr;
    float* dDataOut = nullptr;
    std::vector<float> hModel, nlAcc;   // NL models are assembled on the host

    // geometry lists
    std::vector<DevGeom*> srcG, recG;
    int nSrcRegMax = 1, nRecRegMax = 1;

    // work buffers (per batch of B shots)
    float *pairs = nullptr, *W = nullptr, *pairsQ = nullptr, *Pw = nullptr, *cap = nullptr;
    float *ss = nullptr, *img = nullptr, *dreg = nullptr, *sig = nullptr;
    ShotSlot *dSlots = nullptr, *dSlotsQ = nullptr, *dSlotsC = nullptr, *dSlotsI = nullptr, *dSlotsJ = nullptr;
    const float* mtil = nullptr;        // scaled (extended) model of the forward
    const float* extScale = nullptr;    // per-point scale of the extended model
    const float* rExt = nullptr;        // scaled extended reflectivity / image
    float* acc = nullptr;               // image summed over shots
    StepArgs base{};

    // the batch in progress
    int b0 = 0, nb = 0;
    std::vector<ShotSlot> slots;
    std::vector<float> sigHost;
    int maxSrcRing = 0, maxRecRing = 0;
    StepArgs a{};
    hipEvent_t batch0 = nullptr;
};

// ------------------------------------------------------------ setup stages

void Call::select_tile()
{
    c->tz = is_adj(q.op) ? c->tzAdj : c->tzFwd;
    c->tx = is_adj(q.op) ? c->txAdj : c->txFwd;
    // The forward passes of the extended, tomographic and WEMVA operators
    // run on the generic step kernel, which is faster on 32 x 32 tiles than
    // on the specialized forward kernels' 32 x 16 (2-6 % on MI355X).
    const bool onGeneric = is_ext(q.op) || is_tomo(q.op) || is_wemva(q.op);
    if (!is_adj(q.op) && !c->tileFwdSet && onGeneric) {
        c->tz = 32;
        c->tx = 32;
    }
}

void Call::validate() const
{
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
    const int nRecDev0 = q.rec[0]->g.nDev;
    for (int i = 0; i < q.nRec; ++i)
        if (q.rec[i]->g.nDev != nRecDev0) throw Error(AISO_ERR_GEOM, "receivers per shot must be constant");
    const int nSrcDev0 = q.src[0]->g.nDev;
    for (int i = 0; i < q.nShot; ++i)
        if (q.src[i]->g.nDev != nSrcDev0) throw Error(AISO_ERR_GEOM, "sources per shot must be constant");
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
}

void Call::derive_sizes()
{
    s = c->stream;
    nzx = p.nzx();
    nRecDev = q.rec[0]->g.nDev;
    nSrcDev = q.src[0]->g.nDev;
    nExt = is_ext(q.op) ? p.nExt : 1;
    modelN = is_born(q.op) ? nzx * nExt : (long)q.nShotSig * nSrcDev * p.nts;
    // WEMVA: the velocity perturbation and the extended image are its two
    // sides (model in / out differ in size); both directions read data.
    wemva = is_wemva(q.op);
    nImg = is_ext(q.op) || q.op == Op::WemvaFwd ? p.nExt : 1;   // slices of the per-shot image
    perShotImg = c->imagePerShot && (q.op == Op::BornAdj || q.op == Op::ExtAdj);
    mInN = q.op == Op::WemvaAdj ? nzx * p.nExt : modelN;
    mOutN = q.op == Op::WemvaFwd ? nzx * p.nExt : perShotImg ? (long)q.nShot * nImg * nzx : modelN;
    dataInput = is_adj(q.op) || wemva;
    dataN = (long)q.nShot * nRecDev * p.nts;
    h = p.h();
    f = p.filt.data();
    fsurf = p.freeSurface != 0 && !wemva;
    tomo = is_tomo(q.op) || wemva;   // the tomographic and WEMVA passes share the machinery below
    tlag = (is_ext(q.op) || tomo) && p.extension == AISO_EXT_TIME;
}

void Call::stage_operands()
{
    if (q.mem == AISO_MEM_DEVICE) wait_default_stream(c);
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
        float* m = (float*)c->modelDev.need(std::max(modelN, mOutN) * 4);
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
        float* d = (float*)c->dataDev.need(dataN * 4);
        if (dataInput || q.add) HIPC(hipMemcpyAsync(d, dataInput ? q.dataIn : q.dataOut, dataN * 4, hipMemcpyHostToDevice, s));
        dDataIn = d;
        dDataOut = d;
    }
}

void Call::build_geometry()
{
    srcG.resize(q.nShot);
    recG.resize(q.nRec);
    for (int i = 0; i < q.nShot; ++i) {
        srcG[i] = dev_geom(c, q.src[i]);
        nSrcRegMax = std::max(nSrcRegMax, srcG[i]->nReg);
    }
    for (int i = 0; i < q.nRec; ++i) {
        recG[i] = dev_geom(c, q.rec[i]);
        nRecRegMax = std::max(nRecRegMax, recG[i]->nReg);
    }
}

// Shots per launch and the passes' forms: the Born forward with sub == 2
// runs its two propagations in lockstep (the secondary source then lives in
// a 3-slice ring instead of nts slices); with a free surface the nonlinear
// and Born passes run the surface instances of the specialized kernels
// (k_step_fs, k_fwd_pair_fs), without two steps per launch.
void Call::plan_batches()
{
    lockstep = q.op == Op::BornFwd && p.sub == 2 && c->lockstep;
    // Two fine steps per launch (k_fwd_pair2, k_rec2) when no source or
    // receiver point lies in the FAT ring, whose values those kernels take as 0.
    bool noRing = true;
    for (int i = 0; noRing && i < q.nShot; ++i) noRing = srcG[i]->nRing == 0;
    for (int i = 0; noRing && i < q.nRec; ++i) noRing = recG[i]->nRing == 0;
    pair2 = lockstep && c->pair2 && noRing && !fsurf;
    adj2 = q.op == Op::BornAdj && p.sub == 2 && c->pair2 && noRing && !fsurf;
    pairSlices = pair2 || adj2 ? 4 : 2;
    wSlices = lockstep ? 3 : p.nts;
    needW = is_born(q.op);
    needSS = is_ext(q.op) || tomo || q.op == Op::BornAdj;
    // The offset-extended passes keep a ring of coarse-rate slices per shot:
    // extK samples batched per imaging / secondary-source launch, plus the
    // one the time steps are writing or reading.
    extK = std::min(c->extBatch, p.nts);
    ssSlices = is_ext(q.op) || tomo ? extK + 1 : 2;
    needImg = q.op == Op::BornAdj || q.op == Op::ExtAdj || q.op == Op::TomoAdj || wemva;
    // captured coarse samples kept per shot: the WEMVA forward's imaging and
    // the tomographic / WEMVA adjoints' image accumulation take extK at a time
    capSlices = wemva || q.op == Op::TomoAdj ? extK + 1 : 2;
    // Three consecutive fine slices hold everything a coarse sample needs
    // when sub == 2, so captures and imaging are written once per coarse
    // step instead of accumulated at every fine step.
    win = p.sub == 2;
    // The extended-Born, tomographic and WEMVA operators, and every pass with
    // a free surface, step on the unspecialized kernel.  A GPU run of the
    // specialized kernel on the extended-Born path ended in a memory fault
    // whose cause was never shown; the specialized kernel is not built for
    // the extended modes.
    generic = is_ext(q.op) || tomo;

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
    B = is_adj(q.op) ? c->batchAdj : c->batchFwd;
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
}

void Call::allocate()
{
    pairs = (float*)c->pairs.need((size_t)B * pairSlices * nzx * 4);
    W = needW ? (float*)c->W.need((size_t)B * wSlices * nzx * 4) : nullptr;
    pairsQ = lockstep ? (float*)c->pairsQ.need((size_t)B * pairSlices * nzx * 4) : nullptr;
    dSlotsQ = lockstep || adj2 || tomo ? (ShotSlot*)c->slotsQ.need(sizeof(ShotSlot) * B) : nullptr;
    // Tomographic operator: P holds the scattered wavefield of the forward
    // legs or the receiver wavefield of the adjoint (slots dSlotsQ see it as
    // W); cap holds the captured coarse samples (slots dSlotsC).
    Pw = tomo ? (float*)c->P.need((size_t)B * p.nts * nzx * 4) : nullptr;
    cap = tomo ? (float*)c->cap.need((size_t)B * capSlices * nzx * 4) : nullptr;
    dSlotsC = tomo ? (ShotSlot*)c->slotsC.need(sizeof(ShotSlot) * B) : nullptr;
    // WEMVA forward's batched offset imaging: slots that see (ss: captures,
    // W: W'') and (ss: the receiver wavefield, W: captures)
    dSlotsI = q.op == Op::WemvaFwd ? (ShotSlot*)c->slotsI.need(sizeof(ShotSlot) * B) : nullptr;
    dSlotsJ = q.op == Op::WemvaFwd ? (ShotSlot*)c->slotsJ.need(sizeof(ShotSlot) * B) : nullptr;
    ss = needSS ? (float*)c->ss.need((size_t)B * ssSlices * nzx * 4) : nullptr;
    img = needImg ? (float*)c->img.need((size_t)B * nImg * nzx * 4) : nullptr;
    dreg = (float*)c->dreg.need((size_t)B * nRecRegMax * p.nts * 4);
    sig = (float*)c->sig.need((size_t)B * nSrcRegMax * p.ntw * 4);
    dSlots = (ShotSlot*)c->slots.need(sizeof(ShotSlot) * B);
}

void Call::prepare_model()
{
    if (q.op == Op::BornFwd) {
        float* m = (float*)c->mtil.need(nzx * 4);
        HIPC(launch_scale_model(dModelIn, m, c->v2dt2.as<float>(), c->refl.as<float>(), p.nz, p.nx, s));
        mtil = m;
    }
    // Per-point scale of the extended model: 2/v^3 for offsets; for time
    // lags also v^2 dtw^2, or -v^2 dtw^2 with sloth (which neither the
    // reference's free-surface routines nor its WEMVA take).
    extScale = c->refl.as<float>();
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
    if (q.op == Op::NlAdj && q.nShotSig == 1) nlAcc.assign((size_t)nSrcDev * p.nts, 0.f);
}

void Call::init_step_args()
{
    base = StepArgs{};
    base.slots = dSlots;
    base.v2dt2 = c->v2dt2.as<float>();
    base.cd = c->cd.as<float>();
    base.mtil = mtil;
    base.nz = p.nz; base.nx = p.nx; base.minPad = p.minPad; base.nts = p.nts; base.ntw = p.ntw;
    base.tilesZ = tiles_z(c);
    base.ntiles = base.tilesZ * tiles_x(c);
    for (int k = 0; k < 6; ++k) { base.zc[k] = p.zc[k]; base.xc[k] = p.xc[k]; }
    base.wA = base.wB = base.wImg = 0;
    base.ntW = c->streamNT;
    base.fs = fsurf;
}

// Time lags of a batch: sample k uses slices e in [lo, hi), those whose W
// sample k - dir * 2 (e - hExt) lies in [0, nts)
// (F/src/BornExtShotsGpuFunctions.cu:362-366, 374-378 for dir = +1;
// F/src/tomoExtShotsGpuFunctions.cu:2630-2633 for dir = -1).
void Call::time_lags(ExtBatchArgs& e, int dir) const
{
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
}

// Sparse operations on devices inside the FAT ring (k_ring), after the
// step launch; true if it launched anything.
bool Call::ring(unsigned mode, const StepArgs& sa)
{
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
}

// ------------------------------------------------------------ per batch

// Slot descriptors and source traces of shots [first, first + count).
void Call::begin_batch(int first, int count)
{
    b0 = first;
    nb = count;
    batch0 = nullptr;
    if (c->profile) {
        batch0 = pooled_event(c);
        HIPC(hipEventRecord(batch0, s));
    }
    slots.assign(nb, ShotSlot{});
    maxSrcRing = maxRecRing = 0;
    sigHost.assign((size_t)nb * nSrcRegMax * p.ntw, 0.f);
    for (int k = 0; k < nb; ++k) {
        const int shot = b0 + k;
        const aiso_device* sd = q.src[shot];
        DevGeom* sg = srcG[shot];
        DevGeom* rg = rec_geom(shot);
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
    a = base;
    a.nshots = nb;
    a.gridLogical = a.ntiles * nb;
    c->curLanes = pair2 || adj2 ? 1 : std::max(1, std::min(c->lanes, nb));
    if (c->curLanes > 1) ensure_lanes(c);
}

// Every slot table the batch's passes read, copied once per batch through
// the pinned stage: the shots' own slots, and the views of the same shots
// the lockstep, two-step, tomographic and WEMVA passes launch with.
void Call::upload_tables()
{
    const size_t tb = sizeof(ShotSlot) * nb;
    std::vector<ShotSlot> slotsQ, slotsP, slotsCv, slotsIv, slotsJv;
    if (tomo) {   // the same shots seeing P, or the captured samples, as W
        slotsP = slots;
        slotsCv = slots;
        for (int k = 0; k < nb; ++k) {
            slotsP[k].W = Pw + (size_t)k * p.nts * nzx;
            slotsCv[k].W = cap + (size_t)k * capSlices * nzx;
        }
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
    }
    if (pair2 || adj2) {
        // two-step kernels: source points listed for each tile's extended
        // region; receivers for the owning tile (forward) or the region (adjoint)
        slotsQ = slots;
        for (int k = 0; k < nb; ++k) {
            const int shot = b0 + k;
            DevGeom* sgE = dev_geom_pair2(c, q.src[shot], true);
            DevGeom* rgE = dev_geom_pair2(c, rec_dev(shot), adj2);
            slotsQ[k].srcTile = sgE->tilePtr.as<int>();
            slotsQ[k].srcEnt = sgE->ent.as<int2>();
            slotsQ[k].recTile = rgE->tilePtr.as<int>();
            slotsQ[k].recEnt = rgE->ent.as<int2>();
        }
    } else if (lockstep) {   // the scattered field seen as "the" pair, for single-field launches
        slotsQ = slots;
        for (auto& S : slotsQ) { S.p[0] = S.q[0]; S.p[1] = S.q[1]; }
    }
    const bool sigUp = q.op != Op::NlAdj;
    size_t bytes = stage_size(tb);
    if (!slotsQ.empty()) bytes += stage_size(tb);
    if (tomo) bytes += 2 * stage_size(tb);
    if (q.op == Op::WemvaFwd) bytes += 2 * stage_size(tb);
    if (sigUp) bytes += stage_size(sigHost.size() * 4);
    stage_begin(c, bytes);
    stage_copy(c, dSlots, slots.data(), tb);
    if (!slotsQ.empty()) stage_copy(c, dSlotsQ, slotsQ.data(), tb);
    if (tomo) {
        stage_copy(c, dSlotsQ, slotsP.data(), tb);
        stage_copy(c, dSlotsC, slotsCv.data(), tb);
    }
    if (q.op == Op::WemvaFwd) {
        stage_copy(c, dSlotsI, slotsIv.data(), tb);
        stage_copy(c, dSlotsJ, slotsJv.data(), tb);
    }
    if (sigUp) stage_copy(c, sig, sigHost.data(), sigHost.size() * 4);
    else HIPC(launch_fill(sig, (long)nb * nSrcRegMax * p.ntw, s));
    stage_end(c);
}

// Adjoint (and WEMVA): data traces to the regular receiver grid; forward:
// zeroed regular traces to record into.
void Call::load_receiver_traces()
{
    if (dataInput) {
        for (int k = 0; k < nb; ++k) {
            const int shot = b0 + k;
            DevGeom* rg = rec_geom(shot);
            HIPC(launch_dev_adjoint(dreg + (size_t)k * nRecRegMax * p.nts, dDataIn + (size_t)shot * nRecDev * p.nts,
                                    rg->adjPtr.as<int>(), rg->adjEnt.as<int>(), rg->weight.as<float>(),
                                    rg->nReg, rec_dev(shot)->g.nb, p.nts, s));
        }
    } else {
        HIPC(launch_fill(dreg, (long)nb * nRecRegMax * p.nts, s));
    }
}

// Regular-grid traces of the batch to the caller's data rows (deviceGpu::forward).
void Call::record_data()
{
    for (int k = 0; k < nb; ++k) {
        const int shot = b0 + k;
        DevGeom* rg = rec_geom(shot);
        HIPC(launch_dev_forward(dreg + (size_t)k * nRecRegMax * p.nts, dDataOut + (size_t)shot * nRecDev * p.nts,
                                rg->cornerReg.as<int>(), rg->weight.as<float>(), nRecDev, rec_dev(shot)->g.nb, p.nts,
                                q.add, s));
    }
}

// Lockstep Born forward, two fine steps per launch (k_fwd_pair2): the
// scattered field lags the source field by one launch; SS[k] lives in a
// 3-slice ring of W.
void Call::pass_pair2()
{
    HIPC(launch_fill(pairs, (long)nb * 4 * nzx, s));
    HIPC(launch_fill(pairsQ, (long)nb * 4 * nzx, s));
    auto ring3 = [&](long k) { return (k % 3) * nzx; };
    Pair2Args a2{};
    a2.slots = dSlotsQ;
    a2.v2dt2 = base.v2dt2;
    a2.cd = base.cd;
    a2.mtil = mtil;
    a2.nz = p.nz; a2.nx = p.nx; a2.minPad = p.minPad; a2.nts = p.nts; a2.ntw = p.ntw;
    a2.tilesZ = tiles_z2(c);
    a2.ntiles = a2.tilesZ * tiles_x2(c);
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
    record_data();
}

// Lockstep Born forward (k_fwd_pair): one launch advances the source field
// by fine step t and the scattered field by step t - 2.
void Call::pass_lockstep()
{
    HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
    HIPC(launch_fill(pairsQ, (long)nb * 2 * nzx, s));
    const int T = p.ntw - 1;   // fine steps of each propagation
    auto ring3 = [&](long k) { return (k % 3) * nzx; };
    int curS = 1, curQ = 1;
    a.winOut = WIN_SS;
    a.wa = f[h + 1];
    a.wb = f[h + 2];
    a.wc = f[1];
    hipEvent_t passT = pass_timer_begin(c);
    fork_lanes(c);
    // With sampled profiling on one lane, the pair kernel is timed over
    // windows of profEvery consecutive launches between two plain events
    // (the average time per launch, back to back); a window that another
    // kernel entered is dropped.
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
            launch_timed(c, MODE_SRC_WIN, AISO_K_SRC, x);
            ring(MODE_SRC_WIN, x);
        } else {
            if (windows) close_window();
            StepArgs xq = x;
            xq.slots = dSlotsQ;
            xq.cur = curQ;
            xq.wA = x.wQa;
            xq.wB = x.wQb;
            launch_timed(c, MODE_BORN_SCAT_PRE, AISO_K_BORN_SCAT, xq);
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
    pass_timer_end(c, passT, AISO_K_BORN_PAIR, 2.0 * double(p.interior()) * nb * (T - 2), T - 2);
    record_data();
}

// Born adjoint, two fine steps per launch each: source pass (W[k] for every
// coarse k) then receiver pass with imaging.
void Call::pass_adjoint2()
{
    HIPC(launch_fill(pairs, (long)nb * 4 * nzx, s));
    Pair2Args a2{};
    a2.slots = dSlotsQ;
    a2.v2dt2 = base.v2dt2;
    a2.cd = base.cd;
    a2.mtil = nullptr;                            // raw W
    a2.nz = p.nz; a2.nx = p.nx; a2.minPad = p.minPad; a2.nts = p.nts; a2.ntw = p.ntw;
    a2.tilesZ = tiles_z2(c);
    a2.ntiles = a2.tilesZ * tiles_x2(c);
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

// Source wavefield pass of every Born-type operator: W[k] for every coarse
// sample k (F/src/BornShotsGpuFunctions.cu:296-321); the Born forward's W
// holds the secondary source m~ * W instead.
void Call::pass_source()
{
    HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
    a.cur = 1;
    const unsigned mode = win ? MODE_SRC_WIN : MODE_SRC;
    const int winOut = q.op == Op::BornFwd ? WIN_SS : WIN_W;
    a.winOut = winOut;
    a.wa = f[h + 1];
    a.wb = f[h + 2 < 2 * h ? h + 2 : h + 1];
    a.wc = f[1];
    hipEvent_t passT = pass_timer_begin(c);
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
            launch_timed(c, mode, AISO_K_SRC, a, generic);
            ring(mode, a);
            a.cur ^= 1;
        }
    join_lanes(c);
    pass_timer_end(c, passT, AISO_K_SRC, double(p.interior()) * nb * (p.ntw - 1), p.ntw - 1);
    if (win)   // W[nts-1] = p(last-1)*f[h+1] + p(last)*f[h+2]
        HIPC(launch_win_tail(dSlots, nb, a.cur, winOut, f[h + 1], f[h + 2], (long)(p.nts - 1) * nzx, 0, 0, mtil,
                             p.nz, p.nx, s));
    a.win = 0;
}

// Tomographic and WEMVA operators: both legs, forward or adjoint, after the
// source pass (F/src/tomoExtShotsGpuFunctions.cu:467-572, 912-1020;
// F/src/wemvaExtShotsGpuFunctions.cu:458-574, 689-808).
void Call::passes_tomo()
{
    const bool haveAll = ss && Pw && cap && dSlotsQ && dSlotsC && (!needImg || img) &&
                         (q.op != Op::WemvaFwd || (dSlotsI && dSlotsJ)) &&
                         ((q.op != Op::TomoFwd && q.op != Op::WemvaFwd) || mtil) &&
                         ((!is_tomo(q.op) && q.op != Op::WemvaAdj) || rExt);
    if (!haveAll) throw Error(AISO_ERR_STATE, "internal: a work buffer of the tomographic / WEMVA passes is missing");
    // W <- its coarse second time derivative, in place
    const float d2 = p.dts * p.dts;
    HIPC(launch_dt2_coarse(dSlots, nb, p.nts, (float)(1.0 / (double)d2), (float)(-2.0 / (double)d2), p.nz, p.nx, s));
    // Lanes: inside a pass the batch's shots run on the context's lanes;
    // every launch in the time loop goes to each lane with its own shots.
    bool forked = false;
    auto lanes = [&](const std::function<void(int, int, hipStream_t)>& fn) {
        const int L = forked ? c->curLanes : 1;
        for (int l = 0; l < L; ++l) {
            int k0, cnt;
            lane_range(nb, L, l, k0, cnt);
            if (cnt > 0) fn(k0, cnt, L > 1 ? lane_stream(c, l) : s);
        }
    };
    hipEvent_t passT = nullptr;
    auto fork = [&] {
        passT = pass_timer_begin(c);
        fork_lanes(c);
        forked = c->curLanes > 1;
    };
    auto join = [&] {
        join_lanes(c);
        forked = false;
        pass_timer_end(c, passT, AISO_K_TOMO, double(p.interior()) * nb * (p.ntw - 1), p.ntw - 1);
    };
    enum { SS_EXT, SS_MUL, SS_ADJ };   // ring producers
    auto produce = [&](int kind, const ShotSlot* from, const float* scale, int k0, int n) {
        ExtBatchArgs e{};
        e.slots = from;
        e.v2dt2 = base.v2dt2;
        e.mext = scale;
        e.nz = p.nz; e.nx = p.nx; e.hExt = p.hExt;
        e.nk = n;
        for (int j = 0; j < n; ++j) {
            e.ssOff[j] = ss_slot(k0 + j);
            e.wOff[j] = (long)(k0 + j) * nzx;
        }
        if (kind != SS_MUL && tlag) time_lags(e, kind == SS_EXT ? 1 : -1);
        lanes([&](int s0, int cnt, hipStream_t st) {
            ExtBatchArgs el = e;
            el.slots = from + s0;
            if (kind == SS_MUL) HIPC(launch_mul_batch(el, cnt, st));
            else if (tlag) HIPC(launch_time_ss_batch(el, cnt, st));   // imagingTimeFwdGpu / imagingTimeTomoAdjGpu
            else if (kind == SS_EXT) HIPC(launch_ext_ss_batch(el, cnt, st));
            else HIPC(launch_tomo_adj_ss(el, cnt, st));
        });
    };
    // Forward in time.  ssKind >= 0: inject ring samples made from `from`'s
    // W (scaled by `scale`); capture: 1 = into the launch slots' W at the
    // sample index (nts slices), 2 = into the capture ring; done(k) runs
    // once coarse sample k is complete.
    auto forward_pass = [&](unsigned mode, const ShotSlot* on, int ssKind, const ShotSlot* from, const float* scale,
                            int capture, const std::function<void(int)>& done) {
        HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
        StepArgs x = a;
        x.slots = on;
        x.cur = 1;
        x.win = 0;
        int next = 0;
        auto off = [&](int k) { return capture == 1 ? (long)k * nzx : cap_off(k); };
        // sub == 2: sample its written once at the first fine step of its
        // interval from the old, current and new slices (as the source pass)
        const bool wcap = capture && win && mode == MODE_TOMO_CAP;
        if (wcap) {
            mode = MODE_TOMO_CAP_WIN;
            x.winOut = WIN_W;
            x.wa = f[h + 1];
            x.wb = f[h + 2 < 2 * h ? h + 2 : h + 1];
            x.wc = f[1];
        }
        fork();
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
                    x.ssOffA = ss_slot(its);
                    x.ssOffB = ss_slot(its + 1);
                }
                if (wcap) {
                    x.wA = off(its);
                    x.win = it2 == 1;
                } else if (capture) {
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
                launch_timed(c, mode, AISO_K_TOMO, x, true);
                ring(mode, x);
                x.cur ^= 1;
            }
            if (done) done(its);
        }
        if (wcap)   // sample nts-1 = p(last-1)*f[h+1] + p(last)*f[h+2]
            lanes([&](int s0, int cnt, hipStream_t st) {
                HIPC(launch_win_tail(on + s0, cnt, x.cur, WIN_W, f[h + 1], f[h + 2], off(p.nts - 1), 0, 0, nullptr,
                                     p.nz, p.nx, st));
            });
        if (done) done(p.nts - 1);
        join();
    };
    // Backward in time on the adjoint stepper from ring samples made from
    // `from`'s W, injected with the weights of fine index it2 + 1, captured
    // into the ring; done(k) once coarse sample k is complete.
    auto backward_pass = [&](int ssKind, const ShotSlot* from, const float* scale,
                             const std::function<void(int)>& done) {
        HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
        HIPC(launch_fill(cap, (long)nb * capSlices * nzx, s));
        StepArgs y = a;
        y.slots = dSlotsC;
        y.cur = 1;
        y.win = 0;
        int next = p.nts - 1;   // highest sample not yet in the ring
        // sub == 2: sample its + 1 written once at the first fine step of
        // interval its (in backward order), as the Born receiver pass
        unsigned mode = MODE_TOMO_ADJ_CAP;
        if (win) {
            mode = MODE_TOMO_ADJ_CAP_WIN;
            y.winOut = WIN_W;
            y.wa = f[1];
            y.wb = f[0];
            y.wc = f[h + 1];
        }
        fork();
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
                y.ssOffA = ss_slot(its);
                y.ssOffB = ss_slot(its + 1);
                y.wA = win ? cap_off(its + 1) : cap_off(its);
                y.wB = cap_off(its + 1);
                y.win = it2 == 1;
                y.fe = f[it2];
                y.ff = f[h + it2];
                y.skipA = y.fe == 0.f;
                y.skipB = y.ff == 0.f;
                y.assignA = it2 == p.sub - 1;
                y.assignB = 0;
                launch_timed(c, mode, AISO_K_TOMO, y, true);
                y.cur ^= 1;
            }
            done(its + 1);
        }
        if (win)   // sample 0 = q(1)*f[1] + q(0)*f[0]
            lanes([&](int s0, int cnt, hipStream_t st) {
                HIPC(launch_win_tail(dSlotsC + s0, cnt, y.cur, WIN_W, f[1], f[0], cap_off(0), 0, 0, nullptr, p.nz,
                                     p.nx, st));
            });
        done(0);
        join();
    };
    // img += X[k] * captured sample k, for the samples in the order they
    // complete, extK of them per launch (the capture ring holds extK + 1:
    // the pending samples and the one being captured); `last` is the pass's
    // final sample.
    auto image_with = [&](const ShotSlot* X, int last) {
        auto pend = std::make_shared<std::vector<int>>();
        return [&, X, last, pend](int k) {
            pend->push_back(k);
            if ((int)pend->size() < extK && k != last) return;
            ImgAccArgs g{};
            g.nz = p.nz;
            g.nx = p.nx;
            g.nk = (int)pend->size();
            for (int j = 0; j < g.nk; ++j) {
                g.xOff[j] = (long)(*pend)[j] * nzx;
                g.yOff[j] = cap_off((*pend)[j]);
            }
            pend->clear();
            lanes([&](int s0, int cnt, hipStream_t st) {
                ImgAccArgs h = g;
                h.X = X + s0;
                h.Y = dSlotsC + s0;
                HIPC(launch_img_acc_batch(h, cnt, st));
            });
        };
    };
    if (dataInput) {   // receiver wavefield R into P (computeTomoRecWfld, :1481-1515)
        HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
        StepArgs x = a;
        x.slots = dSlotsQ;
        x.cur = 1;
        x.win = 0;
        // sub == 2: R[its + 1] written once at the first fine step of interval its
        unsigned mode = MODE_TOMO_REC;
        if (win) {
            mode = MODE_REC_WIN;
            x.winOut = WIN_W;
            x.wa = f[1];
            x.wb = f[0];
            x.wc = f[h + 1];
        }
        fork();
        for (int its = p.nts - 2; its >= 0; --its)
            for (int it2 = p.sub - 1; it2 >= 0; --it2) {
                x.its = its;
                x.itwRec = its * p.sub + it2;
                x.fc = f[it2 + 1];
                x.fd = f[h + it2 + 1];
                x.wA = win ? (long)(its + 1) * nzx : (long)its * nzx;
                x.wB = (long)(its + 1) * nzx;
                x.win = it2 == 1;
                x.fe = f[it2];
                x.ff = f[h + it2];
                x.skipA = x.fe == 0.f;
                x.skipB = x.ff == 0.f;
                x.assignA = it2 == p.sub - 1;
                x.assignB = its == p.nts - 2 && it2 == p.sub - 1;
                launch_timed(c, mode, AISO_K_TOMO, x, true);
                ring(mode, x);
                x.cur ^= 1;
            }
        if (win)   // R[0] = q(1)*f[1] + q(0)*f[0]
            lanes([&](int s0, int cnt, hipStream_t st) {
                HIPC(launch_win_tail(dSlotsQ + s0, cnt, x.cur, WIN_W, f[1], f[0], 0, 0, 0, nullptr, p.nz, p.nx, st));
            });
        join();
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
        record_data();
    } else if (q.op == Op::TomoAdj || q.op == Op::WemvaAdj) {
        // The same two passes: forward from the extended scattering of W''
        // imaged against R, and backward from the adjoint extended
        // scattering of R imaged against W''; the tomographic adjoint runs
        // them as legs 1, 2, WEMVA's as legs 2, 1.
        auto fwdLeg = [&] { forward_pass(MODE_TOMO_CAP, dSlotsC, SS_EXT, dSlots, rExt, 2, image_with(dSlotsQ, p.nts - 1)); };
        auto bwdLeg = [&] { backward_pass(SS_ADJ, dSlotsQ, rExt, image_with(dSlots, 0)); };
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
                e.ssOff[j] = rSide ? (long)k * nzx : cap_off(k);   // receiver side: R[k] / sample k
                e.wOff[j] = rSide ? cap_off(k) : (long)k * nzx;    // source side: sample k / W''[k] (R[k])
            }
            if (tlag) time_lags(e, wemvaFwd ? -1 : 1);   // imagingWemvaTimeAdjGpu / imagingTimeAdjGpu
            const ShotSlot* base0 = e.slots;
            lanes([&](int s0, int cnt, hipStream_t st) {
                ExtBatchArgs el = e;
                el.slots = base0 + s0;
                if (tlag) HIPC(launch_time_img_batch(el, cnt, st));
                else HIPC(launch_ext_img_batch(el, cnt, st, wemvaFwd));
            });
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

// Second pass of the forward operators: the nonlinear propagation, or the
// Born / extended-Born scattered field, with the data recorded.
void Call::pass_forward()
{
    HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
    a.cur = 1;
    a.assignA = a.assignB = a.skipA = a.skipB = 0;
    unsigned mode = MODE_NL_FWD;
    int kind = AISO_K_NL_FWD;
    if (q.op == Op::BornFwd) { mode = win ? MODE_BORN_SCAT_PRE : MODE_BORN_SCAT; kind = AISO_K_BORN_SCAT; }
    if (q.op == Op::ExtFwd) { mode = MODE_EXT_SCAT; kind = AISO_K_EXT_SCAT; }
    int L = 0, R = 1;
    // Secondary sources of the offset-extended forward, extK samples per
    // launch into the ring, ahead of their use.
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
                e.ssOff[j] = ss_slot(nextSS + j);
                e.wOff[j] = (long)(nextSS + j) * nzx;
            }
            if (tlag) time_lags(e, 1);
            for (int l = 0; l < c->curLanes; ++l) {   // each lane's shots, on its stream
                int k0, cnt;
                lane_range(nb, c->curLanes, l, k0, cnt);
                if (cnt == 0) continue;
                ExtBatchArgs el = e;
                el.slots = dSlots + k0;
                if (tlag) HIPC(launch_time_ss_batch(el, cnt, lane_stream(c, l)));
                else HIPC(launch_ext_ss_batch(el, cnt, lane_stream(c, l)));
            }
            nextSS += e.nk;
        }
    };
    hipEvent_t passT = pass_timer_begin(c);
    fork_lanes(c);
    for (int its = 0; its < p.nts - 1; ++its) {
        if (q.op == Op::ExtFwd) {
            ext_ss_upto(its + 1);
            L = R = 0;
            a.ssOffA = ss_slot(its);
            a.ssOffB = ss_slot(its + 1);
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
            launch_timed(c, mode, kind, a, generic);
            ring(mode, a);
            a.cur ^= 1;
        }
        std::swap(L, R);
    }
    join_lanes(c);
    pass_timer_end(c, passT, kind, double(p.interior()) * nb * (p.ntw - 1), p.ntw - 1);
    record_data();
}

// Second pass of the adjoint operators: the nonlinear adjoint propagation,
// or the receiver field of the Born / extended-Born adjoint with its imaging
// (F/src/BornShotsGpuFunctions.cu:893-945).
void Call::pass_adjoint()
{
    HIPC(launch_fill(pairs, (long)nb * 2 * nzx, s));
    a.cur = 1;
    a.assignA = a.assignB = a.skipA = a.skipB = 0;
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
    // The stored source wavefield lets the receiver pass run the batch in
    // smaller groups, whose per-shot slices (pair + image) stay in the
    // on-die cache.
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
        // Completed receiver-wavefield samples awaiting the offset imaging,
        // newest last; imaged extK at a time.
        std::vector<int> pend;
        int lanesNow = c->curLanes;
        auto ext_img_flush = [&] {
            if (pend.empty()) return;
            ExtBatchArgs e{};
            e.slots = sl;
            e.v2dt2 = base.v2dt2;
            e.nz = p.nz; e.nx = p.nx; e.hExt = p.hExt;
            e.nk = (int)pend.size();
            for (int j = 0; j < e.nk; ++j) {
                e.ssOff[j] = ss_slot(pend[j]);
                e.wOff[j] = (long)pend[j] * nzx;
            }
            if (tlag) time_lags(e, 1);
            // Inside the time loop each lane images its own shots on its
            // stream; after the lanes join (lanesNow 1) the whole group on s.
            for (int l = 0; l < lanesNow; ++l) {
                int k0, cnt;
                lane_range(nbr, lanesNow, l, k0, cnt);
                if (cnt == 0) continue;
                ExtBatchArgs el = e;
                el.slots = sl + k0;
                hipStream_t ls = lanesNow > 1 ? lane_stream(c, l) : s;
                if (tlag) HIPC(launch_time_img_batch(el, cnt, ls));
                else HIPC(launch_ext_img_batch(el, cnt, ls));
            }
            pend.clear();
        };
        hipEvent_t passT = pass_timer_begin(c);
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
                    r.ssOffA = ss_slot(win ? its + 1 : its);
                    r.ssOffB = ss_slot(its + 1);
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
                launch_timed(c, mode, kind, r, generic);
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
        lanesNow = 1;
        pass_timer_end(c, passT, kind, double(p.interior()) * nbr * (p.ntw - 1), p.ntw - 1);
        if (q.op == Op::BornAdj) {
            if (win) {   // img += W[0] * (q(1)*f[1] + q(0)*f[0]), then the scaling
                HIPC(launch_win_tail(sl, nbr, r.cur, WIN_IMG, f[1], f[0], 0, 0, 0, nullptr, p.nz, p.nx, s));
                HIPC(launch_born_adj_finish(sl, nbr, -1, base.v2dt2, c->refl.as<float>(), p.nz, p.nx, s));
            } else {
                HIPC(launch_born_adj_finish(sl, nbr, R, base.v2dt2, c->refl.as<float>(), p.nz, p.nx, s));
            }
        } else if (q.op == Op::ExtAdj) {
            if (win)
                HIPC(launch_win_tail(sl, nbr, r.cur, WIN_SLICE, f[1], f[0], ss_slot(0), 0, 0, nullptr, p.nz, p.nx, s));
            pend.push_back(0);
            ext_img_flush();
            HIPC(launch_ext_adj_finish(sl, nbr, extScale, p.nz, p.nx, nExt, s));
        }
    }
}

// Per-shot images to the caller (one image per shot), or summed in shot
// order into the call's image.
void Call::reduce_images()
{
    if (!needImg) return;
    if (perShotImg) {
        for (int k = 0; k < nb; ++k)
            HIPC(launch_add_to(dModelOut + (size_t)(b0 + k) * nImg * nzx, img + (size_t)k * nImg * nzx, nzx * nImg,
                               q.add, s));
    } else {
        for (int k = 0; k < nb; ++k) HIPC(launch_accumulate(acc, img + (size_t)k * nImg * nzx, nzx * nImg, s));
    }
}

// Nonlinear adjoint: the regular-grid source traces back through the host
// chain (F/src/nonlinearPropGpu.cpp:114-120): time-interpolation adjoint,
// source scaling, deviceGpu::forward.
void Call::finish_nl_adjoint()
{
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

// Wavefield saving (the reference's forwardWavefield / adjointWavefield for
// shot wavefieldShotNumber, F/src/BornShotsGpu.cpp:145-240, 341-438,
// F/src/nonlinearPropShotsGpu.cpp:153-250, 360-458): shot `wavefieldShot`
// of the call runs once more, alone, on the generic step kernel, capturing
// the coarse wavefields [nts][nx][nz] the reference keeps:
//   nonlinear forward / adjoint    the propagated field (getWfld)
//   Born-type operators            the source wavefield (getSrcWfld)
//   Born forward                   + the scattered wavefield (getSecWfld)
//   Born adjoint                   + the receiver wavefield scaled like
//                                    the image (recordScaleWavefield)
// The call's outputs come from its regular passes; this pass only fills
// the context's wavefield buffers, in the reference's arithmetic.
void Call::save_wavefields()
{
    const int ws = c->wavefieldShot;
    if (ws < 0 || ws >= q.nShot) throw Error(AISO_ERR_PARAM, "wavefieldShotNumber is not a shot of the call");
    c->wfSrcValid = c->wfSecValid = 0;
    const bool secFwd = q.op == Op::BornFwd, secAdj = q.op == Op::BornAdj;
    if (tomo) return;   // the reference's tomographic and WEMVA operators save no wavefields
    float* wsrc = (float*)c->wfSrc.need((size_t)p.nts * nzx * 4);
    float* wsec = secFwd || secAdj ? (float*)c->wfSec.need((size_t)p.nts * nzx * 4) : nullptr;
    float* scr = (float*)c->wfScratch.need((size_t)(secAdj ? 5 : 2) * nzx * 4);   // pair (+ ss pair + image)
    // the captures write the interior; the FAT ring of a saved slice reads 0,
    // as the reference's zeroed wavefield buffers do
    HIPC(launch_fill(wsrc, (long)p.nts * nzx, s));
    if (wsec) HIPC(launch_fill(wsec, (long)p.nts * nzx, s));
    b0 = ws;
    nb = 1;
    const aiso_device* sd = q.src[ws];
    DevGeom* sg = srcG[ws];
    DevGeom* rg = rec_geom(ws);
    ShotSlot S;
    std::memset(&S, 0, sizeof(S));
    S.p[0] = scr;
    S.p[1] = scr + nzx;
    S.W = wsrc;
    if (secAdj) {
        S.ss[0] = scr + 2 * nzx;
        S.ss[1] = scr + 3 * nzx;
        S.img = scr + 4 * nzx;
    }
    if (secFwd) S.img = wsec;
    S.dreg = dreg;
    S.sig = sig;
    S.srcTile = sg->tilePtr.as<int>();
    S.srcEnt = sg->ent.as<int2>();
    S.srcRing = sg->ring.as<int2>();
    S.nSrcRing = sg->nRing;
    S.recTile = rg->tilePtr.as<int>();
    S.recEnt = rg->ent.as<int2>();
    S.recRing = rg->ring.as<int2>();
    S.nRecRing = rg->nRing;
    maxSrcRing = sg->nRing;
    maxRecRing = rg->nRing;
    std::vector<float> fine;
    if (is_born(q.op)) {
        fine = born_source(c, sd->g, q.wavelets + (size_t)(q.nWav == 1 ? 0 : ws) * nSrcDev * p.nts);
    } else if (q.op == Op::NlFwd) {
        fine = nl_source(c, sd->g, hModel.data() + (size_t)(q.nShotSig == 1 ? 0 : ws) * nSrcDev * p.nts);
    }
    stage_begin(c, stage_size(sizeof(ShotSlot)) + stage_size(fine.size() * 4));
    stage_copy(c, dSlots, &S, sizeof(ShotSlot));
    stage_copy(c, sig, fine.data(), fine.size() * 4);
    stage_end(c);
    HIPC(launch_fill(scr, (long)(secAdj ? 5 : 2) * nzx, s));
    if (dataInput) {
        HIPC(launch_dev_adjoint(dreg, dDataIn + (size_t)ws * nRecDev * p.nts, rg->adjPtr.as<int>(), rg->adjEnt.as<int>(),
                                rg->weight.as<float>(), rg->nReg, rec_dev(ws)->g.nb, p.nts, s));
    }
    c->curLanes = 1;
    StepArgs x = base;
    x.slots = dSlots;
    x.nshots = 1;
    x.gridLogical = x.ntiles;
    x.win = 0;
    // a forward pass capturing into W (mode MODE_SRC) or img (MODE_BORN_SCAT_CAP)
    auto forward = [&](unsigned mode) {
        HIPC(launch_fill(scr, 2 * nzx, s));
        x.cur = 1;
        for (int its = 0; its < p.nts - 1; ++its)
            for (int it2 = 1; it2 <= p.sub; ++it2) {
                x.its = its;
                x.itwInj = its * p.sub + it2 - 1;
                x.fa = f[it2 - 1];
                x.fb = f[h + it2 - 1];
                x.wA = (long)its * nzx;
                x.wB = (long)(its + 1) * nzx;
                x.fe = f[it2];
                x.ff = f[h + it2];
                x.skipA = x.fe == 0.f;
                x.skipB = x.ff == 0.f;
                x.assignA = its == 0 && it2 == 1;
                x.assignB = it2 == 1;
                HIPC(launch_step_generic(mode, c->tz, c->tx, x, s));
                ring(mode, x);
                x.cur ^= 1;
            }
    };
    if (q.op != Op::NlAdj) forward(MODE_SRC);          // NL wavefield, or the source wavefield
    if (secFwd) {                                        // scattered wavefield
        x.mtil = mtil;
        forward(MODE_BORN_SCAT_CAP);
    }
    if (q.op == Op::NlAdj || secAdj) {
        // backward in time from the data: the nonlinear adjoint field into W,
        // or the Born receiver field into the ss pair, imaged against W and
        // recorded scaled after each coarse step
        HIPC(launch_fill(scr, 2 * nzx, s));
        const unsigned mode = secAdj ? MODE_BORN_REC : MODE_TOMO_REC;
        x.cur = 1;
        int L = 0, R = 1;
        for (int its = p.nts - 2; its >= 0; --its) {
            for (int it2 = p.sub - 1; it2 >= 0; --it2) {
                x.its = its;
                x.itwRec = its * p.sub + it2;
                x.fc = f[it2 + 1];
                x.fd = f[h + it2 + 1];
                x.fe = f[it2];
                x.ff = f[h + it2];
                x.skipA = x.fe == 0.f;
                x.skipB = x.ff == 0.f;
                if (secAdj) {
                    x.ssA = L;
                    x.ssB = R;
                    x.ssOffA = x.ssOffB = 0;
                    x.assignA = it2 == p.sub - 1;
                    x.assignB = 0;
                    x.image = it2 == 0;
                    x.wImg = (long)(its + 1) * nzx;
                } else {
                    x.wA = (long)its * nzx;
                    x.wB = (long)(its + 1) * nzx;
                    x.assignA = it2 == p.sub - 1;
                    x.assignB = its == p.nts - 2 && it2 == p.sub - 1;
                }
                HIPC(launch_step_generic(mode, c->tz, c->tx, x, s));
                ring(mode, x);
                x.cur ^= 1;
            }
            if (secAdj) {
                HIPC(launch_record_scale(dSlots, R, wsec + (size_t)(its + 1) * nzx, c->refl.as<float>(), base.v2dt2,
                                         p.nz, p.nx, s));
                std::swap(L, R);
            }
        }
        if (secAdj)
            HIPC(launch_record_scale(dSlots, R, wsec, c->refl.as<float>(), base.v2dt2, p.nz, p.nx, s));
    }
    c->wfSrcValid = 1;
    c->wfSecValid = wsec != nullptr;
}

void Call::end_batch()
{
    if (c->profile) {
        hipEvent_t batch1 = pooled_event(c);
        HIPC(hipEventRecord(batch1, s));
        c->batchEv.push_back({batch0, batch1, nb});
    }
    drain_profile(c);
}

void Call::outputs()
{
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

void Call::run()
{
    select_tile();
    validate();
    derive_sizes();
    stage_operands();
    build_geometry();
    plan_batches();
    allocate();
    prepare_model();
    init_step_args();
    for (int first = 0; first < q.nShot; first += B) {
        begin_batch(first, std::min(B, q.nShot - first));
        upload_tables();
        load_receiver_traces();
        if (pair2) pass_pair2();
        else if (lockstep) pass_lockstep();
        if (adj2) pass_adjoint2();
        if (is_born(q.op) && !lockstep && !adj2) pass_source();
        if (tomo) passes_tomo();
        if (!lockstep && !adj2 && !tomo) {
            if (q.op == Op::NlFwd || q.op == Op::BornFwd || q.op == Op::ExtFwd) pass_forward();
            else pass_adjoint();
        }
        reduce_images();
        if (q.op == Op::NlAdj) finish_nl_adjoint();
        end_batch();
    }
    if (c->saveWavefield) save_wavefields();
    outputs();
}

}  // namespace

void run_call(aiso_ctx* c, const Request& q)
{
    Call call(c, q);
    call.run();
}

}  // namespace aiso
