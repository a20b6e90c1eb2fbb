// kernels.hpp -- argument blocks shared by the HIP kernels and their host
// launchers.  Plain structs passed by value (kernel arguments) or through a
// small device array (one ShotSlot per shot advanced by a launch).
#pragma once

#include <hip/hip_runtime.h>

namespace aiso {

// One shot in flight: its wavefield pair and the per-shot buffers the fused
// step kernel reads or updates.  Unused members are null.
struct ShotSlot {
    float* p[4];          // wavefield slices [nx][nz]: the pair p[0], p[1] ("current" picked by parity),
                          // plus p[2], p[3] for the two-step kernel's out-of-place outputs
    float* q[4];          // same for the scattered field of the lockstep Born forward
    float* W;             // coarse source wavefield [nts][nx][nz]
    float* ss[2];         // two coarse-rate accumulators / secondary sources [nx][nz]
    float* img;           // per-shot image [nExt][nx][nz]
    float* dreg;          // regular-grid receiver traces [nRecReg][nts]
    float* sig;           // regular-grid source traces [nSrcReg][ntw]
    const int* srcTile;   // per-tile CSR offsets into srcEnt (tiles+1)
    const int2* srcEnt;   // (local index in tile, trace) for interior source points
    const int* recTile;
    const int2* recEnt;
    const int2* srcRing;  // (grid index, trace) for points in the FAT ring
    const int2* recRing;
    int nSrcRing, nRecRing;
};

// Operating modes of the fused time-step kernel (template bits).
enum : unsigned {
    M_ADJ = 1u << 0,        // adjoint stepper (Laplacian of v2dt2*c)
    M_SS_BORN = 1u << 1,    // + m~*(W[a]*fa) + m~*(W[b]*fb)   (Born scattered source)
    M_SS_SLICE = 1u << 2,   // + ss[a]*fa + ss[b]*fb            (precomputed secondary source)
    M_INJ_SIG = 1u << 3,    // + sig[s][itwInj] at source points
    M_INJ_DATA = 1u << 4,   // + d[r][its]*fc + d[r][its+1]*fd at receiver points
    M_CAPTURE = 1u << 5,    // A (+)= n*fe, B (+)= n*ff after damping
    M_IMAGE = 1u << 6,      // img += Wimg * B (when args.image != 0)
    M_REC_DATA = 1u << 7,   // d[r][its] += n*fe, d[r][its+1] += n*ff at receivers
    M_REC_SIG = 1u << 8,    // sig[s][itwRec] += n at source points
    M_WIN = 1u << 9,        // window capture V = (o*wa + c*wb) + n*wc when args.win != 0
    M_SS_PRE = 1u << 10,    // + SS[a]*fa + SS[b]*fb, SS = m~*W stored in the W array
    M_CAP_W = 1u << 11,     // adjoint stepper: capture into W[wA], W[wB] instead of ss
    M_CAP_ALT = 1u << 12,   // forward stepper: capture into img[wA], img[wB] instead of W
};

// Where a window capture V goes (StepArgs::winOut).
enum : int {
    WIN_W = 0,      // W[wA] = V                (source wavefield)
    WIN_SS = 1,     // W[wA] = m~ * V           (Born secondary source)
    WIN_IMG = 2,    // img += W[wImg] * V       (Born imaging condition)
    WIN_SLICE = 3,  // ss[ssA] = V              (receiver wavefield for the offset imaging)
};

struct StepArgs {
    const ShotSlot* slots;
    const float* v2dt2;
    const float* cd;        // sponge table (minPad entries)
    const float* mtil;      // scaled reflectivity for M_SS_BORN
    int nz, nx, minPad, nts, ntw;
    int tilesZ, ntiles, nshots, gridLogical;
    float zc[6], xc[6];
    int cur;                // index of the current slice in ShotSlot::p
    int its, itwInj, itwRec;
    long wA, wB;            // element offsets of the W slices used (capture / SS_BORN)
    int ssA, ssB;           // ShotSlot::ss indices used as A and B
    long ssOffA, ssOffB;    // element offsets added to them (ring slots of the offset-extended passes)
    float fa, fb;           // pre-damp pointwise weights (b skipped when fb == 0)
    float fc, fd;           // data-injection weights
    float fe, ff;           // capture / record weights
    int assignA, assignB;   // first touch of A / B: write instead of accumulate
    int skipA, skipB;       // zero weight: leave A / B untouched
    int image;              // apply M_IMAGE this step
    long wImg;              // element offset of the W slice used by M_IMAGE / WIN_IMG
    int win, winOut;        // M_WIN: capture at this step, and its destination
    float wa, wb, wc;       // M_WIN weights of the old, current and new slices
    int ntW;                // non-temporal accesses to the W / SS stream
    int fs;                 // free surface at iz = FAT (k_step_generic, k_step_fs, k_fwd_pair_fs)
    // lockstep Born forward (k_fwd_pair): scattered-field parity and the
    // secondary-source slices it reads
    int curQ;
    long wQa, wQb;
};

// Offset-extended Born over a batch of coarse samples (k_ext_ss_tile,
// k_ext_img_tile): sample k's receiver / secondary-source slice is
// ShotSlot::ss[0] + ssOff[k] and its source-wavefield slice W + wOff[k].
constexpr int kExtBatchMax = 16;
struct ExtBatchArgs {
    const ShotSlot* slots;
    const float* v2dt2;
    const float* mext;      // scaled extended model [nExt][nx][nz] (secondary sources)
    int nz, nx, hExt;
    int nk;                 // samples in the batch, 1..kExtBatchMax
    long ssOff[kExtBatchMax];
    long wOff[kExtBatchMax];
    // time-lag extension: sample j uses lag slices [lo[j], hi[j]); eLo..eHi
    // spans them all; slice e reads W at k_j - lagDir * 2 (e - hExt)
    int lo[kExtBatchMax], hi[kExtBatchMax];
    int eLo, eHi;
    int lagDir;             // +1 or -1
    int nts;                // coarse samples of W (bounds check of the time-lag batches)
};

// Image accumulation over a batch of coarse samples (k_img_acc_batch):
// X.img += X.W[xOff[j]] * Y.W[yOff[j]] for j = 0 .. nk-1, in that order.
struct ImgAccArgs {
    const ShotSlot* X;      // image, and the first operand's slices
    const ShotSlot* Y;      // the second operand's slices
    int nz, nx;
    int nk;                 // samples in the batch, 1..kExtBatchMax
    long xOff[kExtBatchMax];
    long yOff[kExtBatchMax];
};

// Two fine steps of both fields of the lockstep Born forward (sub == 2) per
// launch (k_fwd_pair2).  Launch m advances the source field by steps 2m+1,
// 2m+2 and the scattered field by steps 2m-1, 2m.  Each field reads its
// current and old slices and writes the two new ones to the other two slices.
struct Pair2Args {
    const ShotSlot* slots;  // srcTile/srcEnt: source points of each tile's extended region,
                            // recTile/recEnt: receiver points the tile owns
    const float* v2dt2;
    const float* cd;
    const float* mtil;
    int nz, nx, minPad, nts, ntw;
    int tilesZ, ntiles, gridLogical;
    float zc[6], xc[6];
    int sC, sO, sA, sB;     // source field: current, old, new after step A, new after step B
    int qC, qO, qA, qB;     // scattered field: same roles in ShotSlot::q
    int doS, doQ, tailS;    // advance the source / scattered field; write SS[nts-1] after step B
    int injA, injB;         // source-trace columns of steps A and B
    long ssNew, ssPrev, ssTail;  // element offsets in W of SS[m], SS[m-1], SS[nts-1]
    float wa, wb, wc;       // SS[m] = m~ * ((old*wa + current*wb) + newA*wc)
    float ta, tb;           // SS[nts-1] = m~ * (newA*ta + newB*tb)
    int its;                // coarse index of the scattered steps
    float faA, faB, fbB;    // step A: + SS[m-1]*faA;  step B: + (SS[m-1]*faB + SS[m]*fbB)
    float feA, ffA, feB, ffB;  // data record weights of steps A and B
    int ntW;
};

// Born adjoint receiver pass, two fine steps per launch (sub == 2), backward
// in time: launch m runs the steps (its, it2 = 1) and (its, it2 = 0) of
// coarse step its = nts - 2 - m on the adjoint stepper, with the data
// injection of each step and the imaging of the coarse sample at it2 = 1.
struct Rec2Args {
    const ShotSlot* slots;  // recTile/recEnt: receiver points of each tile's extended region
    const float* v2dt2;
    const float* cd;
    int nz, nx, minPad, nts;
    int tilesZ, ntiles, gridLogical;
    float zc[6], xc[6];
    int sC, sO, sA, sB;     // slices: current, old, new after step A, new after step B
    int its;
    float fcA, fdA, fcB, fdB;  // data injection: d[its]*fc + d[its+1]*fd
    float wa, wb, wc;       // img += W[its+1] * ((old*wa + current*wb) + newA*wc)
    long wImg;
    int tail;               // last launch: img += W[0] * (newA*ta + newB*tb)
    float ta, tb;
    int ntW;
};

// Shots on model windows of their own (uneven apertures): window w has its
// own padded grid nz x nx and v^2 dtw^2; every other step parameter is
// common.  One launch advances every tile of every window of the launch.
struct WinDesc {
    const float* v2dt2;     // [nx][nz] of the window
    int nz, nx, tilesZ, ntiles;
};
struct WinStepArgs {
    StepArgs a;             // common fields; a.slots: one ShotSlot per window of the launch
                            // (a.nz, a.nx, a.v2dt2, a.tilesZ, a.ntiles are taken from WinDesc)
    const WinDesc* win;     // per window of the launch
    const int2* tileMap;    // (window, tile in window) of every tile of the launch, window-major
    int ntotal;             // tiles of the launch
};

// Velocity-gradient imaging of the nonlinear operator: img += L(p_{t-1}) * nu_t,
// with p_{t-1} a forward slice stored in W and nu_t the adjoint slice p[nu].
struct GradArgs {
    const ShotSlot* slots;
    int nz, nx, tilesZ, ntiles, gridLogical;
    float zc[6], xc[6];
    long wF;                // element offset in W of p_{t-1}
    int nu;                 // ShotSlot::p index of nu_t
};

}  // namespace aiso

namespace aiso {

// The seven fused-step configurations used by the operators.
constexpr unsigned MODE_NL_FWD = M_INJ_SIG | M_REC_DATA;
constexpr unsigned MODE_NL_ADJ = M_ADJ | M_INJ_DATA | M_REC_SIG;
constexpr unsigned MODE_SRC = M_INJ_SIG | M_CAPTURE;
constexpr unsigned MODE_BORN_SCAT = M_SS_BORN | M_REC_DATA;
constexpr unsigned MODE_BORN_REC = M_ADJ | M_INJ_DATA | M_CAPTURE | M_IMAGE;
constexpr unsigned MODE_EXT_SCAT = M_SS_SLICE | M_REC_DATA;
constexpr unsigned MODE_EXT_REC = M_ADJ | M_INJ_DATA | M_CAPTURE;
// sub == 2 forms: coarse samples computed once from the old/current/new slices.
constexpr unsigned MODE_SRC_WIN = M_INJ_SIG | M_WIN;
constexpr unsigned MODE_BORN_SCAT_PRE = M_SS_PRE | M_REC_DATA;
constexpr unsigned MODE_REC_WIN = M_ADJ | M_INJ_DATA | M_WIN;
// Offset-extended tomographic operator: scattered fields captured at the
// coarse rate (forward, or adjoint into W), and the receiver wavefield.
constexpr unsigned MODE_TOMO_CAP = M_SS_SLICE | M_CAPTURE;
constexpr unsigned MODE_TOMO_ADJ_CAP = M_ADJ | M_SS_SLICE | M_CAPTURE | M_CAP_W;
constexpr unsigned MODE_TOMO_REC = M_ADJ | M_INJ_DATA | M_CAPTURE | M_CAP_W;
// Their sub == 2 forms (window capture into W, winOut WIN_W; the receiver
// wavefield's is MODE_REC_WIN).
constexpr unsigned MODE_TOMO_CAP_WIN = M_SS_SLICE | M_WIN;
constexpr unsigned MODE_TOMO_ADJ_CAP_WIN = M_ADJ | M_SS_SLICE | M_WIN;
// Wavefield saving: the Born scattered field captured at the coarse rate
// into ShotSlot::img (the source wavefield W stays readable).
constexpr unsigned MODE_BORN_SCAT_CAP = M_SS_BORN | M_CAPTURE | M_CAP_ALT;

// spb: shots advanced per workgroup (1 or 2), sharing the tile's v2dt2 and sponge.
// e0/e1 (optional): events timestamped by the dispatch itself.
// occ: 0, or 6 / 8 for the kernels compiled to the register budget of that
// many waves per SIMD (spb 1 only).
hipError_t launch_step(unsigned mode, int tz, int tx, int spb, const StepArgs& a, hipStream_t s,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, int occ = 0);
// The unspecialized step kernel: the extended-Born, tomographic and WEMVA
// operators, and every operator with a free surface (MODE_NL_*, MODE_SRC,
// MODE_SRC_WIN, MODE_BORN_SCAT, MODE_BORN_REC, MODE_EXT_*, MODE_REC_WIN,
// MODE_TOMO_*).
hipError_t launch_step_generic(unsigned mode, int tz, int tx, const StepArgs& a, hipStream_t s,
                               hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
hipError_t launch_ring(unsigned mode, const StepArgs& a, int maxRing, hipStream_t s);
// One fine step of every window of a launch (MODE_NL_FWD).  grid: workgroups,
// a multiple of 8; workgroups beyond the tile count loop over further tiles
// (a persistent launch when grid < w.ntotal).
hipError_t launch_step_windows(unsigned mode, int tz, int tx, int grid, const WinStepArgs& w, hipStream_t s,
                               hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
hipError_t launch_fill(float* p, long n, hipStream_t s);
hipError_t launch_scale_model(const float* m, float* out, const float* v2dt2, const float* refl,
                              int nz, int nx, hipStream_t s);
hipError_t launch_scale_ext(const float* m, float* out, const float* refl, int nz, int nx, int nExt,
                            hipStream_t s);
hipError_t launch_born_adj_finish(const ShotSlot* slots, int nshots, int ssR, const float* v2dt2,
                                  const float* refl, int nz, int nx, hipStream_t s);
hipError_t launch_ext_ss_batch(const ExtBatchArgs& a, int nshots, hipStream_t s);
// wemvaFwd: the WEMVA forward's imaging, (W_k[x-h] * ss_k[x+h]) * v2dt2[x+h]
hipError_t launch_ext_img_batch(const ExtBatchArgs& a, int nshots, hipStream_t s, bool wemvaFwd = false);
hipError_t launch_time_ss_batch(const ExtBatchArgs& a, int nshots, hipStream_t s);
hipError_t launch_time_img_batch(const ExtBatchArgs& a, int nshots, hipStream_t s);
hipError_t launch_time_scale(const float* v2dt2, const float* refl, int sloth, float* out, int nz, int nx,
                             hipStream_t s);
hipError_t launch_ext_adj_finish(const ShotSlot* slots, int nshots, const float* refl, int nz, int nx,
                                 int nExt, hipStream_t s);
// Tomographic operator: W <- second time derivative of W (nts coarse
// slices, in place); ss ring slots <- m~ * W[k] (a.mext is the 2-D m~);
// ss ring slots <- adjoint offset scattering of the receiver wavefield in W
// (a.mext is the scaled extended reflectivity); img += X.W[xOff] * Y.W[yOff].
hipError_t launch_dt2_coarse(const ShotSlot* slots, int nshots, int nts, float cSide, float cCenter, int nz, int nx,
                             hipStream_t s);
hipError_t launch_mul_batch(const ExtBatchArgs& a, int nshots, hipStream_t s);
hipError_t launch_tomo_adj_ss(const ExtBatchArgs& a, int nshots, hipStream_t s);
hipError_t launch_img_acc_batch(const ImgAccArgs& a, int nshots, hipStream_t s);
hipError_t launch_accumulate(float* acc, const float* img, long n, hipStream_t s);
// out = (ss[ssIdx] * refl) * v2dt2 on the interior of one shot slot
// (recordScaleWavefield, F/src/kernelsGpu.cu:145-153, into a zeroed slice).
hipError_t launch_record_scale(const ShotSlot* slot, int ssIdx, float* out, const float* refl, const float* v2dt2,
                               int nz, int nx, hipStream_t s);
hipError_t launch_add_to(float* out, const float* acc, long n, int add, hipStream_t s);
hipError_t launch_dev_forward(const float* reg, float* out, const int* cornerReg, const float* weight,
                              int nDev, int nb, int nt, int add, hipStream_t s);
// One fine step of the source field and, two fine steps behind, the
// scattered field of the Born forward in one launch (sub == 2).
hipError_t launch_fwd_pair(int tz, int tx, const StepArgs& a, hipStream_t s, hipEvent_t e0 = nullptr,
                           hipEvent_t e1 = nullptr);
// Extended region (EZ x EX) of k_fwd_pair2; the output tile is (EZ-10) x (EX-10).
constexpr int kPair2EZ = 64, kPair2EX = 32;
hipError_t launch_fwd_pair2(const Pair2Args& a, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
hipError_t launch_rec2(const Rec2Args& a, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
hipError_t launch_grad(int tz, int tx, const GradArgs& a, hipStream_t s);
hipError_t launch_win_tail(const ShotSlot* slots, int nshots, int cur, int winOut, float wa, float wb,
                           long wA, long wImg, int ssA, const float* mtil, int nz, int nx, hipStream_t s);
hipError_t launch_dev_adjoint(float* reg, const float* irreg, const int* adjPtr, const int* adjEnt,
                              const float* weight, int nReg, int nb, int nt, hipStream_t s);

}  // namespace aiso
