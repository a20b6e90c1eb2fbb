/*
 * aiso.h -- C-ABI of the MI355X-native 2D acoustic isotropic FD engine
 * (libaiso.so).  Plain pointers and sizes, status codes, no C++ or torch types.
 *
 * The entry points replace the reference's C++ operator classes that its
 * pybind11 modules bind (paths relative to the reference root;
 * F/ = acoustic_iso_lib/acoustic_iso_float, PY/ = acoustic_iso_lib/python/python_float):
 *
 *   aiso_ctx_*         fdParam + init*Gpu + allocate*Gpu     F/src/fdParam.cpp:9-140,
 *                                                            F/src/BornShotsGpuFunctions.cu:81-238
 *   aiso_device_*      deviceGpu (3 constructors, fwd/adj)   F/include/deviceGpu.h:30-45,
 *                                                            PY/pyAcoustic_iso_float_nl.cpp:18-25
 *   aiso_nl_*          nonlinearPropShotsGpu fwd/adj         F/include/nonlinearPropShotsGpu.h:18-53,
 *                                                            PY/pyAcoustic_iso_float_nl.cpp:27-40
 *   aiso_born_*        BornShotsGpu fwd/adj                  F/include/BornShotsGpu.h:20-56,
 *                                                            PY/pyAcoustic_iso_float_born.cpp:18-34
 *   aiso_bornext_*     BornExtShotsGpu fwd/adj (offset)      F/include/BornExtShotsGpu.h:20-58
 *   aiso_pad_plus      padding rule                          F/src/padFileGpuMain.cpp:37-66
 *
 * Conventions (same as the reference operators):
 *   - grids are [ix][iz] (z fastest) on the padded grid incl. the FAT=5 ring;
 *   - data are [iShot][iRec][its]; wavelets [iSrcDevice][its]; NL models
 *     [iShot|0][iSrcDevice][its]; extended models [iExt][ix][iz];
 *   - add == 0 overwrites the output, add != 0 accumulates into it;
 *   - every call returns AISO_OK or an error code; aiso_last_error() gives the
 *     message (the reference prints it and throws std::runtime_error);
 *   - with AISO_MEM_DEVICE, work the caller queued on the legacy default
 *     stream before the call is ordered before the call's reads (work on other
 *     streams must be complete); outputs are complete when the call returns.
 * Each context owns all of its state (no globals), so contexts are independent;
 * calls on one context must not overlap.
 */
#ifndef AISO_H
#define AISO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    AISO_OK = 0,
    AISO_ERR_PARAM = 1,    /* inconsistent parameters (fdParam / init checks)  */
    AISO_ERR_GEOM = 2,     /* device out of bounds, bad geometry               */
    AISO_ERR_HIP = 3,      /* HIP runtime failure                              */
    AISO_ERR_MEMORY = 4,   /* not enough device memory                         */
    AISO_ERR_STATE = 5     /* call out of order (e.g. no velocity set)         */
} aiso_status;

enum { AISO_MEM_HOST = 0, AISO_MEM_DEVICE = 1 };
enum { AISO_EXT_NONE = 0, AISO_EXT_OFFSET = 1, AISO_EXT_TIME = 2 };

/* Parfile keys read by fdParam (F/src/fdParam.cpp:14-96).  nz/nx are the
 * padded sizes of the velocity model, oz/ox its origins. */
typedef struct {
    int nz, nx;
    float dz, dx, oz, ox;
    int nts;
    float dts, ots;
    int sub;
    int zPadMinus, zPadPlus, xPadMinus, xPadPlus;
    int fat;            /* must be 5                                         */
    int blockSize;      /* must be 16 (model-size rule of fdParam.cpp:258-268) */
    float alphaCos;     /* default 0.99                                      */
    float fMax;         /* default 1000                                      */
    int freeSurface;    /* 0 or 1 (1 needs zPadMinus == 0)                   */
    int nExt;           /* odd, default 1                                    */
    int extension;      /* AISO_EXT_NONE, AISO_EXT_OFFSET or AISO_EXT_TIME   */
} aiso_fd_params;

typedef struct aiso_ctx aiso_ctx;
typedef struct aiso_device aiso_device;

/* ---------------------------------------------------------------- errors */
/* Message of the last failed call on this thread. */
const char* aiso_last_error(void);

/* --------------------------------------------------------------- context */
/* fdParam + init*Gpu: validates the parameters that do not depend on the
 * velocity (F/src/fdParam.cpp:28-96, :258-268; sub/minPad limits of
 * F/src/BornShotsGpuFunctions.cu:154-186) and binds HIP device `device`. */
aiso_status aiso_ctx_create(const aiso_fd_params* par, int device, aiso_ctx** out);
aiso_status aiso_ctx_destroy(aiso_ctx* ctx);

/* fdParam velocity part + allocate*Gpu: CFL <= 0.45 and dispersion >= 3
 * checks (F/src/fdParam.cpp:228-256), v^2*dtw^2 and 2/v^3 scalings
 * (:122-139), kept resident on the device.  vel: nx*nz floats.  The
 * reference's setVel (F/include/BornShotsGpu.h:54) maps here. */
aiso_status aiso_ctx_set_velocity(aiso_ctx* ctx, const float* vel, int mem);

/* Shots advanced together by one launch (0 = automatic, by free HBM). */
aiso_status aiso_ctx_set_batch(aiso_ctx* ctx, int shotsPerLaunch);
/* Tile shape of the time-step kernel for forward and adjoint calls: tz in
 * {32,64}, tx in {16,32} (default 32 x 16 forward, 32 x 32 adjoint; the
 * extended, tomographic and WEMVA forwards, and forwards with a free
 * surface, default to 32 x 32 until a tile is set). */
aiso_status aiso_ctx_set_tile(aiso_ctx* ctx, int tz, int tx);
/* Engine tuning knobs by name (no counterpart in the reference):
 *   "streamNT" 0/1: non-temporal access to the coarse wavefield stream;
 *   "lockstep" 0/1: Born forward with sub == 2 as one lockstep pass;
 *   "tileFwd", "tileAdj": tile of forward / adjoint calls, tz * 100 + tx;
 *   "pair2" 0/1: two fine steps per launch in the Born forward (lockstep)
 *       and adjoint when sub == 2 and no device lies in the FAT ring;
 *   "batchFwd", "batchAdj": shots per launch of forward / adjoint calls
 *       (0 = aiso_ctx_set_batch's value, or automatic; "batchAdj" also
 *       sets the batch of aiso_nl_velocity_gradient, 16 by default);
 *   "batchRec": group size of the adjoint receiver pass within a batch;
 *   "shotsPerBlock" 1/2: shots advanced by one workgroup of the step kernel;
 *   "occupancy" 0/6/8: the specialized step kernels compiled to the
 *       register budget of that many waves per SIMD, 0 the compiler's own
 *       (shotsPerBlock 1; default 6);
 *   "cacheBudgetMB": on-die cache budget the automatic batch size fills;
 *   "profileEvery" N >= 1: with profiling on, events on every N-th dispatch
 *       of each kernel kind (default 1: every dispatch);
 *   "imagePerShot" 0/1: aiso_born_adjoint / aiso_bornext_adjoint write one
 *       image per shot, model [nShot][nExt][nx][nz], instead of their sum
 *       (for callers that batch independent requests; default 0);
 *   "extBatch" 1..16: coarse samples per offset-imaging / secondary-source
 *       launch of the extended Born operators (default 16);
 *   "lanes" 1..4: streams the shots of a launch are split over, whose
 *       kernels run concurrently (default 2; the extended, tomographic and
 *       WEMVA operators and the two-step kernels run on one);
 *   "leg1", "leg2" 0/1: legs of the tomographic operator (parfile keys
 *       leg1/leg2 of F/src/tomoExtGpu.cpp:10-11, default 1);
 *   "sloth" 0/1: time-lag extended Born model is a slowness-squared
 *       perturbation, scaled by -v^2 dtw^2 (parfile key sloth of
 *       F/src/BornExtGpu.cpp:14, default 0; ignored with a free surface, as
 *       by the reference's free-surface routines);
 *   "saveWavefield" 0/1, "wavefieldShot" N: the next nonlinear and
 *       Born-type calls also save the coarse wavefields of their shot N (see
 *       aiso_ctx_wavefield; parfile keys saveWavefield / wavefieldShotNumber
 *       of F/src/BornShotsGpu.cpp:18-19). */
aiso_status aiso_ctx_set_option(aiso_ctx* ctx, const char* key, int value);
/* The current value of an aiso_ctx_set_option knob (tiles as tz * 100 + tx). */
aiso_status aiso_ctx_get_option(const aiso_ctx* ctx, const char* key, int* value);
/* Derived sizes: ntw, sub, nShotsPerLaunch actually used by the last call. */
aiso_status aiso_ctx_info(const aiso_ctx* ctx, int* ntw, int* batchUsed, size_t* bytesPerShot);
/* The HIP stream (hipStream_t) every kernel of this context runs on. */
aiso_status aiso_ctx_stream(const aiso_ctx* ctx, void** stream);
/* Variants of the fused time-step kernel, for profiling. */
enum {
    AISO_K_ALL = -1,
    AISO_K_NL_FWD = 0,     /* nonlinear forward step                         */
    AISO_K_NL_ADJ = 1,     /* nonlinear adjoint step                         */
    AISO_K_SRC = 2,        /* source-wavefield step with coarse capture      */
    AISO_K_BORN_SCAT = 3,  /* Born scattered-wavefield step                  */
    AISO_K_BORN_REC = 4,   /* Born adjoint receiver step with imaging        */
    AISO_K_EXT_SCAT = 5,   /* extended-Born scattered step                   */
    AISO_K_EXT_REC = 6,    /* extended-Born adjoint receiver step            */
    AISO_K_BORN_PAIR = 7,  /* lockstep Born forward: source + scattered step */
    AISO_K_TOMO = 8,       /* tomographic operator: scattered / receiver steps */
    AISO_K_COUNT = 9
};
/* Per-launch timing of the time-step kernels with HIP events on that stream
 * (enable != 0), reset of the counters.  With more than one lane, a pass is
 * timed whole instead (events before its lanes fork and after they join),
 * and each of its fine steps counts as one launch of all lanes' shots. */
aiso_status aiso_ctx_profile(aiso_ctx* ctx, int enable);
/* Launch count, summed duration (ms) and summed interior point-updates of
 * the time-step kernels of variant `which` (AISO_K_*) timed since profiling
 * was enabled. */
aiso_status aiso_ctx_profile_read(aiso_ctx* ctx, int which, long long* launches, double* ms,
                                  double* pointUpdates);
/* Wall time (ms, HIP events) of every shot batch run since profiling was
 * enabled, with the number of shots in each: a shot's latency is the time of
 * the batch that carried it.  Writes at most maxBatches entries; *count gets
 * the number available. */
aiso_status aiso_ctx_batch_read(aiso_ctx* ctx, float* ms, int* shots, int maxBatches, int* count);

/* Saved wavefields [nts][nx][nz] of the last call made with "saveWavefield"
 * (the reference's forwardWavefield / adjointWavefield and getters:
 * F/src/BornShotsGpu.cpp:145-240, 341-438 with getSrcWavefield /
 * getSecWavefield of F/include/BornShotsGpu.h:50-51;
 * F/src/nonlinearPropShotsGpu.cpp:153-250, 360-458 with getWavefield of
 * F/include/nonlinearPropShotsGpu.h:47; PY/pyAcoustic_iso_float_born.cpp:25-33,
 * PY/pyAcoustic_iso_float_nl.cpp:34-42).  AISO_WAVEFIELD_SOURCE: the
 * nonlinear operator's propagated wavefield, or the source wavefield of a
 * Born-type operator; AISO_WAVEFIELD_SECONDARY: the Born forward's scattered
 * wavefield or the Born adjoint's receiver wavefield scaled by
 * 2/v^3 v^2 dtw^2 (F/src/kernelsGpu.cu:145-153).  out: nts*nx*nz floats,
 * host or device memory by mem. */
enum { AISO_WAVEFIELD_SOURCE = 0, AISO_WAVEFIELD_SECONDARY = 1 };
aiso_status aiso_ctx_wavefield(aiso_ctx* ctx, int which, float* out, int mem);

/* ------------------------------------------------------------ geometry --
 * deviceGpu.  nzVel/nxVel are the padded velocity sizes; nt the trace length. */
/* Constructor #1 (F/src/deviceGpu.cpp:8-84): physical coordinates, bilinear
 * corner weights.  oz/dz/ox/dx are the velocity axes. */
aiso_status aiso_device_create_coords(int nDev, const float* zCoord, const float* xCoord,
                                      int nzVel, int nxVel, float oz, float dz, float ox,
                                      float dx, int nt, int dipole, float zDipoleShift,
                                      float xDipoleShift, aiso_device** out);
/* Constructor #2 (F/src/deviceGpu.cpp:87-139): integer grid positions. */
aiso_status aiso_device_create_ints(int nDev, const int* zGrid, const int* xGrid,
                                    int nzVel, int nxVel, int nt, int dipole,
                                    int zDipoleShift, int xDipoleShift, aiso_device** out);
/* Constructor #3 (F/src/deviceGpu.cpp:142-199): regular o/d/n device grid. */
aiso_status aiso_device_create_grid(int nzDev, int ozDev, int dzDev, int nxDev, int oxDev,
                                    int dxDev, int nzVel, int nxVel, int nt, int dipole,
                                    int zDipoleShift, int xDipoleShift, aiso_device** out);
aiso_status aiso_device_destroy(aiso_device* dev);
/* getNDeviceIrreg / getNDeviceReg / getRegPosUnique (F/include/deviceGpu.h:41-47). */
aiso_status aiso_device_sizes(const aiso_device* dev, int* nDevIrreg, int* nDevReg, int* nbCorner);
aiso_status aiso_device_reg_pos(const aiso_device* dev, int* regPos);
/* Raw corner list (nDevIrreg*nbCorner indices and weights). */
aiso_status aiso_device_corners(const aiso_device* dev, int* gridIdx, float* weight);
/* deviceGpu::forward (regular [nReg][nt] -> irregular [nDev][nt]) and
 * ::adjoint, on host buffers (F/src/deviceGpu.cpp:226-261). */
aiso_status aiso_device_forward(const aiso_device* dev, int add, const float* reg,
                                float* irreg, int nt);
aiso_status aiso_device_adjoint(const aiso_device* dev, int add, float* reg,
                                const float* irreg, int nt);

/* ----------------------------------------------------------- operators --
 * sources: nShot geometries (one per shot).  receivers: 1 (constant geometry)
 * or nShot.  wavelets: nWavelet (1 or nShot) blocks of [nSrcDev][nts].
 * mem: AISO_MEM_HOST or AISO_MEM_DEVICE for model and data. */

/* nonlinearPropShotsGpu::forward/adjoint (F/src/nonlinearPropShotsGpu.cpp:62-152,
 * 257-359).  model [nShotSig][nSrcDev][nts] with nShotSig in {1, nShot}. */
aiso_status aiso_nl_forward(aiso_ctx* ctx, int add, const float* model, int nShotSig,
                            float* data, aiso_device* const* sources, int nShot,
                            aiso_device* const* receivers, int nRecGeom, int mem);
aiso_status aiso_nl_adjoint(aiso_ctx* ctx, int add, float* model, int nShotSig,
                            const float* data, aiso_device* const* sources, int nShot,
                            aiso_device* const* receivers, int nRecGeom, int mem);

/* aiso_nl_forward without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * buffer beyond the inputs and outputs lives in one 64 MiB device array per
 * GPU, held until the process ends; a call whose per-tile lists and one
 * shot's wavefield pair, regular traces and source signal exceed it fails
 * with AISO_ERR_MEMORY (shots beyond what fits run in further batches).
 * model and data are device memory (AISO_MEM_DEVICE).  No free surface.
 * Calls that overlap on one GPU share the array and race: serialize them. */
aiso_status aiso_nl_forward_noalloc(aiso_ctx* ctx, int add, const float* model, int nShotSig, float* data,
                                    aiso_device* const* sources, int nShot, aiso_device* const* receivers,
                                    int nRecGeom);

/* Velocity gradient of a data misfit through the nonlinear operator, by the
 * adjoint-state method with checkpointed forward recomputation.  No
 * counterpart in the reference (its FWI takes the Born adjoint as the
 * gradient, PY/fwiMain.py:123-142); it serves torch.autograd and FWI callers.
 * model: the source signals of aiso_nl_forward ([nShotSig][nSrcDev][nts]);
 * dataGrad: dJ/d(data) [nShot][nRec][nts]; velGrad (out): dJ/d(velocity)
 * on the padded grid [nx][nz], overwritten.  checkpointEvery: fine steps
 * between checkpoints (0 = ceil(sqrt(ntw - 1))).  mem applies to dataGrad
 * and velGrad (the model is read through the host either way). */
aiso_status aiso_nl_velocity_gradient(aiso_ctx* ctx, const float* model, int nShotSig,
                                      const float* dataGrad, float* velGrad,
                                      aiso_device* const* sources, int nShot,
                                      aiso_device* const* receivers, int nRecGeom,
                                      int checkpointEvery, int mem);

/* BornShotsGpu::forward/adjoint (F/src/BornShotsGpu.cpp:60-143, 245-339).
 * model [nx][nz]. */
aiso_status aiso_born_forward(aiso_ctx* ctx, int add, const float* model, float* data,
                              aiso_device* const* sources, int nShot,
                              const float* wavelets, int nWavelet,
                              aiso_device* const* receivers, int nRecGeom, int mem);
aiso_status aiso_born_adjoint(aiso_ctx* ctx, int add, float* model, const float* data,
                              aiso_device* const* sources, int nShot,
                              const float* wavelets, int nWavelet,
                              aiso_device* const* receivers, int nRecGeom, int mem);

/* BornExtShotsGpu::forward/adjoint (F/src/BornExtGpu.cpp:33-134) with
 * extension=offset (-> F/src/BornExtShotsGpuFunctions.cu:582-724, 1848-1989)
 * or extension=time (-> :269-418, 1525-1678; free surface :885-1036,
 * 2152-2304).  model [nExt][nx][nz]. */
aiso_status aiso_bornext_forward(aiso_ctx* ctx, int add, const float* model, float* data,
                                 aiso_device* const* sources, int nShot,
                                 const float* wavelets, int nWavelet,
                                 aiso_device* const* receivers, int nRecGeom, int mem);
aiso_status aiso_bornext_adjoint(aiso_ctx* ctx, int add, float* model, const float* data,
                                 aiso_device* const* sources, int nShot,
                                 const float* wavelets, int nWavelet,
                                 aiso_device* const* receivers, int nRecGeom, int mem);

/* ------------------------------- tomographic (extension=offset or time) */
/* tomoExtShotsGpu (F/include/tomoExtShotsGpu.h:20-58): the extended
 * reflectivity [nExt][nx][nz] the operator linearizes about, the constructor
 * argument and setReflectivityExt of the reference (:37, :55).  Copied; the
 * per-call 2/v^3 scaling follows the current velocity. */
aiso_status aiso_ctx_set_ext_reflectivity(aiso_ctx* ctx, const float* reflectivityExt, int mem);
/* tomoExtShotsGpu::forward / adjoint (F/src/tomoExtShotsGpu.cpp:62-145 and
 * the adjoint below it; F/src/tomoExtGpu.cpp:34-100): velocity perturbation
 * model [nx][nz] <-> data [nShot][nRec][nts], legs selected by the options
 * "leg1" / "leg2".  Needs extension = AISO_EXT_OFFSET or AISO_EXT_TIME (the
 * time-lag routines of F/src/tomoExtShotsGpuFunctions.cu) and
 * aiso_ctx_set_ext_reflectivity first. */
aiso_status aiso_tomo_forward(aiso_ctx* ctx, int add, const float* model, float* data,
                              aiso_device* const* sources, int nShot,
                              const float* wavelets, int nWavelet,
                              aiso_device* const* receivers, int nRecGeom, int mem);
aiso_status aiso_tomo_adjoint(aiso_ctx* ctx, int add, float* model, const float* data,
                              aiso_device* const* sources, int nShot,
                              const float* wavelets, int nWavelet,
                              aiso_device* const* receivers, int nRecGeom, int mem);

/* ------------------------------------- WEMVA (extension=offset or time) */
/* wemvaExtShotsGpu::forward / adjoint (F/include/wemvaExtShotsGpu.h:36-57,
 * F/src/wemvaExtShotsGpu.cpp, F/src/wemvaExtShotsGpuFunctions.cu:458-808):
 * velocity perturbation model [nx][nz] <-> extended image [nExt][nx][nz],
 * linearized about the receiver wavefields of the observed data
 * [nShot][nRec][nts] (the reference's receiversSignalsVector, passed with
 * every call).  The forward writes (add: accumulates into) image, the
 * adjoint model.  Legs by the options "leg1" / "leg2"; needs extension =
 * AISO_EXT_OFFSET or AISO_EXT_TIME (the time-lag routines of
 * F/src/wemvaExtShotsGpuFunctions.cu). */
aiso_status aiso_wemva_forward(aiso_ctx* ctx, int add, const float* model, float* image, const float* data,
                               aiso_device* const* sources, int nShot,
                               const float* wavelets, int nWavelet,
                               aiso_device* const* receivers, int nRecGeom, int mem);
aiso_status aiso_wemva_adjoint(aiso_ctx* ctx, int add, float* model, const float* image, const float* data,
                               aiso_device* const* sources, int nShot,
                               const float* wavelets, int nWavelet,
                               aiso_device* const* receivers, int nRecGeom, int mem);

/* ------------------------------------- batched windows (uneven apertures)
 * No counterpart in the reference, which runs one shot at a time on the full
 * model per GPU (F/src/nonlinearPropShotsGpu.cpp:62-152,
 * F/src/nonlinearShotsGpuFunctions.cu:227-309); this batches many shots,
 * each on a window of its own size, in one launch per fine step.  Window w
 * has its own padded grid nz[w] x nx[w], velocity [nx][nz], source and
 * receiver geometry (built on that window's grid); every other parameter
 * (spacing, time sampling, pads, sponge, fMax) comes from `par` (its nz/nx
 * are ignored).  Each window's data equal those of aiso_nl_forward for that
 * shot alone on a context of that window.  No free surface, no device in a
 * window's FAT ring, the same source count in every window. */
typedef struct aiso_batch aiso_batch;
aiso_status aiso_batch_create(const aiso_fd_params* par, int device, aiso_batch** out);
aiso_status aiso_batch_destroy(aiso_batch* b);
/* Velocity checks (CFL, dispersion: F/src/fdParam.cpp:228-256) per window;
 * vel[w] host or device memory by mem.  Replaces any previous windows. */
aiso_status aiso_batch_set_windows(aiso_batch* b, int nWin, const int* nz, const int* nx, const float* const* vel,
                                   aiso_device* const* sources, aiso_device* const* receivers, int mem);
/* Launch configuration: tile tz in {32,64} x tx in {16,32}; windows per
 * launch (0 = all in one launch; more groups make more launch groups);
 * workgroups per CU (0 = one workgroup per tile, else a persistent launch
 * of that many per CU, each looping over tiles); lanes 1..4: streams the
 * launch groups alternate over, whose kernels run concurrently (default 2);
 * graph != 0: capture each group's time loop in a HIP graph on first use
 * and replay it later. */
aiso_status aiso_batch_set_config(aiso_batch* b, int tz, int tx, int windowsPerLaunch, int blocksPerCU, int lanes,
                                  int graph);
aiso_status aiso_batch_get_config(const aiso_batch* b, int* tz, int* tx, int* windowsPerLaunch, int* blocksPerCU,
                                  int* lanes, int* graph);
/* Autotuner: times `steps` fine steps of every launch group, on its lanes,
 * under each candidate (tile shapes x windows per launch {8,16,32,64,all} x
 * workgroups per CU {0,4,8} x lanes {1,2}), re-times the four fastest three
 * times over 3 x steps, and keeps the fastest per point-update (the graph
 * flag is left as set).  msPerPointUpdate (optional) gets its time. */
aiso_status aiso_batch_tune(aiso_batch* b, int steps, double* msPerPointUpdate);
/* Nonlinear forward of every window: model[k] [nSrcDev][nts] (nModel 1: one
 * signal for all windows, or one per window); data[w] [nRecDev_w][nts]. */
aiso_status aiso_batch_nl_forward(aiso_batch* b, int add, const float* const* model, int nModel,
                                  float* const* data, int mem);
/* Windows, launch groups of the current configuration, interior
 * point-updates of one call, device time of the last call (ms, HIP events
 * around its launches) and host time spent capturing and instantiating
 * graphs so far (ms). */
aiso_status aiso_batch_info(const aiso_batch* b, int* nWin, int* launchGroups, long long* pointUpdates,
                            float* lastMs, double* captureMs);
aiso_status aiso_batch_stream(const aiso_batch* b, void** stream);

/* ---------------------------------------------------------------- misc */
/* padFileGpuMain padding rule (F/src/padFileGpuMain.cpp:49-66, no free
 * surface): returns padPlus for a side of n samples with pad on the minus side. */
int aiso_pad_plus(int n, int pad, int blockSize);
/* The same with a free surface (F/src/padFileGpuMain.cpp:37-47): zPadMinus is
 * 0 and zPadPlus pads the depth axis alone to a multiple of blockSize. */
int aiso_pad_plus_free_surface(int nz, int zPad, int blockSize);

#ifdef __cplusplus
}
#endif
#endif
