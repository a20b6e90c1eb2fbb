/*
 * aiso_oracle.h -- CPU restatement of the 2D acoustic isotropic FD engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle that the HIP library is
 * checked against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it, and only as the checker / CPU baseline.  The
 * product path (acoustic_isotropic_operators_amd) never links or calls it.
 *
 * Every function restates one piece of the reference (paths relative to the
 * reference root, F/ = acoustic_iso_lib/acoustic_iso_float, D/ = the double
 * twin).  Each function exists twice: suffix _f (float, follows F/) and
 * suffix _d (double, follows D/, the FP64 oracle).
 *
 * Layouts follow the reference: 2-D grids are [ix][iz] (z fastest, index
 * ix*nz+iz) including the FAT=5 ring; traces are [trace][it] (time fastest);
 * wavefields are [its][ix][iz]; extended models are [iExt][ix][iz].
 */
#ifndef AISO_ORACLE_H
#define AISO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_FAT 5

typedef struct {
    int nz, nx;      /* padded grid sizes including the FAT ring          */
    int nts, sub;    /* coarse samples, fine steps per coarse sample      */
    int minPad;      /* min(zPadMinus, zPadPlus, xPadMinus, xPadPlus)      */
    int nExt;        /* extension length (odd), 1 = none                  */
    double dz, dx, dts, ots, alphaCos;
    int freeSurface; /* 1: free surface at iz = FAT (no top sponge)        */
} orc_grid;

#define ORC_DECL(REAL, SFX)                                                          \
void orc_tables##SFX(const orc_grid* g, REAL* zc, REAL* xc, REAL* filt, REAL* cd);   \
void orc_vel_scales##SFX(const orc_grid* g, const REAL* vel, REAL* v2dt2,           \
                         REAL* refl);                                                \
void orc_step_fwd##SFX(const orc_grid* g, const REAL* zc, const REAL* xc,           \
                       REAL* o, REAL* c, REAL* n, const REAL* v2dt2);                \
void orc_step_adj##SFX(const orc_grid* g, const REAL* zc, const REAL* xc,           \
                       REAL* o, const REAL* c, REAL* n, const REAL* v2dt2,           \
                       REAL* tmp);                                                   \
void orc_damp##SFX(const orc_grid* g, const REAL* cd, REAL* p1, REAL* p2);          \
void orc_nl_fwd##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* sig,         \
                     const int* srcPos, int nSrc, const int* recPos, int nRec,       \
                     REAL* dreg);                                                    \
void orc_nl_fwd_wfld##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* sig,    \
                          const int* srcPos, int nSrc, const int* recPos, int nRec,  \
                          REAL* dreg, REAL* wfld);                                   \
void orc_nl_adj_wfld##SFX(const orc_grid* g, const REAL* v2dt2, REAL* sig,          \
                          const int* srcPos, int nSrc, const int* recPos, int nRec,  \
                          const REAL* dreg, REAL* wfld);                             \
void orc_born_fwd_wfld##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl, \
                            const REAL* sig, const int* srcPos, int nSrc,            \
                            const int* recPos, int nRec, const REAL* model,          \
                            REAL* dreg, REAL* W, REAL* Wsec);                        \
void orc_born_adj_wfld##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl, \
                            const REAL* sig, const int* srcPos, int nSrc,            \
                            const int* recPos, int nRec, REAL* model,                \
                            const REAL* dreg, REAL* W, REAL* Wsec);                  \
void orc_nl_adj##SFX(const orc_grid* g, const REAL* v2dt2, REAL* sig,               \
                     const int* srcPos, int nSrc, const int* recPos, int nRec,       \
                     const REAL* dreg);                                              \
void orc_nl_vgrad##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* sig,       \
                       const int* srcPos, int nSrc, const int* recPos, int nRec,     \
                       const REAL* dres, REAL* grad, REAL* sigAdj);                  \
void orc_born_fwd##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl,      \
                       const REAL* sig, const int* srcPos, int nSrc,                 \
                       const int* recPos, int nRec, const REAL* model, REAL* dreg,   \
                       REAL* W);                                                     \
void orc_born_adj##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl,      \
                       const REAL* sig, const int* srcPos, int nSrc,                 \
                       const int* recPos, int nRec, REAL* model, const REAL* dreg,   \
                       REAL* W);                                                     \
void orc_bornext_fwd##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl,   \
                          const REAL* sig, const int* srcPos, int nSrc,              \
                          const int* recPos, int nRec, const REAL* model, REAL* dreg,\
                          REAL* W);                                                  \
void orc_bornext_adj##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl,   \
                          const REAL* sig, const int* srcPos, int nSrc,              \
                          const int* recPos, int nRec, REAL* model, const REAL* dreg,\
                          REAL* W);                                                  \
void orc_borntime_fwd##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl,   \
                           int sloth, const REAL* sig, const int* srcPos, int nSrc,  \
                           const int* recPos, int nRec, const REAL* model,           \
                           REAL* dreg, REAL* W);                                     \
void orc_borntime_adj##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl,   \
                           int sloth, const REAL* sig, const int* srcPos, int nSrc,  \
                           const int* recPos, int nRec, REAL* model,                 \
                           const REAL* dreg, REAL* W);                               \
void orc_tomo_fwd##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl,      \
                       const REAL* sig, const int* srcPos, int nSrc,                 \
                       const int* recPos, int nRec, const REAL* model,               \
                       const REAL* extRefl, int leg1, int leg2, int timeLag,         \
                       int sloth, REAL* dreg, REAL* W, REAL* P);                     \
void orc_tomo_adj##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl,      \
                       const REAL* sig, const int* srcPos, int nSrc,                 \
                       const int* recPos, int nRec, REAL* model, const REAL* dreg,   \
                       const REAL* extRefl, int leg1, int leg2, int timeLag,         \
                       int sloth, REAL* W, REAL* P);                                 \
void orc_dt2_coarse##SFX(const orc_grid* g, REAL* W);                                \
void orc_wemva_fwd##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl,     \
                        const REAL* sig, const int* srcPos, int nSrc,                \
                        const int* recPos, int nRec, const REAL* model,              \
                        const REAL* dobs, int leg1, int leg2, int timeLag,           \
                        REAL* image, REAL* W, REAL* P);                              \
void orc_wemva_adj##SFX(const orc_grid* g, const REAL* v2dt2, const REAL* refl,     \
                        const REAL* sig, const int* srcPos, int nSrc,                \
                        const int* recPos, int nRec, REAL* model, const REAL* image, \
                        const REAL* dobs, int leg1, int leg2, int timeLag,           \
                        REAL* W, REAL* P);                                           \
void orc_interp_fwd##SFX(int nts, double dts, double ots, int sub, int nDev,        \
                         int add, const REAL* m, REAL* d);                           \
void orc_interp_adj##SFX(int nts, double dts, double ots, int sub, int nDev,        \
                         int add, REAL* m, const REAL* d);                           \
void orc_d2t_fwd##SFX(int nt, double dt, int nDev, int add, const REAL* m, REAL* d);\
void orc_d2t_adj##SFX(int nt, double dt, int nDev, int add, REAL* m, const REAL* d);\
void orc_scale_source##SFX(const orc_grid* g, const REAL* vel, const int* pos,      \
                           int nPos, int nt, REAL* sig);                             \
int  orc_dev_coords##SFX(int nDev, const REAL* zCoord, const REAL* xCoord,          \
                         int nzVel, int nxVel, double oz, double dz, double ox,      \
                         double dx, int dipole, double zShift, double xShift,        \
                         int* gridIdx, REAL* weight);                                \
void orc_dev_fwd##SFX(int nDev, int nbCorner, const int* cornerReg,                 \
                      const REAL* weight, int nt, int add, const REAL* mReg,         \
                      REAL* dIrreg);                                                 \
void orc_dev_adj##SFX(int nDev, int nbCorner, const int* cornerReg,                 \
                      const REAL* weight, int nt, int add, REAL* mReg,               \
                      const REAL* dIrreg);

ORC_DECL(float, _f)
ORC_DECL(double, _d)

/* Integer geometry (precision independent). */
int orc_dev_grid(int nzDev, int ozDev, int dzDev, int nxDev, int oxDev, int dxDev,
                 int nzVel, int nxVel, int dipole, int zShift, int xShift,
                 int* gridIdx, float* weight);
int orc_dev_ints(int nDev, const int* zGrid, const int* xGrid, int nzVel, int nxVel,
                 int dipole, int zShift, int xShift, int* gridIdx, float* weight);
int orc_unique(int nIdx, const int* gridIdx, int* regPos, int* cornerReg);
int orc_pad_plus(int n, int pad, int blockSize);

#ifdef __cplusplus
}
#endif
#endif
