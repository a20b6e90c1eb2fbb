// kernels.hip -- gfx950 kernels of the 2D acoustic isotropic FD engine.
//
// The hot kernel is k_step: one fine time step of the 10th-order stencil for
// a batch of shots, with the cosine sponge, the sparse injections/records and
// the coarse-rate wavefield captures fused in.  The reference runs the same
// work as 4-5 separate launches per step (e.g. F/src/BornShotsGpuFunctions.cu
// :303-312: stepFwdGpu, injectSource, dampCosineEdge, interpWavefield).
//
// Fusing the sponge: the reference damps the new slice n and the current
// slice c after every step (F/src/kernelsGpu.cu:202-219).  The damped c is
// only ever read again as the "old" slice o of the next step, so k_step
// applies that second damping factor when it reads o, and never writes c
// (other blocks still read c as their halo).  x*d rounded then *d rounded is
// exactly what the reference computes, so this changes no bits.
//
// Layout: slices are [ix][iz] (z fastest).  A block owns a TZ x TX tile of
// interior points; 256 threads = TZ lanes along z x (256/TZ) lanes along x,
// each thread computing PX consecutive x-points with an x-window in
// registers.  The tile and its 5-point halo are staged in LDS.
#include <algorithm>
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace aiso {

constexpr int FAT = 5;
constexpr int NT = 256;
// The step kernels address a slice as a 64-bit base plus a 32-bit element
// offset; their launchers refuse larger slices.
constexpr long kMaxSlicePoints = 0x7fffffffL;

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// Linear workgroup ids go round-robin over the 8 XCDs; give each XCD a
// contiguous run of (shot, tile) work so neighbouring tiles share its L2.
__device__ __forceinline__ int xcd_logical(int bid, int gridPadded)
{
    const int per = gridPadded >> 3;
    return (bid & 7) * per + (bid >> 3);
}

// Block (x, y, z) of an nbx x nby x nbz grid launched as one 1-D grid padded
// to a multiple of 8, x fastest: x-neighbour tiles run on one XCD, so the
// column halos they share are served from its L2.  False for padding blocks.
struct Blk3 {
    int x, y, z;
};
__device__ __forceinline__ bool xcd_block3(int nbx, int nby, int nbz, Blk3& b)
{
    const int l = xcd_logical(blockIdx.x, gridDim.x);
    if (l >= nbx * nby * nbz) return false;
    b.x = l % nbx;
    const int r = l / nbx;
    b.z = r % nbz;
    b.y = r / nbz;
    return true;
}

// Streams touched once per coarse step (W, SS, image) can bypass the caches
// so they do not evict the wavefield slices reused at every fine step.
__device__ __forceinline__ float ld_stream(const float* p, int nt)
{
    return nt ? __builtin_nontemporal_load(p) : *p;
}
__device__ __forceinline__ void st_stream(float* p, float v, int nt)
{
    if (nt) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Global-memory views: pointers read from the ShotSlot table are generic to
// the compiler; casting them lets it emit global (not flat) loads and stores
// with a uniform 64-bit base and a 32-bit per-lane offset.
#define GBL __attribute__((address_space(1)))
using gf = GBL float;
using cgf = const GBL float;
__device__ __forceinline__ gf* G(float* p) { return (gf*)p; }
__device__ __forceinline__ cgf* G(const float* p) { return (cgf*)p; }

__device__ __forceinline__ float sponge(cgf* cd, int gz, int gx, int nz, int nx, int minPad)
{
    int dist = imin(imin(gz - FAT, gx - FAT), imin(nz - gz - 1 - FAT, nx - gx - 1 - FAT));
    return dist < minPad ? cd[dist] : 1.0f;
}

// dampCosineEdgeFs (F/src/kernelsGpu.cu:221-239): no sponge at a free surface.
__device__ __forceinline__ float sponge_fs(cgf* cd, int gz, int gx, int nz, int nx, int minPad)
{
    int dist = imin(gx - FAT, imin(nz - gz - 1 - FAT, nx - gx - 1 - FAT));
    return dist < minPad ? cd[dist] : 1.0f;
}

__device__ __forceinline__ float ldg_s(cgf* p, int nt) { return nt ? __builtin_nontemporal_load(p) : *p; }
__device__ __forceinline__ void stg_s(gf* p, float v, int nt)
{
    if (nt) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <int TZ, int TX>
struct TileShape {
    static constexpr int NXW = NT / TZ, PX = TX / NXW, LZ = TZ + 2 * FAT, LX = TX + 2 * FAT;
    static constexpr int KX = (LX + NXW - 1) / NXW, KZ = (LZ + TZ - 1) / TZ;
    static_assert(PX >= 1 && PX * NXW == TX, "tile shape");
};

// Stage a (TZ+10) x (TX+10) window of src (times mul, when given) into LDS:
// all loads first, then the stores.  With FULL the window lies inside the
// grid and no bounds are tested.
template <int TZ, int TX, bool FULL, bool MUL>
__device__ __forceinline__ void stage_tile(float* sT, cgf* src, const float* sMul, unsigned soff, int nz,
                                           int nx, int x0, int z0, int lane, int wy)
{
    using T = TileShape<TZ, TX>;
    float tv[T::KX][T::KZ];
#pragma unroll
    for (int kx = 0; kx < T::KX; ++kx)
#pragma unroll
        for (int kz = 0; kz < T::KZ; ++kz) {
            const int lx = wy + kx * T::NXW, lz = lane + kz * TZ;
            bool in = (kx < T::KX - 1 || lx < T::LX) && (kz < T::KZ - 1 || lz < T::LZ);
            if (!FULL) in = in && (x0 - FAT + lx < nx) && (z0 - FAT + lz < nz);
            tv[kx][kz] = in ? src[soff + (unsigned)(kx * T::NXW * nz + kz * TZ)] : 0.f;
        }
#pragma unroll
    for (int kx = 0; kx < T::KX; ++kx)
#pragma unroll
        for (int kz = 0; kz < T::KZ; ++kz) {
            const int lx = wy + kx * T::NXW, lz = lane + kz * TZ;
            if ((kx < T::KX - 1 || lx < T::LX) && (kz < T::KZ - 1 || lz < T::LZ)) {
                if constexpr (MUL) sT[lx * T::LZ + lz] = tv[kx][kz] * sMul[lx * T::LZ + lz];
                else sT[lx * T::LZ + lz] = tv[kx][kz];
            }
        }
}

// The adjoint Laplacian below a free surface (stepAdjFsGpu,
// F/src/kernelsGpu.cu:613-700): the z terms, minus those the forward folds
// onto the four rows under the surface, then the x terms, in the reference's
// order.  gz is the point's row; away from the surface the sum equals
// lap_at's.
template <int TZ, int TX, class A>
__device__ __forceinline__ float lap_adj_fs(const float* sT, const float* w, int j, int lx0, int lane, int gz,
                                            const A& a)
{
    using T = TileShape<TZ, TX>;
    const float* s = &sT[(lx0 + j + FAT) * T::LZ + lane + FAT];
    const float cc = w[j + FAT];
    float lap = a.zc[0] * cc
              + a.zc[1] * (s[-1] + s[1])
              + a.zc[2] * (s[-2] + s[2])
              + a.zc[3] * (s[-3] + s[3])
              + a.zc[4] * (s[-4] + s[4])
              + a.zc[5] * (s[-5] + s[5]);
    if (gz == FAT + 1) lap = lap - a.zc[2] * s[0] - a.zc[3] * s[1] - a.zc[4] * s[2] - a.zc[5] * s[3];
    else if (gz == FAT + 2) lap = lap - a.zc[3] * s[-1] - a.zc[4] * s[0] - a.zc[5] * s[1];
    else if (gz == FAT + 3) lap = lap - a.zc[4] * s[-2] - a.zc[5] * s[-1];
    else if (gz == FAT + 4) lap = lap - a.zc[5] * s[-3];
    return lap
         + a.xc[0] * cc
         + a.xc[1] * (w[j + 6] + w[j + 4])
         + a.xc[2] * (w[j + 7] + w[j + 3])
         + a.xc[3] * (w[j + 8] + w[j + 2])
         + a.xc[4] * (w[j + 9] + w[j + 1])
         + a.xc[5] * (w[j + 10] + w[j + 0]);
}

// 10th-order Laplacian at the PX points of a thread, in the reference's
// summation order (z terms, then x terms).  w holds the x-window.
template <int TZ, int TX, class A>
__device__ __forceinline__ float lap_at(const float* sT, const float* w, int j, int lx0, int lane, const A& a)
{
    using T = TileShape<TZ, TX>;
    const float* s = &sT[(lx0 + j + FAT) * T::LZ + lane + FAT];
    const float cc = w[j + FAT];
    return a.zc[0] * cc
         + a.zc[1] * (s[-1] + s[1])
         + a.zc[2] * (s[-2] + s[2])
         + a.zc[3] * (s[-3] + s[3])
         + a.zc[4] * (s[-4] + s[4])
         + a.zc[5] * (s[-5] + s[5])
         + a.xc[0] * cc
         + a.xc[1] * (w[j + 6] + w[j + 4])
         + a.xc[2] * (w[j + 7] + w[j + 3])
         + a.xc[3] * (w[j + 8] + w[j + 2])
         + a.xc[4] * (w[j + 9] + w[j + 1])
         + a.xc[5] * (w[j + 10] + w[j + 0]);
}

// Free surface on a staged top tile (setFsConditionFwdGpu,
// F/src/kernelsGpu.cu:555-567), for the interior columns: the surface row is
// zero and the FAT rows above it mirror the rows below with the opposite
// sign.  The stored slice keeps its surface value; the staged copy and the
// old-slice read take it as zero, as the reference's slice holds it after
// the condition.  The caller synchronizes before and after.
template <int TZ, int TX>
__device__ __forceinline__ void fs_mirror(float* sT, int x0, int nx, int lane, int wy)
{
    using T = TileShape<TZ, TX>;
    if (lane != 0) return;
    for (int lx = wy; lx < T::LX; lx += T::NXW) {
        const int gx = x0 - FAT + lx;
        if (gx >= FAT && gx < nx - FAT) {
            float* col = &sT[lx * T::LZ];
            col[FAT] = 0.f;
#pragma unroll
            for (int k = 0; k < FAT; ++k) col[k] = -col[2 * FAT - k];
        }
    }
}

// Sponge factors of a thread's points; distZ is constant along the thread's
// x-run.  FULL: every point is interior.  FS: a free surface on the top row,
// with no sponge above (dampCosineEdgeFs, F/src/kernelsGpu.cu:221-239).
template <int PX, bool FULL, bool SPONGE, bool FS = false>
__device__ __forceinline__ void tile_masks(bool (&ok)[PX], float (&dmp)[PX], const StepArgs& a, int gz, int gx0)
{
    const int nz = a.nz, nx = a.nx;
    const int distZ = FS ? nz - 1 - FAT - gz : imin(gz - FAT, nz - 1 - FAT - gz);
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        const int gx = gx0 + j;
        ok[j] = FULL || (gz < nz - FAT && gx < nx - FAT);
        float d = 1.f;
        if (SPONGE && ok[j]) {
            const int dist = imin(distZ, imin(gx - FAT, nx - 1 - FAT - gx));
            if (dist < a.minPad) d = a.cd[dist];
        }
        dmp[j] = d;
    }
}

template <int TZ, int TX, unsigned MODE, int SPB, bool FULL, bool SPONGE, bool FS = false>
__device__ __forceinline__ void step_block(const StepArgs& a, int group, int tile, int z0, int x0, float* sC,
                                           float* sV, float* sN)
{
    using T = TileShape<TZ, TX>;
    constexpr int PX = T::PX, NXW = T::NXW;
    constexpr bool ADJ = MODE & M_ADJ;
    constexpr bool PRE_SPARSE = MODE & (M_INJ_SIG | M_INJ_DATA);
    constexpr bool POST_SPARSE = MODE & (M_REC_DATA | M_REC_SIG);
    constexpr bool WIN = MODE & M_WIN;
    constexpr bool SSB = MODE & M_SS_BORN, SSP = MODE & M_SS_PRE, SSS = MODE & M_SS_SLICE;
    constexpr bool HAS_SS = SSB || SSP || SSS;

    const int nz = a.nz, nx = a.nx;
    const int lane = threadIdx.x % TZ, wy = threadIdx.x / TZ;
    const int gz = z0 + lane;
    const int lx0 = wy * PX;
    const unsigned off0 = (unsigned)((x0 + lx0) * nz + gz);                 // thread's first point
    const unsigned soff = (unsigned)((x0 - FAT + wy) * nz + (z0 - FAT + lane));  // first staged element
    const bool useB = a.fb != 0.f;
    cgf* V = G(a.v2dt2);

    bool ok[PX];
    float dmp[PX], vc[ADJ ? 1 : PX];
    tile_masks<PX, FULL, SPONGE, FS>(ok, dmp, a, gz, x0 + lx0);
    const bool surfTile = FS && z0 == FAT;          // block-uniform
    const bool surf = surfTile && gz == FAT;        // the thread's points lie on the surface row
    if constexpr (!ADJ) {
#pragma unroll
        for (int j = 0; j < PX; ++j) vc[j] = ok[j] ? V[off0 + j * nz] : 0.f;
    } else {
        // v2dt2 window shared by the shots; each thread reads back only the
        // entries it wrote itself.
        stage_tile<TZ, TX, FULL, false>(sV, V, nullptr, soff, nz, nx, x0, z0, lane, wy);
    }

    for (int si = 0; si < SPB; ++si) {
        const int shot = group * SPB + si;
        if (shot >= a.nshots) break;                  // block-uniform
        const ShotSlot& S = a.slots[shot];
        cgf* pc = G(S.p[a.cur]);
        gf* po = G(S.p[a.cur ^ 1]);

        // Per-point operands first, so their latency overlaps the staging.
        // (The imaging window's W sample and image come after the staging:
        // loading them here too measured 9 % slower.)
        float os[PX], pv[ADJ ? PX : 1];
        float sA[HAS_SS ? PX : 1], sB[HAS_SS ? PX : 1], mm[SSB ? PX : 1];
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            const unsigned o = off0 + j * nz;
            os[j] = ok[j] ? po[o] : 0.f;
            if constexpr (ADJ) pv[j] = ok[j] ? pc[o] : 0.f;
            if (ok[j]) {
                if constexpr (SSB) {
                    mm[j] = G(a.mtil)[o];
                    sA[j] = G(S.W + a.wA)[o];
                    sB[j] = useB ? G(S.W + a.wB)[o] : 0.f;
                }
                if constexpr (SSP) {
                    sA[j] = ldg_s(G(S.W + a.wA) + o, a.ntW);
                    sB[j] = useB ? ldg_s(G(S.W + a.wB) + o, a.ntW) : 0.f;
                }
                if constexpr (SSS) {
                    sA[j] = G(S.ss[a.ssA] + a.ssOffA)[o];
                    sB[j] = useB ? G(S.ss[a.ssB] + a.ssOffB)[o] : 0.f;
                }
            }
        }
        stage_tile<TZ, TX, FULL, ADJ>(sC, pc, sV, soff, nz, nx, x0, z0, lane, wy);
        __syncthreads();
        if constexpr (FS && !ADJ) {
            if (surfTile) {
                fs_mirror<TZ, TX>(sC, x0, nx, lane, wy);
                __syncthreads();
            }
        }
        // The adjoint receiver pass's imaging operands (the W sample and the
        // image), issued once the tile is staged so their latency overlaps
        // the stencil (17 % less time per imaging launch on MI355X than
        // issuing them after the stencil, and they cost registers only here).
        constexpr bool IMG_EARLY = WIN && ADJ;
        float wImg[IMG_EARLY ? PX : 1], iImg[IMG_EARLY ? PX : 1];
        if constexpr (IMG_EARLY) {
            if (a.win && a.winOut == WIN_IMG) {
#pragma unroll
                for (int j = 0; j < PX; ++j) {
                    const unsigned o = off0 + j * nz;
                    wImg[j] = ok[j] ? ldg_s(G(S.W + a.wImg) + o, a.ntW) : 0.f;
                    iImg[j] = ok[j] ? G(S.img)[o] : 0.f;
                }
            }
        }

        float w[PX + 2 * FAT];
#pragma unroll
        for (int j = 0; j < PX + 2 * FAT; ++j) w[j] = sC[(lx0 + j) * T::LZ + lane + FAT];

        float nv[PX];
        float cur[WIN ? PX : 1];
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            float lap;
            if constexpr (FS && ADJ) lap = surfTile ? lap_adj_fs<TZ, TX>(sC, w, j, lx0, lane, gz, a)
                                                    : lap_at<TZ, TX>(sC, w, j, lx0, lane, a);
            else lap = lap_at<TZ, TX>(sC, w, j, lx0, lane, a);
            const float cc = w[j + FAT];
            float n = 0.f;
            if (ok[j]) {
                // forward at a free surface: the old slice's surface value
                // reads as zero, as the reference zeroed it when it was current
                const float osj = (FS && !ADJ && surf) ? 0.f : os[j];
                const float o = SPONGE ? osj * dmp[j] : osj;       // second damping of the old slice
                if constexpr (ADJ) {
                    const float c = pv[j];
                    n = (FS && surf) ? o : lap + c + c - o;         // the adjoint leaves the surface row
                    if constexpr (WIN) cur[j] = c;
                } else {
                    n = vc[j] * lap + cc + cc - o;
                    // the window takes the slices as stored; at the surface
                    // the staged value is the mirrored one
                    if constexpr (WIN) cur[j] = (FS && surf) ? pc[off0 + j * nz] : cc;
                }
                if constexpr (SSB) {
                    const float ssa = mm[j] * sA[j];
                    if (useB) n += ssa * a.fa + (mm[j] * sB[j]) * a.fb;
                    else n += ssa * a.fa;
                }
                if constexpr (SSP || SSS) {
                    if (useB) n += sA[j] * a.fa + sB[j] * a.fb;
                    else n += sA[j] * a.fa;
                }
            }
            nv[j] = n;
        }

        if constexpr (PRE_SPARSE) {
            const int* tp = (MODE & M_INJ_SIG) ? S.srcTile : S.recTile;
            const int2* te = (MODE & M_INJ_SIG) ? S.srcEnt : S.recEnt;
            const int e0 = tp[tile], e1 = tp[tile + 1];
            if (e1 > e0) {                               // block-uniform
#pragma unroll
                for (int j = 0; j < PX; ++j) sN[(lx0 + j) * TZ + lane] = nv[j];
                __syncthreads();
                for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
                    const int2 en = te[e];
                    float v;
                    if constexpr (MODE & M_INJ_SIG) {
                        v = G(S.sig)[(long)en.y * a.ntw + a.itwInj];
                    } else {
                        cgf* dr = G(S.dreg) + ((long)en.y * a.nts + a.its);
                        v = dr[0] * a.fc + dr[1] * a.fd;
                    }
                    sN[en.x] += v;
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < PX; ++j) nv[j] = sN[(lx0 + j) * TZ + lane];
                __syncthreads();
            }
        }

#pragma unroll
        for (int j = 0; j < PX; ++j) {
            if (!ok[j]) continue;
            const unsigned gi = off0 + j * nz;
            const float n = SPONGE ? nv[j] * dmp[j] : nv[j];
            nv[j] = n;
            if constexpr (MODE & M_CAPTURE) {
                // Forward passes capture the coarse source wavefield W; the
                // adjoint receiver pass accumulates into the two ss slices.
                gf* A = ADJ ? G(S.ss[a.ssA] + a.ssOffA) : G(S.W + a.wA);
                gf* B = ADJ ? G(S.ss[a.ssB] + a.ssOffB) : G(S.W + a.wB);
                // A zero weight leaves the slice as it is (x + n*0 == x); a
                // first touch writes instead of accumulating into a zeroed slice.
                if (!a.skipA) A[gi] = a.assignA ? n * a.fe : A[gi] + n * a.fe;
                else if (a.assignA) A[gi] = 0.f;
                float bnew = 0.f;
                if (!a.skipB) {
                    bnew = a.assignB ? n * a.ff : B[gi] + n * a.ff;
                    B[gi] = bnew;
                } else if (a.assignB) {
                    B[gi] = 0.f;
                } else if constexpr (MODE & M_IMAGE) {
                    bnew = B[gi];
                }
                if constexpr (MODE & M_IMAGE) {
                    if (a.image) G(S.img)[gi] += G(S.W + a.wImg)[gi] * bnew;
                }
            }
            if constexpr (WIN) {
                if (a.win) {
                    // Coarse sample from three consecutive fine slices, accumulated
                    // in the reference's order (old, current, new).
                    const float v = (os[j] * a.wa + cur[j] * a.wb) + n * a.wc;
                    switch (a.winOut) {
                    case WIN_W: stg_s(G(S.W + a.wA) + gi, v, a.ntW); break;
                    case WIN_SS: stg_s(G(S.W + a.wA) + gi, G(a.mtil)[gi] * v, a.ntW); break;
                    case WIN_IMG:
                        if constexpr (IMG_EARLY) G(S.img)[gi] = iImg[j] + wImg[j] * v;
                        else G(S.img)[gi] += ldg_s(G(S.W + a.wImg) + gi, a.ntW) * v;
                        break;
                    default: G(S.ss[a.ssA] + a.ssOffA)[gi] = v; break;
                    }
                }
            }
            po[gi] = n;
        }

        if constexpr (POST_SPARSE) {
            const int* tp = (MODE & M_REC_SIG) ? S.srcTile : S.recTile;
            const int2* te = (MODE & M_REC_SIG) ? S.srcEnt : S.recEnt;
            const int e0 = tp[tile], e1 = tp[tile + 1];
            if (e1 > e0) {                               // block-uniform
#pragma unroll
                for (int j = 0; j < PX; ++j) sN[(lx0 + j) * TZ + lane] = nv[j];
                __syncthreads();
                for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
                    const int2 en = te[e];
                    const float n = sN[en.x];
                    if constexpr (MODE & M_REC_SIG) {
                        G(S.sig)[(long)en.y * a.ntw + a.itwRec] += n;
                    } else {
                        gf* dr = G(S.dreg) + ((long)en.y * a.nts + a.its);
                        dr[0] += n * a.fe;
                        dr[1] += n * a.ff;
                    }
                }
            }
        }
        if (SPB > 1) __syncthreads();                 // sC / sN are reused by the next shot
    }
}

// Tile position, interior / sponge classification of the workgroup.
struct TileInfo {
    int group, tile, z0, x0;
    bool full, sponge;
};

template <int TZ, int TX, bool FS = false>
__device__ __forceinline__ bool tile_info(const StepArgs& a, TileInfo& t)
{
    const int logical = xcd_logical(blockIdx.x, gridDim.x);
    if (logical >= a.gridLogical) return false;      // padding blocks (block-uniform)
    t.group = logical / a.ntiles;
    t.tile = logical - t.group * a.ntiles;
    const int tz = t.tile % a.tilesZ, tx = t.tile / a.tilesZ;
    t.z0 = FAT + tz * TZ;
    t.x0 = FAT + tx * TX;
    t.full = t.z0 + TZ <= a.nz - FAT && t.x0 + TX <= a.nx - FAT;
    const int top = FS ? a.minPad : t.z0 - FAT;      // no sponge under a free surface
    const int edge = imin(imin(top, a.nz - FAT - (t.z0 + TZ)), imin(t.x0 - FAT, a.nx - FAT - (t.x0 + TX)));
    t.sponge = !t.full || edge < a.minPad;
    return true;
}

template <int TZ, int TX, unsigned MODE, int SPB, bool FS = false>
__device__ __forceinline__ void step_tile(const StepArgs& a)
{
    using T = TileShape<TZ, TX>;
    constexpr bool ADJ = MODE & M_ADJ;
    constexpr bool SPARSE = MODE & (M_INJ_SIG | M_INJ_DATA | M_REC_DATA | M_REC_SIG);
    __shared__ float sC[T::LX * T::LZ];
    __shared__ float sV[ADJ ? T::LX * T::LZ : 1];   // v2dt2 window (adjoint), shared by the shots
    __shared__ float sN[SPARSE ? TX * TZ : 1];
    TileInfo t;
    if (!tile_info<TZ, TX, FS>(a, t)) return;
    if (t.full && !t.sponge)
        step_block<TZ, TX, MODE, SPB, true, false, FS>(a, t.group, t.tile, t.z0, t.x0, sC, sV, sN);
    else if (t.full) step_block<TZ, TX, MODE, SPB, true, true, FS>(a, t.group, t.tile, t.z0, t.x0, sC, sV, sN);
    else step_block<TZ, TX, MODE, SPB, false, true, FS>(a, t.group, t.tile, t.z0, t.x0, sC, sV, sN);
}

template <int TZ, int TX, unsigned MODE, int SPB>
__global__ __launch_bounds__(NT) void k_step(const StepArgs a)
{
    step_tile<TZ, TX, MODE, SPB>(a);
}

// The step with a free surface on the top row (the nonlinear and Born
// operators' surface condition, F/src/kernelsGpu.cu:555-700), one shot per
// workgroup, at the compiler's register budget (the adjoint source pass
// spills at 6 waves per SIMD).
template <int TZ, int TX, unsigned MODE>
__global__ __launch_bounds__(NT) void k_step_fs(const StepArgs a)
{
    step_tile<TZ, TX, MODE, 1, true>(a);
}

// The same step with the register budget of W waves per SIMD (W = 6: 80
// VGPRs, W = 8: 64): more workgroups resident to cover the latency of the
// tile staging (the adjoint source pass otherwise compiles to 76 VGPRs, 6
// waves, and the receiver pass with its early imaging operands to 84, 5).
// The budget applies to the 32-row tiles only: the 64-row ones spill in it.
template <int TZ, int TX, unsigned MODE, int W>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(TZ == 32 ? W : 1, 8)))
void k_step_w(const StepArgs a)
{
    step_tile<TZ, TX, MODE, 1>(a);
}

// One tile of a window (k_step's work on that window's own grid): the
// window, the tile and its classification from the tile list entry l.
template <int TZ, int TX, unsigned MODE>
__device__ __forceinline__ void win_tile(const WinStepArgs& w, int l, float* sC, float* sV, float* sN)
{
    const int2 tm = w.tileMap[l];
    const WinDesc d = w.win[tm.x];
    StepArgs a = w.a;
    a.slots = w.a.slots + tm.x;
    a.v2dt2 = d.v2dt2;
    a.nz = d.nz;
    a.nx = d.nx;
    a.tilesZ = d.tilesZ;
    a.ntiles = d.ntiles;
    a.nshots = 1;
    const int z0 = FAT + (tm.y % d.tilesZ) * TZ, x0 = FAT + (tm.y / d.tilesZ) * TX;
    const bool full = z0 + TZ <= d.nz - FAT && x0 + TX <= d.nx - FAT;
    const int edge = imin(imin(z0 - FAT, d.nz - FAT - (z0 + TZ)), imin(x0 - FAT, d.nx - FAT - (x0 + TX)));
    if (full && edge >= a.minPad) step_block<TZ, TX, MODE, 1, true, false>(a, 0, tm.y, z0, x0, sC, sV, sN);
    else if (full) step_block<TZ, TX, MODE, 1, true, true>(a, 0, tm.y, z0, x0, sC, sV, sN);
    else step_block<TZ, TX, MODE, 1, false, true>(a, 0, tm.y, z0, x0, sC, sV, sN);
}

// One fine step of a launch of windows of uneven size.  The tiles of all
// windows form one list; XCD k (workgroup id mod 8) owns the k-th eighth of
// it, so neighbouring tiles of a window share that XCD's L2.  One workgroup
// per tile: the launch has 8 * ceil(ntotal / 8) workgroups.  Every point
// receives the operations k_step gives it, in the same order.  The 32-row
// tiles fit the register budget of 8 waves per SIMD without spilling.
template <int TZ, int TX, unsigned MODE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(TZ == 32 ? 8 : 1, 8)))
void k_step_win(const WinStepArgs w)
{
    using T = TileShape<TZ, TX>;
    constexpr bool ADJ = MODE & M_ADJ;
    constexpr bool SPARSE = MODE & (M_INJ_SIG | M_INJ_DATA | M_REC_DATA | M_REC_SIG);
    __shared__ float sC[T::LX * T::LZ];
    __shared__ float sV[ADJ ? T::LX * T::LZ : 1];
    __shared__ float sN[SPARSE ? TX * TZ : 1];
    const int xcd = blockIdx.x & 7;
    const int chunk = (w.ntotal + 7) >> 3;
    const int l = xcd * chunk + (int)(blockIdx.x >> 3);
    if (l >= imin((xcd + 1) * chunk, w.ntotal)) return;   // block-uniform
    win_tile<TZ, TX, MODE>(w, l, sC, sV, sN);
}

// The persistent form: fewer workgroups than tiles, each looping over the
// tiles of its XCD's eighth of the list.
template <int TZ, int TX, unsigned MODE>
__global__ __launch_bounds__(NT) void k_step_win_loop(const WinStepArgs w)
{
    using T = TileShape<TZ, TX>;
    constexpr bool ADJ = MODE & M_ADJ;
    constexpr bool SPARSE = MODE & (M_INJ_SIG | M_INJ_DATA | M_REC_DATA | M_REC_SIG);
    __shared__ float sC[T::LX * T::LZ];
    __shared__ float sV[ADJ ? T::LX * T::LZ : 1];
    __shared__ float sN[SPARSE ? TX * TZ : 1];
    const int xcd = blockIdx.x & 7, per = gridDim.x >> 3;
    const int chunk = (w.ntotal + 7) >> 3;
    const int end = imin((xcd + 1) * chunk, w.ntotal);
    for (int l = xcd * chunk + (int)(blockIdx.x >> 3); l < end; l += per) {   // block-uniform
        win_tile<TZ, TX, MODE>(w, l, sC, sV, sN);
        __syncthreads();                                  // sC, sV, sN are reused by the next tile
    }
}

// Lockstep Born forward, sub == 2.  One launch advances the source field
// (pair S.p) by fine step t and the scattered field (pair S.q) by fine step
// t-2.  The source step writes the secondary source SS[k] = m~ * W[k] of the
// coarse sample it completes into a 3-slice ring (S.W); the scattered step
// reads the two ring slices of its coarse interval, which earlier launches
// completed.  Velocity, sponge factor and tile bookkeeping are shared by the
// two fields; the arithmetic per field is that of k_step (MODE_SRC_WIN with
// WIN_SS, and MODE_BORN_SCAT_PRE).
template <int TZ, int TX, bool FULL, bool SPONGE, bool FS = false>
__device__ __forceinline__ void pair_block(const StepArgs& a, int shot, int tile, int z0, int x0, float* sA, float* sB,
                                           float* sN)
{
    using T = TileShape<TZ, TX>;
    constexpr int PX = T::PX;
    const int nz = a.nz, nx = a.nx;
    const int lane = threadIdx.x % TZ, wy = threadIdx.x / TZ;
    const int gz = z0 + lane;
    const int lx0 = wy * PX;
    const unsigned off0 = (unsigned)((x0 + lx0) * nz + gz);
    const unsigned soff = (unsigned)((x0 - FAT + wy) * nz + (z0 - FAT + lane));
    const bool useB = a.fb != 0.f;
    const ShotSlot& S = a.slots[shot];
    cgf* pcA = G(S.p[a.cur]);
    gf* poA = G(S.p[a.cur ^ 1]);
    cgf* pcB = G(S.q[a.curQ]);
    gf* poB = G(S.q[a.curQ ^ 1]);
    cgf* V = G(a.v2dt2);
    cgf* Qa = G(S.W + a.wQa);
    cgf* Qb = G(S.W + a.wQb);

    bool ok[PX];
    float dmp[PX];
    tile_masks<PX, FULL, SPONGE, FS>(ok, dmp, a, gz, x0 + lx0);
    const bool surfTile = FS && z0 == FAT;          // block-uniform
    const bool surf = surfTile && gz == FAT;
    float oa[PX], ob[PX], v2[PX], qa[PX], qb[PX], mm[PX];
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        const unsigned o = off0 + j * nz;
        float o1 = 0.f, o2 = 0.f, v = 0.f, s1 = 0.f, s2 = 0.f, m = 0.f;
        if (ok[j]) {
            o1 = poA[o];
            o2 = poB[o];
            v = V[o];
            s1 = ldg_s(Qa + o, a.ntW);
            if (useB) s2 = ldg_s(Qb + o, a.ntW);
            if (a.win) m = G(a.mtil)[o];
        }
        oa[j] = o1;
        ob[j] = o2;
        v2[j] = v;
        qa[j] = s1;
        qb[j] = s2;
        mm[j] = m;
    }
    stage_tile<TZ, TX, FULL, false>(sA, pcA, nullptr, soff, nz, nx, x0, z0, lane, wy);
    stage_tile<TZ, TX, FULL, false>(sB, pcB, nullptr, soff, nz, nx, x0, z0, lane, wy);
    __syncthreads();
    if constexpr (FS) {
        if (surfTile) {
            fs_mirror<TZ, TX>(sA, x0, nx, lane, wy);
            fs_mirror<TZ, TX>(sB, x0, nx, lane, wy);
            __syncthreads();
        }
    }

    float wA[PX + 2 * FAT], wB[PX + 2 * FAT];
#pragma unroll
    for (int j = 0; j < PX + 2 * FAT; ++j) {
        wA[j] = sA[(lx0 + j) * T::LZ + lane + FAT];
        wB[j] = sB[(lx0 + j) * T::LZ + lane + FAT];
    }
    float nA[PX], nB[PX];
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        const float lapA = lap_at<TZ, TX>(sA, wA, j, lx0, lane, a);
        const float lapB = lap_at<TZ, TX>(sB, wB, j, lx0, lane, a);
        const float ca = wA[j + FAT], cb = wB[j + FAT];
        float na = 0.f, nb = 0.f;
        if (ok[j]) {
            // at a free surface the old slices' surface values read as zero
            const float oaj = (FS && surf) ? 0.f : oa[j], obj = (FS && surf) ? 0.f : ob[j];
            na = v2[j] * lapA + ca + ca - (SPONGE ? oaj * dmp[j] : oaj);
            nb = v2[j] * lapB + cb + cb - (SPONGE ? obj * dmp[j] : obj);
            if (useB) nb += qa[j] * a.fa + qb[j] * a.fb;
            else nb += qa[j] * a.fa;
        }
        nA[j] = na;
        nB[j] = nb;
    }

    // source injection into the source field (before damping)
    {
        const int e0 = S.srcTile[tile], e1 = S.srcTile[tile + 1];
        if (e1 > e0) {
#pragma unroll
            for (int j = 0; j < PX; ++j) sN[(lx0 + j) * TZ + lane] = nA[j];
            __syncthreads();
            for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
                const int2 en = S.srcEnt[e];
                sN[en.x] += G(S.sig)[(long)en.y * a.ntw + a.itwInj];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < PX; ++j) nA[j] = sN[(lx0 + j) * TZ + lane];
            __syncthreads();
        }
    }

    gf* Wd = G(S.W + a.wA);
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        if (!ok[j]) continue;
        const unsigned gi = off0 + j * nz;
        const float na = SPONGE ? nA[j] * dmp[j] : nA[j];
        const float nb = SPONGE ? nB[j] * dmp[j] : nB[j];
        nB[j] = nb;
        if (a.win) {
            // the window takes the slices as stored (at a free surface the
            // staged current value is the mirrored one)
            const float cA = (FS && surf) ? pcA[gi] : wA[j + FAT];
            const float v = (oa[j] * a.wa + cA * a.wb) + na * a.wc;
            stg_s(Wd + gi, mm[j] * v, a.ntW);
        }
        poA[gi] = na;
        poB[gi] = nb;
    }

    // data record from the scattered field (after damping)
    {
        const int e0 = S.recTile[tile], e1 = S.recTile[tile + 1];
        if (e1 > e0) {
#pragma unroll
            for (int j = 0; j < PX; ++j) sN[(lx0 + j) * TZ + lane] = nB[j];
            __syncthreads();
            for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
                const int2 en = S.recEnt[e];
                const float n = sN[en.x];
                gf* dr = G(S.dreg) + ((long)en.y * a.nts + a.its);
                dr[0] += n * a.fe;
                dr[1] += n * a.ff;
            }
        }
    }
}

// The 32 x 16 instance (the default tile) is compiled to the register
// budget of 6 waves per SIMD (80 VGPRs, one spilled): the sixth wave covers
// more staging latency than the spill costs (40.85 -> 40.23 us per fine step
// of the Born benchmark on MI355X).  The other tiles spill heavily in it.
template <int TZ, int TX>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(TZ == 32 && TX == 16 ? 6 : 1, 8)))
void k_fwd_pair(const StepArgs a)
{
    using T = TileShape<TZ, TX>;
    __shared__ float sA[T::LX * T::LZ];
    __shared__ float sB[T::LX * T::LZ];
    __shared__ float sN[TX * TZ];
    TileInfo t;
    if (!tile_info<TZ, TX>(a, t)) return;
    if (t.full && !t.sponge) pair_block<TZ, TX, true, false>(a, t.group, t.tile, t.z0, t.x0, sA, sB, sN);
    else if (t.full) pair_block<TZ, TX, true, true>(a, t.group, t.tile, t.z0, t.x0, sA, sB, sN);
    else pair_block<TZ, TX, false, true>(a, t.group, t.tile, t.z0, t.x0, sA, sB, sN);
}

// The lockstep pair with a free surface on the top row.
template <int TZ, int TX>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(TZ == 32 && TX == 16 ? 6 : 1, 8)))
void k_fwd_pair_fs(const StepArgs a)
{
    using T = TileShape<TZ, TX>;
    __shared__ float sA[T::LX * T::LZ];
    __shared__ float sB[T::LX * T::LZ];
    __shared__ float sN[TX * TZ];
    TileInfo t;
    if (!tile_info<TZ, TX, true>(a, t)) return;
    if (t.full && !t.sponge) pair_block<TZ, TX, true, false, true>(a, t.group, t.tile, t.z0, t.x0, sA, sB, sN);
    else if (t.full) pair_block<TZ, TX, true, true, true>(a, t.group, t.tile, t.z0, t.x0, sA, sB, sN);
    else pair_block<TZ, TX, false, true, true>(a, t.group, t.tile, t.z0, t.x0, sA, sB, sN);
}

// Stage the (EZ+10) x (EX+10) window of src whose first element is grid
// point (rz0, rx0); points outside the grid read as 0.
template <int EZ, int EX>
__device__ __forceinline__ void stage_region(float* sT, cgf* src, int nz, int nx, int rz0, int rx0, int lane, int wy)
{
    using T = TileShape<EZ, EX>;
    float tv[T::KX][T::KZ];
#pragma unroll
    for (int kx = 0; kx < T::KX; ++kx)
#pragma unroll
        for (int kz = 0; kz < T::KZ; ++kz) {
            const int lx = wy + kx * T::NXW, lz = lane + kz * EZ;
            const int gx = rx0 + lx, gz = rz0 + lz;
            const bool in = lx < T::LX && lz < T::LZ && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
            tv[kx][kz] = in ? src[(unsigned)(gx * nz + gz)] : 0.f;
        }
#pragma unroll
    for (int kx = 0; kx < T::KX; ++kx)
#pragma unroll
        for (int kz = 0; kz < T::KZ; ++kz) {
            const int lx = wy + kx * T::NXW, lz = lane + kz * EZ;
            if (lx < T::LX && lz < T::LZ) sT[lx * T::LZ + lz] = tv[kx][kz];
        }
}

// Adds the traces of the sparse entries [e0, e1) of a tile (local index in the
// EZ x EX region, trace row) to the thread's values v: column `col` of sig.
template <int EZ, int PX>
__device__ __forceinline__ void inject_region(float (&v)[PX], float* sN, const int2* ent, int e0, int e1,
                                              cgf* sig, int ntw, int col, int lx0, int lane)
{
    if (e1 <= e0) return;                              // block-uniform
#pragma unroll
    for (int j = 0; j < PX; ++j) sN[(lx0 + j) * EZ + lane] = v[j];
    __syncthreads();
    for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
        const int2 en = ent[e];
        sN[en.x] += sig[(long)en.y * ntw + col];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PX; ++j) v[j] = sN[(lx0 + j) * EZ + lane];
    __syncthreads();
}

// d[r][its] += v*fe, d[r][its+1] += v*ff at the receiver points the tile owns.
template <int EZ, int PX>
__device__ __forceinline__ void record_region(const float (&v)[PX], float* sN, const int2* ent, int e0, int e1,
                                              gf* dreg, int nts, int its, float fe, float ff, int lx0, int lane)
{
    if (e1 <= e0) return;                              // block-uniform
#pragma unroll
    for (int j = 0; j < PX; ++j) sN[(lx0 + j) * EZ + lane] = v[j];
    __syncthreads();
    for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
        const int2 en = ent[e];
        const float n = sN[en.x];
        gf* dr = dreg + ((long)en.y * nts + its);
        dr[0] += n * fe;
        dr[1] += n * ff;
    }
    __syncthreads();
}

// Lockstep Born forward, two fine steps per field per launch (sub == 2).
// With mtil == nullptr and doQ == 0 it is the source pass of the Born
// adjoint: W[m] itself is stored (x * 1 == x).
//
// A workgroup owns an (EZ-10) x (EX-10) output tile and computes step A on
// the EZ x EX region around it (the tile plus the 5-point halo step B
// needs), then step B on the tile from step A's region kept in LDS.  Each
// field therefore reads its current slice (with a 10-point halo) and its old
// slice once per two steps and writes both new slices, out of place: the old
// slice is read with a halo too.  Per point and step the arithmetic is that
// of k_fwd_pair: damping of the old slice on read, injection before damping,
// SS[m] from the old/current/new slices of step A, records after damping.
// Points of the FAT ring are never computed and read as 0 (callers use this
// kernel only when no source or receiver lies in the ring).
template <int EZ, int EX>
__global__ __launch_bounds__(NT) void k_fwd_pair2(const Pair2Args a)
{
    using T = TileShape<EZ, EX>;
    constexpr int PX = T::PX, OZ = EZ - 2 * FAT, OX = EX - 2 * FAT;
    __shared__ float sX[T::LX * T::LZ];   // current slice, region plus 5-point halo
    __shared__ float sY[T::LX * T::LZ];   // step A's new slice, same frame
    __shared__ float sN[EX * EZ];         // sparse injection / record staging

    const int logical = xcd_logical(blockIdx.x, gridDim.x);
    if (logical >= a.gridLogical) return;             // padding blocks (block-uniform)
    const int shot = logical / a.ntiles, tile = logical - shot * a.ntiles;
    const int z0 = FAT + (tile % a.tilesZ) * OZ, x0 = FAT + (tile / a.tilesZ) * OX;
    const int ez0 = z0 - FAT, ex0 = x0 - FAT;         // region origin
    const int lane = threadIdx.x % EZ, wy = threadIdx.x / EZ, lx0 = wy * PX;
    const int nz = a.nz, nx = a.nx, gz = ez0 + lane;
    const ShotSlot& S = a.slots[shot];
    cgf* V = G(a.v2dt2);

    bool inA[PX], inT[PX];
    unsigned gi[PX];
    float dmp[PX], v2[PX];
    {
        const bool zin = gz >= FAT && gz < nz - FAT;
        const bool zT = lane >= FAT && lane < EZ - FAT;
        const int distZ = imin(gz - FAT, nz - 1 - FAT - gz);
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            const int gx = ex0 + lx0 + j;
            inA[j] = zin && gx >= FAT && gx < nx - FAT;
            inT[j] = inA[j] && zT && lx0 + j >= FAT && lx0 + j < EX - FAT;
            gi[j] = inA[j] ? (unsigned)(gx * nz + gz) : 0u;
            float d = 1.f;
            if (inA[j]) {
                const int dist = imin(distZ, imin(gx - FAT, nx - 1 - FAT - gx));
                if (dist < a.minPad) d = a.cd[dist];
            }
            dmp[j] = d;
            v2[j] = inA[j] ? V[gi[j]] : 0.f;
        }
    }
    const int e0s = S.srcTile[tile], e1s = S.srcTile[tile + 1];
    const int e0r = S.recTile[tile], e1r = S.recTile[tile + 1];
    gf* W = G(S.W);

    float ssm[PX];                                     // SS[m] at the tile's points
    if (a.doS) {                                       // block-uniform
        cgf* pC = G(S.p[a.sC]);
        cgf* pO = G(S.p[a.sO]);
        float o[PX], mm[PX];
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            o[j] = inA[j] ? pO[gi[j]] : 0.f;
            mm[j] = inT[j] ? (a.mtil ? G(a.mtil)[gi[j]] : 1.f) : 0.f;   // raw W without m~
        }
        stage_region<EZ, EX>(sX, pC, nz, nx, ez0 - FAT, ex0 - FAT, lane, wy);
        __syncthreads();
        float w[PX + 2 * FAT];
#pragma unroll
        for (int j = 0; j < PX + 2 * FAT; ++j) w[j] = sX[(lx0 + j) * T::LZ + lane + FAT];
        float nA[PX];
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            const float lap = lap_at<EZ, EX>(sX, w, j, lx0, lane, a);
            const float cc = w[j + FAT];
            nA[j] = inA[j] ? v2[j] * lap + cc + cc - o[j] * dmp[j] : 0.f;
        }
        inject_region<EZ, PX>(nA, sN, S.srcEnt, e0s, e1s, G(S.sig), a.ntw, a.injA, lx0, lane);
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            nA[j] = nA[j] * dmp[j];
            ssm[j] = 0.f;
            if (inT[j]) {
                const float v = (o[j] * a.wa + w[j + FAT] * a.wb) + nA[j] * a.wc;
                ssm[j] = mm[j] * v;
                stg_s(W + a.ssNew + gi[j], ssm[j], a.ntW);
            }
            sY[(lx0 + j + FAT) * T::LZ + lane + FAT] = nA[j];
        }
        __syncthreads();
        float wB[PX + 2 * FAT];
#pragma unroll
        for (int j = 0; j < PX + 2 * FAT; ++j) wB[j] = sY[(lx0 + j) * T::LZ + lane + FAT];
        float nB[PX];
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            const float lap = lap_at<EZ, EX>(sY, wB, j, lx0, lane, a);
            const float cc = wB[j + FAT];
            nB[j] = inT[j] ? v2[j] * lap + cc + cc - w[j + FAT] * dmp[j] : 0.f;
        }
        inject_region<EZ, PX>(nB, sN, S.srcEnt, e0s, e1s, G(S.sig), a.ntw, a.injB, lx0, lane);
        gf* pA = G(S.p[a.sA]);
        gf* pB = G(S.p[a.sB]);
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            if (!inT[j]) continue;
            const float nb = nB[j] * dmp[j];
            pA[gi[j]] = nA[j];
            pB[gi[j]] = nb;
            if (a.tailS) stg_s(W + a.ssTail + gi[j], mm[j] * (nA[j] * a.ta + nb * a.tb), a.ntW);
        }
        __syncthreads();                               // sX, sY reused by the scattered field
    }

    if (a.doQ) {                                       // block-uniform
        cgf* qC = G(S.q[a.qC]);
        cgf* qO = G(S.q[a.qO]);
        float o[PX], sp[PX];
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            o[j] = inA[j] ? qO[gi[j]] : 0.f;
            sp[j] = inA[j] ? ldg_s(W + a.ssPrev + gi[j], a.ntW) : 0.f;
            // without a source step in this launch SS[m] was completed earlier
            if (!a.doS) ssm[j] = inT[j] ? ldg_s(W + a.ssNew + gi[j], a.ntW) : 0.f;
        }
        stage_region<EZ, EX>(sX, qC, nz, nx, ez0 - FAT, ex0 - FAT, lane, wy);
        __syncthreads();
        float w[PX + 2 * FAT];
#pragma unroll
        for (int j = 0; j < PX + 2 * FAT; ++j) w[j] = sX[(lx0 + j) * T::LZ + lane + FAT];
        float nA[PX];
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            const float lap = lap_at<EZ, EX>(sX, w, j, lx0, lane, a);
            const float cc = w[j + FAT];
            float n = 0.f;
            if (inA[j]) {
                n = v2[j] * lap + cc + cc - o[j] * dmp[j];
                n += sp[j] * a.faA;
            }
            nA[j] = n * dmp[j];
            sY[(lx0 + j + FAT) * T::LZ + lane + FAT] = nA[j];
        }
        record_region<EZ, PX>(nA, sN, S.recEnt, e0r, e1r, G(S.dreg), a.nts, a.its, a.feA, a.ffA, lx0, lane);
        __syncthreads();
        float wB[PX + 2 * FAT];
#pragma unroll
        for (int j = 0; j < PX + 2 * FAT; ++j) wB[j] = sY[(lx0 + j) * T::LZ + lane + FAT];
        float nB[PX];
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            const float lap = lap_at<EZ, EX>(sY, wB, j, lx0, lane, a);
            const float cc = wB[j + FAT];
            float n = 0.f;
            if (inT[j]) {
                n = v2[j] * lap + cc + cc - w[j + FAT] * dmp[j];
                n += sp[j] * a.faB + ssm[j] * a.fbB;
            }
            nB[j] = n * dmp[j];
        }
        record_region<EZ, PX>(nB, sN, S.recEnt, e0r, e1r, G(S.dreg), a.nts, a.its, a.feB, a.ffB, lx0, lane);
        gf* qA = G(S.q[a.qA]);
        gf* qB = G(S.q[a.qB]);
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            if (!inT[j]) continue;
            qA[gi[j]] = nA[j];
            qB[gi[j]] = nB[j];
        }
    }
}

// Stage the (EZ+10) x (EX+10) window of src * mul (element-wise) whose first
// element is grid point (rz0, rx0); points outside the grid read as 0.
template <int EZ, int EX>
__device__ __forceinline__ void stage_region_mul(float* sT, cgf* src, cgf* mul, int nz, int nx, int rz0, int rx0,
                                                 int lane, int wy)
{
    using T = TileShape<EZ, EX>;
    float tv[T::KX][T::KZ], tm[T::KX][T::KZ];
#pragma unroll
    for (int kx = 0; kx < T::KX; ++kx)
#pragma unroll
        for (int kz = 0; kz < T::KZ; ++kz) {
            const int lx = wy + kx * T::NXW, lz = lane + kz * EZ;
            const int gx = rx0 + lx, gz = rz0 + lz;
            const bool in = lx < T::LX && lz < T::LZ && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
            const unsigned g = in ? (unsigned)(gx * nz + gz) : 0u;
            tv[kx][kz] = in ? src[g] : 0.f;
            tm[kx][kz] = in ? mul[g] : 0.f;
        }
#pragma unroll
    for (int kx = 0; kx < T::KX; ++kx)
#pragma unroll
        for (int kz = 0; kz < T::KZ; ++kz) {
            const int lx = wy + kx * T::NXW, lz = lane + kz * EZ;
            if (lx < T::LX && lz < T::LZ) sT[lx * T::LZ + lz] = tv[kx][kz] * tm[kx][kz];
        }
}

// d[r][its]*fc + d[r][its+1]*fd added at the receiver points of the region.
template <int EZ, int PX>
__device__ __forceinline__ void inject_data_region(float (&v)[PX], float* sN, const int2* ent, int e0, int e1,
                                                   cgf* dreg, int nts, int its, float fc, float fd, int lx0,
                                                   int lane)
{
    if (e1 <= e0) return;                              // block-uniform
#pragma unroll
    for (int j = 0; j < PX; ++j) sN[(lx0 + j) * EZ + lane] = v[j];
    __syncthreads();
    for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
        const int2 en = ent[e];
        cgf* dr = dreg + ((long)en.y * nts + its);
        sN[en.x] += dr[0] * fc + dr[1] * fd;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PX; ++j) v[j] = sN[(lx0 + j) * EZ + lane];
    __syncthreads();
}

// Born adjoint receiver pass, two fine steps per launch (sub == 2).  Same
// tiling as k_fwd_pair2; the adjoint stepper takes the Laplacian of
// v2dt2 * current, so step A stages that product with a 10-point halo and
// step B stages v2dt2 * (step A's new slice) from registers.  Per point and
// step the arithmetic is that of k_step in MODE_REC_WIN.
template <int EZ, int EX>
__global__ __launch_bounds__(NT) void k_rec2(const Rec2Args a)
{
    using T = TileShape<EZ, EX>;
    constexpr int PX = T::PX, OZ = EZ - 2 * FAT, OX = EX - 2 * FAT;
    __shared__ float sX[T::LX * T::LZ];
    __shared__ float sY[T::LX * T::LZ];
    __shared__ float sN[EX * EZ];

    const int logical = xcd_logical(blockIdx.x, gridDim.x);
    if (logical >= a.gridLogical) return;             // padding blocks (block-uniform)
    const int shot = logical / a.ntiles, tile = logical - shot * a.ntiles;
    const int z0 = FAT + (tile % a.tilesZ) * OZ, x0 = FAT + (tile / a.tilesZ) * OX;
    const int ez0 = z0 - FAT, ex0 = x0 - FAT;
    const int lane = threadIdx.x % EZ, wy = threadIdx.x / EZ, lx0 = wy * PX;
    const int nz = a.nz, nx = a.nx, gz = ez0 + lane;
    const ShotSlot& S = a.slots[shot];
    cgf* V = G(a.v2dt2);

    bool inA[PX], inT[PX];
    unsigned gi[PX];
    float dmp[PX], v2[PX];
    {
        const bool zin = gz >= FAT && gz < nz - FAT;
        const bool zT = lane >= FAT && lane < EZ - FAT;
        const int distZ = imin(gz - FAT, nz - 1 - FAT - gz);
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            const int gx = ex0 + lx0 + j;
            inA[j] = zin && gx >= FAT && gx < nx - FAT;
            inT[j] = inA[j] && zT && lx0 + j >= FAT && lx0 + j < EX - FAT;
            gi[j] = inA[j] ? (unsigned)(gx * nz + gz) : 0u;
            float d = 1.f;
            if (inA[j]) {
                const int dist = imin(distZ, imin(gx - FAT, nx - 1 - FAT - gx));
                if (dist < a.minPad) d = a.cd[dist];
            }
            dmp[j] = d;
            v2[j] = inA[j] ? V[gi[j]] : 0.f;
        }
    }
    const int e0 = S.recTile[tile], e1 = S.recTile[tile + 1];
    cgf* pC = G(S.p[a.sC]);
    cgf* pO = G(S.p[a.sO]);
    float o[PX], c[PX];
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        o[j] = inA[j] ? pO[gi[j]] : 0.f;
        c[j] = inA[j] ? pC[gi[j]] : 0.f;
    }
    stage_region_mul<EZ, EX>(sX, pC, V, nz, nx, ez0 - FAT, ex0 - FAT, lane, wy);
    __syncthreads();
    float w[PX + 2 * FAT];
#pragma unroll
    for (int j = 0; j < PX + 2 * FAT; ++j) w[j] = sX[(lx0 + j) * T::LZ + lane + FAT];
    float nA[PX];
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        const float lap = lap_at<EZ, EX>(sX, w, j, lx0, lane, a);
        nA[j] = inA[j] ? lap + c[j] + c[j] - o[j] * dmp[j] : 0.f;
    }
    inject_data_region<EZ, PX>(nA, sN, S.recEnt, e0, e1, G(S.dreg), a.nts, a.its, a.fcA, a.fdA, lx0, lane);
    gf* img = G(S.img);
    cgf* W = G(S.W);
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        nA[j] = nA[j] * dmp[j];
        if (inT[j]) {
            const float v = (o[j] * a.wa + c[j] * a.wb) + nA[j] * a.wc;
            img[gi[j]] += ldg_s(W + a.wImg + gi[j], a.ntW) * v;
        }
        sY[(lx0 + j + FAT) * T::LZ + lane + FAT] = nA[j] * v2[j];
    }
    __syncthreads();
    float wB[PX + 2 * FAT];
#pragma unroll
    for (int j = 0; j < PX + 2 * FAT; ++j) wB[j] = sY[(lx0 + j) * T::LZ + lane + FAT];
    float nB[PX];
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        const float lap = lap_at<EZ, EX>(sY, wB, j, lx0, lane, a);
        nB[j] = inT[j] ? lap + nA[j] + nA[j] - c[j] * dmp[j] : 0.f;
    }
    inject_data_region<EZ, PX>(nB, sN, S.recEnt, e0, e1, G(S.dreg), a.nts, a.its, a.fcB, a.fdB, lx0, lane);
    gf* pA = G(S.p[a.sA]);
    gf* pB = G(S.p[a.sB]);
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        if (!inT[j]) continue;
        const float nb = nB[j] * dmp[j];
        pA[gi[j]] = nA[j];
        pB[gi[j]] = nb;
        if (a.tail) img[gi[j]] += ldg_s(W + gi[j], a.ntW) * (nA[j] * a.ta + nb * a.tb);
    }
}

// Velocity-gradient imaging of the nonlinear operator, one (tile, shot) per
// workgroup: img += L(p_{t-1}) * nu_t at the tile's points, the Laplacian in
// the stepper's order on an LDS-staged window of the forward slice.
template <int TZ, int TX>
__global__ __launch_bounds__(NT) void k_grad(const GradArgs a)
{
    using T = TileShape<TZ, TX>;
    constexpr int PX = T::PX;
    __shared__ float sC[T::LX * T::LZ];
    const int logical = xcd_logical(blockIdx.x, gridDim.x);
    if (logical >= a.gridLogical) return;             // padding blocks (block-uniform)
    const int shot = logical / a.ntiles, tile = logical - shot * a.ntiles;
    const int z0 = FAT + (tile % a.tilesZ) * TZ, x0 = FAT + (tile / a.tilesZ) * TX;
    const int nz = a.nz, nx = a.nx;
    const int lane = threadIdx.x % TZ, wy = threadIdx.x / TZ, lx0 = wy * PX, gz = z0 + lane;
    const ShotSlot& S = a.slots[shot];
    const unsigned soff = (unsigned)((x0 - FAT + wy) * nz + (z0 - FAT + lane));
    stage_tile<TZ, TX, false, false>(sC, G(S.W + a.wF), nullptr, soff, nz, nx, x0, z0, lane, wy);
    __syncthreads();
    float w[PX + 2 * FAT];
#pragma unroll
    for (int j = 0; j < PX + 2 * FAT; ++j) w[j] = sC[(lx0 + j) * T::LZ + lane + FAT];
    cgf* nu = G(S.p[a.nu]);
    gf* img = G(S.img);
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        const int gx = x0 + lx0 + j;
        if (gz >= nz - FAT || gx >= nx - FAT) continue;
        const unsigned gi = (unsigned)(gx * nz + gz);
        const float lap = lap_at<TZ, TX>(sC, w, j, lx0, lane, a);
        img[gi] += lap * nu[gi];
    }
}

// Unspecialized form of the fused step: one shot per workgroup, bounds and
// sponge tested at every point.  Every slice is addressed as a 64-bit base
// (read once per workgroup) plus a 32-bit element offset within the slice
// (launch_step_generic refuses grids of 2^31 points or more).  The operands
// are read through global address-space views, so their loads are not
// counted with the LDS reads of the stencil.  FS: the free-surface
// conditions, for launches with a surface; without it they are compiled
// out.  The extended-Born, tomographic and WEMVA operators run on it.
template <int TZ, int TX, unsigned MODE, bool FS>
__global__ __launch_bounds__(NT) void k_step_generic(const StepArgs a)
{
    constexpr bool ADJ = MODE & M_ADJ;
    constexpr bool PRE_SPARSE = MODE & (M_INJ_SIG | M_INJ_DATA);
    constexpr bool POST_SPARSE = MODE & (M_REC_DATA | M_REC_SIG);
    constexpr int NXW = NT / TZ, PX = TX / NXW, LZ = TZ + 2 * FAT, LX = TX + 2 * FAT;
    static_assert(PX >= 1 && PX * NXW == TX, "tile shape");

    __shared__ float sC[LX * LZ];
    __shared__ float sN[(PRE_SPARSE || POST_SPARSE) ? TX * TZ : 1];

    const int logical = xcd_logical(blockIdx.x, gridDim.x);
    if (logical >= a.gridLogical) return;             // padding blocks (block-uniform)
    const int shot = logical / a.ntiles, tile = logical - shot * a.ntiles;
    const int tz = tile % a.tilesZ, tx = tile / a.tilesZ;
    const ShotSlot& S = a.slots[shot];
    const int nz = a.nz, nx = a.nx;
    const int z0 = FAT + tz * TZ, x0 = FAT + tx * TX;
    const int lane = threadIdx.x % TZ, wy = threadIdx.x / TZ;
    cgf* pc = G(S.p[a.cur]);
    gf* po = G(S.p[a.cur ^ 1]);
    cgf* v2dt2 = G(a.v2dt2);

    // Per-point operands.  The forward modes load them before staging the
    // tile, so that their latency overlaps it; the adjoint modes, whose
    // staging reads the same current slice, load them after it (measured
    // faster each way on MI355X).  The values are the same either way.
    const int gz = z0 + lane;
    const int lx0 = wy * PX;
    const bool zok = gz < nz - FAT;
    constexpr bool SSB = MODE & M_SS_BORN, SSP = MODE & M_SS_PRE, SSS = MODE & M_SS_SLICE;
    constexpr bool HAS_SS = SSB || SSP || SSS;
    const bool useB = a.fb != 0.f;
    bool ok[PX];
    float os[PX], pv[PX], dmp[PX];
    float sA[HAS_SS ? PX : 1], sB[HAS_SS ? PX : 1], mm[SSB ? PX : 1];
    // the secondary-source slices
    cgf* ssa = nullptr;
    cgf* ssb = nullptr;
    if constexpr (SSB || SSP) {
        ssa = G(S.W) + a.wA;
        if (useB) ssb = G(S.W) + a.wB;
    }
    if constexpr (SSS) {
        ssa = G(S.ss[a.ssA]) + a.ssOffA;
        if (useB) ssb = G(S.ss[a.ssB]) + a.ssOffB;
    }
    auto load_operands = [&] {
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            const int gx = x0 + lx0 + j;
            ok[j] = zok && gx < nx - FAT;
            os[j] = 0.f;
            pv[j] = 0.f;
            dmp[j] = 1.f;
            if constexpr (HAS_SS) { sA[j] = 0.f; sB[j] = 0.f; }
            if constexpr (SSB) mm[j] = 0.f;
            if (ok[j]) {
                const unsigned gi = (unsigned)(gx * nz + gz);
                os[j] = po[gi];
                pv[j] = ADJ ? pc[gi] : v2dt2[gi];
                // sponge factor, its table read issued with the operands
                dmp[j] = FS ? sponge_fs(G(a.cd), gz, gx, nz, nx, a.minPad) : sponge(G(a.cd), gz, gx, nz, nx, a.minPad);
                if constexpr (SSB) mm[j] = G(a.mtil)[gi];
                if constexpr (HAS_SS) {
                    sA[j] = ssa[gi];
                    if (useB) sB[j] = ssb[gi];
                }
            }
        }
    };
    if constexpr (!ADJ) load_operands();

    // Stage the current slice (times v2dt2 for the adjoint) with its halo:
    // every load of the thread's share first, then the LDS stores, so the
    // loads are in flight together.  Points outside the grid stage as zero.
    {
        constexpr int KX = (LX + NXW - 1) / NXW, KZ = (LZ + TZ - 1) / TZ;
        float tv[KX][KZ], tm[ADJ ? KX : 1][ADJ ? KZ : 1];
        const unsigned soff = (unsigned)((x0 - FAT + wy) * nz + (z0 - FAT + lane));   // first staged element
#pragma unroll
        for (int kx = 0; kx < KX; ++kx)
#pragma unroll
            for (int kz = 0; kz < KZ; ++kz) {
                const int lx = wy + kx * NXW, lz = lane + kz * TZ;
                const bool in = lx < LX && lz < LZ && x0 - FAT + lx < nx && z0 - FAT + lz < nz;
                const unsigned gi = soff + (unsigned)(kx * NXW * nz + kz * TZ);
                tv[kx][kz] = in ? pc[gi] : 0.f;
                if constexpr (ADJ) tm[kx][kz] = in ? v2dt2[gi] : 0.f;
            }
#pragma unroll
        for (int kx = 0; kx < KX; ++kx)
#pragma unroll
            for (int kz = 0; kz < KZ; ++kz) {
                const int lx = wy + kx * NXW, lz = lane + kz * TZ;
                if (lx < LX && lz < LZ) {
                    if constexpr (ADJ) sC[lx * LZ + lz] = tv[kx][kz] * tm[kx][kz];
                    else sC[lx * LZ + lz] = tv[kx][kz];
                }
            }
    }
    __syncthreads();
    if constexpr (!ADJ) {
        // Free surface (setFsConditionFwdGpu, F/src/kernelsGpu.cu:555-567), on
        // the staged copy of a top tile's interior columns: the surface row
        // is zero and the rows above it mirror the rows below with the
        // opposite sign.  The stored slice keeps its surface value; the
        // old-slice read below takes it as zero, as the reference's slice
        // holds it after the condition.
        if (FS && z0 == FAT) {
            for (int lx = wy; lx < LX; lx += NXW) {
                const int gx = x0 - FAT + lx;
                if (gx >= FAT && gx < nx - FAT && lane == 0) {
                    float* col = &sC[lx * LZ];
                    col[FAT] = 0.f;
#pragma unroll
                    for (int k = 0; k < FAT; ++k) col[k] = -col[2 * FAT - k];
                }
            }
            __syncthreads();
        }
    }

    if constexpr (ADJ) load_operands();
    float w[PX + 2 * FAT];
#pragma unroll
    for (int j = 0; j < PX + 2 * FAT; ++j) w[j] = sC[(lx0 + j) * LZ + lane + FAT];

    float nv[PX];
    float ost[(MODE & M_WIN) ? PX : 1], cur[(MODE & M_WIN) ? PX : 1];
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        const int gx = x0 + lx0 + j;
        const float* s = &sC[(lx0 + j + FAT) * LZ + lane + FAT];
        const float cc = w[j + FAT];
        float lap = a.zc[0] * cc
                  + a.zc[1] * (s[-1] + s[1])
                  + a.zc[2] * (s[-2] + s[2])
                  + a.zc[3] * (s[-3] + s[3])
                  + a.zc[4] * (s[-4] + s[4])
                  + a.zc[5] * (s[-5] + s[5]);
        if constexpr (ADJ) {
            // stepAdjFsGpu (F/src/kernelsGpu.cu:613-700): below a free
            // surface the four rows drop the z terms the forward folds onto them.
            if (FS && gz <= FAT + 4) {
                if (gz == FAT + 1) lap = lap - a.zc[2] * s[0] - a.zc[3] * s[1] - a.zc[4] * s[2] - a.zc[5] * s[3];
                else if (gz == FAT + 2) lap = lap - a.zc[3] * s[-1] - a.zc[4] * s[0] - a.zc[5] * s[1];
                else if (gz == FAT + 3) lap = lap - a.zc[4] * s[-2] - a.zc[5] * s[-1];
                else if (gz == FAT + 4) lap = lap - a.zc[5] * s[-3];
            }
        }
        lap = lap
            + a.xc[0] * cc
            + a.xc[1] * (w[j + 6] + w[j + 4])
            + a.xc[2] * (w[j + 7] + w[j + 3])
            + a.xc[3] * (w[j + 8] + w[j + 2])
            + a.xc[4] * (w[j + 9] + w[j + 1])
            + a.xc[5] * (w[j + 10] + w[j + 0]);
        float n = 0.f;
        if (ok[j]) {
            const bool surf = FS && gz == FAT;
            const float d = dmp[j];
            // forward: the reference zeroed the old slice's surface value when
            // it was the current slice
            const float osj = (!ADJ && surf) ? 0.f : os[j];
            const float o = osj * d;                 // second damping of the old slice
            if constexpr (ADJ) {
                const float c = pv[j];
                n = surf ? o : lap + c + c - o;      // the adjoint leaves the surface row as it is
                if constexpr (MODE & M_WIN) cur[j] = c;
            } else {
                n = pv[j] * lap + cc + cc - o;
                // The window takes the slices as stored (their values when
                // they were new); at a free surface the staged current value
                // is the mirrored one, so the stored one is read there.
                if constexpr (MODE & M_WIN) cur[j] = surf ? pc[(unsigned)(gx * nz + gz)] : cc;
            }
            if constexpr (MODE & M_WIN) ost[j] = os[j];
            if constexpr (SSB) {
                const float sa = mm[j] * sA[j];
                if (useB) {
                    const float sb = mm[j] * sB[j];
                    n += sa * a.fa + sb * a.fb;
                } else {
                    n += sa * a.fa;
                }
            }
            if constexpr (SSP || SSS) {
                if (useB) n += sA[j] * a.fa + sB[j] * a.fb;
                else n += sA[j] * a.fa;
            }
        }
        nv[j] = n;
    }

    const int t = tile;
    // Receiver / source entries of the tile for the records after the
    // stores, read before them (see the buffer pointers below).
    int postE0 = 0, postE1 = 0;
    const int2* postTe = nullptr;
    gf* postSig = nullptr;
    gf* postDreg = nullptr;
    if constexpr (POST_SPARSE) {
        const int* tp = (MODE & M_REC_SIG) ? S.srcTile : S.recTile;
        postTe = (MODE & M_REC_SIG) ? S.srcEnt : S.recEnt;
        postE0 = tp[t];
        postE1 = tp[t + 1];
        postSig = G(S.sig);
        postDreg = G(S.dreg);
    }
    if constexpr (PRE_SPARSE) {
        const int* tp = (MODE & M_INJ_SIG) ? S.srcTile : S.recTile;
        const int2* te = (MODE & M_INJ_SIG) ? S.srcEnt : S.recEnt;
        const int e0 = tp[t], e1 = tp[t + 1];
        if (e1 > e0) {                               // block-uniform
#pragma unroll
            for (int j = 0; j < PX; ++j) sN[(lx0 + j) * TZ + lane] = nv[j];
            __syncthreads();
            for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
                const int2 en = te[e];
                float v;
                if constexpr (MODE & M_INJ_SIG) {
                    v = G(S.sig)[(long)en.y * a.ntw + a.itwInj];
                } else {
                    cgf* dr = G(S.dreg) + ((long)en.y * a.nts + a.its);
                    v = dr[0] * a.fc + dr[1] * a.fd;
                }
                sN[en.x] += v;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < PX; ++j) nv[j] = sN[(lx0 + j) * TZ + lane];
            __syncthreads();
        }
    }

    // The shot's buffers written below, read from its slot before the first
    // store: the compiler cannot tell that the stores leave the slot table
    // alone, and would read them again before every store.
    constexpr bool CAP_SS = ADJ && !(MODE & M_CAP_W);
    constexpr bool CAP_ALT = !ADJ && (MODE & M_CAP_ALT);
    gf* capA = nullptr;
    gf* capB = nullptr;
    long capOffA = 0, capOffB = 0;
    if constexpr (MODE & M_CAPTURE) {
        // Forward passes capture the coarse source wavefield W; the
        // adjoint receiver pass accumulates into the two ss slices.
        capA = G(CAP_SS ? S.ss[a.ssA] : CAP_ALT ? S.img : S.W);
        capB = G(CAP_SS ? S.ss[a.ssB] : CAP_ALT ? S.img : S.W);
        capOffA = CAP_SS ? a.ssOffA : a.wA;
        capOffB = CAP_SS ? a.ssOffB : a.wB;
    }
    gf* imgI = nullptr;       // image of the imaging condition and its W slice
    cgf* imgW = nullptr;
    if constexpr (MODE & M_IMAGE) {
        if (a.image) {
            imgI = G(S.img);
            imgW = G(S.W) + a.wImg;
        }
    }
    gf* winDst = nullptr;     // window capture: destination, and the W slice of WIN_IMG
    cgf* winW = nullptr;
    if constexpr (MODE & M_WIN) {
        if (a.win) {
            if (a.winOut == WIN_W || a.winOut == WIN_SS) {
                winDst = G(S.W) + a.wA;
            } else if (a.winOut == WIN_IMG) {
                winDst = G(S.img);
                winW = G(S.W) + a.wImg;
            } else {
                winDst = G(S.ss[a.ssA]) + a.ssOffA;
            }
        }
    }
    gf* const A = capA ? capA + capOffA : nullptr;
    gf* const B = capB ? capB + capOffB : nullptr;
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        if (!ok[j]) continue;
        const int gx = x0 + lx0 + j;
        const unsigned gi = (unsigned)(gx * nz + gz);
        const float n = nv[j] * dmp[j];
        nv[j] = n;
        if constexpr (MODE & M_CAPTURE) {
            // A zero weight leaves the slice as it is (x + n*0 == x); a
            // first touch writes instead of accumulating into a zeroed slice.
            if (!a.skipA) A[gi] = a.assignA ? n * a.fe : A[gi] + n * a.fe;
            else if (a.assignA) A[gi] = 0.f;
            float bnew = 0.f;
            if (!a.skipB) {
                bnew = a.assignB ? n * a.ff : B[gi] + n * a.ff;
                B[gi] = bnew;
            } else if (a.assignB) {
                B[gi] = 0.f;
            } else if constexpr (MODE & M_IMAGE) {
                bnew = B[gi];
            }
            if constexpr (MODE & M_IMAGE) {
                if (a.image) imgI[gi] += imgW[gi] * bnew;
            }
        }
        if constexpr (MODE & M_WIN) {
            if (a.win) {
                // Coarse sample from three consecutive fine slices, accumulated
                // in the reference's order (old, current, new).
                const float v = (ost[j] * a.wa + cur[j] * a.wb) + n * a.wc;
                switch (a.winOut) {
                case WIN_W: winDst[gi] = v; break;
                case WIN_SS: winDst[gi] = G(a.mtil)[gi] * v; break;
                case WIN_IMG: winDst[gi] += winW[gi] * v; break;
                default: winDst[gi] = v; break;
                }
            }
        }
        po[gi] = n;
    }

    if constexpr (POST_SPARSE) {
        const int2* te = postTe;
        const int e0 = postE0, e1 = postE1;
        if (e1 > e0) {                               // block-uniform
#pragma unroll
            for (int j = 0; j < PX; ++j) sN[(lx0 + j) * TZ + lane] = nv[j];
            __syncthreads();
            for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
                const int2 en = te[e];
                const float n = sN[en.x];
                if constexpr (MODE & M_REC_SIG) {
                    postSig[(long)en.y * a.ntw + a.itwRec] += n;
                } else {
                    gf* dr = postDreg + ((long)en.y * a.nts + a.its);
                    dr[0] += n * a.fe;
                    dr[1] += n * a.ff;
                }
            }
        }
    }
}

// Sparse operations on points of the FAT ring, which the stepper never
// updates (the reference still injects/records there).  One thread per entry,
// launched after k_step; block y = shot.
template <unsigned MODE>
__global__ void k_ring(const StepArgs a)
{
    const ShotSlot& S = a.slots[blockIdx.y];
    float* p = S.p[a.cur ^ 1];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (MODE & (M_INJ_SIG | M_REC_SIG)) {
        if (i < S.nSrcRing) {
            const int2 en = S.srcRing[i];
            if constexpr (MODE & M_INJ_SIG) p[en.x] += S.sig[(long)en.y * a.ntw + a.itwInj];
            if constexpr (MODE & M_REC_SIG) S.sig[(long)en.y * a.ntw + a.itwRec] += p[en.x];
        }
    }
    if constexpr (MODE & (M_INJ_DATA | M_REC_DATA)) {
        if (i < S.nRecRing) {
            const int2 en = S.recRing[i];
            float* dr = S.dreg + (long)en.y * a.nts + a.its;
            if constexpr (MODE & M_INJ_DATA) p[en.x] += dr[0] * a.fc + dr[1] * a.fd;
            if constexpr (MODE & M_REC_DATA) {
                const float n = p[en.x];
                dr[0] += n * a.fe;
                dr[1] += n * a.ff;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Coarse-rate and per-shot helper kernels.  Thread per interior point, block
// y = shot slot unless noted.

__global__ void k_fill(float* __restrict__ p, long n, float v)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = v;
}

// scaleReflectivity (F/src/kernelsGpu.cu:246-252): out = m * (v2dt2 * refl) on
// the interior, 0 on the FAT ring.
__global__ void k_scale_model(const float* __restrict__ m, float* __restrict__ out,
                              const float* __restrict__ v2dt2, const float* __restrict__ refl,
                              int nz, int nx)
{
    const long n = (long)nz * nx;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int iz = int(i % nz), ix = int(i / nz);
        const bool in = iz >= FAT && iz < nz - FAT && ix >= FAT && ix < nx - FAT;
        out[i] = in ? m[i] * (v2dt2[i] * refl[i]) : 0.f;
    }
}

// scaleReflectivityLinExt (F/src/kernelsGpu.cu:256-267): out[e] = m[e]*refl on
// the interior of every extension slice, 0 on the ring.
__global__ void k_scale_ext(const float* __restrict__ m, float* __restrict__ out,
                            const float* __restrict__ refl, int nz, int nx, int nExt)
{
    const long nzx = (long)nz * nx, n = nzx * nExt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long s = i % nzx;
        const int iz = int(s % nz), ix = int(s / nz);
        const bool in = iz >= FAT && iz < nz - FAT && ix >= FAT && ix < nx - FAT;
        out[i] = in ? m[i] * refl[s] : 0.f;
    }
}

// End of the Born adjoint receiver pass for every shot slot: the last
// imagingAdjGpu (W[0] * ssR) then scaleReflectivity
// (F/src/BornShotsGpuFunctions.cu:942-945).
__global__ void k_born_adj_finish(const ShotSlot* slots, int ssR, const float* __restrict__ v2dt2,
                                  const float* __restrict__ refl, int nz, int nx)
{
    const ShotSlot& S = slots[blockIdx.y];
    const long nzx = (long)nz * nx;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nzx; i += (long)gridDim.x * blockDim.x) {
        const int iz = int(i % nz), ix = int(i / nz);
        if (iz < FAT || iz >= nz - FAT || ix < FAT || ix >= nx - FAT) continue;
        float m = S.img[i];
        if (ssR >= 0) m += S.W[i] * S.ss[ssR][i];
        S.img[i] = m * (v2dt2[i] * refl[i]);
    }
}

// ---------------------------------------------------------------------------
// Offset-extended Born over a batch of coarse samples.  One launch handles
// a.nk samples; every output element sees the same sequence of float
// operations as one launch per sample taken in the listed order, so the
// batching changes the HBM traffic and not the bits.
//
// Tiles are kExtTZ rows by kExtTX columns; 256 threads, each owning one row
// and kExtCX consecutive columns.  HX >= hExt bounds the offsets at compile
// time so that the per-thread offset loops index registers only.

constexpr int kExtTZ = 64, kExtTX = 16, kExtCX = 4;
constexpr int kExtSsSpb = 2;   // shots per block of k_ext_ss_tile: the model registers serve 2 * nk tiles
static_assert(kExtTX == 4 * kExtCX, "four column groups of kExtCX columns");

// imagingOffsetFwdGpu + scaleSecondarySourceFd (F/src/kernelsGpu.cu:381-396,
// 270-275): for each shot and sample k,
//   ss_k[x] = v2dt2[x] * sum_{h = lo..hi} m~[h][x-h] * W_k[x-2h]
//   lo = max(-hExt, x+1+FAT+hExt-nx),  hi = min(x-hExt-FAT, hExt)
// written into ss[0] + ssOff[k].  A block walks (shot, sample) pairs of its
// shot group; the scaled model values of the thread's columns stay in
// registers for all of them, and the next pair's W tile is fetched into
// registers while the current one is computed.
template <int HX>
__global__ __launch_bounds__(256) void k_ext_ss_tile(ExtBatchArgs a, int nshots, int nbx, int nby, int nbz)
{
    constexpr int NE = 2 * HX + 1;
    constexpr int LW = kExtTX + 4 * HX;        // LDS columns at hExt == HX
    constexpr int WW = kExtCX + 4 * HX;        // register window of W_k columns
    constexpr int PF = LW / 4;                 // tile columns fetched per thread (its own row)
    static_assert(LW % 4 == 0, "four column groups fetch the tile");
    __shared__ float Wl[LW * kExtTZ];
    Blk3 blk;
    if (!xcd_block3(nbx, nby, nbz, blk)) return;     // padding blocks (block-uniform)
    const int hExt = a.hExt, ne = 2 * hExt + 1, lw = kExtTX + 4 * hExt;
    const long nz = a.nz, nzx = nz * a.nx;
    const int tz = threadIdx.x, g = threadIdx.y;
    const int z = FAT + blk.z * kExtTZ + tz;
    const bool zok = z < a.nz - FAT;
    const int x0 = FAT + blk.x * kExtTX;        // first output column of the tile
    const int xb = x0 + kExtCX * g;
    int lo[kExtCX], hi[kExtCX];
    float mv[NE][kExtCX];
#pragma unroll
    for (int c = 0; c < kExtCX; ++c) {
        const int x = xb + c;
        lo[c] = max(-hExt, x + 1 + FAT + hExt - a.nx);
        hi[c] = min(x - hExt - FAT, hExt);
        if (!zok || x >= a.nx - FAT) hi[c] = lo[c] - 1;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int h = e - hExt;
            mv[e][c] = (e < ne && h >= lo[c] && h <= hi[c]) ? G(a.mext)[nzx * e + (long)(x - h) * nz + z] : 0.f;
        }
    }
    const int sh0 = blk.y * kExtSsSpb;
    const int nsh = max(0, min(kExtSsSpb, nshots - sh0));
    const int nIt = nsh * a.nk;
    // The shots' W and ss bases and the thread's v2dt2 values are loaded
    // once: in the sample loop the only global loads are the next pair's W
    // tile, so waiting for the current tile never waits for the next one.
    cgf* W0 = G(nsh > 0 ? a.slots[sh0].W : nullptr);
    cgf* W1 = G(nsh > 1 ? a.slots[sh0 + 1].W : nullptr);
    gf* S0 = G(nsh > 0 ? a.slots[sh0].ss[0] : nullptr);
    gf* S1 = G(nsh > 1 ? a.slots[sh0 + 1].ss[0] : nullptr);
    float vs[kExtCX];
#pragma unroll
    for (int c = 0; c < kExtCX; ++c) {
        const int x = xb + c;
        vs[c] = (zok && x < a.nx - FAT) ? G(a.v2dt2)[(long)x * nz + z] : 0.f;
    }
    // Wl column j holds global column x0 - 2*hExt + j; thread (g, tz)
    // fetches columns j = 4u + g of its own row.  Columns outside the
    // interior are taken as 0: the in-range offsets never reach them, and
    // the out-of-range ones carry a zero model value.
    float pre[PF];
    auto fetch = [&](int q, int k) {
        cgf* W = (q ? W1 : W0) + a.wOff[k];
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int j = 4 * u + g, gx = x0 - 2 * hExt + j;
            pre[u] = (zok && j < lw && gx >= FAT && gx < a.nx - FAT) ? W[(long)gx * nz + z] : 0.f;
        }
    };
    // Window entry i holds column xb + i - 4*HX + 2*hExt, so W_k[x - 2h] of
    // column c and offset index e is entry c - 2e + 4*HX.
    const int jb = xb - x0 - 4 * HX + 4 * hExt;
    int q = 0, k = 0;
    if (nIt > 0) fetch(0, 0);
    for (int it = 0; it < nIt; ++it) {
        int qn = q, kn = k + 1;
        if (kn == a.nk) {
            kn = 0;
            ++qn;
        }
#pragma unroll
        for (int u = 0; u < PF; ++u) Wl[(4 * u + g) * kExtTZ + tz] = pre[u];
        __syncthreads();
        if (it + 1 < nIt) fetch(qn, kn);
        float ww[WW];
#pragma unroll
        for (int i = 0; i < WW; ++i) {
            const int j = jb + i;
            ww[i] = (j >= 0 && j < lw) ? Wl[j * kExtTZ + tz] : 0.f;
        }
        __syncthreads();
        if (zok) {
            gf* ss = (q ? S1 : S0) + a.ssOff[k];
#pragma unroll
            for (int c = 0; c < kExtCX; ++c) {
                const int x = xb + c;
                if (x >= a.nx - FAT) continue;
                // Offsets outside [lo, hi] have mv == 0 and a finite W, so
                // their products add a zero: the sum is that over [lo, hi]
                // in the same order, without a mask per product.
                float acc = 0.f;
#pragma unroll
                for (int e = 0; e < NE; ++e) acc += mv[e][c] * ww[c - 2 * e + 4 * HX];
                ss[(long)x * nz + z] = acc * vs[c];
            }
        }
        q = qn;
        k = kn;
    }
}

// scaleSecondarySourceFd + imagingOffsetAdjGpu (F/src/kernelsGpu.cu:270-275,
// 398-411): for each shot and each sample k in order,
//   img[e][x] += (ss_k[x+h] * v2dt2[x+h]) * W_k[x-h],  h = e - hExt,
// for x in [FAT+hExt, nx-FAT-hExt).  With WF (imagingOffsetWemvaScaleFwdGpu,
// :429-443, the WEMVA forward) the product is (W_k[x-h] * ss_k[x+h]) *
// v2dt2[x+h] instead.  The thread's image values of a shot stay in
// registers across the batch, so the extended image is read and written
// once per batch; the two operands pass through LDS, the next pair's tiles
// fetched into registers while the current one is computed.  A workgroup
// images NG of the offsets (offset group = fastest block index, so the
// groups of one tile run side by side and share its tiles in L2): fewer
// image registers per thread, more waves in flight.
// 21 offsets (HX 10) in groups of 7, imaging (not WF): compiled to 4 waves
// per SIMD (128 VGPRs, two registers spilled), faster on MI355X than the
// compiler's own budget (129 VGPRs, 3 waves).  The WF form would spill six
// at that budget and keeps the compiler's.
template <int HX, int NG, bool WF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HX == 10 && !WF ? 4 : 1, 8))) void k_ext_img_tile(ExtBatchArgs a, int nbx, int ngrp, int nby, int nbz)
{
    constexpr int LW = kExtTX + 2 * HX;
    constexpr int RW = kExtCX + NG - 1;        // register windows of R and W columns
    constexpr int PF = (LW + 3) / 4;
    __shared__ float Rl[LW * kExtTZ];
    __shared__ float Wl[LW * kExtTZ];
    Blk3 blk;                                   // blk.x: x tile * ngrp + offset group, blk.y: the shot
    if (!xcd_block3(nbx * ngrp, nby, nbz, blk)) return;     // padding blocks (block-uniform)
    const int grp = blk.x % ngrp, bx = blk.x / ngrp;
    const int e0 = grp * NG;                    // the group's first offset index
    const int hExt = a.hExt, ne = 2 * hExt + 1, lw = kExtTX + 2 * hExt;
    const long nz = a.nz, nzx = nz * a.nx;
    const int tz = threadIdx.x, g = threadIdx.y;
    const int z = FAT + blk.z * kExtTZ + tz;
    const int x0 = FAT + hExt + bx * kExtTX;
    const int xEnd = a.nx - FAT - hExt;
    const bool zok = z < a.nz - FAT;
    const int xb = x0 + kExtCX * g;
    // The shot's bases are loaded once: in the sample loop the only global
    // loads are the next sample's tiles, so waiting for the current tiles
    // never waits for the next ones.
    const ShotSlot& S = a.slots[blk.y];
    cgf* ssB = G(S.ss[0]);
    cgf* WB = G(S.W);
    gf* img = G(S.img);
    // Rl/Wl column j holds global column x0 - hExt + j; thread (g, tz)
    // fetches columns j = 4u + g of its own row.
    float vv[PF], pr[PF], pw[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int j = 4 * u + g, gx = x0 - hExt + j;
        vv[u] = (!WF && zok && j < lw && gx < a.nx) ? G(a.v2dt2)[(long)gx * nz + z] : 0.f;
    }
    auto fetch = [&](int k) {
        cgf* ss = ssB + a.ssOff[k];
        cgf* W = WB + a.wOff[k];
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int j = 4 * u + g, gx = x0 - hExt + j;
            const bool in = zok && j < lw && gx < a.nx;
            pr[u] = in ? ss[(long)gx * nz + z] : 0.f;
            pw[u] = in ? W[(long)gx * nz + z] : 0.f;
        }
    };
    // Offset e = e0 + q (q < NG), column xb + c: R at xb + c + h is R window
    // entry c + q (column xb + e0 + i - hExt for entry i); W at xb + c - h is
    // W window entry c - q + NG - 1 (column xb - e0 - NG + 1 + i + hExt).
    const int jr = xb - x0 + e0;
    const int jw = xb - x0 + 2 * hExt - e0 - NG + 1;
    float vw[WF ? RW : 1];                       // WF: v2dt2 on the R window
    if constexpr (WF) {
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const int gx = xb + e0 + i - hExt;
            vw[i] = (zok && gx < a.nx) ? G(a.v2dt2)[(long)gx * nz + z] : 0.f;
        }
    }
    float m[NG][kExtCX];
#pragma unroll
    for (int q = 0; q < NG; ++q)
#pragma unroll
        for (int c = 0; c < kExtCX; ++c)
            m[q][c] = (e0 + q < ne && zok && xb + c < xEnd) ? img[nzx * (e0 + q) + (long)(xb + c) * nz + z] : 0.f;
    fetch(0);
    for (int k = 0; k < a.nk; ++k) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            if (4 * u + g < LW) {
                Rl[(4 * u + g) * kExtTZ + tz] = WF ? pr[u] : pr[u] * vv[u];
                Wl[(4 * u + g) * kExtTZ + tz] = pw[u];
            }
        }
        __syncthreads();
        if (k + 1 < a.nk) fetch(k + 1);
        float rw[RW], ww[RW];
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const int j0 = jr + i, j1 = jw + i;
            rw[i] = j0 < lw ? Rl[j0 * kExtTZ + tz] : 0.f;
            ww[i] = j1 >= 0 ? Wl[j1 * kExtTZ + tz] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NG; ++q)
#pragma unroll
            for (int c = 0; c < kExtCX; ++c)
                if (e0 + q < ne) {
                    if constexpr (WF) m[q][c] += (ww[c - q + NG - 1] * rw[c + q]) * vw[c + q];
                    else m[q][c] += rw[c + q] * ww[c - q + NG - 1];
                }
    }
    // The store addresses are formed again here rather than kept live
    // across the sample loop from the loads above.
    float* imgOut = S.img;
    asm volatile("" : "+v"(imgOut));
#pragma unroll
    for (int q = 0; q < NG; ++q)
#pragma unroll
        for (int c = 0; c < kExtCX; ++c)
            if (e0 + q < ne && zok && xb + c < xEnd) G(imgOut)[nzx * (e0 + q) + (long)(xb + c) * nz + z] = m[q][c];
}

// The same two operations for offsets beyond the register-tiled kernels:
// one thread per grid point, operands read through the caches.
__global__ void k_ext_ss_any(ExtBatchArgs a)
{
    const ShotSlot& S = a.slots[blockIdx.y];
    const long nz = a.nz, nzx = nz * a.nx;
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + blockIdx.z;
    if (iz >= a.nz - FAT) return;
    const int hExt = a.hExt;
    const int lo = max(-hExt, ix + 1 + FAT + hExt - a.nx);
    const int hi = min(ix - hExt - FAT, hExt);
    const long i = (long)ix * nz + iz;
    for (int k = 0; k < a.nk; ++k) {
        const float* W = S.W + a.wOff[k];
        float acc = 0.f;
        for (int h = lo; h <= hi; ++h)
            acc += a.mext[nzx * (h + hExt) + (long)(ix - h) * nz + iz] * W[(long)(ix - 2 * h) * nz + iz];
        S.ss[0][a.ssOff[k] + i] = acc * a.v2dt2[i];
    }
}

template <bool WF>
__global__ void k_ext_img_any(ExtBatchArgs a)
{
    const ShotSlot& S = a.slots[blockIdx.y];
    const long nz = a.nz, nzx = nz * a.nx;
    const int hExt = a.hExt;
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + hExt + blockIdx.z;
    if (iz >= a.nz - FAT || ix >= a.nx - FAT - hExt) return;
    for (int e = 0; e <= 2 * hExt; ++e) {
        const int h = e - hExt;
        const long ir = (long)(ix + h) * nz + iz, iw = (long)(ix - h) * nz + iz;
        float* mp = S.img + nzx * e + (long)ix * nz + iz;
        float m = *mp;
        for (int k = 0; k < a.nk; ++k) {
            if constexpr (WF) m += (S.W[a.wOff[k] + iw] * S.ss[0][a.ssOff[k] + ir]) * a.v2dt2[ir];
            else m += (S.ss[0][a.ssOff[k] + ir] * a.v2dt2[ir]) * S.W[a.wOff[k] + iw];
        }
        *mp = m;
    }
}

// Final ext-adjoint scaling by 2/v^3 (scaleReflectivityLinExt) per slot.
__global__ void k_ext_adj_finish(const ShotSlot* slots, const float* __restrict__ refl,
                                 int nz, int nx, int nExt)
{
    const ShotSlot& S = slots[blockIdx.y];
    const long nzx = (long)nz * nx, n = nzx * nExt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long s = i % nzx;
        const int iz = int(s % nz), ix = int(s / nz);
        if (iz < FAT || iz >= nz - FAT || ix < FAT || ix >= nx - FAT) continue;
        S.img[i] *= refl[s];
    }
}

// ---------------------------------------------------------------------------
// Time-lag extension (extension=time of F/src/BornExtShotsGpuFunctions.cu,
// tomoExtShotsGpuFunctions.cu, wemvaExtShotsGpuFunctions.cu).  Lag slice e
// pairs sample k with W sample k - lagDir * 2 (e - hExt).  Both kernels take
// one thread per interior point of one shot and walk the lag slices in
// ascending order; the nk samples of the batch live in registers, so each
// model (or image) value is touched once per batch.  The grid is (z blocks,
// shots, x columns): workgroups dispatched together share the model slices
// of one column across shots (a shot-major grid measured slower).

// imagingTimeFwdGpu (F/src/kernelsGpu.cu:337-350, lagDir = +1) and
// imagingTimeTomoAdjGpu (:445-458, lagDir = -1) for nk samples:
// ss_j = sum over e in [lo_j, hi_j), ascending, of mext[e] * W[k_j - lagDir * 2 (e - hExt)].
__global__ void k_time_ss(ExtBatchArgs a)
{
    const ShotSlot& S = a.slots[blockIdx.y];
    const long nz = a.nz, nzx = nz * a.nx;
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + blockIdx.z;
    if (iz >= a.nz - FAT) return;
    const long i = (long)ix * nz + iz;
    cgf* W = G(S.W) + i;
    cgf* M = G(a.mext) + i;
    gf* SS = G(S.ss[0]) + i;
    float acc[kExtBatchMax];
#pragma unroll
    for (int j = 0; j < kExtBatchMax; ++j) acc[j] = 0.f;
    for (int e = a.eLo; e < a.eHi; ++e) {
        const float mv = M[nzx * e];
        const long sh = 2L * a.lagDir * (e - a.hExt) * nzx;
#pragma unroll
        for (int j = 0; j < kExtBatchMax; ++j)
            if (j < a.nk && e >= a.lo[j] && e < a.hi[j]) acc[j] += mv * W[a.wOff[j] - sh];
    }
#pragma unroll
    for (int j = 0; j < kExtBatchMax; ++j)
        if (j < a.nk) SS[a.ssOff[j]] = acc[j];
}

// imagingTimeAdjGpu (F/src/kernelsGpu.cu:352-364, lagDir = +1) and
// imagingWemvaTimeAdjGpu (:366-378, lagDir = -1) for nk samples, in batch
// order: img[e] += ss_j * W[k_j - lagDir * 2 (e - hExt)] for e in [lo_j, hi_j).
__global__ void k_time_img(ExtBatchArgs a)
{
    const ShotSlot& S = a.slots[blockIdx.y];
    const long nz = a.nz, nzx = nz * a.nx;
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + blockIdx.z;
    if (iz >= a.nz - FAT) return;
    const long i = (long)ix * nz + iz;
    cgf* W = G(S.W) + i;
    cgf* SS = G(S.ss[0]) + i;
    gf* IMG = G(S.img) + i;
    float sv[kExtBatchMax];
#pragma unroll
    for (int j = 0; j < kExtBatchMax; ++j) sv[j] = j < a.nk ? SS[a.ssOff[j]] : 0.f;
    for (int e = a.eLo; e < a.eHi; ++e) {
        const long sh = 2L * a.lagDir * (e - a.hExt) * nzx;
        float m = IMG[nzx * e];
#pragma unroll
        for (int j = 0; j < kExtBatchMax; ++j)
            if (j < a.nk && e >= a.lo[j] && e < a.hi[j]) m += sv[j] * W[a.wOff[j] - sh];
        IMG[nzx * e] = m;
    }
}

// The same two operations for a full interior batch: NK consecutive samples
// (k_j = k0 + DIRK * j) with every lag |l| <= hExt <= HX usable.  The W
// samples the batch touches, k0 - 2 hExt ... k0 + NK - 1 + 2 hExt, are read
// once into a register window; slot r of the window holds W sample
// kmin - 2 HX + r, so that every (sample, lag) pair has a compile-time
// register index.  Each output receives the general kernels' operations in
// their order (lags ascending; samples in batch order).
template <int HX, int NK, int DIRK, int LAGDIR>
__device__ __forceinline__ void time_window(const ExtBatchArgs& a, cgf* W, long nzx, float (&w)[NK + 4 * HX])
{
    const long kmin = DIRK > 0 ? a.wOff[0] / nzx : a.wOff[0] / nzx - (NK - 1);
    const int h = a.hExt;
#pragma unroll
    for (int r = 0; r < NK + 4 * HX; ++r) {
        const bool need = r >= 2 * (HX - h) && r < NK + 2 * HX + 2 * h;
        w[r] = need ? W[(kmin - 2 * HX + r) * nzx] : 0.f;
    }
}

template <int HX, int NK, int LAGDIR>
__global__ void __launch_bounds__(64) k_time_ss_win(ExtBatchArgs a)
{
    const ShotSlot& S = a.slots[blockIdx.y];
    const long nz = a.nz, nzx = nz * a.nx;
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + blockIdx.z;
    if (iz >= a.nz - FAT) return;
    const long i = (long)ix * nz + iz;
    float w[NK + 4 * HX];
    time_window<HX, NK, 1, LAGDIR>(a, G(S.W) + i, nzx, w);
    const int h = a.hExt;
    cgf* M = G(a.mext) + i;
    float mv[2 * HX + 1];   // every lag's model value in flight before the arithmetic
#pragma unroll
    for (int l = -HX; l <= HX; ++l) mv[l + HX] = (l >= -h && l <= h) ? M[nzx * (l + h)] : 0.f;
    float acc[NK];
#pragma unroll
    for (int j = 0; j < NK; ++j) acc[j] = 0.f;
#pragma unroll
    for (int l = -HX; l <= HX; ++l) {
        if (l < -h || l > h) continue;
#pragma unroll
        for (int j = 0; j < NK; ++j) acc[j] += mv[l + HX] * w[j - 2 * LAGDIR * l + 2 * HX];
    }
    gf* SS = G(S.ss[0]) + i;
#pragma unroll
    for (int j = 0; j < NK; ++j) SS[a.ssOff[j]] = acc[j];
}

template <int HX, int NK, int DIRK, int LAGDIR>
__global__ void __launch_bounds__(64) k_time_img_win(ExtBatchArgs a)
{
    const ShotSlot& S = a.slots[blockIdx.y];
    const long nz = a.nz, nzx = nz * a.nx;
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + blockIdx.z;
    if (iz >= a.nz - FAT) return;
    const long i = (long)ix * nz + iz;
    float w[NK + 4 * HX];
    time_window<HX, NK, DIRK, LAGDIR>(a, G(S.W) + i, nzx, w);
    cgf* SS = G(S.ss[0]) + i;
    gf* IMG = G(S.img) + i;
    float sv[NK];
#pragma unroll
    for (int j = 0; j < NK; ++j) sv[j] = SS[a.ssOff[j]];
    const int h = a.hExt;
    float m[2 * HX + 1];    // every lag's image value in flight before the arithmetic
#pragma unroll
    for (int l = -HX; l <= HX; ++l) m[l + HX] = (l >= -h && l <= h) ? IMG[nzx * (l + h)] : 0.f;
#pragma unroll
    for (int l = -HX; l <= HX; ++l) {
        if (l < -h || l > h) continue;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            const int jj = DIRK > 0 ? j : NK - 1 - j;   // k_j - kmin
            m[l + HX] += sv[j] * w[jj - 2 * LAGDIR * l + 2 * HX];
        }
        IMG[nzx * (l + h)] = m[l + HX];
    }
}

// Per-point scale of the time-lag model: scaleReflectivityExt
// (F/src/kernelsGpu.cu:277-288), 2/v^3 * v^2 dtw^2, or with sloth
// scaleReflectivitySlothExt (:290-302), -v^2 dtw^2.
__global__ void k_time_scale(const float* __restrict__ v2dt2, const float* __restrict__ refl, int sloth,
                             float* __restrict__ out, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = sloth ? -v2dt2[i] : refl[i] * v2dt2[i];
}

// ---------------------------------------------------------------------------
// Offset-extended tomographic operator (F/src/tomoExtShotsGpuFunctions.cu).

// computeTomoSrcWfldDt2 + srcWfldSecondTimeDerivative (:1251-1360,
// F/src/kernelsGpu.cu:188-197): W[k] <- cSide*(W[k-1] + W[k+1]) +
// cCenter*W[k], zero beyond both ends, in place; one thread walks the nts
// samples of one interior point.  The samples are read in chunks of 16, all
// loads of a chunk (samples k0+1 .. k0+16, still unwritten) before its
// stores (k0 .. k0+15), so a chunk waits for its loads once.
__global__ void k_dt2_coarse(const ShotSlot* slots, int nts, float cSide, float cCenter, int nz, int nx)
{
    constexpr int D = 16;
    const ShotSlot& S = slots[blockIdx.y];
    const long nzx = (long)nz * nx;
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + blockIdx.z;
    if (iz >= nz - FAT) return;
    gf* w = G(S.W) + (long)ix * nz + iz;
    float prev = 0.f, cur = w[0];
    for (int k0 = 0; k0 < nts; k0 += D) {
        float nxt[D];
#pragma unroll
        for (int j = 0; j < D; ++j) nxt[j] = k0 + 1 + j < nts ? w[(long)(k0 + 1 + j) * nzx] : 0.f;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            if (k0 + j >= nts) break;
            w[(long)(k0 + j) * nzx] = cSide * (prev + nxt[j]) + cCenter * cur;
            prev = cur;
            cur = nxt[j];
        }
    }
}

// imagingFwdGpu (F/src/kernelsGpu.cu:309-316) into ring slots:
// ss[0][ssOff[k]] = m~ * W[wOff[k]] at interior points.
__global__ void k_mul_batch(ExtBatchArgs a)
{
    const ShotSlot& S = a.slots[blockIdx.y];
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + blockIdx.z;
    if (iz >= a.nz - FAT) return;
    const long i = (long)ix * a.nz + iz;
    for (int k = 0; k < a.nk; ++k) S.ss[0][a.ssOff[k] + i] = a.mext[i] * S.W[a.wOff[k] + i];
}

// imagingOffsetTomoAdjGpu (F/src/kernelsGpu.cu:460-475) into ring slots:
// ss[x] = sum_e r[e][x+e] * R[k][x+2e] * v2dt2[x+2e], R in W[wOff[k]].
__global__ void k_tomo_adj_ss(ExtBatchArgs a)
{
    const ShotSlot& S = a.slots[blockIdx.y];
    const long nz = a.nz, nzx = nz * a.nx;
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + blockIdx.z;
    if (iz >= a.nz - FAT) return;
    const int hExt = a.hExt;
    const int lo = max(-hExt, FAT + hExt - ix);
    const int hi = min(a.nx - 1 - FAT - ix - hExt, hExt);
    const long i = (long)ix * nz + iz;
    for (int k = 0; k < a.nk; ++k) {
        const float* R = S.W + a.wOff[k];
        float acc = 0.f;
        for (int e = lo; e <= hi; ++e) {
            const long iRec = (long)(ix + 2 * e) * nz + iz;
            acc += a.mext[nzx * (e + hExt) + (long)(ix + e) * nz + iz] * R[iRec] * a.v2dt2[iRec];
        }
        S.ss[0][a.ssOff[k] + i] = acc;
    }
}

// k_tomo_adj_ss on full tiles: the same sums, in the same order, with the
// thread's model values r[e][x+h] in registers for all the samples of its
// shots, v2dt2 staged in LDS once, and each sample's R tile staged through
// LDS (the next one fetched into registers while the current one is
// computed).
// Window entry i of column c is column xb + i - 2*hExt, so the terms of
// offset index e (h = e - hExt) use entry c + 2e.
template <int HX>
__global__ __launch_bounds__(256) void k_tomo_adj_ss_tile(ExtBatchArgs a, int nshots, int nbx, int nby, int nbz)
{
    constexpr int NE = 2 * HX + 1;
    constexpr int LW = kExtTX + 4 * HX;        // LDS columns at hExt == HX
    constexpr int WN = kExtCX + 4 * HX;        // register windows of R and v2dt2 columns
    constexpr int PF = LW / 4;
    static_assert(LW % 4 == 0, "four column groups fetch the tile");
    __shared__ float Rl[LW * kExtTZ];
    __shared__ float Vl[LW * kExtTZ];            // v2dt2 on the same columns, staged once
    Blk3 blk;
    if (!xcd_block3(nbx, nby, nbz, blk)) return;     // padding blocks (block-uniform)
    const int hExt = a.hExt, ne = 2 * hExt + 1, lw = kExtTX + 4 * hExt;
    const long nz = a.nz, nzx = nz * a.nx;
    const int tz = threadIdx.x, g = threadIdx.y;
    const int z = FAT + blk.z * kExtTZ + tz;
    const bool zok = z < a.nz - FAT;
    const int x0 = FAT + blk.x * kExtTX;
    const int xb = x0 + kExtCX * g;
    float mv[NE][kExtCX];
#pragma unroll
    for (int c = 0; c < kExtCX; ++c) {
        const int x = xb + c;
        int lo = max(-hExt, FAT + hExt - x);
        const int hi = min(a.nx - 1 - FAT - x - hExt, hExt);
        if (!zok || x >= a.nx - FAT) lo = hi + 1;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int h = e - hExt;
            mv[e][c] = (e < ne && h >= lo && h <= hi) ? G(a.mext)[nzx * e + (long)(x + h) * nz + z] : 0.f;
        }
    }
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int j = 4 * u + g, gx = x0 - 2 * hExt + j;
        Vl[j * kExtTZ + tz] = (zok && gx >= 0 && gx < a.nx) ? G(a.v2dt2)[(long)gx * nz + z] : 0.f;
    }
    const int sh0 = blk.y * kExtSsSpb;
    const int nsh = max(0, min(kExtSsSpb, nshots - sh0));
    const int nIt = nsh * a.nk;
    cgf* R0 = G(nsh > 0 ? a.slots[sh0].W : nullptr);
    cgf* R1 = G(nsh > 1 ? a.slots[sh0 + 1].W : nullptr);
    gf* S0 = G(nsh > 0 ? a.slots[sh0].ss[0] : nullptr);
    gf* S1 = G(nsh > 1 ? a.slots[sh0 + 1].ss[0] : nullptr);
    // Rl column j holds global column x0 - 2*hExt + j; columns outside the
    // interior are taken as 0 (the in-range offsets never reach them).
    float pre[PF];
    auto fetch = [&](int q, int k) {
        cgf* R = (q ? R1 : R0) + a.wOff[k];
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int j = 4 * u + g, gx = x0 - 2 * hExt + j;
            pre[u] = (zok && j < lw && gx >= FAT && gx < a.nx - FAT) ? R[(long)gx * nz + z] : 0.f;
        }
    };
    int q = 0, k = 0;
    if (nIt > 0) fetch(0, 0);
    for (int it = 0; it < nIt; ++it) {
        int qn = q, kn = k + 1;
        if (kn == a.nk) {
            kn = 0;
            ++qn;
        }
#pragma unroll
        for (int u = 0; u < PF; ++u) Rl[(4 * u + g) * kExtTZ + tz] = pre[u];
        __syncthreads();
        if (it + 1 < nIt) fetch(qn, kn);
        float rw[WN];
#pragma unroll
        for (int i = 0; i < WN; ++i) {
            const int j = kExtCX * g + i;
            rw[i] = j < lw ? Rl[j * kExtTZ + tz] : 0.f;
        }
        __syncthreads();
        if (zok) {
            gf* ss = (q ? S1 : S0) + a.ssOff[k];
            const float* vl = Vl + kExtCX * g * kExtTZ + tz;   // v2dt2 window entry i at vl[i * kExtTZ]
#pragma unroll
            for (int c = 0; c < kExtCX; ++c) {
                const int x = xb + c;
                if (x >= a.nx - FAT) continue;
                // Offsets outside [lo, hi] have mv == 0 and finite R and
                // v2dt2: they add zeros to the same ordered sum.
                float acc = 0.f;
#pragma unroll
                for (int e = 0; e < NE; ++e) acc += (mv[e][c] * rw[c + 2 * e]) * vl[(c + 2 * e) * kExtTZ];
                ss[(long)x * nz + z] = acc;
            }
        }
        q = qn;
        k = kn;
    }
}

// imagingAdjGpu (F/src/kernelsGpu.cu:318-325), img += X.W[xOff] * Y.W[yOff]
// at interior points with img from X's slot, over nk coarse samples in the
// order given: the image element stays in a register across the samples
// (the same float operations in the same order as one sample at a time), so
// the image is read and written once per batch.  All the batch's operand
// loads are issued first.
__global__ __launch_bounds__(64) void k_img_acc_batch(const ImgAccArgs a)
{
    const ShotSlot& SX = a.X[blockIdx.y];
    const ShotSlot& SY = a.Y[blockIdx.y];
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + blockIdx.z;
    if (iz >= a.nz - FAT) return;
    const unsigned i = (unsigned)(ix * a.nz + iz);
    cgf* xw = G(SX.W);
    cgf* yw = G(SY.W);
    gf* img = G(SX.img);
    float xv[kExtBatchMax], yv[kExtBatchMax];
#pragma unroll
    for (int j = 0; j < kExtBatchMax; ++j) {
        if (j < a.nk) {
            xv[j] = xw[a.xOff[j] + i];
            yv[j] = yw[a.yOff[j] + i];
        }
    }
    float v = img[i];
#pragma unroll
    for (int j = 0; j < kExtBatchMax; ++j)
        if (j < a.nk) v += xv[j] * yv[j];
    img[i] = v;
}

// acc += img (one shot, in shot order).
__global__ void k_record_scale(const ShotSlot* slot, int ssIdx, float* __restrict__ out,
                               const float* __restrict__ refl, const float* __restrict__ v2dt2, int nz, int nx)
{
    const float* ss = slot->ss[ssIdx];
    const long nzx = (long)nz * nx;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nzx; i += (long)gridDim.x * blockDim.x) {
        const int iz = int(i % nz), ix = int(i / nz);
        if (iz < FAT || iz >= nz - FAT || ix < FAT || ix >= nx - FAT) continue;
        out[i] = (ss[i] * refl[i]) * v2dt2[i];
    }
}

__global__ void k_accumulate(float* __restrict__ acc, const float* __restrict__ img, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        acc[i] += img[i];
}

// out = (add ? out : 0) + acc, used for model += modelTemp.
__global__ void k_add_to(float* __restrict__ out, const float* __restrict__ acc, long n, int add)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (add ? out[i] : 0.f) + acc[i];
}

// deviceGpu::forward on the device (F/src/deviceGpu.cpp:226-243), one shot:
// out[dev][t] = (add ? out : 0) + sum_c w[dev,c] * reg[cornerReg[dev,c]][t].
__global__ void k_dev_forward(const float* __restrict__ reg, float* __restrict__ out,
                              const int* __restrict__ cornerReg, const float* __restrict__ weight,
                              int nDev, int nb, int nt, int add)
{
    const long n = (long)nDev * nt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int dv = int(i / nt), t = int(i - (long)dv * nt);
        float acc = 0.f;
        for (int c = 0; c < nb; ++c) {
            const int k = dv * nb + c;
            acc += weight[k] * reg[(long)cornerReg[k] * nt + t];
        }
        out[i] = (add ? out[i] : 0.f) + acc;
    }
}

// deviceGpu::adjoint on the device (F/src/deviceGpu.cpp:245-261), one shot:
// reg[r][t] = sum over the (device, corner) entries of r, in order, of w*d.
__global__ void k_dev_adjoint(float* __restrict__ reg, const float* __restrict__ irreg,
                              const int* __restrict__ adjPtr, const int* __restrict__ adjEnt,
                              const float* __restrict__ weight, int nReg, int nb, int nt)
{
    const long n = (long)nReg * nt;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int r = int(i / nt), t = int(i - (long)r * nt);
        float acc = 0.f;
        for (int e = adjPtr[r]; e < adjPtr[r + 1]; ++e) {
            const int k = adjEnt[e];
            acc += weight[k] * irreg[(long)(k / nb) * nt + t];
        }
        reg[i] = acc;
    }
}

// Last window capture after a pass: V = o*wa + c*wb from the final pair.
__global__ void k_win_tail(const ShotSlot* slots, int cur, int winOut, float wa, float wb, long wA,
                           long wImg, int ssA, const float* __restrict__ mtil, int nz, int nx)
{
    const ShotSlot& S = slots[blockIdx.y];
    const long nzx = (long)nz * nx;
    const float* c = S.p[cur];
    const float* o = S.p[cur ^ 1];
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nzx; i += (long)gridDim.x * blockDim.x) {
        const int iz = int(i % nz), ix = int(i / nz);
        if (iz < FAT || iz >= nz - FAT || ix < FAT || ix >= nx - FAT) continue;
        const float v = o[i] * wa + c[i] * wb;
        switch (winOut) {
        case WIN_W: S.W[wA + i] = v; break;
        case WIN_SS: S.W[wA + i] = mtil[i] * v; break;
        case WIN_IMG: S.img[i] += S.W[wImg + i] * v; break;
        default: S.ss[ssA][wA + i] = v; break;   // WIN_SLICE: wA offsets the ss slice
        }
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers (declared in kernels.hpp).

using StepKernel = void (*)(const StepArgs);

template <unsigned MODE, int SPB>
static StepKernel pick_tile_spb(int tz, int tx)
{
    if (tz == 64 && tx == 16) return k_step<64, 16, MODE, SPB>;
    if (tz == 64 && tx == 32) return k_step<64, 32, MODE, SPB>;
    if (tz == 32 && tx == 16) return k_step<32, 16, MODE, SPB>;
    if (tz == 32 && tx == 32) return k_step<32, 32, MODE, SPB>;
    return nullptr;
}

template <unsigned MODE, int W>
static StepKernel pick_tile_w(int tz, int tx)
{
    if (tz == 64 && tx == 16) return k_step_w<64, 16, MODE, W>;
    if (tz == 64 && tx == 32) return k_step_w<64, 32, MODE, W>;
    if (tz == 32 && tx == 16) return k_step_w<32, 16, MODE, W>;
    if (tz == 32 && tx == 32) return k_step_w<32, 32, MODE, W>;
    return nullptr;
}

// occ: 0 (the compiler's register budget), 6 or 8 waves per SIMD.  The
// adjoint receiver passes with imaging run at 6 when 8 is asked for: they
// spill at 64 VGPRs.
template <unsigned MODE>
static StepKernel pick_tile(int tz, int tx, int spb, int occ)
{
    if (spb == 2) return pick_tile_spb<MODE, 2>(tz, tx);
    if (occ == 6) return pick_tile_w<MODE, 6>(tz, tx);
    if constexpr (MODE == MODE_REC_WIN || MODE == MODE_BORN_REC) {
        if (occ == 8) return pick_tile_w<MODE, 6>(tz, tx);
    } else {
        if (occ == 8) return pick_tile_w<MODE, 8>(tz, tx);
    }
    return pick_tile_spb<MODE, 1>(tz, tx);
}

template <unsigned MODE>
static StepKernel pick_tile_fs(int tz, int tx)
{
    if (tz == 64 && tx == 16) return k_step_fs<64, 16, MODE>;
    if (tz == 64 && tx == 32) return k_step_fs<64, 32, MODE>;
    if (tz == 32 && tx == 16) return k_step_fs<32, 16, MODE>;
    if (tz == 32 && tx == 32) return k_step_fs<32, 32, MODE>;
    return nullptr;
}

// The nonlinear and Born modes with a free surface.
static StepKernel pick_step_fs(unsigned mode, int tz, int tx)
{
    switch (mode) {
    case MODE_NL_FWD: return pick_tile_fs<MODE_NL_FWD>(tz, tx);
    case MODE_NL_ADJ: return pick_tile_fs<MODE_NL_ADJ>(tz, tx);
    case MODE_SRC: return pick_tile_fs<MODE_SRC>(tz, tx);
    case MODE_BORN_SCAT: return pick_tile_fs<MODE_BORN_SCAT>(tz, tx);
    case MODE_BORN_REC: return pick_tile_fs<MODE_BORN_REC>(tz, tx);
    case MODE_SRC_WIN: return pick_tile_fs<MODE_SRC_WIN>(tz, tx);
    case MODE_BORN_SCAT_PRE: return pick_tile_fs<MODE_BORN_SCAT_PRE>(tz, tx);
    case MODE_REC_WIN: return pick_tile_fs<MODE_REC_WIN>(tz, tx);
    default: return nullptr;
    }
}

// The extended-Born modes (MODE_EXT_SCAT, MODE_EXT_REC) are not instantiated
// here: they run on k_step_generic only, and launch_step refuses them.
static StepKernel pick_step(unsigned mode, int tz, int tx, int spb, int occ)
{
    switch (mode) {
    case MODE_NL_FWD: return pick_tile<MODE_NL_FWD>(tz, tx, spb, occ);
    case MODE_NL_ADJ: return pick_tile<MODE_NL_ADJ>(tz, tx, spb, occ);
    case MODE_SRC: return pick_tile<MODE_SRC>(tz, tx, spb, occ);
    case MODE_BORN_SCAT: return pick_tile<MODE_BORN_SCAT>(tz, tx, spb, occ);
    case MODE_BORN_REC: return pick_tile<MODE_BORN_REC>(tz, tx, spb, occ);
    case MODE_SRC_WIN: return pick_tile<MODE_SRC_WIN>(tz, tx, spb, occ);
    case MODE_BORN_SCAT_PRE: return pick_tile<MODE_BORN_SCAT_PRE>(tz, tx, spb, occ);
    case MODE_REC_WIN: return pick_tile<MODE_REC_WIN>(tz, tx, spb, occ);
    default: return nullptr;
    }
}

// With a start/stop event pair the dispatch itself carries the timestamps
// (hipExtLaunchKernelGGL), so the elapsed time is the kernel's execution.
static hipError_t launch_k(StepKernel k, int grid, const StepArgs& b, hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
    if (e0 && e1) hipExtLaunchKernelGGL(k, dim3(grid), dim3(NT), 0, s, e0, e1, 0, b);
    else hipLaunchKernelGGL(k, dim3(grid), dim3(NT), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_step(unsigned mode, int tz, int tx, int spb, const StepArgs& a, hipStream_t s, hipEvent_t e0,
                       hipEvent_t e1, int occ)
{
    if (occ != 0 && occ != 6 && occ != 8) return hipErrorInvalidValue;
    if (a.fs) spb = 1;                              // the surface instances advance one shot per workgroup
    StepKernel k = a.fs ? pick_step_fs(mode, tz, tx) : pick_step(mode, tz, tx, spb, occ);
    if (!k) return hipErrorInvalidValue;
    if ((mode & M_WIN) && a.winOut == WIN_SLICE) return hipErrorInvalidValue;   // extended-Born capture: generic only
    if ((long)a.nz * a.nx > kMaxSlicePoints) return hipErrorInvalidValue;
    // One workgroup per (tile, group of spb shots).
    StepArgs b = a;
    b.gridLogical = a.ntiles * ((a.nshots + spb - 1) / spb);
    return launch_k(k, (b.gridLogical + 7) / 8 * 8, b, s, e0, e1);
}

using WinKernel = void (*)(const WinStepArgs);

template <unsigned MODE>
static WinKernel pick_win_tile(int tz, int tx, bool loop)
{
    if (tz == 64 && tx == 16) return loop ? k_step_win_loop<64, 16, MODE> : k_step_win<64, 16, MODE>;
    if (tz == 64 && tx == 32) return loop ? k_step_win_loop<64, 32, MODE> : k_step_win<64, 32, MODE>;
    if (tz == 32 && tx == 16) return loop ? k_step_win_loop<32, 16, MODE> : k_step_win<32, 16, MODE>;
    if (tz == 32 && tx == 32) return loop ? k_step_win_loop<32, 32, MODE> : k_step_win<32, 32, MODE>;
    return nullptr;
}

hipError_t launch_step_windows(unsigned mode, int tz, int tx, int grid, const WinStepArgs& w, hipStream_t s,
                               hipEvent_t e0, hipEvent_t e1)
{
    if (grid < 8 || grid % 8 != 0 || w.ntotal < 1) return hipErrorInvalidValue;
    // one workgroup per tile when the grid covers the list, else the persistent loop
    const bool loop = (long)grid < (long)(w.ntotal + 7) / 8 * 8;
    WinKernel k = mode == MODE_NL_FWD ? pick_win_tile<MODE_NL_FWD>(tz, tx, loop) : nullptr;
    if (!k) return hipErrorInvalidValue;
    if (e0 && e1) hipExtLaunchKernelGGL(k, dim3(grid), dim3(NT), 0, s, e0, e1, 0, w);
    else hipLaunchKernelGGL(k, dim3(grid), dim3(NT), 0, s, w);
    return hipGetLastError();
}

template <unsigned MODE, bool FS>
static StepKernel pick_generic_fs(int tz, int tx)
{
    if (tz == 64 && tx == 16) return k_step_generic<64, 16, MODE, FS>;
    if (tz == 64 && tx == 32) return k_step_generic<64, 32, MODE, FS>;
    if (tz == 32 && tx == 16) return k_step_generic<32, 16, MODE, FS>;
    if (tz == 32 && tx == 32) return k_step_generic<32, 32, MODE, FS>;
    return nullptr;
}

template <unsigned MODE>
static StepKernel pick_generic_tile(int tz, int tx, bool fs)
{
    return fs ? pick_generic_fs<MODE, true>(tz, tx) : pick_generic_fs<MODE, false>(tz, tx);
}

hipError_t launch_step_generic(unsigned mode, int tz, int tx, const StepArgs& a, hipStream_t s, hipEvent_t e0,
                               hipEvent_t e1)
{
    StepKernel k = nullptr;
    const bool fs = a.fs != 0;
    switch (mode) {
    case MODE_NL_FWD: k = pick_generic_tile<MODE_NL_FWD>(tz, tx, fs); break;
    case MODE_NL_ADJ: k = pick_generic_tile<MODE_NL_ADJ>(tz, tx, fs); break;
    case MODE_BORN_REC: k = pick_generic_tile<MODE_BORN_REC>(tz, tx, fs); break;
    case MODE_SRC: k = pick_generic_tile<MODE_SRC>(tz, tx, fs); break;
    case MODE_SRC_WIN: k = pick_generic_tile<MODE_SRC_WIN>(tz, tx, fs); break;
    case MODE_BORN_SCAT: k = pick_generic_tile<MODE_BORN_SCAT>(tz, tx, fs); break;
    case MODE_BORN_SCAT_PRE: k = pick_generic_tile<MODE_BORN_SCAT_PRE>(tz, tx, fs); break;
    case MODE_EXT_SCAT: k = pick_generic_tile<MODE_EXT_SCAT>(tz, tx, fs); break;
    case MODE_TOMO_CAP: k = pick_generic_tile<MODE_TOMO_CAP>(tz, tx, fs); break;
    case MODE_TOMO_ADJ_CAP: k = pick_generic_tile<MODE_TOMO_ADJ_CAP>(tz, tx, fs); break;
    case MODE_TOMO_REC: k = pick_generic_tile<MODE_TOMO_REC>(tz, tx, fs); break;
    case MODE_TOMO_CAP_WIN: k = pick_generic_tile<MODE_TOMO_CAP_WIN>(tz, tx, fs); break;
    case MODE_TOMO_ADJ_CAP_WIN: k = pick_generic_tile<MODE_TOMO_ADJ_CAP_WIN>(tz, tx, fs); break;
    case MODE_EXT_REC: k = pick_generic_tile<MODE_EXT_REC>(tz, tx, fs); break;
    case MODE_REC_WIN: k = pick_generic_tile<MODE_REC_WIN>(tz, tx, fs); break;
    case MODE_BORN_SCAT_CAP: k = pick_generic_tile<MODE_BORN_SCAT_CAP>(tz, tx, fs); break;
    default: break;
    }
    if (!k) return hipErrorInvalidValue;
    if ((long)a.nz * a.nx > kMaxSlicePoints) return hipErrorInvalidValue;
    StepArgs b = a;
    b.gridLogical = a.ntiles * a.nshots;               // one workgroup per (tile, shot)
    return launch_k(k, (b.gridLogical + 7) / 8 * 8, b, s, e0, e1);
}

hipError_t launch_fwd_pair(int tz, int tx, const StepArgs& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
    StepKernel k = nullptr;
    if (tz == 64 && tx == 16) k = a.fs ? k_fwd_pair_fs<64, 16> : k_fwd_pair<64, 16>;
    else if (tz == 64 && tx == 32) k = a.fs ? k_fwd_pair_fs<64, 32> : k_fwd_pair<64, 32>;
    else if (tz == 32 && tx == 16) k = a.fs ? k_fwd_pair_fs<32, 16> : k_fwd_pair<32, 16>;
    else if (tz == 32 && tx == 32) k = a.fs ? k_fwd_pair_fs<32, 32> : k_fwd_pair<32, 32>;
    if (!k) return hipErrorInvalidValue;
    if ((long)a.nz * a.nx > kMaxSlicePoints) return hipErrorInvalidValue;
    return launch_k(k, (a.gridLogical + 7) / 8 * 8, a, s, e0, e1);
}

hipError_t launch_fwd_pair2(const Pair2Args& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
    const int grid = (a.gridLogical + 7) / 8 * 8;
    auto k = k_fwd_pair2<kPair2EZ, kPair2EX>;
    if (e0 && e1) hipExtLaunchKernelGGL(k, dim3(grid), dim3(NT), 0, s, e0, e1, 0, a);
    else hipLaunchKernelGGL(k, dim3(grid), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_rec2(const Rec2Args& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
    const int grid = (a.gridLogical + 7) / 8 * 8;
    auto k = k_rec2<kPair2EZ, kPair2EX>;
    if (e0 && e1) hipExtLaunchKernelGGL(k, dim3(grid), dim3(NT), 0, s, e0, e1, 0, a);
    else hipLaunchKernelGGL(k, dim3(grid), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_grad(int tz, int tx, const GradArgs& a, hipStream_t s)
{
    using GradKernel = void (*)(const GradArgs);
    GradKernel g = nullptr;
    if (tz == 64 && tx == 16) g = k_grad<64, 16>;
    else if (tz == 64 && tx == 32) g = k_grad<64, 32>;
    else if (tz == 32 && tx == 16) g = k_grad<32, 16>;
    else if (tz == 32 && tx == 32) g = k_grad<32, 32>;
    if (!g) return hipErrorInvalidValue;
    hipLaunchKernelGGL(g, dim3((a.gridLogical + 7) / 8 * 8), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ring(unsigned mode, const StepArgs& a, int maxRing, hipStream_t s)
{
    if (maxRing <= 0) return hipSuccess;
    dim3 grid((maxRing + 127) / 128, a.nshots);
    const unsigned sparse = mode & (M_INJ_SIG | M_INJ_DATA | M_REC_DATA | M_REC_SIG);
    switch (sparse) {
    case M_INJ_SIG | M_REC_DATA: hipLaunchKernelGGL(k_ring<M_INJ_SIG | M_REC_DATA>, grid, dim3(128), 0, s, a); break;
    case M_INJ_DATA | M_REC_SIG: hipLaunchKernelGGL(k_ring<M_INJ_DATA | M_REC_SIG>, grid, dim3(128), 0, s, a); break;
    case M_INJ_SIG: hipLaunchKernelGGL(k_ring<M_INJ_SIG>, grid, dim3(128), 0, s, a); break;
    case M_REC_DATA: hipLaunchKernelGGL(k_ring<M_REC_DATA>, grid, dim3(128), 0, s, a); break;
    case M_INJ_DATA: hipLaunchKernelGGL(k_ring<M_INJ_DATA>, grid, dim3(128), 0, s, a); break;
    default: return hipSuccess;
    }
    return hipGetLastError();
}

static dim3 flat_grid(long n)
{
    long b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

hipError_t launch_fill(float* p, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, flat_grid(n), dim3(256), 0, s, p, n, 0.f);
    return hipGetLastError();
}

hipError_t launch_scale_model(const float* m, float* out, const float* v2dt2, const float* refl,
                              int nz, int nx, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_model, flat_grid((long)nz * nx), dim3(256), 0, s, m, out, v2dt2, refl, nz, nx);
    return hipGetLastError();
}

hipError_t launch_scale_ext(const float* m, float* out, const float* refl, int nz, int nx, int nExt,
                            hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_ext, flat_grid((long)nz * nx * nExt), dim3(256), 0, s, m, out, refl, nz, nx, nExt);
    return hipGetLastError();
}

hipError_t launch_born_adj_finish(const ShotSlot* slots, int nshots, int ssR, const float* v2dt2,
                                  const float* refl, int nz, int nx, hipStream_t s)
{
    dim3 g = flat_grid((long)nz * nx);
    g.y = nshots;
    hipLaunchKernelGGL(k_born_adj_finish, g, dim3(256), 0, s, slots, ssR, v2dt2, refl, nz, nx);
    return hipGetLastError();
}

// Register-tiled variants by the bound HX >= hExt; larger offsets take the
// untiled kernels.
// Blocks of a grid mapped by xcd_block3: the count rounded up to a multiple of 8.
static unsigned xcd_padded(int blocks) { return (unsigned)((blocks + 7) / 8 * 8); }

hipError_t launch_ext_ss_batch(const ExtBatchArgs& a, int nshots, hipStream_t s)
{
    if (a.nk < 1 || a.nk > kExtBatchMax) return hipErrorInvalidValue;
    const int nzi = a.nz - 2 * FAT, nxi = a.nx - 2 * FAT;
    if (nzi <= 0 || nxi <= 0) return hipSuccess;
    if (a.hExt <= 20) {
        const int bx = (nxi + kExtTX - 1) / kExtTX, by = (nshots + kExtSsSpb - 1) / kExtSsSpb;
        const int bz = (nzi + kExtTZ - 1) / kExtTZ;
        const dim3 g(xcd_padded(bx * by * bz)), b(kExtTZ, 4);
        if (a.hExt <= 2) hipLaunchKernelGGL(k_ext_ss_tile<2>, g, b, 0, s, a, nshots, bx, by, bz);
        else if (a.hExt <= 5) hipLaunchKernelGGL(k_ext_ss_tile<5>, g, b, 0, s, a, nshots, bx, by, bz);
        else if (a.hExt <= 10) hipLaunchKernelGGL(k_ext_ss_tile<10>, g, b, 0, s, a, nshots, bx, by, bz);
        else hipLaunchKernelGGL(k_ext_ss_tile<20>, g, b, 0, s, a, nshots, bx, by, bz);
    } else {
        dim3 g((nzi + 63) / 64, nshots, nxi);
        hipLaunchKernelGGL(k_ext_ss_any, g, dim3(64), 0, s, a);
    }
    return hipGetLastError();
}

template <bool WF>
static void ext_img_launch(const ExtBatchArgs& a, int nshots, hipStream_t s, int nzi, int nxr)
{
    if (a.hExt <= 20) {
        // offsets per workgroup: all of them up to 11, else groups of 7 (21 offsets) or 14
        const int ng = a.hExt <= 5 ? 2 * a.hExt + 1 : a.hExt <= 10 ? 7 : 14;
        const int ngrp = (2 * a.hExt + 1 + ng - 1) / ng;
        const int bx = (nxr + kExtTX - 1) / kExtTX, by = nshots, bz = (nzi + kExtTZ - 1) / kExtTZ;
        const dim3 g(xcd_padded(bx * ngrp * by * bz)), b(kExtTZ, 4);
        if (a.hExt <= 2) hipLaunchKernelGGL((k_ext_img_tile<2, 5, WF>), g, b, 0, s, a, bx, 1, by, bz);
        else if (a.hExt <= 5) hipLaunchKernelGGL((k_ext_img_tile<5, 11, WF>), g, b, 0, s, a, bx, 1, by, bz);
        else if (a.hExt <= 10) hipLaunchKernelGGL((k_ext_img_tile<10, 7, WF>), g, b, 0, s, a, bx, ngrp, by, bz);
        else hipLaunchKernelGGL((k_ext_img_tile<20, 14, WF>), g, b, 0, s, a, bx, ngrp, by, bz);
    } else {
        dim3 g((nzi + 63) / 64, nshots, nxr);
        hipLaunchKernelGGL(k_ext_img_any<WF>, g, dim3(64), 0, s, a);
    }
}

hipError_t launch_ext_img_batch(const ExtBatchArgs& a, int nshots, hipStream_t s, bool wemvaFwd)
{
    if (a.nk < 1 || a.nk > kExtBatchMax) return hipErrorInvalidValue;
    const int nzi = a.nz - 2 * FAT, nxr = a.nx - 2 * FAT - 2 * a.hExt;
    if (nzi <= 0 || nxr <= 0) return hipSuccess;
    if (wemvaFwd) ext_img_launch<true>(a, nshots, s, nzi, nxr);
    else ext_img_launch<false>(a, nshots, s, nzi, nxr);
    return hipGetLastError();
}

static dim3 column_grid(int nz, int nx, int nshots) { return dim3((nz - 2 * FAT + 63) / 64, nshots, nx - 2 * FAT); }

hipError_t launch_dt2_coarse(const ShotSlot* slots, int nshots, int nts, float cSide, float cCenter, int nz, int nx,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_dt2_coarse, column_grid(nz, nx, nshots), dim3(64), 0, s, slots, nts, cSide, cCenter, nz, nx);
    return hipGetLastError();
}

hipError_t launch_mul_batch(const ExtBatchArgs& a, int nshots, hipStream_t s)
{
    if (a.nk < 1 || a.nk > kExtBatchMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mul_batch, column_grid(a.nz, a.nx, nshots), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_tomo_adj_ss(const ExtBatchArgs& a, int nshots, hipStream_t s)
{
    if (a.nk < 1 || a.nk > kExtBatchMax) return hipErrorInvalidValue;
    const int nzi = a.nz - 2 * FAT, nxi = a.nx - 2 * FAT;
    if (nzi <= 0 || nxi <= 0) return hipSuccess;
    if (a.hExt <= 10) {
        const int bx = (nxi + kExtTX - 1) / kExtTX, by = (nshots + kExtSsSpb - 1) / kExtSsSpb;
        const int bz = (nzi + kExtTZ - 1) / kExtTZ;
        const dim3 g(xcd_padded(bx * by * bz)), b(kExtTZ, 4);
        if (a.hExt <= 2) hipLaunchKernelGGL(k_tomo_adj_ss_tile<2>, g, b, 0, s, a, nshots, bx, by, bz);
        else if (a.hExt <= 5) hipLaunchKernelGGL(k_tomo_adj_ss_tile<5>, g, b, 0, s, a, nshots, bx, by, bz);
        else hipLaunchKernelGGL(k_tomo_adj_ss_tile<10>, g, b, 0, s, a, nshots, bx, by, bz);
    } else {
        hipLaunchKernelGGL(k_tomo_adj_ss, column_grid(a.nz, a.nx, nshots), dim3(64), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_img_acc_batch(const ImgAccArgs& a, int nshots, hipStream_t s)
{
    if (a.nk < 1 || a.nk > kExtBatchMax) return hipErrorInvalidValue;
    if ((long)a.nz * a.nx > kMaxSlicePoints) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_img_acc_batch, column_grid(a.nz, a.nx, nshots), dim3(64), 0, s, a);
    return hipGetLastError();
}

// The lag ranges must lie inside [0, 2 hExt] and reach W samples inside the
// stored [0, nts) only (W is read at k - lagDir * 2 (e - hExt), extreme at
// e = lo and e = hi - 1).
static bool time_batch_ok(const ExtBatchArgs& a)
{
    if (a.nk < 1 || a.nk > kExtBatchMax || a.hExt < 0 || (a.lagDir != 1 && a.lagDir != -1)) return false;
    if (a.eLo < 0 || a.eHi > 2 * a.hExt + 1 || a.eLo > a.eHi) return false;
    const long nzx = (long)a.nz * a.nx;
    for (int j = 0; j < a.nk; ++j) {
        if (a.lo[j] < a.eLo || a.hi[j] > a.eHi || a.lo[j] > a.hi[j]) return false;
        if (a.wOff[j] % nzx != 0) return false;
        const long k = a.wOff[j] / nzx;
        if (k < 0 || k >= a.nts) return false;
        if (a.lo[j] < a.hi[j]) {
            const long s0 = k - 2L * a.lagDir * (a.lo[j] - a.hExt), s1 = k - 2L * a.lagDir * (a.hi[j] - 1 - a.hExt);
            if (std::min(s0, s1) < 0 || std::max(s0, s1) >= a.nts) return false;
        }
    }
    return true;
}

// A batch the register-window kernels take: kExtBatchMax samples k0 + dirK * j
// with every lag usable (checked by time_batch_ok to stay inside [0, nts)).
static bool time_batch_full(const ExtBatchArgs& a, int dirK)
{
    if (a.nk != kExtBatchMax || a.hExt > 20 || a.eLo != 0 || a.eHi != 2 * a.hExt + 1) return false;
    const long nzx = (long)a.nz * a.nx;
    for (int j = 0; j < a.nk; ++j)
        if (a.lo[j] != 0 || a.hi[j] != 2 * a.hExt + 1 || a.wOff[j] != a.wOff[0] + dirK * j * nzx) return false;
    return true;
}

template <int HX>
static void time_ss_win(const ExtBatchArgs& a, dim3 g, hipStream_t s)
{
    if (a.lagDir > 0) hipLaunchKernelGGL((k_time_ss_win<HX, kExtBatchMax, 1>), g, dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_time_ss_win<HX, kExtBatchMax, -1>), g, dim3(64), 0, s, a);
}

template <int HX, int DIRK>
static void time_img_win(const ExtBatchArgs& a, dim3 g, hipStream_t s)
{
    if (a.lagDir > 0) hipLaunchKernelGGL((k_time_img_win<HX, kExtBatchMax, DIRK, 1>), g, dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_time_img_win<HX, kExtBatchMax, DIRK, -1>), g, dim3(64), 0, s, a);
}

hipError_t launch_time_ss_batch(const ExtBatchArgs& a, int nshots, hipStream_t s)
{
    if (!time_batch_ok(a)) return hipErrorInvalidValue;
    const dim3 g = column_grid(a.nz, a.nx, nshots);
    if (!time_batch_full(a, 1)) hipLaunchKernelGGL(k_time_ss, g, dim3(64), 0, s, a);
    else if (a.hExt <= 5) time_ss_win<5>(a, g, s);
    else if (a.hExt <= 10) time_ss_win<10>(a, g, s);
    else time_ss_win<20>(a, g, s);
    return hipGetLastError();
}

hipError_t launch_time_img_batch(const ExtBatchArgs& a, int nshots, hipStream_t s)
{
    if (!time_batch_ok(a)) return hipErrorInvalidValue;
    const dim3 g = column_grid(a.nz, a.nx, nshots);
    const int dirK = time_batch_full(a, 1) ? 1 : time_batch_full(a, -1) ? -1 : 0;
    if (dirK == 0) hipLaunchKernelGGL(k_time_img, g, dim3(64), 0, s, a);
    else if (a.hExt <= 5) dirK > 0 ? time_img_win<5, 1>(a, g, s) : time_img_win<5, -1>(a, g, s);
    else if (a.hExt <= 10) dirK > 0 ? time_img_win<10, 1>(a, g, s) : time_img_win<10, -1>(a, g, s);
    else dirK > 0 ? time_img_win<20, 1>(a, g, s) : time_img_win<20, -1>(a, g, s);
    return hipGetLastError();
}

hipError_t launch_time_scale(const float* v2dt2, const float* refl, int sloth, float* out, int nz, int nx,
                             hipStream_t s)
{
    const long n = (long)nz * nx;
    hipLaunchKernelGGL(k_time_scale, flat_grid(n), dim3(256), 0, s, v2dt2, refl, sloth, out, n);
    return hipGetLastError();
}

hipError_t launch_ext_adj_finish(const ShotSlot* slots, int nshots, const float* refl, int nz, int nx,
                                 int nExt, hipStream_t s)
{
    dim3 g = flat_grid((long)nz * nx * nExt);
    g.y = nshots;
    hipLaunchKernelGGL(k_ext_adj_finish, g, dim3(256), 0, s, slots, refl, nz, nx, nExt);
    return hipGetLastError();
}

hipError_t launch_record_scale(const ShotSlot* slot, int ssIdx, float* out, const float* refl, const float* v2dt2,
                               int nz, int nx, hipStream_t s)
{
    hipLaunchKernelGGL(k_record_scale, flat_grid((long)nz * nx), dim3(256), 0, s, slot, ssIdx, out, refl, v2dt2, nz, nx);
    return hipGetLastError();
}

hipError_t launch_accumulate(float* acc, const float* img, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_accumulate, flat_grid(n), dim3(256), 0, s, acc, img, n);
    return hipGetLastError();
}

hipError_t launch_add_to(float* out, const float* acc, long n, int add, hipStream_t s)
{
    hipLaunchKernelGGL(k_add_to, flat_grid(n), dim3(256), 0, s, out, acc, n, add);
    return hipGetLastError();
}

hipError_t launch_dev_forward(const float* reg, float* out, const int* cornerReg, const float* weight,
                              int nDev, int nb, int nt, int add, hipStream_t s)
{
    hipLaunchKernelGGL(k_dev_forward, flat_grid((long)nDev * nt), dim3(256), 0, s, reg, out, cornerReg,
                       weight, nDev, nb, nt, add);
    return hipGetLastError();
}

hipError_t launch_win_tail(const ShotSlot* slots, int nshots, int cur, int winOut, float wa, float wb,
                           long wA, long wImg, int ssA, const float* mtil, int nz, int nx, hipStream_t s)
{
    dim3 g = flat_grid((long)nz * nx);
    g.y = nshots;
    hipLaunchKernelGGL(k_win_tail, g, dim3(256), 0, s, slots, cur, winOut, wa, wb, wA, wImg, ssA, mtil, nz, nx);
    return hipGetLastError();
}

hipError_t launch_dev_adjoint(float* reg, const float* irreg, const int* adjPtr, const int* adjEnt,
                              const float* weight, int nReg, int nb, int nt, hipStream_t s)
{
    hipLaunchKernelGGL(k_dev_adjoint, flat_grid((long)nReg * nt), dim3(256), 0, s, reg, irreg, adjPtr,
                       adjEnt, weight, nReg, nb, nt);
    return hipGetLastError();
}

}  // namespace aiso
