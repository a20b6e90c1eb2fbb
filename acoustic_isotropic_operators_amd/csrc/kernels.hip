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
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace aiso {

constexpr int FAT = 5;
constexpr int NT = 256;

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// Linear workgroup ids go round-robin over the 8 XCDs; give each XCD a
// contiguous run of (shot, tile) work so neighbouring tiles share its L2.
__device__ __forceinline__ int xcd_logical(int bid, int gridPadded)
{
    const int per = gridPadded >> 3;
    return (bid & 7) * per + (bid >> 3);
}

__device__ __forceinline__ float sponge(const float* __restrict__ cd, int gz, int gx,
                                        int nz, int nx, int minPad)
{
    int dist = imin(imin(gz - FAT, gx - FAT), imin(nz - gz - 1 - FAT, nx - gx - 1 - FAT));
    return dist < minPad ? cd[dist] : 1.0f;
}

template <int TZ, int TX, unsigned MODE>
__global__ __launch_bounds__(NT) void k_step(const StepArgs a)
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
    const float* __restrict__ pc = S.p[a.cur];
    float* __restrict__ po = S.p[a.cur ^ 1];
    const float* __restrict__ v2dt2 = a.v2dt2;

    // Stage the current slice (times v2dt2 for the adjoint) with its halo.
    for (int lx = wy; lx < LX; lx += NXW) {
        const int gx = x0 - FAT + lx;
        for (int lz = lane; lz < LZ; lz += TZ) {
            const int gz = z0 - FAT + lz;
            float v = 0.f;
            if (gx < nx && gz < nz) {
                const long gi = (long)gx * nz + gz;
                v = pc[gi];
                if constexpr (ADJ) v = v * v2dt2[gi];
            }
            sC[lx * LZ + lz] = v;
        }
    }
    __syncthreads();

    const int gz = z0 + lane;
    const int lx0 = wy * PX;
    const bool zok = gz < nz - FAT;
    float w[PX + 2 * FAT];
#pragma unroll
    for (int j = 0; j < PX + 2 * FAT; ++j) w[j] = sC[(lx0 + j) * LZ + lane + FAT];

    float nv[PX], dmp[PX];
    float ost[(MODE & M_WIN) ? PX : 1], cur[(MODE & M_WIN) ? PX : 1];
    bool ok[PX];
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        const int gx = x0 + lx0 + j;
        ok[j] = zok && gx < nx - FAT;
        const float* s = &sC[(lx0 + j + FAT) * LZ + lane + FAT];
        const float cc = w[j + FAT];
        const float lap = a.zc[0] * cc
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
        float n = 0.f, d = 1.f;
        if (ok[j]) {
            const long gi = (long)gx * nz + gz;
            d = sponge(a.cd, gz, gx, nz, nx, a.minPad);
            const float os = po[gi];
            const float o = os * d;                  // second damping of the old slice
            if constexpr (ADJ) {
                const float c = pc[gi];
                n = lap + c + c - o;
                if constexpr (MODE & M_WIN) cur[j] = c;
            } else {
                n = v2dt2[gi] * lap + cc + cc - o;
                if constexpr (MODE & M_WIN) cur[j] = cc;
            }
            if constexpr (MODE & M_WIN) ost[j] = os;
            if constexpr (MODE & M_SS_BORN) {
                const float m = a.mtil[gi];
                const float sa = m * S.W[a.wA + gi];
                if (a.fb != 0.f) {
                    const float sb = m * S.W[a.wB + gi];
                    n += sa * a.fa + sb * a.fb;
                } else {
                    n += sa * a.fa;
                }
            }
            if constexpr (MODE & M_SS_PRE) {
                const float sa = S.W[a.wA + gi];
                if (a.fb != 0.f) n += sa * a.fa + S.W[a.wB + gi] * a.fb;
                else n += sa * a.fa;
            }
            if constexpr (MODE & M_SS_SLICE) {
                const float sa = S.ss[a.ssA][gi];
                if (a.fb != 0.f) n += sa * a.fa + S.ss[a.ssB][gi] * a.fb;
                else n += sa * a.fa;
            }
        }
        nv[j] = n;
        dmp[j] = d;
    }

    const int t = tile;
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
                    v = S.sig[(long)en.y * a.ntw + a.itwInj];
                } else {
                    const float* dr = S.dreg + (long)en.y * a.nts + a.its;
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
        const int gx = x0 + lx0 + j;
        const long gi = (long)gx * nz + gz;
        const float n = nv[j] * dmp[j];
        nv[j] = n;
        if constexpr (MODE & M_CAPTURE) {
            // Forward passes capture the coarse source wavefield W; the
            // adjoint receiver pass accumulates into the two ss slices.
            float* A = ADJ ? S.ss[a.ssA] : S.W + a.wA;
            float* B = ADJ ? S.ss[a.ssB] : S.W + a.wB;
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
                if (a.image) S.img[gi] += S.W[a.wImg + gi] * bnew;
            }
        }
        if constexpr (MODE & M_WIN) {
            if (a.win) {
                // Coarse sample from three consecutive fine slices, accumulated
                // in the reference's order (old, current, new).
                const float v = (ost[j] * a.wa + cur[j] * a.wb) + n * a.wc;
                switch (a.winOut) {
                case WIN_W: S.W[a.wA + gi] = v; break;
                case WIN_SS: S.W[a.wA + gi] = a.mtil[gi] * v; break;
                case WIN_IMG: S.img[gi] += S.W[a.wImg + gi] * v; break;
                default: S.ss[a.ssA][gi] = v; break;
                }
            }
        }
        po[gi] = n;
    }

    if constexpr (POST_SPARSE) {
        const int* tp = (MODE & M_REC_SIG) ? S.srcTile : S.recTile;
        const int2* te = (MODE & M_REC_SIG) ? S.srcEnt : S.recEnt;
        const int e0 = tp[t], e1 = tp[t + 1];
        if (e1 > e0) {                               // block-uniform
#pragma unroll
            for (int j = 0; j < PX; ++j) sN[(lx0 + j) * TZ + lane] = nv[j];
            __syncthreads();
            for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
                const int2 en = te[e];
                const float n = sN[en.x];
                if constexpr (MODE & M_REC_SIG) {
                    S.sig[(long)en.y * a.ntw + a.itwRec] += n;
                } else {
                    float* dr = S.dreg + (long)en.y * a.nts + a.its;
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

// imagingOffsetFwdGpu + scaleSecondarySourceFd (F/src/kernelsGpu.cu:381-396,
// 270-275): ss[x] = v2dt2 * sum_h m~[h][x-h] * W[k][x-2h], written (not added)
// into the slot's ss[dst].
__global__ void k_ext_ss(const ShotSlot* slots, int dst, long wk, const float* __restrict__ mext,
                         const float* __restrict__ v2dt2, int nz, int nx, int hExt)
{
    const ShotSlot& S = slots[blockIdx.y];
    const long nzx = (long)nz * nx;
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + blockIdx.z;
    if (iz >= nz - FAT) return;
    const int lo = max(-hExt, ix + 1 + FAT + hExt - nx);
    const int hi = min(ix - hExt - FAT, hExt);
    const float* Wk = S.W + wk;
    float acc = 0.f;
    for (int e = lo; e <= hi; ++e)
        acc += mext[nzx * (e + hExt) + (long)nz * (ix - e) + iz] * Wk[(long)(ix - 2 * e) * nz + iz];
    const long i = (long)ix * nz + iz;
    S.ss[dst][i] = acc * v2dt2[i];
}

// scaleSecondarySourceFd + imagingOffsetAdjGpu (F/src/kernelsGpu.cu:270-275,
// 398-411): img[e][x] += (ss[x+h]*v2dt2[x+h]) * W[k][x-h], h = e - hExt,
// for x in [FAT+hExt, nx-FAT-hExt).
__global__ void k_ext_img(const ShotSlot* slots, int src, long wk, const float* __restrict__ v2dt2,
                          int nz, int nx, int nExt)
{
    const ShotSlot& S = slots[blockIdx.y];
    const int hExt = (nExt - 1) / 2;
    const long nzx = (long)nz * nx;
    const int iz = FAT + blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = FAT + hExt + blockIdx.z;
    if (iz >= nz - FAT || ix >= nx - FAT - hExt) return;
    const float* ss = S.ss[src];
    const float* Wk = S.W + wk;
    for (int e = 0; e < nExt; ++e) {
        const int h = e - hExt;
        const long ir = (long)(ix + h) * nz + iz;
        const float r = ss[ir] * v2dt2[ir];
        float* m = S.img + nzx * e + (long)ix * nz + iz;
        *m += r * Wk[(long)(ix - h) * nz + iz];
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

// acc += img (one shot, in shot order).
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
        default: S.ss[ssA][i] = v; break;
        }
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers (declared in kernels.hpp).

using StepKernel = void (*)(const StepArgs);

template <unsigned MODE>
static StepKernel pick_tile(int tz, int tx)
{
    if (tz == 64) {
        if (tx == 8) return k_step<64, 8, MODE>;
        if (tx == 16) return k_step<64, 16, MODE>;
        if (tx == 32) return k_step<64, 32, MODE>;
    } else if (tz == 32) {
        if (tx == 8) return k_step<32, 8, MODE>;
        if (tx == 16) return k_step<32, 16, MODE>;
        if (tx == 32) return k_step<32, 32, MODE>;
    }
    return nullptr;
}

static StepKernel pick_step(unsigned mode, int tz, int tx)
{
    switch (mode) {
    case MODE_NL_FWD: return pick_tile<MODE_NL_FWD>(tz, tx);
    case MODE_NL_ADJ: return pick_tile<MODE_NL_ADJ>(tz, tx);
    case MODE_SRC: return pick_tile<MODE_SRC>(tz, tx);
    case MODE_BORN_SCAT: return pick_tile<MODE_BORN_SCAT>(tz, tx);
    case MODE_BORN_REC: return pick_tile<MODE_BORN_REC>(tz, tx);
    case MODE_EXT_SCAT: return pick_tile<MODE_EXT_SCAT>(tz, tx);
    case MODE_EXT_REC: return pick_tile<MODE_EXT_REC>(tz, tx);
    case MODE_SRC_WIN: return pick_tile<MODE_SRC_WIN>(tz, tx);
    case MODE_BORN_SCAT_PRE: return pick_tile<MODE_BORN_SCAT_PRE>(tz, tx);
    case MODE_REC_WIN: return pick_tile<MODE_REC_WIN>(tz, tx);
    default: return nullptr;
    }
}

hipError_t launch_step(unsigned mode, int tz, int tx, const StepArgs& a, hipStream_t s)
{
    StepKernel k = pick_step(mode, tz, tx);
    if (!k) return hipErrorInvalidValue;
    const int grid = (a.gridLogical + 7) / 8 * 8;
    hipLaunchKernelGGL(k, dim3(grid), dim3(NT), 0, s, a);
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

hipError_t launch_ext_ss(const ShotSlot* slots, int nshots, int dst, long wk, const float* mext,
                         const float* v2dt2, int nz, int nx, int hExt, hipStream_t s)
{
    dim3 g((nz - 2 * FAT + 63) / 64, nshots, nx - 2 * FAT);
    hipLaunchKernelGGL(k_ext_ss, g, dim3(64), 0, s, slots, dst, wk, mext, v2dt2, nz, nx, hExt);
    return hipGetLastError();
}

hipError_t launch_ext_img(const ShotSlot* slots, int nshots, int src, long wk, const float* v2dt2,
                          int nz, int nx, int nExt, hipStream_t s)
{
    const int hExt = (nExt - 1) / 2;
    const int nxr = nx - 2 * FAT - 2 * hExt;
    if (nxr <= 0) return hipSuccess;
    dim3 g((nz - 2 * FAT + 63) / 64, nshots, nxr);
    hipLaunchKernelGGL(k_ext_img, g, dim3(64), 0, s, slots, src, wk, v2dt2, nz, nx, nExt);
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
