/*
 * aiso_oracle.c -- CPU oracle, float (_f) and double (_d) instantiations plus
 * the integer geometry.  TEST INFRASTRUCTURE ONLY (see aiso_oracle.h).
 * Built by oracle/Makefile into oracle/liborc.so.
 */
#define _USE_MATH_DEFINES
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "aiso_oracle.h"

#define REAL float
#define SFX _f
#include "aiso_oracle_body.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX _d
#include "aiso_oracle_body.h"
#undef REAL
#undef SFX

/* Corner layout shared by deviceGpu constructors #2 and #3
 * (F/src/deviceGpu.cpp:100-137 and :155-197): the device point gets weight
 * 1, its three bilinear neighbours weight 0 (they are still registered). */
static void grid_corners(int i1, int base, int nzVel, int dipole, int dipBase,
                         int* gridIdx, float* weight)
{
    gridIdx[i1] = base;              weight[i1] = 1.0f;
    gridIdx[i1 + 1] = base + 1;      weight[i1 + 1] = 0.0f;
    gridIdx[i1 + 2] = base + nzVel;  weight[i1 + 2] = 0.0f;
    gridIdx[i1 + 3] = base + nzVel + 1; weight[i1 + 3] = 0.0f;
    if (dipole == 1) {
        gridIdx[i1 + 4] = dipBase;              weight[i1 + 4] = -1.0f;
        gridIdx[i1 + 5] = dipBase + 1;          weight[i1 + 5] = 0.0f;
        gridIdx[i1 + 6] = dipBase + nzVel;      weight[i1 + 6] = 0.0f;
        gridIdx[i1 + 7] = dipBase + nzVel + 1;  weight[i1 + 7] = 0.0f;
    }
}

/* deviceGpu constructor #3, F/src/deviceGpu.cpp:142-199 (x outer, z inner);
 * bounds check :286-294.  Returns the device count or -1 when out of bounds. */
int orc_dev_grid(int nzDev, int ozDev, int dzDev, int nxDev, int oxDev, int dxDev,
                 int nzVel, int nxVel, int dipole, int zShift, int xShift,
                 int* gridIdx, float* weight)
{
    int zMax = ozDev + (nzDev - 1) * dzDev, xMax = oxDev + (nxDev - 1) * dxDev;
    if (zMax >= nzVel || xMax >= nxVel) return -1;
    const int nb = dipole == 1 ? 8 : 4;
    int iDev = -1;
    for (int ix = 0; ix < nxDev; ++ix) {
        int xd = oxDev + ix * dxDev;
        for (int iz = 0; iz < nzDev; ++iz) {
            int zd = ozDev + iz * dzDev;
            ++iDev;
            grid_corners(iDev * nb, xd * nzVel + zd, nzVel, dipole,
                         (xd + xShift) * nzVel + zd + zShift, gridIdx, weight);
        }
    }
    return iDev + 1;
}

/* deviceGpu constructor #2, F/src/deviceGpu.cpp:87-139; bounds :276-284. */
int orc_dev_ints(int nDev, const int* zGrid, const int* xGrid, int nzVel, int nxVel,
                 int dipole, int zShift, int xShift, int* gridIdx, float* weight)
{
    int zm = zGrid[0], xm = xGrid[0];
    for (int i = 1; i < nDev; ++i) {
        if (zGrid[i] > zm) zm = zGrid[i];
        if (xGrid[i] > xm) xm = xGrid[i];
    }
    if (zm >= nzVel || xm >= nxVel) return -1;
    const int nb = dipole == 1 ? 8 : 4;
    for (int i = 0; i < nDev; ++i)
        grid_corners(i * nb, xGrid[i] * nzVel + zGrid[i], nzVel, dipole,
                     (xGrid[i] + xShift) * nzVel + zGrid[i] + zShift, gridIdx, weight);
    return nDev;
}

/* deviceGpu::convertIrregToReg, F/src/deviceGpu.cpp:201-224: unique grid
 * points in first-seen order; cornerReg maps every (device, corner) entry to
 * its unique index.  Returns the unique count.  Quadratic search: the oracle
 * only sees test-sized geometries. */
int orc_unique(int nIdx, const int* gridIdx, int* regPos, int* cornerReg)
{
    int n = 0;
    for (int k = 0; k < nIdx; ++k) {
        int found = -1;
        for (int j = 0; j < n; ++j)
            if (regPos[j] == gridIdx[k]) { found = j; break; }
        if (found < 0) { regPos[n] = gridIdx[k]; found = n++; }
        cornerReg[k] = found;
    }
    return n;
}

/* Padding rule of F/src/padFileGpuMain.cpp:49-66 (no free surface):
 * padPlus = ceil((2*pad+n)/blockSize)*blockSize - n - pad, ceil in float. */
int orc_pad_plus(int n, int pad, int blockSize)
{
    long total = (long)pad * 2 + n;
    float ratio = (float)total / (float)blockSize;
    ratio = ceilf(ratio);
    long nb = (long)ratio;
    return (int)(nb * blockSize - n - pad);
}
