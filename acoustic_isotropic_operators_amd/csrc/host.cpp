// host.cpp -- FD parameters, tables, geometry and the per-shot signal chain.
// Arithmetic (including the float/double rounding points) follows the
// reference's host code; each function cites what it restates.
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace aiso {

static void fail(aiso_status c, const std::string& m) { throw Error(c, "**** ERROR: " + m + " ****"); }

FdParams make_params(const aiso_fd_params& a)
{
    FdParams p;
    p.nz = a.nz; p.nx = a.nx; p.dz = a.dz; p.dx = a.dx; p.oz = a.oz; p.ox = a.ox;
    p.nts = a.nts; p.dts = a.dts; p.ots = a.ots; p.sub = a.sub;
    p.zPadMinus = a.zPadMinus; p.zPadPlus = a.zPadPlus;
    p.xPadMinus = a.xPadMinus; p.xPadPlus = a.xPadPlus;
    p.alphaCos = a.alphaCos; p.fMax = a.fMax;
    p.nExt = a.nExt; p.extension = a.extension;

    if (a.nts < 2) fail(AISO_ERR_PARAM, "nts must be at least 2");
    if (a.sub < 1) fail(AISO_ERR_PARAM, "sub must be at least 1");
    if (a.dts <= 0.f || a.dz <= 0.f || a.dx <= 0.f) fail(AISO_ERR_PARAM, "dts, dz and dx must be positive");
    if (a.fat != kFat) fail(AISO_ERR_PARAM, "fat must be 5 for the 10th-order stencil");
    // fdParam.cpp:29-45: a free surface sits on the first interior row and
    // comes with no top padding; an absorbing top needs some.
    p.freeSurface = a.freeSurface == 1;
    if ((a.zPadMinus == 0 && !p.freeSurface) || (a.zPadMinus > 0 && p.freeSurface))
        fail(AISO_ERR_PARAM, "[fdParam]: zPadMinus and freeSurface flag are inconsistent");
    if (a.zPadMinus < 0 || a.zPadPlus < 0 || a.xPadMinus < 0 || a.xPadPlus < 0)
        fail(AISO_ERR_PARAM, "padding must be non-negative");
    // fdParam.cpp:69-70
    if (a.nExt < 1 || a.nExt % 2 == 0) fail(AISO_ERR_PARAM, "Length of extended axis must be an uneven number");
    p.hExt = (a.nExt - 1) / 2;
    if (a.extension != AISO_EXT_NONE && a.extension != AISO_EXT_OFFSET && a.extension != AISO_EXT_TIME)
        fail(AISO_ERR_PARAM, "extension must be none, offset or time");
    // fdParam.cpp:18-25
    p.ntw = (p.nts - 1) * p.sub + 1;
    p.dtw = p.dts / float(p.sub);
    // fdParam.cpp:33-36,57,93
    int zPad = p.freeSurface ? p.zPadPlus : std::min(p.zPadMinus, p.zPadPlus);
    int xPad = std::min(p.xPadMinus, p.xPadPlus);
    p.minPad = std::min(zPad, xPad);
    // BornShotsGpuFunctions.cu:155,171,185
    if (p.sub > kSubMax) fail(AISO_ERR_PARAM, "Subsampling parameter is too high");
    if (p.minPad > kPadMax) fail(AISO_ERR_PARAM, "Padding value is too high");
    if (a.blockSize != 16) fail(AISO_ERR_PARAM, "Block size for time stepper is not consistent with parfile");
    // fdParam.cpp:258-268
    if ((p.nz - 2 * kFat) <= 0 || (p.nz - 2 * kFat) % a.blockSize != 0) fail(AISO_ERR_PARAM, "nz not a multiple of block size");
    if ((p.nx - 2 * kFat) <= 0 || (p.nx - 2 * kFat) % a.blockSize != 0) fail(AISO_ERR_PARAM, "nx not a multiple of block size");
    if (p.nExt > 1 && p.extension == AISO_EXT_OFFSET && 2 * p.hExt >= p.nx - 2 * kFat)
        fail(AISO_ERR_PARAM, "extension longer than the model");
    // The kernels address a slice as a 64-bit base plus a 32-bit element offset.
    if ((long)p.nz * p.nx > 0x7fffffffL) fail(AISO_ERR_PARAM, "grid of 2^31 points or more");

    // Laplacian coefficients: double literal / float(d*d), stored as float.
    static const double C[6] = {-2.927222222, 1.666666667, -0.238095238,
                                0.039682539, -0.004960317, 0.000317460};
    for (int k = 0; k < 6; ++k) {
        p.zc[k] = float(C[k] / double(p.dz * p.dz));
        p.xc[k] = float(C[k] / double(p.dx * p.dx));
    }
    // Linear time-interpolation filter scaled by 1/sqrt(ntw/nts).
    const int h = p.h();
    p.filt.assign(2 * h, 0.f);
    const double s = 1.0 / std::sqrt(double(float(p.ntw) / float(p.nts)));
    for (int i = 0; i < h; ++i) {
        float left = float(1.0 - 1.0 * i / p.sub);
        float right = float(1.0 - double(left));
        p.filt[i] = float(double(left) * s);
        p.filt[i + h] = float(double(right) * s);
    }
    // Cosine sponge: cd[j] = alpha + (1-alpha) cos(pi (minPad-j)/minPad).
    p.cd.assign(std::max(p.minPad, 1), 1.f);
    for (int j = 0; j < p.minPad; ++j) {
        float arg = float(M_PI / (1.0 * p.minPad) * 1.0 * (p.minPad - j));
        p.cd[j] = float(double(p.alphaCos) + (1.0 - double(p.alphaCos)) * std::cos(double(arg)));
    }
    return p;
}

// fdParam::checkFdStability / checkFdDispersion (F/src/fdParam.cpp:228-256)
// on the velocity without its FAT ring (:102-110).
void check_velocity(const FdParams& p, const float* vel)
{
    float vmax = -INFINITY, vmin = INFINITY;
    for (int ix = kFat; ix < p.nx - kFat; ++ix)
        for (int iz = kFat; iz < p.nz - kFat; ++iz) {
            float v = vel[(long)ix * p.nz + iz];
            if (!std::isfinite(v) || v <= 0.f) fail(AISO_ERR_PARAM, "velocity must be finite and positive");
            vmax = std::max(vmax, v);
            vmin = std::min(vmin, v);
        }
    float courant = vmax * p.dtw / std::min(p.dz, p.dx);
    if (courant > 0.45f) fail(AISO_ERR_PARAM, "Courant is too big: " + std::to_string(courant));
    float disp = vmin / (p.fMax * std::max(p.dz, p.dx));
    if (disp < 3.0f) fail(AISO_ERR_PARAM, "Dispersion is too small: " + std::to_string(disp) + " > 3");
}

void velocity_scales(const FdParams& p, const float* vel, float* v2dt2, float* refl)
{
    for (long i = 0; i < p.nzx(); ++i) {
        float v = vel[i];
        v2dt2[i] = v * v * p.dtw * p.dtw;
        refl[i] = float(2.0 / double(v * v * v));
    }
}

// ------------------------------------------------------------- geometry --
void Geometry::finalize()
{
    const int n = nDev * nb;
    regPos.clear();
    cornerReg.assign(n, 0);
    std::unordered_map<int, int> seen;
    seen.reserve(n * 2);
    for (int k = 0; k < n; ++k) {
        auto it = seen.find(gridIdx[k]);
        if (it == seen.end()) {
            int id = int(regPos.size());
            seen.emplace(gridIdx[k], id);
            regPos.push_back(gridIdx[k]);
            cornerReg[k] = id;
        } else {
            cornerReg[k] = it->second;
        }
    }
    const int nReg = int(regPos.size());
    adjPtr.assign(nReg + 1, 0);
    for (int k = 0; k < n; ++k) adjPtr[cornerReg[k] + 1]++;
    for (int r = 0; r < nReg; ++r) adjPtr[r + 1] += adjPtr[r];
    adjEnt.assign(n, 0);
    std::vector<int> fill(adjPtr.begin(), adjPtr.end() - 1);
    for (int k = 0; k < n; ++k) adjEnt[fill[cornerReg[k]]++] = k;
}

// deviceGpu::forward: d[dev][t] += w * m[reg][t] over corners in order.
void Geometry::forward(bool add, const float* reg, float* irreg, int ntr) const
{
    if (!add) std::memset(irreg, 0, sizeof(float) * (size_t)nDev * ntr);
    for (int i = 0; i < nDev; ++i)
        for (int c = 0; c < nb; ++c) {
            int k = i * nb + c;
            const float* m = reg + (long)cornerReg[k] * ntr;
            float* d = irreg + (long)i * ntr;
            for (int t = 0; t < ntr; ++t) d[t] += weight[k] * m[t];
        }
}

// deviceGpu::adjoint: m[reg][t] += w * d[dev][t] over (device, corner) in order.
void Geometry::adjoint(bool add, float* reg, const float* irreg, int ntr) const
{
    if (!add) std::memset(reg, 0, sizeof(float) * regPos.size() * ntr);
    for (int i = 0; i < nDev; ++i)
        for (int c = 0; c < nb; ++c) {
            int k = i * nb + c;
            float* m = reg + (long)cornerReg[k] * ntr;
            const float* d = irreg + (long)i * ntr;
            for (int t = 0; t < ntr; ++t) m[t] += weight[k] * d[t];
        }
}

static void grid_corners(Geometry& g, int i, int base, int nzVel, int dipole, int dipBase)
{
    int k = i * g.nb;
    g.gridIdx[k] = base;             g.weight[k] = 1.f;
    g.gridIdx[k + 1] = base + 1;     g.weight[k + 1] = 0.f;
    g.gridIdx[k + 2] = base + nzVel; g.weight[k + 2] = 0.f;
    g.gridIdx[k + 3] = base + nzVel + 1; g.weight[k + 3] = 0.f;
    if (dipole == 1) {
        g.gridIdx[k + 4] = dipBase;             g.weight[k + 4] = -1.f;
        g.gridIdx[k + 5] = dipBase + 1;         g.weight[k + 5] = 0.f;
        g.gridIdx[k + 6] = dipBase + nzVel;     g.weight[k + 6] = 0.f;
        g.gridIdx[k + 7] = dipBase + nzVel + 1; g.weight[k + 7] = 0.f;
    }
}

static void check_points(const Geometry& g, int nzVel, int nxVel)
{
    long n = (long)nzVel * nxVel;
    for (int v : g.gridIdx)
        if (v < 0 || v >= n) fail(AISO_ERR_GEOM, "One of the device is out of bounds");
}

// Constructor #1, F/src/deviceGpu.cpp:8-84 with the bounds check of :263-274.
Geometry geom_coords(int nDev, const float* z, const float* x, int nzVel, int nxVel,
                     float oz, float dz, float ox, float dx, int nt, int dipole,
                     float zShift, float xShift)
{
    if (nDev <= 0) fail(AISO_ERR_GEOM, "empty device list");
    Geometry g;
    g.nDev = nDev; g.nb = dipole == 1 ? 8 : 4; g.nt = nt;
    g.gridIdx.assign(nDev * g.nb, 0);
    g.weight.assign(nDev * g.nb, 0.f);
    float zMax = oz + (nzVel - 1) * dz, xMax = ox + (nxVel - 1) * dx;
    for (int i = 0; i < nDev; ++i)
        if (z[i] >= zMax || x[i] >= xMax) fail(AISO_ERR_GEOM, "One of the device is out of bounds");
    auto corners = [&](float zc, float xc, float sign, int k0) {
        float wz = (zc - oz) / dz, wx = (xc - ox) / dx;
        int zr = int(wz);
        wz = wz - zr;
        wz = float(1.0 - double(wz));
        int xr = int(wx);
        wx = wx - xr;
        wx = float(1.0 - double(wx));
        int b = xr * nzVel + zr;
        g.gridIdx[k0] = b;
        g.gridIdx[k0 + 1] = b + 1;
        g.gridIdx[k0 + 2] = b + nzVel;
        g.gridIdx[k0 + 3] = b + nzVel + 1;
        if (sign > 0) {
            g.weight[k0] = wz * wx;
            g.weight[k0 + 1] = float((1.0 - double(wz)) * double(wx));
            g.weight[k0 + 2] = float(double(wz) * (1.0 - double(wx)));
            g.weight[k0 + 3] = float((1.0 - double(wz)) * (1.0 - double(wx)));
        } else {
            g.weight[k0] = float(-1.0 * double(wz) * double(wx));
            g.weight[k0 + 1] = float(-1.0 * (1.0 - double(wz)) * double(wx));
            g.weight[k0 + 2] = float(-1.0 * double(wz) * (1.0 - double(wx)));
            g.weight[k0 + 3] = float(-1.0 * (1.0 - double(wz)) * (1.0 - double(wx)));
        }
    };
    for (int i = 0; i < nDev; ++i) {
        corners(z[i], x[i], 1.f, i * g.nb);
        if (dipole == 1) corners(z[i] + zShift, x[i] + xShift, -1.f, i * g.nb + 4);
    }
    check_points(g, nzVel, nxVel);
    g.finalize();
    return g;
}

// Constructor #2, F/src/deviceGpu.cpp:87-139 (bounds :276-284).
Geometry geom_ints(int nDev, const int* z, const int* x, int nzVel, int nxVel, int nt,
                   int dipole, int zShift, int xShift)
{
    if (nDev <= 0) fail(AISO_ERR_GEOM, "empty device list");
    int zm = *std::max_element(z, z + nDev), xm = *std::max_element(x, x + nDev);
    if (zm >= nzVel || xm >= nxVel) fail(AISO_ERR_GEOM, "One of the device is out of bounds");
    Geometry g;
    g.nDev = nDev; g.nb = dipole == 1 ? 8 : 4; g.nt = nt;
    g.gridIdx.assign(nDev * g.nb, 0);
    g.weight.assign(nDev * g.nb, 0.f);
    for (int i = 0; i < nDev; ++i)
        grid_corners(g, i, x[i] * nzVel + z[i], nzVel, dipole, (x[i] + xShift) * nzVel + z[i] + zShift);
    check_points(g, nzVel, nxVel);
    g.finalize();
    return g;
}

// Constructor #3, F/src/deviceGpu.cpp:142-199 (x outer loop, z inner).
Geometry geom_grid(int nzDev, int ozDev, int dzDev, int nxDev, int oxDev, int dxDev,
                   int nzVel, int nxVel, int nt, int dipole, int zShift, int xShift)
{
    if (nzDev <= 0 || nxDev <= 0) fail(AISO_ERR_GEOM, "empty device grid");
    int zMax = ozDev + (nzDev - 1) * dzDev, xMax = oxDev + (nxDev - 1) * dxDev;
    if (zMax >= nzVel || xMax >= nxVel) fail(AISO_ERR_GEOM, "One of the device is out of bounds");
    Geometry g;
    g.nDev = nzDev * nxDev; g.nb = dipole == 1 ? 8 : 4; g.nt = nt;
    g.gridIdx.assign(g.nDev * g.nb, 0);
    g.weight.assign(g.nDev * g.nb, 0.f);
    int i = 0;
    for (int ix = 0; ix < nxDev; ++ix) {
        int xd = oxDev + ix * dxDev;
        for (int iz = 0; iz < nzDev; ++iz, ++i) {
            int zd = ozDev + iz * dzDev;
            grid_corners(g, i, xd * nzVel + zd, nzVel, dipole, (xd + xShift) * nzVel + zd + zShift);
        }
    }
    check_points(g, nzVel, nxVel);
    g.finalize();
    return g;
}

// ------------------------------------------------------- signal chain --
static float interp_scale(const FdParams& p)
{
    float s = float(p.ntw) / float(p.nts);
    return float(1.0 / std::sqrt(double(s)));
}

// interpTimeLinTbb::forward, F/src/interpTimeLinTbb.cpp:26-62.
void interp_time_fwd(const FdParams& p, int nDev, const float* m, float* d)
{
    const float dtw = p.dts / p.sub, otw = p.ots, sc = interp_scale(p);
    for (int dv = 0; dv < nDev; ++dv) {
        const float* md = m + (long)dv * p.nts;
        float* dd = d + (long)dv * p.ntw;
        for (int itw = 0; itw < p.ntw - 1; ++itw) {
            float tw = otw + itw * dtw;
            float w = (tw - p.ots) / p.dts;
            int k = int(w);
            w = w - k;
            dd[itw] = float(((double)md[k] * (1.0 - (double)w) + (double)(md[k + 1] * w)) * (double)sc);
        }
        dd[p.ntw - 1] = sc * md[p.nts - 1];
    }
}

// interpTimeLinTbb::adjoint, F/src/interpTimeLinTbb.cpp:64-102.
void interp_time_adj(const FdParams& p, int nDev, float* m, const float* d)
{
    const float dtw = p.dts / p.sub, otw = p.ots, sc = interp_scale(p);
    for (int dv = 0; dv < nDev; ++dv) {
        float* md = m + (long)dv * p.nts;
        const float* dd = d + (long)dv * p.ntw;
        std::fill(md, md + p.nts, 0.f);
        for (int itw = 0; itw < p.ntw - 1; ++itw) {
            float tw = otw + itw * dtw;
            float w = (tw - p.ots) / p.dts;
            int k = int(w);
            w = w - k;
            md[k] = float((double)md[k] + (double)(sc * dd[itw]) * (1.0 - (double)w));
            md[k + 1] += sc * dd[itw] * w;
        }
        md[p.nts - 1] += sc * dd[p.ntw - 1];
    }
}

// secondTimeDerivative::forward with add=false, F/src/secondTimeDerivative.cpp:16-28.
void second_time_derivative(const FdParams& p, int nDev, const float* m, float* d)
{
    const int nt = p.nts;
    const float dt2 = float(1.0 / double(p.dts * p.dts));
    for (int dv = 0; dv < nDev; ++dv) {
        const float* x = m + (long)dv * nt;
        float* y = d + (long)dv * nt;
        y[0] = float((double)dt2 * ((double)x[1] - 2.0 * (double)x[0]));
        for (int it = 1; it < nt - 1; ++it)
            y[it] = float((double)dt2 * ((double)x[it + 1] - 2.0 * (double)x[it] + (double)x[it - 1]));
        y[nt - 1] = float((double)dt2 * ((double)x[nt - 2] - 2.0 * (double)x[nt - 1]));
    }
}

void scale_source(const FdParams& p, const float* vel, const std::vector<int>& pos, int nt, float* sig)
{
    for (size_t i = 0; i < pos.size(); ++i) {
        float sc = p.dtw * p.dtw * vel[pos[i]] * vel[pos[i]];
        for (int t = 0; t < nt; ++t) sig[i * nt + t] = sig[i * nt + t] * sc;
    }
}

// F/src/padFileGpuMain.cpp:49-66.
int pad_plus(int n, int pad, int blockSize)
{
    long total = (long)pad * 2 + n;
    float ratio = std::ceil(float(total) / float(blockSize));
    return int((long)ratio * blockSize - n - pad);
}

// F/src/padFileGpuMain.cpp:37-47 (free surface: no top padding, all of it below).
int pad_plus_free_surface(int nz, int zPad, int blockSize)
{
    long total = (long)zPad + nz;
    float ratio = std::ceil(float(total) / float(blockSize));
    return int((long)ratio * blockSize - nz);
}

}  // namespace aiso
