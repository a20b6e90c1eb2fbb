// host.hpp -- host-side parts of the engine: FD parameters and their checks,
// coefficient tables, acquisition geometry and the per-shot signal chain.
// These are the pieces the reference runs on the CPU before each shot
// (fdParam, deviceGpu, interpTimeLinTbb, secondTimeDerivative,
// scaleSeismicSource).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "aiso.h"

namespace aiso {

constexpr int kFat = 5;
constexpr int kSubMax = 100;   // F/include/varDeclare.h:12
constexpr int kPadMax = 200;   // F/include/varDeclare.h:11

struct Error : std::runtime_error {
    aiso_status code;
    Error(aiso_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// fdParam restated (F/src/fdParam.cpp:9-140) for the padded grid.
struct FdParams {
    int nz = 0, nx = 0;
    float dz = 0, dx = 0, oz = 0, ox = 0;
    int nts = 0, sub = 0, ntw = 0;
    float dts = 0, ots = 0, dtw = 0;
    int zPadMinus = 0, zPadPlus = 0, xPadMinus = 0, xPadPlus = 0;
    int minPad = 0;
    float alphaCos = 0.99f, fMax = 1000.f;
    int nExt = 1, hExt = 0, extension = AISO_EXT_NONE;
    int freeSurface = 0;       // 1: free surface on the first interior row, no top sponge

    // Derived tables, computed like initBornGpu (F/src/BornShotsGpuFunctions.cu:131-182).
    float zc[6] = {0}, xc[6] = {0};
    std::vector<float> filt;   // 2*(sub+1)
    std::vector<float> cd;     // minPad
    int h() const { return sub + 1; }
    long nzx() const { return (long)nz * nx; }
    long interior() const { return (long)(nz - 2 * kFat) * (nx - 2 * kFat); }
};

FdParams make_params(const aiso_fd_params& p);

// Velocity-dependent part of fdParam: checks + v^2 dtw^2 and 2/v^3.
void check_velocity(const FdParams& p, const float* vel);
void velocity_scales(const FdParams& p, const float* vel, float* v2dt2, float* refl);

// deviceGpu restated (F/src/deviceGpu.cpp).
struct Geometry {
    int nDev = 0, nb = 4, nt = 0;
    std::vector<int> gridIdx;     // nDev*nb raw corner indices
    std::vector<float> weight;    // nDev*nb
    std::vector<int> regPos;      // unique points in first-seen order
    std::vector<int> cornerReg;   // nDev*nb -> unique index
    // Transposed lists for the adjoint: for each unique point, the
    // (device*nb+corner) entries that map to it, in (device, corner) order.
    std::vector<int> adjPtr, adjEnt;
    void finalize();
    void forward(bool add, const float* reg, float* irreg, int nt) const;
    void adjoint(bool add, float* reg, const float* irreg, int nt) const;
};

Geometry geom_coords(int nDev, const float* z, const float* x, int nzVel, int nxVel,
                     float oz, float dz, float ox, float dx, int nt, int dipole,
                     float zShift, float xShift);
Geometry geom_ints(int nDev, const int* z, const int* x, int nzVel, int nxVel, int nt,
                   int dipole, int zShift, int xShift);
Geometry geom_grid(int nzDev, int ozDev, int dzDev, int nxDev, int oxDev, int dxDev,
                   int nzVel, int nxVel, int nt, int dipole, int zShift, int xShift);

// interpTimeLinTbb (F/src/interpTimeLinTbb.cpp): [nDev][nts] <-> [nDev][ntw].
void interp_time_fwd(const FdParams& p, int nDev, const float* m, float* d);
void interp_time_adj(const FdParams& p, int nDev, float* m, const float* d);
// secondTimeDerivative::forward (F/src/secondTimeDerivative.cpp:16-28).
void second_time_derivative(const FdParams& p, int nDev, const float* m, float* d);
// scaleSeismicSource (F/include/seismicOperator2D.cpp:73-86), in place.
void scale_source(const FdParams& p, const float* vel, const std::vector<int>& pos,
                  int nt, float* sig);

int pad_plus(int n, int pad, int blockSize);
int pad_plus_free_surface(int nz, int zPad, int blockSize);

}  // namespace aiso
