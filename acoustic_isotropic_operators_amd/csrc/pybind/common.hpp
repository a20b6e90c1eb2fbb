// common.hpp -- shared pieces of the pybind11 modules that mirror the
// reference's pyAcoustic_iso_float_{nl,born,born_ext} bindings
// (PY/pyAcoustic_iso_float_nl.cpp, _born.cpp, _born_ext.cpp) on top of the
// C-ABI of libaiso.so.
//
// Vectors are duck-typed: anything with getNdArray() (the package's
// FloatVector, SepVector-like objects) or a float32 numpy array.  Outputs are
// written in place, as the reference writes into its SEP vectors.  The
// parameter object is duck-typed too: getInt/getFloat/getString/getInts with
// an optional default, like SEP's paramObj.
#pragma once

#include <pybind11/iostream.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "aiso.h"

namespace py = pybind11;

namespace aisopy {

using F32 = py::array_t<float, py::array::c_style>;

inline void check(aiso_status s)
{
    if (s != AISO_OK) throw std::runtime_error(aiso_last_error());
}

// The ndarray behind a vector object, without copying; must be float32 and
// C-contiguous so in-place writes reach the caller.
inline F32 nd(py::handle v, const char* what)
{
    py::object o = py::reinterpret_borrow<py::object>(v);
    if (py::hasattr(o, "getNdArray")) o = o.attr("getNdArray")();
    if (!py::isinstance<py::array>(o)) throw std::runtime_error(std::string(what) + ": not an array-backed vector");
    py::array a = o.cast<py::array>();
    if (!a.dtype().is(py::dtype::of<float>())) throw std::runtime_error(std::string(what) + ": vector must be float32");
    if (!(a.flags() & py::array::c_style)) throw std::runtime_error(std::string(what) + ": vector must be C-contiguous");
    return py::reinterpret_borrow<F32>(a);
}

struct Axis {
    int n = 1;
    double o = 0.0, d = 1.0;
};

// Axes of a vector (SEP order: axis 1 is the fastest).
inline std::vector<Axis> axes(py::handle v)
{
    std::vector<Axis> out;
    py::object o = py::reinterpret_borrow<py::object>(v);
    if (py::hasattr(o, "getHyper")) {
        py::object ax = o.attr("getHyper")().attr("axes");
        for (py::handle a : ax) {
            Axis x;
            x.n = a.attr("n").cast<int>();
            x.o = a.attr("o").cast<double>();
            x.d = a.attr("d").cast<double>();
            out.push_back(x);
        }
        return out;
    }
    F32 a = nd(v, "vector");
    for (ssize_t k = a.ndim() - 1; k >= 0; --k) {
        Axis x;
        x.n = (int)a.shape(k);
        out.push_back(x);
    }
    return out;
}

// paramObj access with the reference's defaults.
struct Par {
    py::object p;
    explicit Par(py::object o) : p(std::move(o)) {}
    int i(const char* k) const { return p.attr("getInt")(k).cast<int>(); }
    int i(const char* k, int d) const { return p.attr("getInt")(k, d).cast<int>(); }
    float f(const char* k) const { return p.attr("getFloat")(k).cast<float>(); }
    float f(const char* k, float d) const { return p.attr("getFloat")(k, d).cast<float>(); }
    std::string s(const char* k, const char* d) const { return p.attr("getString")(k, d).cast<std::string>(); }
    std::vector<int> ints(const char* k, std::vector<int> d) const
    {
        return p.attr("getInts")(k, d).cast<std::vector<int>>();
    }
};

// The modules can be imported as top-level names (like the reference's) or
// through the package; both names must map to the one module object, or
// pybind11 would register the same types twice.
inline void alias_module(py::module_& m, const char* shortName)
{
    py::dict mods = py::module_::import("sys").attr("modules");
    mods[shortName] = m;
    mods[(std::string("acoustic_isotropic_operators_amd.") + shortName).c_str()] = m;
}

// deviceGpu handle shared by Python objects.
struct Device {
    aiso_device* d = nullptr;
    explicit Device(aiso_device* x) : d(x) {}
    ~Device() { aiso_device_destroy(d); }
    int nIrreg() const { int a, b, c; check(aiso_device_sizes(d, &a, &b, &c)); return a; }
    int nReg() const { int a, b, c; check(aiso_device_sizes(d, &a, &b, &c)); return b; }
    std::vector<int> regPos() const
    {
        std::vector<int> r(nReg());
        check(aiso_device_reg_pos(d, r.data()));
        return r;
    }
};
using DevPtr = std::shared_ptr<Device>;

inline void bind_device(py::module_& m)
{
    py::class_<Device, DevPtr>(m, "deviceGpu")
        // Overloads are tried in order: the integer forms first, so that
        // coordinate vectors fall through to constructor #1.
        // Constructor #3: o/d/n grid of devices (PY/pyAcoustic_iso_float_nl.cpp:23).
        .def(py::init([](int nzD, int ozD, int dzD, int nxD, int oxD, int dxD, py::object vel, int nt, int dipole,
                         int zs, int xs) {
                 auto ax = axes(vel);
                 aiso_device* d = nullptr;
                 check(aiso_device_create_grid(nzD, ozD, dzD, nxD, oxD, dxD, ax[0].n, ax[1].n, nt, dipole, zs, xs, &d));
                 return std::make_shared<Device>(d);
             }),
             py::arg("nzDevice"), py::arg("ozDevice"), py::arg("dzDevice"), py::arg("nxDevice"), py::arg("oxDevice"),
             py::arg("dxDevice"), py::arg("vel"), py::arg("nt"), py::arg("dipole") = 0, py::arg("zDipoleShift") = 0,
             py::arg("xDipoleShift") = 0)
        // Constructor #2: integer grid positions (PY/pyAcoustic_iso_float_nl.cpp:21).
        .def(py::init([](std::vector<int> z, std::vector<int> x, py::object vel, int nt, int dipole, int zs, int xs) {
                 auto ax = axes(vel);
                 if (z.size() != x.size()) throw std::runtime_error("zGridVector and xGridVector differ in length");
                 aiso_device* d = nullptr;
                 check(aiso_device_create_ints((int)z.size(), z.data(), x.data(), ax[0].n, ax[1].n, nt, dipole, zs, xs, &d));
                 return std::make_shared<Device>(d);
             }),
             py::arg("zGridVector").noconvert(), py::arg("xGridVector").noconvert(), py::arg("vel"), py::arg("nt"), py::arg("dipole") = 0,
             py::arg("zDipoleShift") = 0, py::arg("xDipoleShift") = 0)
        // Constructor #1: coordinates + velocity axes (PY/pyAcoustic_iso_float_nl.cpp:19).
        .def(py::init([](py::object z, py::object x, py::object vel, int nt, int dipole, double zs, double xs) {
                 F32 za = nd(z, "zCoord"), xa = nd(x, "xCoord");
                 auto ax = axes(vel);
                 if (za.size() != xa.size()) throw std::runtime_error("zCoord and xCoord differ in length");
                 aiso_device* d = nullptr;
                 check(aiso_device_create_coords((int)za.size(), za.data(), xa.data(), ax[0].n, ax[1].n, (float)ax[0].o,
                                                 (float)ax[0].d, (float)ax[1].o, (float)ax[1].d, nt, dipole, (float)zs,
                                                 (float)xs, &d));
                 return std::make_shared<Device>(d);
             }),
             py::arg("zCoord"), py::arg("xCoord"), py::arg("vel"), py::arg("nt"), py::arg("dipole") = 0,
             py::arg("zDipoleShift") = 0.0, py::arg("xDipoleShift") = 0.0)
        .def("getNDeviceReg", &Device::nReg)
        .def("getNDeviceIrreg", &Device::nIrreg)
        .def("getRegPosUnique", &Device::regPos)
        .def("getSizePosUnique", &Device::nReg)
        .def("getRegPos", [](const Device& self) {
            int n, r, nb;
            check(aiso_device_sizes(self.d, &n, &r, &nb));
            std::vector<int> g((size_t)n * nb);
            check(aiso_device_corners(self.d, g.data(), nullptr));
            return g;
        })
        .def("getWeights", [](const Device& self) {
            int n, r, nb;
            check(aiso_device_sizes(self.d, &n, &r, &nb));
            std::vector<float> w((size_t)n * nb);
            check(aiso_device_corners(self.d, nullptr, w.data()));
            return w;
        })
        // deviceGpu::forward / ::adjoint (F/src/deviceGpu.cpp:226-261).
        .def("forward", [](const Device& self, bool add, py::object reg, py::object irreg) {
            F32 a = nd(reg, "signalReg"), b = nd(irreg, "signalIrreg");
            int nt = (int)a.shape(a.ndim() - 1);
            if (a.size() != (ssize_t)self.nReg() * nt || b.size() != (ssize_t)self.nIrreg() * nt)
                throw std::runtime_error("deviceGpu.forward: shape mismatch");
            check(aiso_device_forward(self.d, add, a.data(), b.mutable_data(), nt));
        })
        .def("adjoint", [](const Device& self, bool add, py::object reg, py::object irreg) {
            F32 a = nd(reg, "signalReg"), b = nd(irreg, "signalIrreg");
            int nt = (int)a.shape(a.ndim() - 1);
            if (a.size() != (ssize_t)self.nReg() * nt || b.size() != (ssize_t)self.nIrreg() * nt)
                throw std::runtime_error("deviceGpu.adjoint: shape mismatch");
            check(aiso_device_adjoint(self.d, add, a.mutable_data(), b.data(), nt));
        });
}

// Shared state of every *ShotsGpu operator: parameters, velocity, geometry
// and the engine context.  The reference rebuilds fdParam (and runs its
// checks) at each forward/adjoint call; the context here is (re)validated
// lazily at the first call after construction or setVel, so errors surface
// at the same point.
struct ShotsBase {
    Par par;
    py::object vel;
    std::vector<float> velHost;
    std::vector<DevPtr> sources, receivers;
    aiso_ctx* ctx = nullptr;
    bool dirty = true;
    int ctxGen = 0;                      // bumped each time ensure() makes a new context
    int nShot = 0, gpu = 0;
    aiso_fd_params fp{};

    ShotsBase(py::object v, py::object p, std::vector<DevPtr> s, std::vector<DevPtr> r)
        : par(std::move(p)), vel(std::move(v)), sources(std::move(s)), receivers(std::move(r))
    {
        nShot = par.i("nShot");
        // createGpuIdList (F/src/BornShotsGpu.cpp:26-58): one process drives one
        // GPU here; the first listed id is used.
        int nGpu = par.i("nGpu", -1);
        std::vector<int> ids = par.ints("iGpu", {-1});
        if (nGpu <= 0 && (ids.empty() || ids[0] < 0))
            throw std::runtime_error("**** ERROR: Please provide a list of GPUs to be used ****");
        gpu = (!ids.empty() && ids[0] >= 0) ? ids[0] : 0;
        if ((int)sources.size() != nShot)
            throw std::runtime_error("number of source geometries differs from nShot");
        if (receivers.size() != 1 && (int)receivers.size() != nShot)
            throw std::runtime_error("receiver geometries must be 1 or nShot");
        take_vel(vel);
    }
    ~ShotsBase()
    {
        if (ctx) aiso_ctx_destroy(ctx);
    }

    void take_vel(py::object v)
    {
        F32 a = nd(v, "velocity");
        velHost.assign(a.data(), a.data() + a.size());
        vel = v;
        dirty = true;
    }

    // fdParam keys (F/src/fdParam.cpp:14-96) plus its vel/parfile consistency checks.
    void ensure()
    {
        if (!dirty) return;
        auto ax = axes(vel);
        aiso_fd_params q{};
        q.nz = par.i("nz");
        q.nx = par.i("nx");
        if (q.nz != ax[0].n) throw std::runtime_error("**** ERROR [fdParam]: nz from vel tag not consistent with nz from parfile ****");
        if (q.nx != ax[1].n) throw std::runtime_error("**** ERROR [fdParam]: nx from vel tag not consistent with nx from parfile ****");
        q.dz = par.f("dz", -1.f);
        q.dx = par.f("dx", -1.f);
        q.oz = (float)ax[0].o;
        q.ox = (float)ax[1].o;
        q.nts = par.i("nts");
        q.dts = par.f("dts", 0.f);
        q.ots = par.f("ots", 0.f);
        q.sub = par.i("sub");
        q.zPadMinus = par.i("zPadMinus");
        q.zPadPlus = par.i("zPadPlus");
        q.xPadMinus = par.i("xPadMinus");
        q.xPadPlus = par.i("xPadPlus");
        q.fat = par.i("fat");
        q.blockSize = par.i("blockSize");
        q.alphaCos = par.f("alphaCos", 0.99f);
        q.fMax = par.f("fMax", 1000.f);
        q.freeSurface = par.i("freeSurface", 0);
        q.nExt = par.i("nExt", 1);
        std::string ext = par.s("extension", "none");
        q.extension = ext == "offset" ? AISO_EXT_OFFSET : ext == "time" ? AISO_EXT_TIME : AISO_EXT_NONE;
        if (ctx) {
            aiso_ctx_destroy(ctx);
            ctx = nullptr;
        }
        check(aiso_ctx_create(&q, gpu, &ctx));
        ++ctxGen;
        check(aiso_ctx_set_velocity(ctx, velHost.data(), AISO_MEM_HOST));
        // Engine-specific keys (the reference ignores them): shots advanced
        // per launch and the step-kernel tile.
        int batch = par.i("shotsPerLaunch", 0);
        if (batch > 0) check(aiso_ctx_set_batch(ctx, batch));
        const int tz = par.i("tileZ", 0), tx = par.i("tileX", 0);
        if (tz > 0 || tx > 0) check(aiso_ctx_set_tile(ctx, tz > 0 ? tz : 32, tx > 0 ? tx : 32));
        check(aiso_ctx_set_option(ctx, "lockstep", par.i("lockstep", 1)));
        check(aiso_ctx_set_option(ctx, "pair2", par.i("twoStep", 0)));
        // BornExtGpu's sloth key (F/src/BornExtGpu.cpp:14)
        check(aiso_ctx_set_option(ctx, "sloth", par.i("sloth", 0)));
        fp = q;
        dirty = false;
    }

    std::vector<aiso_device*> raw(const std::vector<DevPtr>& v) const
    {
        std::vector<aiso_device*> r;
        for (auto& d : v) r.push_back(d->d);
        return r;
    }

    // forwardWavefield / adjointWavefield (F/src/BornShotsGpu.cpp:145-240,
    // 341-438): the operator call, saving the wavefields of shot
    // wavefieldShotNumber (parfile key, default 0) on the way; they are
    // copied to the host, as the reference keeps them in host vectors.
    py::object savedSrc = py::none(), savedSec = py::none();

    template <class F>
    void with_wavefield(F&& call)
    {
        ensure();
        check(aiso_ctx_set_option(ctx, "wavefieldShot", par.i("wavefieldShotNumber", 0)));
        check(aiso_ctx_set_option(ctx, "saveWavefield", 1));
        try {
            call();
        } catch (...) {
            aiso_ctx_set_option(ctx, "saveWavefield", 0);
            throw;
        }
        check(aiso_ctx_set_option(ctx, "saveWavefield", 0));
        savedSrc = fetch(AISO_WAVEFIELD_SOURCE);
        savedSec = fetch(AISO_WAVEFIELD_SECONDARY);
    }

    // A saved wavefield [nts][nx][nz] from the context, or None if the last
    // call saved none of that kind.
    py::object fetch(int which)
    {
        F32 out({(ssize_t)fp.nts, (ssize_t)fp.nx, (ssize_t)fp.nz});
        float* o = out.mutable_data();
        aiso_status st;
        {
            py::gil_scoped_release nogil;
            st = aiso_ctx_wavefield(ctx, which, o, AISO_MEM_HOST);
        }
        if (st == AISO_ERR_STATE) return py::none();
        check(st);
        return std::move(out);
    }

    // getSrcWavefield / getSecWavefield / getWavefield of the reference
    // classes (F/include/BornShotsGpu.h:50-51, nonlinearPropShotsGpu.h:47).
    py::object wavefield(int which) const
    {
        const py::object& w = which == AISO_WAVEFIELD_SOURCE ? savedSrc : savedSec;
        if (w.is_none())
            throw std::runtime_error(which == AISO_WAVEFIELD_SOURCE
                                         ? "**** ERROR: no wavefield saved: call forwardWavefield or adjointWavefield first ****"
                                         : "**** ERROR: no secondary wavefield saved (Born forwardWavefield / adjointWavefield) ****");
        return w;
    }
};

// Operator::dotTest (F/include/operator.h:50-88) with uniform random vectors.
template <class Fwd, class Adj>
bool dot_test(std::vector<ssize_t> dshape, std::vector<ssize_t> rshape, Fwd fwd, Adj adj, bool verbose, float maxError)
{
    std::mt19937 gen(1234);
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    F32 d1(dshape), d2(dshape), r1(rshape), r2(rshape);
    for (ssize_t i = 0; i < d1.size(); ++i) d1.mutable_data()[i] = U(gen);
    for (ssize_t i = 0; i < r1.size(); ++i) r1.mutable_data()[i] = U(gen);
    auto dot = [](const F32& a, const F32& b) {
        double s = 0;
        for (ssize_t i = 0; i < a.size(); ++i) s += (double)a.data()[i] * b.data()[i];
        return s;
    };
    fwd(false, d1, r2);
    adj(false, d2, r1);
    double dot1 = dot(r1, r2), dot2 = dot(d1, d2);
    double e1 = std::fabs((dot1 - dot2) / dot2);
    fwd(true, d1, r2);
    adj(true, d2, r1);
    double dot3 = dot(r1, r2), dot4 = dot(d1, d2);
    double e2 = std::fabs((dot3 - dot4) / dot4);
    if (verbose) {
        py::print("WITHOUT ADD:\ndot1:", dot1, "\ndot2:", dot2, "\nerror WITHOUT ADD:", e1);
        py::print("WITH ADD:\ndot3:", dot3, "\ndot4:", dot4, "\nerror with ADD:", e2);
    }
    return e1 <= maxError && e2 <= maxError;
}

}  // namespace aisopy
