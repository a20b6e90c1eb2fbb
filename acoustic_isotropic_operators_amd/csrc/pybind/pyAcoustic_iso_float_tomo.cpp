// pyAcoustic_iso_float_tomo -- tomoExtShotsGpu (subsurface-offset extension)
// with the names and methods of PY/pyAcoustic_iso_float_tomo.cpp:14-35,
// backed by libaiso.so.
#include "born_impl.hpp"

namespace aisopy {

// Velocity perturbation [nx][nz] <-> data [nShot][nReceiver][nts] about the
// extended reflectivity [nExt][nx][nz] given at construction or by
// setReflectivityExt (F/include/tomoExtShotsGpu.h:37-58); parfile keys
// leg1 / leg2 (F/src/tomoExtGpu.cpp:10-11).
struct TomoShots : BornShots<false> {
    std::vector<float> refl;
    py::object reflObj;
    int reflGen = -1;                    // context generation the reflectivity was given to

    TomoShots(py::object v, py::object p, std::vector<DevPtr> s, std::vector<py::object> sig, std::vector<DevPtr> r,
              py::object reflExt)
        : BornShots<false>(std::move(v), std::move(p), std::move(s), std::move(sig), std::move(r))
    {
        take_refl(std::move(reflExt));
    }

    void take_refl(py::object o)
    {
        F32 a = nd(o, "reflectivityExt");
        refl.assign(a.data(), a.data() + a.size());
        reflObj = std::move(o);
        reflGen = -1;
    }

    void run(bool adj, bool add, py::object model, py::object data)
    {
        ensure();
        if (par.s("extension", "none") != "offset")
            throw std::runtime_error("**** ERROR: the tomographic operator needs extension=offset (time-lag is not supported) ****");
        if ((ssize_t)wav.size() != (ssize_t)nWav * sources[0]->nIrreg() * fp.nts)
            throw std::runtime_error("**** [Seismic source file] ERROR: nts not consistent with parfile ****");
        if ((ssize_t)refl.size() != (ssize_t)fp.nExt * fp.nx * fp.nz)
            throw std::runtime_error("**** [Reflectivity file] ERROR: size not consistent with parfile ****");
        F32 m = nd(model, "model"), d = nd(data, "data");
        int nRec = receivers[0]->nIrreg();
        if (m.size() != (ssize_t)fp.nx * fp.nz) throw std::runtime_error("model size not consistent with parfile");
        if (d.size() != (ssize_t)nShot * nRec * fp.nts) throw std::runtime_error("data must be [nShot][nReceiver][nts]");
        if (reflGen != ctxGen) {
            check(aiso_ctx_set_ext_reflectivity(ctx, refl.data(), AISO_MEM_HOST));
            reflGen = ctxGen;
        }
        check(aiso_ctx_set_option(ctx, "leg1", par.i("leg1", 1)));
        check(aiso_ctx_set_option(ctx, "leg2", par.i("leg2", 1)));
        auto s = raw(sources), r = raw(receivers);
        float* mp = m.mutable_data();
        float* dp = d.mutable_data();
        aiso_status st;
        {
            py::gil_scoped_release nogil;
            st = adj ? aiso_tomo_adjoint(ctx, add, mp, dp, s.data(), nShot, wav.data(), nWav, r.data(), (int)r.size(), AISO_MEM_HOST)
                     : aiso_tomo_forward(ctx, add, mp, dp, s.data(), nShot, wav.data(), nWav, r.data(), (int)r.size(), AISO_MEM_HOST);
        }
        check(st);
    }
};

}  // namespace aisopy

PYBIND11_MODULE(pyAcoustic_iso_float_tomo, m)
{
    using aisopy::TomoShots;
    using aisopy::DevPtr;
    aisopy::import_nl();
    py::add_ostream_redirect(m, "ostream_redirect");
    auto unsupported = [](TomoShots&, bool, py::object, py::object) {
        throw std::runtime_error("**** ERROR: wavefield saving is not supported by this engine ****");
    };
    py::class_<TomoShots, std::shared_ptr<TomoShots>>(m, "tomoExtShotsGpu")
        .def(py::init<py::object, py::object, std::vector<DevPtr>, std::vector<py::object>, std::vector<DevPtr>,
                      py::object>())
        .def("forward", [](TomoShots& s, bool add, py::object mo, py::object da) { s.run(false, add, mo, da); })
        .def("adjoint", [](TomoShots& s, bool add, py::object mo, py::object da) { s.run(true, add, mo, da); })
        .def("forwardWavefield", unsupported)
        .def("adjointWavefield", unsupported)
        .def("setVel", [](TomoShots& s, py::object v) { s.take_vel(v); })
        .def("setReflectivityExt", [](TomoShots& s, py::object r) { s.take_refl(r); })
        .def("dotTest", [](TomoShots& s, bool verbose, float maxError) {
            s.ensure();
            int nRec = s.receivers[0]->nIrreg();
            return aisopy::dot_test({s.fp.nx, s.fp.nz}, {s.nShot, nRec, s.fp.nts},
                                    [&](bool add, aisopy::F32& a, aisopy::F32& b) { s.run(false, add, a, b); },
                                    [&](bool add, aisopy::F32& a, aisopy::F32& b) { s.run(true, add, a, b); },
                                    verbose, maxError);
        }, py::arg("verbose") = false, py::arg("maxError") = 1e-5f);
    aisopy::alias_module(m, "pyAcoustic_iso_float_tomo");
}
