// pyAcoustic_iso_float_wemva -- wemvaExtShotsGpu (subsurface-offset
// extension) with the names and methods of
// PY/pyAcoustic_iso_float_wemva.cpp:14-33, backed by libaiso.so.
#include "born_impl.hpp"

namespace aisopy {

// Velocity perturbation [nx][nz] <-> extended image [nExt][nx][nz] about the
// observed data of each shot (receiversSignalsVector, [nReceiver][nts] per
// shot; F/include/wemvaExtShotsGpu.h:36-57); parfile keys leg1 / leg2.
struct WemvaShots : BornShots<false> {
    std::vector<float> dobs;             // [nShot][nReceiver][nts]

    WemvaShots(py::object v, py::object p, std::vector<DevPtr> s, std::vector<py::object> sig, std::vector<DevPtr> r,
               std::vector<py::object> recSig)
        : BornShots<false>(std::move(v), std::move(p), std::move(s), std::move(sig), std::move(r))
    {
        take_data(recSig);
    }

    void take_data(const std::vector<py::object>& recSig)
    {
        if ((int)recSig.size() != nShot) throw std::runtime_error("receiversSignalsVector must hold nShot data slices");
        dobs.clear();
        for (auto& o : recSig) {
            F32 a = nd(o, "receiversSignals");
            dobs.insert(dobs.end(), a.data(), a.data() + a.size());
        }
    }

    void run(bool adj, bool add, py::object model, py::object data)
    {
        ensure();
        if (par.s("extension", "none") != "offset")
            throw std::runtime_error("**** ERROR: the WEMVA operator needs extension=offset (time-lag is not supported) ****");
        if ((ssize_t)wav.size() != (ssize_t)nWav * sources[0]->nIrreg() * fp.nts)
            throw std::runtime_error("**** [Seismic source file] ERROR: nts not consistent with parfile ****");
        const int nRec = receivers[0]->nIrreg();
        if ((ssize_t)dobs.size() != (ssize_t)nShot * nRec * fp.nts)
            throw std::runtime_error("**** [Seismic data file] ERROR: data must be [nShot][nReceiver][nts] ****");
        F32 m = nd(model, "model"), im = nd(data, "data");
        if (m.size() != (ssize_t)fp.nx * fp.nz) throw std::runtime_error("model size not consistent with parfile");
        if (im.size() != (ssize_t)fp.nExt * fp.nx * fp.nz)
            throw std::runtime_error("extended image must be [nExt][nx][nz]");
        check(aiso_ctx_set_option(ctx, "leg1", par.i("leg1", 1)));
        check(aiso_ctx_set_option(ctx, "leg2", par.i("leg2", 1)));
        auto s = raw(sources), r = raw(receivers);
        float* mp = m.mutable_data();
        float* ip = im.mutable_data();
        aiso_status st;
        {
            py::gil_scoped_release nogil;
            st = adj ? aiso_wemva_adjoint(ctx, add, mp, ip, dobs.data(), s.data(), nShot, wav.data(), nWav, r.data(),
                                          (int)r.size(), AISO_MEM_HOST)
                     : aiso_wemva_forward(ctx, add, mp, ip, dobs.data(), s.data(), nShot, wav.data(), nWav, r.data(),
                                          (int)r.size(), AISO_MEM_HOST);
        }
        check(st);
    }
};

}  // namespace aisopy

PYBIND11_MODULE(pyAcoustic_iso_float_wemva, m)
{
    using aisopy::DevPtr;
    using aisopy::WemvaShots;
    aisopy::import_nl();
    py::add_ostream_redirect(m, "ostream_redirect");
    auto unsupported = [](WemvaShots&, bool, py::object, py::object) {
        throw std::runtime_error("**** ERROR: wavefield saving is not supported by this engine ****");
    };
    py::class_<WemvaShots, std::shared_ptr<WemvaShots>>(m, "wemvaExtShotsGpu")
        .def(py::init<py::object, py::object, std::vector<DevPtr>, std::vector<py::object>, std::vector<DevPtr>,
                      std::vector<py::object>>())
        .def("forward", [](WemvaShots& s, bool add, py::object mo, py::object da) { s.run(false, add, mo, da); })
        .def("adjoint", [](WemvaShots& s, bool add, py::object mo, py::object da) { s.run(true, add, mo, da); })
        .def("forwardWavefield", unsupported)
        .def("adjointWavefield", unsupported)
        .def("setVel", [](WemvaShots& s, py::object v) { s.take_vel(v); })
        .def("setReceiversSignalsVector", [](WemvaShots& s, std::vector<py::object> d) { s.take_data(d); })
        .def("dotTest", [](WemvaShots& s, bool verbose, float maxError) {
            s.ensure();
            return aisopy::dot_test({s.fp.nx, s.fp.nz}, {s.fp.nExt, s.fp.nx, s.fp.nz},
                                    [&](bool add, aisopy::F32& a, aisopy::F32& b) { s.run(false, add, a, b); },
                                    [&](bool add, aisopy::F32& a, aisopy::F32& b) { s.run(true, add, a, b); },
                                    verbose, maxError);
        }, py::arg("verbose") = false, py::arg("maxError") = 1e-5f);
    aisopy::alias_module(m, "pyAcoustic_iso_float_wemva");
}
