// pyAcoustic_iso_float_nl -- deviceGpu and nonlinearPropShotsGpu, with the
// names and methods of PY/pyAcoustic_iso_float_nl.cpp:14-40 (the reference's
// nonlinear module), backed by libaiso.so.
#include "common.hpp"

using namespace aisopy;

namespace {

struct NlShots : ShotsBase {
    using ShotsBase::ShotsBase;

    // model [nShotSig][nSrcDev][nts], data [nShot][nRec][nts]
    // (F/include/nonlinearPropShotsGpu.h:27-31).
    void run(bool adj, bool add, py::object model, py::object data)
    {
        ensure();
        F32 m = nd(model, "model"), d = nd(data, "data");
        int nSrcDev = sources[0]->nIrreg(), nRec = receivers[0]->nIrreg();
        if (m.ndim() != 3 || m.shape(1) != nSrcDev || m.shape(2) != fp.nts || (m.shape(0) != 1 && m.shape(0) != nShot))
            throw std::runtime_error("model must be [1 or nShot][nSource][nts]");
        if (d.size() != (ssize_t)nShot * nRec * fp.nts) throw std::runtime_error("data must be [nShot][nReceiver][nts]");
        auto s = raw(sources), r = raw(receivers);
        int nSig = (int)m.shape(0);
        float* mp = m.mutable_data();
        float* dp = d.mutable_data();
        aiso_status st;
        {
            py::gil_scoped_release nogil;
            st = adj ? aiso_nl_adjoint(ctx, add, mp, nSig, dp, s.data(), nShot, r.data(), (int)r.size(), AISO_MEM_HOST)
                     : aiso_nl_forward(ctx, add, mp, nSig, dp, s.data(), nShot, r.data(), (int)r.size(), AISO_MEM_HOST);
        }
        check(st);
    }
};

}  // namespace

PYBIND11_MODULE(pyAcoustic_iso_float_nl, m)
{
    py::add_ostream_redirect(m, "ostream_redirect");
    bind_device(m);
    py::class_<NlShots, std::shared_ptr<NlShots>>(m, "nonlinearPropShotsGpu")
        .def(py::init<py::object, py::object, std::vector<DevPtr>, std::vector<DevPtr>>(),
             "Initialize a nonlinearPropShotsGpu")
        .def("forward", [](NlShots& s, bool add, py::object mo, py::object da) { s.run(false, add, mo, da); })
        .def("adjoint", [](NlShots& s, bool add, py::object mo, py::object da) { s.run(true, add, mo, da); })
        .def("forwardWavefield", [](NlShots& s, bool add, py::object mo, py::object da) {
            s.with_wavefield([&] { s.run(false, add, mo, da); });
        })
        .def("adjointWavefield", [](NlShots& s, bool add, py::object mo, py::object da) {
            s.with_wavefield([&] { s.run(true, add, mo, da); });
        })
        .def("setVel", [](NlShots& s, py::object v) { s.take_vel(v); })
        .def("getVel", [](NlShots& s) { return s.vel; })
        .def("dotTest", [](NlShots& s, bool verbose, float maxError) {
            s.ensure();
            int nSrc = s.sources[0]->nIrreg(), nRec = s.receivers[0]->nIrreg();
            return dot_test({1, nSrc, s.fp.nts}, {s.nShot, nRec, s.fp.nts},
                            [&](bool add, F32& a, F32& b) { s.run(false, add, a, b); },
                            [&](bool add, F32& a, F32& b) { s.run(true, add, a, b); }, verbose, maxError);
        }, py::arg("verbose") = false, py::arg("maxError") = 1e-5f)
        .def("getWfld", [](NlShots& s) { return s.wavefield(AISO_WAVEFIELD_SOURCE); });
    alias_module(m, "pyAcoustic_iso_float_nl");
}
