// born_impl.hpp -- BornShotsGpu / BornExtShotsGpu binding shared by the two
// Born modules (PY/pyAcoustic_iso_float_born.cpp:18-34,
// PY/pyAcoustic_iso_float_born_ext.cpp:18-35).
#pragma once
#include "common.hpp"

namespace aisopy {

// deviceGpu lives in the nonlinear module; make sure it is loaded (once).
inline void import_nl()
{
    py::dict mods = py::module_::import("sys").attr("modules");
    if (mods.contains("pyAcoustic_iso_float_nl")) return;
    try {
        py::module_::import("acoustic_isotropic_operators_amd.pyAcoustic_iso_float_nl");
    } catch (py::error_already_set&) {
        py::module_::import("pyAcoustic_iso_float_nl");
    }
}

template <bool EXT>
struct BornShots : ShotsBase {
    std::vector<float> wav;
    int nWav = 0;

    BornShots(py::object v, py::object p, std::vector<DevPtr> s, std::vector<py::object> sig, std::vector<DevPtr> r)
        : ShotsBase(std::move(v), std::move(p), std::move(s), std::move(r))
    {
        // sourcesSignalsVector: one wavelet [nSrcDev][nts] for all shots or one per shot.
        if (sig.size() != 1 && (int)sig.size() != nShot)
            throw std::runtime_error("sourcesSignalsVector must hold 1 or nShot wavelets");
        nWav = (int)sig.size();
        int nSrc = sources[0]->nIrreg();
        for (auto& o : sig) {
            F32 a = nd(o, "sourcesSignals");
            if (a.size() % nSrc) throw std::runtime_error("wavelet size does not match the source devices");
            wav.insert(wav.end(), a.data(), a.data() + a.size());
        }
    }

    int nExtOf() const { return EXT ? fp.nExt : 1; }

    void run(bool adj, bool add, py::object model, py::object data)
    {
        ensure();
        if ((ssize_t)wav.size() != (ssize_t)nWav * sources[0]->nIrreg() * fp.nts)
            throw std::runtime_error("**** [Seismic source file] ERROR: nts not consistent with parfile ****");
        F32 m = nd(model, "model"), d = nd(data, "data");
        int nRec = receivers[0]->nIrreg();
        if (m.size() != (ssize_t)nExtOf() * fp.nx * fp.nz) throw std::runtime_error("model size not consistent with parfile");
        if (d.size() != (ssize_t)nShot * nRec * fp.nts) throw std::runtime_error("data must be [nShot][nReceiver][nts]");
        auto s = raw(sources), r = raw(receivers);
        float* mp = m.mutable_data();
        float* dp = d.mutable_data();
        aiso_status st;
        {
            py::gil_scoped_release nogil;
            if (EXT)
                st = adj ? aiso_bornext_adjoint(ctx, add, mp, dp, s.data(), nShot, wav.data(), nWav, r.data(), (int)r.size(), AISO_MEM_HOST)
                         : aiso_bornext_forward(ctx, add, mp, dp, s.data(), nShot, wav.data(), nWav, r.data(), (int)r.size(), AISO_MEM_HOST);
            else
                st = adj ? aiso_born_adjoint(ctx, add, mp, dp, s.data(), nShot, wav.data(), nWav, r.data(), (int)r.size(), AISO_MEM_HOST)
                         : aiso_born_forward(ctx, add, mp, dp, s.data(), nShot, wav.data(), nWav, r.data(), (int)r.size(), AISO_MEM_HOST);
        }
        check(st);
    }
};

template <bool EXT>
void bind_born(py::module_& m, const char* name)
{
    using T = BornShots<EXT>;
    py::class_<T, std::shared_ptr<T>>(m, name)
        .def(py::init<py::object, py::object, std::vector<DevPtr>, std::vector<py::object>, std::vector<DevPtr>>())
        .def("forward", [](T& s, bool add, py::object mo, py::object da) { s.run(false, add, mo, da); })
        .def("adjoint", [](T& s, bool add, py::object mo, py::object da) { s.run(true, add, mo, da); })
        .def("forwardWavefield", [](T& s, bool add, py::object mo, py::object da) {
            s.with_wavefield([&] { s.run(false, add, mo, da); });
        })
        .def("adjointWavefield", [](T& s, bool add, py::object mo, py::object da) {
            s.with_wavefield([&] { s.run(true, add, mo, da); });
        })
        .def("setVel", [](T& s, py::object v) { s.take_vel(v); })
        .def("getVel", [](T& s) { return s.vel; })
        .def("dotTest", [](T& s, bool verbose, float maxError) {
            s.ensure();
            int nRec = s.receivers[0]->nIrreg();
            std::vector<ssize_t> dom = EXT ? std::vector<ssize_t>{s.fp.nExt, s.fp.nx, s.fp.nz}
                                           : std::vector<ssize_t>{s.fp.nx, s.fp.nz};
            return dot_test(dom, {s.nShot, nRec, s.fp.nts},
                            [&](bool add, F32& a, F32& b) { s.run(false, add, a, b); },
                            [&](bool add, F32& a, F32& b) { s.run(true, add, a, b); }, verbose, maxError);
        }, py::arg("verbose") = false, py::arg("maxError") = 1e-5f)
        .def("getSrcWfld", [](T& s) { return s.wavefield(AISO_WAVEFIELD_SOURCE); })
        .def("getSecWfld", [](T& s) { return s.wavefield(AISO_WAVEFIELD_SECONDARY); });
}

}  // namespace aisopy
