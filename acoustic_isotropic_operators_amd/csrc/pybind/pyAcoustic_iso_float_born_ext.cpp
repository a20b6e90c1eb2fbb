// pyAcoustic_iso_float_born_ext -- BornExtShotsGpu (subsurface-offset or time-lag
// extension) with the names and methods of
// PY/pyAcoustic_iso_float_born_ext.cpp:14-35, backed by libaiso.so.
#include "born_impl.hpp"

PYBIND11_MODULE(pyAcoustic_iso_float_born_ext, m)
{
    aisopy::import_nl();
    py::add_ostream_redirect(m, "ostream_redirect");
    aisopy::bind_born<true>(m, "BornExtShotsGpu");
    aisopy::alias_module(m, "pyAcoustic_iso_float_born_ext");
}
