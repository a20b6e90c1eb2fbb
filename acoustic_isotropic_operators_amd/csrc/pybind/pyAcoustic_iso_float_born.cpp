// pyAcoustic_iso_float_born -- BornShotsGpu with the names and methods of
// PY/pyAcoustic_iso_float_born.cpp:14-35, backed by libaiso.so.
#include "born_impl.hpp"

PYBIND11_MODULE(pyAcoustic_iso_float_born, m)
{
    aisopy::import_nl();
    py::add_ostream_redirect(m, "ostream_redirect");
    aisopy::bind_born<false>(m, "BornShotsGpu");
    aisopy::alias_module(m, "pyAcoustic_iso_float_born");
}
