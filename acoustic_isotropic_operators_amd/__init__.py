"""MI355X-native 2D acoustic isotropic finite-difference engine.

Nonlinear, Born and subsurface-offset extended-Born operators (2nd order in
time, 10th order in space, cosine sponge) with the reference operator API of
gbarnier/acoustic_isotropic_operators:

* ``pyAcoustic_iso_float_nl`` / ``_born`` / ``_born_ext``: pybind11 modules with
  the reference's module, class and method names (C++ over ``libaiso.so``);
* ``Acoustic_iso_float``: the Python operator layer (``*OpInitFloat`` builders,
  ``Operator`` subclasses);
* ``include/aiso.h``: the C-ABI of ``libaiso.so`` (HIP kernels for gfx950).

The native libraries are built in-tree by ``acoustic_isotropic_operators_amd.build``.
"""
import os

__all__ = ["PKG_DIR", "lib_path"]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))


def lib_path():
    """Path of the built C-ABI library; raises if it has not been built."""
    p = os.path.join(PKG_DIR, "libaiso.so")
    if not os.path.exists(p):
        raise ImportError("libaiso.so is not built: run python -m acoustic_isotropic_operators_amd.build")
    return p
