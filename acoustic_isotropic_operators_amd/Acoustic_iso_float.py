"""Operator layer with the reference's Python API
(PY/Acoustic_iso_float.py:94-1175): parameter handling, acquisition builders,
``*OpInitFloat`` builders and the ``Operator`` subclasses that wrap the
pybind11 modules.

Differences from the reference, by design:

* vectors are ``sep.FloatVector`` (numpy-backed, SEP axis order) and the
  parameter object is ``sep.ParamObj``; files are SEP ``.H`` headers read with
  ``sep.read`` -- genericIO/SepVector are external to the reference and absent;
* the Dask multi-node path (``client=``) is not provided: one process drives
  one GPU, and shots are batched on it by the engine.

Wavefield saving follows the reference: ``forwardWavefield`` /
``adjointWavefield`` run the operator and keep the coarse wavefields
[nts][nx][nz] of shot ``wavefieldShotNumber`` (parfile key, default 0);
``getWfld`` / ``getSrcWfld`` / ``getSecWfld`` return them as vectors on the
velocity's z and x axes and the time axis (PY/Acoustic_iso_float.py:663-699,
927-969).
"""
from __future__ import annotations

import numpy as np

from . import sep
from .pyOperator import Operator

try:  # the modules sit next to this file; the package dir may also be on sys.path
    from . import pyAcoustic_iso_float_nl, pyAcoustic_iso_float_born, pyAcoustic_iso_float_born_ext
    from . import pyAcoustic_iso_float_tomo, pyAcoustic_iso_float_wemva
except ImportError:  # pragma: no cover
    import pyAcoustic_iso_float_nl
    import pyAcoustic_iso_float_born
    import pyAcoustic_iso_float_born_ext
    import pyAcoustic_iso_float_tomo
    import pyAcoustic_iso_float_wemva

deviceGpu = pyAcoustic_iso_float_nl.deviceGpu
parfile2pars = sep.parfile2pars


def create_parObj(args):
    """Parameter object from ``key=value`` arguments (``par=`` files expanded)."""
    return sep.ParamObj(args)


def _no_client(client):
    if client is not None:
        raise NotImplementedError("the Dask client path is not provided: one process drives one GPU")


def _cpp(v):
    return v.getCpp() if hasattr(v, "getCpp") else v


def _time_axis(par):
    return sep.axis(n=par.getInt("nts", -1), o=par.getFloat("ots", 0.0), d=par.getFloat("dts", -1.0))


def _read_vel(par, key="vel"):
    f = par.getString(key, "noVelFile")
    if f == "noVelFile":
        raise IOError("**** ERROR: User did not provide velocity file ****")
    return sep.read(f, ndims=2)


def _read_sources(par):
    f = par.getString("sources", "noSourcesFile")
    if f == "noSourcesFile":
        raise IOError("**** ERROR: User did not provide seismic sources file ****")
    return sep.read(f, ndims=2)


def _dipole(par):
    return par.getInt("dipole", 0), par.getInt("zDipoleShift", 2), par.getInt("xDipoleShift", 0)


def buildSourceGeometry(parObject, vel, client=None):
    """One deviceGpu per shot (PY/Acoustic_iso_float.py:237-284): from
    ``sourceGeomFile`` (bilinear constructor on physical coordinates) or from
    zSource/xSource/spacingShots on the padded grid (iz = zSource-1+zPadMinus+fat).

    The geometry file is read as the reference's code reads it
    (PY/Acoustic_iso_float.py:250-259): its array is [3][nShot] (SEP axis 1 =
    shot, axis 2 = coordinate x, y, z), x at [0, shot] and z at [2, shot]."""
    _no_client(client)
    nShot = parObject.getInt("nShot")
    nts = parObject.getInt("nts")
    dip, zs, xs = _dipole(parObject)
    geomFile = parObject.getString("sourceGeomFile", "None")
    sources = []
    if geomFile != "None":
        g = sep.read(geomFile, ndims=2).getNdArray()          # [3][nShot]
        if g.shape[1] != nShot:
            raise ValueError(f"Number of shots (#shot={nShot}) not consistent with geometry file (#shots={g.shape[1]})!")
        for s in range(nShot):
            z = np.array([g[2, s]], np.float32)
            x = np.array([g[0, s]], np.float32)
            sources.append(deviceGpu(z, x, _cpp(vel), nts, dip, float(zs), float(xs)))
        return sources, sep.axis(n=nShot, o=1.0, d=1.0)
    xax = vel.getHyper().axes[1]
    fat = parObject.getInt("fat", 5)
    oz = parObject.getInt("zSource") - 1 + parObject.getInt("zPadMinus") + fat
    ox = parObject.getInt("xSource") - 1 + parObject.getInt("xPadMinus") + fat
    spacing = parObject.getInt("spacingShots")
    axis = sep.axis(n=nShot, o=xax.o + ox * xax.d, d=spacing * xax.d)
    for _ in range(nShot):
        sources.append(deviceGpu(1, oz, 1, 1, ox, 1, _cpp(vel), nts, dip, zs, xs))
        ox += spacing
    return sources, axis


def buildReceiversGeometry(parObject, vel, client=None):
    """Receivers (PY/Acoustic_iso_float.py:386-456): one deviceGpu per shot
    from ``receiverGeomFile``, or one constant regular line from
    depthReceiver/nReceiver/oReceiver/dReceiver.

    The geometry file is read as the reference's code reads it
    (PY/Acoustic_iso_float.py:414-431): its array is [3][nRec][nShot] (SEP
    axis 1 = shot, axis 2 = receiver, axis 3 = coordinate x, y, z), x at
    [0, :, shot] and z at [2, :, shot]."""
    _no_client(client)
    nts = parObject.getInt("nts")
    dip, zs, xs = _dipole(parObject)
    geomFile = parObject.getString("receiverGeomFile", "None")
    receivers = []
    if geomFile != "None":
        g = sep.read(geomFile, ndims=3).getNdArray()          # [3][nRec][nShot]
        nShot = parObject.getInt("nShot", -1)
        if nShot != g.shape[2]:
            raise ValueError(f"**** ERROR [buildReceiversGeometry]: Number of shots from parfile (#shot={nShot}) "
                             f"not consistent with receivers' geometry file (#shots={g.shape[2]}) ****")
        nRec = parObject.getInt("nReceiverPerShot", -1)
        if nRec != g.shape[1]:
            raise ValueError(f"**** ERROR [buildReceiversGeometry]: Number of receivers from parfile "
                             f"(#nReceiverPerShot={nRec}) not consistent with receivers' geometry file "
                             f"(#recs={g.shape[1]}) ****")
        for s in range(nShot):
            z = np.ascontiguousarray(g[2, :, s], np.float32)
            x = np.ascontiguousarray(g[0, :, s], np.float32)
            receivers.append(deviceGpu(z, x, _cpp(vel), nts, dip, float(zs), float(xs)))
        return receivers, sep.axis(n=nRec, o=1.0, d=1.0)
    xax = vel.getHyper().axes[1]
    fat = parObject.getInt("fat", 5)
    oz = parObject.getInt("depthReceiver") - 1 + parObject.getInt("zPadMinus") + fat
    n = parObject.getInt("nReceiver")
    ox = parObject.getInt("oReceiver") - 1 + parObject.getInt("xPadMinus") + fat
    dx = parObject.getInt("dReceiver")
    axis = sep.axis(n=n, o=xax.o + ox * xax.d, d=dx * xax.d)
    receivers.append(deviceGpu(1, oz, 1, n, ox, dx, _cpp(vel), nts, dip, zs, xs))
    return receivers, axis


def _data_vector(time, recAxis, srcAxis):
    return sep.getSepVector(axes=[time, recAxis, srcAxis])


# ------------------------------------------------------------- nonlinear --
def nonlinearOpInitFloat(args, client=None):
    """(model, data, vel, parObject, sources, receivers, modelLocal); model is
    one wavelet [1][1][nts] (PY/Acoustic_iso_float.py:558-622)."""
    _no_client(client)
    par = create_parObj(args)
    vel = _read_vel(par)
    time = _time_axis(par)
    one = sep.axis(n=1)
    model = sep.getSepVector(axes=[time, one, one])
    sources, srcAxis = buildSourceGeometry(par, vel)
    receivers, recAxis = buildReceiversGeometry(par, vel)
    data = _data_vector(time, recAxis, srcAxis)
    return model, data, vel, par, sources, receivers, model


class _Wrapped(Operator):
    """Shared forward/adjoint plumbing over a pybind11 operator object."""

    _mod = None
    _name = ""

    def __str__(self):
        return self._name

    def forward(self, add, model, data):
        with self._mod.ostream_redirect():
            self.pyOp.forward(add, _cpp(model), _cpp(data))

    def adjoint(self, add, model, data):
        with self._mod.ostream_redirect():
            self.pyOp.adjoint(add, _cpp(model), _cpp(data))

    def forwardWavefield(self, add, model, data):
        with self._mod.ostream_redirect():
            self.pyOp.forwardWavefield(add, _cpp(model), _cpp(data))

    def adjointWavefield(self, add, model, data):
        with self._mod.ostream_redirect():
            self.pyOp.adjointWavefield(add, _cpp(model), _cpp(data))

    def _keep_axes(self, velocity, paramP):
        """Axes of the saved wavefields: the velocity's z and x, and time."""
        vax = velocity.getHyper().axes if hasattr(velocity, "getHyper") else None
        self._wfld_axes = None
        if vax is not None:
            self._wfld_axes = [vax[0], vax[1], _time_axis(paramP)]

    def _wfld_vector(self, arr):
        hyper = sep.hypercube(axes=self._wfld_axes) if getattr(self, "_wfld_axes", None) else \
            sep.hypercube(ns=[arr.shape[2], arr.shape[1], arr.shape[0]])
        return sep.FloatVector(hyper, arr)

    def setVel(self, vel):
        with self._mod.ostream_redirect():
            self.pyOp.setVel(_cpp(vel))

    def dotTestCpp(self, verb=False, maxError=.00001):
        """The C++ class dot-product test (F/include/operator.h:50-88)."""
        with self._mod.ostream_redirect():
            return self.pyOp.dotTest(verb, maxError)


class nonlinearPropShotsGpu(_Wrapped):
    """Nonlinear propagator: wavelet(s) -> data (PY/Acoustic_iso_float.py:624-700)."""

    _mod = pyAcoustic_iso_float_nl
    _name = " NLOper "

    def __init__(self, domain, range, velocity, paramP, sourceVector, receiversVector):
        self.setDomainRange(domain, range)
        self.pyOp = pyAcoustic_iso_float_nl.nonlinearPropShotsGpu(_cpp(velocity), _cpp(paramP).param, sourceVector,
                                                                  receiversVector)
        self._keep_axes(velocity, paramP)

    def getWfld(self):
        """Wavefield saved by forwardWavefield / adjointWavefield."""
        with self._mod.ostream_redirect():
            return self._wfld_vector(self.pyOp.getWfld())


# ------------------------------------------------------------------ Born --
def BornOpInitFloat(args, client=None):
    """(model, data, vel, parObject, sources, sourcesSignals, receivers,
    modelLocal); model has the velocity's hypercube
    (PY/Acoustic_iso_float.py:809-883)."""
    _no_client(client)
    par = create_parObj(args)
    vel = _read_vel(par)
    model = sep.getSepVector(vel.getHyper())
    time = _time_axis(par)
    signals = _read_sources(par)
    sources, srcAxis = buildSourceGeometry(par, vel)
    receivers, recAxis = buildReceiversGeometry(par, vel)
    data = _data_vector(time, recAxis, srcAxis)
    return model, data, vel, par, sources, [signals], receivers, model


class BornShotsGpu(_Wrapped):
    """Born operator: reflectivity -> data (PY/Acoustic_iso_float.py:885-971)."""

    _mod = pyAcoustic_iso_float_born
    _name = " BornOp "

    def __init__(self, domain, range, velocity, paramP, sourceVector, sourcesSignalsVector, receiversVector):
        self.setDomainRange(domain, range)
        sig = [_cpp(s) for s in sourcesSignalsVector]
        self.pyOp = pyAcoustic_iso_float_born.BornShotsGpu(_cpp(velocity), _cpp(paramP).param, sourceVector, sig,
                                                           receiversVector)
        self._keep_axes(velocity, paramP)

    def getSrcWfld(self):
        """Source wavefield saved by forwardWavefield / adjointWavefield."""
        with self._mod.ostream_redirect():
            return self._wfld_vector(self.pyOp.getSrcWfld())

    def getSecWfld(self):
        """Scattered (forward) or scaled receiver (adjoint) wavefield."""
        with self._mod.ostream_redirect():
            return self._wfld_vector(self.pyOp.getSecWfld())


# --------------------------------------------------------- extended Born --
def _ext_axis(par):
    """Extension axis centred on zero: d = dts for time lags, d = dx for
    subsurface offsets (PY/Acoustic_iso_float.py:996-1020)."""
    extension = par.getString("extension", "noExtensionType")
    if extension == "noExtensionType":
        raise ValueError("**** ERROR: User did not provide extension type ****")
    if extension not in ("offset", "time"):
        raise ValueError(f"extension={extension}: the extension is offset or time")
    nExt = par.getInt("nExt", -1)
    if nExt == -1:
        raise ValueError("**** ERROR: User did not provide size of extension axis ****")
    if nExt % 2 == 0:
        raise ValueError("Length of extension axis must be an uneven number")
    dExt = par.getFloat("dts" if extension == "time" else "dx", -1.0)
    hExt = (nExt - 1) // 2
    return sep.axis(n=nExt, o=-dExt * hExt, d=dExt)


def BornExtOpInitFloat(args, client=None):
    """Extended model [nExt][nx][nz] with the extension axis centred on zero
    (offset: d = dx; PY/Acoustic_iso_float.py:973-1078)."""
    _no_client(client)
    par = create_parObj(args)
    vel = _read_vel(par)
    zAxis, xAxis = vel.getHyper().axes[0], vel.getHyper().axes[1]
    time = _time_axis(par)
    extAxis = _ext_axis(par)
    signals = _read_sources(par)
    model = sep.getSepVector(axes=[zAxis, xAxis, extAxis])
    sources, srcAxis = buildSourceGeometry(par, vel)
    receivers, recAxis = buildReceiversGeometry(par, vel)
    data = _data_vector(time, recAxis, srcAxis)
    return model, data, vel, par, sources, [signals], receivers, model


class BornExtShotsGpu(_Wrapped):
    """Extended Born operator, subsurface offsets or time lags (PY/Acoustic_iso_float.py:1080-1174)."""

    _mod = pyAcoustic_iso_float_born_ext
    _name = " BornExtOp "

    def __init__(self, domain, range, velocity, paramP, sourceVector, sourcesSignalsVector, receiversVector):
        self.setDomainRange(domain, range)
        sig = [_cpp(s) for s in sourcesSignalsVector]
        self.pyOp = pyAcoustic_iso_float_born_ext.BornExtShotsGpu(_cpp(velocity), _cpp(paramP).param, sourceVector,
                                                                  sig, receiversVector)
        self._keep_axes(velocity, paramP)

    def getSrcWfld(self):
        """Source wavefield saved by forwardWavefield / adjointWavefield
        (the reference's extended class keeps it without a Python getter)."""
        with self._mod.ostream_redirect():
            return self._wfld_vector(self.pyOp.getSrcWfld())


# ------------------------------------------------------- tomographic op --
def tomoExtOpInitFloat(args, client=None):
    """(model, data, vel, parObject, sources, sourcesSignals, receivers,
    reflectivityExt, modelLocal): model is a velocity perturbation on the
    velocity's hypercube; reflectivityExt [nExt][nx][nz] is read from the
    ``reflectivity`` parfile key, zero when absent
    (PY/Acoustic_iso_float.py:1306-1422)."""
    _no_client(client)
    par = create_parObj(args)
    vel = _read_vel(par)
    zAxis, xAxis = vel.getHyper().axes[0], vel.getHyper().axes[1]
    time = _time_axis(par)
    extAxis = _ext_axis(par)
    signals = _read_sources(par)
    reflFile = par.getString("reflectivity", "None")
    if reflFile == "None":
        refl = sep.getSepVector(axes=[zAxis, xAxis, extAxis])
        refl.scale(0.0)
    else:
        refl = sep.read(reflFile, ndims=3)
    model = sep.getSepVector(vel.getHyper())
    sources, srcAxis = buildSourceGeometry(par, vel)
    receivers, recAxis = buildReceiversGeometry(par, vel)
    data = _data_vector(time, recAxis, srcAxis)
    return model, data, vel, par, sources, [signals], receivers, refl, model


class tomoExtShotsGpu(_Wrapped):
    """Subsurface-offset tomographic operator: velocity perturbation -> data
    about a fixed extended reflectivity (PY/Acoustic_iso_float.py:1424-1510)."""

    _mod = pyAcoustic_iso_float_tomo
    _name = " TomoOp "

    def __init__(self, domain, range, velocity, paramP, sourceVector, sourcesSignalsVector, receiversVector,
                 reflectivityExt):
        self.setDomainRange(domain, range)
        sig = [_cpp(s) for s in sourcesSignalsVector]
        self.pyOp = pyAcoustic_iso_float_tomo.tomoExtShotsGpu(_cpp(velocity), _cpp(paramP).param, sourceVector, sig,
                                                              receiversVector, _cpp(reflectivityExt))

    def setReflectivityExt(self, reflectivityExt):
        with self._mod.ostream_redirect():
            self.pyOp.setReflectivityExt(_cpp(reflectivityExt))


# ------------------------------------------------------------------ WEMVA --
def wemvaExtOpInitFloat(args):
    """(model, data, vel, parObject, sources, sourcesSignals, receivers,
    receiversSignals): model is a velocity perturbation, data the extended
    image [nExt][nx][nz]; receiversSignals are the observed data of each shot
    [nReceiver][nts], read from the ``seismicData`` file [nShot][nReceiver][nts]
    (PY/Acoustic_iso_float.py:1512-1603)."""
    par = create_parObj(args)
    vel = _read_vel(par)
    sources, srcAxis = buildSourceGeometry(par, vel)
    receivers, recAxis = buildReceiversGeometry(par, vel)
    zAxis, xAxis = vel.getHyper().axes[0], vel.getHyper().axes[1]
    time = _time_axis(par)
    extAxis = _ext_axis(par)
    signals = _read_sources(par)
    dataFile = par.getString("seismicData", "noWemvaDataFile")
    if dataFile == "noWemvaDataFile":
        raise ValueError("**** ERROR: User did not provide wemva seismic data file ****")
    allData = sep.read(dataFile, ndims=3).getNdArray()
    recSignals = []
    for s in range(srcAxis.n):
        sl = sep.getSepVector(axes=[time, recAxis])
        sl.getNdArray()[...] = allData[s]
        recSignals.append(sl)
    data = sep.getSepVector(axes=[zAxis, xAxis, extAxis])
    model = sep.getSepVector(vel.getHyper())
    return model, data, vel, par, sources, [signals], receivers, recSignals


class wemvaExtShotsGpu(_Wrapped):
    """Subsurface-offset WEMVA operator: velocity perturbation -> extended
    image perturbation about the observed data (PY/Acoustic_iso_float.py:1605-1684)."""

    _mod = pyAcoustic_iso_float_wemva
    _name = " WemvaOp "

    def __init__(self, domain, range, velocity, paramP, sourceVector, sourcesSignalsVector, receiversVector,
                 receiversSignalsVector):
        self.setDomainRange(domain, range)
        sig = [_cpp(s) for s in sourcesSignalsVector]
        rsig = [_cpp(s) for s in receiversSignalsVector]
        self.pyOp = pyAcoustic_iso_float_wemva.wemvaExtShotsGpu(_cpp(velocity), _cpp(paramP).param, sourceVector, sig,
                                                                receiversVector, rsig)

    def setReceiversSignalsVector(self, receiversSignalsVector):
        with self._mod.ostream_redirect():
            self.pyOp.setReceiversSignalsVector([_cpp(s) for s in receiversSignalsVector])
