"""ctypes binding of the C-ABI in include/aiso.h (libaiso.so).

Used where buffers already live on the device (torch tensors, the serving
loop, bench.py): pointers are passed with ``mem=MEM_DEVICE``.  Raises
ImportError if the library has not been built, and RuntimeError carrying
``aiso_last_error()`` on any failed call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib_path

MEM_HOST, MEM_DEVICE = 0, 1
WAVEFIELD_SOURCE, WAVEFIELD_SECONDARY = 0, 1
EXT_NONE, EXT_OFFSET, EXT_TIME = 0, 1, 2
K_ALL, K_NL_FWD, K_NL_ADJ, K_SRC, K_BORN_SCAT, K_BORN_REC, K_EXT_SCAT, K_EXT_REC, K_BORN_PAIR, K_TOMO = range(-1, 9)


class FdParams(C.Structure):
    _fields_ = [("nz", C.c_int), ("nx", C.c_int), ("dz", C.c_float), ("dx", C.c_float),
                ("oz", C.c_float), ("ox", C.c_float), ("nts", C.c_int), ("dts", C.c_float),
                ("ots", C.c_float), ("sub", C.c_int), ("zPadMinus", C.c_int), ("zPadPlus", C.c_int),
                ("xPadMinus", C.c_int), ("xPadPlus", C.c_int), ("fat", C.c_int), ("blockSize", C.c_int),
                ("alphaCos", C.c_float), ("fMax", C.c_float), ("freeSurface", C.c_int),
                ("nExt", C.c_int), ("extension", C.c_int)]


_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_PP = C.POINTER(C.c_void_p)

_SIGS = {
    "aiso_last_error": (C.c_char_p, []),
    "aiso_ctx_create": (_I, [C.POINTER(FdParams), _I, _PP]),
    "aiso_ctx_destroy": (_I, [_P]),
    "aiso_ctx_set_velocity": (_I, [_P, _P, _I]),
    "aiso_ctx_set_batch": (_I, [_P, _I]),
    "aiso_ctx_set_tile": (_I, [_P, _I, _I]),
    "aiso_ctx_set_option": (_I, [_P, C.c_char_p, _I]),
    "aiso_ctx_get_option": (_I, [_P, C.c_char_p, C.POINTER(_I)]),
    "aiso_ctx_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(C.c_size_t)]),
    "aiso_ctx_stream": (_I, [_P, _PP]),
    "aiso_ctx_profile": (_I, [_P, _I]),
    "aiso_ctx_profile_read": (_I, [_P, _I, C.POINTER(C.c_longlong), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "aiso_ctx_batch_read": (_I, [_P, _P, _P, _I, C.POINTER(_I)]),
    "aiso_device_create_coords": (_I, [_I, _P, _P, _I, _I, _F, _F, _F, _F, _I, _I, _F, _F, _PP]),
    "aiso_device_create_ints": (_I, [_I, _P, _P, _I, _I, _I, _I, _I, _I, _PP]),
    "aiso_device_create_grid": (_I, [_I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _PP]),
    "aiso_device_destroy": (_I, [_P]),
    "aiso_device_sizes": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "aiso_device_reg_pos": (_I, [_P, _P]),
    "aiso_device_corners": (_I, [_P, _P, _P]),
    "aiso_device_forward": (_I, [_P, _I, _P, _P, _I]),
    "aiso_device_adjoint": (_I, [_P, _I, _P, _P, _I]),
    "aiso_nl_forward": (_I, [_P, _I, _P, _I, _P, _PP, _I, _PP, _I, _I]),
    "aiso_nl_adjoint": (_I, [_P, _I, _P, _I, _P, _PP, _I, _PP, _I, _I]),
    "aiso_nl_velocity_gradient": (_I, [_P, _P, _I, _P, _P, _PP, _I, _PP, _I, _I, _I]),
    "aiso_nl_forward_noalloc": (_I, [_P, _I, _P, _I, _P, _PP, _I, _PP, _I]),
    "aiso_born_forward": (_I, [_P, _I, _P, _P, _PP, _I, _P, _I, _PP, _I, _I]),
    "aiso_born_adjoint": (_I, [_P, _I, _P, _P, _PP, _I, _P, _I, _PP, _I, _I]),
    "aiso_bornext_forward": (_I, [_P, _I, _P, _P, _PP, _I, _P, _I, _PP, _I, _I]),
    "aiso_bornext_adjoint": (_I, [_P, _I, _P, _P, _PP, _I, _P, _I, _PP, _I, _I]),
    "aiso_ctx_set_ext_reflectivity": (_I, [_P, _P, _I]),
    "aiso_ctx_wavefield": (_I, [_P, _I, _P, _I]),
    "aiso_tomo_forward": (_I, [_P, _I, _P, _P, _PP, _I, _P, _I, _PP, _I, _I]),
    "aiso_tomo_adjoint": (_I, [_P, _I, _P, _P, _PP, _I, _P, _I, _PP, _I, _I]),
    "aiso_wemva_forward": (_I, [_P, _I, _P, _P, _P, _PP, _I, _P, _I, _PP, _I, _I]),
    "aiso_wemva_adjoint": (_I, [_P, _I, _P, _P, _P, _PP, _I, _P, _I, _PP, _I, _I]),
    "aiso_batch_create": (_I, [C.POINTER(FdParams), _I, _PP]),
    "aiso_batch_destroy": (_I, [_P]),
    "aiso_batch_set_windows": (_I, [_P, _I, _P, _P, _PP, _PP, _PP, _I]),
    "aiso_batch_set_config": (_I, [_P, _I, _I, _I, _I, _I, _I]),
    "aiso_batch_get_config": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I),
                                   C.POINTER(_I)]),
    "aiso_batch_tune": (_I, [_P, _I, C.POINTER(C.c_double)]),
    "aiso_batch_nl_forward": (_I, [_P, _I, _PP, _I, _PP, _I]),
    "aiso_batch_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(C.c_longlong), C.POINTER(_F),
                             C.POINTER(C.c_double)]),
    "aiso_batch_stream": (_I, [_P, _PP]),
    "aiso_pad_plus": (_I, [_I, _I, _I]),
    "aiso_pad_plus_free_surface": (_I, [_I, _I, _I]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        L = C.CDLL(lib_path())
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_SIGS)


def _check(st):
    if st != 0:
        raise RuntimeError(lib().aiso_last_error().decode(errors="replace"))


def _ptr(x):
    """Raw pointer of a numpy array or a torch tensor (host or device)."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return C.c_void_p(x.data_ptr())
    if isinstance(x, np.ndarray):
        return x.ctypes.data_as(C.c_void_p)
    return C.c_void_p(int(x))


def _nelem(x):
    return int(x.numel()) if hasattr(x, "numel") else int(x.size)


def _check_arg(x, n, mem, name, writable=False):
    """Refuse an operand the engine would misread or overrun: it must be a
    float32, C-contiguous array (numpy or a CPU tensor for MEM_HOST, a GPU tensor for
    MEM_DEVICE) of exactly ``n`` elements, writable when it is an output."""
    if x is None:
        raise ValueError(f"{name}: missing operand")
    if mem == MEM_DEVICE:
        if not hasattr(x, "data_ptr") or not hasattr(x, "is_cuda"):
            raise ValueError(f"{name}: MEM_DEVICE operands must be GPU tensors")
        if not x.is_cuda:
            raise ValueError(f"{name}: tensor is not on the GPU")
        import torch
        if x.dtype != torch.float32:
            raise ValueError(f"{name}: dtype {x.dtype}, expected float32")
        if not x.is_contiguous():
            raise ValueError(f"{name}: tensor is not contiguous")
    elif mem == MEM_HOST and hasattr(x, "data_ptr") and hasattr(x, "is_cuda"):
        import torch
        if x.is_cuda:
            raise ValueError(f"{name}: MEM_HOST tensor is on the GPU")
        if x.dtype != torch.float32:
            raise ValueError(f"{name}: dtype {x.dtype}, expected float32")
        if not x.is_contiguous():
            raise ValueError(f"{name}: tensor is not contiguous")
    elif mem == MEM_HOST:
        if not isinstance(x, np.ndarray):
            raise ValueError(f"{name}: MEM_HOST operands must be numpy arrays or CPU tensors")
        if x.dtype != np.float32:
            raise ValueError(f"{name}: dtype {x.dtype}, expected float32")
        if not x.flags.c_contiguous:
            raise ValueError(f"{name}: array is not C-contiguous")
        if writable and not x.flags.writeable:
            raise ValueError(f"{name}: output array is read-only")
    else:
        raise ValueError(f"mem must be MEM_HOST or MEM_DEVICE, got {mem}")
    if _nelem(x) != n:
        raise ValueError(f"{name}: {_nelem(x)} elements, expected {n}")


def pad_plus(n, pad, block=16):
    return lib().aiso_pad_plus(n, pad, block)


def pad_plus_free_surface(nz, zPad, block=16):
    return lib().aiso_pad_plus_free_surface(nz, zPad, block)


class Device:
    """deviceGpu handle (aiso_device*)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def grid(cls, nzDev, ozDev, dzDev, nxDev, oxDev, dxDev, nzVel, nxVel, nt, dipole=0, zShift=0, xShift=0):
        h = C.c_void_p()
        _check(lib().aiso_device_create_grid(nzDev, ozDev, dzDev, nxDev, oxDev, dxDev, nzVel, nxVel, nt,
                                             dipole, zShift, xShift, C.byref(h)))
        return cls(h)

    @classmethod
    def ints(cls, z, x, nzVel, nxVel, nt, dipole=0, zShift=0, xShift=0):
        z = np.ascontiguousarray(z, np.int32)
        x = np.ascontiguousarray(x, np.int32)
        h = C.c_void_p()
        _check(lib().aiso_device_create_ints(len(z), _ptr(z), _ptr(x), nzVel, nxVel, nt, dipole, zShift, xShift,
                                             C.byref(h)))
        return cls(h)

    @classmethod
    def coords(cls, z, x, nzVel, nxVel, oz, dz, ox, dx, nt, dipole=0, zShift=0.0, xShift=0.0):
        z = np.ascontiguousarray(z, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        h = C.c_void_p()
        _check(lib().aiso_device_create_coords(len(z), _ptr(z), _ptr(x), nzVel, nxVel, oz, dz, ox, dx, nt, dipole,
                                               zShift, xShift, C.byref(h)))
        return cls(h)

    def sizes(self):
        a, b, c = _I(), _I(), _I()
        _check(lib().aiso_device_sizes(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    @property
    def nDev(self):
        return self.sizes()[0]

    @property
    def nReg(self):
        return self.sizes()[1]

    def reg_pos(self):
        out = np.zeros(self.nReg, np.int32)
        _check(lib().aiso_device_reg_pos(self.h, _ptr(out)))
        return out

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.aiso_device_destroy(self.h)
            self.h = None


def _devarray(devs):
    arr = (C.c_void_p * len(devs))(*[d.h for d in devs])
    return C.cast(arr, _PP), arr


class Context:
    """aiso_ctx: one padded grid + velocity resident on one HIP device."""

    def __init__(self, params: dict, device=0):
        fp = FdParams()
        defaults = dict(oz=0.0, ox=0.0, ots=0.0, fat=5, blockSize=16, alphaCos=0.99, fMax=1000.0,
                        freeSurface=0, nExt=1, extension=EXT_NONE)
        merged = {**defaults, **params}
        for k, _ in FdParams._fields_:
            setattr(fp, k, merged[k])
        self.params = merged
        self.options = {}
        self.h = C.c_void_p()
        _check(lib().aiso_ctx_create(C.byref(fp), device, C.byref(self.h)))

    def close(self):
        if getattr(self, "h", None):
            lib().aiso_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        if _LIB is not None:
            self.close()

    def set_velocity(self, vel, mem=MEM_HOST):
        if mem == MEM_HOST:
            vel = np.ascontiguousarray(vel, np.float32)
        _check_arg(vel, self._nzx(), mem, "velocity")
        _check(lib().aiso_ctx_set_velocity(self.h, _ptr(vel), mem))

    def set_batch(self, n):
        _check(lib().aiso_ctx_set_batch(self.h, n))

    def set_tile(self, tz, tx):
        _check(lib().aiso_ctx_set_tile(self.h, tz, tx))

    def set_option(self, key, value):
        _check(lib().aiso_ctx_set_option(self.h, key.encode(), int(value)))
        self.options[key] = int(value)

    def get_option(self, key):
        v = _I()
        _check(lib().aiso_ctx_get_option(self.h, key.encode(), C.byref(v)))
        return v.value

    # operand sizes ---------------------------------------------------------
    def _nzx(self):
        return int(self.params["nz"]) * int(self.params["nx"])

    def _nts(self):
        return int(self.params["nts"])

    def _data_n(self, sources, receivers):
        return len(sources) * receivers[0].nDev * self._nts()

    def _image_n(self, nShot, ext):
        n = self._nzx() * (int(self.params.get("nExt", 1)) if ext else 1)
        return n * nShot if self.options.get("imagePerShot", 0) else n

    def info(self):
        a, b, c = _I(), _I(), C.c_size_t()
        _check(lib().aiso_ctx_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(ntw=a.value, batch=b.value, bytes_per_shot=c.value)

    def stream(self):
        s = C.c_void_p()
        _check(lib().aiso_ctx_stream(self.h, C.byref(s)))
        return s.value

    def profile(self, enable=True):
        _check(lib().aiso_ctx_profile(self.h, int(enable)))

    def profile_read(self, which=K_ALL):
        n, ms, pts = C.c_longlong(), C.c_double(), C.c_double()
        _check(lib().aiso_ctx_profile_read(self.h, which, C.byref(n), C.byref(ms), C.byref(pts)))
        return dict(launches=n.value, ms=ms.value, point_updates=pts.value)

    def batch_read(self):
        n = _I()
        _check(lib().aiso_ctx_batch_read(self.h, None, None, 0, C.byref(n)))
        ms = np.zeros(max(n.value, 1), np.float32)
        sh = np.zeros(max(n.value, 1), np.int32)
        _check(lib().aiso_ctx_batch_read(self.h, _ptr(ms), _ptr(sh), n.value, C.byref(n)))
        return ms[:n.value], sh[:n.value]

    # operators ------------------------------------------------------------
    def nl_forward(self, model, data, sources, receivers, add=False, mem=MEM_HOST):
        nSig = model.shape[0]
        if nSig not in (1, len(sources)):
            raise ValueError(f"model: shot axis {nSig}, expected 1 or {len(sources)}")
        _check_arg(model, nSig * sources[0].nDev * self._nts(), mem, "model")
        _check_arg(data, self._data_n(sources, receivers), mem, "data", writable=True)
        s, _ks = _devarray(sources)
        r, _kr = _devarray(receivers)
        _check(lib().aiso_nl_forward(self.h, int(add), _ptr(model), model.shape[0], _ptr(data), s, len(sources),
                                     r, len(receivers), mem))

    def nl_forward_noalloc(self, model, data, sources, receivers, add=False):
        """aiso_nl_forward_noalloc: model and data are GPU tensors."""
        nSig = model.shape[0]
        if nSig not in (1, len(sources)):
            raise ValueError(f"model: shot axis {nSig}, expected 1 or {len(sources)}")
        _check_arg(model, nSig * sources[0].nDev * self._nts(), MEM_DEVICE, "model")
        _check_arg(data, self._data_n(sources, receivers), MEM_DEVICE, "data", writable=True)
        s, _ks = _devarray(sources)
        r, _kr = _devarray(receivers)
        _check(lib().aiso_nl_forward_noalloc(self.h, int(add), _ptr(model), nSig, _ptr(data), s, len(sources), r,
                                             len(receivers)))

    def nl_adjoint(self, model, data, sources, receivers, add=False, mem=MEM_HOST):
        nSig = model.shape[0]
        if nSig not in (1, len(sources)):
            raise ValueError(f"model: shot axis {nSig}, expected 1 or {len(sources)}")
        _check_arg(model, nSig * sources[0].nDev * self._nts(), mem, "model", writable=True)
        _check_arg(data, self._data_n(sources, receivers), mem, "data")
        s, _ks = _devarray(sources)
        r, _kr = _devarray(receivers)
        _check(lib().aiso_nl_adjoint(self.h, int(add), _ptr(model), model.shape[0], _ptr(data), s, len(sources),
                                     r, len(receivers), mem))

    def nl_velocity_gradient(self, model, data_grad, vel_grad, sources, receivers, checkpoint_every=0,
                             mem=MEM_HOST):
        """dJ/d(velocity) [nx][nz] given dJ/d(data) of nl_forward(model)."""
        s, _ks = _devarray(sources)
        r, _kr = _devarray(receivers)
        if mem == MEM_HOST:
            model = np.ascontiguousarray(model, np.float32)
        nSig = model.shape[0]
        if nSig not in (1, len(sources)):
            raise ValueError(f"model: shot axis {nSig}, expected 1 or {len(sources)}")
        _check_arg(model, nSig * sources[0].nDev * self._nts(), mem, "model")
        _check_arg(data_grad, self._data_n(sources, receivers), mem, "data_grad")
        _check_arg(vel_grad, self._nzx(), mem, "vel_grad", writable=True)
        _check(lib().aiso_nl_velocity_gradient(self.h, _ptr(model), model.shape[0], _ptr(data_grad), _ptr(vel_grad),
                                               s, len(sources), r, len(receivers), int(checkpoint_every), mem))

    def _born(self, fn, model, data, sources, wavelets, receivers, add, mem, adjoint, ext):
        if adjoint:
            _check_arg(model, self._image_n(len(sources), ext), mem, "model", writable=True)
            _check_arg(data, self._data_n(sources, receivers), mem, "data")
        else:
            _check_arg(model, self._nzx() * (int(self.params.get("nExt", 1)) if ext else 1), mem, "model")
            _check_arg(data, self._data_n(sources, receivers), mem, "data", writable=True)
        s, _ks = _devarray(sources)
        r, _kr = _devarray(receivers)
        w = self._wavelets(wavelets, sources)
        nwav = w.shape[0] if w.ndim == 3 else 1
        _check(fn(self.h, int(add), _ptr(model), _ptr(data), s, len(sources), _ptr(w), nwav, r, len(receivers), mem))

    def _wavelets(self, wavelets, sources):
        """Host wavelets [nSrcDev][nts] or [nWavelet][nSrcDev][nts], nWavelet in {1, nShot}."""
        w = np.ascontiguousarray(wavelets, np.float32)
        nwav = w.shape[0] if w.ndim == 3 else 1
        if nwav not in (1, len(sources)):
            raise ValueError(f"wavelets: {nwav} blocks, expected 1 or {len(sources)}")
        n = nwav * sources[0].nDev * self._nts()
        if w.size != n:
            raise ValueError(f"wavelets: {w.size} elements, expected {n}")
        return w

    def born_forward(self, model, data, sources, wavelets, receivers, add=False, mem=MEM_HOST):
        self._born(lib().aiso_born_forward, model, data, sources, wavelets, receivers, add, mem, False, False)

    def born_adjoint(self, model, data, sources, wavelets, receivers, add=False, mem=MEM_HOST):
        self._born(lib().aiso_born_adjoint, model, data, sources, wavelets, receivers, add, mem, True, False)

    def bornext_forward(self, model, data, sources, wavelets, receivers, add=False, mem=MEM_HOST):
        self._born(lib().aiso_bornext_forward, model, data, sources, wavelets, receivers, add, mem, False, True)

    def bornext_adjoint(self, model, data, sources, wavelets, receivers, add=False, mem=MEM_HOST):
        self._born(lib().aiso_bornext_adjoint, model, data, sources, wavelets, receivers, add, mem, True, True)

    def wavefield(self, which=WAVEFIELD_SOURCE):
        """A saved wavefield [nts][nx][nz] (host copy) of the last call made
        with the option saveWavefield set."""
        p = self.params
        out = np.zeros((int(p["nts"]), int(p["nx"]), int(p["nz"])), np.float32)
        _check(lib().aiso_ctx_wavefield(self.h, int(which), _ptr(out), MEM_HOST))
        return out

    def set_ext_reflectivity(self, refl, mem=MEM_HOST):
        """Extended reflectivity [nExt][nx][nz] of the tomographic operator."""
        if mem == MEM_HOST:
            refl = np.ascontiguousarray(refl, np.float32)
        _check_arg(refl, self._nzx() * int(self.params.get("nExt", 1)), mem, "reflectivity")
        self._ext_refl = refl
        _check(lib().aiso_ctx_set_ext_reflectivity(self.h, _ptr(refl), mem))

    def tomo_forward(self, model, data, sources, wavelets, receivers, add=False, mem=MEM_HOST):
        self._born(lib().aiso_tomo_forward, model, data, sources, wavelets, receivers, add, mem, False, False)

    def tomo_adjoint(self, model, data, sources, wavelets, receivers, add=False, mem=MEM_HOST):
        self._born(lib().aiso_tomo_adjoint, model, data, sources, wavelets, receivers, add, mem, True, False)

    def _wemva(self, fn, model, image, data, sources, wavelets, receivers, add, mem, adjoint):
        nExt = int(self.params.get("nExt", 1))
        _check_arg(model, self._nzx(), mem, "model", writable=adjoint)
        _check_arg(image, self._nzx() * nExt, mem, "image", writable=not adjoint)
        _check_arg(data, self._data_n(sources, receivers), mem, "data")
        s, _ks = _devarray(sources)
        r, _kr = _devarray(receivers)
        w = self._wavelets(wavelets, sources)
        nwav = w.shape[0] if w.ndim == 3 else 1
        _check(fn(self.h, int(add), _ptr(model), _ptr(image), _ptr(data), s, len(sources), _ptr(w), nwav, r,
                  len(receivers), mem))

    def wemva_forward(self, model, image, data, sources, wavelets, receivers, add=False, mem=MEM_HOST):
        """Velocity perturbation [nx][nz] -> extended image [nExt][nx][nz] about
        the observed data [nShot][nRec][nts]."""
        self._wemva(lib().aiso_wemva_forward, model, image, data, sources, wavelets, receivers, add, mem, False)

    def wemva_adjoint(self, model, image, data, sources, wavelets, receivers, add=False, mem=MEM_HOST):
        self._wemva(lib().aiso_wemva_adjoint, model, image, data, sources, wavelets, receivers, add, mem, True)


def _fd_params(params):
    fp = FdParams()
    defaults = dict(nz=0, nx=0, oz=0.0, ox=0.0, ots=0.0, fat=5, blockSize=16, alphaCos=0.99, fMax=1000.0,
                    freeSurface=0, nExt=1, extension=EXT_NONE)
    merged = {**defaults, **params}
    for k, _ in FdParams._fields_:
        setattr(fp, k, merged[k])
    return fp, merged


def _ptrarray(xs):
    arr = (C.c_void_p * len(xs))(*[_ptr(x).value for x in xs])
    return C.cast(arr, _PP), arr


class Batch:
    """aiso_batch: many shots, each on a model window of its own size,
    advanced together (one launch per fine step over every window of a
    launch group, the time loop of each group captured in a HIP graph).

    ``params`` holds the keys every window shares (spacing, time sampling,
    pads, sponge, fMax); ``set_windows`` gives each window's padded velocity
    [nx][nz] and its source and receiver devices built on that window's
    grid."""

    def __init__(self, params: dict, device=0):
        fp, self.params = _fd_params(params)
        self.h = C.c_void_p()
        self.windows = []
        _check(lib().aiso_batch_create(C.byref(fp), device, C.byref(self.h)))

    def close(self):
        if getattr(self, "h", None):
            lib().aiso_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        if _LIB is not None:
            self.close()

    def set_windows(self, vels, sources, receivers, mem=MEM_HOST):
        n = len(vels)
        if len(sources) != n or len(receivers) != n:
            raise ValueError("one source and one receiver device per window")
        if mem == MEM_HOST:
            vels = [np.ascontiguousarray(v, np.float32) for v in vels]
        nz = np.array([v.shape[1] for v in vels], np.int32)
        nx = np.array([v.shape[0] for v in vels], np.int32)
        for k, v in enumerate(vels):
            _check_arg(v, int(nz[k]) * int(nx[k]), mem, f"velocity of window {k}")
        vp, _kv = _ptrarray(vels)
        s, _ks = _devarray(sources)
        r, _kr = _devarray(receivers)
        _check(lib().aiso_batch_set_windows(self.h, n, _ptr(nz), _ptr(nx), vp, s, r, mem))
        self.windows = [(int(nz[k]), int(nx[k]), sources[k], receivers[k]) for k in range(n)]
        self._keep = (sources, receivers)

    def set_config(self, tz=32, tx=16, windows_per_launch=0, blocks_per_cu=0, graph=True, lanes=2):
        _check(lib().aiso_batch_set_config(self.h, tz, tx, windows_per_launch, blocks_per_cu, lanes, int(graph)))

    def config(self):
        v = [_I() for _ in range(6)]
        _check(lib().aiso_batch_get_config(self.h, *[C.byref(x) for x in v]))
        return dict(tz=v[0].value, tx=v[1].value, windows_per_launch=v[2].value, blocks_per_cu=v[3].value,
                    lanes=v[4].value, graph=bool(v[5].value))

    def tune(self, steps=16, cache=None):
        """Pick tile, windows per launch, workgroups per CU and lanes by timing
        (aiso_batch_tune).  With ``cache`` (a JSON file path) the choice is
        looked up by the windows' signature first and stored after tuning."""
        key = self.signature()
        if cache is not None:
            import json
            import os
            table = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    table = json.load(f)
            if key in table:
                c = table[key]
                self.set_config(c["tz"], c["tx"], c["windows_per_launch"], c["blocks_per_cu"],
                                self.config()["graph"], c.get("lanes", 2))
                return dict(c, cached=True)
        t = C.c_double()
        _check(lib().aiso_batch_tune(self.h, int(steps), C.byref(t)))
        c = self.config()
        out = dict(tz=c["tz"], tx=c["tx"], windows_per_launch=c["windows_per_launch"],
                   blocks_per_cu=c["blocks_per_cu"], lanes=c["lanes"], ns_per_point_update=t.value * 1e6)
        if cache is not None:
            table[key] = out
            tmp = cache + ".tmp"
            with open(tmp, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return dict(out, cached=False)

    def signature(self):
        """Operator and window sizes: the autotuner's cache key."""
        import hashlib
        dims = ",".join(f"{nz}x{nx}" for nz, nx, _, _ in self.windows)
        return "nl_fwd:" + hashlib.sha1(dims.encode()).hexdigest()[:16] + f":{len(self.windows)}"

    def info(self):
        n, g, pts, ms, cap = _I(), _I(), C.c_longlong(), _F(), C.c_double()
        _check(lib().aiso_batch_info(self.h, C.byref(n), C.byref(g), C.byref(pts), C.byref(ms), C.byref(cap)))
        return dict(windows=n.value, launch_groups=g.value, point_updates=pts.value, last_ms=ms.value,
                    capture_ms=cap.value)

    def stream(self):
        s = C.c_void_p()
        _check(lib().aiso_batch_stream(self.h, C.byref(s)))
        return s.value

    def nl_forward(self, models, datas, add=False, mem=MEM_HOST):
        """models: one source signal [nSrcDev][nts] for all windows, or one per
        window; datas[w]: [nRecDev_w][nts] outputs."""
        n = len(self.windows)
        if len(datas) != n:
            raise ValueError(f"data: {len(datas)} arrays, expected {n}")
        if len(models) not in (1, n):
            raise ValueError(f"models: {len(models)} arrays, expected 1 or {n}")
        nts = int(self.params["nts"])
        nsrc = self.windows[0][2].nDev
        for k, m in enumerate(models):
            _check_arg(m, nsrc * nts, mem, f"model {k}")
        for k, d in enumerate(datas):
            _check_arg(d, self.windows[k][3].nDev * nts, mem, f"data {k}", writable=True)
        mp, _km = _ptrarray(models)
        dp, _kd = _ptrarray(datas)
        _check(lib().aiso_batch_nl_forward(self.h, int(add), mp, len(models), dp, mem))
