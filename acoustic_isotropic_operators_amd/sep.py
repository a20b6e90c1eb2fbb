"""SEP-style vectors, hypercubes, parameter objects and .H file IO.

The reference's Python layer passes SepVector objects (numpy-backed arrays with
a Hypercube of axes, axis 1 fastest) and a genericIO parameter object (parfile /
argv ``key=value`` pairs) to its operators (PY/Acoustic_iso_float.py:94-114,
237-456).  Those libraries are external to the reference and absent here, so
this module provides the subset the operators need, with the same method names:

* ``axis`` / ``hypercube`` (n, o, d per axis, axis 1 fastest);
* ``FloatVector`` with getNdArray/getHyper/getCpp/clone/cloneSpace/zero/scale/
  scaleAdd/dot/norm/rand/set/copy;
* ``ParamObj`` with getInt/getFloat/getString/getInts/getFloats, reading
  ``key=value`` arguments and ``par=`` files like genericIO;
* ``read``/``write`` for SEP ``.H`` headers with ``in=`` binary (xdr_float
  big-endian or native float), e.g. notebooks/velocityMarmousi.H.
"""
from __future__ import annotations

import os
import re
import shlex

import numpy as np


class axis:
    def __init__(self, n=1, o=0.0, d=1.0, label="", unit=""):
        self.n, self.o, self.d, self.label, self.unit = int(n), float(o), float(d), label, unit

    def __repr__(self):
        return f"axis(n={self.n}, o={self.o}, d={self.d})"


class hypercube:
    """Axes in SEP order (axes[0] is the fastest)."""

    def __init__(self, axes=None, ns=None, os_=None, ds=None):
        if axes is None:
            ns = list(ns)
            os_ = list(os_) if os_ is not None else [0.0] * len(ns)
            ds = list(ds) if ds is not None else [1.0] * len(ns)
            axes = [axis(n, o, d) for n, o, d in zip(ns, os_, ds)]
        self.axes = [axis(a.n, a.o, a.d, a.label, a.unit) for a in axes]

    def getNdim(self):
        return len(self.axes)

    def getAxis(self, i):
        """1-based, like the C++ hypercube."""
        return self.axes[i - 1]

    def getNs(self):
        return [a.n for a in self.axes]

    def shape(self):
        """numpy shape (slowest first)."""
        return tuple(a.n for a in reversed(self.axes))

    def checkSame(self, other):
        return len(self.axes) == len(other.axes) and all(
            a.n == b.n and abs(a.o - b.o) < 1e-6 and abs(a.d - b.d) < 1e-6 for a, b in zip(self.axes, other.axes))


class FloatVector:
    """float32 array + hypercube; ``getCpp()`` returns the vector itself, which
    is what the pybind11 modules accept."""

    def __init__(self, hyper: hypercube, arr=None):
        self._hyper = hyper
        shape = hyper.shape()
        if arr is None:
            self._arr = np.zeros(shape, np.float32)
        else:
            a = np.ascontiguousarray(arr, dtype=np.float32)
            if a.size != int(np.prod(shape)):
                raise ValueError("array size does not match the hypercube")
            self._arr = a.reshape(shape)

    # accessors
    def getNdArray(self):
        return self._arr

    def getHyper(self):
        return self._hyper

    def getCpp(self):
        return self

    def getNdim(self):
        return self._hyper.getNdim()

    # vector-space operations (pyVector/SepVector names)
    def clone(self):
        return FloatVector(self._hyper, self._arr.copy())

    def cloneSpace(self):
        return FloatVector(self._hyper)

    def zero(self):
        self._arr[...] = 0.0
        return self

    def set(self, val):
        self._arr[...] = val
        return self

    def scale(self, sc):
        self._arr *= np.float32(sc)
        return self

    def scaleAdd(self, other, sc1=1.0, sc2=1.0):
        self._arr[...] = np.float32(sc1) * self._arr + np.float32(sc2) * other.getNdArray()
        return self

    def copy(self, other):
        self._arr[...] = other.getNdArray()
        return self

    def dot(self, other):
        return float(np.dot(self._arr.ravel().astype(np.float64), other.getNdArray().ravel().astype(np.float64)))

    def norm(self, N=2):
        return float(np.linalg.norm(self._arr.ravel().astype(np.float64), N))

    def rand(self, seed=None):
        rng = np.random.default_rng(seed)
        self._arr[...] = rng.uniform(-1.0, 1.0, self._arr.shape).astype(np.float32)
        return self

    def checkSame(self, other):
        return self._hyper.checkSame(other.getHyper())

    def __repr__(self):
        return f"FloatVector({self._hyper.getNs()})"


def getSepVector(hyper=None, ns=None, axes=None):
    """SepVector.getSepVector subset: from a hypercube, a list of axes or ns."""
    if hyper is None:
        hyper = hypercube(axes=axes) if axes is not None else hypercube(ns=ns)
    return FloatVector(hyper)


# ------------------------------------------------------------ parameters --
def parfile2pars(args):
    """Expand the last ``par=file`` argument into its lines
    (PY/Acoustic_iso_float.py:94-114)."""
    args = list(args)
    par_arg = None
    for a in args:
        m = re.match(r"^par=(.*)", a)
        if m:
            par_arg = a
    if par_arg:
        with open(par_arg.split("=", 1)[1]) as fid:
            lines = fid.read().splitlines()
        idx = args.index(par_arg)
        args = args[:idx] + lines + args[idx + 1:]
    return args


class ParamObj:
    """genericIO-style parameter object over ``key=value`` strings.  The last
    occurrence of a key wins; ``par=`` files are expanded in place; a missing
    key without a default raises KeyError."""

    def __init__(self, args=None, **kw):
        self._p = {}
        for a in parfile2pars(args or []):
            for tok in shlex.split(a, comments=True):
                if "=" in tok:
                    k, v = tok.split("=", 1)
                    self._p[k.strip()] = v.strip()
        for k, v in kw.items():
            self._p[k] = v if isinstance(v, str) else str(v) if not isinstance(v, (list, tuple)) else ",".join(map(str, v))

    @property
    def param(self):
        """The object handed to the pybind11 modules (``paramP.param``)."""
        return self

    def getCpp(self):
        return self

    def has(self, k):
        return k in self._p

    def _get(self, k, d, conv):
        if k in self._p:
            return conv(self._p[k])
        if d is None:
            raise KeyError(f"missing parameter {k}")
        return d

    def getInt(self, k, d=None):
        return self._get(k, d, lambda s: int(float(s)))

    def getFloat(self, k, d=None):
        return self._get(k, d, float)

    def getString(self, k, d=None):
        return self._get(k, d, str)

    def getInts(self, k, d=None):
        return self._get(k, d, lambda s: [int(float(x)) for x in s.split(",") if x != ""])

    def getFloats(self, k, d=None):
        return self._get(k, d, lambda s: [float(x) for x in s.split(",") if x != ""])

    def set(self, k, v):
        self._p[k] = str(v)


# ------------------------------------------------------------------- IO --
def _parse_header(path):
    kv = {}
    with open(path, "r", errors="replace") as fid:
        for line in fid:
            for tok in shlex.split(line, comments=True):
                if "=" in tok:
                    k, v = tok.split("=", 1)
                    kv[k] = v.strip('"')
    return kv


def read(path, ndims=None):
    """Read a SEP .H file (header + binary named by ``in=``) into a FloatVector."""
    kv = _parse_header(path)
    ns, os_, ds = [], [], []
    i = 1
    while f"n{i}" in kv:
        ns.append(int(kv[f"n{i}"]))
        os_.append(float(kv.get(f"o{i}", 0.0)))
        ds.append(float(kv.get(f"d{i}", 1.0)))
        i += 1
    if ndims is not None:
        while len(ns) < ndims:
            ns.append(1); os_.append(0.0); ds.append(1.0)
        while len(ns) > ndims and ns[-1] == 1:
            ns.pop(); os_.pop(); ds.pop()
    else:   # trailing axes of length 1 (n3=1 ... n9=1 in history blocks) are not dimensions
        while len(ns) > 1 and ns[-1] == 1:
            ns.pop(); os_.pop(); ds.pop()
    # The last in= of the history wins; a relative path is taken next to the header.
    binpath = kv.get("in", path + "@")
    if not os.path.isabs(binpath) and not os.path.exists(binpath):
        binpath = os.path.join(os.path.dirname(path), os.path.basename(binpath))
    fmt = kv.get("data_format", "xdr_float")
    if fmt not in ("xdr_float", "native_float") or int(kv.get("esize", 4)) != 4:
        raise ValueError(f"{path}: only 4-byte xdr_float / native_float SEP data are supported (data_format={fmt})")
    dt = ">f4" if fmt == "xdr_float" else "<f4"
    count = int(np.prod(ns))
    arr = np.fromfile(binpath, dtype=dt, count=count).astype(np.float32)
    if arr.size != count:
        raise ValueError(f"{binpath}: holds {arr.size} samples, the header describes {count}")
    return FloatVector(hypercube(ns=ns, os_=os_, ds=ds), arr)


def write(path, vec: FloatVector, fmt="xdr_float"):
    """Write a FloatVector as a SEP .H header plus ``path@`` binary."""
    h = vec.getHyper()
    lines = [f"n{i+1}={a.n} o{i+1}={a.o} d{i+1}={a.d}" for i, a in enumerate(h.axes)]
    binpath = path + "@"
    with open(path, "w") as fid:
        fid.write("\n".join(lines) + f"\nesize=4 data_format=\"{fmt}\"\nin=\"{binpath}\"\n")
    dt = ">f4" if fmt == "xdr_float" else "<f4"
    vec.getNdArray().astype(dt).tofile(binpath)
