"""Minimal linear-operator base with the method names the operator classes of
``Acoustic_iso_float`` rely on (``setDomainRange``, ``getDomain``,
``getRange``, ``forward``, ``adjoint``, ``dotTest``).

The reference imports this base from pySolver (``pyOperator``), an
un-vendored dependency (.gitmodules:4-6) that is absent here; its solver-level
behaviour is not restated, only the operator contract:

* ``forward(add, model, data)``: data = (add ? data : 0) + A model;
* ``adjoint(add, model, data)``: model = (add ? model : 0) + A^T data;
* ``dotTest``: <A x, y> against <x, A^T y> for random x, y, with add=False and
  with add=True, relative to the larger magnitude (the criterion of the
  reference's C++ Operator::dotTest, F/include/operator.h:50-88).
"""
from __future__ import annotations


class Operator:
    def setDomainRange(self, domain, range):
        self.domain = domain.cloneSpace()
        self.range = range.cloneSpace()

    def getDomain(self):
        return self.domain

    def getRange(self):
        return self.range

    def forward(self, add, model, data):
        raise NotImplementedError("forward must be defined by the operator")

    def adjoint(self, add, model, data):
        raise NotImplementedError("adjoint must be defined by the operator")

    def __str__(self):
        return "Operator"

    def dotTest(self, verb=False, maxError=1e-4, seed=0):
        """True when both relative errors are below maxError."""
        x = self.domain.cloneSpace().rand(seed)
        y = self.range.cloneSpace().rand(seed + 1)
        errs = []
        for add in (False, True):
            d = self.range.cloneSpace().rand(seed + 2) if add else self.range.cloneSpace()
            m = self.domain.cloneSpace().rand(seed + 3) if add else self.domain.cloneSpace()
            d0, m0 = d.clone(), m.clone()
            self.forward(add, x, d)
            self.adjoint(add, m, y)
            if add:
                d.scaleAdd(d0, 1.0, -1.0)
                m.scaleAdd(m0, 1.0, -1.0)
            dy, xm = d.dot(y), x.dot(m)
            err = abs(dy - xm) / max(abs(dy), abs(xm), 1e-300)
            errs.append(err)
            if verb:
                print(f"dot test (add={add}): <Ax,y>={dy:.9e} <x,A'y>={xm:.9e} rel.err={err:.3e}")
        return max(errs) <= maxError
