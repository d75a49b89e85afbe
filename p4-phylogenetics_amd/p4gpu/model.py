"""Model components and their C-side mirror (p4/model.py, and the Comp /
RMatrix / Gdasrv / PInvar classes of p4/tree.py:80-220).

Comp values, gamma val/freqs/rates and bQETneedsReset are numpy arrays whose
pointers the likelihood library keeps; they are always mutated in place.
"""
import math

import numpy

from . import pf
from .p4exceptions import P4Error
from .var import var


class Comp(object):
    def __init__(self):
        self.num = -1
        self.partNum = None
        self.free = None
        self.spec = None
        self.symbol = None
        self.val = None      # numpy float64 [dim]
        self.nNodes = 0

    def __repr__(self):
        return "<Comp part %s num %s %s>" % (self.partNum, self.num, self.val)


class RMatrix(object):
    def __init__(self):
        self.num = -1
        self.partNum = None
        self.free = None
        self.spec = None
        self.symbol = None
        self.val = None      # numpy float64 [dim*(dim-1)/2], or [kappa] for '2p'
        self.nNodes = 0


class Gdasrv(object):
    """Discrete gamma among-site rate variation (p4/tree.py:80-128)."""

    def __init__(self):
        self.num = -1
        self.partNum = None
        self.free = None
        self.symbol = None
        self._val = numpy.zeros(1, dtype=numpy.double)
        self.freqs = None
        self.rates = None
        self.nGammaCat = None
        self.c = None
        self.nNodes = 0

    def _getVal(self):
        return self._val

    def _setVal(self, theVal):
        if theVal < 1.e-16:
            raise P4Error(["Gdasrv._setVal()", "Attempt to set Gdasrv.val (ie alpha) to %g" % theVal])
        self._val[0] = theVal
        self.calcRates()

    val = property(_getVal, _setVal)

    def calcRates(self):
        if self.c:
            pf.gdasrvCalcRates(self.c)
        else:
            pf.gdasrvCalcRates_np(self.nGammaCat, self._val[0], self.freqs, self.rates)


class PInvar(object):
    def __init__(self):
        self.partNum = None
        self.free = 0
        self.val = 0.0


class ModelPart(object):
    def __init__(self):
        self.num = -1
        self.dim = 0
        self.comps = []
        self.rMatrices = []
        self.gdasrvs = []
        self.pInvar = None
        self.relRate = 1.0
        self.nGammaCat = 1
        self.isHet = 0
        self.bQETneedsReset = None
        self.ndch2 = False
        self.ndrh2 = False

    @property
    def nComps(self):
        return len(self.comps)

    @property
    def nRMatrices(self):
        return len(self.rMatrices)

    @property
    def nGdasrvs(self):
        return len(self.gdasrvs)

    def setCStuff(self, theModel):
        """Poke R values, kappa, pInvar and relRate into C
        (p4/model.py:153-207).  Comps are aliased, not poked."""
        for mt in self.comps:
            if mt.val is None:
                raise P4Error("ModelPart.setCStuff(): comp %i of part %i has no val" % (mt.num, self.num))
            assert numpy.min(mt.val) >= var.PIVEC_MIN
        for mNum, mt in enumerate(self.rMatrices):
            if mt.spec == '2p':
                pf.p4_setKappa(theModel.cModel, self.num, mNum, mt.val[0])
            elif mt.free or mt.spec == 'specified':
                k = 0
                if var.rMatrixNormalizeTo1:
                    for i in range(self.dim - 1):
                        for j in range(i + 1, self.dim):
                            pf.p4_setRMatrixBigR(theModel.cModel, self.num, mNum, i, j, mt.val[k])
                            k += 1
                else:
                    for i in range(self.dim - 2):
                        for j in range(i + 1, self.dim):
                            pf.p4_setRMatrixBigR(theModel.cModel, self.num, mNum, i, j, mt.val[k])
                            k += 1
        pf.p4_setPInvarVal(theModel.cModel, self.num, self.pInvar.val)
        pf.p4_setRelRateVal(theModel.cModel, self.num, self.relRate)

    def copyValsTo(self, other):
        """In-place copy of the free values into another ModelPart
        (p4/model.py:209-268)."""
        for a, b in zip(self.comps, other.comps):
            b.val[:] = a.val
        for a, b in zip(self.rMatrices, other.rMatrices):
            if a.val is not None:
                b.val[:] = a.val
        for a, b in zip(self.gdasrvs, other.gdasrvs):
            b._val[0] = a._val[0]
            b.freqs[:] = a.freqs
            b.rates[:] = a.rates
        other.pInvar.val = self.pInvar.val
        other.relRate = self.relRate

    def copyBQETneedsResetTo(self, other):
        other.bQETneedsReset[:] = self.bQETneedsReset


class Model(object):
    def __init__(self, nParts):
        self.parts = []
        for i in range(nParts):
            mp = ModelPart()
            mp.num = i
            self.parts.append(mp)
        self.cModel = None
        self.doRelRates = 0
        self.relRatesAreFree = 0
        self.nFreePrams = 0
        self.isHet = 0

    @property
    def nParts(self):
        return len(self.parts)

    def __del__(self):
        if getattr(self, "cModel", None):
            try:
                pf.p4_freeModel(self.cModel)
            except Exception:
                pass
            self.cModel = None

    def allocCStuff(self):
        """p4/model.py:735-824."""
        if self.cModel:
            raise P4Error("Model.allocCStuff(): cModel already exists")
        self.cModel = pf.p4_newModel(self.nParts, self.doRelRates, self.relRatesAreFree,
                                     int(self.nFreePrams), self.isHet, var._rMatrixNormalizeTo1,
                                     var._PINVAR_MIN, var._PINVAR_MAX, var._KAPPA_MIN,
                                     var._KAPPA_MAX, var._GAMMA_SHAPE_MIN, var._GAMMA_SHAPE_MAX,
                                     var._PIVEC_MIN, var._PIVEC_MAX, var._RATE_MIN,
                                     var._RATE_MAX, var._RELRATE_MIN, var._RELRATE_MAX,
                                     var._BRLEN_MIN, var._BRLEN_MAX)
        for pNum, mp in enumerate(self.parts):
            mp.bQETneedsReset = numpy.ones((mp.nComps, mp.nRMatrices), numpy.int32)
            pf.p4_newModelPart(self.cModel, pNum, mp.dim, mp.nComps, mp.nRMatrices,
                               mp.nGdasrvs, mp.nGammaCat, mp.pInvar.free, mp.bQETneedsReset)
            for mNum, mt in enumerate(mp.comps):
                assert mt.val is not None
                pf.p4_newComp(self.cModel, pNum, mNum, mt.free, mt.val)
            for mNum, mt in enumerate(mp.rMatrices):
                if mt.spec == 'ones':
                    spec = 100
                elif mt.spec in ('specified', 'optimized'):
                    spec = 20
                elif mt.spec == '2p':
                    spec = 5
                else:
                    raise P4Error("rMatrix spec '%s' is not supported by this engine" % mt.spec)
                pf.p4_newRMatrix(self.cModel, pNum, mNum, mt.free, spec)
            for mNum, mt in enumerate(mp.gdasrvs):
                mt.c = pf.p4_newGdasrv(self.cModel, pNum, mNum, mt.nGammaCat, mt.free, mt._val,
                                       mt.freqs, mt.rates)
                mt.calcRates()

    def setCStuff(self, partNum=None):
        if partNum is None:
            for mp in self.parts:
                mp.setCStuff(self)
        else:
            self.parts[partNum].setCStuff(self)

    def copyValsTo(self, other):
        for a, b in zip(self.parts, other.parts):
            a.copyValsTo(b)

    def copyBQETneedsResetTo(self, other):
        for a, b in zip(self.parts, other.parts):
            a.copyBQETneedsResetTo(b)
