"""Global settings (p4/var.py).

The numeric bounds are 1-element numpy arrays because the likelihood library
keeps pointers to them (p4/var.py:283-305, Pf/p4_model.c:34-48); the
properties read and write the arrays in place.
"""
import numpy


class Var(object):
    NO_ORDER = -10000
    GAP_CODE = -1
    QMARK_CODE = -2
    N_LIKE = -3
    EQUATES_BASE = -64

    def __init__(self):
        self._rMatrixNormalizeTo1 = numpy.array([1], numpy.int32)
        self._PIVEC_MIN = numpy.array([1.0e-13])
        self._PIVEC_MAX = numpy.array([0.999])
        self._RATE_MIN = numpy.array([1.0e-13])
        self._RATE_MAX = numpy.array([0.9999999])
        self._GAMMA_SHAPE_MIN = numpy.array([0.1])
        self._GAMMA_SHAPE_MAX = numpy.array([300.0])
        self._PINVAR_MIN = numpy.array([0.0])
        self._PINVAR_MAX = numpy.array([0.99])
        self._RELRATE_MIN = numpy.array([1.0e-8])
        self._RELRATE_MAX = numpy.array([1.0e8])
        self._KAPPA_MIN = numpy.array([0.000001])
        self._KAPPA_MAX = numpy.array([100.0])
        self._BRLEN_MIN = numpy.array([1.0e-8])
        self._BRLEN_MAX = numpy.array([3.0])
        self._newtAndBrentPowellOptPassLimit = numpy.array([50], numpy.int32)
        self.doMcmcSp = True
        self.mcmc_swapVector = True
        self.verboseRead = False
        self.rMatrixProteinSpecs = ['cpREV', 'd78', 'jtt', 'mtREV24', 'mtmam', 'wag', 'rtRev',
                                    'tmjtt94', 'tmlg99', 'lg', 'blosum62', 'hivb', 'mtart',
                                    'mtzoa', 'gcpREV', 'stmtREV', 'vt', 'pmb']
        self.rMatrixSpecs = ['ones', '2p', 'optimized', 'specified'] + self.rMatrixProteinSpecs
        self.compSpecs = ['equal', 'empirical', 'specified', 'wag']
        self.modelSymbols = list('=@#$%&*+abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ123456789')
        # Partials precision of new trees: 'fp64' (the reference's numerics)
        # or 'fp32' (FP32 partials with power-of-two scale exponents).
        self.precision = 'fp64'


def _bound(name):
    def g(self):
        return float(getattr(self, "_" + name)[0])

    def s(self, v):
        getattr(self, "_" + name)[0] = v
    return property(g, s)


for _n in ['PIVEC_MIN', 'PIVEC_MAX', 'RATE_MIN', 'RATE_MAX', 'GAMMA_SHAPE_MIN', 'GAMMA_SHAPE_MAX',
           'PINVAR_MIN', 'PINVAR_MAX', 'RELRATE_MIN', 'RELRATE_MAX', 'KAPPA_MIN', 'KAPPA_MAX',
           'BRLEN_MIN', 'BRLEN_MAX']:
    setattr(Var, _n, _bound(_n))
Var.rMatrixNormalizeTo1 = property(lambda self: int(self._rMatrixNormalizeTo1[0]),
                                   lambda self, v: self._rMatrixNormalizeTo1.__setitem__(0, int(v)))

var = Var()
