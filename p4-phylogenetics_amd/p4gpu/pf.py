"""ctypes shim over libp4gpu.so with the names and argument order of p4's
C extension ``p4.pf`` (Pf/pfmodule.c, method table 2873-3031), for the
likelihood / MCMC subset.

Handles are returned as Python ints, as the reference does.  Numpy arrays
that the reference aliases (comp values, gamma val/freqs/rates,
bQETneedsReset, partLikes, pre/postOrder, the var bounds) are passed by
pointer and must stay alive and be mutated in place by the caller.

Where the reference prints and calls exit(1), these functions raise
P4Error with the library's message.  There is no CPU fallback: if the
library or a GPU is missing, evaluation calls raise.
"""
import ctypes
import os

import numpy

from .p4exceptions import P4Error

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libp4gpu.so")

_lib = None

c_int = ctypes.c_int
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
P_int = ctypes.POINTER(ctypes.c_int)
P_double = ctypes.POINTER(ctypes.c_double)


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


def lib():
    """Load libp4gpu.so (once).  Raises P4Error if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise P4Error(["libp4gpu.so is not built.",
                       "Run: python p4-phylogenetics_amd/build.py"])
    _lib = ctypes.CDLL(LIB_PATH)
    _declare()
    _applyTuning()
    return _lib


TUNING_PATH = os.path.join(_HERE, "tuning.json")


def _applyTuning():
    """Apply the autotuner's choices (p4gpu/tuning.json) if present."""
    import json
    if not os.path.exists(TUNING_PATH):
        return
    try:
        with open(TUNING_PATH) as fh:
            tun = json.load(fh)
        for key, fp32 in (("fp64", 0), ("fp32", 1)):
            if key in tun:
                _lib.p4g_setPruneTuning(fp32, int(tun[key]["opsPerGroup"]), int(tun[key]["stagePInLds"]))
    except (OSError, ValueError, KeyError):
        pass


def _declare():
    V, I, D = c_void_p, c_int, c_double
    _sig("p4g_lastError", c_char_p)
    _sig("p4g_deviceAvailable", I)
    _sig("p4g_setKernelTiming", None, I)
    _sig("p4g_setPruneTuning", None, I, I, I)
    _sig("p4g_setPmatRoute", None, I)
    _sig("p4g_setAaTuning", None, I, I)
    _sig("p4g_setWavePreloadLds", None, I)
    _sig("p4g_setAaLdsSlots", None, I)
    _sig("p4g_lastKernelMicroseconds", D, V)
    _sig("p4g_newData", V, I, I)
    _sig("p4g_freeData", None, V)
    _sig("p4g_pokePartInData", I, V, V, I)
    _sig("p4g_newPart", V, I, I, c_char_p, I, c_char_p, I)
    _sig("p4g_freePart", None, V)
    _sig("p4g_pokeEquatesTable", I, V, c_char_p)
    _sig("p4g_pokeSequences", I, V, c_char_p)
    _sig("p4g_makePatterns", I, V)
    _sig("p4g_makePatternsDevice", I, V)
    _sig("p4g_setGlobalInvarSitesVec", I, V)
    _sig("p4g_partNPatterns", I, V)
    _sig("p4g_partNChar", I, V)
    _sig("p4g_partNTax", I, V)
    _sig("p4g_partDim", I, V)
    _sig("p4g_partGetPatterns", I, V, V, V, V, V, V)
    _sig("p4g_partSetPatterns", I, V, I, V, V)
    _sig("p4g_getSiteLikes", I, V, V)
    _sig("p4g_newModel", V, I, I, I, I, I, V, V, V, V, V, V, V, V, V, V, V, V, V, V, V)
    _sig("p4g_freeModel", None, V)
    _sig("p4g_newModelPart", I, V, I, I, I, I, I, I, I, V)
    _sig("p4g_newComp", I, V, I, I, I, V)
    _sig("p4g_newRMatrix", I, V, I, I, I, I)
    _sig("p4g_newGdasrv", V, V, I, I, I, I, V, V, V)
    _sig("p4g_setRMatrixBigR", I, V, I, I, I, I, D)
    _sig("p4g_setKappa", I, V, I, I, D)
    _sig("p4g_gdasrvCalcRates", I, V)
    _sig("p4g_gdasrvCalcRates_np", I, I, D, V, V)
    _sig("p4g_setPInvarVal", I, V, I, D)
    _sig("p4g_setRelRateVal", I, V, I, D)
    _sig("p4g_resetBQET", I, V, I, I, I)
    _sig("p4g_newTree", V, I, I, V, V, V, V, V)
    _sig("p4g_freeTree", None, V)
    _sig("p4g_newNode", V, I, V, I, I, I)
    _sig("p4g_freeNode", None, V)
    _sig("p4g_setNodeRelation", I, V, I, I)
    _sig("p4g_setTreeRoot", I, V, V)
    _sig("p4g_setBrLen", I, V, D)
    _sig("p4g_getBrLen", D, V)
    _sig("p4g_getTreeLen", D, V)
    _sig("p4g_setCompNum", I, V, I, I)
    _sig("p4g_setRMatrixNum", I, V, I, I)
    _sig("p4g_setGdasrvNum", I, V, I, I)
    _sig("p4g_setPrecision", I, V, I)
    _sig("p4g_getPrecision", I, V)
    _sig("p4g_setPrams", I, V, I)
    _sig("p4g_treeLogLike", I, V, I, P_double)
    _sig("p4g_treeLogLikeNoAlloc", I, V, P_double)
    _sig("p4g_partLogLike", I, V, V, I, I, P_double)
    _sig("p4g_calculateBigPDecks", I, V)
    _sig("p4g_calculateBigPDecksPart", I, V, I)
    _sig("p4g_calculateAllBigPDecksAllParts", I, V)
    _sig("p4g_setConditionalLikelihoodsOfInternalNodePart", I, V, I)
    _sig("p4g_setConditionalLikelihoodsOfInternalNode", I, V)
    _sig("p4g_copyCondLikes", I, V, V, I)
    _sig("p4g_copyBigPDecks", I, V, V, I)
    _sig("p4g_copyModelPrams", I, V, V)
    _sig("p4g_verifyIdentityOfTwoTrees", I, V, V)
    _sig("p4g_getBigPDeck", I, V, I, V)
    _sig("p4g_getCondLikes", I, V, I, V)
    _sig("p4g_branchDerivatives", I, V, I, V, V, V)
    _sig("p4g_getCondLikes2", I, V, I, V)
    _sig("p4g_getBigPDeck1stD", I, V, I, V)
    _sig("p4g_getBigPDeck2ndD", I, V, I, V)
    _sig("p4g_markPath", I, V, I, I)
    _sig("p4g_markAll", I, V, I)
    _sig("p4g_syncChanged", I, V, V)
    _sig("p4g_cpoAccumulate", I, V, V)
    _sig("p4g_cpoGetSums", I, V, V, V)
    _sig("p4g_cpoReset", I, V)
    _sig("p4g_partNSites", I, V)
    _sig("p4g_drawAncStates", I, V, I, I, I, V, ctypes.c_ulonglong, V)
    _sig("p4g_philox4x64", None, V, V, V)
    _sig("p4g_simulate", I, V, V, ctypes.c_ulonglong)
    _sig("p4g_symbolSequences", I, V, c_char_p)
    _sig("p4g_partSimCats", I, V, V)
    _sig("p4g_partSeqStats", I, V, V, V, V, V)
    _sig("p4g_partBigXSquared", D, V)
    _sig("p4g_getUnconstrainedLogLike", D, V)
    _sig("p4g_newBatch", V, I, V)
    _sig("p4g_mlpNew", V, I, I, I, ctypes.c_ulonglong)
    _sig("p4g_mlpFree", None, V)
    _sig("p4g_mlpNumParams", I, V)
    _sig("p4g_mlpGetParams", I, V, V)
    _sig("p4g_mlpSetParams", I, V, V)
    _sig("p4g_mlpForward", I, V, V, I, V)
    _sig("p4g_mlpGradients", I, V, V, V, V, I, V, V)
    _sig("p4g_mlpTrainStep", I, V, V, V, V, I, ctypes.c_float, V)
    _sig("p4g_freeBatch", None, V)
    _sig("p4g_batchLogLike", I, V, V)


def _err(what):
    msg = lib().p4g_lastError().decode(errors="replace")
    raise P4Error([what, msg or "(no message)"])


def _ok(status, what):
    if status != 0:
        _err(what)


def _ptr(a, dtype):
    if not isinstance(a, numpy.ndarray) or a.dtype != dtype or not a.flags["C_CONTIGUOUS"]:
        raise P4Error(f"expected a C-contiguous numpy {numpy.dtype(dtype).name} array, got {type(a)}")
    return a.ctypes.data


def _dptr(a):
    return _ptr(a, numpy.float64)


def _iptr(a):
    return _ptr(a, numpy.int32)


def _h(handle, what):
    if not handle:
        _err(what)
    return handle


# ------------------------------------------------------------------ status
def deviceAvailable():
    return bool(lib().p4g_deviceAvailable())


def setKernelTiming(on):
    lib().p4g_setKernelTiming(1 if on else 0)


def setPruneTuning(fp32, opsPerGroup, stagePInLds):
    lib().p4g_setPruneTuning(1 if fp32 else 0, int(opsPerGroup), 1 if stagePInLds else 0)


def setPmatRoute(direct):
    """4-state P matrices straight from exp(Q t) (True, the default) or from
    the eigensystem (False)."""
    lib().p4g_setPmatRoute(1 if direct else 0)


AA_CAT_MIN_OPS = 32


def setWavePreloadLds(nBytes):
    """Cap (bytes) on the 4-state kernel's LDS when a dirty-path program
    preloads clean children; 0: no cap below the slot budget."""
    lib().p4g_setWavePreloadLds(int(nBytes))


def setAaLdsSlots(maxSlots):
    """LDS slots the 20-state category kernel may forward children through
    (default 5; 0: every child from HBM)."""
    lib().p4g_setAaLdsSlots(int(maxSlots))


def setAaTuning(categoryWaves, minOps=AA_CAT_MIN_OPS):
    """20-state FP32 pruning with one wavefront per (tile, category) for
    programs of at least minOps node ops (categoryWaves True, the default),
    else one wavefront per tile."""
    lib().p4g_setAaTuning(1 if categoryWaves else 0, int(minOps))


def lastKernelMicroseconds(cTree):
    return lib().p4g_lastKernelMicroseconds(cTree)


# -------------------------------------------------------------------- data
def newData(nTax, nParts):
    return _h(lib().p4g_newData(nTax, nParts), "newData")


def freeData(cData):
    lib().p4g_freeData(cData)


def pokePartInData(cPart, cData, partNum):
    _ok(lib().p4g_pokePartInData(cPart, cData, partNum), "pokePartInData")


def newPart(nTax, nChar, equateSymbols, nEquates, symbols, dim):
    return _h(lib().p4g_newPart(nTax, nChar, equateSymbols.encode(), nEquates,
                                symbols.encode(), dim), "newPart")


def freePart(cPart):
    lib().p4g_freePart(cPart)


def pokeEquatesTable(cPart, theString):
    _ok(lib().p4g_pokeEquatesTable(cPart, theString.encode()), "pokeEquatesTable")


def pokeSequences(cPart, theString):
    _ok(lib().p4g_pokeSequences(cPart, theString.encode()), "pokeSequences")


def makePatternsDevice(cPart):
    """makePatterns on the GPU: same patterns, order, counts and site map."""
    _ok(lib().p4g_makePatternsDevice(cPart), "makePatternsDevice")


def makePatterns(cPart):
    _ok(lib().p4g_makePatterns(cPart), "makePatterns")


def setGlobalInvarSitesVec(cPart):
    _ok(lib().p4g_setGlobalInvarSitesVec(cPart), "setGlobalInvarSitesVec")


def partPatternCount(cPart):
    return lib().p4g_partNPatterns(cPart)


def partGetPatterns(cPart, nTax):
    """(patterns[nTax][nPat], counts[nPat], seqPosPatternIndex[nChar],
    invarVec[nPat], invarArray[dim][nPat]) as numpy int32 arrays."""
    L = lib()
    nP, nC, dim = L.p4g_partNPatterns(cPart), L.p4g_partNChar(cPart), L.p4g_partDim(cPart)
    pats = numpy.zeros((nTax, nP), numpy.int32)
    cnts = numpy.zeros(nP, numpy.int32)
    spi = numpy.zeros(nC, numpy.int32)
    iv = numpy.zeros(nP, numpy.int32)
    ia = numpy.zeros((dim, nP), numpy.int32)
    _ok(L.p4g_partGetPatterns(cPart, pats.ctypes.data, cnts.ctypes.data, spi.ctypes.data,
                              iv.ctypes.data, ia.ctypes.data), "partGetPatterns")
    return pats, cnts, spi, iv, ia


def partSetPatterns(cPart, patterns, counts):
    patterns = numpy.ascontiguousarray(patterns, dtype=numpy.int32)
    counts = numpy.ascontiguousarray(counts, dtype=numpy.int32)
    _ok(lib().p4g_partSetPatterns(cPart, int(counts.shape[0]), patterns.ctypes.data,
                                  counts.ctypes.data), "partSetPatterns")


def getSiteLikes(cPart):
    n = lib().p4g_partNChar(cPart)
    out = numpy.zeros(n, numpy.float64)
    _ok(lib().p4g_getSiteLikes(cPart, out.ctypes.data), "getSiteLikes")
    return list(out)


# ------------------------------------------------------------------- model
def p4_newModel(nParts, doRelRates, relRatesAreFree, nFreePrams, isHet, rMatrixNormalizeTo1,
                PINVAR_MIN, PINVAR_MAX, KAPPA_MIN, KAPPA_MAX, GAMMA_SHAPE_MIN, GAMMA_SHAPE_MAX,
                PIVEC_MIN, PIVEC_MAX, RATE_MIN, RATE_MAX, RELRATE_MIN, RELRATE_MAX, BRLEN_MIN,
                BRLEN_MAX):
    return _h(lib().p4g_newModel(
        nParts, doRelRates, relRatesAreFree, int(nFreePrams), isHet, _iptr(rMatrixNormalizeTo1),
        _dptr(PINVAR_MIN), _dptr(PINVAR_MAX), _dptr(KAPPA_MIN), _dptr(KAPPA_MAX),
        _dptr(GAMMA_SHAPE_MIN), _dptr(GAMMA_SHAPE_MAX), _dptr(PIVEC_MIN), _dptr(PIVEC_MAX),
        _dptr(RATE_MIN), _dptr(RATE_MAX), _dptr(RELRATE_MIN), _dptr(RELRATE_MAX),
        _dptr(BRLEN_MIN), _dptr(BRLEN_MAX)), "p4_newModel")


def p4_freeModel(cModel):
    lib().p4g_freeModel(cModel)


def p4_newModelPart(cModel, pNum, dim, nComps, nRMatrices, nGdasrvs, nCat, pInvarFree,
                    bQETneedsReset):
    _ok(lib().p4g_newModelPart(cModel, pNum, dim, nComps, nRMatrices, nGdasrvs, nCat,
                               pInvarFree, _iptr(bQETneedsReset)), "p4_newModelPart")


def p4_newComp(cModel, pNum, mNum, free, val):
    _ok(lib().p4g_newComp(cModel, pNum, mNum, free, _dptr(val)), "p4_newComp")


def p4_newRMatrix(cModel, pNum, mNum, free, spec):
    _ok(lib().p4g_newRMatrix(cModel, pNum, mNum, free, spec), "p4_newRMatrix")


def p4_newGdasrv(cModel, pNum, mNum, nCat, free, val, freqs, rates):
    return _h(lib().p4g_newGdasrv(cModel, pNum, mNum, nCat, free, _dptr(val), _dptr(freqs),
                                  _dptr(rates)), "p4_newGdasrv")


def p4_setRMatrixBigR(cModel, pNum, mNum, i, j, val):
    _ok(lib().p4g_setRMatrixBigR(cModel, pNum, mNum, i, j, float(val)), "p4_setRMatrixBigR")


def p4_setKappa(cModel, pNum, mNum, val):
    _ok(lib().p4g_setKappa(cModel, pNum, mNum, float(val)), "p4_setKappa")


def gdasrvCalcRates(cGdasrv):
    _ok(lib().p4g_gdasrvCalcRates(cGdasrv), "gdasrvCalcRates")


def gdasrvCalcRates_np(nCat, val, freqs, rates):
    _ok(lib().p4g_gdasrvCalcRates_np(nCat, float(val), _dptr(freqs), _dptr(rates)),
        "gdasrvCalcRates_np")


def p4_setPInvarVal(cModel, pNum, val):
    _ok(lib().p4g_setPInvarVal(cModel, pNum, float(val)), "p4_setPInvarVal")


def p4_setRelRateVal(cModel, pNum, val):
    _ok(lib().p4g_setRelRateVal(cModel, pNum, float(val)), "p4_setRelRateVal")


def p4_resetBQET(cModel, pNum, compNum, rMatrixNum):
    _ok(lib().p4g_resetBQET(cModel, pNum, compNum, rMatrixNum), "p4_resetBQET")


# -------------------------------------------------------------------- tree
def p4_newTree(nNodes, nLeaves, preOrder, postOrder, passLimit, partLikes, cData, cModel):
    """Same argument list as the reference (Pf/pfmodule.c:1383-1406);
    passLimit (newtAndBrentPowellOptPassLimit) is accepted and unused."""
    return _h(lib().p4g_newTree(nNodes, nLeaves, _iptr(preOrder), _iptr(postOrder),
                                _dptr(partLikes), cData, cModel), "p4_newTree")


def p4_freeTree(cTree):
    lib().p4g_freeTree(cTree)


def p4_newNode(nodeNum, cTree, seqNum, isLeaf, inTree):
    return _h(lib().p4g_newNode(nodeNum, cTree, seqNum, isLeaf, inTree), "p4_newNode")


def p4_freeNode(cNode):
    lib().p4g_freeNode(cNode)


def p4_setNodeRelation(cNode, relation, relNum):
    _ok(lib().p4g_setNodeRelation(cNode, relation, relNum), "p4_setNodeRelation")


def p4_setTreeRoot(cTree, cNode):
    _ok(lib().p4g_setTreeRoot(cTree, cNode), "p4_setTreeRoot")


def p4_setBrLen(cNode, brLen):
    _ok(lib().p4g_setBrLen(cNode, float(brLen)), "p4_setBrLen")


def p4_getBrLen(cNode):
    return lib().p4g_getBrLen(cNode)


def p4_getTreeLen(cTree):
    return lib().p4g_getTreeLen(cTree)


def p4_setCompNum(cNode, pNum, val):
    _ok(lib().p4g_setCompNum(cNode, pNum, val), "p4_setCompNum")


def p4_setRMatrixNum(cNode, pNum, val):
    _ok(lib().p4g_setRMatrixNum(cNode, pNum, val), "p4_setRMatrixNum")


def p4_setGdasrvNum(cNode, pNum, val):
    _ok(lib().p4g_setGdasrvNum(cNode, pNum, val), "p4_setGdasrvNum")


def p4_setPrecision(cTree, fp32):
    _ok(lib().p4g_setPrecision(cTree, 1 if fp32 else 0), "p4_setPrecision")


# -------------------------------------------------------------- evaluation
def p4_setPrams(cTree, pNum):
    _ok(lib().p4g_setPrams(cTree, pNum), "p4_setPrams")


def p4_treeLogLike(cTree, getSiteLikes):
    out = c_double(0.0)
    _ok(lib().p4g_treeLogLike(cTree, getSiteLikes, ctypes.byref(out)), "p4_treeLogLike")
    return out.value


def treeLogLikeNoAlloc(cTree):
    """p4_treeLogLike(cTree, 0) with no device allocation (the tree must
    have been evaluated once); no counterpart in p4."""
    out = c_double(0.0)
    _ok(lib().p4g_treeLogLikeNoAlloc(cTree, ctypes.byref(out)), "treeLogLikeNoAlloc")
    return out.value


def p4_partLogLike(cTree, cPart, pNum, getSiteLikes):
    out = c_double(0.0)
    _ok(lib().p4g_partLogLike(cTree, cPart, pNum, getSiteLikes, ctypes.byref(out)),
        "p4_partLogLike")
    return out.value


def p4_calculateBigPDecks(cNode):
    _ok(lib().p4g_calculateBigPDecks(cNode), "p4_calculateBigPDecks")


def p4_calculateBigPDecksPart(cNode, pNum):
    _ok(lib().p4g_calculateBigPDecksPart(cNode, pNum), "p4_calculateBigPDecksPart")


def p4_calculateAllBigPDecksAllParts(cTree):
    _ok(lib().p4g_calculateAllBigPDecksAllParts(cTree), "p4_calculateAllBigPDecksAllParts")


def p4_setConditionalLikelihoodsOfInternalNodePart(cNode, pNum):
    _ok(lib().p4g_setConditionalLikelihoodsOfInternalNodePart(cNode, pNum),
        "p4_setConditionalLikelihoodsOfInternalNodePart")


def p4_setConditionalLikelihoodsOfInternalNode(cNode):
    _ok(lib().p4g_setConditionalLikelihoodsOfInternalNode(cNode),
        "p4_setConditionalLikelihoodsOfInternalNode")


def p4_copyCondLikes(cTreeA, cTreeB, doAll):
    _ok(lib().p4g_copyCondLikes(cTreeA, cTreeB, doAll), "p4_copyCondLikes")


def p4_copyBigPDecks(cTreeA, cTreeB, doAll):
    _ok(lib().p4g_copyBigPDecks(cTreeA, cTreeB, doAll), "p4_copyBigPDecks")


def p4_copyModelPrams(cTreeA, cTreeB):
    _ok(lib().p4g_copyModelPrams(cTreeA, cTreeB), "p4_copyModelPrams")


def p4_verifyIdentityOfTwoTrees(cTreeA, cTreeB):
    return lib().p4g_verifyIdentityOfTwoTrees(cTreeA, cTreeB)


def p4_getBigPDeck(cNode, pNum, nCat, dim):
    out = numpy.zeros((nCat, dim, dim), numpy.float64)
    _ok(lib().p4g_getBigPDeck(cNode, pNum, out.ctypes.data), "p4_getBigPDeck")
    return out


def branchDerivatives(cTree, pNum, nNodes):
    """(lnLEdge, d1, d2), arrays indexed by node number: the log likelihood
    across each non-root node's branch and its first and second derivatives
    with respect to that branch length (part pNum, or all parts for -1).
    What Pf's Newton optimiser computes one node at a time
    (Pf/p4_treeNewt.c:210-596); FP64 trees only."""
    lnl = numpy.zeros(nNodes)
    d1 = numpy.zeros(nNodes)
    d2 = numpy.zeros(nNodes)
    _ok(lib().p4g_branchDerivatives(cTree, pNum, lnl.ctypes.data, d1.ctypes.data, d2.ctypes.data),
        "branchDerivatives")
    return lnl, d1, d2


def p4_getCondLikes2(cNode, pNum, nCat, dim, nPatterns):
    out = numpy.zeros((nCat, dim, nPatterns), numpy.float64)
    _ok(lib().p4g_getCondLikes2(cNode, pNum, out.ctypes.data), "p4_getCondLikes2")
    return out


def p4_getBigPDeck1stD(cNode, pNum, nCat, dim):
    out = numpy.zeros((nCat, dim, dim), numpy.float64)
    _ok(lib().p4g_getBigPDeck1stD(cNode, pNum, out.ctypes.data), "p4_getBigPDeck1stD")
    return out


def p4_getBigPDeck2ndD(cNode, pNum, nCat, dim):
    out = numpy.zeros((nCat, dim, dim), numpy.float64)
    _ok(lib().p4g_getBigPDeck2ndD(cNode, pNum, out.ctypes.data), "p4_getBigPDeck2ndD")
    return out


# ---------------------------------------------------------- MCMC state sync
def markPath(cTree, nodeNum, pNum=-1):
    """Node nodeNum and its ancestors marked for recomputation in part pNum
    (or every part, -1), in one call."""
    _ok(lib().p4g_markPath(cTree, int(nodeNum), int(pNum)), "markPath")


def markAll(cTree, pNum=-1):
    _ok(lib().p4g_markAll(cTree, int(pNum)), "markAll")


def syncChanged(cTreeSrc, cTreeDst):
    """dst takes src's partials and P matrices where either changed since
    their last sync (one copy launch for all parts)."""
    _ok(lib().p4g_syncChanged(cTreeSrc, cTreeDst), "syncChanged")


# ------------------------------------------------------------- CPO / LPML
def cpoAccumulate(cTree):
    """One CPO sample of tree cTree (p4/mcmc.py:3286-3297 without the host
    round trip): every part's per-pattern sum of 1/like grows in GPU
    memory.  Returns the tree's log likelihood."""
    out = ctypes.c_double()
    _ok(lib().p4g_cpoAccumulate(cTree, ctypes.byref(out)), "cpoAccumulate")
    return out.value


def cpoGetSums(cPart):
    """(per-site sums of 1/siteLike, number of samples) of one part."""
    n = lib().p4g_partNSites(cPart)
    out = numpy.zeros(max(n, 0), numpy.float64)
    ns = ctypes.c_int()
    _ok(lib().p4g_cpoGetSums(cPart, out.ctypes.data, ctypes.byref(ns)), "cpoGetSums")
    return out, ns.value


def cpoReset(cPart):
    _ok(lib().p4g_cpoReset(cPart), "cpoReset")


def partNSites(cPart):
    return lib().p4g_partNSites(cPart)


# ------------------------------------------------------- ancestral states
def drawAncStates(cTree, pNum, first, n, u=None, seed=0):
    """Root ancestral-state draws for sites [first, first + n) of part pNum:
    an int32 array [n][4] of (chStNum, catNum, isInvar, invarChNum) as
    pf.p4_drawAncState fills for one site.  u: n uniforms in [0, 1), or
    None for the GPU's Philox stream keyed by (seed, site)."""
    draw = numpy.zeros((n, 4), numpy.int32)
    if u is not None:
        u = numpy.ascontiguousarray(u, numpy.float64)
        if u.shape != (n,):
            raise P4Error("drawAncStates: u must hold n values")
    _ok(lib().p4g_drawAncStates(cTree, pNum, first, n, None if u is None else u.ctypes.data,
                                ctypes.c_ulonglong(seed & 0xFFFFFFFFFFFFFFFF), draw.ctypes.data),
        "drawAncStates")
    return draw


def philox4x64(counter, key):
    """Philox4x64-10 (the GPU draws' generator) of counter[4], key[2]."""
    c = numpy.ascontiguousarray(counter, numpy.uint64)
    k = numpy.ascontiguousarray(key, numpy.uint64)
    out = numpy.zeros(4, numpy.uint64)
    lib().p4g_philox4x64(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def p4_drawAncState(cTree, partNum, seqPos, draw):
    """pf.p4_drawAncState (Pf/pfmodule.c:2353-2367): one site's draw into
    the int32 array draw[4], with a fresh 64-bit seed from Python's
    random module."""
    import random
    d = drawAncStates(cTree, partNum, seqPos, 1, seed=random.getrandbits(64))
    draw[:4] = d[0]


# --------------------------------------------------------------- simulation
def p4_simulate(cTree, refCTree, seed):
    """pf.p4_simulate (Pf/p4_treeSim.c:14-590) on the GPU: new sequences
    for every part of the tree's data; seed keys the Philox stream."""
    _ok(lib().p4g_simulate(cTree, refCTree, ctypes.c_ulonglong(seed & 0xFFFFFFFFFFFFFFFF)), "p4_simulate")


def partNTax(cPart):
    return lib().p4g_partNTax(cPart)


def symbolSequences(cPart):
    L = lib()
    n = L.p4g_partNChar(cPart) * partNTax(cPart)
    buf = ctypes.create_string_buffer(n + 1)
    _ok(L.p4g_symbolSequences(cPart, buf), "symbolSequences")
    return buf.raw[:n].decode()


def partSimCats(cPart):
    """Per site of the last simulation: gamma category, -1 for invariant."""
    out = numpy.zeros(lib().p4g_partNChar(cPart), numpy.int32)
    _ok(lib().p4g_partSimCats(cPart, out.ctypes.data), "partSimCats")
    return out


def partSeqStats(cPart):
    """(compCounts [nTax][dim], nConstant, meanNChars, nNonState)."""
    L = lib()
    counts = numpy.zeros((partNTax(cPart), L.p4g_partDim(cPart)), numpy.int32)
    nc, other = ctypes.c_int(), ctypes.c_int()
    mean = ctypes.c_double()
    _ok(L.p4g_partSeqStats(cPart, counts.ctypes.data, ctypes.byref(nc), ctypes.byref(mean), ctypes.byref(other)),
        "partSeqStats")
    return counts, nc.value, mean.value, other.value


def partBigXSquared(cPart):
    v = lib().p4g_partBigXSquared(cPart)
    if v == -3.0:
        _err("partBigXSquared")
    return v


def partSimpleConstantSitesCount(cPart):
    return partSeqStats(cPart)[1]


def partMeanNCharsPerSite(cPart):
    return partSeqStats(cPart)[2]


def getUnconstrainedLogLike(cPart):
    if lib().p4g_partNPatterns(cPart) < 1:
        _err("getUnconstrainedLogLike")
    return lib().p4g_getUnconstrainedLogLike(cPart)


def p4_getCondLikes(cNode, pNum, nCat, dim, nPatterns):
    out = numpy.zeros((nCat, dim, nPatterns), numpy.float64)
    _ok(lib().p4g_getCondLikes(cNode, pNum, out.ctypes.data), "p4_getCondLikes")
    return out


# ----------------------------------------------------------------- batches
class Batch:
    """All pending partials and site sums of several trees in one launch
    (per state count and precision)."""

    def __init__(self, cTrees):
        self.n = len(cTrees)
        arr = (c_void_p * self.n)(*cTrees)
        self._arr = arr
        self.c = _h(lib().p4g_newBatch(self.n, ctypes.cast(arr, c_void_p)), "newBatch")
        self.out = numpy.zeros(self.n, numpy.float64)

    def logLikes(self):
        _ok(lib().p4g_batchLogLike(self.c, self.out.ctypes.data), "batchLogLike")
        return self.out.copy()

    def __del__(self):
        try:
            if self.c and _lib is not None:
                _lib.p4g_freeBatch(self.c)
        except Exception:
            pass


class Mlp:
    """The learned proposal's network on the GPU (p4g_mlp*): nIn -> nHidden
    -> nHidden -> nOut with ReLU, BF16 weights on the MFMA units, FP32
    accumulation, master weights and Adam."""

    def __init__(self, nIn, nHidden, nOut, seed=0):
        self.nIn, self.nHidden, self.nOut = nIn, nHidden, nOut
        self.c = _h(lib().p4g_mlpNew(nIn, nHidden, nOut, seed), "mlpNew")
        self.nParams = _lib.p4g_mlpNumParams(self.c)

    @staticmethod
    def _f32(a, cols=None):
        a = numpy.ascontiguousarray(a, dtype=numpy.float32)
        if cols is not None and (a.ndim != 2 or a.shape[1] != cols):
            raise P4Error("Mlp: expected an array of shape [n][%d], got %s" % (cols, a.shape))
        return a

    def getParams(self):
        out = numpy.empty(self.nParams, numpy.float32)
        _ok(_lib.p4g_mlpGetParams(self.c, out.ctypes.data), "mlpGetParams")
        return out

    def setParams(self, theta):
        theta = numpy.ascontiguousarray(theta, dtype=numpy.float32)
        if theta.size != self.nParams:
            raise P4Error("Mlp.setParams: %d values, expected %d" % (theta.size, self.nParams))
        _ok(_lib.p4g_mlpSetParams(self.c, theta.ctypes.data), "mlpSetParams")

    def forward(self, x):
        x = self._f32(x, self.nIn)
        out = numpy.empty((x.shape[0], self.nOut), numpy.float32)
        _ok(_lib.p4g_mlpForward(self.c, x.ctypes.data, x.shape[0], out.ctypes.data), "mlpForward")
        return out

    def _batch(self, x, y, mask):
        D = self.nOut // 2
        return self._f32(x, self.nIn), self._f32(y, D), self._f32(mask, D)

    def gradients(self, x, y, mask):
        """(loss summed over rows, gradient of loss / n)."""
        x, y, mask = self._batch(x, y, mask)
        g = numpy.empty(self.nParams, numpy.float32)
        loss = numpy.zeros(1, numpy.float32)
        _ok(_lib.p4g_mlpGradients(self.c, x.ctypes.data, y.ctypes.data, mask.ctypes.data, x.shape[0],
                                  g.ctypes.data, loss.ctypes.data), "mlpGradients")
        return float(loss[0]), g

    def trainStep(self, x, y, mask, lr):
        x, y, mask = self._batch(x, y, mask)
        loss = numpy.zeros(1, numpy.float32)
        _ok(_lib.p4g_mlpTrainStep(self.c, x.ctypes.data, y.ctypes.data, mask.ctypes.data, x.shape[0],
                                  float(lr), loss.ctypes.data), "mlpTrainStep")
        return float(loss[0])

    def __del__(self):
        try:
            if self.c and _lib is not None:
                _lib.p4g_mlpFree(self.c)
        except Exception:
            pass
