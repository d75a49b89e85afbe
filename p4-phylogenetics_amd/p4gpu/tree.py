"""Tree: topology, model placement, and the likelihood entry points
(p4/tree.py).  Kept to the hot-path subset: building trees, placing model
components on nodes, Tree.calcLogLike(), Tree.getSiteLikes(), and the C
mirror that Chain/Mcmc drive incrementally.

The C mirror is the GPU library (libp4gpu.so via p4gpu.pf): node relations,
branch lengths and model numbers are pushed by setCStuff(); partials, P
matrices and eigensystems live in GPU memory.
"""
import math
import random

import numpy

from . import pf
from .model import Comp, Gdasrv, Model, PInvar, RMatrix
from .node import Node, NodeBranchPart, NodePart
from .p4exceptions import P4Error
from .var import var


class Tree(object):
    def __init__(self):
        self.name = None
        self.nodes = []
        self.root = None
        self._data = None
        self._model = None
        self.cTree = None
        self.preOrder = None
        self.postOrder = None
        self.preAndPostOrderAreValid = 0
        self.partLikes = None
        self.logLike = None
        self.precision = var.precision
        self._taxNames = None

    # ------------------------------------------------------------- basics
    def __del__(self):
        try:
            self.deleteCStuff()
        except Exception:
            pass

    @property
    def nNodes(self):
        return len(self.nodes)

    @property
    def taxNames(self):
        if self._taxNames is None:
            self._taxNames = [n.name for n in self.iterLeavesNoRoot()]
        return self._taxNames

    @taxNames.setter
    def taxNames(self, v):
        self._taxNames = list(v)

    def node(self, spec):
        if isinstance(spec, Node):
            return spec
        if isinstance(spec, int):
            return self.nodes[spec]
        for n in self.nodes:
            if n.name == spec:
                return n
        raise P4Error("Tree.node(): no node %s" % spec)

    def iterNodes(self):
        for n in self.nodes:
            if n.nodeNum != var.NO_ORDER:
                yield n

    def iterNodesNoRoot(self):
        for n in self.iterNodes():
            if n is not self.root:
                yield n

    def iterLeavesNoRoot(self):
        for n in self.iterNodesNoRoot():
            if n.isLeaf:
                yield n

    def iterInternals(self):
        for n in self.iterNodes():
            if not n.isLeaf:
                yield n

    def iterInternalsNoRoot(self):
        for n in self.iterInternals():
            if n is not self.root:
                yield n

    def iterPreOrder(self):
        if not self.preAndPostOrderAreValid:
            self.setPreAndPostOrder()
        for i in self.preOrder:
            if i != var.NO_ORDER:
                yield self.nodes[i]

    def iterPostOrder(self):
        if not self.preAndPostOrderAreValid:
            self.setPreAndPostOrder()
        for i in self.postOrder:
            if i != var.NO_ORDER:
                yield self.nodes[i]

    def iterInternalsPostOrder(self):
        for n in self.iterPostOrder():
            if not n.isLeaf:
                yield n

    def setPreAndPostOrder(self):
        """Fill the int32 preOrder / postOrder arrays (NO_ORDER padded), in
        place when they already exist, since the C side aliases them."""
        pre = []
        stack = [self.root]
        while stack:
            n = stack.pop()
            pre.append(n.nodeNum)
            kids = list(n.iterChildren())
            for c in reversed(kids):
                stack.append(c)
        post = []
        stack = [(self.root, False)]
        while stack:
            n, done = stack.pop()
            if done or not n.leftChild:
                post.append(n.nodeNum)
                continue
            stack.append((n, True))
            for c in reversed(list(n.iterChildren())):
                stack.append((c, False))
        N = len(self.nodes)
        if self.preOrder is None or len(self.preOrder) != N:
            self.preOrder = numpy.full(N, var.NO_ORDER, numpy.int32)
            self.postOrder = numpy.full(N, var.NO_ORDER, numpy.int32)
        self.preOrder[:] = var.NO_ORDER
        self.postOrder[:] = var.NO_ORDER
        self.preOrder[:len(pre)] = pre
        self.postOrder[:len(post)] = post
        self.preAndPostOrderAreValid = 1

    def getLen(self):
        return sum(n.br.len for n in self.iterNodesNoRoot())

    def getNodeNumsAbove(self, theNode, leavesOnly=0):
        out = []
        stack = list(theNode.iterChildren())
        while stack:
            n = stack.pop()
            if not leavesOnly or n.isLeaf:
                out.append(n.nodeNum)
            stack.extend(n.iterChildren())
        return out

    # ----------------------------------------------------------- newick
    def writeNewick(self, withBrLens=True, precision=8):
        def rec(n):
            if n.isLeaf and n is not self.root:
                s = n.name or str(n.nodeNum)
            else:
                s = "(" + ",".join(rec(c) for c in n.iterChildren()) + ")"
                if n.isLeaf:
                    s = (n.name or str(n.nodeNum)) + s
            if withBrLens and n is not self.root:
                s += ":%.*g" % (precision, n.br.len)
            return s
        return rec(self.root) + ";"

    # ------------------------------------------------------------- data
    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, theData):
        """Attach data; leaves get seqNums by name and a fresh Model with
        one ModelPart per data part is made (p4/tree.py:6790-6860)."""
        self.deleteCStuff()
        self._data = theData
        if theData is None:
            return
        names = theData.taxNames
        for n in self.iterLeavesNoRoot():
            if n.name not in names:
                raise P4Error("Tree.data: leaf %s is not in the data" % n.name)
            n.seqNum = names.index(n.name)
        if self.root.isLeaf:
            self.root.seqNum = names.index(self.root.name)
        self._model = Model(theData.nParts)
        for pNum, dp in enumerate(theData.parts):
            mp = self._model.parts[pNum]
            mp.dim = dp.dim
            mp.pInvar = PInvar()
            mp.pInvar.partNum = pNum
        for n in self.nodes:
            n.parts = [NodePart() for _ in range(theData.nParts)]
            n.br.parts = [NodeBranchPart() for _ in range(theData.nParts)]

    @property
    def model(self):
        return self._model

    # ------------------------------------------------------ model setup
    def _checkModelThing(self, partNum, symbol, head):
        if not self._model:
            raise P4Error([head, "Set the data first, eg myTree.data = theData"])
        if partNum < 0 or partNum >= self._model.nParts:
            raise P4Error([head, "partNum %i is out of range" % partNum])

    def newComp(self, partNum=0, free=0, spec='empirical', val=None, symbol=None):
        """p4/tree.py:6901-7044.  spec: 'equal', 'empirical' or 'specified'
        (dim or dim-1 values)."""
        head = 'Tree.newComp()'
        self._checkModelThing(partNum, symbol, head)
        self.deleteCStuff()
        mp = self._model.parts[partNum]
        mt = Comp()
        mt.partNum = partNum
        mt.free = free
        if spec not in ('equal', 'empirical', 'specified'):
            raise P4Error([head, "spec should be one of 'equal', 'empirical', 'specified'"])
        mt.spec = spec
        mt.num = len(mp.comps)
        mt.symbol = symbol or (var.modelSymbols[mt.num] if mt.num < len(var.modelSymbols) else '-')
        dim = mp.dim
        if spec == 'equal':
            mt.val = numpy.ones(dim, dtype=numpy.double) / dim
        elif spec == 'empirical':
            mt.val = None
        else:
            try:
                val = [float(x) for x in val]
            except TypeError:
                raise P4Error([head, "Comp is 'specified', but bad 'val' arg."])
            if len(val) == dim - 1:
                last = 1.0 - sum(val)
                if last <= 0.0:
                    raise P4Error([head, "Bad comp vals %s" % val])
                val = val + [last]
            elif len(val) == dim:
                if math.fabs(sum(val) - 1.0) > 9.e-16:
                    raise P4Error([head, "Bad comp vals %s" % val, "does not sum to 1.0"])
            else:
                raise P4Error([head, "Bad length for val arg (%i); dim is %i" % (len(val), dim)])
            needsNorm = False
            for i in range(len(val)):
                if val[i] < var.PIVEC_MIN:
                    val[i] = (1.5 + random.random()) * var.PIVEC_MIN
                    needsNorm = True
            if needsNorm:
                s = sum(val)
                val = [v / s for v in val]
            mt.val = numpy.array(val, dtype=numpy.double)
        mp.comps.append(mt)
        return mt

    def newRMatrix(self, partNum=0, free=0, spec='ones', val=None, symbol=None):
        """p4/tree.py:7046-7171.  spec 'ones', '2p' (val = kappa) or
        'specified' (dim*(dim-1)/2 values, normalised to sum 1)."""
        head = 'Tree.newRMatrix()'
        self._checkModelThing(partNum, symbol, head)
        self.deleteCStuff()
        mp = self._model.parts[partNum]
        mt = RMatrix()
        mt.partNum = partNum
        mt.free = free
        if spec not in ('ones', '2p', 'specified', 'optimized'):
            if spec in var.rMatrixProteinSpecs:
                raise P4Error([head, "empirical protein rate matrices are not built into this engine;",
                               "give the 190 exchangeabilities with spec='specified'"])
            raise P4Error([head, "Got unknown rMatrix spec '%s'." % spec])
        mt.spec = spec
        mt.num = len(mp.rMatrices)
        mt.symbol = symbol or (var.modelSymbols[mt.num] if mt.num < len(var.modelSymbols) else '-')
        dim = mp.dim
        goodLen = (dim * dim - dim) // 2 if var.rMatrixNormalizeTo1 else (dim * dim - dim) // 2 - 1
        if spec == 'specified':
            if val is None or len(val) != goodLen:
                raise P4Error([head, "Bad length for arg val; should be %i" % goodLen])
            v = numpy.array(val, dtype=numpy.double)
            if var.rMatrixNormalizeTo1:
                v /= v.sum()
        elif spec == 'ones' or spec == 'optimized':
            v = numpy.ones(goodLen, dtype=numpy.double)
            if var.rMatrixNormalizeTo1:
                v /= v.sum()
        else:
            try:
                k = float(val)
            except (TypeError, ValueError):
                k = 2.0
            k = min(max(k, var.KAPPA_MIN), var.KAPPA_MAX)
            v = numpy.array([k], dtype=numpy.double)
        mt.val = v
        mp.rMatrices.append(mt)
        return mt

    def newGdasrv(self, partNum=0, free=0, val=None, symbol=None):
        """p4/tree.py:7173-7225."""
        head = 'Tree.newGdasrv()'
        self._checkModelThing(partNum, symbol, head)
        self.deleteCStuff()
        mp = self._model.parts[partNum]
        if mp.nGammaCat == 1:
            raise P4Error([head, "nGammaCat is 1 for part %i; call setNGammaCat first" % partNum])
        if val is None:
            raise P4Error([head, "Please specify a val, a positive float."])
        v = float(val)
        if v <= 0.000001 or v >= 300.0:
            raise P4Error([head, "Arg val must be between 0.000001 and 300.  Got %f" % v])
        mt = Gdasrv()
        mt.nGammaCat = mp.nGammaCat
        mt.partNum = partNum
        mt.free = free
        mt.num = len(mp.gdasrvs)
        mt.symbol = symbol or var.modelSymbols[mt.num]
        mt.freqs = numpy.zeros(mt.nGammaCat, dtype=numpy.double)
        mt.rates = numpy.zeros(mt.nGammaCat, dtype=numpy.double)
        mt._val[0] = v
        mt.calcRates()
        mp.gdasrvs.append(mt)
        return mt

    def setPInvar(self, partNum=0, free=0, val=0.0):
        head = 'Tree.setPInvar()'
        v = float(val)
        if v < 0.0 or v >= 1.0:
            raise P4Error([head, "Arg val must be zero or more, and less than 1.  Got %f" % v])
        self._checkModelThing(partNum, None, head)
        self.deleteCStuff()
        mt = PInvar()
        mt.partNum = partNum
        mt.free = free
        mt.val = v
        self._model.parts[partNum].pInvar = mt

    def setRelRate(self, partNum=0, val=0.0):
        head = 'Tree.setRelRate()'
        v = float(val)
        if v <= 0.0 or v >= 1000.0:
            raise P4Error([head, "Arg val must be more than zero, and less than 1000.  Got %f" % v])
        self._checkModelThing(partNum, None, head)
        self.deleteCStuff()
        self._model.parts[partNum].relRate = v

    def setNGammaCat(self, partNum=0, nGammaCat=1):
        head = 'Tree.setNGammaCat()'
        self._checkModelThing(partNum, None, head)
        if not isinstance(nGammaCat, int) or nGammaCat < 1:
            raise P4Error([head, "nGammaCat should be a positive int"])
        self.deleteCStuff()
        self._model.parts[partNum].nGammaCat = nGammaCat

    def setModelComponentOnNode(self, theModelComponent, node=None, clade=1):
        """p4/tree.py:7273-7339."""
        if not isinstance(theModelComponent, (Comp, RMatrix, Gdasrv)):
            raise P4Error("Expecting a model component instance (comp, rMatrix or gdasrv)")
        self.deleteCStuff()
        pNum = theModelComponent.partNum
        theNode = self.root if node is None else self.node(node)
        targets = [theNode]
        if clade:
            targets += [self.nodes[i] for i in self.getNodeNumsAbove(theNode)]
        for n in targets:
            if isinstance(theModelComponent, Comp):
                n.parts[pNum].compNum = theModelComponent.num
            elif n is not self.root:
                if isinstance(theModelComponent, RMatrix):
                    n.br.parts[pNum].rMatrixNum = theModelComponent.num
                else:
                    n.br.parts[pNum].gdasrvNum = theModelComponent.num

    def setModelComponentsOnNodesRandomly(self, forceRepresentation=2, rng=None):
        """p4/tree.py:7347-7485: each component gets at least
        forceRepresentation nodes, the rest are random."""
        rng = rng or random
        head = 'Tree.setModelComponentsOnNodesRandomly()'
        if not isinstance(forceRepresentation, int) or forceRepresentation < 1:
            raise P4Error([head, "Arg 'forceRepresentation' should be 1 or more."])
        self.deleteCStuff()
        if not self.preAndPostOrderAreValid:
            self.setPreAndPostOrder()
        for pNum, mp in enumerate(self._model.parts):
            if mp.nComps == 1:
                for n in self.nodes:
                    n.parts[pNum].compNum = 0
            elif mp.nComps > 1:
                if mp.nComps * forceRepresentation > len(self.nodes):
                    raise P4Error([head, "not enough nodes for %i comps" % mp.nComps])
                nList = self.nodes[:]
                rng.shuffle(nList)
                for mt in range(mp.nComps):
                    for _ in range(forceRepresentation):
                        nList.pop().parts[pNum].compNum = mt
                for n in nList:
                    n.parts[pNum].compNum = rng.randrange(mp.nComps)
            else:
                raise P4Error([head, "No comps in part %i" % pNum])
            for attr, count in (('rMatrixNum', mp.nRMatrices),
                                ('gdasrvNum', mp.nGdasrvs if mp.nGammaCat > 1 else 0)):
                if count == 0:
                    if attr == 'rMatrixNum':
                        raise P4Error([head, "No rMatrices in part %i" % pNum])
                    continue
                nList = [n for n in self.nodes if n is not self.root]
                if count == 1:
                    for n in nList:
                        setattr(n.br.parts[pNum], attr, 0)
                    continue
                if count * forceRepresentation > len(nList):
                    raise P4Error([head, "not enough nodes for %i components" % count])
                rng.shuffle(nList)
                for mt in range(count):
                    for _ in range(forceRepresentation):
                        setattr(nList.pop().br.parts[pNum], attr, mt)
                for n in nList:
                    setattr(n.br.parts[pNum], attr, rng.randrange(count))

    def modelSanityCheck(self, resetEmpiricalComps=True):
        """Fill defaults and check completeness (p4/tree.py:7647-7790)."""
        head = 'Tree.modelSanityCheck()'
        if not self._model:
            raise P4Error([head, "No model (attach data first)"])
        isHet = 0
        for pNum, mp in enumerate(self._model.parts):
            if not mp.comps:
                raise P4Error([head, "Part %i has no comps" % pNum])
            if not mp.rMatrices:
                raise P4Error([head, "Part %i has no rMatrices" % pNum])
            if mp.nGammaCat > 1 and not mp.gdasrvs:
                raise P4Error([head, "Part %i has nGammaCat %i but no gdasrv" % (pNum, mp.nGammaCat)])
            if mp.nGammaCat == 1 and mp.gdasrvs:
                raise P4Error([head, "Part %i has gdasrvs but nGammaCat is 1" % pNum])
            if mp.pInvar is None:
                mp.pInvar = PInvar()
            for mt in mp.comps:
                if mt.spec == 'empirical' and (mt.val is None or resetEmpiricalComps):
                    emp = self._empiricalComp(pNum)
                    if mt.val is None:
                        mt.val = emp
                    else:
                        mt.val[:] = emp
            mp.isHet = int(mp.nComps > 1 or mp.nRMatrices > 1 or mp.nGdasrvs > 1)
            isHet |= mp.isHet
            for n in self.iterNodes():
                if mp.nComps == 1:
                    n.parts[pNum].compNum = 0
                elif n.parts[pNum].compNum < 0:
                    raise P4Error([head, "node %i has no comp in part %i" % (n.nodeNum, pNum)])
                if n is self.root:
                    continue
                if mp.nRMatrices == 1:
                    n.br.parts[pNum].rMatrixNum = 0
                elif n.br.parts[pNum].rMatrixNum < 0:
                    raise P4Error([head, "node %i has no rMatrix in part %i" % (n.nodeNum, pNum)])
                if mp.nGdasrvs == 1:
                    n.br.parts[pNum].gdasrvNum = 0
                elif mp.nGdasrvs > 1 and n.br.parts[pNum].gdasrvNum < 0:
                    raise P4Error([head, "node %i has no gdasrv in part %i" % (n.nodeNum, pNum)])
        self._model.isHet = isHet

    def _empiricalComp(self, pNum):
        pats, cnts, _, _, _ = self._data.parts[pNum].patterns()
        dim = self._model.parts[pNum].dim
        c = numpy.zeros(dim)
        for s in range(dim):
            c[s] = ((pats == s) * cnts[None, :]).sum()
        c = numpy.maximum(c / max(c.sum(), 1.0), var.PIVEC_MIN)
        return c / c.sum()

    # -------------------------------------------------------- C mirror
    def deleteCStuff(self):
        """p4/tree.py:9233-9253."""
        for n in self.nodes:
            if n.cNode:
                n.cNode = None
        if self.cTree:
            pf.p4_freeTree(self.cTree)
            self.cTree = None
        if self._model is not None and self._model.cModel:
            pf.p4_freeModel(self._model.cModel)
            self._model.cModel = None

    def _allocCStuff(self, resetEmpiricalComps=True):
        """p4/tree.py:9255-9320."""
        for i, n in enumerate(self.nodes):
            if n.nodeNum != i:
                raise P4Error("Tree._allocCStuff(): nodes must be numbered consecutively from zero")
        self.modelSanityCheck(resetEmpiricalComps=resetEmpiricalComps)
        if not self._data.cData:
            self._data._setCStuff()
        if not self._model.cModel:
            self._model.allocCStuff()
        if not self.preAndPostOrderAreValid:
            self.setPreAndPostOrder()
        nLeaves = sum(1 for n in self.iterNodes() if n.isLeaf)
        self.partLikes = numpy.zeros(self._model.nParts, dtype=numpy.double)
        self.cTree = pf.p4_newTree(len(self.nodes), nLeaves, self.preOrder, self.postOrder,
                                   var._newtAndBrentPowellOptPassLimit, self.partLikes,
                                   self._data.cData, self._model.cModel)
        pf.p4_setPrecision(self.cTree, self.precision == 'fp32')
        inTree = set(int(i) for i in self.preOrder if i != var.NO_ORDER)
        for n in self.nodes:
            n.cNode = pf.p4_newNode(n.nodeNum, self.cTree, n.seqNum, n.isLeaf,
                                    1 if n.nodeNum in inTree else 0)

    def setCStuff(self):
        """Push relations, root, branch lengths and model numbers
        (p4/tree.py:9322-9377)."""
        for n in self.iterNodes():
            pf.p4_setNodeRelation(n.cNode, 0, n.parent.nodeNum if n.parent else -1)
            pf.p4_setNodeRelation(n.cNode, 1, n.leftChild.nodeNum if n.leftChild else -1)
            pf.p4_setNodeRelation(n.cNode, 2, n.sibling.nodeNum if n.sibling else -1)
        pf.p4_setTreeRoot(self.cTree, self.root.cNode)
        for n in self.iterNodesNoRoot():
            pf.p4_setBrLen(n.cNode, n.br.len)
        if self._model.isHet:
            for pNum, mp in enumerate(self._model.parts):
                if mp.isHet:
                    for n in self.iterNodes():
                        pf.p4_setCompNum(n.cNode, pNum, n.parts[pNum].compNum)
                        if n is not self.root:
                            pf.p4_setRMatrixNum(n.cNode, pNum, n.br.parts[pNum].rMatrixNum)
                            pf.p4_setGdasrvNum(n.cNode, pNum, max(n.br.parts[pNum].gdasrvNum, 0))
        if not self.preAndPostOrderAreValid:
            self.setPreAndPostOrder()

    def _commonCStuff(self, resetEmpiricalComps=True):
        """p4/tree.py:9379-9404."""
        if not self._data:
            raise P4Error(["Tree (_commonCStuff)", "This tree has no data attached."])
        if not self.cTree:
            self._allocCStuff(resetEmpiricalComps=resetEmpiricalComps)
        self._model.setCStuff()
        self.setCStuff()
        pf.p4_setPrams(self.cTree, -1)

    def calcLogLike(self, verbose=1, resetEmpiricalComps=True):
        """Calculate the likelihood of the tree, without optimization
        (p4/tree.py:9406-9414)."""
        self._commonCStuff(resetEmpiricalComps=resetEmpiricalComps)
        self.logLike = pf.p4_treeLogLike(self.cTree, 0)
        if verbose:
            print("Tree.calcLogLike(). %f" % self.logLike)
        return self.logLike

    def getSiteLikes(self):
        """Per-site likelihoods of every part (p4/tree.py:9679-9686)."""
        self._commonCStuff()
        self.logLike = pf.p4_treeLogLike(self.cTree, 1)
        out = []
        for p in self._data.parts:
            out += pf.getSiteLikes(p.cPart)
        return out

    def simulate(self, calculatePatterns=True, resetSequences=True, refTree=None, seed=None):
        """Simulate into the attached data (p4/tree.py:9527-9627), every part
        on the GPU, one thread per site.  With refTree (a tree, model and
        other data of the same shape), each site's root state, category and
        invariance are drawn from refTree's root conditional likelihoods, as
        the reference's PhyloBayes-style option does.  seed keys the
        counter-based generator (default: from Python's random module).
        calculatePatterns makes the parts' patterns again; resetSequences
        brings the new sequences into the alignments."""
        if refTree is not None:
            if refTree.data is None or refTree.model is None:
                raise P4Error("Tree.simulate(): refTree needs data and a model")
            if refTree.data is self.data:
                raise P4Error("Tree.simulate(): refTree must have its own data")
            if refTree.data.nParts != self.data.nParts or refTree.data.nTax != self.data.nTax:
                raise P4Error("Tree.simulate(): refTree's data differs in shape")
            refTree.calcLogLike(verbose=0)
        self._commonCStuff()
        if seed is None:
            seed = random.getrandbits(64)
        pf.p4_simulate(self.cTree, refTree.cTree if refTree is not None else None, int(seed))
        if calculatePatterns:
            for p in self.data.parts:
                pf.makePatternsDevice(p.cPart)   # the sequences are still on the device
                pf.setGlobalInvarSitesVec(p.cPart)
                p.nPatterns = pf.partPatternCount(p.cPart)
        if resetSequences and self.data.alignments:
            self.data.resetSequencesFromParts()

    def ancestralStateDraw(self, seed=None):
        """One draw of the root's character states, all sites, as a string
        (p4/tree.py:9640-9676): for each site a (gamma category, state) or
        an invariant state is drawn from the root's conditional likelihoods
        and the root composition, and the state's symbol is kept.  All
        sites of a part are drawn in one GPU launch; seed (default: from
        Python's random module) keys the GPU's Philox stream."""
        self._commonCStuff()
        self.logLike = pf.p4_treeLogLike(self.cTree, 0)
        out = []
        for pNum, dp in enumerate(self._data.parts):
            n = pf.partNSites(dp.cPart)
            s = random.getrandbits(64) if seed is None else (int(seed) + pNum) & 0xFFFFFFFFFFFFFFFF
            d = pf.drawAncStates(self.cTree, pNum, 0, n, seed=s)
            st = numpy.where(d[:, 1] >= 0, d[:, 0], d[:, 3])
            if (st < 0).any():
                raise P4Error("Tree.ancestralStateDraw(): a site got no draw")
            sym = numpy.frombuffer(dp.symbols.encode(), numpy.uint8)
            out.append(sym[st].tobytes().decode())
        return ''.join(out)

    def branchDerivatives(self, pNum=-1):
        """(lnLEdge, d1, d2) by node number: the log likelihood computed
        across each non-root node's branch (it equals the tree's) and its
        first and second derivatives with respect to that branch length,
        for part pNum or summed over all parts (-1).  FP64 only."""
        self._commonCStuff()
        return pf.branchDerivatives(self.cTree, pNum, len(self.nodes))

    def optBrLensNewton(self, epsilon=1.0e-5, likeDelta=1.0e-7, maxIter=100, verbose=0):
        """Maximise the likelihood over all branch lengths (model fixed) by
        Newton-Raphson on every branch at once, from the derivatives of
        branchDerivatives().  Each branch's step follows p4_newtNode
        (Pf/p4_treeNewt.c:543-590): t - d1/d2, or t/5 where the curvature
        is not negative, clipped to [BRLEN_MIN, BRLEN_MAX] and to five times
        the length before the step.  The reference moves one branch at a time; here
        all move together, and a step that lowers the likelihood is halved
        until it does not.  Stops when every |d1| < epsilon, or when a full
        (unhalved) step gains less than likeDelta.  Returns the final log
        likelihood."""
        lo, hi = var.BRLEN_MIN, var.BRLEN_MAX
        nodes = [n for n in self.iterNodesNoRoot()]
        lnl, d1, d2 = self.branchDerivatives()
        cur = float(lnl[nodes[0].nodeNum])
        for it in range(maxIter):
            t = numpy.array([n.br.len for n in nodes])
            g1 = numpy.array([d1[n.nodeNum] for n in nodes])
            g2 = numpy.array([d2[n.nodeNum] for n in nodes])
            free = ~(((t <= lo) & (g1 < 0)) | ((t >= hi) & (g1 > 0)))
            if not (numpy.abs(g1[free]) >= epsilon).any():
                break
            with numpy.errstate(divide='ignore', invalid='ignore'):
                nt = numpy.where(g2 < 0.0, t - g1 / g2, t / 5.0)
            nt = numpy.clip(numpy.minimum(nt, 5.0 * t), lo, hi)
            step = nt - t
            halved = 0
            for halved in range(30):
                for n, v in zip(nodes, t + step):
                    n.br.len = float(v)
                new = self.calcLogLike(verbose=0)
                if new >= cur - 1e-12 * abs(cur):
                    break
                step *= 0.5
            else:
                for n, v in zip(nodes, t):
                    n.br.len = float(v)
                new = self.calcLogLike(verbose=0)
                break
            gain = new - cur
            cur = new
            if verbose:
                print("optBrLensNewton %3d  lnL %.8f  gain %.3g" % (it, cur, gain))
            lnl, d1, d2 = self.branchDerivatives()
            if halved == 0 and abs(gain) < likeDelta:
                break
        self.logLike = self.calcLogLike(verbose=0)
        return self.logLike

    def setPrecision(self, precision):
        """'fp64' (default, the reference's numerics) or 'fp32' (FP32
        partials with power-of-two scale exponents)."""
        if precision not in ('fp64', 'fp32'):
            raise P4Error("precision should be 'fp64' or 'fp32'")
        self.precision = precision
        if self.cTree:
            pf.p4_setPrecision(self.cTree, precision == 'fp32')

    # ------------------------------------------------------------ copies
    def dupe(self):
        """A structural copy (no C state), with its own model values."""
        t = Tree()
        t.name = self.name
        t.precision = self.precision
        t.nodes = []
        for n in self.nodes:
            m = Node()
            m.name, m.nodeNum, m.isLeaf, m.seqNum = n.name, n.nodeNum, n.isLeaf, n.seqNum
            m.br.len = n.br.len
            t.nodes.append(m)
        for n, m in zip(self.nodes, t.nodes):
            m.parent = t.nodes[n.parent.nodeNum] if n.parent else None
            m.leftChild = t.nodes[n.leftChild.nodeNum] if n.leftChild else None
            m.sibling = t.nodes[n.sibling.nodeNum] if n.sibling else None
        t.root = t.nodes[self.root.nodeNum]
        t._taxNames = list(self.taxNames)
        if self._data is not None:
            t._data = self._data
            t._model = _copyModel(self._model)
            for n, m in zip(self.nodes, t.nodes):
                m.parts = [NodePart() for _ in n.parts]
                m.br.parts = [NodeBranchPart() for _ in n.br.parts]
                for a, b in zip(n.parts, m.parts):
                    b.compNum = a.compNum
                for a, b in zip(n.br.parts, m.br.parts):
                    b.rMatrixNum, b.gdasrvNum = a.rMatrixNum, a.gdasrvNum
        t.setPreAndPostOrder()
        return t

    def copyToTree(self, other):
        """Copy topology, branch lengths and node model numbers into a
        tree with the same nodes (the MCMC cur/prop pair), in place."""
        for n, m in zip(self.nodes, other.nodes):
            m.parent = other.nodes[n.parent.nodeNum] if n.parent else None
            m.leftChild = other.nodes[n.leftChild.nodeNum] if n.leftChild else None
            m.sibling = other.nodes[n.sibling.nodeNum] if n.sibling else None
            m.br.len = n.br.len
            for a, b in zip(n.parts, m.parts):
                b.compNum = a.compNum
            for a, b in zip(n.br.parts, m.br.parts):
                b.rMatrixNum, b.gdasrvNum = a.rMatrixNum, a.gdasrvNum
        other.root = other.nodes[self.root.nodeNum]
        other.preOrder[:] = self.preOrder
        other.postOrder[:] = self.postOrder
        other.preAndPostOrderAreValid = self.preAndPostOrderAreValid


def _copyModel(m):
    import copy
    nm = Model(m.nParts)
    nm.doRelRates, nm.relRatesAreFree, nm.nFreePrams, nm.isHet = (
        m.doRelRates, m.relRatesAreFree, m.nFreePrams, m.isHet)
    for a, b in zip(m.parts, nm.parts):
        b.dim, b.nGammaCat, b.isHet, b.relRate = a.dim, a.nGammaCat, a.isHet, a.relRate
        b.pInvar = copy.copy(a.pInvar)
        for c in a.comps:
            d = copy.copy(c)
            d.val = c.val.copy() if c.val is not None else None
            b.comps.append(d)
        for r in a.rMatrices:
            d = copy.copy(r)
            d.val = r.val.copy() if r.val is not None else None
            b.rMatrices.append(d)
        for g in a.gdasrvs:
            d = Gdasrv()
            d.num, d.partNum, d.free, d.symbol, d.nGammaCat = g.num, g.partNum, g.free, g.symbol, g.nGammaCat
            d._val = g._val.copy()
            d.freqs = g.freqs.copy()
            d.rates = g.rates.copy()
            b.gdasrvs.append(d)
    return nm
