"""Metropolis-coupled MCMC over the GPU likelihood (p4/mcmc.py, p4/chain.py).

Structure follows p4: each Chain holds a curTree and a propTree, each with
its own GPU state; a proposal mutates propTree, the affected P matrices are
rebuilt and the affected partials marked, the proposal is scored, accepted
or rejected by Metropolis-Hastings with the chain's heat, and the winner's
state is copied over the loser's (p4/chain.py:1138-1704).

Differences that keep the same stationary distribution but suit the GPU:
  * the pending proposals of all chains are scored in ONE batched launch
    (p4g_batchLogLike) instead of one chain after another
    (p4/mcmc.py:2861-2950);
  * partial likelihoods are recomputed in one launch per evaluation for
    all marked nodes, in postorder, at p4_partLogLike time.

Proposals: 'brLen' (multiplier with log reflection, p4/chain.py:3091-3141),
'allBrLens' (sliding window, p4/chain.py:3143-3196), 'compDir'
(Dirichlet centred on the current comp, p4/chain.py:1798-1866),
'allCompsDir' (every comp of a part at once, p4/chain.py:2259-2289),
'rMatrixDir' (same for the exchangeabilities), 'gdasrv' (multiplier on the
gamma shape), 'pInvar' (sliding window), 'nni' (nearest-neighbour
interchange, a symmetric topology move), and optionally 'learned' (see
p4gpu.learned).  Priors: branch lengths Exponential(10), flat Dirichlet on
comps and exchangeabilities, uniform gamma shape, pInvar and topology.
"""
import math
import random
import time

import numpy

from . import pf
from .func import effectiveSampleSize
from .p4exceptions import P4Error
from .var import var


class Proposal(object):
    """A proposal type with per-temperature tuning and counters
    (p4/mcmc.py:170-283)."""

    def __init__(self, name, weight, tuning, nTemps, pNum=-1):
        self.name = name
        self.weight = weight
        self.pNum = pNum
        self.tuning = [tuning] * nTemps
        self.nProposals = [0] * nTemps
        self.nAcceptances = [0] * nTemps
        self.tnNSamples = [0] * nTemps
        self.tnNAccepts = [0] * nTemps

    def tune(self, tempNum, target=(0.1, 0.7)):
        """Move the tuning toward an acceptance rate inside target
        (p4/mcmc.py:250-283)."""
        n = self.tnNSamples[tempNum]
        if n < 50 or self.name == 'learned':
            return
        acc = self.tnNAccepts[tempNum] / n
        t = self.tuning[tempNum]
        multiplier_like = self.name in ('brLen', 'gdasrv', 'allBrLens', 'pInvar', 'nni')
        if acc < target[0]:
            t = t / 1.5 if multiplier_like else t * 1.5
        elif acc > target[1]:
            t = t * 1.5 if multiplier_like else t / 1.5
        self.tuning[tempNum] = t
        self.tnNSamples[tempNum] = 0
        self.tnNAccepts[tempNum] = 0


def _dirichlet_logpdf(x, alpha):
    return (math.lgamma(alpha.sum()) - sum(math.lgamma(a) for a in alpha) +
            float(((alpha - 1.0) * numpy.log(x)).sum()))


def heatedLogRatio(beta, lnlProp, lnlCur, logPriorRatio, logProposalRatio, samplePrior=False):
    """Log Metropolis-Hastings ratio of a move in a chain of heat
    beta = 1 / (1 + T): Metropolis-coupled MCMC powers the likelihood ratio
    and the prior ratio by beta, not the proposal ratio (p4/chain.py:1060-1122,
    1129).  samplePrior leaves the likelihood ratio out (the commented
    logLikeRatio = 0.0 of p4/chain.py:1065-1066).  A proposed lnL at the
    -1e99 sentinel (a site likelihood <= 0) is a rejection."""
    if samplePrior:
        return beta * logPriorRatio + logProposalRatio
    if lnlProp <= -1.0e99 * 0.5:
        return -1.0e300
    return beta * (lnlProp - lnlCur) + beta * logPriorRatio + logProposalRatio


def acceptProbability(logR):
    """p4/chain.py:1310-1320: 0 below -100, 1 at or above 0, else exp."""
    if logR < -100.0:
        return 0.0
    if logR >= 0.0:
        return 1.0
    return math.exp(logR)


def swapLogRatio(beta1, beta2, lnl1, lnl2):
    """Log ratio of exchanging the states of two chains of heats beta1 and
    beta2 with current log likelihoods lnl1 and lnl2, as p4/mcmc.py:3315-3322
    computes it: likelihoods only (the reference leaves the heated priors out
    of the swap, which is exact when the prior is flat)."""
    return beta1 * lnl2 + beta2 * lnl1 - beta1 * lnl1 - beta2 * lnl2


class Chain(object):
    def __init__(self, mcmc, tempNum, tree):
        self.mcmc = mcmc
        self.tempNum = tempNum
        self.curTree = tree
        self.propTree = tree.dupe()
        self.propTree.precision = tree.precision
        self.propTree.calcLogLike(verbose=0)
        self.curTree.calcLogLike(verbose=0)
        self.rng = mcmc.rng
        self.logPriorRatio = 0.0
        self.logProposalRatio = 0.0
        self.touched = set()
        self.touchedParts = set()
        self.topologyChanged = False

    # ------------------------------------------------------------ helpers
    def _markPath(self, tree, node, pNums=None):
        """Mark node and its ancestors for recomputation (one library call
        per part list, not one per node and part)."""
        if pNums is None:
            pf.markPath(tree.cTree, node.nodeNum, -1)
        else:
            for p in pNums:
                pf.markPath(tree.cTree, node.nodeNum, p)

    def _markAll(self, tree, pNum):
        pf.markAll(tree.cTree, pNum)

    # ---------------------------------------------------------- proposals
    def propose(self, prop):
        self.logPriorRatio = 0.0
        self.logProposalRatio = 0.0
        self.touched = set()
        self.touchedParts = set()
        self.topologyChanged = False
        getattr(self, "_propose_" + prop.name)(prop)

    def _propose_brLen(self, prop):
        t = self.propTree
        rnd = self.rng
        n = t.nodes[rnd.randrange(1, len(t.nodes))] if t.root is t.nodes[0] else \
            rnd.choice([x for x in t.iterNodesNoRoot()])
        old = n.br.len
        new = old * math.exp(prop.tuning[self.tempNum] * (rnd.random() - 0.5))
        while new < var.BRLEN_MIN or new > var.BRLEN_MAX:
            if new < var.BRLEN_MIN:
                new = var.BRLEN_MIN * var.BRLEN_MIN / new
            else:
                new = var.BRLEN_MAX * var.BRLEN_MAX / new
        n.br.len = new
        self.logProposalRatio = math.log(new / old)
        self.logPriorRatio = self.mcmc.brLenPriorLambda * (old - new)
        pf.p4_setBrLen(n.cNode, new)
        pf.p4_calculateBigPDecks(n.cNode)
        self._markPath(t, n.parent)
        self.touched.add(n)

    def _propose_allBrLens(self, prop):
        t = self.propTree
        rnd = self.rng
        lam = self.mcmc.brLenPriorLambda
        for n in t.iterNodesNoRoot():
            old = n.br.len
            new = old + prop.tuning[self.tempNum] * (rnd.random() - 0.5)
            while new < var.BRLEN_MIN or new > var.BRLEN_MAX:
                if new < var.BRLEN_MIN:
                    new = (var.BRLEN_MIN - new) + var.BRLEN_MIN
                else:
                    new = var.BRLEN_MAX - (new - var.BRLEN_MAX)
            n.br.len = new
            self.logPriorRatio += lam * (old - new)
            pf.p4_setBrLen(n.cNode, new)
            self.touched.add(n)
        pf.p4_calculateAllBigPDecksAllParts(t.cTree)
        for p in range(t.model.nParts):
            self._markAll(t, p)

    def _dirichletMove(self, cur, a, lo):
        """A draw from Dirichlet(a * cur) (p4/chain.py:1798-1866) and the log
        Hastings ratio log q(cur | new) - log q(new | cur), or None when a
        component falls below lo.  The reference nudges such components up
        and keeps the Dirichlet densities, which are then not the density of
        what was proposed; here the draw is rejected instead (the prior is
        zero there), which keeps the move exact."""
        new = self.mcmc.nprng.dirichlet(a * cur)
        if new.min() < lo:
            return None, 0.0
        new = new / new.sum()
        return new, _dirichlet_logpdf(cur, a * new) - _dirichlet_logpdf(new, a * cur)

    def _reject(self):
        self.logProposalRatio = -1.0e300

    def _propose_compDir(self, prop):
        t = self.propTree
        mp = t.model.parts[prop.pNum]
        c = mp.comps[self.rng.randrange(mp.nComps)]
        new, lr = self._dirichletMove(c.val.copy(), prop.tuning[self.tempNum], var.PIVEC_MIN)
        if new is None:
            self._reject()
            return
        self.logProposalRatio = lr
        c.val[:] = new
        self._modelChanged(prop.pNum)

    def _propose_allCompsDir(self, prop):
        """Every comp of the part in one move, each from a Dirichlet centred
        on its current value; the log proposal ratios add
        (p4/chain.py:2259-2289)."""
        t = self.propTree
        mp = t.model.parts[prop.pNum]
        news, lr = [], 0.0
        for c in mp.comps:
            new, l1 = self._dirichletMove(c.val.copy(), prop.tuning[self.tempNum], var.PIVEC_MIN)
            if new is None:
                self._reject()
                return
            news.append(new)
            lr += l1
        self.logProposalRatio = lr
        for c, new in zip(mp.comps, news):
            c.val[:] = new
        self._modelChanged(prop.pNum)

    def _propose_rMatrixDir(self, prop):
        t = self.propTree
        mp = t.model.parts[prop.pNum]
        r = mp.rMatrices[self.rng.randrange(mp.nRMatrices)]
        new, lr = self._dirichletMove(r.val / r.val.sum(), prop.tuning[self.tempNum], var.RATE_MIN)
        if new is None:
            self._reject()
            return
        self.logProposalRatio = lr
        r.val[:] = new
        mp.setCStuff(t.model)
        self._modelChanged(prop.pNum, pushModel=False)

    def _propose_gdasrv(self, prop):
        t = self.propTree
        mp = t.model.parts[prop.pNum]
        g = mp.gdasrvs[self.rng.randrange(mp.nGdasrvs)]
        old = float(g.val[0])
        new = old * math.exp(prop.tuning[self.tempNum] * (self.rng.random() - 0.5))
        lo, hi = var.GAMMA_SHAPE_MIN, var.GAMMA_SHAPE_MAX
        while new < lo or new > hi:
            new = lo * lo / new if new < lo else hi * hi / new
        g.val = new
        self.logProposalRatio = math.log(new / old)
        self._modelChanged(prop.pNum)

    def _propose_pInvar(self, prop):
        t = self.propTree
        mp = t.model.parts[prop.pNum]
        old = mp.pInvar.val
        new = old + prop.tuning[self.tempNum] * (self.rng.random() - 0.5)
        hi = var.PINVAR_MAX
        while new < 0.0 or new > hi:
            new = -new if new < 0.0 else hi - (new - hi)
        mp.pInvar.val = new
        self._modelChanged(prop.pNum)

    def _modelChanged(self, pNum, pushModel=True):
        t = self.propTree
        if pushModel:
            t.model.parts[pNum].setCStuff(t.model)
        pf.p4_setPrams(t.cTree, pNum)
        self._markAll(t, pNum)
        self.touchedParts.add(pNum)

    def _propose_nni(self, prop):
        """Nearest-neighbour interchange across a random internal edge
        u -- parent(u): swap a child of u with a sibling of u.  The move is
        its own inverse and the numbers of choices are the same in both
        directions, so the proposal ratio is 1."""
        t = self.propTree
        rnd = self.rng
        cands = self.mcmc._nniCandidates
        u = t.nodes[rnd.choice(cands)]
        p = u.parent
        xs = list(u.iterChildren())
        ss = [c for c in p.iterChildren() if c is not u]
        x = xs[rnd.randrange(len(xs))]
        s = ss[rnd.randrange(len(ss))]
        _swapSubtrees(x, u, s, p)
        t.setPreAndPostOrder()
        for n in (u, p, x, s):
            self.touched.add(n)
        for n in list(u.iterChildren()) + list(p.iterChildren()) + [u, p]:
            _pushRelations(n)
            self.touched.add(n)
        self._markPath(t, u)
        self.topologyChanged = True

    # ---------------------------------------------------- accept / reject
    def sync(self, src, dst):
        """Make dst's Python and GPU state equal to src's after a
        proposal (p4/chain.py:1351-1531)."""
        # the touched set names nodes by number; src and dst have the same
        # node numbering
        for num in set(n.nodeNum for n in self.touched):
            n = src.nodes[num]
            m = dst.nodes[num]
            m.br.len = n.br.len
            m.parent = dst.nodes[n.parent.nodeNum] if n.parent else None
            m.leftChild = dst.nodes[n.leftChild.nodeNum] if n.leftChild else None
            m.sibling = dst.nodes[n.sibling.nodeNum] if n.sibling else None
            pf.p4_setBrLen(m.cNode, m.br.len)
            if self.topologyChanged:
                _pushRelations(m)
        if self.topologyChanged:
            dst.preOrder[:] = src.preOrder
            dst.postOrder[:] = src.postOrder
        for pNum in self.touchedParts:
            src.model.parts[pNum].copyValsTo(dst.model.parts[pNum])
            dst.model.parts[pNum].setCStuff(dst.model)
        # the partial slots and P blocks either tree changed since the last
        # sync: the proposal's recomputed path, not the whole tree
        pf.syncChanged(src.cTree, dst.cTree)
        if self.touchedParts:
            pf.p4_copyModelPrams(src.cTree, dst.cTree)
            for pNum in self.touchedParts:
                dst.model.parts[pNum].bQETneedsReset[:] = src.model.parts[pNum].bQETneedsReset
        dst.logLike = src.logLike


def _pushRelations(n):
    pf.p4_setNodeRelation(n.cNode, 0, n.parent.nodeNum if n.parent else -1)
    pf.p4_setNodeRelation(n.cNode, 1, n.leftChild.nodeNum if n.leftChild else -1)
    pf.p4_setNodeRelation(n.cNode, 2, n.sibling.nodeNum if n.sibling else -1)


def _relink(parent, old, new, newSibling):
    """In parent's child list put `new` where `old` was."""
    if parent.leftChild is old:
        parent.leftChild = new
    else:
        c = parent.leftChild
        while c.sibling is not old:
            c = c.sibling
        c.sibling = new
    new.sibling = newSibling
    new.parent = parent


def _swapSubtrees(x, u, s, p):
    """x (child of u) and s (child of p, sibling of u) trade places."""
    xs, ss = x.sibling, s.sibling
    _relink(u, x, s, xs)
    _relink(p, s, x, ss)


class McmcTunings(object):
    def __init__(self):
        self.default = {'brLen': 2.0 * math.log(2.0), 'allBrLens': 2.0 * math.log(1.02),
                        'compDir': 100.0, 'allCompsDir': 500.0, 'rMatrixDir': 200.0,
                        'gdasrv': 2.0 * math.log(1.5), 'pInvar': 0.5, 'nni': 1.0, 'learned': 1.0}


class Mcmc(object):
    """Mcmc(tree, nChains=4, ...).run(nGens) (p4/mcmc.py:672, 2496).

    chainTemps: temperatures T_i, heat beta_i = 1 / (1 + T_i); the default
    gives T = 0, 0.1, 0.2, ... as with p4's chainTempDiffs, untuned."""

    def __init__(self, aTree, nChains=4, runNum=0, sampleInterval=100, chainTemps=None,
                 seed=None, verbose=True, writeFiles=False, proposalWeights=None, simulate=None):
        if aTree.data is None or aTree.model is None:
            raise P4Error("Mcmc: the tree needs data and a model")
        self.rng = random.Random(seed)
        self.nprng = numpy.random.default_rng(seed)
        self.nChains = nChains
        self.runNum = runNum
        self.sampleInterval = sampleInterval
        self.verbose = verbose
        self.writeFiles = writeFiles
        self.chainTemps = list(chainTemps) if chainTemps is not None else [0.1 * i for i in range(nChains)]
        if len(self.chainTemps) != nChains:
            raise P4Error("Mcmc: chainTemps must have nChains values")
        self.brLenPriorLambda = 10.0
        self.tunings = McmcTunings()
        self.gen = -1
        aTree.calcLogLike(verbose=0)
        self.tree = aTree
        self.chains = []
        for i in range(nChains):
            t = aTree if i == 0 else aTree.dupe()
            t.precision = aTree.precision
            self.chains.append(Chain(self, i, t))
        self._nniCandidates = [n.nodeNum for n in aTree.iterInternalsNoRoot()]
        w = dict(brLen=1.0, nni=0.5 if self._nniCandidates else 0.0, allBrLens=0.0)
        for pNum, mp in enumerate(aTree.model.parts):
            if any(c.free for c in mp.comps):
                w['compDir.%d' % pNum] = 0.2
            if any(r.free and r.spec != '2p' for r in mp.rMatrices):
                w['rMatrixDir.%d' % pNum] = 0.2
            if any(g.free for g in mp.gdasrvs):
                w['gdasrv.%d' % pNum] = 0.1
            if mp.pInvar.free:
                w['pInvar.%d' % pNum] = 0.1
        if proposalWeights:
            w.update(proposalWeights)
        self.proposals = []
        for key, weight in w.items():
            if weight <= 0:
                continue
            name, _, p = key.partition('.')
            self.proposals.append(Proposal(name, weight, self.tunings.default[name], nChains,
                                           int(p) if p else -1))
        tot = sum(p.weight for p in self.proposals)
        self._cum = numpy.cumsum([p.weight / tot for p in self.proposals])
        self.learned = None
        self.swapMatrix = numpy.zeros((nChains, nChains), int)
        self.samples = dict(gen=[], lnL=[], treeLen=[], splits=[], prams=[])
        # callables f(tree) called with the cold chain's tree at every sample
        self.sampleHooks = []
        self._batch = pf.Batch([ch.propTree.cTree for ch in self.chains])
        self.tuneInterval = 250
        self.timing = dict(seconds=0.0, gens=0)
        # True: the likelihood ratio is left out of the acceptance ratio, so
        # the chain samples the prior (a check of the proposal ratios).
        self.samplePrior = False
        # CPO (p4/mcmc.py:918-921): with doCpo, every sample after
        # cpo_startGen adds 1/siteLike of the cold chain's tree to per-site
        # sums held in GPU memory by the data parts.
        self.doCpo = False
        self.cpo_startGen = 0
        for dp in aTree.data.parts:
            pf.cpoReset(dp.cPart)
        # Posterior-predictive simulation at every sample (p4/mcmc.py:936,
        # 2654-2657, 3461-3514): a bit mask of the statistics recorded for a
        # data set simulated on the cold chain's tree and model -- 1 the
        # unconstrained log likelihood, 2 the compositional chi-square,
        # 4 the mean number of states per site, 16 the constant sites, each
        # per part, into samples['sim'] rows [gen + 1, ...].
        self.simulate = int(simulate or 0)
        if self.simulate & 8:
            raise P4Error("Mcmc: simulate bit 8 (compStatFromCharFreqs) is not provided")
        self.simTree = None
        self.samples['sim'] = []
        if self.simulate:
            self.simTree = aTree.dupe()
            self.simTree._data = aTree.data.dupe()
            self.simTree.calcLogLike(verbose=0)

    def addLearnedProposal(self, learned, weight):
        """Add the learned joint branch-length move (p4gpu.learned)."""
        self.learned = learned
        self.proposals.append(Proposal('learned', weight, 1.0, self.nChains))
        tot = sum(p.weight for p in self.proposals)
        self._cum = numpy.cumsum([p.weight / tot for p in self.proposals])

    def coldChain(self):
        for ch in self.chains:
            if ch.tempNum == 0:
                return ch

    def _choose(self):
        return self.proposals[int(numpy.searchsorted(self._cum, self.rng.random(), side='right'))
                              if self._cum[-1] > 0 else 0]

    def run(self, nGens, verbose=None):
        verbose = self.verbose if verbose is None else verbose
        t0 = time.perf_counter()
        for g in range(nGens):
            self.gen += 1
            props = []
            for ch in self.chains:
                prop = self._choose()
                if prop.name == 'learned':
                    self.learned.propose(ch, prop)
                else:
                    ch.propose(prop)
                props.append(prop)
            lnLs = self._batch.logLikes()
            for ch, prop, lnl in zip(self.chains, props, lnLs):
                ch.propTree.logLike = lnl
                beta = 1.0 / (1.0 + self.chainTemps[ch.tempNum])
                pRet = heatedLogRatio(beta, lnl, ch.curTree.logLike, ch.logPriorRatio, ch.logProposalRatio,
                                      self.samplePrior)
                r = acceptProbability(pRet)
                accept = r == 1.0 or self.rng.random() < r
                prop.nProposals[ch.tempNum] += 1
                prop.tnNSamples[ch.tempNum] += 1
                if accept:
                    prop.nAcceptances[ch.tempNum] += 1
                    prop.tnNAccepts[ch.tempNum] += 1
                    ch.sync(ch.propTree, ch.curTree)
                else:
                    ch.sync(ch.curTree, ch.propTree)
                if self.learned is not None and prop.name == 'learned':
                    self.learned.afterMove(ch, accept)
            if self.nChains > 1:
                self._swap()
            if self.gen % self.tuneInterval == 0 and self.gen > 0:
                for prop in self.proposals:
                    for ti in range(self.nChains):
                        prop.tune(ti)
            if self.gen % self.sampleInterval == 0:
                self._sample()
        el = time.perf_counter() - t0
        self.timing['seconds'] += el
        self.timing['gens'] += nGens
        if verbose:
            print("Mcmc.run(): %d generations in %.2f s (%.0f gens/s)" % (nGens, el, nGens / el))

    def _swap(self):
        """Adjacent-temperature swap (p4/mcmc.py:3299-3420)."""
        t1 = self.rng.randrange(self.nChains - 1)
        t2 = t1 + 1
        c1 = next(ch for ch in self.chains if ch.tempNum == t1)
        c2 = next(ch for ch in self.chains if ch.tempNum == t2)
        self.swapMatrix[t1][t2] += 1
        b1 = 1.0 / (1.0 + self.chainTemps[t1])
        b2 = 1.0 / (1.0 + self.chainTemps[t2])
        lnR = swapLogRatio(b1, b2, c1.curTree.logLike, c2.curTree.logLike)
        # p4/mcmc.py:3324-3332: r = 0 below -100, 1 at or above 0
        r = acceptProbability(lnR)
        if self.rng.random() < r:
            self.swapMatrix[t2][t1] += 1
            c1.tempNum, c2.tempNum = t2, t1

    def _sample(self):
        ch = self.coldChain()
        t = ch.curTree
        self.samples['gen'].append(self.gen)
        self.samples['lnL'].append(t.logLike)
        self.samples['treeLen'].append(t.getLen())
        self.samples['splits'].append(splitKeys(t))
        pr = []
        for mp in t.model.parts:
            for c in mp.comps:
                pr.extend(c.val.tolist())
            for g in mp.gdasrvs:
                pr.append(float(g.val[0]))
            pr.append(mp.pInvar.val)
        self.samples['prams'].append(pr)
        if self.doCpo and self.gen + 1 > self.cpo_startGen:
            pf.cpoAccumulate(t.cTree)
        if self.simulate:
            self._doSimulate(t)
        for f in self.sampleHooks:
            f(t)

    def _doSimulate(self, curTree):
        """p4/mcmc.py:3461-3514: simulate on the cold chain's tree and
        model, then the statistics of the simulate bit mask."""
        curTree.copyToTree(self.simTree)
        curTree.model.copyValsTo(self.simTree.model)
        self.simTree.simulate(resetSequences=False, seed=self.rng.getrandbits(64))
        d = self.simTree.data
        row = [self.gen + 1]
        if 1 & self.simulate:
            row += [pf.getUnconstrainedLogLike(p.cPart) for p in d.parts]
        if 2 & self.simulate:
            row += d.simpleBigXSquared()
        if 4 & self.simulate:
            row += d.meanNCharsPerSite()
        if 16 & self.simulate:
            row += d.simpleConstantSitesCount()
        self.samples['sim'].append(row)

    # ------------------------------------------------------------- CPO
    @property
    def cpo_nSamples(self):
        return pf.cpoGetSums(self.tree.data.parts[0].cPart)[1]

    @property
    def cpo_sumsOfInverseSiteLikes(self):
        """Per site, all parts in order: the sum over CPO samples of
        1/siteLike (p4/mcmc.py:3292-3296)."""
        return numpy.concatenate([pf.cpoGetSums(dp.cPart)[0] for dp in self.tree.data.parts])

    def lpml(self, verbose=True):
        """Log pseudo-marginal likelihood from the CPO sums
        (p4/mcmc.py:3730-3748): sum over sites of -log(mean of 1/siteLike)."""
        if not self.doCpo:
            print("Mcmc.lpml().  doCpo was not on.")
            return 0.0
        n = self.cpo_nSamples
        if not n:
            print("Mcmc.lpml().  no cpo_nSamples?")
            return 0.0
        a = self.cpo_sumsOfInverseSiteLikes / n
        v = float((-numpy.log(a)).sum())
        if verbose:
            print(f"cpo_nSamples: {n}, lpml: {v}")
        return v

    # ------------------------------------------------------------ judges
    def splitFrequencies(self, burnin=0):
        ss = self.samples['splits'][burnin:]
        counts = {}
        for s in ss:
            for k in s:
                counts[k] = counts.get(k, 0) + 1
        n = float(len(ss))
        return {k: v / n for k, v in counts.items()}

    def ess(self, key='lnL', burnin=0):
        x = numpy.asarray(self.samples[key][burnin:], dtype=float)
        return effectiveSampleSize(x)

    def acceptanceRates(self):
        out = {}
        for p in self.proposals:
            n = p.nProposals[0]
            out[p.name + ('.%d' % p.pNum if p.pNum >= 0 else '')] = (p.nAcceptances[0] / n) if n else None
        return out


def splitKeys(tree):
    """Splits of an unrooted tree as integer bit sets over seqNums, each
    normalised to exclude taxon 0."""
    nTax = len(tree.taxNames)
    full = (1 << nTax) - 1
    index = {name: i for i, name in enumerate(tree.taxNames)}

    def bit(n):
        return 1 << (n.seqNum if n.seqNum >= 0 else index[n.name])
    below = {}
    keys = []
    for n in tree.iterPostOrder():
        if n.isLeaf and n is not tree.root:
            below[n.nodeNum] = bit(n)
        else:
            b = 0
            for c in n.iterChildren():
                b |= below[c.nodeNum]
            if n.isLeaf:
                b |= bit(n)
            below[n.nodeNum] = b
            if n is not tree.root and n.parent is not None:
                k = b if not (b & 1) else full ^ b
                if 1 < bin(k).count("1") < nTax - 1:
                    keys.append(k)
    return frozenset(keys)


def compareSplitFrequencies(fa, fb, minFreq=0.1):
    """Maximum and average standard deviation of split frequencies between
    two runs (the ASDOSS / max-difference criterion of
    p4/func.py:3471 and p4/mcmccheckpointreader.py:146-189)."""
    keys = set(k for k, v in fa.items() if v >= minFreq) | set(k for k, v in fb.items() if v >= minFreq)
    if not keys:
        return 0.0, 0.0
    diffs = [abs(fa.get(k, 0.0) - fb.get(k, 0.0)) for k in keys]
    sds = [numpy.std([fa.get(k, 0.0), fb.get(k, 0.0)]) for k in keys]
    return max(diffs), float(numpy.mean(sds))
