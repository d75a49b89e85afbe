"""Learned joint proposals for Mcmc: the gamma-shape / tree-length ridge,
blocks of branch lengths, and compositions.

The reference's proposals are hand-designed (p4/chain.py:3091-3196 for
branch lengths, 1798-1866 and 2259-2289 for compositions, 535-545 for the
gamma shape).  This module adds moves that are learned: a 3-layer MLP
(p4g_mlp*, BF16 weights on the MFMA units, FP32 accumulation; trained on
the GPU by the library's own forward/backward/Adam kernels) maps per-row
features to Gaussians.  Row types:

  branch (edge above node n)  log brLen'   features: current log brLen,
                              log tree length, log gamma shape, Fitch
                              parsimony summaries of the child's state sets
                              (changes per site on the edge, ambiguous and
                              single-state fractions of the down-pass sets:
                              0/1 partials), clade size, leaf flag, depth,
                              child count, the node's composition assignment
                              (one-hot) and that composition's values
  composition c               alr(pi')     log tree length, log gamma
                              shape, assignment one-hot, empirical comp
  gamma                       log alpha'   empirical composition only
  tree                        log T'       log alpha it is conditioned on
  (alr(pi) = log(pi[:3] / pi[3]); every row also carries the empirical
  composition.)

Moves, chosen with state-independent probabilities:
  'ridge'  the gamma shape and a common scale of all branch lengths: alpha'
           from the gamma row, then the tree length T' from the tree row
           given alpha', every branch multiplied by c = T'/T.  Tree length
           and gamma shape lie on a narrow posterior ridge that one-branch
           and one-parameter moves cross slowly.  In z = (log alpha,
           log b_1..log b_n) the move is a shift with density
           q_a(log alpha') q_T(log T' | alpha'); its reverse draws
           (log alpha, log T); the shift has unit Jacobian, so
           log r = log pi(x')/pi(x) + log(alpha'/alpha) + n log c
                   + log q_a(log alpha) + log q_T(log T | alpha)
                   - log q_a(log alpha') - log q_T(log T' | alpha').
  'block'  a random block of branches, jointly, each log brLen' from its
           row (features at x); the reverse density uses the rows at x':
           q(b') = prod N(log b'_i) / b'_i.
  'comp'   every free composition of the part, jointly; its rows do not
           depend on the compositions, so forward and reverse use one pass:
           q(pi') = N(alr pi') / prod_j pi'_j.
The branch-length prior ratio is included; compositions and the gamma
shape have flat priors.  Proposals outside the parameter bounds are rejected
(the target is zero there).  So the stationary distribution is the
posterior whatever the network outputs; training only changes how well the
chain mixes.

Training data: samples of the cold chain of a plain-proposal run
(Mcmc.sampleHooks).  Each training row pairs the features of sample i with
the value in a random sample j of the same topology (the same split for
branch rows; branch rows conditioned on sample j's gamma shape, tree rows on
the alpha of the same sample j), so the network learns posterior marginals
given the conditioning values, and the moves behave like independence
samplers.
"""
import math

import numpy

from . import pf
from .p4exceptions import P4Error
from .var import var

NFEAT = 24
NCOORD = 3          # output coordinates: means [0:3], log sds [3:6]
MAX_COMPS = 4
BRANCH, COMP, GAMMA, TREE = 0, 1, 2, 3
_LOG2PI = math.log(2.0 * math.pi)


def _popcount8(a):
    a = a - ((a >> 1) & 0x55)
    a = (a & 0x33) + ((a >> 2) & 0x33)
    return (a + (a >> 4)) & 0x0F


def _alr(p):
    p = numpy.asarray(p, dtype=numpy.float64)
    return numpy.log(p[..., :3] / p[..., 3:4])


def _alrInv(z):
    z = numpy.asarray(z, dtype=numpy.float64)
    e = numpy.exp(numpy.concatenate([z, numpy.zeros(z.shape[:-1] + (1,))], axis=-1))
    return e / e.sum(axis=-1, keepdims=True)


def _logNormal(y, mu, sd):
    r = (y - mu) / sd
    return float((-0.5 * r * r - numpy.log(sd) - 0.5 * _LOG2PI).sum())


def edgeKey(bits, full):
    """Split of an edge as a bit set over taxa, normalised to exclude
    taxon 0 (leaf edges included)."""
    return bits if not (bits & 1) else full ^ bits


class _Topo(object):
    """Static per-topology data: branch order, static feature columns,
    composition index per branch, split keys."""
    __slots__ = ('nums', 'static', 'compIdx', 'splits', 'splitIndex')


class LearnedProposal(object):
    """The learned moves of one DNA part.  weights: relative probabilities
    of 'ridge', 'block' and 'comp' (moves whose parameters are not free
    are dropped)."""

    def __init__(self, tree, pNum=0, hidden=256, seed=0, weights=None, blockSize=4, scale=1.0):
        if tree.data is None or tree.model is None:
            raise P4Error("LearnedProposal: the tree needs data and a model")
        mp = tree.model.parts[pNum]
        if mp.dim != 4:
            raise P4Error("LearnedProposal: DNA (4-state) parts only")
        if mp.nComps > MAX_COMPS:
            raise P4Error("LearnedProposal: at most %d compositions" % MAX_COMPS)
        if mp.nGdasrvs > 1:
            raise P4Error("LearnedProposal: at most one gamma shape per part")
        self.pNum = pNum
        self.hasComp = any(c.free for c in mp.comps)
        self.hasGamma = mp.nGdasrvs == 1 and bool(mp.gdasrvs[0].free) and mp.nGammaCat > 1
        w = dict(ridge=0.4, block=0.3, comp=0.3)
        if weights:
            w.update(weights)
        if not self.hasGamma:
            w['ridge'] = 0.0
        if not self.hasComp:
            w['comp'] = 0.0
        tot = sum(w.values())
        if tot <= 0:
            raise P4Error("LearnedProposal: no move has a positive weight")
        self.kinds = [k for k in ('ridge', 'block', 'comp') if w[k] > 0]
        self.cum = numpy.cumsum([w[k] / tot for k in self.kinds])
        self.blockSize = blockSize
        self.scale = scale
        self.rng = numpy.random.default_rng(seed)
        self.net = pf.Mlp(NFEAT, hidden, 2 * NCOORD, seed)
        self.trained = False
        self.xMean = numpy.zeros(NFEAT)
        self.xSd = numpy.ones(NFEAT)
        # output scalings per row type
        self.yMean = numpy.zeros((4, NCOORD))
        self.ySd = numpy.ones((4, NCOORD))
        self._gammaOut = None
        part = tree.data.parts[pNum]
        pats, counts = part.patterns()[:2]
        codes = numpy.asarray(pats)
        # leaf state sets as bit masks; gaps, missing and ambiguity codes
        # count as the full set
        sets = numpy.where(codes >= 0, numpy.left_shift(1, numpy.clip(codes, 0, 7)), 0x0F)
        self.leafSets = sets.astype(numpy.uint8)
        self.w = numpy.asarray(counts, dtype=numpy.float64)
        self.nChar = float(self.w.sum())
        emp = numpy.array([(self.w * (codes == s)).sum() for s in range(4)], dtype=float)
        emp = emp / max(emp.sum(), 1.0)
        self.empAlr = _alr(numpy.maximum(emp, 1e-3))
        self.taxIndex = {name: i for i, name in enumerate(tree.taxNames)}
        self.nTax = len(tree.taxNames)
        self._topo = {}
        self.records = []
        self.stats = {}
        for k in ('ridge', 'block', 'comp'):
            self.stats[k] = 0
            self.stats[k + 'Accepted'] = 0
        self.stats.update(proposed=0, accepted=0, outOfBounds=0)
        self._lastKind = None
        self.trainLog = []

    # ------------------------------------------------------------ features
    def _key(self, tree):
        return (tuple(n.parent.nodeNum if n.parent is not None else -1 for n in tree.nodes),
                tuple(n.parts[self.pNum].compNum for n in tree.nodes))

    def _topology(self, tree):
        key = self._key(tree)
        got = self._topo.get(key)
        if got is not None:
            return got
        nN = len(tree.nodes)
        down = [None] * nN
        bits = [0] * nN
        size = [0] * nN
        for n in tree.iterPostOrder():
            kids = sorted(n.iterChildren(), key=lambda c: c.nodeNum)
            if n.isLeaf:
                si = n.seqNum if n.seqNum >= 0 else self.taxIndex[n.name]
                s = self.leafSets[si].copy()
                b, sz = 1 << si, 1
            else:
                s, b, sz = None, 0, 0
            for c in kids:
                if s is None:
                    s = down[c.nodeNum].copy()
                else:
                    inter = s & down[c.nodeNum]
                    s = numpy.where(inter != 0, inter, s | down[c.nodeNum])
                b |= bits[c.nodeNum]
                sz += size[c.nodeNum]
            down[n.nodeNum] = s
            bits[n.nodeNum] = b
            size[n.nodeNum] = sz
        final = [None] * nN
        depth = [0] * nN
        changes = numpy.zeros(nN)
        for n in tree.iterPreOrder():
            if n.parent is None:
                final[n.nodeNum] = down[n.nodeNum]
                continue
            inter = down[n.nodeNum] & final[n.parent.nodeNum]
            empty = inter == 0
            final[n.nodeNum] = numpy.where(empty, down[n.nodeNum], inter)
            changes[n.nodeNum] = float((self.w * empty).sum()) / self.nChar
            depth[n.nodeNum] = depth[n.parent.nodeNum] + 1
        full = (1 << self.nTax) - 1
        nodes = [n for n in tree.iterNodesNoRoot()]
        T = _Topo()
        T.nums = numpy.array([n.nodeNum for n in nodes], dtype=numpy.int64)
        T.static = numpy.zeros((len(nodes), NFEAT))
        T.compIdx = numpy.array([n.parts[self.pNum].compNum for n in nodes], dtype=numpy.int64)
        T.splits = [edgeKey(bits[n.nodeNum], full) for n in nodes]
        T.splitIndex = {k: i for i, k in enumerate(T.splits)}
        for i, n in enumerate(nodes):
            pc = _popcount8(down[n.nodeNum].astype(numpy.int32))
            x = T.static[i]
            x[BRANCH] = 1.0
            x[7] = changes[n.nodeNum]
            x[8] = float((self.w * (pc > 1)).sum()) / self.nChar
            x[9] = float((self.w * (pc == 1)).sum()) / self.nChar
            x[10] = size[n.nodeNum] / float(self.nTax)
            x[11] = 1.0 if n.isLeaf else 0.0
            x[12] = depth[n.nodeNum] / 10.0
            x[13] = math.log(changes[n.nodeNum] + 1e-3)
            x[14] = sum(1 for _ in n.iterChildren()) / 3.0
            x[15 + T.compIdx[i]] = 1.0
            x[21:24] = self.empAlr
        if len(self._topo) > 20000:
            self._topo.clear()
        self._topo[key] = T
        return T

    # column layout: 0-3 row type; 4 log brLen; 5 log tree length; 6 log
    # gamma shape (current, or the one conditioned on); 7-14 topology
    # features; 15-18 composition one-hot; 19-20 and 21-23 composition
    # values (alr[0:2] of the node's comp) and the empirical alr
    def _branchRows(self, T, lens, treeLen, comps, logAlpha, sel=None):
        if sel is None:
            sel = numpy.arange(len(lens))
        X = T.static[sel].copy()
        X[:, 4] = numpy.log(lens[sel])
        X[:, 5] = math.log(treeLen)
        X[:, 6] = logAlpha
        X[:, 19:21] = _alr(comps)[T.compIdx[sel]][:, :2]
        return X

    def _compRows(self, compNums, treeLen, logAlpha):
        X = numpy.zeros((len(compNums), NFEAT))
        X[:, COMP] = 1.0
        X[:, 5] = math.log(treeLen)
        X[:, 6] = logAlpha
        for r, cn in enumerate(compNums):
            X[r, 15 + cn] = 1.0
        X[:, 21:24] = self.empAlr
        return X

    def _gammaRow(self):
        X = numpy.zeros((1, NFEAT))
        X[0, GAMMA] = 1.0
        X[0, 21:24] = self.empAlr
        return X

    def _treeRows(self, logAlphas):
        X = numpy.zeros((len(logAlphas), NFEAT))
        X[:, TREE] = 1.0
        X[:, 6] = logAlphas
        X[:, 21:24] = self.empAlr
        return X

    def _gauss(self, X, kinds):
        """Means and sds in the original coordinates for feature rows X of
        row types kinds."""
        out = self.net.forward((X - self.xMean) / self.xSd).astype(numpy.float64)
        mu = out[:, :NCOORD]
        ls = numpy.clip(out[:, NCOORD:], -8.0, 4.0)
        yM, yS = self.yMean[kinds], self.ySd[kinds]
        return yM + yS * mu, yS * numpy.exp(ls) * self.scale

    def _gammaGauss(self):
        # the gamma row's features are constant: one pass per network
        if self._gammaOut is None:
            mu, sd = self._gauss(self._gammaRow(), numpy.array([GAMMA]))
            self._gammaOut = (float(mu[0, 0]), float(sd[0, 0]))
        return self._gammaOut

    # ------------------------------------------------------------ state
    def _state(self, tree):
        mp = tree.model.parts[self.pNum]
        T = self._topology(tree)
        lens = numpy.array([tree.nodes[k].br.len for k in T.nums])
        comps = numpy.array([c.val for c in mp.comps])
        alpha = float(mp.gdasrvs[0].val[0]) if mp.nGdasrvs == 1 else 1.0
        return T, lens, comps, alpha

    def recordState(self, tree):
        """Sample hook: remember the cold chain's state for training."""
        T, lens, comps, alpha = self._state(tree)
        self.records.append((self._key(tree), frozenset(T.splits), lens, comps.copy(), alpha))

    def attach(self, mcmc):
        """Record the cold chain at every sample of mcmc."""
        mcmc.sampleHooks.append(self.recordState)

    def _trainingSet(self, minGroup=20):
        """(X, Y, mask, kind) with a fresh random partner per sample."""
        groups = {}
        for i, r in enumerate(self.records):
            groups.setdefault(r[1], []).append(i)
        Xs, Ys, Ms, Ks = [], [], [], []

        def add(X, Y, M, kind):
            Xs.append(X)
            Ys.append(Y)
            Ms.append(M)
            Ks.append(numpy.full(len(X), kind, numpy.int64))
        for key, idx in groups.items():
            if len(idx) < minGroup:
                continue
            partner = self.rng.permutation(idx)
            for i, j in zip(idx, partner):
                ki, _, li, ci, ai = self.records[i]
                kj, _, lj, cj, aj = self.records[j]
                Ti, Tj = self._topo[ki], self._topo[kj]
                perm = numpy.array([Tj.splitIndex[s] for s in Ti.splits])
                Y = numpy.zeros((len(li), NCOORD))
                Y[:, 0] = numpy.log(lj[perm])
                M = numpy.zeros_like(Y)
                M[:, 0] = 1.0
                add(self._branchRows(Ti, li, li.sum(), ci, math.log(aj)), Y, M, BRANCH)
                if self.hasComp:
                    cns = numpy.arange(len(ci))
                    add(self._compRows(cns, li.sum(), math.log(ai)), _alr(cj), numpy.ones((len(ci), NCOORD)), COMP)
                if self.hasGamma:
                    one = numpy.array([[1.0, 0.0, 0.0]])
                    add(self._gammaRow(), numpy.array([[math.log(aj), 0.0, 0.0]]), one, GAMMA)
                    add(self._treeRows([math.log(aj)]), numpy.array([[math.log(lj.sum()), 0.0, 0.0]]), one, TREE)
        if not Xs:
            raise P4Error("LearnedProposal: no topology has %d or more samples" % minGroup)
        return numpy.concatenate(Xs), numpy.concatenate(Ys), numpy.concatenate(Ms), numpy.concatenate(Ks)

    def train(self, epochs=30, batch=4096, lr=1e-3, minGroup=20, verbose=False):
        """Fit the network to the recorded samples on the GPU (Adam on the
        Gaussian negative log likelihood)."""
        X, Y, M, K = self._trainingSet(minGroup)
        self.xMean = X.mean(axis=0)
        self.xSd = X.std(axis=0)
        self.xSd[self.xSd < 1e-6] = 1.0
        for k in range(4):
            sel = K == k
            if sel.any():
                self.yMean[k] = Y[sel].mean(axis=0)
                self.ySd[k] = numpy.maximum(Y[sel].std(axis=0), 1e-3)
        self.trainLog = []
        for ep in range(epochs):
            if ep > 0:
                X, Y, M, K = self._trainingSet(minGroup)
            Yn = (Y - self.yMean[K]) / self.ySd[K]
            Xn = (X - self.xMean) / self.xSd
            order = self.rng.permutation(len(X))
            tot = 0.0
            for s in range(0, len(order), batch):
                sel = order[s:s + batch]
                tot += self.net.trainStep(Xn[sel], Yn[sel], M[sel], lr)
            self.trainLog.append(tot / M.sum())
            if verbose:
                print("epoch %d  loss/coord %.4f" % (ep, self.trainLog[-1]), flush=True)
        self._gammaOut = None
        self.trained = True
        return self.trainLog

    def save(self, path):
        """Write the trained network and its input/output scalings (.npz)."""
        numpy.savez(path, params=self.net.getParams(), xMean=self.xMean, xSd=self.xSd, yMean=self.yMean,
                    ySd=self.ySd, hidden=numpy.int64(self.net.nHidden), trainLog=numpy.asarray(self.trainLog))

    def load(self, path):
        with numpy.load(path, allow_pickle=False) as z:
            if int(z['hidden']) != self.net.nHidden:
                raise P4Error("LearnedProposal.load: hidden width %d, expected %d"
                              % (int(z['hidden']), self.net.nHidden))
            self.net.setParams(z['params'])
            for k in ('xMean', 'xSd', 'yMean', 'ySd'):
                setattr(self, k, numpy.array(z[k], dtype=numpy.float64))
            self.trainLog = list(z['trainLog'])
        self._gammaOut = None
        self.trained = True

    # ------------------------------------------------------------ moves
    def propose(self, ch, prop):
        """Mutate ch.propTree (equal to ch.curTree on entry) and set the
        chain's log prior and proposal ratios."""
        ch.logPriorRatio = 0.0
        ch.logProposalRatio = 0.0
        ch.touched = set()
        ch.touchedParts = set()
        ch.topologyChanged = False
        if not self.trained:
            raise P4Error("LearnedProposal.propose: train() first")
        self.stats['proposed'] += 1
        kind = self.kinds[int(numpy.searchsorted(self.cum, self.rng.random(), side='right'))
                          if len(self.kinds) > 1 else 0]
        self._lastKind = kind
        self.stats[kind] += 1
        getattr(self, '_' + kind)(ch)

    def _reject(self, ch):
        self.stats['outOfBounds'] += 1
        ch.logProposalRatio = -1.0e300

    def _setLens(self, ch, T, new):
        t = ch.propTree
        for k, num in enumerate(T.nums):
            n = t.nodes[num]
            n.br.len = float(new[k])
            pf.p4_setBrLen(n.cNode, n.br.len)
            ch.touched.add(n)

    def _ridge(self, ch):
        t = ch.propTree
        mp = t.model.parts[self.pNum]
        T, lens, comps, alpha = self._state(t)
        la = math.log(alpha)
        tl = float(lens.sum())
        muA, sdA = self._gammaGauss()
        la2 = muA + sdA * self.rng.standard_normal()
        mu, sd = self._gauss(self._treeRows([la2, la]), numpy.array([TREE, TREE]))
        lt2 = mu[0, 0] + sd[0, 0] * self.rng.standard_normal()
        logc = lt2 - math.log(tl)
        new = lens * math.exp(logc)
        alpha2 = math.exp(la2)
        if new.min() < var.BRLEN_MIN or new.max() > var.BRLEN_MAX or \
                not (var.GAMMA_SHAPE_MIN <= alpha2 <= var.GAMMA_SHAPE_MAX):
            self._reject(ch)
            return
        logQf = _logNormal(la2, muA, sdA) + _logNormal(lt2, mu[0, 0], sd[0, 0])
        logQr = _logNormal(la, muA, sdA) + _logNormal(math.log(tl), mu[1, 0], sd[1, 0])
        ch.logProposalRatio = logQr - logQf + (la2 - la) + len(lens) * logc
        ch.logPriorRatio = ch.mcmc.brLenPriorLambda * (tl - float(new.sum()))
        self._setLens(ch, T, new)
        mp.gdasrvs[0].val = alpha2
        ch._modelChanged(self.pNum)
        # every branch length changed: the other parts' P matrices and
        # partials are stale too (their model parameters are not)
        for p in range(t.model.nParts):
            if p != self.pNum:
                pf.p4_setPrams(t.cTree, p)
                ch._markAll(t, p)

    def _block(self, ch):
        t = ch.propTree
        T, lens, comps, alpha = self._state(t)
        la = math.log(alpha)
        k = min(self.blockSize, len(lens))
        sel = numpy.sort(self.rng.choice(len(lens), size=k, replace=False))
        kinds = numpy.zeros(k, numpy.int64)
        mu, sd = self._gauss(self._branchRows(T, lens, float(lens.sum()), comps, la, sel), kinds)
        y = mu[:, 0] + sd[:, 0] * self.rng.standard_normal(k)
        new = lens.copy()
        new[sel] = numpy.exp(y)
        if new[sel].min() < var.BRLEN_MIN or new[sel].max() > var.BRLEN_MAX:
            self._reject(ch)
            return
        logQf = _logNormal(y, mu[:, 0], sd[:, 0]) - float(y.sum())
        mu2, sd2 = self._gauss(self._branchRows(T, new, float(new.sum()), comps, la, sel), kinds)
        yOld = numpy.log(lens[sel])
        logQr = _logNormal(yOld, mu2[:, 0], sd2[:, 0]) - float(yOld.sum())
        ch.logProposalRatio = logQr - logQf
        ch.logPriorRatio = ch.mcmc.brLenPriorLambda * float(lens[sel].sum() - new[sel].sum())
        t2 = ch.propTree
        for q in sel:
            n = t2.nodes[T.nums[q]]
            n.br.len = float(new[q])
            pf.p4_setBrLen(n.cNode, n.br.len)
            pf.p4_calculateBigPDecks(n.cNode)
            ch._markPath(t2, n.parent)
            ch.touched.add(n)

    def _comp(self, ch):
        t = ch.propTree
        mp = t.model.parts[self.pNum]
        T, lens, comps, alpha = self._state(t)
        cns = numpy.array([c.num for c in mp.comps if c.free], dtype=numpy.int64)
        kinds = numpy.full(len(cns), COMP, numpy.int64)
        mu, sd = self._gauss(self._compRows(cns, float(lens.sum()), math.log(alpha)), kinds)
        y = mu + sd * self.rng.standard_normal(mu.shape)
        pi = _alrInv(y)
        if pi.min() < var.PIVEC_MIN or pi.max() > var.PIVEC_MAX:
            self._reject(ch)
            return
        old = comps[cns]
        logQf = _logNormal(y, mu, sd) - float(numpy.log(pi).sum())
        logQr = _logNormal(_alr(old), mu, sd) - float(numpy.log(old).sum())
        ch.logProposalRatio = logQr - logQf
        for r, cn in enumerate(cns):
            mp.comps[cn].val[:] = pi[r]
        ch._modelChanged(self.pNum)

    def afterMove(self, ch, accept):
        if accept:
            self.stats['accepted'] += 1
            if self._lastKind:
                self.stats[self._lastKind + 'Accepted'] += 1
