"""Learned joint branch-length / composition proposal for Mcmc.

The reference's proposals are hand-designed (p4/chain.py:3091-3196 for
branch lengths, 1798-1866 and 2259-2289 for compositions).  This module
adds one move that is learned: a 3-layer MLP (p4g_mlp*, BF16 weights on the
MFMA units, FP32 accumulation; trained on the GPU by the library's own
forward/backward/Adam kernels) maps per-branch and per-composition features
to a Gaussian:

  branch row (edge above node n):  log brLen'  ~ N(mu, sigma)
  composition row (comp c):         alr(pi')    ~ N(mu[0:3], diag sigma[0:3])
                                    alr(pi) = log(pi[:3] / pi[3])

Features of a branch row: its current log length, the current log tree
length, Fitch parsimony summaries of the child's state sets (changes per
site on the edge, ambiguous and single-state fractions of the child's
down-pass sets -- a 0/1 "partial"), clade size, leaf flag, depth, and the
node's composition assignment (one-hot) with that composition's values.
A composition row carries the current and the empirical composition.

One move updates a random block of `blockSize` branches and, with
probability `compProb`, the part's free compositions, all jointly.  It is
Metropolis-Hastings corrected: the reverse density q(x | x') is evaluated
with a second forward pass on the features of x', and

  log r = beta (lnL' - lnL) + log prior ratio + log q(x | x') - log q(x' | x),

with the Jacobians of the log and additive-log-ratio maps:
q(b') = N(log b') / b'  and  q(pi') = N(alr pi') / prod_j pi'_j.
Proposals outside [BRLEN_MIN, BRLEN_MAX] or with a composition value below
PIVEC_MIN are rejected (the target is zero there).  The block and the
decision to include compositions do not depend on the state, so their
probabilities cancel.  The stationary distribution is therefore the
posterior whatever the network outputs; training only changes how well the
chain mixes.

Training data: samples of the cold chain of a plain-proposal run
(Mcmc.sampleHooks).  Each training row pairs the features of sample i with
the value of the same split (or composition) in a random sample j of the
same topology, so the network learns the posterior marginal of each branch
given its context, and the move behaves like a blocked independence
sampler.
"""
import math

import numpy

from . import pf
from .p4exceptions import P4Error
from .var import var

NFEAT = 24
NCOORD = 3          # output coordinates: means [0:3], log sds [3:6]
MAX_COMPS = 4


def _popcount8(a):
    a = a - ((a >> 1) & 0x55)
    a = (a & 0x33) + ((a >> 2) & 0x33)
    return (a + (a >> 4)) & 0x0F


def _alr(p):
    p = numpy.asarray(p, dtype=numpy.float64)
    return numpy.log(p[:3] / p[3])


def _alrInv(z):
    e = numpy.exp(numpy.concatenate([z, [0.0]]))
    return e / e.sum()


def _logNormal(y, mu, sd):
    r = (y - mu) / sd
    return float((-0.5 * r * r - numpy.log(sd) - 0.5 * math.log(2.0 * math.pi)).sum())


def edgeKey(bits, full):
    """Split of an edge as a bit set over taxa, normalised to exclude
    taxon 0 (leaf edges included)."""
    return bits if not (bits & 1) else full ^ bits


class LearnedBrLenCompProposal(object):
    def __init__(self, tree, pNum=0, hidden=256, seed=0, blockSize=6, compProb=0.3, scale=1.0):
        if tree.data is None or tree.model is None:
            raise P4Error("LearnedBrLenCompProposal: the tree needs data and a model")
        mp = tree.model.parts[pNum]
        if mp.dim != 4:
            raise P4Error("LearnedBrLenCompProposal: DNA (4-state) parts only")
        if mp.nComps > MAX_COMPS:
            raise P4Error("LearnedBrLenCompProposal: at most %d compositions" % MAX_COMPS)
        self.pNum = pNum
        self.blockSize = blockSize
        self.compProb = compProb if any(c.free for c in mp.comps) else 0.0
        self.scale = scale
        self.rng = numpy.random.default_rng(seed)
        self.net = pf.Mlp(NFEAT, hidden, 2 * NCOORD, seed)
        self.trained = False
        self.xMean = numpy.zeros(NFEAT)
        self.xSd = numpy.ones(NFEAT)
        self.yMean = numpy.zeros(NCOORD)
        self.ySd = numpy.ones(NCOORD)
        self.yMeanComp = numpy.zeros(NCOORD)
        self.ySdComp = numpy.ones(NCOORD)
        part = tree.data.parts[pNum]
        pats, counts = part.patterns()[:2]
        # leaf state sets as bit masks; gaps, missing and ambiguity codes
        # count as the full set
        codes = numpy.asarray(pats)
        sets = numpy.where(codes >= 0, numpy.left_shift(1, numpy.clip(codes, 0, 7)), 0x0F)
        self.leafSets = sets.astype(numpy.uint8)
        self.w = numpy.asarray(counts, dtype=numpy.float64)
        self.nChar = float(self.w.sum())
        emp = numpy.zeros(4)
        for s in range(4):
            emp[s] = (self.w * (codes == s)).sum()
        emp = emp / max(emp.sum(), 1.0)
        self.empAlr = _alr(numpy.maximum(emp, 1e-3))
        self.taxIndex = {name: i for i, name in enumerate(tree.taxNames)}
        self.nTax = len(tree.taxNames)
        self._topo = {}
        self.records = []
        self.stats = dict(proposed=0, accepted=0, outOfBounds=0)
        self.trainLog = []

    # ------------------------------------------------------------ features
    def _structKey(self, tree):
        return tuple(n.parent.nodeNum if n.parent is not None else -1 for n in tree.nodes)

    def _topology(self, tree):
        """Per-node topology features (Fitch summaries, clade size, depth)
        and edge split keys, cached by the tree's parent structure."""
        key = self._structKey(tree)
        got = self._topo.get(key)
        if got is not None:
            return got
        nN = len(tree.nodes)
        down = [None] * nN
        bits = [0] * nN
        size = [0] * nN
        post = list(tree.iterPostOrder())
        for n in post:
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
            fp = final[n.parent.nodeNum]
            inter = down[n.nodeNum] & fp
            empty = inter == 0
            final[n.nodeNum] = numpy.where(empty, down[n.nodeNum], inter)
            changes[n.nodeNum] = float((self.w * empty).sum()) / self.nChar
            depth[n.nodeNum] = depth[n.parent.nodeNum] + 1
        full = (1 << self.nTax) - 1
        feats = {}
        splits = {}
        for n in tree.iterNodesNoRoot():
            d = down[n.nodeNum]
            pc = _popcount8(d.astype(numpy.int32))
            amb = float((self.w * (pc > 1)).sum()) / self.nChar
            single = float((self.w * (pc == 1)).sum()) / self.nChar
            nk = sum(1 for _ in n.iterChildren())
            feats[n.nodeNum] = (changes[n.nodeNum], amb, single, size[n.nodeNum] / float(self.nTax),
                                1.0 if n.isLeaf else 0.0, depth[n.nodeNum] / 10.0, nk / 3.0)
            splits[n.nodeNum] = edgeKey(bits[n.nodeNum], full)
        got = (feats, splits)
        if len(self._topo) > 20000:
            self._topo.clear()
        self._topo[key] = got
        return got

    def _branchRow(self, topo, nodeNum, brLen, treeLen, compNum, compVal):
        ch, amb, single, clade, leaf, depth, nk = topo[nodeNum]
        x = numpy.zeros(NFEAT)
        x[0] = 1.0
        x[2] = math.log(brLen)
        x[3] = math.log(treeLen)
        x[4] = ch
        x[5] = amb
        x[6] = clade
        x[7] = leaf
        x[8] = depth
        x[9] = math.log(ch + 1e-3)
        x[10 + compNum] = 1.0
        x[14:17] = _alr(compVal)
        x[17:20] = self.empAlr
        x[20] = single
        x[21] = nk
        return x

    def _compRow(self, compNum, compVal, treeLen):
        x = numpy.zeros(NFEAT)
        x[1] = 1.0
        x[3] = math.log(treeLen)
        x[10 + compNum] = 1.0
        x[14:17] = _alr(compVal)
        x[17:20] = self.empAlr
        return x

    def _rows(self, tree, topo, nodeNums, brLens, compVals, withComps):
        """Feature rows for the given nodes (brLens: nodeNum -> length) and,
        if withComps, one row per free composition."""
        treeLen = sum(brLens.values())
        rows = []
        for num in nodeNums:
            n = tree.nodes[num]
            cn = n.parts[self.pNum].compNum
            rows.append(self._branchRow(topo, num, brLens[num], treeLen, cn, compVals[cn]))
        if withComps:
            mp = tree.model.parts[self.pNum]
            for c in mp.comps:
                if c.free:
                    rows.append(self._compRow(c.num, compVals[c.num], treeLen))
        return numpy.array(rows)

    def _gauss(self, X, nBranch):
        """Means and sds in the original coordinates for feature rows X."""
        out = self.net.forward((X - self.xMean) / self.xSd).astype(numpy.float64)
        mu = out[:, :NCOORD]
        ls = numpy.clip(out[:, NCOORD:], -8.0, 4.0)
        yM = numpy.where(numpy.arange(len(X))[:, None] < nBranch, self.yMean, self.yMeanComp)
        yS = numpy.where(numpy.arange(len(X))[:, None] < nBranch, self.ySd, self.ySdComp)
        return yM + yS * mu, yS * numpy.exp(ls) * self.scale

    # ------------------------------------------------------------ sampling
    def recordState(self, tree):
        """Sample hook: remember the cold chain's state for training."""
        feats, splits = self._topology(tree)
        mp = tree.model.parts[self.pNum]
        self.records.append(dict(
            key=self._structKey(tree),
            topoKey=frozenset(splits.values()),
            brLens={n.nodeNum: n.br.len for n in tree.iterNodesNoRoot()},
            comps=[c.val.copy() for c in mp.comps],
            compNums={n.nodeNum: n.parts[self.pNum].compNum for n in tree.iterNodesNoRoot()},
            freeComps=[c.num for c in mp.comps if c.free]))

    def attach(self, mcmc):
        """Record the cold chain at every sample of mcmc."""
        mcmc.sampleHooks.append(self.recordState)

    def _trainingSet(self, minGroup=20):
        """(X, Y, mask, isBranch) with a fresh random partner per row."""
        groups = {}
        for i, r in enumerate(self.records):
            groups.setdefault(r['topoKey'], []).append(i)
        X, Y, M, B = [], [], [], []
        for key, idx in groups.items():
            if len(idx) < minGroup:
                continue
            partner = self.rng.permutation(idx)
            for i, j in zip(idx, partner):
                ri, rj = self.records[i], self.records[j]
                feats, splits = self._topo[ri['key']]
                # split -> length in sample j
                fj, sj = self._topo[rj['key']]
                lenBySplit = {sj[num]: l for num, l in rj['brLens'].items()}
                treeLen = sum(ri['brLens'].values())
                for num, l in ri['brLens'].items():
                    cn = ri['compNums'][num]
                    X.append(self._branchRow(feats, num, l, treeLen, cn, ri['comps'][cn]))
                    Y.append([math.log(lenBySplit[splits[num]]), 0.0, 0.0])
                    M.append([1.0, 0.0, 0.0])
                    B.append(True)
                for cn in ri['freeComps']:
                    X.append(self._compRow(cn, ri['comps'][cn], treeLen))
                    Y.append(_alr(rj['comps'][cn]).tolist())
                    M.append([1.0, 1.0, 1.0])
                    B.append(False)
        if not X:
            raise P4Error("LearnedBrLenCompProposal: no topology has %d or more samples" % minGroup)
        return numpy.array(X), numpy.array(Y), numpy.array(M), numpy.array(B)

    def train(self, epochs=30, batch=4096, lr=1e-3, minGroup=20, verbose=False):
        """Fit the network to the recorded samples on the GPU (Adam on the
        Gaussian negative log likelihood)."""
        X, Y, M, B = self._trainingSet(minGroup)
        self.xMean = X.mean(axis=0)
        self.xSd = X.std(axis=0)
        self.xSd[self.xSd < 1e-6] = 1.0
        self.yMean = numpy.array([Y[B, 0].mean(), 0.0, 0.0])
        self.ySd = numpy.array([max(Y[B, 0].std(), 1e-3), 1.0, 1.0])
        if (~B).any():
            self.yMeanComp = Y[~B].mean(axis=0)
            self.ySdComp = numpy.maximum(Y[~B].std(axis=0), 1e-3)
        self.trainLog = []
        for ep in range(epochs):
            if ep > 0:
                X, Y, M, B = self._trainingSet(minGroup)
            Yn = numpy.where(B[:, None], (Y - self.yMean) / self.ySd, (Y - self.yMeanComp) / self.ySdComp)
            Xn = (X - self.xMean) / self.xSd
            order = self.rng.permutation(len(X))
            tot = 0.0
            for s in range(0, len(order), batch):
                sel = order[s:s + batch]
                tot += self.net.trainStep(Xn[sel], Yn[sel], M[sel], lr)
            self.trainLog.append(tot / M.sum())
            if verbose:
                print("epoch %d  loss/coord %.4f" % (ep, self.trainLog[-1]))
        self.trained = True
        return self.trainLog

    def save(self, path):
        """Write the trained network and its input/output scalings (.npz)."""
        numpy.savez(path, params=self.net.getParams(), xMean=self.xMean, xSd=self.xSd, yMean=self.yMean,
                    ySd=self.ySd, yMeanComp=self.yMeanComp, ySdComp=self.ySdComp,
                    hidden=numpy.int64(self.net.nHidden), trainLog=numpy.asarray(self.trainLog))

    def load(self, path):
        with numpy.load(path, allow_pickle=False) as z:
            if int(z['hidden']) != self.net.nHidden:
                raise P4Error("LearnedBrLenCompProposal.load: hidden width %d, expected %d"
                              % (int(z['hidden']), self.net.nHidden))
            self.net.setParams(z['params'])
            for k in ('xMean', 'xSd', 'yMean', 'ySd', 'yMeanComp', 'ySdComp'):
                setattr(self, k, numpy.array(z[k], dtype=numpy.float64))
            self.trainLog = list(z['trainLog'])
        self.trained = True

    # ------------------------------------------------------------ the move
    def propose(self, ch, prop):
        """Mutate ch.propTree (equal to ch.curTree on entry) and set the
        chain's log prior and proposal ratios."""
        ch.logPriorRatio = 0.0
        ch.logProposalRatio = 0.0
        ch.touched = set()
        ch.touchedParts = set()
        ch.topologyChanged = False
        if not self.trained:
            raise P4Error("LearnedBrLenCompProposal.propose: train() first")
        t = ch.propTree
        mp = t.model.parts[self.pNum]
        self.stats['proposed'] += 1
        topo = self._topology(t)[0]
        nodes = [n.nodeNum for n in t.iterNodesNoRoot()]
        k = min(self.blockSize, len(nodes))
        block = [nodes[i] for i in self.rng.choice(len(nodes), size=k, replace=False)]
        withComps = self.compProb > 0.0 and self.rng.random() < self.compProb
        brLens = {n.nodeNum: n.br.len for n in t.iterNodesNoRoot()}
        compVals = [c.val.copy() for c in mp.comps]
        freeComps = [c.num for c in mp.comps if c.free] if withComps else []

        X = self._rows(t, topo, block, brLens, compVals, withComps)
        mu, sd = self._gauss(X, k)
        eps = self.rng.standard_normal(mu.shape)
        y = mu + sd * eps
        newLens = dict(brLens)
        logQf = 0.0
        yOld = []
        for i, num in enumerate(block):
            yOld.append(math.log(brLens[num]))
            newLens[num] = math.exp(y[i, 0])
            logQf += _logNormal(y[i, 0], mu[i, 0], sd[i, 0]) - y[i, 0]
        newComps = [v.copy() for v in compVals]
        for j, cn in enumerate(freeComps):
            r = k + j
            pi = _alrInv(y[r])
            newComps[cn] = pi
            logQf += _logNormal(y[r], mu[r], sd[r]) - float(numpy.log(pi).sum())
        ok = all(var.BRLEN_MIN <= newLens[num] <= var.BRLEN_MAX for num in block) and \
            all(newComps[cn].min() >= var.PIVEC_MIN for cn in freeComps)
        if not ok:
            self.stats['outOfBounds'] += 1
            ch.logProposalRatio = -1.0e300
            return
        X2 = self._rows(t, topo, block, newLens, newComps, withComps)
        mu2, sd2 = self._gauss(X2, k)
        logQr = 0.0
        for i, num in enumerate(block):
            logQr += _logNormal(yOld[i], mu2[i, 0], sd2[i, 0]) - yOld[i]
        for j, cn in enumerate(freeComps):
            r = k + j
            logQr += _logNormal(_alr(compVals[cn]), mu2[r], sd2[r]) - float(numpy.log(compVals[cn]).sum())
        ch.logProposalRatio = logQr - logQf
        lam = ch.mcmc.brLenPriorLambda
        for num in block:
            n = t.nodes[num]
            ch.logPriorRatio += lam * (brLens[num] - newLens[num])
            n.br.len = newLens[num]
            pf.p4_setBrLen(n.cNode, n.br.len)
            pf.p4_calculateBigPDecks(n.cNode)
            ch._markPath(t, n.parent)
            ch.touched.add(n)
        if freeComps:
            for cn in freeComps:
                mp.comps[cn].val[:] = newComps[cn]
            ch._modelChanged(self.pNum)

    def afterMove(self, ch, accept):
        if accept:
            self.stats['accepted'] += 1
