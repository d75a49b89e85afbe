"""MCMC over the GPU likelihood.  CPU tests cover the topology move and the
split bookkeeping; GPU tests check that every proposal's incremental
likelihood equals a full recomputation (the reference's own runtime check,
p4/chain.py:264-286), that the current and proposed trees stay identical
between generations (p4/chain.py:1603-1617), and that a prior-only run
recovers the prior."""
import math
import random

import numpy
import pytest

from p4gpu import func, pf, synth
from p4gpu import mcmc as mcmcmod
from p4gpu.mcmc import Mcmc, _swapSubtrees, compareSplitFrequencies, splitKeys


def _nni_once(t, rng):
    cands = [n for n in t.iterInternalsNoRoot()]
    u = rng.choice(cands)
    p = u.parent
    x = rng.choice(list(u.iterChildren()))
    s = rng.choice([c for c in p.iterChildren() if c is not u])
    _swapSubtrees(x, u, s, p)
    t.setPreAndPostOrder()
    return x, u, s, p


def _check_tree(t):
    seen = set()
    for n in t.iterPreOrder():
        seen.add(n.nodeNum)
        for c in n.iterChildren():
            assert c.parent is n
    assert len(seen) == len(t.nodes)
    assert sum(1 for n in t.nodes if n.isLeaf) == len(t.taxNames)


def test_nni_keeps_a_valid_tree_and_changes_one_split():
    rng = random.Random(3)
    t = func.randomTree(nTax=20, seed=3)
    for _ in range(200):
        before = splitKeys(t)
        x, u, s, p = _nni_once(t, rng)
        _check_tree(t)
        after = splitKeys(t)
        assert len(before ^ after) in (0, 2)
        assert len(after) == 20 - 3


def test_nni_is_its_own_inverse():
    rng = random.Random(4)
    t = func.randomTree(nTax=12, seed=4)
    nw = t.writeNewick()
    x, u, s, p = _nni_once(t, rng)
    _swapSubtrees(s, u, x, p)
    t.setPreAndPostOrder()
    assert splitKeys(t) == splitKeys(func.readNewick(nw))


def test_split_comparison():
    a = {1: 1.0, 2: 0.5, 3: 0.05}
    b = {1: 1.0, 2: 0.4}
    mx, asd = compareSplitFrequencies(a, b)
    assert abs(mx - 0.1) < 1e-12
    assert abs(asd - numpy.mean([0.0, 0.05])) < 1e-12


# ------------------------------------------------- MC^3 formulas (CPU)
def test_heated_ratio_is_the_references():
    """p4/chain.py:1104-1129: with nChains > 1 the likelihood ratio and the
    prior ratio are multiplied by beta = 1 / (1 + T); the proposal ratio is
    not.  Checked against that formula written out term by term."""
    rng = numpy.random.default_rng(1)
    for _ in range(200):
        T = float(rng.choice([0.0, 0.1, 0.2, 0.3, 1.5]))
        beta = 1.0 / (1.0 + T)
        lp, lc = rng.normal(-5000, 50, 2)
        pr, qr = rng.normal(0, 3, 2)
        logLikeRatio = (lp - lc) * beta
        logPriorRatio = pr * beta
        want = logLikeRatio + qr + logPriorRatio
        got = mcmcmod.heatedLogRatio(beta, lp, lc, pr, qr)
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want))
        # a prior-only run drops the likelihood ratio
        assert abs(mcmcmod.heatedLogRatio(beta, lp, lc, pr, qr, samplePrior=True) - (pr * beta + qr)) < 1e-12
    # the -1e99 sentinel of a site likelihood <= 0 is never accepted
    assert mcmcmod.acceptProbability(mcmcmod.heatedLogRatio(1.0, -1.0e99, -10.0, 0.0, 50.0)) == 0.0


def test_accept_probability_cutoffs():
    """p4/chain.py:1310-1320."""
    assert mcmcmod.acceptProbability(-100.0001) == 0.0
    assert mcmcmod.acceptProbability(0.0) == 1.0
    assert mcmcmod.acceptProbability(3.0) == 1.0
    assert mcmcmod.acceptProbability(-1.0) == math.exp(-1.0)
    assert mcmcmod.acceptProbability(-99.0) == math.exp(-99.0)


def test_swap_ratio_is_the_references():
    """p4/mcmc.py:3315-3322, term by term: lnR = b1*L2 + b2*L1 - b1*L1 -
    b2*L2 with b = 1/(1+T) -- likelihoods only, no prior."""
    rng = numpy.random.default_rng(2)
    temps = [0.0, 0.1, 0.2, 0.3]
    for _ in range(200):
        i = int(rng.integers(0, 3))
        l1, l2 = rng.normal(-1000, 20, 2)
        lnR = (1.0 / (1.0 + temps[i])) * l2
        lnR += (1.0 / (1.0 + temps[i + 1])) * l1
        lnR -= (1.0 / (1.0 + temps[i])) * l1
        lnR -= (1.0 / (1.0 + temps[i + 1])) * l2
        got = mcmcmod.swapLogRatio(1.0 / (1.0 + temps[i]), 1.0 / (1.0 + temps[i + 1]), l1, l2)
        assert abs(got - lnR) <= 1e-12 * max(1.0, abs(lnR))
    # equal heats or equal likelihoods: always accepted
    assert mcmcmod.swapLogRatio(0.8, 0.8, -10.0, -50.0) == 0.0
    assert mcmcmod.swapLogRatio(1.0, 0.5, -10.0, -10.0) == 0.0


class _Stub(object):
    pass


def test_swap_move_statistics_and_bookkeeping():
    """Mcmc._swap on stand-in chains: an adjacent pair is drawn uniformly,
    attempts go to the upper triangle of swapMatrix and accepted swaps to
    the lower one (p4/mcmc.py:3312, 3352), the chains exchange temperatures,
    and the acceptance frequency of each pair is min(1, exp(lnR))."""
    m = _Stub()
    m.nChains = 4
    m.chainTemps = [0.0, 0.1, 0.2, 0.3]
    m.rng = random.Random(5)
    m.swapMatrix = numpy.zeros((4, 4), int)
    lnls = [-1000.0, -1003.0, -1001.5, -1010.0]
    m.chains = []
    for i in range(4):
        ch = _Stub()
        ch.tempNum = i
        ch.curTree = _Stub()
        ch.curTree.logLike = lnls[i]
        m.chains.append(ch)
    n = 60000
    acc = numpy.zeros(3)
    att = numpy.zeros(3)
    for _ in range(n):
        before = {ch.tempNum: ch for ch in m.chains}
        mcmcmod.Mcmc._swap(m)
        after = {ch.tempNum: ch for ch in m.chains}
        moved = [t for t in range(4) if before[t] is not after[t]]
        assert len(moved) in (0, 2)
        if moved:
            assert moved[1] == moved[0] + 1
        # put the chains back so every attempt sees the same pair states
        for t, ch in before.items():
            ch.tempNum = t
    for t1 in range(3):
        att[t1] = m.swapMatrix[t1][t1 + 1]
        acc[t1] = m.swapMatrix[t1 + 1][t1]
    assert att.sum() == n
    assert numpy.allclose(att / n, 1.0 / 3.0, atol=0.01)
    for t1 in range(3):
        b1, b2 = 1.0 / (1.0 + m.chainTemps[t1]), 1.0 / (1.0 + m.chainTemps[t1 + 1])
        p = min(1.0, math.exp(mcmcmod.swapLogRatio(b1, b2, lnls[t1], lnls[t1 + 1])))
        se = math.sqrt(p * (1 - p) / att[t1]) + 1e-12
        assert abs(acc[t1] / att[t1] - p) < 5 * se + 1e-9, (t1, acc[t1] / att[t1], p)


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_incremental_equals_full_for_every_proposal(prec):
    tree, _ = synth.make_workload("het64", seed=21, nSites=1500, precision=prec)
    m = Mcmc(tree, nChains=2, seed=5, verbose=False)
    tol = 1e-10 if prec == "fp64" else 1e-7
    names = set()
    for g in range(120):
        m.run(1, verbose=False)
        for ch in m.chains:
            full = ch.curTree.dupe()
            full.precision = prec
            v = full.calcLogLike(verbose=0)
            assert abs(v - ch.curTree.logLike) <= tol * abs(v), (g, v, ch.curTree.logLike)
            assert pf.p4_verifyIdentityOfTwoTrees(ch.curTree.cTree, ch.propTree.cTree) == 0
            if g % 15 == 0:
                # the sync leaves prop's partials and P equal to cur's, bit for bit
                mp = ch.curTree.model.parts[0]
                nP = ch.curTree.data.parts[0].nPatterns
                for n, m2 in zip(ch.curTree.iterInternalsNoRoot(), ch.propTree.iterInternalsNoRoot()):
                    a = pf.p4_getCondLikes(n.cNode, 0, mp.nGammaCat, mp.dim, nP)
                    b = pf.p4_getCondLikes(m2.cNode, 0, mp.nGammaCat, mp.dim, nP)
                    assert (a == b).all(), (g, n.nodeNum)
                for n, m2 in zip(ch.curTree.iterNodesNoRoot(), ch.propTree.iterNodesNoRoot()):
                    assert (pf.p4_getBigPDeck(n.cNode, 0, mp.nGammaCat, mp.dim) ==
                            pf.p4_getBigPDeck(m2.cNode, 0, mp.nGammaCat, mp.dim)).all(), (g, n.nodeNum)
        names |= {p.name for p in m.proposals if p.nProposals[0] + p.nProposals[1] > 0}
    assert {"brLen", "nni", "compDir", "rMatrixDir", "gdasrv", "pInvar"} <= names


@pytest.mark.gpu
def test_prior_only_run_recovers_branch_length_prior():
    tree, _ = synth.make_workload("gtr16", seed=22, nSites=300)
    m = Mcmc(tree, nChains=1, seed=6, sampleInterval=5, verbose=False,
             proposalWeights={'nni': 0.3, 'compDir.0': 0.0, 'rMatrixDir.0': 0.0, 'gdasrv.0': 0.0})
    m.samplePrior = True
    m.run(6000, verbose=False)
    tl = numpy.array(m.samples['treeLen'][200:])
    nBr = len(tree.nodes) - 1
    # Exponential(10) per branch: mean tree length nBr / 10
    assert abs(tl.mean() / (nBr / 10.0) - 1.0) < 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("G,prec", [(-3, "fp32"), (-1, "fp32"), (-2, "fp32"), (0, "fp32"),
                                    (4, "fp32"), (-3, "fp64"), (-1, "fp64")])
def test_incremental_equals_full_for_every_kernel_variant(G, prec):
    """Dirty-path evaluations (preloaded clean children, the pair kernel's
    generic and fast ops, junction slots of the two-stream variant, the
    program-order P stream, the grouped kernel) equal a full recomputation,
    generation after generation."""
    fp32 = prec == "fp32"
    tree, _ = synth.make_workload("het64", seed=31, nSites=1200, precision=prec)
    try:
        pf.setPruneTuning(fp32, G, 1)
        m = Mcmc(tree, nChains=2, seed=9, verbose=False)
        for g in range(60):
            m.run(1, verbose=False)
            for ch in m.chains:
                full = ch.curTree.dupe()
                full.precision = prec
                v = full.calcLogLike(verbose=0)
                tol = 1e-7 if fp32 else 1e-11
                assert abs(v - ch.curTree.logLike) <= tol * abs(v), (G, g, v, ch.curTree.logLike)
    finally:
        pf.setPruneTuning(fp32, -3, 1)   # the library default, then the tuned choice
        pf._applyTuning()


def _prior_run(proposalWeights, nGens, seed, sampleInterval=10, tunings=None):
    tree, _ = synth.make_workload("gtr16", seed=40 + seed, nSites=200)
    tree.setPInvar(partNum=0, free=1, val=0.2)
    w = {'brLen': 0.0, 'nni': 0.0, 'allBrLens': 0.0, 'compDir.0': 0.0, 'allCompsDir.0': 0.0,
         'rMatrixDir.0': 0.0, 'gdasrv.0': 0.0, 'pInvar.0': 0.0}
    w.update(proposalWeights)
    m = Mcmc(tree, nChains=1, seed=seed, sampleInterval=sampleInterval, verbose=False, proposalWeights=w)
    m.samplePrior = True
    m.tuneInterval = 10 ** 9      # fixed tuning: an exact Markov chain
    for p in m.proposals:         # broad moves mix fast enough for a short run
        if tunings and p.name in tunings:
            p.tuning = [tunings[p.name]] * m.nChains
    m.run(nGens, verbose=False)
    return tree, m


def _near(x, want, cap):
    """Mean of a trace within 4 Monte Carlo standard errors (Tracer ESS)."""
    x = numpy.asarray(x, dtype=float)
    se = x.std() / math.sqrt(max(func.effectiveSampleSize(x), 1.0))
    return abs(x.mean() - want) < min(4.0 * se + 1e-12, cap), (x.mean(), want, se)


@pytest.mark.gpu
@pytest.mark.parametrize("move", ["compDir", "allCompsDir"])
def test_prior_only_comp_moves_recover_the_flat_dirichlet(move):
    tree, m = _prior_run({move + '.0': 1.0}, 30000, seed=1 if move == "compDir" else 2, tunings={move: 10.0})
    prams = numpy.array(m.samples['prams'][200:])[:, :4]   # the one comp
    for j in range(4):
        ok, info = _near(prams[:, j], 0.25, 0.03)
        assert ok, (move, j, info)
    # Dirichlet(1,1,1,1): variance 3/80 per component
    assert abs(prams.var(axis=0).mean() / (3.0 / 80.0) - 1.0) < 0.15, prams.var(axis=0)


@pytest.mark.gpu
def test_prior_only_rmatrix_move_recovers_the_flat_dirichlet():
    tree, m = _prior_run({'rMatrixDir.0': 1.0}, 30000, seed=3, tunings={'rMatrixDir': 20.0})
    vals = []
    rm = m.coldChain().curTree.model.parts[0].rMatrices[0]

    def hook(t):
        r = t.model.parts[0].rMatrices[0].val
        vals.append(numpy.array(r) / numpy.sum(r))
    m.sampleHooks.append(hook)
    m.run(30000, verbose=False)
    v = numpy.array(vals[50:])
    # flat Dirichlet on 6 exchangeabilities: mean 1/6, variance (1/6)(5/6)/7
    for j in range(6):
        ok, info = _near(v[:, j], 1.0 / 6.0, 0.02)
        assert ok, (j, info)
    assert abs(v.var(axis=0).mean() / (5.0 / 252.0) - 1.0) < 0.15, v.var(axis=0)
    assert rm is not None


@pytest.mark.gpu
def test_prior_only_gamma_shape_move_recovers_the_uniform_prior():
    tree, m = _prior_run({'gdasrv.0': 1.0}, 40000, seed=4)
    a = numpy.array(m.samples['prams'][200:])[:, 4]
    lo, hi = 0.1, 300.0
    ok, info = _near(a, 0.5 * (lo + hi), 15.0)
    assert ok, info
    # uniform: a quarter of the mass below lo + (hi - lo) / 4
    q = (a < lo + 0.25 * (hi - lo)).mean()
    assert abs(q - 0.25) < 0.06, q


@pytest.mark.gpu
def test_prior_only_pinvar_move_recovers_the_uniform_prior():
    tree, m = _prior_run({'pInvar.0': 1.0}, 30000, seed=5)
    pinv = numpy.array(m.samples['prams'][200:])[:, 5]
    ok, info = _near(pinv, 0.99 / 2.0, 0.03)
    assert ok, info
    assert abs(pinv.var() / (0.99 ** 2 / 12.0) - 1.0) < 0.15, pinv.var()


@pytest.mark.gpu
def test_prior_only_all_branch_lengths_move_recovers_the_exponential_prior():
    tree, m = _prior_run({'allBrLens': 1.0}, 40000, seed=6, tunings={'allBrLens': 0.2})
    tl = numpy.array(m.samples['treeLen'][200:])
    nBr = len(tree.nodes) - 1
    # Exponential(10) per branch (restricted to [1e-8, 3], a negligible cut)
    ok, info = _near(tl, nBr / 10.0, 0.1 * nBr / 10.0)
    assert ok, info
    assert abs(tl.var() / (nBr / 100.0) - 1.0) < 0.2, tl.var()


@pytest.mark.gpu
def test_four_heated_chains_swap_and_keep_the_cold_chain_exact():
    """Four Metropolis-coupled chains (T = 0, 0.1, 0.2, 0.3) with data:
    swaps are attempted and accepted between every adjacent pair, the
    temperatures stay a permutation, and every chain's incremental lnL
    equals a full recomputation after the swaps (the swap exchanges
    temperatures, not tree states, as in p4/mcmc.py:3352-3353)."""
    tree, _ = synth.make_workload("gtr16", seed=47, nSites=200)
    m = Mcmc(tree, nChains=4, chainTemps=[0.0, 0.1, 0.2, 0.3], seed=8, sampleInterval=10, verbose=False,
             proposalWeights={'compDir.0': 0.0, 'rMatrixDir.0': 0.0, 'gdasrv.0': 0.0, 'nni': 0.2})
    m.run(300, verbose=False)
    for ch in m.chains:
        full = ch.curTree.dupe()
        v = full.calcLogLike(verbose=0)
        assert abs(v - ch.curTree.logLike) <= 1e-10 * abs(v)
    assert sorted(ch.tempNum for ch in m.chains) == [0, 1, 2, 3]
    for t1 in range(3):
        assert m.swapMatrix[t1][t1 + 1] > 0
        assert m.swapMatrix[t1 + 1][t1] > 0
