"""The batched workload at its full shape: 48 taxa, 24 partitions (16 DNA
GTR+G4 and 8 protein 20-state+G4, sites log-uniform in [200, 40,000],
120,391 patterns), four Metropolis-coupled chains (T = 0, 0.1, 0.2, 0.3),
every chain's pending work of every partition in one p4g_batchLogLike call
(the replacement of the per-chain, per-part loop of p4/mcmc.py:2861-2950).

The cold chain's per-part log likelihoods are checked against the oracle
(FP64 within 1e-10, FP32 within 1e-6 relative), the other chains against
their own single-tree FP64 evaluation; both for full recomputation and for
the dirty path of one branch move per chain."""
import numpy
import pytest

import oracle
from p4gpu import pf, synth
from p4gpu.mcmc import Mcmc

pytestmark = pytest.mark.gpu

TEMPS = [0.0, 0.1, 0.2, 0.3]
TOL = {"fp64": 1e-10, "fp32": 1e-6}


def rel(a, b):
    return abs(a - b) / abs(b)


def _chains(prec):
    base, data = synth.make_workload("multi48", seed=1, precision=prec)
    assert data.nParts == 24 and sum(p.nPatterns for p in data.parts) == 120391
    trees = [base] + [base.dupe() for _ in range(3)]
    for i, t in enumerate(trees):
        t.precision = prec
        for n in t.iterNodesNoRoot():   # four different chain states
            n.br.len *= 1.0 + 0.04 * i
        t.calcLogLike(verbose=0)
    return trees, data


def _oracle_parts(t):
    return numpy.array([oracle.part_loglike(oracle.flat_from_tree(t, p)) for p in range(t.model.nParts)])


def _single_fp64(t):
    d = t.dupe()
    d.precision = "fp64"
    return d.calcLogLike(verbose=0)


def _mark_path(t, n):
    m = n.parent
    while m is not None:
        for p in range(t.model.nParts):
            pf.p4_setConditionalLikelihoodsOfInternalNodePart(m.cNode, p)
        m = m.parent


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_four_chains_full_recompute_in_one_batch(prec):
    trees, data = _chains(prec)
    batch = pf.Batch([t.cTree for t in trees])
    for t in trees:
        pf.p4_setPrams(t.cTree, -1)
        for n in t.iterInternals():
            pf.p4_setConditionalLikelihoodsOfInternalNode(n.cNode)
    got = batch.logLikes()
    want = _oracle_parts(trees[0])
    parts = numpy.array(trees[0].partLikes)
    for p in range(24):
        assert rel(parts[p], want[p]) < TOL[prec], (prec, p, parts[p], want[p])
    assert rel(got[0], want.sum()) < TOL[prec]
    for i in range(1, 4):
        assert rel(got[i], _single_fp64(trees[i])) < TOL[prec], (prec, i)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_four_chains_dirty_path_in_one_batch(prec):
    trees, data = _chains(prec)
    batch = pf.Batch([t.cTree for t in trees])
    for i, t in enumerate(trees):
        nodes = list(t.iterNodesNoRoot())
        n = nodes[(5 + 11 * i) % len(nodes)]     # a different branch per chain
        n.br.len *= 1.3
        pf.p4_setBrLen(n.cNode, n.br.len)
        pf.p4_calculateBigPDecks(n.cNode)
        _mark_path(t, n)
    got = batch.logLikes()
    want = _oracle_parts(trees[0])
    parts = numpy.array(trees[0].partLikes)
    for p in range(24):
        assert rel(parts[p], want[p]) < TOL[prec], (prec, p, parts[p], want[p])
    for i in range(1, 4):
        assert rel(got[i], _single_fp64(trees[i])) < TOL[prec], (prec, i)


def test_four_heated_chains_mcmc_on_the_full_workload():
    """A few generations of Mcmc with four heated chains on the full
    workload (every generation one batched call for the four proposals):
    afterwards the cold chain matches the oracle part by part and every
    chain its own full recomputation."""
    base, data = synth.make_workload("multi48", seed=2, precision="fp32")
    m = Mcmc(base, nChains=4, chainTemps=TEMPS, seed=4, verbose=False)
    m.run(6, verbose=False)
    cold = m.coldChain().curTree
    want = _oracle_parts(cold)
    assert rel(cold.logLike, want.sum()) < TOL["fp32"], (cold.logLike, want.sum())
    for ch in m.chains:
        assert rel(ch.curTree.logLike, _single_fp64(ch.curTree)) < TOL["fp32"]
        assert pf.p4_verifyIdentityOfTwoTrees(ch.curTree.cTree, ch.propTree.cTree) == 0
