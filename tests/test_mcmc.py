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
@pytest.mark.parametrize("G", [-3, -1, -2, 0, 4])
def test_incremental_equals_full_for_every_kernel_variant(G):
    """Dirty-path evaluations (preloaded clean children, the pair kernel's
    generic and fast ops, junction slots of the two-stream variant, the
    program-order P stream, the grouped kernel) equal a full recomputation,
    generation after generation."""
    tree, _ = synth.make_workload("het64", seed=31, nSites=1200, precision="fp32")
    try:
        pf.setPruneTuning(True, G, 1)
        m = Mcmc(tree, nChains=2, seed=9, verbose=False)
        for g in range(60):
            m.run(1, verbose=False)
            for ch in m.chains:
                full = ch.curTree.dupe()
                full.precision = "fp32"
                v = full.calcLogLike(verbose=0)
                assert abs(v - ch.curTree.logLike) <= 1e-7 * abs(v), (G, g, v, ch.curTree.logLike)
    finally:
        pf.setPruneTuning(True, -3, 1)   # the library default, then the tuned choice
        pf._applyTuning()
