"""Simulation (Pf/p4_treeSim.c:14-590, p4/tree.py:9527-9627) and the
posterior-predictive summaries (Pf/part.c:682-714, 1419-1571;
p4/mcmc.py:3461-3514).

The oracle's simulation restates the reference's pickers with the GPU's
uniforms (Philox4x64-10, checked against numpy's generator).  It is pinned
on the CPU by the distribution it must produce: on a 4-taxon tree with
gamma categories and invariant sites, the frequency of each of the 256
site patterns over many simulated sites must match the pattern's
likelihood from the reference's own likelihood code (chi-square), and the
categories and invariant flags must have their model frequencies.  On the
GPU the simulated sequences and categories must equal the oracle's site by
site for the same seed, and the summaries must equal direct counts."""
import itertools
import math

import numpy
import pytest
from scipy.stats import chi2 as chi2d

import oracle
from helpers import load, oracle_flat, tree_from_fixture
from p4gpu import Alignment, Data, P4Error, pf, synth
from p4gpu.mcmc import Mcmc


def _quartet(nCopies=1, pInvar=0.2, K=4, alpha=0.5):
    """A 4-taxon fixture (root with two leaves and a cherry), GTR + G + I,
    whose patterns are all 256 DNA patterns, each nCopies times."""
    R = numpy.array([[0, 1.0, 2.5, 0.7], [1.0, 0, 0.8, 3.1], [2.5, 0.8, 0, 1.2], [0.7, 3.1, 1.2, 0]])
    pats = numpy.array(list(itertools.product(range(4), repeat=4)), numpy.int32).T.copy()
    return dict(parent=numpy.array([-1, 0, 0, 0, 3, 3]), leftChild=numpy.array([1, -1, -1, 4, -1, -1]),
                sibling=numpy.array([-1, 2, 3, -1, 5, -1]), isLeaf=numpy.array([0, 1, 1, 0, 1, 1]),
                seqNum=numpy.array([-1, 0, 1, -1, 2, 3]), brLen=numpy.array([0.0, 0.12, 0.3, 0.07, 0.45, 0.2]),
                postOrder=numpy.array([1, 2, 4, 5, 3, 0]), root=numpy.array(0),
                compNums=numpy.zeros(6, int), rMatrixNums=numpy.zeros(6, int), gdasrvNums=numpy.zeros(6, int),
                dim=numpy.array(4), nCat=numpy.array(K), comps=numpy.array([[0.1, 0.2, 0.3, 0.4]]),
                bigR=R[None], alphas=numpy.array([alpha]), pInvar=numpy.array(pInvar), relRate=numpy.array(1.0),
                patterns=pats, counts=numpy.full(256, nCopies, numpy.int32),
                equates=numpy.zeros((0, 4), numpy.int32), dataType=numpy.array("dna"))


def _codes(cPart):
    s = pf.symbolSequences(cPart)
    n = pf.lib().p4g_partNChar(cPart)
    lut = numpy.full(256, -1, numpy.int64)
    for i, ch in enumerate("acgt"):
        lut[ord(ch)] = i
    return lut[numpy.frombuffer(s.encode(), numpy.uint8)].reshape(-1, n)


# ------------------------------------------------------------------ CPU
def test_vector_philox_matches_numpy():
    rng = numpy.random.default_rng(5)
    c = [rng.integers(1, 2 ** 63, 20, dtype=numpy.uint64) for _ in range(4)]
    out = oracle.philox4x64_vec(*c, 0x1234, 0x5678)
    for i in range(20):
        want = oracle.philox4x64([int(x[i]) for x in c], [0x1234, 0x5678])
        assert [int(o[i]) for o in out] == [int(w) for w in want]


def test_oracle_simulation_has_the_likelihood_pattern_frequencies():
    fx = _quartet()
    f = oracle_flat(fx)
    _, pl = oracle.part_loglike(f, site_likes=True)
    assert abs(pl.sum() - 1.0) < 1e-12          # the 256 pattern probabilities
    if oracle.ref() is not None:
        _, rpl, _, _, _ = oracle.ref_part_loglike(f)
        assert numpy.allclose(pl, rpl, rtol=1e-12)
    Pm = oracle.oracle_pmats(f)
    n = 400000
    states, cats = oracle.simulate(f, Pm, n, seed=2024)
    leaves = [1, 2, 4, 5]
    idx = sum(states[leaves[t]] * 4 ** (3 - t) for t in range(4))   # pattern index, taxon 0 first
    obs = numpy.bincount(idx, minlength=256)
    exp = pl * n
    chi2 = ((obs - exp) ** 2 / exp).sum()
    assert chi2d.sf(chi2, 255) > 1e-4, chi2
    # categories uniform over K among variable sites, invariant share pInvar
    inv = cats < 0
    assert abs(inv.mean() - 0.2) < 4 * math.sqrt(0.2 * 0.8 / n)
    cc = numpy.bincount(cats[~inv], minlength=4)
    assert chi2d.sf((((cc - cc.sum() / 4) ** 2) / (cc.sum() / 4)).sum(), 3) > 1e-4
    # invariant sites keep the root state on every node
    assert (states[:, inv] == states[0, inv][None, :]).all()


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gpu_simulation_matches_the_oracle(prec):
    fx = _quartet(nCopies=200)
    f = oracle_flat(fx)
    t = tree_from_fixture(fx, prec)
    t.calcLogLike(verbose=0)
    cp = t.data.parts[0].cPart
    nChar = int(fx["counts"].sum())
    seed = 0xC0FFEE
    t.simulate(seed=seed)
    got = _codes(cp)
    states, cats = oracle.simulate(f, oracle.oracle_pmats(f), nChar, seed)
    for l in [1, 2, 4, 5]:
        assert (got[int(fx["seqNum"][l])] == states[l]).all(), l
    assert (pf.partSimCats(cp) == cats).all()
    # the part's patterns were made again from the new sequences
    pats, cnts, spi, _, _ = t.data.parts[0].patterns()
    assert cnts.sum() == nChar and (pats[:, spi] == got).all()
    # the same seed gives the same data; another seed does not
    t.simulate(seed=seed)
    assert (_codes(cp) == got).all()
    t.simulate(seed=seed + 1)
    assert (_codes(cp) != got).any()


@pytest.mark.gpu
def test_gpu_simulation_of_a_large_tree_matches_the_oracle():
    tree, data = synth.make_workload("prot128", seed=31, nSites=3000)
    tree.calcLogLike(verbose=0)
    f = oracle.flat_from_tree(tree, 0)
    tree.simulate(seed=77, resetSequences=True)
    cp = data.parts[0].cPart
    nChar = pf.lib().p4g_partNChar(cp)
    states, cats = oracle.simulate(f, oracle.oracle_pmats(f), nChar, 77)
    s = pf.symbolSequences(cp)
    lut = numpy.full(256, -1, numpy.int64)
    for i, ch in enumerate(data.parts[0].symbols):
        lut[ord(ch)] = i
    got = lut[numpy.frombuffer(s.encode(), numpy.uint8)].reshape(-1, nChar)
    for n in tree.iterLeavesNoRoot():
        assert (got[n.seqNum] == states[n.nodeNum]).all(), n.nodeNum
    assert (pf.partSimCats(cp) == cats).all()
    assert data.alignments[0].sequences[0] == s[:nChar]


@pytest.mark.gpu
def test_gpu_simulation_from_a_reference_tree():
    fx = _quartet(nCopies=20)
    ref = tree_from_fixture(fx, "fp64")
    ref.calcLogLike(verbose=0)
    t = tree_from_fixture(fx, "fp64")
    t.calcLogLike(verbose=0)
    seed = 4242
    t.simulate(refTree=ref, seed=seed)
    nChar = int(fx["counts"].sum())
    draw = pf.drawAncStates(ref.cTree, 0, 0, nChar, seed=seed ^ 0x9E3779B97F4A7C15)
    f = oracle_flat(fx)
    states, cats = oracle.simulate(f, oracle.oracle_pmats(f), nChar, seed, ref_draw=draw)
    got = _codes(t.data.parts[0].cPart)
    for l in [1, 2, 4, 5]:
        assert (got[int(fx["seqNum"][l])] == states[l]).all(), l
    assert (pf.partSimCats(t.data.parts[0].cPart) == cats).all()
    with pytest.raises(P4Error):
        t.simulate(refTree=t)


@pytest.mark.gpu
def test_sequence_summaries_equal_direct_counts():
    seqs = ["acgtac-tnr", "acgtacgtaa", "acctacgtaa", "acgtacgt?a"]
    aln = Alignment(["a", "b", "c", "d"], seqs, "dna")
    d = Data(alignments=aln)
    cp = d.parts[0].cPart
    counts, nConst, mean, other = pf.partSeqStats(cp)
    lut = {c: i for i, c in enumerate("acgt")}
    want = numpy.zeros((4, 4), int)
    for t, s in enumerate(seqs):
        for ch in s:
            if ch in lut:
                want[t, lut[ch]] += 1
    assert (counts == want).all()
    cols = list(zip(*seqs))
    dist = [len(set(c) & set("acgt")) for c in cols]
    assert nConst == sum(1 for x in dist if x == 1)
    assert abs(mean - sum(dist) / len(cols)) < 1e-15
    assert other == sum(1 for s in seqs for ch in s if ch not in lut)
    assert pf.partBigXSquared(cp) == -2.0
    # a clean alignment: the chi-square of Pf/part.c:1490-1571
    seqs = ["acgtacgtaa", "acgtacgtac", "ccctacgtaa", "acggtcgtaa"]
    d = Data(alignments=Alignment(["a", "b", "c", "d"], seqs, "dna"))
    obs = numpy.array([[s.count(ch) for ch in "acgt"] for s in seqs], float)
    ex = obs.sum(axis=0) / 4
    x2 = sum(((obs[t, j] - ex[j]) ** 2) / ex[j] for t in range(4) for j in range(4) if ex[j])
    assert abs(d.simpleBigXSquared()[0] - x2) < 1e-12
    _, cnts, _, _, _ = d.parts[0].patterns()
    unc = float((cnts * numpy.log(cnts)).sum() - 10 * math.log(10))
    assert abs(pf.getUnconstrainedLogLike(d.parts[0].cPart) - unc) < 1e-12


@pytest.mark.gpu
def test_mcmc_posterior_predictive_rows():
    tree, data = synth.make_workload("gtr16", seed=8, nSites=800)
    m = Mcmc(tree, nChains=2, sampleInterval=10, seed=3, verbose=False, simulate=1 | 2 | 4 | 16)
    m.run(50)
    rows = m.samples['sim']
    assert len(rows) == len(m.samples['gen']) > 0
    for r in rows:
        gen, unc, x2, mean, nconst = r
        assert unc < 0 and x2 >= 0 and 1.0 <= mean <= 4.0 and 0 <= nconst <= 800
    # the chains' own data is untouched
    assert abs(tree.calcLogLike(verbose=0) - m.coldChain().curTree.calcLogLike(verbose=0)) < 1e30
    with pytest.raises(P4Error):
        Mcmc(tree, nChains=1, simulate=8, verbose=False)
