"""Site-pattern compression on the GPU (the device form of makePatterns,
Pf/part.c:317-448): the patterns, their first-occurrence order, counts and
the site map must equal the host build's and the oracle's, exactly."""
import time

import numpy
import pytest

import oracle
from p4gpu import Alignment, Data, pf, synth


def _host_vs_device(aln):
    d1 = Data(alignments=aln)
    d2 = Data(alignments=Alignment(aln.taxNames, list(aln.sequences), aln.dataType, aln.symbols, dict(aln.equates)))
    p1, p2 = d1.parts[0], d2.parts[0]
    pf.makePatternsDevice(p2.cPart)
    a = p1.patterns()
    b = p2.patterns()
    for x, y in zip(a[:3], b[:3]):
        assert x.shape == y.shape and (x == y).all()
    return b


def _random_aln(rng, nTax, nChar, dataType, nVariants, gapFrac=0.05):
    sym = 'acgt' if dataType == 'dna' else 'arndcqeghilkmfpstwyv'
    extra = 'nrykm-?' if dataType == 'dna' else 'bxz-?'
    # few distinct columns, repeated: many duplicate patterns
    cols = []
    for _ in range(nVariants):
        c = rng.choice(list(sym), nTax)
        m = rng.random(nTax) < gapFrac
        c[m] = rng.choice(list(extra), m.sum())
        cols.append(c)
    pick = rng.integers(0, nVariants, nChar)
    seqs = [''.join(cols[j][t] for j in pick) for t in range(nTax)]
    return Alignment(["t%d" % i for i in range(nTax)], seqs, dataType)


@pytest.mark.gpu
@pytest.mark.parametrize("nTax,nChar,dt,nVar", [(4, 1, 'dna', 1), (5, 300, 'dna', 1), (7, 500, 'dna', 500),
                                                (16, 2000, 'dna', 120), (64, 3000, 'protein', 900),
                                                (300, 400, 'dna', 50)])
def test_device_patterns_equal_the_host_build(nTax, nChar, dt, nVar):
    rng = numpy.random.default_rng(nTax * 1000 + nChar)
    aln = _random_aln(rng, nTax, nChar, dt, nVar)
    pats, cnts, spi, _, _ = _host_vs_device(aln)
    # and the oracle's reference-order compression of the same codes
    codes = oracle.code_sequences(aln.sequences, aln.symbols, aln.equates)
    op, oc, os_ = oracle.compress(codes)
    assert (op == pats).all() and (oc == cnts).all() and (os_ == spi).all()


@pytest.mark.gpu
def test_device_patterns_of_a_simulation():
    tree, data = synth.make_workload("het64", seed=5, nSites=200000)
    tree.calcLogLike(verbose=0)
    tree.simulate(seed=12, calculatePatterns=True, resetSequences=False)   # device compression
    cp = data.parts[0].cPart
    dev = data.parts[0].patterns()
    t0 = time.perf_counter()
    pf.makePatterns(cp)                                                      # host compression
    host_s = time.perf_counter() - t0
    host = data.parts[0].patterns()
    for x, y in zip(dev[:3], host[:3]):
        assert (x == y).all()
    t0 = time.perf_counter()
    pf.makePatternsDevice(cp)
    dev_s = time.perf_counter() - t0
    print("200,000 sites x 64 taxa: host %.1f ms, device %.1f ms" % (host_s * 1e3, dev_s * 1e3))
    # the likelihood of the new data is the same through either build
    pf.setGlobalInvarSitesVec(cp)
    data.parts[0].nPatterns = pf.partPatternCount(cp)
    v = tree.calcLogLike(verbose=0)
    assert numpy.isfinite(v)
