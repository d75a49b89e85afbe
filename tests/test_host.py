"""Host-side checks of libp4gpu that need no GPU: the C-ABI exports every
declared symbol, data preparation (coding, pattern compression, invariant
sites) and discrete-gamma rates agree with the oracle and the reference's
golden vectors, the Python drop-in builds models as p4 does, and evaluation
fails loudly without a GPU.  CPU only."""
import ctypes
import os
import re

import numpy
import pytest

import oracle
from helpers import GOLDEN, load
from p4gpu import P4Error, func, pf, synth
from p4gpu.data import DNA_EQUATES, DNA_SYMBOLS, PROTEIN_EQUATES, PROTEIN_SYMBOLS, Alignment, Data

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "p4gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(p4g_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pf.LIB_PATH)
    names = _declared()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_gpu_means_loud_failure():
    if pf.deviceAvailable():
        pytest.skip("a GPU is present")
    tree, _ = synth.make_workload("gtr16", seed=1, nSites=100)
    with pytest.raises(P4Error, match="GPU"):
        tree.calcLogLike(verbose=0)


@pytest.mark.parametrize("dataType,symbols,equates", [
    ("dna", DNA_SYMBOLS, DNA_EQUATES), ("protein", PROTEIN_SYMBOLS, PROTEIN_EQUATES)])
def test_coding_and_compression_match_oracle(dataType, symbols, equates):
    rng = numpy.random.default_rng(3)
    alphabet = list(symbols) + ['-', '?'] + list(equates.keys())
    probs = numpy.array([1.0] * len(symbols) + [0.05] * (2 + len(equates)))
    seqs = []
    base = rng.choice(alphabet, 700, p=probs / probs.sum())
    for i in range(9):
        s = base.copy()
        flip = rng.random(700) < 0.15
        s[flip] = rng.choice(alphabet, flip.sum(), p=probs / probs.sum())
        seqs.append(''.join(s))
    d = Data(Alignment(["x%d" % i for i in range(9)], seqs, dataType))
    pats, cnts, spi, iv, ia = d.parts[0].patterns()
    codes = oracle.code_sequences(seqs, symbols, equates)
    opats, ocnts, ospi = oracle.compress(codes)
    assert numpy.array_equal(pats, opats)
    assert numpy.array_equal(cnts, ocnts)
    assert numpy.array_equal(spi, ospi)
    oiv, oia = oracle.invar_sites(opats, len(symbols), oracle.equates_table(symbols, equates))
    assert numpy.array_equal(iv, oiv) and numpy.array_equal(ia, oia)


def test_compression_at_full_size_is_consistent():
    # 64 taxa x 50,000 sites: expansion round trip and counts
    tree, data = synth.make_workload("het64", seed=2)
    p = data.parts[0]
    pats, cnts, spi, _, _ = p.patterns()
    aln = data.alignments[0]
    codes = oracle.code_sequences(aln.sequences[:3], aln.symbols, aln.equates)
    assert cnts.sum() == 50000
    assert numpy.array_equal(pats[:3, spi], codes)
    first = numpy.full(pats.shape[1], -1)
    seen = numpy.zeros(pats.shape[1], bool)
    order = []
    for i, k in enumerate(spi):
        if not seen[k]:
            seen[k] = True
            order.append(k)
    assert order == list(range(pats.shape[1]))


def test_product_gamma_rates_match_reference_bitwise():
    g = load("gamma_rates")
    for i, a in enumerate(g["alphas"]):
        for j, K in enumerate(g["Ks"]):
            K = int(K)
            freqs, rates = numpy.zeros(K), numpy.zeros(K)
            pf.gdasrvCalcRates_np(K, float(a), freqs, rates)
            assert numpy.array_equal(rates, g["rates"][i, j, :K]), (a, K)
            assert numpy.allclose(freqs, 1.0 / K)


def test_bad_sequence_character_raises():
    with pytest.raises(P4Error, match="neither a symbol"):
        Data(Alignment(["a", "b"], ["acgtx", "acgta"], "dna"))


def test_newick_round_trip_and_orders():
    t = func.readNewick("((a:0.1,b:0.2):0.05,(c:0.3,(d:0.1,e:0.4):0.2):0.1,f:0.7);")
    assert t.writeNewick() == "((a:0.1,b:0.2):0.05,(c:0.3,(d:0.1,e:0.4):0.2):0.1,f:0.7);"
    post = [i for i in t.postOrder if i != -10000]
    pos = {n: i for i, n in enumerate(post)}
    for n in t.nodes:
        if n.parent:
            assert pos[n.nodeNum] < pos[n.parent.nodeNum]
    assert post[-1] == t.root.nodeNum


def test_random_tree_is_unrooted_yule():
    t = func.randomTree(nTax=64, seed=4)
    assert sum(n.isLeaf for n in t.nodes) == 64
    assert len(t.nodes) == 2 * 64 - 2
    assert t.root.getNChildren() == 3
    for n in t.nodes:
        if not n.isLeaf and n is not t.root:
            assert n.getNChildren() == 2


def test_model_components_placed_with_representation():
    import random
    tree, _ = synth.make_workload("het64", seed=7, nSites=200)
    mp = tree.model.parts[0]
    comps = [n.parts[0].compNum for n in tree.nodes]
    rmats = [n.br.parts[0].rMatrixNum for n in tree.nodes if n is not tree.root]
    for c in range(mp.nComps):
        assert comps.count(c) >= 2
    for r in range(mp.nRMatrices):
        assert rmats.count(r) >= 2


def test_specified_rmatrix_is_normalised():
    tree, _ = synth.make_workload("gtr16", seed=2, nSites=50)
    r = tree.model.parts[0].rMatrices[0]
    assert abs(r.val.sum() - 1.0) < 1e-15


def test_effective_sample_size_ar1():
    # AR(1) with rho: ESS ~ N (1 - rho) / (1 + rho)
    rng = numpy.random.default_rng(5)
    N, rho = 200000, 0.8
    x = numpy.zeros(N)
    e = rng.normal(size=N)
    for i in range(1, N):
        x[i] = rho * x[i - 1] + e[i]
    ess = func.effectiveSampleSize(x)
    want = N * (1 - rho) / (1 + rho)
    assert abs(ess / want - 1) < 0.1
    iid = rng.normal(size=50000)
    assert abs(func.effectiveSampleSize(iid) / 50000 - 1) < 0.1
