"""GPU parity of libp4gpu against the reference's golden vectors and the
oracle.  Tolerances: FP64 path 1e-10 relative on lnL (the project target),
FP32 path with scale exponents 1e-6 relative on lnL."""
import numpy
import pytest

import oracle
from helpers import LNL_CASES, load, oracle_flat, tree_from_fixture
from p4gpu import Alignment, Data, P4Error, func, pf, synth

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_TOL = 1e-6


def rel(a, b):
    return abs(a - b) / abs(b)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert pf.deviceAvailable(), "libp4gpu sees no GPU"


# --------------------------------------------------------- golden vectors
@pytest.mark.parametrize("case", LNL_CASES)
def test_golden_fp64(case):
    fx = load(case)
    t = tree_from_fixture(fx, "fp64")
    v = t.calcLogLike(verbose=0)
    assert rel(v, float(fx["lnL"])) < FP64_TOL, (v, float(fx["lnL"]))
    sl = numpy.array(t.getSiteLikes())
    # a part built from patterns lists its sites in pattern order
    want = numpy.repeat(fx["patternLikes"], fx["counts"])
    assert numpy.allclose(sl, want, rtol=1e-9, atol=0)


@pytest.mark.parametrize("case", LNL_CASES)
def test_golden_fp32(case):
    fx = load(case)
    t = tree_from_fixture(fx, "fp32")
    v = t.calcLogLike(verbose=0)
    assert rel(v, float(fx["lnL"])) < FP32_TOL, (v, float(fx["lnL"]))


@pytest.mark.parametrize("case", ["gtr16_500", "het64_800", "prot128_150"])
def test_golden_p_matrices(case):
    fx = load(case)
    t = tree_from_fixture(fx, "fp64")
    t.calcLogLike(verbose=0)
    K, S = int(fx["nCat"]), int(fx["dim"])
    for n in range(1, len(fx["P"])):
        if n == int(fx["root"]):
            continue
        got = pf.p4_getBigPDeck(t.nodes[n].cNode, 0, K, S)
        assert numpy.abs(got - fx["P"][n]).max() < 2e-14, n


@pytest.mark.parametrize("case,prec,tol", [("het64_800", "fp64", 1e-10), ("het64_800", "fp32", 1e-4),
                                           ("prot128_150", "fp64", 1e-10), ("prot128_150", "fp32", 1e-4),
                                           ("prot128_150", "fp32-tile", 1e-4)])
def test_partials_match_oracle(case, prec, tol):
    """Every stored partial, unscaled, against the oracle's: 4-state (wave
    or lean kernel) and 20-state (FP32 on MFMA with one wavefront per
    (tile, category), or per tile; FP64 on the grouped kernel).  The error
    bound is relative to the largest entry of the same pattern (FP32 entries
    far below that largest one lose relative precision, harmlessly)."""
    fx = load(case)
    tile = prec == "fp32-tile"
    prec = "fp32" if tile else prec
    if tile:
        pf.setAaTuning(False)
    try:
        t = tree_from_fixture(fx, prec)
        t.calcLogLike(verbose=0)
    finally:
        pf.setAaTuning(True)
    f = oracle_flat(fx)
    _, _, cl = oracle.oracle_cl(f)
    K, S, nP = int(fx["nCat"]), int(fx["dim"]), f.patterns.shape[1]
    for n in t.nodes:
        if n.isLeaf or n is t.root:
            continue
        got = pf.p4_getCondLikes(n.cNode, 0, K, S, nP)
        want = cl[n.nodeNum]
        scale = want.max(axis=(0, 1))[None, None, :]
        assert (numpy.abs(got - want) <= tol * scale).all(), n.nodeNum


@pytest.mark.parametrize("cw,slots", [(False, 5), (True, 5), (True, 2), (True, 0)])
def test_protein_ambiguity_partials_match_oracle(cw, slots):
    """20-state tips with partial ambiguities (b = d|n, z = e|q), full ones
    (x, gaps, '?') and pInvar 0.2, on a tree whose root has three children:
    every stored partial against the oracle's, and lnL, for the tile kernel
    and for category waves with 5, 2 and 0 LDS slots (0: every child read
    back from HBM).  These take the kernels' rare paths (an ambiguous tip's
    summed P rows, HBM children, a third root child)."""
    fx = dict(load("prot128_150"))
    rng = numpy.random.default_rng(11)
    pat = fx["patterns"].copy()
    hit = rng.random(pat.shape) < 0.12
    # equate codes in the fixture's order (b, x, z), then gap and '?'
    pat[hit] = rng.choice(numpy.array([-64, -62, -63, -1, -2]), size=int(hit.sum()))
    fx["patterns"] = pat
    fx["pInvar"] = numpy.array(0.2)
    assert int((fx["parent"] == int(fx["root"])).sum()) == 3
    f = oracle_flat(fx)
    wantL, _, cl = oracle.oracle_cl(f)
    K, S, nP = int(fx["nCat"]), int(fx["dim"]), f.patterns.shape[1]
    try:
        pf.setAaTuning(cw, 0)
        pf.setAaLdsSlots(slots)
        t = tree_from_fixture(fx, "fp32")
        v = t.calcLogLike(verbose=0)
        got = {n.nodeNum: pf.p4_getCondLikes(n.cNode, 0, K, S, nP) for n in t.nodes
               if not (n.isLeaf or n is t.root)}
    finally:
        pf.setAaTuning(True)
        pf.setAaLdsSlots(5)
    assert rel(v, wantL) < FP32_TOL, (v, wantL)
    for num, g in got.items():
        want = cl[num]
        scale = want.max(axis=(0, 1))[None, None, :]
        assert (numpy.abs(g - want) <= 1e-4 * scale).all(), num


# -------------------------------------------------- full-size workloads
@pytest.mark.parametrize("name,nSites", [("gtr16", None), ("het64", None), ("prot128", 4000)])
def test_workloads_vs_oracle(name, nSites):
    tree, _ = synth.make_workload(name, seed=11, nSites=nSites)
    want = oracle.part_loglike(oracle.flat_from_tree(tree, 0))
    v64 = tree.calcLogLike(verbose=0)
    tree.setPrecision("fp32")
    v32 = tree.calcLogLike(verbose=0)
    tree.setPrecision("fp64")
    again = tree.calcLogLike(verbose=0)
    assert rel(v64, want) < FP64_TOL
    assert rel(v32, want) < FP32_TOL
    assert again == v64


def test_full_size_protein_fp32_vs_fp64():
    tree, _ = synth.make_workload("prot128", seed=12)
    v64 = tree.calcLogLike(verbose=0)
    tree.setPrecision("fp32")
    v32 = tree.calcLogLike(verbose=0)
    assert rel(v32, v64) < FP32_TOL


def test_multipart_tree():
    tree, data = synth.make_workload("multi48", seed=3, nSites=4000)
    v = tree.calcLogLike(verbose=0)
    parts = [oracle.part_loglike(oracle.flat_from_tree(tree, p)) for p in range(data.nParts)]
    assert rel(v, sum(parts)) < FP64_TOL
    assert numpy.allclose(tree.partLikes, parts, rtol=FP64_TOL)


# ----------------------------------------------- incremental evaluation
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_dirty_path_equals_full(prec):
    tree, _ = synth.make_workload("het64", seed=4, nSites=3000, precision=prec)
    tree.calcLogLike(verbose=0)
    c = tree.cTree
    rng = numpy.random.default_rng(0)
    for _ in range(5):
        n = tree.nodes[int(rng.integers(1, len(tree.nodes)))]
        n.br.len *= float(rng.uniform(0.5, 2.0))
        pf.p4_setBrLen(n.cNode, n.br.len)
        pf.p4_calculateBigPDecks(n.cNode)
        p = n.parent
        while p is not None:
            pf.p4_setConditionalLikelihoodsOfInternalNodePart(p.cNode, 0)
            p = p.parent
        inc = pf.p4_partLogLike(c, tree.data.parts[0].cPart, 0, 0)
        full = pf.p4_treeLogLike(c, 0)
        assert rel(inc, full) < (1e-13 if prec == "fp64" else 1e-9), (inc, full)


def test_batch_equals_single():
    trees = []
    for seed, prec in [(1, "fp64"), (2, "fp32"), (3, "fp32")]:
        t, _ = synth.make_workload("gtr16", seed=seed, nSites=1500, precision=prec)
        t.calcLogLike(verbose=0)
        trees.append(t)
    t4, _ = synth.make_workload("multi48", seed=5, nSites=2000, precision="fp32")
    t4.calcLogLike(verbose=0)
    trees.append(t4)
    single = [pf.p4_treeLogLike(t.cTree, 0) for t in trees]
    for t in trees:
        for n in t.iterInternals():
            pf.p4_setConditionalLikelihoodsOfInternalNode(n.cNode)
    b = pf.Batch([t.cTree for t in trees])
    got = b.logLikes()
    for g, s in zip(got, single):
        assert rel(g, s) < 1e-13


def test_copy_and_verify():
    a, _ = synth.make_workload("gtr16", seed=8, nSites=800)
    a.calcLogLike(verbose=0)
    b = a.dupe()
    b.calcLogLike(verbose=0)
    assert pf.p4_verifyIdentityOfTwoTrees(a.cTree, b.cTree) == 0
    n = b.nodes[3]
    n.br.len *= 1.7
    b.calcLogLike(verbose=0)
    assert pf.p4_verifyIdentityOfTwoTrees(a.cTree, b.cTree) == 1
    # copy a's state into b: identical again
    n.br.len /= 1.7
    pf.p4_setBrLen(n.cNode, n.br.len)
    pf.p4_copyCondLikes(a.cTree, b.cTree, 1)
    pf.p4_copyBigPDecks(a.cTree, b.cTree, 1)
    pf.p4_copyModelPrams(a.cTree, b.cTree)
    assert pf.p4_verifyIdentityOfTwoTrees(a.cTree, b.cTree) == 0
    assert pf.p4_partLogLike(b.cTree, b.data.parts[0].cPart, 0, 0) == pf.p4_treeLogLike(a.cTree, 0)


# ------------------------------------------------------------ error paths
def test_comp_below_pivec_min_raises():
    tree, _ = synth.make_workload("gtr16", seed=9, nSites=200)
    tree.calcLogLike(verbose=0)
    c = tree.model.parts[0].comps[0]
    c.val[:] = [1e-15, 0.5, 0.25, 0.25 - 1e-15]
    with pytest.raises(P4Error, match="Bad"):
        pf.p4_setPrams(tree.cTree, -1)


def test_comp_not_summing_to_one_raises():
    tree, _ = synth.make_workload("gtr16", seed=9, nSites=200)
    tree.calcLogLike(verbose=0)
    tree.model.parts[0].comps[0].val[:] = [0.3, 0.3, 0.3, 0.3]
    with pytest.raises(P4Error, match="sum to 1"):
        pf.p4_setPrams(tree.cTree, -1)


def test_underflow_sentinel_fp64_and_scaling_fp32():
    """FP64 without scaling underflows on a 300-taxon, long-branch protein
    tree and returns the reference's -1e99; FP32 with scale exponents
    returns a finite value."""
    rng = numpy.random.default_rng(1)
    t = func.randomTree(nTax=300, seed=1, meanBrLen=2.5)
    names = t.taxNames
    seqs = [''.join(rng.choice(list('arndcqeghilkmfpstwyv'), 40)) for _ in names]
    t.data = Data(Alignment(names, seqs, 'protein'))
    t.newComp(free=1, spec='equal')
    t.newRMatrix(free=1, spec='ones')
    v = t.calcLogLike(verbose=0)
    assert v == -1.0e99
    t.setPrecision("fp32")
    v32 = t.calcLogLike(verbose=0)
    assert -1e7 < v32 < 0


# ------------------------------------------------ other state counts, K
@pytest.mark.parametrize("symbols", ["01", "012345"])
def test_standard_data_generic_state_count(symbols):
    rng = numpy.random.default_rng(len(symbols))
    t = func.randomTree(nTax=10, seed=len(symbols), meanBrLen=0.2)
    names = t.taxNames
    seqs = [''.join(rng.choice(list(symbols) + ['-'], 300)) for _ in names]
    t.data = Data(Alignment(names, seqs, 'standard', symbols=symbols))
    S = len(symbols)
    t.newComp(free=1, spec='equal')
    t.newRMatrix(free=1, spec='specified', val=list(rng.lognormal(0, 0.5, S * (S - 1) // 2)))
    t.setNGammaCat(nGammaCat=3)
    t.newGdasrv(free=1, val=0.8)
    t.setPInvar(free=1, val=0.1)
    want = oracle.part_loglike(oracle.flat_from_tree(t, 0))
    assert rel(t.calcLogLike(verbose=0), want) < FP64_TOL
    t.setPrecision("fp32")
    assert rel(t.calcLogLike(verbose=0), want) < FP32_TOL


def test_no_gamma_single_category():
    tree, _ = synth.make_workload("gtr16", seed=6, nSites=600)
    mp = tree.model.parts[0]
    tree.deleteCStuff()
    mp.gdasrvs = []
    tree.setNGammaCat(nGammaCat=1)
    for n in tree.nodes:
        n.br.parts[0].gdasrvNum = -1
    tree.setPInvar(free=0, val=0.15)
    want = oracle.part_loglike(oracle.flat_from_tree(tree, 0))
    assert rel(tree.calcLogLike(verbose=0), want) < FP64_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
def test_wave_kernel_variants_match_golden(prec, tol):
    """The pair kernel, every variant of the wave kernel (P by scalar or
    vector loads, one or two streams) and the grouped kernel give the
    reference lnL (the autotuner may pick any of them)."""
    fp32 = prec == "fp32"
    try:
        for case in ("het64_800", "edge_dna_gaps_pinv"):   # the second has gaps and partial ambiguities
            fx = load(case)
            for G in (-3, -1, -2, 0, 4):
                pf.setPruneTuning(fp32, G, 1)
                t = tree_from_fixture(fx, prec)
                v = t.calcLogLike(verbose=0)
                assert rel(v, float(fx["lnL"])) < tol, (case, G, v, float(fx["lnL"]))
    finally:
        pf.setPruneTuning(fp32, -3, 1)   # the library default, then the tuned choice
        pf._applyTuning()


def _bigq_normalised(pi, R):
    S = len(pi)
    Q = R * pi[None, :]
    numpy.fill_diagonal(Q, 0.0)
    numpy.fill_diagonal(Q, -Q.sum(1))
    return Q / -(pi * numpy.diag(Q)).sum()


@pytest.mark.parametrize("case", ["gtr16_500", "het64_800", "prot128_150"])
def test_p_routes_match_golden(case):
    """4- and 20-state P matrices straight from exp(Q t) (the default) and
    from the eigensystem both give the reference's P matrices and lnL; switching
    from the direct route to the eigen route mid-life recomputes the
    eigensystems skipped meanwhile."""
    fx = load(case)
    K, S = int(fx["nCat"]), int(fx["dim"])
    try:
        for route in (True, False, True):
            pf.setPmatRoute(route)
            t = tree_from_fixture(fx, "fp64")
            assert rel(t.calcLogLike(verbose=0), float(fx["lnL"])) < FP64_TOL, route
            for n in range(1, len(fx["P"])):
                if n == int(fx["root"]):
                    continue
                got = pf.p4_getBigPDeck(t.nodes[n].cNode, 0, K, S)
                assert numpy.abs(got - fx["P"][n]).max() < 2e-14, (route, n)
        # a full evaluation on the direct route, then a one-branch change on
        # the eigen route (chosen, or taken by 20-state parts for a flush of
        # few matrices): it needs the eigensystems that the direct route skipped
        for second in (False, True):
            pf.setPmatRoute(True)
            t = tree_from_fixture(fx, "fp64")
            t.calcLogLike(verbose=0)
            pf.setPmatRoute(second)
            nd = t.nodes[1]
            pf.p4_setBrLen(nd.cNode, 0.2)
            pf.p4_calculateBigPDecks(nd.cNode)
            got = pf.p4_getBigPDeck(nd.cNode, 0, K, S)
            c, r = int(fx["compNums"][1]), int(fx["rMatrixNums"][1])
            Q = _bigq_normalised(numpy.asarray(fx["comps"][c], float), numpy.asarray(fx["bigR"][r], float))
            g = int(fx["gdasrvNums"][1])
            for k in range(K):
                tk = 0.2 * float(fx["rates"][g][k]) * float(fx["relRate"]) / (1.0 - float(fx["pInvar"]))
                want = scipy_expm(Q * tk)
                assert numpy.abs(got[k] - want).max() < 1e-13, (second, k)
    finally:
        pf.setPmatRoute(True)


def scipy_expm(A):
    import scipy.linalg
    return scipy.linalg.expm(A)


def test_direct_p_extreme_branch_lengths():
    """exp(Q t) by the scaled Taylor series against scipy's expm and the
    eigen route, from 1e-9 to 200 substitutions per site; short branches
    keep their off-diagonal entries to full relative precision."""
    fx = load("gtr16_500")
    K, S = int(fx["nCat"]), int(fx["dim"])
    lens = [1e-9, 1e-5, 1e-3, 0.05, 0.7, 3.0, 20.0, 200.0]
    pi = numpy.asarray(fx["comps"][0], float)
    Q = _bigq_normalised(pi, numpy.asarray(fx["bigR"][0], float))
    rates = numpy.asarray(fx["rates"][0], float) * float(fx["relRate"]) / (1.0 - float(fx["pInvar"]))
    decks = {}
    try:
        for route in (True, False):
            pf.setPmatRoute(route)
            t = tree_from_fixture(fx, "fp64")
            t.calcLogLike(verbose=0)
            nodes = [n for n in range(1, len(fx["parent"])) if n != int(fx["root"])][:len(lens)]
            for n, L in zip(nodes, lens):
                pf.p4_setBrLen(t.nodes[n].cNode, L)
                pf.p4_calculateBigPDecks(t.nodes[n].cNode)
            decks[route] = [pf.p4_getBigPDeck(t.nodes[n].cNode, 0, K, S) for n in nodes]
    finally:
        pf.setPmatRoute(True)
    off = ~numpy.eye(S, dtype=bool)
    for i, L in enumerate(lens):
        for k in range(K):
            want = scipy_expm(Q * L * rates[k])
            got = decks[True][i][k]
            tol = 1e-14 if L * rates[k] < 5 else 1e-11
            assert numpy.abs(got - want).max() < tol, (L, k)
            assert numpy.abs(got - decks[False][i][k]).max() < max(tol, 1e-12), (L, k)
            assert numpy.allclose(got.sum(1), 1.0, rtol=0, atol=1e-13)
            if L * rates[k] < 1e-2:
                assert (numpy.abs(got - want)[off] / want[off]).max() < 1e-12, (L, k)


@pytest.mark.parametrize("nCat", [4, 3])
def test_protein_category_waves_match_tile_waves(nCat):
    """The 20-state MFMA kernel with one wavefront per (tile, category)
    stores the same partials and exponents, bit for bit, as the one with a
    wavefront per tile holding every category; lnL agrees to 1e-12 (only
    the FP64 root sum over categories is added in another order) and with
    the FP64 path to the FP32 tolerance.  nCat 3: 192-thread workgroups."""
    fx = dict(load("prot128_150"))
    fx["nCat"] = numpy.array(nCat)
    K, S, nP = nCat, int(fx["dim"]), fx["patterns"].shape[1]
    want = tree_from_fixture(fx, "fp64").calcLogLike(verbose=0)
    if nCat == 4:
        assert rel(want, float(fx["lnL"])) < FP64_TOL
    res = {}
    try:
        # tile waves; category waves with children from HBM; category waves
        # forwarding children through the LDS slot file (the default)
        for cw, slots in ((False, 5), (True, 0), (True, 5), (True, 2)):
            pf.setAaTuning(cw, 0)
            pf.setAaLdsSlots(slots)
            t = tree_from_fixture(fx, "fp32")
            v = t.calcLogLike(verbose=0)
            assert rel(v, want) < FP32_TOL, (cw, slots, v, want)
            cls = [pf.p4_getCondLikes(n.cNode, 0, K, S, nP) for n in t.nodes if not (n.isLeaf or n is t.root)]
            res[(cw, slots)] = (v, cls)
    finally:
        pf.setAaTuning(True)
        pf.setAaLdsSlots(5)
    for key in ((True, 0), (True, 5), (True, 2)):
        assert rel(res[key][0], res[(False, 5)][0]) < 1e-12, key
        for a, b in zip(res[key][1], res[(False, 5)][1]):
            assert numpy.array_equal(a, b), key


def test_batch_mixes_protein_kernels():
    """One batched call where one tree's protein parts run long programs
    (category waves) and another's short dirty paths (tile kernel): every
    lnL equals the tree's full evaluation."""
    trees = [synth.make_workload("multi48", seed=s, nSites=1500, precision="fp32")[0] for s in (6, 7)]
    want = [t.calcLogLike(verbose=0) for t in trees]
    for n in trees[0].iterInternals():
        pf.p4_setConditionalLikelihoodsOfInternalNode(n.cNode)
    n = next(iter(trees[1].iterLeavesNoRoot())).parent
    while n is not None:
        pf.p4_setConditionalLikelihoodsOfInternalNode(n.cNode)
        n = n.parent
    got = pf.Batch([t.cTree for t in trees]).logLikes()
    for g, w in zip(got, want):
        assert rel(g, w) < 1e-12, (g, w)


def test_batch_direct_p_in_one_launch():
    """Every branch length of several multi-partition chains scaled by 1.3,
    then setPrams and one batched call: the 4- and 20-state P matrices of
    every chain come from one launch per state count.  They must equal, bit
    for bit, those of the same trees evaluated one by one after the same
    change (so a batched P launch that skipped or misrouted an item would
    leave the old, different matrices), and every lnL must equal the
    single-tree evaluation."""
    def make():
        return [synth.make_workload("multi48", seed=s, nSites=1500, precision="fp32")[0] for s in (8, 9, 10)]

    def scale(t):
        for n in t.iterNodesNoRoot():
            n.br.len *= 1.3
            pf.p4_setBrLen(n.cNode, n.br.len)

    A, B = make(), make()
    before = []
    for t in A + B:
        t.calcLogLike(verbose=0)
    for t in A:
        mp0 = {p: mp for p, mp in enumerate(t.model.parts)}
        before.append({(p, n.nodeNum): pf.p4_getBigPDeck(n.cNode, p, mp.nGammaCat, mp.dim)
                       for p, mp in mp0.items() for n in t.iterNodesNoRoot()})
    for t in A:
        scale(t)
        pf.p4_setPrams(t.cTree, -1)
        for n in t.iterInternals():
            pf.p4_setConditionalLikelihoodsOfInternalNode(n.cNode)
    got = pf.Batch([t.cTree for t in A]).logLikes()
    want = []
    for t in B:
        scale(t)
        pf.p4_setPrams(t.cTree, -1)
        want.append(pf.p4_treeLogLike(t.cTree, 0))
    for g, w in zip(got, want):
        assert rel(g, w) < 1e-12, (g, w)
    for ta, tb, b4 in zip(A, B, before):
        dims = set()
        for (p, nn), old in b4.items():
            mp = ta.model.parts[p]
            pa = pf.p4_getBigPDeck(ta.nodes[nn].cNode, p, mp.nGammaCat, mp.dim)
            pb = pf.p4_getBigPDeck(tb.nodes[nn].cNode, p, mp.nGammaCat, mp.dim)
            assert numpy.array_equal(pa, pb), (p, nn)
            assert not numpy.array_equal(pa, old), (p, nn)
            dims.add(mp.dim)
        assert dims == {4, 20}


def test_batch_eigen_route_p_in_one_launch():
    """A branch move on every chain of a multi-partition batch: 20-state
    parts recompute that branch's P on the eigen route (a flush of few
    matrices), all items in one launch; P decks equal, bit for bit, and lnL
    agrees with the same trees evaluated one by one."""
    def make():
        return [synth.make_workload("multi48", seed=s, nSites=1500, precision="fp32")[0] for s in (11, 12)]

    def move(t):
        n = list(t.iterNodesNoRoot())[5]
        pf.p4_setBrLen(n.cNode, n.br.len * 1.3)
        pf.p4_calculateBigPDecks(n.cNode)
        m = n.parent
        while m is not None:
            for p in range(t.model.nParts):
                pf.p4_setConditionalLikelihoodsOfInternalNodePart(m.cNode, p)
            m = m.parent
        return n.nodeNum

    A, B = make(), make()
    for t in A + B:
        t.calcLogLike(verbose=0)
    moved = [move(t) for t in A]
    for t in B:
        move(t)
    got = pf.Batch([t.cTree for t in A]).logLikes()
    want = [pf.p4_treeLogLike(t.cTree, 0) for t in B]
    for g, w in zip(got, want):
        assert rel(g, w) < 1e-12, (g, w)
    for ta, tb, nn in zip(A, B, moved):
        dims = set()
        for p, mp in enumerate(ta.model.parts):
            pa = pf.p4_getBigPDeck(ta.nodes[nn].cNode, p, mp.nGammaCat, mp.dim)
            pb = pf.p4_getBigPDeck(tb.nodes[nn].cNode, p, mp.nGammaCat, mp.dim)
            assert numpy.array_equal(pa, pb), (p, nn)
            dims.add(mp.dim)
        assert dims == {4, 20}


def test_batch_error_keeps_staged_p_matrices():
    """A batched call whose second tree fails to stage (a postorder holding
    an out-of-range node) still computes the P matrices it deferred for the
    first tree: they equal those of the same move evaluated alone."""
    def make():
        return synth.make_workload("multi48", seed=13, nSites=1500, precision="fp32")[0]

    def move(t):
        n = list(t.iterNodesNoRoot())[5]
        pf.p4_setBrLen(n.cNode, n.br.len * 1.7)
        pf.p4_calculateBigPDecks(n.cNode)
        return n.nodeNum

    a, ctrl = make(), make()
    bad = synth.make_workload("gtr16", seed=2, nSites=300, precision="fp32")[0]
    for t in (a, ctrl, bad):
        t.calcLogLike(verbose=0)
    nn = move(a)
    move(ctrl)
    saved = int(bad.postOrder[0])
    bad.postOrder[0] = 10 ** 6
    try:
        with pytest.raises(P4Error, match="postOrder"):
            pf.Batch([a.cTree, bad.cTree]).logLikes()
    finally:
        bad.postOrder[0] = saved
    pf.p4_treeLogLike(ctrl.cTree, 0)
    for p, mp in enumerate(a.model.parts):
        got = pf.p4_getBigPDeck(a.nodes[nn].cNode, p, mp.nGammaCat, mp.dim)
        want = pf.p4_getBigPDeck(ctrl.nodes[nn].cNode, p, mp.nGammaCat, mp.dim)
        assert numpy.array_equal(got, want), p
