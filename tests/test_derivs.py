"""Branch-length derivatives and up-pass partials (the reference's Newton
path: Pf/eig.c:294-395, Pf/p4_node.c:442-540 and 858-1018,
Pf/p4_treeNewt.c:210-620).

The golden vectors (tests/golden/derivs_*.npz) come from the reference's
own P', P'' and cl2 code; on the CPU they are pinned by finite differences
of the oracle's log likelihood and by the identity that the likelihood
computed across any branch is the tree's.  On the GPU every branch's
derivatives, P', P'' and cl2 are checked against them, the identity is
checked on every fixture (a leaf root included), and finite differences of
the GPU's own likelihood pin the rest."""
import numpy
import pytest

import oracle
from helpers import LNL_CASES, load, oracle_flat, tree_from_fixture
from p4gpu import P4Error, pf, synth

DERIV_CASES = ["gtr16_500", "het64_800", "prot128_150", "edge_dna_gaps_pinv", "multifurcation_root4"]


def _fd(lnl_at, t, h):
    v = [lnl_at(t + k * h) for k in (-2, -1, 0, 1, 2)]
    d1 = (-v[4] + 8 * v[3] - 8 * v[1] + v[0]) / (12 * h)
    d2 = (-v[4] + 16 * v[3] - 30 * v[2] + 16 * v[1] - v[0]) / (12 * h * h)
    return d1, d2


# ------------------------------------------------------------------- CPU
@pytest.mark.parametrize("case", DERIV_CASES)
def test_golden_derivatives_are_those_of_the_oracle_likelihood(case):
    fx = load(case)
    g = load("derivs_" + case)
    f = oracle_flat(fx)
    v = oracle.part_loglike(f)
    root = int(fx["root"])
    nonroot = [n for n in range(len(fx["parent"])) if n != root]
    # the likelihood across every branch is the tree's
    assert max(abs(g["lnLEdge"][n] - v) for n in nonroot) <= 1e-11 * abs(v)
    for n in nonroot[:5]:
        def lnl_at(x, n=n):
            kw = dict(f.kwargs)
            bl = numpy.array(kw["brLen"], float)
            bl[n] = x
            kw["brLen"] = bl
            return oracle.part_loglike(oracle.Flat(**kw))
        t = float(fx["brLen"][n])
        d1, d2 = _fd(lnl_at, t, 1e-4 * max(t, 1e-2))
        assert abs(d1 - g["d1"][n]) <= 1e-5 * max(1.0, abs(g["d1"][n])), (n, d1, g["d1"][n])
        assert abs(d2 - g["d2"][n]) <= 1e-2 * max(1.0, abs(g["d2"][n])), (n, d2, g["d2"][n])


def test_golden_p_derivatives_are_q_times_p():
    """P' = r Q P and P'' = r^2 Q^2 P for the reference's own P' and P''
    (gtr16_500: one comp, one rMatrix, rates from its gamma shape)."""
    fx = load("gtr16_500")
    g = load("derivs_gtr16_500")
    f = oracle_flat(fx)
    _, _, Pm, _, rates = oracle.ref_part_loglike(f)
    Q, _ = oracle.ref_pmat(fx["bigR"][0], fx["comps"][0], numpy.array([0.1]))
    for n in range(1, len(g["P1"])):
        if n == int(fx["root"]):
            continue
        for k in range(int(fx["nCat"])):
            r = rates[0][k]
            assert numpy.abs(g["P1"][n][k] - r * Q @ Pm[n][k]).max() < 1e-13
            assert numpy.abs(g["P2"][n][k] - r * r * Q @ Q @ Pm[n][k]).max() < 1e-13


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("case", DERIV_CASES)
def test_gpu_derivatives_match_the_reference(case):
    fx = load(case)
    g = load("derivs_" + case)
    t = tree_from_fixture(fx, "fp64")
    lnl, d1, d2 = t.branchDerivatives()
    root = int(fx["root"])
    nonroot = [n for n in range(len(fx["parent"])) if n != root]
    for n in nonroot:
        assert abs(lnl[n] - g["lnLEdge"][n]) <= 1e-10 * abs(g["lnLEdge"][n]), n
        assert abs(d1[n] - g["d1"][n]) <= 1e-9 * max(1.0, abs(g["d1"][n])), (n, d1[n], g["d1"][n])
        assert abs(d2[n] - g["d2"][n]) <= 1e-9 * max(1.0, abs(g["d2"][n])), (n, d2[n], g["d2"][n])
    K, S = int(fx["nCat"]), int(fx["dim"])
    for n in range(1, len(g["P1"])):
        if n == root:
            continue
        p1 = pf.p4_getBigPDeck1stD(t.nodes[n].cNode, 0, K, S)
        p2 = pf.p4_getBigPDeck2ndD(t.nodes[n].cNode, 0, K, S)
        s1 = max(1.0, numpy.abs(g["P1"][n]).max())
        s2 = max(1.0, numpy.abs(g["P2"][n]).max())
        assert numpy.abs(p1 - g["P1"][n]).max() <= 1e-12 * s1, n
        assert numpy.abs(p2 - g["P2"][n]).max() <= 1e-12 * s2, n
    nP = fx["patterns"].shape[1]
    for n, want in zip(g["cl2Nodes"], g["cl2"]):
        got = pf.p4_getCondLikes2(t.nodes[int(n)].cNode, 0, K, S, nP)
        scale = want.max(axis=(0, 1))[None, None, :]
        assert (numpy.abs(got - want) <= 1e-11 * scale).all(), int(n)


@pytest.mark.gpu
@pytest.mark.parametrize("case", LNL_CASES)
def test_likelihood_across_every_branch_is_the_trees(case):
    """cl2 x P x cl summed over states, categories and patterns gives the
    tree's log likelihood at every branch -- a leaf root included, whose
    data the up pass takes in (the reference's up pass starts from the root
    frequencies alone, Pf/p4_node.c:858-879, which leaves a leaf root's own
    data out)."""
    fx = load(case)
    t = tree_from_fixture(fx, "fp64")
    v = t.calcLogLike(verbose=0)
    lnl, _, _ = t.branchDerivatives()
    for n in range(len(t.nodes)):
        if t.nodes[n] is t.root:
            continue
        assert abs(lnl[n] - v) <= 1e-10 * abs(v), (case, n, lnl[n], v)


@pytest.mark.gpu
@pytest.mark.parametrize("name,nSites", [("het64", 3000), ("prot128", 400)])
def test_derivatives_are_finite_differences_of_the_gpu_likelihood(name, nSites):
    tree, _ = synth.make_workload(name, seed=61, nSites=nSites)
    lnl, d1, d2 = tree.branchDerivatives()
    nodes = [n for n in tree.iterNodesNoRoot()]
    for n in nodes[:: max(1, len(nodes) // 6)]:
        t0 = n.br.len

        def at(x):
            n.br.len = x
            return tree.calcLogLike(verbose=0)
        f1, f2 = _fd(at, t0, 1e-4 * max(t0, 1e-2))
        n.br.len = t0
        assert abs(f1 - d1[n.nodeNum]) <= 1e-5 * max(1.0, abs(d1[n.nodeNum])), (n.nodeNum, f1, d1[n.nodeNum])
        assert abs(f2 - d2[n.nodeNum]) <= 1e-2 * max(1.0, abs(d2[n.nodeNum])), (n.nodeNum, f2, d2[n.nodeNum])


@pytest.mark.gpu
def test_multipart_derivatives_sum_over_parts():
    tree, data = synth.make_workload("multi48", seed=62, nSites=1500)
    allp = tree.branchDerivatives(-1)
    parts = [tree.branchDerivatives(p) for p in range(data.nParts)]
    for k in range(3):
        s = sum(pd[k] for pd in parts)
        assert numpy.allclose(allp[k], s, rtol=1e-12, atol=1e-9)
    v = tree.calcLogLike(verbose=0)
    assert abs(allp[0][1] - v) <= 1e-10 * abs(v)


@pytest.mark.gpu
def test_newton_on_all_branches_reaches_a_stationary_point():
    tree, _ = synth.make_workload("gtr16", seed=63, nSites=2000)
    for n in tree.iterNodesNoRoot():   # start away from the simulated lengths
        n.br.len = 0.05
    before = tree.calcLogLike(verbose=0)
    after = tree.optBrLensNewton(epsilon=1e-4)
    assert after > before
    lnl, d1, d2 = tree.branchDerivatives()
    for n in tree.iterNodesNoRoot():
        if n.br.len > 2e-8:
            # at a maximum along every branch: the Newton gain left is tiny
            g1, g2 = d1[n.nodeNum], d2[n.nodeNum]
            assert g2 < 0.0, (n.nodeNum, g2)
            assert g1 * g1 / (2.0 * -g2) < 1e-6, (n.nodeNum, n.br.len, g1, g2)


@pytest.mark.gpu
def test_derivatives_need_fp64_partials():
    tree, _ = synth.make_workload("gtr16", seed=64, nSites=300, precision="fp32")
    tree.calcLogLike(verbose=0)
    with pytest.raises(P4Error, match="FP64"):
        tree.branchDerivatives()
