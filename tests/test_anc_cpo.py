"""Root-side consumers of the partials: ancestral-state draws
(Pf/p4_treeSim.c:591-857, p4/tree.py:9640-9676) and CPO / LPML
(p4/mcmc.py:3286-3297, 3730-3748).

The oracle's draw restates the reference's picker; on the CPU it is pinned
by the reference's own root partials and site likelihoods (the table total
is the site likelihood) and by a brute-force sum over every assignment of
internal-node states on small trees (the table entries are the joint
probabilities of the root's state and category).  On the GPU the draws
for given uniforms must equal the oracle's, site by site, except where the
target falls within rounding of an entry boundary; the generator's draws
must equal the oracle's with Philox4x64-10 uniforms from numpy; the CPO
sums must equal the sums of 1/siteLike of the oracle's site likelihoods."""
import itertools

import numpy
import pytest

import oracle
from helpers import LNL_CASES, load, oracle_flat, tree_from_fixture
from p4gpu import P4Error, pf, synth
from p4gpu.mcmc import Mcmc


def _site_pat(fx):
    return numpy.repeat(numpy.arange(fx["patterns"].shape[1]), fx["counts"])


def _tipvec(code, S, equates):
    if code >= 0:
        v = numpy.zeros(S)
        v[code] = 1.0
        return v
    if code in (-1, -2):
        return numpy.ones(S)
    return numpy.array(equates[code + 64], float)


def _brute_joint(f):
    """[nPat][K][S]: sum over every assignment of internal-node states of
    comp[root state] * prod over branches of P, with the root's state and
    the category fixed -- the joint probability of the data and the root's
    (category, state), not through any conditional likelihood."""
    Pm = oracle.oracle_pmats(f)
    K, S, N = f.nCat, f.dim, len(f.parent)
    nP = f.patterns.shape[1]
    root = int(f.root)
    internal = [n for n in range(N) if not f.isLeaf[n] and n != root]
    leaves = [n for n in range(N) if f.isLeaf[n] and n != root]
    comp = f.comps[f.compNums[root]]
    out = numpy.zeros((nP, K, S))
    for k in range(K):
        # message of each leaf given its parent's state: [S][nP]
        lm = {}
        for l in leaves:
            tv = numpy.array([_tipvec(int(c), S, f.equates) for c in f.patterns[f.seqNum[l]]]).T   # [S][nP]
            lm[l] = Pm[l][k] @ tv
        for sr in range(S):
            tot = numpy.zeros(nP)
            for assign in itertools.product(range(S), repeat=len(internal)):
                st = dict(zip(internal, assign))
                st[root] = sr
                w = numpy.ones(nP)
                for n in internal:
                    w = w * Pm[n][k][st[f.parent[n]], st[n]]
                for l in leaves:
                    w = w * lm[l][st[f.parent[l]]]
                tot += w
            out[:, k, sr] = comp[sr] * tot
    return out


# ------------------------------------------------------------------ CPU
@pytest.mark.parametrize("case", ["multifurcation_root4", "edge_dna_gaps_pinv"])
def test_oracle_table_is_the_brute_force_joint(case):
    fx = load(case)
    f = oracle_flat(fx)
    _, _, cl = oracle.oracle_cl(f)
    W = oracle.anc_table(f, cl[int(f.root)])
    K, S = f.nCat, f.dim
    J = _brute_joint(f) * ((1.0 - f.pInvar) / K)
    got = W[:, :K * S].reshape(-1, K, S)
    assert numpy.abs(got - J).max() <= 1e-12 * J.max()


@pytest.mark.parametrize("case", [c for c in LNL_CASES if c != "leaf_root"])
def test_oracle_table_total_is_the_reference_site_likelihood(case):
    fx = load(case)
    f = oracle_flat(fx)
    if oracle.ref() is None:
        pytest.skip("reference subset not built")
    _, pl, _, cl, _ = oracle.ref_part_loglike(f, want_cl=True)
    W = oracle.anc_table(f, cl[int(f.root)])   # the reference's own root->cl
    assert numpy.allclose(W.sum(axis=1), pl, rtol=1e-12, atol=0)
    _, _, ocl = oracle.oracle_cl(f)
    r = int(f.root)
    scale = cl[r].max(axis=(0, 1))[None, None, :]   # per pattern
    assert (numpy.abs(ocl[r] - cl[r]) <= 1e-10 * scale).all()


def test_oracle_draw_order_and_frequencies():
    fx = load("edge_dna_gaps_pinv")
    f = oracle_flat(fx)
    _, _, cl = oracle.oracle_cl(f)
    K, S = f.nCat, f.dim
    W = oracle.anc_table(f, cl[int(f.root)])
    p = int(numpy.argmax(W[:, K * S:].sum(axis=1) > 0))   # a pattern with invariant entries
    nz = numpy.nonzero(W[p])[0]
    d, _ = oracle.anc_draw(f, cl[int(f.root)], [p, p], [0.0, 1.0 - 2.0 ** -53])
    first, last = nz[0], nz[-1]
    assert (d[0, 0], d[0, 1]) == (first % S, first // S)          # category-major, states inner
    assert d[1, 2] == 1 and d[1, 3] == last - K * S               # invariant states last
    # stratified uniforms reproduce the table's proportions
    M = 20000
    u = (numpy.arange(M) + 0.5) / M
    d, _ = oracle.anc_draw(f, cl[int(f.root)], numpy.full(M, p), u)
    idx = numpy.where(d[:, 2] == 1, K * S + d[:, 3], d[:, 1] * S + d[:, 0])
    freq = numpy.bincount(idx, minlength=K * S + S) / M
    assert numpy.abs(freq - W[p] / W[p].sum()).max() <= 2.0 / M


def test_host_philox_matches_numpy():
    assert [int(x) for x in pf.philox4x64([0, 0, 0, 0], [0, 0])] == \
        [0x16554d9eca36314c, 0xdb20fe9d672d0fdc, 0xd7e772cee186176b, 0x7e68b68aec7ba23b]
    rng = numpy.random.default_rng(3)
    for _ in range(50):
        c = rng.integers(0, 2 ** 63, 4, dtype=numpy.uint64) * numpy.uint64(2) + numpy.uint64(1)
        k = rng.integers(0, 2 ** 63, 2, dtype=numpy.uint64)
        assert (pf.philox4x64(c, k) == oracle.philox4x64(c, k)).all()


def test_oracle_cpo_and_lpml():
    sl = [numpy.array([0.5, 0.25]), numpy.array([0.25, 0.125])]
    s = oracle.cpo_sums(sl)
    assert numpy.allclose(s, [6.0, 12.0])
    assert abs(oracle.lpml(s, 2) - (-numpy.log(3.0) - numpy.log(6.0))) < 1e-15


# ------------------------------------------------------------------ GPU
def _check_draws(got, want, cum, u, tol):
    """Equal site by site, except where the target is within tol * total of
    an entry boundary (rounding of the table's running sums)."""
    bad = numpy.nonzero((got != want).any(axis=1))[0]
    for i in bad:
        t = u[i] * cum[i, -1]
        assert numpy.abs(cum[i] - t).min() <= tol * cum[i, -1], (i, got[i], want[i])
    assert len(bad) <= max(2, len(u) // 1000)


@pytest.mark.gpu
@pytest.mark.parametrize("case,prec", [(c, "fp64") for c in LNL_CASES] +
                         [("het64_800", "fp32"), ("prot128_150", "fp32"), ("edge_dna_gaps_pinv", "fp32")])
def test_gpu_draws_match_the_oracle(case, prec):
    fx = load(case)
    f = oracle_flat(fx)
    _, _, cl = oracle.oracle_cl(f)
    sp = _site_pat(fx)
    u = numpy.random.default_rng(11).random(len(sp))
    want, cum = oracle.anc_draw(f, cl[int(f.root)], sp, u)
    t = tree_from_fixture(fx, prec)
    t.calcLogLike(verbose=0)
    got = pf.drawAncStates(t.cTree, 0, 0, len(sp), u=u)
    _check_draws(got, want, cum, u, 1e-12 if prec == "fp64" else 1e-5)
    # a sub-range draws the same sites the same way
    a, b = len(sp) // 3, len(sp) // 3 + 37
    sub = pf.drawAncStates(t.cTree, 0, a, b - a, u=u[a:b])
    assert (sub == got[a:b]).all()


@pytest.mark.gpu
def test_gpu_generator_draws_match_the_oracle():
    fx = load("gtr16_500")
    f = oracle_flat(fx)
    _, _, cl = oracle.oracle_cl(f)
    sp = _site_pat(fx)
    seed = 0x5eed1234abcd
    u = oracle.anc_uniforms(seed, range(len(sp)))
    want, cum = oracle.anc_draw(f, cl[int(f.root)], sp, u)
    t = tree_from_fixture(fx, "fp64")
    t.calcLogLike(verbose=0)
    got = pf.drawAncStates(t.cTree, 0, 0, len(sp), seed=seed)
    _check_draws(got, want, cum, u, 1e-12)


@pytest.mark.gpu
def test_gpu_draw_frequencies_at_scale():
    """50,000 sites of the DNA workload drawn by the generator: the counts
    of each (category, state) / invariant-state class agree with the
    table's probabilities summed over sites (chi-square, 1e-4 level)."""
    tree, data = synth.make_workload("het64", seed=17, nSites=50000)
    tree.calcLogLike(verbose=0)
    n = pf.partNSites(data.parts[0].cPart)
    d = pf.drawAncStates(tree.cTree, 0, 0, n, seed=99)
    f = oracle.flat_from_tree(tree, 0)
    _, _, cl = oracle.oracle_cl(f)
    _, _, spi, _, _ = data.parts[0].patterns()
    K, S = f.nCat, f.dim
    W = oracle.anc_table(f, cl[int(f.root)])
    prob = W / W.sum(axis=1, keepdims=True)
    expect = prob[spi].sum(axis=0)
    idx = numpy.where(d[:, 2] == 1, K * S + d[:, 3], d[:, 1] * S + d[:, 0])
    obs = numpy.bincount(idx, minlength=K * S + S)
    keep = expect > 5
    chi2 = (((obs - expect) ** 2)[keep] / expect[keep]).sum()
    from scipy.stats import chi2 as chi2d
    assert chi2d.sf(chi2, keep.sum() - 1) > 1e-4, (chi2, keep.sum())
    assert obs[~keep].sum() <= max(20, 3 * expect[~keep].sum())


@pytest.mark.gpu
def test_ancestral_state_draw_string():
    fx = load("edge_dna_gaps_pinv")
    t = tree_from_fixture(fx, "fp64")
    s = t.ancestralStateDraw(seed=5)
    assert len(s) == int(fx["counts"].sum())
    assert set(s) <= set("acgt")
    assert s == t.ancestralStateDraw(seed=5)
    draw = numpy.zeros(4, numpy.int32)
    pf.p4_drawAncState(t.cTree, 0, 3, draw)
    assert (draw[1] >= 0 and draw[0] >= 0) or (draw[2] == 1 and draw[3] >= 0)
    with pytest.raises(P4Error, match="past"):
        pf.drawAncStates(t.cTree, 0, len(s) - 2, 5)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cpo_sums_are_the_oracle_site_likelihood_sums(prec):
    fx = load("edge_dna_gaps_pinv")
    t = tree_from_fixture(fx, prec)
    t.calcLogLike(verbose=0)
    cp = t.data.parts[0].cPart
    pf.cpoReset(cp)
    sp = _site_pat(fx)
    samples = []
    for i, bl in enumerate([None, 0.05, 0.2]):
        if bl is not None:
            t.nodes[1].br.len = bl
            t.calcLogLike(verbose=0)   # pushes the new length to the C tree
        lnl = pf.cpoAccumulate(t.cTree)
        f = oracle.flat_from_tree(t, 0)
        v, pl = oracle.part_loglike(f, site_likes=True)
        assert abs(lnl - v) <= (1e-10 if prec == "fp64" else 1e-6) * abs(v)
        samples.append(pl[sp])
    sums, n = pf.cpoGetSums(cp)
    assert n == 3
    want = oracle.cpo_sums(samples)
    assert numpy.allclose(sums, want, rtol=1e-11 if prec == "fp64" else 1e-5, atol=0)
    pf.cpoReset(cp)
    sums, n = pf.cpoGetSums(cp)
    assert n == 0 and not sums.any()


@pytest.mark.gpu
def test_mcmc_lpml_matches_the_sampled_site_likelihoods():
    tree, data = synth.make_workload("multi48", seed=23, nSites=600)
    m = Mcmc(tree, nChains=2, sampleInterval=5, seed=4, verbose=False)
    m.doCpo = True
    m.cpo_startGen = 10
    seen = []

    def hook(t):
        if m.gen + 1 > m.cpo_startGen:
            seen.append(numpy.array(t.getSiteLikes()))
    m.sampleHooks.append(hook)
    m.run(60)
    assert m.cpo_nSamples == len(seen) > 0
    want = oracle.cpo_sums(seen)
    assert numpy.allclose(m.cpo_sumsOfInverseSiteLikes, want, rtol=1e-11, atol=0)
    assert abs(m.lpml(verbose=False) - oracle.lpml(want, len(seen))) <= 1e-9 * abs(oracle.lpml(want, len(seen)))
