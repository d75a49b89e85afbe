"""The learned proposal: its network's kernels against the CPU oracle
(oracle/mlp_oracle.py), the oracle against the mathematics, the move's
exactness (Metropolis-Hastings with the reverse density) and what the
trained network gains over plain proposals."""
import math

import numpy
import pytest
from scipy import stats

import mlp_oracle as mo
from p4gpu import learned, pf, synth
from p4gpu.mcmc import Mcmc

F, H, O = learned.NFEAT, 256, 2 * learned.NCOORD


def _batch(rng, n, F=F, D=O // 2):
    X = rng.normal(size=(n, F)).astype(numpy.float32)
    Y = rng.normal(size=(n, D)).astype(numpy.float32)
    M = (rng.random((n, D)) < 0.7).astype(numpy.float32)
    return X, Y, M


# ------------------------------------------------------------------- CPU
def test_oracle_gradients_finite_difference():
    rng = numpy.random.default_rng(0)
    f, h, o, n = 5, 16, 6, 9
    th = mo.init_like_library(f, h, o, rng)
    X, Y, M = _batch(rng, n, f, o // 2)
    L, g = mo.loss_and_grad(th, X, Y, M, f, h, o)
    eps = 1e-6
    fd = numpy.empty_like(th)
    for i in range(th.size):
        a = th.copy()
        a[i] += eps
        b = th.copy()
        b[i] -= eps
        fd[i] = (mo.loss_and_grad(a, X, Y, M, f, h, o)[0] - mo.loss_and_grad(b, X, Y, M, f, h, o)[0]) / (2 * eps) / n
    assert numpy.abs(fd - g).max() < 1e-7 * max(1.0, numpy.abs(g).max())


def test_oracle_loss_is_gaussian_nll():
    rng = numpy.random.default_rng(1)
    D = 3
    OUT = rng.normal(size=(11, 2 * D))
    Y = rng.normal(size=(11, D))
    M = (rng.random((11, D)) < 0.6).astype(float)
    # -log N(y; mu, e^s) = s + (y - mu)^2 e^{-2s} / 2 + log(2 pi) / 2
    direct = -(M * stats.norm.logpdf(Y, OUT[:, :D], numpy.exp(OUT[:, D:]))).sum()
    assert abs(mo.nll(OUT, Y, M) + 0.5 * math.log(2 * math.pi) * M.sum() - direct) < 1e-10


def test_oracle_adam_first_step():
    rng = numpy.random.default_rng(2)
    th = rng.normal(size=50)
    g = rng.normal(size=50)
    g[3] = 0.0
    new, m, v = mo.adam_step(th, g, numpy.zeros(50), numpy.zeros(50), 1, 1e-3)
    step = th - new
    # first step: m/(1-b1) = g, v/(1-b2) = g^2, so the step is lr g / (|g| + eps)
    nz = g != 0
    assert numpy.allclose(step[nz], 1e-3 * numpy.sign(g[nz]), rtol=1e-6)
    assert step[3] == 0.0


def test_bf16_rounding():
    x = numpy.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 3.0e38, 1e-30],
                    dtype=numpy.float32)
    r = mo.bf16_round(x)
    # 8 significant bits: ties to even
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 4 * 2.0 ** -8 and r[3] == 1.0
    assert r[4] == -2.5
    assert abs(r[5] / 3.0e38 - 1.0) < 2.0 ** -8
    assert abs(r[6] / 1e-30 - 1.0) < 2.0 ** -8


def test_alr_jacobian_is_product_of_components():
    # the learned move's density on the simplex is N(alr pi) / prod_j pi_j:
    # check the determinant of d pi[0:3] / d z numerically
    z = numpy.array([0.3, -1.2, 0.7])
    pi = learned._alrInv(z)
    J = numpy.empty((3, 3))
    e = 1e-6
    for j in range(3):
        a = z.copy()
        a[j] += e
        b = z.copy()
        b[j] -= e
        J[:, j] = (learned._alrInv(a)[:3] - learned._alrInv(b)[:3]) / (2 * e)
    assert abs(numpy.linalg.det(J) / numpy.prod(pi) - 1.0) < 1e-8
    assert numpy.allclose(learned._alr(pi), z)


def test_edge_keys_exclude_taxon_zero():
    full = (1 << 6) - 1
    assert learned.edgeKey(0b000110, full) == 0b000110
    assert learned.edgeKey(0b111001, full) == 0b000110


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_forward_matches_oracle():
    rng = numpy.random.default_rng(3)
    net = pf.Mlp(F, H, O, seed=3)
    th = net.getParams().astype(numpy.float64)
    X = _batch(rng, 301)[0]
    out = net.forward(X)
    ref, _ = mo.forward(th, X, F, H, O, bf16=True)
    assert numpy.abs(out - ref).max() < 3e-4 * max(1.0, numpy.abs(ref).max())


@pytest.mark.gpu
def test_gradients_match_oracle():
    rng = numpy.random.default_rng(4)
    net = pf.Mlp(F, H, O, seed=4)
    th = net.getParams().astype(numpy.float64)
    X, Y, M = _batch(rng, 517)
    L, g = net.gradients(X, Y, M)
    Lo, go = mo.loss_and_grad(th, X, Y, M, F, H, O, bf16=True)
    assert abs(L - Lo) < 1e-4 * abs(Lo)
    # same BF16 operand rounding as the kernels: float32 accumulation order
    # and rare rounding-boundary flips remain
    assert numpy.linalg.norm(g - go) < 2e-3 * numpy.linalg.norm(go)
    off = 0
    for shape in ((H, F), (H,), (H, H), (H,), (O, H), (O,)):
        k = int(numpy.prod(shape))
        assert numpy.linalg.norm(g[off:off + k] - go[off:off + k]) < 5e-3 * numpy.linalg.norm(go[off:off + k]) + 1e-7
        off += k
    # against exact float64 arithmetic: the BF16 operands cost ~2^-8 per term
    L64, g64 = mo.loss_and_grad(th, X, Y, M, F, H, O, bf16=False)
    assert numpy.linalg.norm(g - g64) < 5e-2 * numpy.linalg.norm(g64)


@pytest.mark.gpu
def test_train_step_is_adam_on_the_gradients():
    rng = numpy.random.default_rng(5)
    net = pf.Mlp(F, H, O, seed=5)
    th = net.getParams().astype(numpy.float64)
    X, Y, M = _batch(rng, 200)
    L, g = net.gradients(X, Y, M)
    L2 = net.trainStep(X, Y, M, 1e-3)
    assert L2 == L
    new = net.getParams().astype(numpy.float64)
    want, _, _ = mo.adam_step(th, g.astype(numpy.float64), numpy.zeros_like(th), numpy.zeros_like(th), 1, 1e-3)
    assert numpy.abs(new - want).max() < 2e-6


@pytest.mark.gpu
def test_training_steps_follow_the_oracle():
    """Five Adam steps on the GPU against the same steps on the oracle's
    BF16-rounded gradients."""
    rng = numpy.random.default_rng(7)
    net = pf.Mlp(F, H, O, seed=7)
    th0 = net.getParams().astype(numpy.float64)
    th, m, v = th0.copy(), numpy.zeros_like(th0), numpy.zeros_like(th0)
    for step in range(1, 6):
        X, Y, M = _batch(rng, 333)
        net.trainStep(X, Y, M, 1e-3)
        _, g = mo.loss_and_grad(th, X, Y, M, F, H, O, bf16=True)
        th, m, v = mo.adam_step(th, g, m, v, step, 1e-3)
    got = net.getParams().astype(numpy.float64)
    assert numpy.linalg.norm(got - th) < 0.02 * numpy.linalg.norm(th - th0)


@pytest.mark.gpu
def test_training_fits_a_heteroscedastic_gaussian():
    rng = numpy.random.default_rng(6)
    n = 4096
    X = rng.uniform(-1, 1, size=(n, F)).astype(numpy.float32)
    X[:, 3:] = 0.0
    mu = numpy.sin(2 * X[:, 0]) + 0.5 * X[:, 1]
    sd = 0.05 + 0.2 * (X[:, 2] > 0)
    Y = numpy.zeros((n, 3), numpy.float32)
    Y[:, 0] = mu + sd * rng.standard_normal(n)
    M = numpy.zeros((n, 3), numpy.float32)
    M[:, 0] = 1.0
    net = pf.Mlp(F, H, O, seed=6)
    for ep in range(100):
        for s in range(0, n, 1024):
            net.trainStep(X[s:s + 1024], Y[s:s + 1024], M[s:s + 1024], 2e-3)
    out = net.forward(X).astype(numpy.float64)
    # the oracle trained the same way reaches 0.025 and sds within 6%
    assert numpy.sqrt(((out[:, 0] - mu) ** 2).mean()) < 0.05
    ls = out[:, 3]
    assert abs(numpy.exp(numpy.median(ls[X[:, 2] > 0])) / 0.25 - 1) < 0.25
    assert abs(numpy.exp(numpy.median(ls[X[:, 2] <= 0])) / 0.05 - 1) < 0.25


def _untrained_move(tree, seed):
    """Moves with an untrained network and broad centrings: the proposals
    are state-dependent random functions, a strict test of the reverse
    densities and Jacobians."""
    L = learned.LearnedProposal(tree, seed=seed, weights=dict(ridge=0.3, block=0.4, comp=0.3))
    L.yMean[learned.BRANCH] = [math.log(0.07), 0.0, 0.0]
    L.yMean[learned.TREE] = [math.log(2.9), 0.0, 0.0]
    L.yMean[learned.GAMMA] = [math.log(30.0), 0.0, 0.0]
    L.ySd[:] = 1.0
    L.ySd[learned.TREE] = 0.5
    L.ySd[learned.GAMMA] = 2.0
    L.trained = True
    return L


@pytest.mark.gpu
def test_learned_moves_recover_the_prior():
    tree, _ = synth.make_workload("gtr16", seed=23, nSites=300)
    m = Mcmc(tree, nChains=1, seed=7, sampleInterval=10, verbose=False,
             proposalWeights={'brLen': 0.0, 'nni': 0.2, 'compDir.0': 0.0, 'rMatrixDir.0': 0.0,
                              'gdasrv.0': 0.0})
    L = _untrained_move(tree, seed=7)
    assert L.hasComp and L.hasGamma
    m.addLearnedProposal(L, 1.0)
    m.samplePrior = True
    m.run(40000, verbose=False)
    from p4gpu.func import effectiveSampleSize

    def near(x, want, cap):
        # within 4 Monte Carlo standard errors (Tracer ESS) and a fixed cap
        x = numpy.asarray(x, dtype=float)
        se = x.std() / math.sqrt(max(effectiveSampleSize(x), 1.0))
        return abs(x.mean() - want) < min(4.0 * se + 1e-12, cap), (x.mean(), want, se)
    b = 400
    nBr = len(tree.nodes) - 1
    # Exponential(10) per branch
    ok, info = near(m.samples['treeLen'][b:], nBr / 10.0, 0.1 * nBr / 10.0)
    assert ok, info
    prams = numpy.array(m.samples['prams'][b:])
    # flat Dirichlet(1,1,1,1): mean 1/4, variance 3/80
    for j in range(4):
        ok, info = near(prams[:, j], 0.25, 0.05)
        assert ok, (j, info)
    assert abs(prams[:, :4].var(axis=0).mean() / (3.0 / 80.0) - 1.0) < 0.25
    # gamma shape uniform on [GAMMA_SHAPE_MIN, GAMMA_SHAPE_MAX]
    lo, hi = float(learned.var.GAMMA_SHAPE_MIN), float(learned.var.GAMMA_SHAPE_MAX)
    ok, info = near(prams[:, 4], 0.5 * (lo + hi), 0.15 * 0.5 * (lo + hi))
    assert ok, info
    for kind in ('ridge', 'block', 'comp'):
        acc = L.stats[kind + 'Accepted'] / L.stats[kind]
        assert 0.01 < acc < 0.99, (kind, acc, L.stats)


@pytest.mark.gpu
def test_trained_moves_gain_ess_over_plain():
    """Plain proposals only vs plain + trained learned moves, the same
    number of generations from the same start: the learned run has a larger
    ESS per generation of the tree length and the gamma shape, and the same
    posterior means."""
    tree, _ = synth.make_workload("gtr16", seed=1)
    base = tree.dupe()
    plain = Mcmc(tree, nChains=1, seed=11, sampleInterval=10, verbose=False)
    L = learned.LearnedProposal(tree, seed=11)
    assert L.hasComp and L.hasGamma
    L.attach(plain)
    plain.run(20000, verbose=False)
    L.records = L.records[200:]
    L.train(epochs=20)
    assert L.trainLog[-1] < L.trainLog[0]

    withL = Mcmc(base, nChains=1, seed=12, sampleInterval=10, verbose=False)
    withL.addLearnedProposal(L, 1.0)
    withL.run(20000, verbose=False)
    assert L.stats['ridgeAccepted'] / L.stats['ridge'] > 0.1, L.stats
    b = 200
    ratios = {}
    for key, col in (('treeLen', None), ('gammaShape', 4)):
        if col is None:
            xp, xl = numpy.array(plain.samples[key][b:]), numpy.array(withL.samples[key][b:])
        else:
            xp = numpy.array(plain.samples['prams'][b:])[:, col]
            xl = numpy.array(withL.samples['prams'][b:])[:, col]
        from p4gpu.func import effectiveSampleSize
        eP, eL = effectiveSampleSize(xp), effectiveSampleSize(xl)
        ratios[key] = eL / eP
        se = math.hypot(xp.std() / math.sqrt(eP), xl.std() / math.sqrt(eL))
        assert abs(xl.mean() - xp.mean()) < 5.0 * se, (key, xl.mean(), xp.mean(), se)
    assert min(ratios.values()) > 2.0, (ratios, L.stats)


@pytest.mark.gpu
def test_ridge_move_on_a_multipart_tree_refreshes_every_part():
    """The ridge move scales every branch length.  On a tree with several
    partitions the batched lnL of the proposed tree must equal a full
    recomputation of every part (setPrams(-1) + treeLogLike), so the parts
    the move's own part does not cover are refreshed too."""
    tree, _ = synth.make_workload("multi48", seed=4, nSites=800, precision="fp64")
    m = Mcmc(tree, nChains=1, seed=3, sampleInterval=10, verbose=False)
    L = _untrained_move(tree, seed=3)
    L.kinds = ['ridge']
    L.cum = numpy.array([1.0])
    ch = m.chains[0]
    n = 0
    for _ in range(20):
        L.propose(ch, None)
        if ch.logProposalRatio <= -1.0e299:   # out of bounds: nothing changed
            ch.sync(ch.curTree, ch.propTree)
            continue
        got = m._batch.logLikes()[0]
        t = ch.propTree
        pf.p4_setPrams(t.cTree, -1)
        want = pf.p4_treeLogLike(t.cTree, 0)
        assert abs(got - want) <= 1e-10 * abs(want), (got, want)
        ch.sync(ch.curTree, ch.propTree)
        n += 1
    assert n >= 5
