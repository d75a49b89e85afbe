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
def test_training_fits_a_heteroscedastic_gaussian():
    rng = numpy.random.default_rng(6)
    n = 4096
    X = rng.uniform(-1, 1, size=(n, F)).astype(numpy.float32)
    mu = numpy.sin(2 * X[:, 0]) + 0.5 * X[:, 1]
    sd = 0.05 + 0.2 * (X[:, 2] > 0)
    Y = numpy.zeros((n, 3), numpy.float32)
    Y[:, 0] = mu + sd * rng.standard_normal(n)
    M = numpy.zeros((n, 3), numpy.float32)
    M[:, 0] = 1.0
    net = pf.Mlp(F, H, O, seed=6)
    first = None
    for ep in range(150):
        for s in range(0, n, 1024):
            l = net.trainStep(X[s:s + 1024], Y[s:s + 1024], M[s:s + 1024], 2e-3)
            first = l / 1024 if first is None else first
    out = net.forward(X).astype(numpy.float64)
    # the fitted mean is close to the truth and the fitted sds separate the
    # two noise levels
    assert numpy.sqrt(((out[:, 0] - mu) ** 2).mean()) < 0.06
    ls = out[:, 3]
    assert abs(numpy.exp(numpy.median(ls[X[:, 2] > 0])) / 0.25 - 1) < 0.3
    assert abs(numpy.exp(numpy.median(ls[X[:, 2] <= 0])) / 0.05 - 1) < 0.3


def _untrained_move(tree, seed, blockSize=6, compProb=0.5):
    """A move with an untrained network and a broad centring: its proposal
    is a state-dependent random function, a strict test of the reverse
    density."""
    L = learned.LearnedBrLenCompProposal(tree, seed=seed, blockSize=blockSize, compProb=compProb)
    L.yMean = numpy.array([math.log(0.07), 0.0, 0.0])
    L.ySd = numpy.array([1.0, 1.0, 1.0])
    L.trained = True
    return L


@pytest.mark.gpu
def test_learned_move_recovers_the_prior():
    tree, _ = synth.make_workload("gtr16", seed=23, nSites=300)
    tree.model.parts[0].comps[0].free = 1
    m = Mcmc(tree, nChains=1, seed=7, sampleInterval=5, verbose=False,
             proposalWeights={'brLen': 0.0, 'nni': 0.2, 'compDir.0': 0.0, 'rMatrixDir.0': 0.0,
                              'gdasrv.0': 0.0})
    L = _untrained_move(tree, seed=7)
    m.addLearnedProposal(L, 1.0)
    m.samplePrior = True
    m.run(8000, verbose=False)
    tl = numpy.array(m.samples['treeLen'][100:])
    nBr = len(tree.nodes) - 1
    assert abs(tl.mean() / (nBr / 10.0) - 1.0) < 0.08
    comps = numpy.array([p[:4] for p in m.samples['prams'][100:]])
    # flat Dirichlet(1,1,1,1): mean 1/4, variance 3/80
    assert numpy.abs(comps.mean(axis=0) - 0.25).max() < 0.03
    assert abs(comps.var(axis=0).mean() / (3.0 / 80.0) - 1.0) < 0.2
    acc = L.stats['accepted'] / L.stats['proposed']
    assert 0.02 < acc < 0.98


@pytest.mark.gpu
def test_trained_move_gains_ess_over_plain():
    """Plain proposals only vs plain + trained learned move, same number of
    generations from the same start: the learned run has a larger ESS of the
    tree length per generation and the same posterior mean."""
    tree, _ = synth.make_workload("gtr16", seed=1)
    tree.model.parts[0].comps[0].free = 1
    base = tree.dupe()
    plain = Mcmc(tree, nChains=1, seed=11, sampleInterval=5, verbose=False)
    L = learned.LearnedBrLenCompProposal(tree, seed=11)
    L.attach(plain)
    plain.run(6000, verbose=False)
    L.records = L.records[100:]
    L.train(epochs=15)
    assert L.trainLog[-1] < L.trainLog[0]
    essPlain = plain.ess('treeLen', burnin=100)
    meanPlain = numpy.mean(plain.samples['treeLen'][100:])

    t2 = base.dupe()
    t2.model.parts[0].comps[0].free = 1
    withL = Mcmc(t2, nChains=1, seed=12, sampleInterval=5, verbose=False)
    withL.addLearnedProposal(L, 0.5)
    withL.run(6000, verbose=False)
    essL = withL.ess('treeLen', burnin=100)
    tlL = numpy.array(withL.samples['treeLen'][100:])
    acc = L.stats['accepted'] / max(1, L.stats['proposed'])
    assert acc > 0.1, acc
    assert essL > 2.0 * essPlain, (essL, essPlain)
    se = tlL.std() / math.sqrt(max(essL, 1.0))
    assert abs(tlL.mean() - meanPlain) < 5.0 * se + 0.01 * meanPlain
