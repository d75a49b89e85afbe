"""Pins the CPU oracle (oracle/p4oracle.c) before it is trusted as the
checker: golden vectors from the compiled reference subset, the reference's
own known answers, closed forms, library routines, brute force and
invariants.  CPU only."""
import itertools

import numpy
import pytest
import scipy.linalg

import oracle
from helpers import LNL_CASES, load, oracle_flat


# ------------------------------------------------------------------ gamma
def test_gamma_known_answers():
    # The reference's own stated values (Pf/p4_treeSim.c:185-197).
    kat = {(0.5, 4): [0.033388, 0.251916, 0.820268, 2.894428],
           (1.5, 4): [0.225323, 0.588556, 1.050422, 2.135699],
           (0.5, 2): [0.142652, 1.857348]}
    for (a, K), want in kat.items():
        got = oracle.discrete_gamma(a, K)
        assert numpy.allclose(got, want, atol=6e-7), (a, K, got)


def test_gamma_matches_reference_bitwise():
    g = load("gamma_rates")
    for i, a in enumerate(g["alphas"]):
        for j, K in enumerate(g["Ks"]):
            got = oracle.discrete_gamma(float(a), int(K))
            want = g["rates"][i, j, :K]
            assert numpy.array_equal(got, want), (a, K, got - want)


def test_gamma_mean_rate_is_one():
    for a in (0.1, 0.37, 1.0, 3.3, 250.0):
        for K in (2, 4, 7):
            r = oracle.discrete_gamma(a, K)
            assert abs(r.mean() - 1.0) < 1e-14
            assert numpy.all(numpy.diff(r) > 0)


# -------------------------------------------------------------- Q and P
def _rand_model(S, seed):
    rng = numpy.random.default_rng(seed)
    pi = rng.dirichlet([2.0] * S)
    pi[-1] = 1.0 - pi[:-1].sum()
    v = rng.lognormal(0, 1, S * (S - 1) // 2)
    R = numpy.zeros((S, S))
    R[numpy.triu_indices(S, 1)] = v
    R = R + R.T
    return R, pi


@pytest.mark.parametrize("S", [4, 20])
def test_bigq_invariants(S):
    R, pi = _rand_model(S, S)
    Q = oracle.bigq(R, pi)
    assert numpy.allclose(Q.sum(1), 0.0, atol=1e-14)
    F = pi[:, None] * Q
    assert numpy.allclose(F, F.T, atol=1e-15)
    assert abs(-(pi * numpy.diag(Q)).sum() - 1.0) < 1e-14


def test_bigq_and_pmat_match_reference():
    g = load("pmat")
    ts = g["ts"]
    for k in ("dna0", "dna1", "aa0", "aa1"):
        pi, R, Q, P = g[k + "_pi"], g[k + "_R"], g[k + "_Q"], g[k + "_P"]
        assert numpy.allclose(oracle.bigq(R, pi), Q, rtol=1e-15, atol=0)
        mine = oracle.pmat(R, pi, ts)
        # the eigen routes differ (Jacobi vs EISPACK); P(t) is unique
        assert numpy.abs(mine - P).max() < 2e-14, (k, numpy.abs(mine - P).max())


def test_pmat_jc_closed_form():
    R = numpy.ones((4, 4))
    pi = numpy.full(4, 0.25)
    for t in (1e-6, 0.05, 0.3, 2.0):
        P = oracle.pmat(R, pi, [t])[0]
        e = numpy.exp(-4.0 * t / 3.0)
        want = numpy.full((4, 4), 0.25 - 0.25 * e)
        numpy.fill_diagonal(want, 0.25 + 0.75 * e)
        assert numpy.abs(P - want).max() < 1e-15


def test_pmat_f81_closed_form():
    R = numpy.ones((4, 4))
    pi = numpy.array([0.1, 0.2, 0.3, 0.4])
    beta = 1.0 / (1.0 - (pi * pi).sum())
    for t in (0.01, 0.4, 1.7):
        P = oracle.pmat(R, pi, [t])[0]
        e = numpy.exp(-beta * t)
        want = pi[None, :] * (1 - e) + numpy.eye(4) * e
        assert numpy.abs(P - want).max() < 1e-15


@pytest.mark.parametrize("S", [4, 20])
def test_pmat_vs_expm_and_semigroup(S):
    R, pi = _rand_model(S, 100 + S)
    Q = oracle.bigq(R, pi)
    for t in (1e-3, 0.2, 1.5):
        P = oracle.pmat(R, pi, [t])[0]
        assert numpy.abs(P - scipy.linalg.expm(Q * t)).max() < 1e-13
        assert numpy.allclose(P.sum(1), 1.0, atol=1e-13)
    a, b = oracle.pmat(R, pi, [0.3, 0.45])
    ab = oracle.pmat(R, pi, [0.75])[0]
    assert numpy.abs(a @ b - ab).max() < 1e-13


# ----------------------------------------------------------- likelihood
@pytest.mark.parametrize("case", LNL_CASES)
def test_loglike_matches_reference(case):
    fx = load(case)
    f = oracle_flat(fx)
    v, pl = oracle.part_loglike(f, site_likes=True)
    assert abs(v - fx["lnL"]) <= 1e-12 * abs(fx["lnL"]), (v, fx["lnL"])
    # per-pattern likelihoods as small as 1e-110 inherit the ~1e-15 absolute
    # difference of the two eigen routes through small P entries
    assert numpy.allclose(pl, fx["patternLikes"], rtol=2e-11, atol=0)


def test_all_gap_site_like_is_one():
    fx = load("edge_dna_gaps_pinv")
    f = oracle_flat(fx)
    _, pl = oracle.part_loglike(f, site_likes=True)
    allgap = numpy.all(f.patterns == -1, axis=0)
    assert allgap.any()
    # pInvar > 0: (1 - pinv) * 1 + pinv * sum(pi) = 1
    assert numpy.allclose(pl[allgap], 1.0, atol=1e-15)


def _brute_pattern_likes(f, P):
    """Sum over all internal-state assignments (tiny trees only)."""
    N, K, S = len(f.parent), f.nCat, f.dim
    internal = [i for i in range(N) if not f.isLeaf[i] or i == f.root]
    pi = f.comps[f.compNums[f.root]]
    eq = f.equates
    out = numpy.zeros(f.patterns.shape[1])
    for p in range(f.patterns.shape[1]):
        tot = 0.0
        for k in range(K):
            for assign in itertools.product(range(S), repeat=len(internal)):
                st = dict(zip(internal, assign))
                w = pi[st[f.root]]
                if f.isLeaf[f.root]:
                    c = f.patterns[f.seqNum[f.root], p]
                    allowed = _allowed(c, S, eq)
                    if st[f.root] not in allowed:
                        continue
                for n in range(N):
                    if n == f.root:
                        continue
                    par = st[f.parent[n]]
                    if f.isLeaf[n]:
                        c = f.patterns[f.seqNum[n], p]
                        w *= sum(P[n, k, par, j] for j in _allowed(c, S, eq))
                    else:
                        w *= P[n, k, par, st[n]]
                tot += w
        like = tot / K
        if f.pInvar > 0:
            like = like * (1 - f.pInvar) + f.pInvar * sum(
                pi[s] for s in range(S) if f.invarArray[s, p])
        out[p] = like
    return out


def _allowed(c, S, eq):
    if c >= 0:
        return [c]
    if c in (-1, -2, -3):
        return list(range(S))
    return [s for s in range(S) if eq[c + 64, s]]


def test_brute_force_enumeration():
    fx = load("multifurcation_root4")
    f = oracle_flat(fx)
    P = oracle.oracle_pmats(f)
    _, pl = oracle.part_loglike(f, site_likes=True)
    sel = numpy.arange(0, f.patterns.shape[1], 13)
    brute = _brute_pattern_likes(_sub(f, sel), P)
    assert numpy.allclose(pl[sel], brute, rtol=1e-13, atol=0)


def test_brute_force_leaf_root_with_pinvar():
    fx = load("leaf_root")
    f = oracle_flat(fx)
    P = oracle.oracle_pmats(f)
    _, pl = oracle.part_loglike(f, site_likes=True)
    sel = numpy.arange(0, f.patterns.shape[1], 17)
    brute = _brute_pattern_likes(_sub(f, sel), P)
    assert numpy.allclose(pl[sel], brute, rtol=1e-13, atol=0)


def test_brute_force_catches_transposed_p():
    """A planted mistake (P transposed) must fail the brute-force check."""
    fx = load("multifurcation_root4")
    f = oracle_flat(fx)
    P = oracle.oracle_pmats(f)
    PT = numpy.ascontiguousarray(numpy.swapaxes(P, 2, 3))
    import ctypes
    cl = numpy.zeros((len(f.parent), f.nCat, f.dim, f.patterns.shape[1]))
    pl = numpy.zeros(f.patterns.shape[1])
    oracle.lib().or_part_loglike(ctypes.byref(f.tr), ctypes.byref(f.mp), ctypes.byref(f.dp),
                                 PT.ctypes.data, cl.ctypes.data, pl.ctypes.data)
    sel = numpy.arange(0, f.patterns.shape[1], 13)
    brute = _brute_pattern_likes(_sub(f, sel), P)
    assert not numpy.allclose(pl[sel], brute, rtol=1e-6)


def _sub(f, sel):
    import copy
    g = copy.copy(f)
    g.patterns = numpy.ascontiguousarray(f.patterns[:, sel])
    g.counts = numpy.ascontiguousarray(f.counts[sel])
    g.invarVec = numpy.ascontiguousarray(f.invarVec[sel])
    g.invarArray = numpy.ascontiguousarray(f.invarArray[:, sel])
    return g


def test_two_taxon_pulley_closed_form():
    """Root with two leaves, reversible model: site like =
    pi_x P_xy(t1 + t2) (the pulley principle)."""
    R, pi = _rand_model(4, 7)
    t1, t2 = 0.13, 0.41
    pats = numpy.array([[0, 1, 2, 3, 0, 2], [0, 2, 2, 1, 3, 2]], numpy.int32)
    f = oracle.Flat(parent=[-1, 0, 0], leftChild=[1, -1, -1], sibling=[-1, 2, -1],
                    isLeaf=[0, 1, 1], seqNum=[-1, 0, 1], brLen=[0, t1, t2],
                    postOrder=[1, 2, 0], root=0, compNums=[0, 0, 0], rMatrixNums=[0, 0, 0],
                    gdasrvNums=[0, 0, 0], dim=4, nCat=1, comps=pi[None, :], bigR=R[None],
                    rates=numpy.ones((1, 1)), alphas=numpy.zeros(1), pInvar=0.0, relRate=1.0,
                    patterns=pats, counts=numpy.ones(6, numpy.int32),
                    equates=numpy.zeros((0, 4), numpy.int32),
                    invarVec=numpy.zeros(6, numpy.int32), invarArray=numpy.zeros((4, 6), numpy.int32))
    _, pl = oracle.part_loglike(f, site_likes=True)
    P = oracle.pmat(R, pi, [t1 + t2])[0]
    want = [pi[a] * P[a, b] for a, b in pats.T]
    assert numpy.allclose(pl, want, rtol=1e-13)


def test_reroot_invariance_homogeneous():
    """Homogeneous reversible model: lnL does not depend on the root."""
    fx = load("gtr16_500")
    f = oracle_flat(fx)
    v0 = oracle.part_loglike(f)
    import copy
    # reroot at each internal child of the root in turn
    N = len(f.parent)
    for newRoot in [i for i in range(N) if not f.isLeaf[i] and f.parent[i] == f.root][:2]:
        g = _reroot(f, newRoot)
        assert abs(oracle.part_loglike(g) - v0) < 1e-9 * abs(v0)


def _reroot(f, r):
    import copy
    N = len(f.parent)
    adj = {i: set() for i in range(N)}
    bl = {}
    for i in range(N):
        if f.parent[i] >= 0:
            adj[i].add(f.parent[i])
            adj[f.parent[i]].add(i)
            bl[frozenset((i, f.parent[i]))] = f.brLen[i]
    parent = [-1] * N
    order = [r]
    seen = {r}
    for v in order:
        for w in sorted(adj[v]):
            if w not in seen:
                seen.add(w)
                parent[w] = v
                order.append(w)
    kids = {i: [w for w in sorted(adj[i]) if parent[w] == i] for i in range(N)}
    lc = [kids[i][0] if kids[i] else -1 for i in range(N)]
    sib = [-1] * N
    for i in range(N):
        for a, b in zip(kids[i], kids[i][1:]):
            sib[a] = b
    brl = [bl[frozenset((i, parent[i]))] if parent[i] >= 0 else 0.0 for i in range(N)]
    post = order[::-1]
    g = copy.copy(f)
    g.parent, g.leftChild, g.sibling, g.brLen = parent, lc, sib, brl
    g.postOrder, g.root = post, r
    g._build()
    return g


# ---------------------------------------------------------- data prep
def test_pattern_compression_round_trip():
    fx = load("edge_dna_gaps_pinv")
    from p4gpu.data import DNA_EQUATES, DNA_SYMBOLS
    codes = oracle.code_sequences([str(s) for s in fx["sequences"]], DNA_SYMBOLS, DNA_EQUATES)
    pats, cnts, spi = oracle.compress(codes)
    assert cnts.sum() == codes.shape[1]
    assert numpy.array_equal(pats[:, spi], codes)
    first = [int(numpy.argmax(spi == k)) for k in range(pats.shape[1])]
    assert first == sorted(first)


def test_invariant_sites_known_answers():
    # Pf/part.c:736-760 worked examples: 'hra' is invariant (a), 'yaa' is not,
    # 'raa' is invariant (a), all gaps allow every state.
    from p4gpu.data import DNA_EQUATES, DNA_SYMBOLS
    cols = ["hra", "yaa", "raa", "---", "ccc", "rrg"]
    seqs = ["".join(c[i] for c in cols) for i in range(3)]
    codes = oracle.code_sequences(seqs, DNA_SYMBOLS, DNA_EQUATES)
    eq = oracle.equates_table(DNA_SYMBOLS, DNA_EQUATES)
    iv, ia = oracle.invar_sites(codes, 4, eq)
    assert list(iv) == [1, 0, 1, 4, 1, 1]
    assert list(ia[:, 0]) == [1, 0, 0, 0] and list(ia[:, 4]) == [0, 1, 0, 0]
    assert list(ia[:, 5]) == [0, 0, 1, 0]
