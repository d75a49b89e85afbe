"""ctypes wrapper of the CPU oracle (liboracle.so) and, where it was built,
of the compiled reference subset (_ref/libp4ref.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.

The oracle takes a flat description of one partition: topology arrays,
per-node model numbers, model values and coded patterns.  flat_from_tree()
reads these from a p4gpu Tree (its Python objects, not the GPU library).
"""
import ctypes
import os
import subprocess

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libp4ref.so")

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


class _Part(ctypes.Structure):
    _fields_ = [("nTax", I), ("dim", I), ("nPatterns", I), ("patterns", P),
                ("patternCounts", P), ("nEquates", I), ("equates", P), ("invarVec", P),
                ("invarArray", P)]


class _ModelPart(ctypes.Structure):
    _fields_ = [("dim", I), ("nCat", I), ("nComps", I), ("comps", P), ("nRMatrices", I),
                ("bigR", P), ("nGdasrvs", I), ("gdasrvRates", P), ("pInvar", D),
                ("relRate", D)]


class _Tree(ctypes.Structure):
    _fields_ = [("nNodes", I), ("root", I), ("parent", P), ("leftChild", P), ("sibling", P),
                ("isLeaf", P), ("seqNum", P), ("brLen", P), ("postOrder", P),
                ("compNums", P), ("rMatrixNums", P), ("gdasrvNums", P)]


_lib = None
_ref = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _lib = ctypes.CDLL(ORACLE_SO)
        _lib.or_discrete_gamma.argtypes = [P, P, D, D, I, I]
        _lib.or_discrete_gamma.restype = I
        for f in ("or_point_chi2", "or_incomplete_gamma"):
            getattr(_lib, f).restype = D
        _lib.or_point_chi2.argtypes = [D, D]
        _lib.or_incomplete_gamma.argtypes = [D, D, D]
        _lib.or_ln_gamma.argtypes = [D]
        _lib.or_ln_gamma.restype = D
        _lib.or_point_normal.argtypes = [D]
        _lib.or_point_normal.restype = D
        _lib.or_bigq.argtypes = [I, P, P, P]
        _lib.or_bigq.restype = I
        _lib.or_eigen_reversible.argtypes = [I, P, P, P, P, P]
        _lib.or_eigen_reversible.restype = I
        _lib.or_pmat.argtypes = [I, P, P, P, D, P]
        _lib.or_pmat.restype = None
        _lib.or_set_prams_part.argtypes = [P, P, P]
        _lib.or_set_prams_part.restype = I
        _lib.or_part_loglike.argtypes = [P, P, P, P, P, P]
        _lib.or_part_loglike.restype = D
        _lib.or_tree_loglike_part.argtypes = [P, P, P, P]
        _lib.or_tree_loglike_part.restype = D
        _lib.or_make_patterns.argtypes = [I, I, P, P, P, P]
        _lib.or_make_patterns.restype = I
        _lib.or_invar_sites.argtypes = [I, I, I, P, I, P, P, P]
        _lib.or_invar_sites.restype = None
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            return None
        _ref = ctypes.CDLL(REF_SO)
        _ref.ref_discrete_gamma.argtypes = [D, I, P, P]
        _ref.ref_discrete_gamma.restype = I
        _ref.ref_point_chi2.argtypes = [D, D]
        _ref.ref_point_chi2.restype = D
        _ref.ref_incomplete_gamma.argtypes = [D, D, D]
        _ref.ref_incomplete_gamma.restype = D
        _ref.ref_ln_gamma.argtypes = [D]
        _ref.ref_ln_gamma.restype = D
        _ref.ref_point_normal.argtypes = [D]
        _ref.ref_point_normal.restype = D
        _ref.ref_bigq_pmat.argtypes = [I, P, P, I, P, P, P]
        _ref.ref_bigq_pmat.restype = I
        _ref.ref_part_loglike.argtypes = [P, P, P, P, P, P, P, P]
        _ref.ref_part_loglike.restype = D
        _ref.ref_part_derivs.argtypes = [P] * 11
        _ref.ref_part_derivs.restype = I
    return _ref


def _p(a):
    return a.ctypes.data if a is not None else None


# ------------------------------------------------------------- scalar API
def discrete_gamma(alpha, K, use_ref=False):
    rates = numpy.zeros(K)
    freqs = numpy.zeros(K)
    if use_ref:
        ref().ref_discrete_gamma(alpha, K, _p(rates), _p(freqs))
    else:
        lib().or_discrete_gamma(_p(freqs), _p(rates), alpha, alpha, K, 0)
    return rates


def bigq(R, pi):
    R = numpy.ascontiguousarray(R, dtype=numpy.float64)
    pi = numpy.ascontiguousarray(pi, dtype=numpy.float64)
    Q = numpy.zeros_like(R)
    if lib().or_bigq(len(pi), _p(R), _p(pi), _p(Q)):
        raise ValueError("comp does not sum to 1")
    return Q


def pmat(R, pi, ts):
    """P(t) for each t: [len(ts)][dim][dim] via the oracle's eigen route."""
    Q = bigq(R, pi)
    pi = numpy.ascontiguousarray(pi, dtype=numpy.float64)
    n = len(pi)
    V = numpy.zeros((n, n))
    Vi = numpy.zeros((n, n))
    lam = numpy.zeros(n)
    lib().or_eigen_reversible(n, _p(Q), _p(pi), _p(V), _p(Vi), _p(lam))
    out = numpy.zeros((len(ts), n, n))
    for i, t in enumerate(ts):
        buf = numpy.zeros((n, n))
        lib().or_pmat(n, _p(V), _p(Vi), _p(lam), float(t), _p(buf))
        out[i] = buf
    return out


def ref_pmat(R, pi, ts):
    R = numpy.ascontiguousarray(R, dtype=numpy.float64)
    pi = numpy.ascontiguousarray(pi, dtype=numpy.float64)
    ts = numpy.ascontiguousarray(ts, dtype=numpy.float64)
    n = len(pi)
    Q = numpy.zeros((n, n))
    out = numpy.zeros((len(ts), n, n))
    if ref().ref_bigq_pmat(n, _p(R), _p(pi), len(ts), _p(ts), _p(Q), _p(out)):
        raise ValueError("reference eigensystem failed")
    return Q, out


# ------------------------------------------------------- flat descriptions
class Flat(object):
    """Arrays describing one partition of a tree, kept alive together."""

    def __init__(self, **kw):
        self.kwargs = dict(kw)   # picklable: Flat(**f.kwargs) rebuilds it in another process
        self.__dict__.update(kw)
        self._build()

    def _build(self):
        i32 = lambda a: numpy.ascontiguousarray(a, dtype=numpy.int32)
        f64 = lambda a: numpy.ascontiguousarray(a, dtype=numpy.float64)
        for k in ("parent", "leftChild", "sibling", "isLeaf", "seqNum", "postOrder",
                  "compNums", "rMatrixNums", "gdasrvNums", "patterns", "counts", "equates",
                  "invarVec", "invarArray"):
            setattr(self, k, i32(getattr(self, k)))
        for k in ("brLen", "comps", "bigR", "rates", "alphas"):
            setattr(self, k, f64(getattr(self, k)))
        self.tr = _Tree(len(self.parent), int(self.root), _p(self.parent), _p(self.leftChild),
                        _p(self.sibling), _p(self.isLeaf), _p(self.seqNum), _p(self.brLen),
                        _p(self.postOrder), _p(self.compNums), _p(self.rMatrixNums),
                        _p(self.gdasrvNums))
        self.mp = _ModelPart(self.dim, self.nCat, self.comps.shape[0], _p(self.comps),
                             self.bigR.shape[0], _p(self.bigR), self.rates.shape[0] if self.nCat > 1 else 0,
                             _p(self.rates), float(self.pInvar), float(self.relRate))
        self.dp = _Part(self.patterns.shape[0], self.dim, self.patterns.shape[1],
                        _p(self.patterns), _p(self.counts), self.equates.shape[0],
                        _p(self.equates) if self.equates.size else None, _p(self.invarVec),
                        _p(self.invarArray))


def code_sequences(seqs, symbols, equates):
    """Code sequence strings as Pf/part.c:127-275 does: symbols -> index,
    '-' -> -1, '?' -> -2, equate i -> -64 + i (equates in dict order)."""
    eq = list(equates.keys())
    lut = {}
    for i, e in enumerate(eq):
        lut[e] = -64 + i
    lut['-'] = -1
    lut['?'] = -2
    for i, s in enumerate(symbols):
        lut[s] = i
    return numpy.array([[lut[c] for c in s] for s in seqs], dtype=numpy.int32)


def equates_table(symbols, equates):
    return numpy.array([[1 if s in equates[e] else 0 for s in symbols] for e in equates],
                       dtype=numpy.int32).reshape(len(equates), len(symbols))


def compress(codes):
    """Pattern compression with the oracle's reference-order algorithm."""
    codes = numpy.ascontiguousarray(codes, dtype=numpy.int32)
    nTax, nChar = codes.shape
    pats = numpy.zeros((nTax, nChar), numpy.int32)
    cnt = numpy.zeros(nChar, numpy.int32)
    spi = numpy.zeros(nChar, numpy.int32)
    nP = lib().or_make_patterns(nTax, nChar, _p(codes), _p(pats), _p(cnt), _p(spi))
    return numpy.ascontiguousarray(pats[:, :nP]), cnt[:nP].copy(), spi


def invar_sites(patterns, dim, equates):
    patterns = numpy.ascontiguousarray(patterns, dtype=numpy.int32)
    nTax, nP = patterns.shape
    equates = numpy.ascontiguousarray(equates, dtype=numpy.int32)
    iv = numpy.zeros(nP, numpy.int32)
    ia = numpy.zeros((dim, nP), numpy.int32)
    lib().or_invar_sites(nTax, dim, nP, _p(patterns), equates.shape[0],
                         _p(equates) if equates.size else None, _p(iv), _p(ia))
    return iv, ia


def flat_from_tree(tree, pNum=0, patterns=None, counts=None, own_compression=False):
    """Flat description of part pNum of a p4gpu Tree.  Patterns come from
    (a) the arguments, (b) the oracle's own compression of the alignment's
    sequences (own_compression=True), or (c) the Part's patterns."""
    from p4gpu.synth import bigR_from_vals  # R-vector layout of p4/model.py:185-190
    tree.modelSanityCheck(resetEmpiricalComps=False)  # fills node model numbers
    mp = tree.model.parts[pNum]
    dp = tree.data.parts[pNum]
    dim = mp.dim
    eqt = equates_table(dp.symbols, dp.equates)
    if patterns is None:
        if own_compression and tree.data.alignments:
            aln = tree.data.alignments[pNum]
            patterns, counts, _ = compress(code_sequences(aln.sequences, aln.symbols, aln.equates))
        else:
            patterns, counts, _, _, _ = dp.patterns()
    iv, ia = invar_sites(patterns, dim, eqt)
    N = len(tree.nodes)
    par = [n.parent.nodeNum if n.parent else -1 for n in tree.nodes]
    lc = [n.leftChild.nodeNum if n.leftChild else -1 for n in tree.nodes]
    sib = [n.sibling.nodeNum if n.sibling else -1 for n in tree.nodes]
    if not tree.preAndPostOrderAreValid:
        tree.setPreAndPostOrder()
    comps = numpy.array([c.val for c in mp.comps])
    Rs = []
    for r in mp.rMatrices:
        if r.spec == '2p':
            al, k = 1.0 / 3.0, r.val[0]
            be = al * k
            Rs.append([[0, al, be, al], [al, 0, al, be], [be, al, 0, al], [al, be, al, 0]])
        else:
            Rs.append(bigR_from_vals(r.val, dim))
    alphas = numpy.array([g.val[0] for g in mp.gdasrvs]) if mp.gdasrvs else numpy.zeros(1)
    K = mp.nGammaCat
    rates = (numpy.array([discrete_gamma(a, K) for a in alphas]) if mp.gdasrvs
             else numpy.ones((1, 1)))
    return Flat(parent=par, leftChild=lc, sibling=sib,
                isLeaf=[n.isLeaf for n in tree.nodes], seqNum=[n.seqNum for n in tree.nodes],
                brLen=[n.br.len if n is not tree.root else 0.0 for n in tree.nodes],
                postOrder=tree.postOrder, root=tree.root.nodeNum,
                compNums=[n.parts[pNum].compNum for n in tree.nodes],
                rMatrixNums=[max(n.br.parts[pNum].rMatrixNum, 0) for n in tree.nodes],
                gdasrvNums=[max(n.br.parts[pNum].gdasrvNum, 0) for n in tree.nodes],
                dim=dim, nCat=K, comps=comps, bigR=numpy.array(Rs), rates=rates,
                alphas=alphas, pInvar=mp.pInvar.val, relRate=mp.relRate,
                patterns=patterns, counts=counts, equates=eqt, invarVec=iv, invarArray=ia)


def part_loglike(f, site_likes=False):
    pl = numpy.zeros(f.patterns.shape[1]) if site_likes else None
    v = lib().or_tree_loglike_part(ctypes.byref(f.tr), ctypes.byref(f.mp), ctypes.byref(f.dp),
                                   _p(pl))
    if v <= -2.0e99:
        raise ValueError("oracle rejected the input (model numbers, comps or allocation)")
    return (v, pl) if site_likes else v


def oracle_pmats(f):
    K, dim, N = f.nCat, f.dim, len(f.parent)
    Pm = numpy.zeros((N, K, dim, dim))
    if lib().or_set_prams_part(ctypes.byref(f.tr), ctypes.byref(f.mp), _p(Pm)):
        raise ValueError("oracle setPrams failed")
    return Pm


def oracle_cl(f):
    """(lnL, P [N][K][dim][dim], cl [N][K][dim][nPat]) from the oracle."""
    K, dim, N, nP = f.nCat, f.dim, len(f.parent), f.patterns.shape[1]
    Pm = oracle_pmats(f)
    cl = numpy.zeros((N, K, dim, nP))
    v = lib().or_part_loglike(ctypes.byref(f.tr), ctypes.byref(f.mp), ctypes.byref(f.dp),
                              _p(Pm), _p(cl), None)
    return v, Pm, cl


def ref_part_derivs(f, want_cl2=()):
    """Branch-length derivatives of the partition's lnL through the compiled
    reference subset (its P', P'' and up-pass conditional likelihoods, and
    the restated per-site sums of Pf/p4_treeNewt.c:257-500).  Returns
    (lnLEdge [N], d1 [N], d2 [N], P1 [N][K][S][S], P2, cl2 {node: [K][S][nP]})."""
    r = ref()
    if r is None:
        raise RuntimeError("reference subset not built")
    K, dim, N, nP = f.nCat, f.dim, len(f.parent), f.patterns.shape[1]
    want = numpy.zeros(N, numpy.int32)
    for n in want_cl2:
        want[n] = 1
    cl2 = numpy.zeros((N, K, dim, nP)) if len(want_cl2) else None
    P1 = numpy.zeros((N, K, dim, dim))
    P2 = numpy.zeros((N, K, dim, dim))
    lnl, d1, d2 = numpy.zeros(N), numpy.zeros(N), numpy.zeros(N)
    r.ref_part_derivs(ctypes.byref(f.tr), ctypes.byref(f.mp), ctypes.byref(f.dp), _p(f.alphas), _p(want),
                      _p(cl2), _p(P1), _p(P2), _p(lnl), _p(d1), _p(d2))
    return lnl, d1, d2, P1, P2, ({n: cl2[n] for n in want_cl2} if cl2 is not None else {})


def ref_part_loglike(f, want_cl=False):
    """Same partition through the compiled reference subset.  Returns
    (lnL, patternLikes, P, cl or None, rates)."""
    r = ref()
    if r is None:
        raise RuntimeError("reference subset not built")
    K, dim, N, nP = f.nCat, f.dim, len(f.parent), f.patterns.shape[1]
    Pm = numpy.zeros((N, K, dim, dim))
    cl = numpy.zeros((N, K, dim, nP)) if want_cl else None
    rates = numpy.zeros((max(len(f.alphas), 1), K))
    pl = numpy.zeros(nP)
    v = r.ref_part_loglike(ctypes.byref(f.tr), ctypes.byref(f.mp), ctypes.byref(f.dp),
                           _p(f.alphas), _p(Pm), _p(cl), _p(rates), _p(pl))
    return v, pl, Pm, cl, rates


# ------------------------------------------------- root-side consumers
def anc_table(f, cl_root):
    """The picker table of Pf/p4_treeSim.c:665-692 for every pattern:
    entries (cat, state), category-major, each comp[s] * cl[cat][s][p] *
    (1 - pInvar)/nCat with comp the root's composition, then one entry per
    state, comp[s] * pInvar where the pattern is invariant in that state
    (0 elsewhere; the reference skips those, which leaves the running sums
    unchanged).  cl_root: [K][S][nPat].  Returns [nPat][K*S + S] weights."""
    K, S = f.nCat, f.dim
    comp = f.comps[f.compNums[f.root]]
    fac = (1.0 - f.pInvar) / K
    nP = cl_root.shape[2]
    W = numpy.zeros((nP, K * S + S))
    for k in range(K):
        for s in range(S):
            W[:, k * S + s] = (comp[s] * cl_root[k, s, :]) * fac
    if f.pInvar != 0.0:
        for s in range(S):
            W[:, K * S + s] = numpy.where(f.invarArray[s] != 0, comp[s] * f.pInvar, 0.0)
    return W


def anc_draw(f, cl_root, site_pat, u):
    """Pf/p4_treeSim.c:591-845 for each site: theRanDouble = u * sLike, the
    first table entry whose running sum exceeds it.  Returns [nSites][4]
    (chStNum, catNum, isInvar, invarChNum), and the running sums
    [nSites][K*S + S] for the checks on ties."""
    K, S = f.nCat, f.dim
    W = anc_table(f, cl_root)[numpy.asarray(site_pat)]
    cum = numpy.cumsum(W, axis=1)      # left to right: the reference's running sum
    target = numpy.asarray(u) * cum[:, -1]
    hit = target[:, None] < cum
    idx = numpy.where(hit.any(axis=1), hit.argmax(axis=1), -1)
    out = numpy.full((len(idx), 4), -1, numpy.int32)
    out[:, 2] = 0
    var = (idx >= 0) & (idx < K * S)
    out[var, 0] = idx[var] % S
    out[var, 1] = idx[var] // S
    inv = idx >= K * S
    out[inv, 2] = 1
    out[inv, 3] = idx[inv] - K * S
    return out, cum


def philox4x64(counter, key):
    """Philox4x64-10 of one counter (4 words) and key (2 words), through
    numpy's Philox bit generator (which increments its counter before each
    block of four outputs)."""
    c = numpy.array(counter, dtype=numpy.uint64) - numpy.array([1, 0, 0, 0], dtype=numpy.uint64)
    if int(counter[0]) == 0:   # borrow through the words, as the generator carries
        c = numpy.array(counter, dtype=numpy.uint64)
        for i in range(4):
            if c[i] != 0:
                c[i] -= numpy.uint64(1)
                break
            c[i] = numpy.uint64(0xFFFFFFFFFFFFFFFF)
    bg = numpy.random.Philox(key=numpy.array(key, dtype=numpy.uint64), counter=c)
    return bg.random_raw(4)


def anc_uniforms(seed, sites):
    """The uniforms the GPU draws take when none are given: the first word of
    Philox4x64-10 at counter (site, 'anc', 0, 0), key (seed, 'root'), top
    53 bits."""
    return numpy.array([float(int(philox4x64([s, 0x616e63, 0, 0], [seed, 0x726f6f74])[0]) >> 11) * 2.0 ** -53
                        for s in sites])


def cpo_sums(site_like_samples):
    """p4/mcmc.py:3286-3297: per site, the sum over samples of 1/siteLike."""
    acc = numpy.zeros(len(site_like_samples[0]))
    for sl in site_like_samples:
        for i, v in enumerate(sl):
            acc[i] += 1.0 / v
    return acc


def lpml(sums, nSamples):
    """p4/mcmc.py:3742-3746."""
    a = numpy.array(sums) / nSamples
    return float((-numpy.log(a)).sum())


# ------------------------------------------------------------ simulation
_M32 = numpy.uint64(0xFFFFFFFF)


def _mulhilo64(a, b):
    """(high, low) 64-bit words of the 128-bit products of uint64 arrays."""
    a0, a1 = a & _M32, a >> numpy.uint64(32)
    b0, b1 = b & _M32, b >> numpy.uint64(32)
    p00, p01, p10, p11 = a0 * b0, a0 * b1, a1 * b0, a1 * b1
    mid = (p00 >> numpy.uint64(32)) + (p01 & _M32) + (p10 & _M32)
    hi = p11 + (p01 >> numpy.uint64(32)) + (p10 >> numpy.uint64(32)) + (mid >> numpy.uint64(32))
    with numpy.errstate(over='ignore'):
        lo = a * b
    return hi, lo


def philox4x64_vec(c0, c1, c2, c3, k0, k1):
    """Philox4x64-10 (Salmon et al., SC 2011) over arrays of counters; the
    scalar form is checked against numpy's Philox generator."""
    M0, M1 = numpy.uint64(0xD2E7470EE14C6C93), numpy.uint64(0xCA5A826395121157)
    W0, W1 = numpy.uint64(0x9E3779B97F4A7C15), numpy.uint64(0xBB67AE8584CAA73B)
    c = [numpy.asarray(x, dtype=numpy.uint64) for x in (c0, c1, c2, c3)]
    n = max(x.size for x in c)
    c = [numpy.broadcast_to(x, (n,)).copy() for x in c]
    k0, k1 = numpy.uint64(k0), numpy.uint64(k1)
    with numpy.errstate(over='ignore'):
        for _ in range(10):
            hi0, lo0 = _mulhilo64(M0, c[0])
            hi1, lo1 = _mulhilo64(M1, c[2])
            c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
            k0 = k0 + W0
            k1 = k1 + W1
    return c


def _u53(x):
    return (x >> numpy.uint64(11)).astype(numpy.float64) * 2.0 ** -53


SIM_CTR, SIM_KEY_ROOT, SIM_KEY_NODE = 0x73696d, 0x726f6f74, 0x6e6f6465


def simulate(f, Pm, nChar, seed, pNum=0, ref_draw=None):
    """Pf/p4_treeSim.c:14-590 for one part with the GPU's uniforms (Philox
    counted by (site, node, part)): per site the category floor(K u), the
    root state from the root comp's running sums (last entry 1.0) and the
    invariant flag u < pInvar -- or those from ref_draw [nChar][4] -- then
    each non-root node in preorder from the running sums of P[cat][parent
    state] (the picker decks of Pf/p4_node.c:545-586).  Pm: [N][K][S][S]
    P[s][j].  Returns (states [N][nChar], cats [nChar], -1 for invariant)."""
    K, S, N = f.nCat, f.dim, len(f.parent)
    root = int(f.root)
    sites = numpy.arange(nChar, dtype=numpy.uint64)
    states = numpy.zeros((N, nChar), numpy.int64)
    if ref_draw is not None:
        ref_draw = numpy.asarray(ref_draw)
        inv = ref_draw[:, 2] != 0
        cat = numpy.where(inv, 0, ref_draw[:, 1])
        states[root] = numpy.where(inv, ref_draw[:, 3], ref_draw[:, 0])
    else:
        r = philox4x64_vec(sites, 0, pNum, SIM_CTR, seed, SIM_KEY_ROOT)
        cat = numpy.minimum(numpy.floor(K * _u53(r[0])).astype(numpy.int64), K - 1) if K > 1 \
            else numpy.zeros(nChar, numpy.int64)
        comp = f.comps[f.compNums[root]]
        picker = numpy.cumsum(comp)
        picker[S - 1] = 1.0
        u = _u53(r[1])
        states[root] = numpy.argmax(u[:, None] < picker[None, :], axis=1)
        inv = (_u53(r[2]) < f.pInvar) if f.pInvar > 0 else numpy.zeros(nChar, bool)
    # preorder: the reverse of the postorder
    post = [int(x) for x in f.postOrder if int(x) >= 0]
    for n in reversed(post):
        if n == root:
            continue
        ps = states[f.parent[n]]
        u = _u53(philox4x64_vec(sites, n + 1, pNum, SIM_CTR, seed, SIM_KEY_NODE)[0])
        deck = numpy.cumsum(Pm[n], axis=2)     # [K][S][S] running sums of each row
        deck[:, :, S - 1] = 1.0
        rows = deck[cat, ps]                   # [nChar][S]
        new = numpy.argmax(u[:, None] < rows, axis=1)
        states[n] = numpy.where(inv, ps, new)
    return states, numpy.where(inv, -1, cat)
