"""Seeded synthetic workloads: random trees, model parameters, and sequence
simulation under the tree-heterogeneous model.

Simulation follows the semantics of p4's Tree.simulate / p4_simulate
(Pf/p4_treeSim.c:14-590): a gamma category drawn uniformly per site, the
root state drawn from the root composition, a site made invariant with
probability pInvar (it then keeps the root state everywhere), and other
sites mutated down the tree in preorder by each branch's P matrix.  The
random stream is numpy's, not GSL's, so data match p4's in distribution
only.

Workloads (sizes as named in the benchmark definitions):
  gtr16      16 taxa, 2,000 sites, GTR+G4 (alpha 0.5), one composition
  het64      64 taxa, 50,000 sites, 3 comps x 2 GTR matrices placed on
             nodes, G4 alpha ~ U(0.3, 2), pInvar 0.2
  prot128    128 taxa, 20,000 sites, 20 states, 190 exchangeabilities
             ~ LogNormal(0, 1), 2 comps ~ Dirichlet(1 x 20), G4
  multi48    48 taxa, 24 partitions (16 DNA + 8 protein) with sites
             log-uniform in [200, 40,000], each with its own model, G4
"""
import math
import random

import numpy

from .data import Alignment, Data
from .func import randomTree
from .var import var


def _qmatrix(R, pi):
    """Q = R diag(pi), rows sum to zero, mean rate 1 (Pf/util.c:44-115)."""
    S = len(pi)
    Q = R * pi[None, :]
    numpy.fill_diagonal(Q, 0.0)
    numpy.fill_diagonal(Q, -Q.sum(1))
    Q /= -(pi * numpy.diag(Q)).sum()
    return Q


def pmatrices(R, pi, ts):
    """P(t) for each t via the symmetrised eigensystem (host, FP64)."""
    Q = _qmatrix(R, pi)
    sq = numpy.sqrt(pi)
    A = sq[:, None] * Q / sq[None, :]
    A = 0.5 * (A + A.T)
    lam, U = numpy.linalg.eigh(A)
    V = U / sq[:, None]
    Vi = U.T * sq[None, :]
    ts = numpy.atleast_1d(ts)
    P = numpy.einsum('ik,tk,kj->tij', V, numpy.exp(numpy.outer(ts, lam)), Vi)
    return numpy.clip(P, 0.0, None)


def bigR_from_vals(vals, dim):
    R = numpy.zeros((dim, dim))
    k = 0
    for i in range(dim - 1):
        for j in range(i + 1, dim):
            R[i, j] = R[j, i] = vals[k]
            k += 1
    return R


def simulate_part(tree, pNum, nSites, rng):
    """Simulated codes [nTax][nSites] for model part pNum of tree."""
    mp = tree.model.parts[pNum]
    S, K = mp.dim, mp.nGammaCat
    nTax = len(tree.taxNames)
    cats = rng.integers(0, K, nSites)
    pinv = mp.pInvar.val if mp.pInvar else 0.0
    invar = rng.random(nSites) < pinv
    rootPi = mp.comps[tree.root.parts[pNum].compNum].val
    states = {tree.root.nodeNum: rng.choice(S, size=nSites, p=rootPi / rootPi.sum())}
    Rs = [bigR_from_vals(r.val, S) if r.spec != '2p' else None for r in mp.rMatrices]
    for n in tree.iterPreOrder():
        if n is tree.root:
            continue
        c = n.parts[pNum].compNum
        r = n.br.parts[pNum].rMatrixNum
        rates = mp.gdasrvs[n.br.parts[pNum].gdasrvNum].rates if K > 1 else numpy.ones(1)
        ts = n.br.len * numpy.asarray(rates) * mp.relRate
        if pinv > 0:
            ts = ts / (1.0 - pinv)
        P = pmatrices(Rs[r], mp.comps[c].val, ts)           # [K][S][S]
        cum = numpy.cumsum(P, axis=2)
        cum[:, :, -1] = 1.0
        par = states[n.parent.nodeNum]
        rows = cum[cats, par]                                 # [nSites][S]
        u = rng.random(nSites)
        new = (u[:, None] > rows).sum(1)
        new = numpy.minimum(new, S - 1)
        states[n.nodeNum] = numpy.where(invar, par, new)
    out = numpy.zeros((nTax, nSites), numpy.int32)
    for n in tree.iterLeavesNoRoot():
        out[n.seqNum] = states[n.nodeNum]
    return out


def _dirichlet(rng, alpha, dim):
    v = rng.dirichlet([alpha] * dim)
    v = numpy.maximum(v, var.PIVEC_MIN * 10)
    return v / v.sum()


def _attach(tree, parts_spec, rng, pyrng):
    """parts_spec: list of dicts with dataType, nSites, nComps, nRMat,
    exch sigma, comp alpha, gamma alpha, pInvar.  Returns (tree, Data)."""
    names = tree.taxNames
    # placeholder data so the model exists, then real data after simulation
    alns = []
    for ps in parts_spec:
        dt = ps['dataType']
        dim = 4 if dt == 'dna' else 20
        alns.append(Alignment.fromCodes(names, numpy.zeros((len(names), 1), numpy.int32), dt))
    tree.data = Data(alns)
    for pNum, ps in enumerate(parts_spec):
        dim = 4 if ps['dataType'] == 'dna' else 20
        for c in range(ps['nComps']):
            tree.newComp(partNum=pNum, free=1, spec='specified',
                         val=list(_dirichlet(rng, ps['compAlpha'], dim)))
        nEx = dim * (dim - 1) // 2
        for r in range(ps['nRMat']):
            tree.newRMatrix(partNum=pNum, free=1, spec='specified',
                            val=list(rng.lognormal(0.0, ps['exchSigma'], nEx)))
        tree.setNGammaCat(partNum=pNum, nGammaCat=ps.get('nCat', 4))
        if ps.get('nCat', 4) > 1:
            tree.newGdasrv(partNum=pNum, free=1, val=ps['alpha'])
        tree.setPInvar(partNum=pNum, free=1 if ps.get('pInvar', 0.0) > 0 else 0,
                       val=ps.get('pInvar', 0.0))
    tree.setModelComponentsOnNodesRandomly(forceRepresentation=2, rng=pyrng)
    codes = [simulate_part(tree, pNum, ps['nSites'], rng) for pNum, ps in enumerate(parts_spec)]
    # real data; the model objects are kept (Tree.data would replace them)
    model = tree.model
    node_parts = [(n.parts, n.br.parts) for n in tree.nodes]
    data = Data([Alignment.fromCodes(names, c, ps['dataType']) for c, ps in zip(codes, parts_spec)])
    tree.data = data
    tree._model = model
    for n, (a, b) in zip(tree.nodes, node_parts):
        n.parts, n.br.parts = a, b
    return tree, data


def make_workload(name, seed=1, nSites=None, precision='fp64'):
    """Build (tree, data) for a named workload.  nSites overrides the site
    count of single-partition workloads (smaller parity cases)."""
    rng = numpy.random.default_rng(seed)
    pyrng = random.Random(seed)
    if name == 'gtr16':
        tree = randomTree(nTax=16, rng=pyrng, meanBrLen=0.1)
        spec = [dict(dataType='dna', nSites=nSites or 2000, nComps=1, nRMat=1, compAlpha=1.0,
                     exchSigma=0.5, alpha=0.5, pInvar=0.0)]
    elif name == 'het64':
        tree = randomTree(nTax=64, rng=pyrng, meanBrLen=0.1)
        spec = [dict(dataType='dna', nSites=nSites or 50000, nComps=3, nRMat=2, compAlpha=2.0,
                     exchSigma=0.5, alpha=float(rng.uniform(0.3, 2.0)), pInvar=0.2)]
    elif name == 'prot128':
        tree = randomTree(nTax=128, rng=pyrng, meanBrLen=0.1)
        spec = [dict(dataType='protein', nSites=nSites or 20000, nComps=2, nRMat=1, compAlpha=1.0,
                     exchSigma=1.0, alpha=float(rng.uniform(0.3, 2.0)), pInvar=0.0)]
    elif name == 'multi48':
        tree = randomTree(nTax=48, rng=pyrng, meanBrLen=0.1)
        spec = []
        for i in range(24):
            dt = 'dna' if i < 16 else 'protein'
            ns = int(round(math.exp(rng.uniform(math.log(200), math.log(40000)))))
            if nSites:
                ns = max(50, ns * nSites // 40000)
            spec.append(dict(dataType=dt, nSites=ns, nComps=1, nRMat=1, compAlpha=1.0,
                             exchSigma=0.5 if dt == 'dna' else 1.0,
                             alpha=float(rng.uniform(0.3, 2.0)), pInvar=0.0))
    else:
        raise ValueError("unknown workload %s" % name)
    tree, data = _attach(tree, spec, rng, pyrng)
    tree.precision = precision
    tree.workload = dict(name=name, seed=seed, spec=spec)
    return tree, data
