"""Shared test helpers: golden fixtures and rebuilding trees from them."""
import os

import numpy

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

LNL_CASES = ["gtr16_500", "het64_800", "prot128_150", "edge_dna_gaps_pinv",
             "multifurcation_root4", "leaf_root"]


def load(name):
    return dict(numpy.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def flat_from_fixture(fx):
    import oracle
    return oracle.Flat(parent=fx["parent"], leftChild=fx["leftChild"], sibling=fx["sibling"],
                       isLeaf=fx["isLeaf"], seqNum=fx["seqNum"], brLen=fx["brLen"],
                       postOrder=fx["postOrder"], root=int(fx["root"]), compNums=fx["compNums"],
                       rMatrixNums=fx["rMatrixNums"], gdasrvNums=fx["gdasrvNums"],
                       dim=int(fx["dim"]), nCat=int(fx["nCat"]), comps=fx["comps"],
                       bigR=fx["bigR"], rates=_rates(fx), alphas=fx["alphas"],
                       pInvar=float(fx["pInvar"]), relRate=float(fx["relRate"]),
                       patterns=fx["patterns"], counts=fx["counts"], equates=fx["equates"],
                       invarVec=numpy.zeros(1), invarArray=numpy.zeros(1))


def _rates(fx):
    import oracle
    K = int(fx["nCat"])
    if K == 1:
        return numpy.ones((1, 1))
    return numpy.array([oracle.discrete_gamma(float(a), K) for a in fx["alphas"]])


def oracle_flat(fx):
    """Flat description with invariant sites computed by the oracle."""
    import oracle
    f = flat_from_fixture(fx)
    iv, ia = oracle.invar_sites(f.patterns, f.dim, f.equates)
    kw = dict(f.kwargs)
    kw.update(invarVec=iv, invarArray=ia)
    return oracle.Flat(**kw)


def tree_from_fixture(fx, precision="fp64"):
    """A p4gpu Tree with the fixture's topology, model and patterns."""
    from p4gpu import Data, Part, Tree
    from p4gpu.node import Node
    N = len(fx["parent"])
    t = Tree()
    for i in range(N):
        n = Node()
        n.nodeNum = i
        n.isLeaf = int(fx["isLeaf"][i])
        n.seqNum = int(fx["seqNum"][i])
        if n.isLeaf:
            n.name = "t%d" % n.seqNum
        n.br.len = float(fx["brLen"][i]) if i != int(fx["root"]) else 0.1
        t.nodes.append(n)
    for i, n in enumerate(t.nodes):
        p, l, s = int(fx["parent"][i]), int(fx["leftChild"][i]), int(fx["sibling"][i])
        n.parent = t.nodes[p] if p >= 0 else None
        n.leftChild = t.nodes[l] if l >= 0 else None
        n.sibling = t.nodes[s] if s >= 0 else None
    t.root = t.nodes[int(fx["root"])]
    t.setPreAndPostOrder()
    dt = str(fx["dataType"])
    nTax = fx["patterns"].shape[0]
    part = Part.fromPatterns(fx["patterns"], fx["counts"], dt)
    t.data = Data(parts=[part], taxNames=["t%d" % i for i in range(nTax)])
    dim = int(fx["dim"])
    for c in fx["comps"]:
        t.newComp(free=1, spec="specified", val=list(c))
    iu = numpy.triu_indices(dim, 1)
    for R in fx["bigR"]:
        t.newRMatrix(free=1, spec="specified", val=list(R[iu]))
    K = int(fx["nCat"])
    t.setNGammaCat(nGammaCat=K)
    if K > 1:
        for a in fx["alphas"]:
            t.newGdasrv(free=1, val=float(a))
    t.setPInvar(free=1, val=float(fx["pInvar"]))
    if float(fx["relRate"]) != 1.0:
        t.setRelRate(val=float(fx["relRate"]))
    for i, n in enumerate(t.nodes):
        n.parts[0].compNum = int(fx["compNums"][i])
        if n is not t.root:
            n.br.parts[0].rMatrixNum = int(fx["rMatrixNums"][i])
            n.br.parts[0].gdasrvNum = int(fx["gdasrvNums"][i]) if K > 1 else -1
    t.precision = precision
    return t
