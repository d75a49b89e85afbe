"""Helpers (p4/func.py subset): newick reading, seeded random trees, and the
MCMC judges (effective sample size)."""
import random
import re

import numpy

from .node import Node
from .p4exceptions import P4Error
from .tree import Tree
from .var import var

_TOKEN = re.compile(r"\s*([(),;:]|[^(),;:\s]+)")


def readNewick(s, name=None):
    """Parse a newick string into a Tree.  Nodes are numbered in the order
    they are met (the root is node 0), as p4 numbers nodes read from a
    newick description."""
    toks = _TOKEN.findall(s)
    if not toks or toks[-1] != ';':
        raise P4Error("readNewick: missing ';'")
    t = Tree()
    t.name = name
    pos = [0]

    def new():
        n = Node()
        n.nodeNum = len(t.nodes)
        t.nodes.append(n)
        return n

    def peek():
        return toks[pos[0]]

    def take():
        tok = toks[pos[0]]
        pos[0] += 1
        return tok

    def parse(parent):
        n = new()
        n.parent = parent
        if peek() == '(':
            take()
            prev = None
            while True:
                c = parse(n)
                if prev is None:
                    n.leftChild = c
                else:
                    prev.sibling = c
                prev = c
                tok = take()
                if tok == ')':
                    break
                if tok != ',':
                    raise P4Error("readNewick: unexpected %s" % tok)
            if peek() not in (',', ')', ';', ':'):
                n.name = take()
        else:
            n.name = take()
            n.isLeaf = 1
        if peek() == ':':
            take()
            n.br.len = float(take())
        return n

    t.root = parse(None)
    if take() != ';':
        raise P4Error("readNewick: trailing tokens")
    t.setPreAndPostOrder()
    return t


def randomTree(taxNames=None, nTax=None, seed=None, rng=None, meanBrLen=0.1, yule=True):
    """A seeded random tree: Yule (pure-birth) topology grown from a rooted
    pair, then unrooted so that the root is a trifurcating internal node;
    branch lengths ~ Exponential(mean meanBrLen), clamped to
    [BRLEN_MIN, BRLEN_MAX].  Nodes are renumbered in preorder (root 0)."""
    rng = rng or random.Random(seed)
    if taxNames is None:
        taxNames = ["t%i" % i for i in range(nTax)]
    n = len(taxNames)
    if n < 3:
        raise P4Error("randomTree needs at least 3 taxa")
    # grow a rooted Yule tree as child lists
    children = {0: []}
    leaves = [0]
    nxt = 1
    while len(leaves) < n:
        k = rng.randrange(len(leaves)) if yule else len(leaves) - 1
        v = leaves.pop(k)
        a, b = nxt, nxt + 1
        nxt += 2
        children[v] = [a, b]
        children[a] = []
        children[b] = []
        leaves += [a, b]
    # unroot: move the root into its first internal child (trifurcation)
    r0 = 0
    a, b = children[r0]
    if not children[a]:
        a, b = b, a
    root = a
    children[root] = children[root] + [b]
    del children[r0]
    rng.shuffle(leaves)
    leafName = {v: taxNames[i] for i, v in enumerate(leaves)}
    t = Tree()
    order = []
    stack = [root]
    while stack:
        v = stack.pop()
        order.append(v)
        stack.extend(reversed(children[v]))
    num = {v: i for i, v in enumerate(order)}
    for v in order:
        nd = Node()
        nd.nodeNum = num[v]
        if not children[v]:
            nd.isLeaf = 1
            nd.name = leafName[v]
        t.nodes.append(nd)
    for v in order:
        nd = t.nodes[num[v]]
        prev = None
        for c in children[v]:
            cn = t.nodes[num[c]]
            cn.parent = nd
            if prev is None:
                nd.leftChild = cn
            else:
                prev.sibling = cn
            prev = cn
    t.root = t.nodes[0]
    for nd in t.nodes[1:]:
        bl = rng.expovariate(1.0 / meanBrLen)
        nd.br.len = min(max(bl, var.BRLEN_MIN), var.BRLEN_MAX)
    t.setPreAndPostOrder()
    return t


def effectiveSampleSize(data, mean=None):
    """Effective sample size as Tracer v1.4 computes it (p4/func.py:2953-3021,
    Pf/pfmodule.c:2631-2698): autocovariances up to maxLag = min(N/3, 1000),
    summed in adjacent-lag pairs while the pair sum stays positive."""
    data = numpy.asarray(data, dtype=numpy.float64)
    N = len(data)
    if mean is None:
        mean = data.mean()
    mean = float(numpy.asarray(mean).reshape(-1)[0])
    maxLag = min(int(N / 3), 1000)
    d = data - mean
    prev = 0.0
    var0 = 0.0
    varStat = 0.0
    lag = 0
    while lag < maxLag:
        g = float(numpy.dot(d[:N - lag], d[lag:])) / (N - lag)
        if lag == 0:
            varStat = g
            var0 = g
        elif lag % 2 == 0:
            if prev + g > 0.0:
                varStat += 2.0 * (prev + g)
            else:
                break
        lag += 1
        prev = g
    if varStat == 0:
        return 0.0
    return N * (var0 / varStat)
