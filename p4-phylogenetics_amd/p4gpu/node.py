"""Tree vertices (p4/node.py): Node, its branch, and per-part model numbers."""


class NodeBranchPart(object):
    def __init__(self):
        self.rMatrixNum = -1
        self.gdasrvNum = -1


class NodeBranch(object):
    def __init__(self):
        self.len = 0.1
        self.parts = []
        self.lenChanged = False
        self.splitKey = None
        self.rawSplitKey = None


class NodePart(object):
    def __init__(self):
        self.compNum = -1


class Node(object):
    """A vertex of a Tree.  Pointers to parent, leftChild and sibling may be
    None; all nodes but the root have a branch (br)."""

    def __init__(self):
        self.name = None
        self.nodeNum = -1
        self.parent = None
        self.leftChild = None
        self.sibling = None
        self.isLeaf = 0
        self.cNode = None
        self.seqNum = -1
        self.br = NodeBranch()
        self.parts = []
        self.flag = 0

    def __repr__(self):
        return "<Node %i %s>" % (self.nodeNum, self.name or "")

    def iterChildren(self):
        c = self.leftChild
        while c:
            yield c
            c = c.sibling

    def getNChildren(self):
        return sum(1 for _ in self.iterChildren())

    def rightmostChild(self):
        c = self.leftChild
        while c and c.sibling:
            c = c.sibling
        return c
