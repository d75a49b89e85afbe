"""p4gpu -- MI355X likelihood engine and MCMC loop behind p4's
Tree.calcLogLike() and Mcmc, as a drop-in for that path.

    import sys; sys.path.insert(0, "<repo>/p4-phylogenetics_amd")
    from p4gpu import func, var, Data, Alignment
    t = func.readNewick("((a:0.1,b:0.2):0.05,c:0.3,d:0.1);")
    t.data = Data(Alignment(names, seqs, 'dna'))
    t.newComp(free=1, spec='empirical'); t.newRMatrix(free=1, spec='specified', val=[...])
    t.setNGammaCat(nGammaCat=4); t.newGdasrv(free=1, val=0.5); t.setPInvar()
    t.calcLogLike()

All likelihood arithmetic runs in libp4gpu.so (HIP kernels for gfx950);
there is no CPU fallback.
"""
from .p4exceptions import P4Error
from .var import var
from .node import Node
from .model import Comp, RMatrix, Gdasrv, PInvar, Model, ModelPart
from .data import Alignment, Data, Part
from .tree import Tree
from . import func, pf

__all__ = ["P4Error", "var", "Node", "Comp", "RMatrix", "Gdasrv", "PInvar", "Model",
           "ModelPart", "Alignment", "Data", "Part", "Tree", "func", "pf"]
