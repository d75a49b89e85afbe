"""The entry point that allocates nothing (p4g_treeLogLikeNoAlloc) against
the ordinary one (p4g_treeLogLike) through the C-ABI: two calls in a row,
the second on a smaller input, both equal to the ordinary entry point's
result bit for bit (the same kernels reduce in the same order); a tree
with no device state yet is refused."""
import numpy
import pytest

from p4gpu import P4Error, pf, synth

pytestmark = pytest.mark.gpu


def test_no_alloc_entry_point_matches_tree_log_like():
    assert pf.deviceAvailable()
    big, _ = synth.make_workload("het64", seed=21, nSites=6000, precision="fp32")
    small, _ = synth.make_workload("gtr16", seed=22, nSites=700, precision="fp64")
    for t in (big, small):
        t.calcLogLike(verbose=0)          # sets the tree up (allocates)
    for t in (big, small):
        want = pf.p4_treeLogLike(t.cTree, 0)
        got = pf.treeLogLikeNoAlloc(t.cTree)
        assert got == want, (got, want)
        # the scratch array keeps the last call's contents: a second call
        # after a parameter change still gives the ordinary result
        n = t.nodes[2]
        n.br.len *= 1.25
        pf.p4_setBrLen(n.cNode, n.br.len)
        pf.p4_setPrams(t.cTree, -1)
        got2 = pf.treeLogLikeNoAlloc(t.cTree)
        pf.p4_setPrams(t.cTree, -1)
        want2 = pf.p4_treeLogLike(t.cTree, 0)
        assert got2 == want2 and got2 != got, (got2, want2, got)
        assert numpy.isfinite(got2)


def test_no_alloc_entry_point_refuses_a_tree_without_device_state():
    t, _ = synth.make_workload("gtr16", seed=23, nSites=300, precision="fp64")
    t.calcLogLike(verbose=0)
    t.setPrecision("fp32")   # the FP32 partials do not exist yet
    with pytest.raises(P4Error, match="evaluate it once"):
        pf.treeLogLikeNoAlloc(t.cTree)
    v = t.calcLogLike(verbose=0)
    pf.p4_setPrams(t.cTree, -1)
    assert pf.treeLogLikeNoAlloc(t.cTree) == v
