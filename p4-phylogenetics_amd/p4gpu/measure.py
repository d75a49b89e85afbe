"""Measurements beside the headline benchmark (bench.py measures het64).

  prot128   128 taxa, 20,000 sites, 20 states, 2 comps, G4: one full FP32
            evaluation on the MFMA kernel per step; reports the kernel's
            device time, its HBM fraction (algorithmic bytes 80,132 B per
            pattern) and MFMA utilisation (algorithmic flops
            K*2*S^2*E_int per pattern against 157.3 TFLOP/s FP32).
  multi48   48 taxa, 24 partitions (16 DNA, 8 protein, sites log-uniform
            in [200, 40,000]), 4 Metropolis-coupled chains: one unit of work
            = the pending evaluations of 4 chains in one batched call
            (p4g_batchLogLike); full recompute and dirty-path variants.

  het64     the 64-taxon, 50,000-site workload with one 4-state pruning
            kernel variant forced (opsPerGroup as in p4gpu.autotune: -3
            lean, -1 wave with scalar P, ...): full evaluations, the
            pruning kernel's median device time; for counter runs.

    python -m p4gpu.measure prot128 [--steps 50]
    python -m p4gpu.measure multi48 [--steps 50]
    python -m p4gpu.measure het64 --variant -3 --precision fp32 [--steps 4]
    python -m p4gpu.measure rows [--steps 10]
    python -m p4gpu.measure mcmc --workload het64 --precision fp32 [--steps 200]

  rows      het64 FP64: wall time per call of the branch derivatives, the
            root ancestral draws, a CPO sample, a simulation and the
            sequence summaries.
  mcmc      Mcmc.run generations per second, 4 chains, on a workload.
"""
import argparse
import json
import random
import time

from . import pf, synth

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def prot128(steps, warmup):
    tree, data = synth.make_workload("prot128", seed=1, precision="fp32")
    tree.calcLogLike(verbose=0)
    c = tree.cTree
    for _ in range(warmup):
        pf.p4_setPrams(c, -1)
        pf.p4_treeLogLike(c, 0)
    pf.setKernelTiming(True)
    kus = []
    t0 = time.perf_counter()
    for _ in range(steps):
        pf.p4_setPrams(c, -1)
        lnl = pf.p4_treeLogLike(c, 0)
        kus.append(pf.lastKernelMicroseconds(c))
    sec = (time.perf_counter() - t0) / steps
    pf.setKernelTiming(False)
    mp = tree.model.parts[0]
    S, K = mp.dim, mp.nGammaCat
    n = sum(1 for _ in tree.iterLeavesNoRoot())
    nInt = sum(1 for _ in tree.iterInternals())
    eInt = nInt - 1
    P = data.parts[0].nPatterns
    bytes_per = 2 * eInt * S * K * 4 + n + 4
    flops_per = K * 2 * S * S * eInt
    ku = sum(kus) / len(kus)
    tree.setPrecision("fp64")
    lnl64 = tree.calcLogLike(verbose=0)
    return dict(workload="prot128: 128 taxa, 20,000 sites (%d patterns), 20 states, 2 comps, G4, FP32 MFMA" % P,
                evals_per_s=1.0 / sec, kernel_us=ku,
                hbm_GBs=bytes_per * P / (ku * 1e-6) / 1e9, hbm_frac=bytes_per * P / (ku * 1e-6) / 1e9 / HBM_PEAK_GBS,
                mfma_TFLOPs=flops_per * P / (ku * 1e-6) / 1e12,
                mfma_util=flops_per * P / (ku * 1e-6) / 1e12 / FP32_PEAK_TFLOPS,
                bytes_per_pattern=bytes_per, flops_per_pattern=flops_per,
                lnL_fp32=lnl, lnL_fp64=lnl64, rel_err=abs(lnl - lnl64) / abs(lnl64))


def multi48(steps, warmup):
    base, data = synth.make_workload("multi48", seed=1, precision="fp32")
    base.calcLogLike(verbose=0)
    trees = [base] + [base.dupe() for _ in range(3)]
    for t in trees[1:]:
        t.precision = "fp32"
        t.calcLogLike(verbose=0)
    batch = pf.Batch([t.cTree for t in trees])
    rng = random.Random(3)

    def full():
        for t in trees:
            pf.p4_setPrams(t.cTree, -1)
        return batch.logLikes()

    def dirty():
        # one branch-length change per chain, its path marked in every part
        for t in trees:
            nodes = [n for n in t.iterNodesNoRoot()]
            n = nodes[rng.randrange(len(nodes))]
            pf.p4_setBrLen(n.cNode, n.br.len * (1.0 + 0.01 * (rng.random() - 0.5)))
            pf.p4_calculateBigPDecks(n.cNode)
            pf.markPath(t.cTree, n.parent.nodeNum, -1)
        return batch.logLikes()

    out = dict(workload="multi48: 48 taxa, 24 partitions (16 DNA + 8 protein, %d patterns), 4 chains, FP32"
               % sum(p.nPatterns for p in data.parts))
    # FP32 batch against FP64 single-tree evaluation of the same four states
    v32 = full()
    errs = []
    for t, v in zip(trees, v32):
        d = t.dupe()
        d.precision = "fp64"
        v64 = d.calcLogLike(verbose=0)
        errs.append(abs(v - v64) / abs(v64))
    out["lnL_rel_err_fp32_vs_fp64"] = max(errs)
    for name, fn in (("full", full), ("dirty_path", dirty)):
        for _ in range(warmup):
            fn()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        sec = (time.perf_counter() - t0) / steps
        out[name] = dict(units_per_s=1.0 / sec, ms_per_unit=sec * 1e3)
    return out


def het64(steps, variant, precision):
    pf.setPruneTuning(precision == "fp32", variant, 1)
    tree, _ = synth.make_workload("het64", seed=1, precision=precision)
    tree.calcLogLike(verbose=0)
    c = tree.cTree
    pf.setKernelTiming(True)
    us = []
    for _ in range(steps):
        pf.p4_setPrams(c, -1)
        lnl = pf.p4_treeLogLike(c, 0)
        us.append(pf.lastKernelMicroseconds(c))
    pf.setKernelTiming(False)
    us.sort()
    return dict(workload="het64", variant=variant, precision=precision, lnL=lnl,
                kernel_us_median=us[len(us) // 2], kernel_us_min=us[0])


def _wall(fn, steps, warmup=2):
    for _ in range(warmup):
        fn()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    return (time.perf_counter() - t0) / steps


def rows(steps):
    """Wall time per call of the root-side and derivative paths on het64
    (64 taxa, 50,000 sites, G4 + pInvar, FP64): every branch's d1/d2 with
    the up-pass partials, one root ancestral draw of every site, one CPO
    sample, one simulation of every site (sequences to the host included;
    patterns not remade)."""
    tree, data = synth.make_workload("het64", seed=1, precision="fp64")
    tree.calcLogLike(verbose=0)
    c = tree.cTree
    cp = data.parts[0].cPart
    n = pf.partNSites(cp)
    out = dict(workload="het64: 64 taxa, 50,000 sites (%d patterns), FP64" % data.parts[0].nPatterns,
               nodes=len(tree.nodes), sites=n)
    out["branch_derivatives_ms"] = 1e3 * _wall(lambda: pf.branchDerivatives(c, -1, len(tree.nodes)), steps)
    out["anc_draw_all_sites_ms"] = 1e3 * _wall(lambda: pf.drawAncStates(c, 0, 0, n, seed=7), steps)
    pf.cpoReset(cp)
    out["cpo_sample_ms"] = 1e3 * _wall(lambda: pf.cpoAccumulate(c), steps)
    sim, sdata = synth.make_workload("het64", seed=2, precision="fp64")
    sim.calcLogLike(verbose=0)
    seeds = iter(range(10 ** 6))
    out["simulate_ms"] = 1e3 * _wall(lambda: pf.p4_simulate(sim.cTree, None, next(seeds)), steps)
    out["seq_stats_ms"] = 1e3 * _wall(lambda: pf.partSeqStats(sdata.parts[0].cPart), steps)
    return out


def mcmc(gens, warmup, workload, precision, nChains=4):
    """Generations per second of Mcmc.run: nChains Metropolis-coupled
    chains, the default proposal mix, no sampling inside the timed
    generations."""
    from .mcmc import Mcmc
    tree, data = synth.make_workload(workload, seed=1, precision=precision)
    m = Mcmc(tree, nChains=nChains, seed=1, verbose=False, sampleInterval=10 ** 9)
    m.run(warmup, verbose=False)
    t0 = time.perf_counter()
    m.run(gens, verbose=False)
    sec = time.perf_counter() - t0
    acc = {p.name + ("" if p.pNum < 0 else ".%d" % p.pNum): round(sum(p.nAcceptances) / max(1, sum(p.nProposals)), 3)
           for p in m.proposals}
    return dict(workload="%s, %d chains, %s" % (workload, nChains, precision),
                patterns=sum(p.nPatterns for p in data.parts), gens_per_s=gens / sec, ms_per_gen=1e3 * sec / gens,
                acceptance=acc)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["prot128", "multi48", "het64", "rows", "mcmc"])
    ap.add_argument("--workload", default="het64")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--variant", type=int, default=-3)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    a = ap.parse_args(argv)
    if a.what == "prot128":
        res = prot128(a.steps, a.warmup)
    elif a.what == "multi48":
        res = multi48(a.steps, a.warmup)
    elif a.what == "rows":
        res = rows(a.steps)
    elif a.what == "mcmc":
        res = mcmc(a.steps, a.warmup, a.workload, a.precision)
    else:
        res = het64(a.steps, a.variant, a.precision)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
