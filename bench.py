"""Benchmark: full log-likelihood evaluations per second on one MI355X.

One step = one full evaluation, the equivalent of p4's
pf.p4_setPrams(tree, -1) + pf.p4_treeLogLike(tree, 0): discrete-gamma
rates, Q + eigensystems, every P matrix, every partial, and the weighted
site sum.  Workload: 64 taxa, 50,000 simulated sites, tree-heterogeneous
DNA (3 compositions x 2 GTR matrices on nodes), G4 + pInvar 0.2, FP32
partials with power-of-two scale exponents (the FP64 path is reported
beside it).  Inputs are resident in GPU memory when timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "p4-phylogenetics_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(tree, nPat, elt):
    """Bytes one full evaluation must move: each non-root internal partial
    written once and read once, tip codes (1 B), weights (4 B), invariant
    masks (4 B when pInvar > 0), and the P matrices."""
    mp = tree.model.parts[0]
    S, K = mp.dim, mp.nGammaCat
    nLeaves = sum(1 for n in tree.iterLeavesNoRoot())
    nInt = sum(1 for n in tree.iterInternals())
    eInt = nInt - 1          # edges into internal, non-root nodes
    per_pattern = 2 * eInt * S * K * elt + nLeaves * 1 + 4 + (4 if mp.pInvar.val > 0 else 0)
    nBranches = len(tree.nodes) - 1
    return nPat * per_pattern + nBranches * K * S * S * elt, per_pattern


def run_gpu(tree, steps, warmup, pf):
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
    t1 = time.perf_counter()
    pf.setKernelTiming(False)
    return (t1 - t0) / steps, sum(kus) / len(kus), lnl


def cpu_baseline(tree, target_seconds=12.0):
    """The reference's own Pf CPU path (its P, eigen, gamma and
    conditional-likelihood functions compiled from Pf/*.c into
    oracle/_ref/libp4ref.so, driven by oracle/ref_driver.c), one core, on a
    bounded sample of the same workload's patterns.  Without that library,
    the oracle (scalar FP64 restatement of the same loop nest)."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import numpy
    import oracle
    pats, cnts, _, _, _ = tree.data.parts[0].patterns()
    nPat = pats.shape[1]
    sample = min(nPat, 2000)
    f = oracle.flat_from_tree(tree, 0, patterns=numpy.ascontiguousarray(pats[:, :sample]),
                              counts=cnts[:sample])
    if oracle.ref() is not None:
        kind, run = "reference", lambda: oracle.ref_part_loglike(f)
    else:
        kind, run = "port", lambda: oracle.part_loglike(f)
    run()  # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        run()
        n += 1
        el = time.perf_counter() - t0
        if el > target_seconds:
            break
    per_eval_sample = el / n
    per_eval_full = per_eval_sample * nPat / sample
    return dict(value=1.0 / per_eval_full, unit="evals/s", cores=1, kind=kind,
                sample="%d of %d patterns, %d evaluations in %.1f s (P, eigen, gamma and pruning each time); "
                       "scaled to all patterns" % (sample, nPat, n, el))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (site patterns would be "
              "sharded across GPUs with one all-reduce of the part sums); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from p4gpu import pf, synth

    tree, data = synth.make_workload("het64", seed=args.seed, precision="fp32")
    tree.calcLogLike(verbose=0)
    nPat = data.parts[0].nPatterns
    sec32, kus32, lnl32 = run_gpu(tree, args.steps, args.warmup, pf)

    tree.setPrecision("fp64")
    tree.calcLogLike(verbose=0)
    sec64, kus64, lnl64 = run_gpu(tree, max(args.steps // 2, 1), max(args.warmup // 2, 1), pf)
    rel32 = abs(lnl32 - lnl64) / abs(lnl64)

    algo32, perpat32 = algorithmic_bytes(tree, nPat, 4)
    algo64, _ = algorithmic_bytes(tree, nPat, 8)
    achieved = algo32 / (kus32 * 1e-6) / 1e9

    kname = "k_prune<4,float,true>"
    try:
        with open(pf.TUNING_PATH) as fh:
            if int(json.load(fh)["fp32"]["opsPerGroup"]) <= 0:
                kname = "k_pruneWave<float,true>"
    except (OSError, ValueError, KeyError):
        kname = "k_pruneWave<float,true>"
    traffic = None
    tf = os.path.join(HERE, "profiles", "traffic_het64_fp32.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    out = {
        "metric": "log-like evals/s (50k-site DNA, 1 GPU)",
        "value": 1.0 / sec32,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded simulation under the model)",
        "config": {"workload": "het64: 64 taxa, 50,000 sites (%d patterns), 3 comps x 2 GTR "
                               "node-assigned, G4 + pInvar 0.2, FP32 partials + scale exponents" % nPat,
                   "seed": args.seed, "patterns": nPat},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "kernel_us": kus32,
                     "algorithmic_bytes_per_launch": algo32, "bytes_per_pattern": perpat32},
        "lnL_fp32": lnl32,
        "lnL_fp64": lnl64,
        "lnL_rel_err_fp32_vs_fp64": rel32,
        "fp64_path": {"value": 1.0 / sec64, "unit": "evals/s", "kernel_us": kus64,
                      "achieved_GBs": algo64 / (kus64 * 1e-6) / 1e9},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tree)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
