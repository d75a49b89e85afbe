"""Benchmark: full log-likelihood evaluations per second on one MI355X.

One step = one full evaluation, the equivalent of p4's
pf.p4_setPrams(tree, -1) + pf.p4_treeLogLike(tree, 0): discrete-gamma
rates, Q and every P matrix, every partial, and the weighted site sum.
Workload: 64 taxa, 50,000 simulated sites, tree-heterogeneous DNA (3
compositions x 2 GTR matrices on nodes), G4 + pInvar 0.2.

The headline is the FP64 path (the reference's precision); the FP32 path
with power-of-two scale exponents is reported beside it with its own
roofline.  Inputs are resident in GPU memory when timing starts.  The CPU
baseline is the reference's own Pf code (compiled from Pf/*.c into
oracle/_ref) evaluating the same workload, all patterns, on one core and on
N cores (N independent processes), timed before the GPU is touched.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.
"""
import argparse
import json
import multiprocessing
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "p4-phylogenetics_amd"))

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(HERE, "profiles", "traffic_het64.json")


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


def kernel_name(prec):
    """The pruning kernel the tuning applies for this precision."""
    from p4gpu import pf
    G = None
    try:
        with open(pf.TUNING_PATH) as fh:
            G = int(json.load(fh)[prec]["opsPerGroup"])
    except (OSError, ValueError, KeyError):
        pass
    t = "float, true" if prec == "fp32" else "double, false"
    if G is None or G == -3:
        return "k_pruneLean<%s>" % t
    if G <= 0:
        return "k_pruneWave<%s, %s>" % (t, "true" if G == 0 else "false")
    return "k_prune<4, %s>" % t


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


# ------------------------------------------------------------ CPU baseline
def _cpu_worker(flat_kwargs, seconds, out):
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle
    f = oracle.Flat(**flat_kwargs)
    run = (lambda: oracle.ref_part_loglike(f)) if oracle.ref() is not None else (lambda: oracle.part_loglike(f))
    run()
    n, t0 = 0, time.perf_counter()
    while True:
        run()
        n += 1
        el = time.perf_counter() - t0
        if el > seconds:
            break
    out.put(n / el)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(tree, seconds=10.0):
    """The reference's own Pf CPU path (its P, eigen, gamma and
    conditional-likelihood functions compiled from Pf/*.c into
    oracle/_ref/libp4ref.so, driven by oracle/ref_driver.c) on the whole
    workload (all patterns), one core; then N cores running independent
    evaluations in N processes.  Without that library, the oracle (FP64
    restatement of the same loop nest)."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle
    f = oracle.flat_from_tree(tree, 0)
    kind = "reference" if oracle.ref() is not None else "port"
    ctx = multiprocessing.get_context("spawn")
    args = f.kwargs
    q = ctx.Queue()
    wait = 3 * seconds + 120
    p = ctx.Process(target=_cpu_worker, args=(args, seconds, q))
    p.start()
    one = q.get(timeout=wait)
    p.join()
    try:
        nCores = len(os.sched_getaffinity(0))
    except AttributeError:
        nCores = os.cpu_count() or 1
    nCores = max(1, min(nCores, 16))   # the GPU box's CPU share per GPU
    procs = [ctx.Process(target=_cpu_worker, args=(args, seconds, q)) for _ in range(nCores)]
    for pr in procs:
        pr.start()
    rates = [q.get(timeout=wait) for _ in procs]
    for pr in procs:
        pr.join()
    nPat = int(f.patterns.shape[1])
    return dict(value=one, unit="evals/s", cores=1, kind=kind,
                sample="all %d patterns, FP64, repeated full evaluations (P, eigen, gamma and pruning each time) "
                       "for %.0f s" % (nPat, seconds),
                n_core={"value": sum(rates), "unit": "evals/s", "cores": nCores,
                        "per_process": rates},
                cpu_model=cpu_model())


def cpu_rows(tree):
    """CPU timings beside the GPU rows (--rows), one core: the branch
    derivatives through the reference's own Pf code (P', P'', up-pass
    partials; the per-site sums restated), and the oracle's numpy
    restatements of the root draws and the simulation (kind 'port')."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import numpy
    import oracle
    f = oracle.flat_from_tree(tree, 0)
    out = {}
    if oracle.ref() is not None:
        t0 = time.perf_counter()
        oracle.ref_part_derivs(f)
        out["branch_derivatives_ms"] = {"value": 1e3 * (time.perf_counter() - t0), "kind": "reference", "cores": 1}
    _, _, cl = oracle.oracle_cl(f)
    _, _, spi, _, _ = tree.data.parts[0].patterns()
    u = numpy.random.default_rng(1).random(len(spi))
    t0 = time.perf_counter()
    oracle.anc_draw(f, cl[int(f.root)], spi, u)
    out["anc_draw_all_sites_ms"] = {"value": 1e3 * (time.perf_counter() - t0), "kind": "port", "cores": 1,
                                    "note": "root partials precomputed"}
    Pm = oracle.oracle_pmats(f)
    t0 = time.perf_counter()
    oracle.simulate(f, Pm, len(spi), seed=3)
    out["simulate_ms"] = {"value": 1e3 * (time.perf_counter() - t0), "kind": "port", "cores": 1}
    return out


def traffic_for(kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic_het64.json
    (rocprofv3 FETCH_SIZE + WRITE_SIZE of the same kernel on this workload,
    with the commit they were measured at); None if that file does not hold
    this kernel."""
    try:
        with open(TRAFFIC_FILE) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None, None
    k = d.get("kernels", {}).get(kernel)
    if not k:
        return None, None
    return k.get("hbm_bytes_per_launch"), d.get("commit")


def roofline(kernel, algo, kus):
    achieved = algo / (kus * 1e-6) / 1e9
    traffic, commit = traffic_for(kernel)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_measured_at": commit,
            "kernel": kernel, "kernel_us": kus, "algorithmic_bytes_per_launch": algo}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="seconds of CPU work per baseline leg (1 core, then N cores)")
    ap.add_argument("--rows", action="store_true",
                    help="also time the derivative, ancestral-draw, CPO and simulation paths (GPU and CPU)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (site patterns would be "
              "sharded across GPUs with one all-reduce of the part sums); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from p4gpu import pf, synth

    tree, data = synth.make_workload("het64", seed=args.seed, precision="fp64")
    nPat = data.parts[0].nPatterns
    cpu = None if args.no_cpu else cpu_baseline(tree, args.cpu_seconds)   # before the GPU is touched

    tree.calcLogLike(verbose=0)
    sec64, kus64, lnl64 = run_gpu(tree, args.steps, args.warmup, pf)
    tree.setPrecision("fp32")
    tree.calcLogLike(verbose=0)
    sec32, kus32, lnl32 = run_gpu(tree, args.steps, args.warmup, pf)
    rel32 = abs(lnl32 - lnl64) / abs(lnl64)

    algo64, perpat64 = algorithmic_bytes(tree, nPat, 8)
    algo32, perpat32 = algorithmic_bytes(tree, nPat, 4)
    rl64 = roofline(kernel_name("fp64"), algo64, kus64)
    rl64["bytes_per_pattern"] = perpat64
    rl32 = roofline(kernel_name("fp32"), algo32, kus32)
    rl32["bytes_per_pattern"] = perpat32

    out = {
        "metric": "log-like evals/s (50k-site DNA, 1 GPU)",
        "value": 1.0 / sec64,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded simulation under the model)",
        "config": {"workload": "het64: 64 taxa, 50,000 sites (%d patterns), 3 comps x 2 GTR "
                               "node-assigned, G4 + pInvar 0.2, FP64 partials" % nPat,
                   "seed": args.seed, "patterns": nPat},
        "roofline": rl64,
        "lnL_fp64": lnl64,
        "fp32_path": {"value": 1.0 / sec32, "unit": "evals/s", "ms_per_step": sec32 * 1e3,
                      "dtype": "f32 partials + int exponents", "roofline": rl32, "lnL": lnl32,
                      "lnL_rel_err_vs_fp64": rel32, "steps": args.steps},
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    if args.rows:
        from p4gpu import measure
        tree.setPrecision("fp64")
        out["rows"] = {"gpu": measure.rows(10), "cpu": cpu_rows(tree)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
