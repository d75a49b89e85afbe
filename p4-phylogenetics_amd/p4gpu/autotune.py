"""Autotuner for the 4-state pruning kernel.

Times the lean kernel (the wave kernel's layout with precomputed op records
and tip P^T staged in LDS), the wave-per-category kernel
and every (ops per group, P staged in LDS) variant of the grouped kernel on
a workload, for FP32 and FP64 partials, and writes the fastest to
p4gpu/tuning.json, which the library applies when it is loaded
(opsPerGroup -3 = the lean kernel; 0 / -1 / -2 = the wave kernel with its
P^T read by vector / scalar loads / scalar loads with the tree split into
two streams of waves).  The grouped
variants compute identical results; the wave kernel keeps FP32 scale
exponents per category, which changes FP32 results only at rounding level.

    python -m p4gpu.autotune [--workload het64] [--reps 30]
"""
import argparse
import json
import os

from . import pf, synth


def time_variant(tree, reps):
    c = tree.cTree
    pf.setKernelTiming(True)
    for _ in range(3):
        pf.p4_setPrams(c, -1)
        pf.p4_treeLogLike(c, 0)
    us = []
    for _ in range(reps):
        pf.p4_setPrams(c, -1)
        pf.p4_treeLogLike(c, 0)
        us.append(pf.lastKernelMicroseconds(c))
    pf.setKernelTiming(False)
    us.sort()
    return us[len(us) // 2]


def autotune(workload="het64", reps=30, seed=1, write=True, verbose=True):
    out = {}
    for prec in ("fp32", "fp64"):
        tree, _ = synth.make_workload(workload, seed=seed, precision=prec)
        tree.calcLogLike(verbose=0)
        res = {}
        for G, label in ((-3, "lean"), (-1, "wave, scalar P"), (-2, "wave, scalar P, 2 streams"),
                         (0, "wave, vector P")):
            pf.setPruneTuning(prec == "fp32", G, 1)
            res[(G, 1)] = time_variant(tree, reps)
            if verbose:
                print("%s %s  %.1f us" % (prec, label, res[(G, 1)]), flush=True)
        for G in (1, 4):
            for lds in (0, 1):
                pf.setPruneTuning(prec == "fp32", G, lds)
                res[(G, lds)] = time_variant(tree, reps)
                if verbose:
                    print("%s G=%d lds=%d  %.1f us" % (prec, G, lds, res[(G, lds)]), flush=True)
        (G, lds), best = min(res.items(), key=lambda kv: kv[1])
        out[prec] = {"opsPerGroup": G, "stagePInLds": lds, "kernel_us": best,
                     "workload": workload,
                     "all": {({0: "wave_vectorP", -1: "wave_scalarP", -2: "wave_scalarP_2streams",
                                  -3: "lean"}.get(k[0]) or "G%d_lds%d" % k): v
                             for k, v in res.items()}}
        pf.setPruneTuning(prec == "fp32", G, lds)
    if write:
        with open(pf.TUNING_PATH, "w") as fh:
            json.dump(out, fh, indent=1)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="het64")
    ap.add_argument("--reps", type=int, default=30)
    args = ap.parse_args()
    print(json.dumps(autotune(args.workload, args.reps), indent=1))
