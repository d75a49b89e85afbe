"""Learned vs plain proposal MCMC: ESS per second and split frequencies.

Workload: 16 taxa, 2,000 simulated sites, GTR+G4 (the seeded gtr16 data),
free composition, exchangeabilities and gamma shape, topology moves on.

Phases (each writes JSON; they may run as separate processes):
  train    plain-proposal run of --gens generations sampled every
           --interval; the cold chain's samples train the network on the
           GPU; writes the network (--net) and a summary.
  plain    plain proposals only, --gens generations.
  learned  plain proposals + the learned move (weight --weight), loading --net.
  report   combines the plain and learned summaries: ESS/s ratios for lnL,
           tree length and the learned move's parameters (composition), and
           the split-frequency agreement (max difference and ASDOSS).

  python -m p4gpu.learned_bench --phase train --gens 100000 --interval 5 --net net.npz
  python -m p4gpu.learned_bench --phase plain --gens 1000000 --out plain.json
  python -m p4gpu.learned_bench --phase learned --gens 1000000 --net net.npz --out learned.json
  python -m p4gpu.learned_bench --phase report --plain plain.json --learned learned.json
"""
import argparse
import json
import sys
import time

import numpy

from . import synth
from .func import effectiveSampleSize
from .learned import LearnedProposal
from .mcmc import Mcmc, compareSplitFrequencies


def _tree(seed):
    tree, _ = synth.make_workload("gtr16", seed=seed)
    return tree


def _summary(m, burnFrac, wall, extra=None):
    n = len(m.samples['lnL'])
    b = int(burnFrac * n)
    prams = numpy.asarray(m.samples['prams'][b:], dtype=float)
    ess = {
        'lnL': effectiveSampleSize(numpy.asarray(m.samples['lnL'][b:])),
        'treeLen': effectiveSampleSize(numpy.asarray(m.samples['treeLen'][b:])),
    }
    for j in range(4):
        ess['comp%d' % j] = effectiveSampleSize(prams[:, j])
    ess['gammaShape'] = effectiveSampleSize(prams[:, 4])
    freqs = m.splitFrequencies(burnin=b)
    out = dict(gens=m.timing['gens'], seconds=wall, gensPerSecond=m.timing['gens'] / wall, samples=n - b,
               ess=ess, essPerSecond={k: v / wall for k, v in ess.items()},
               meanTreeLen=float(numpy.mean(m.samples['treeLen'][b:])),
               meanLnL=float(numpy.mean(m.samples['lnL'][b:])),
               splitFreqs={str(k): v for k, v in freqs.items()},
               acceptance={k: v for k, v in m.acceptanceRates().items()})
    if extra:
        out.update(extra)
    return out


def _progress(m, total, step, t0):
    done = 0
    while done < total:
        k = min(step, total - done)
        m.run(k, verbose=False)
        done += k
        print("  %d / %d generations, %.0f s" % (done, total, time.perf_counter() - t0), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--phase', required=True, choices=['train', 'plain', 'learned', 'report'])
    ap.add_argument('--gens', type=int, default=100000)
    ap.add_argument('--interval', type=int, default=100)
    ap.add_argument('--chains', type=int, default=1)
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--dataSeed', type=int, default=1)
    ap.add_argument('--net', default='learned_net.npz')
    ap.add_argument('--weight', type=float, default=1.0)
    ap.add_argument('--wRidge', type=float, default=0.4)
    ap.add_argument('--wBlock', type=float, default=0.3)
    ap.add_argument('--wComp', type=float, default=0.3)
    ap.add_argument('--epochs', type=int, default=30)
    ap.add_argument('--burnin', type=float, default=0.1)
    ap.add_argument('--out', default=None)
    ap.add_argument('--plain', default=None)
    ap.add_argument('--learned', default=None)
    a = ap.parse_args(argv)

    if a.phase == 'report':
        with open(a.plain) as fh:
            P = json.load(fh)
        with open(a.learned) as fh:
            L = json.load(fh)
        ratio = {k: L['essPerSecond'][k] / P['essPerSecond'][k] for k in P['essPerSecond']}
        fa = {int(k): v for k, v in P['splitFreqs'].items()}
        fb = {int(k): v for k, v in L['splitFreqs'].items()}
        mx, asd = compareSplitFrequencies(fa, fb, minFreq=0.0)
        rep = dict(essPerSecondRatio=ratio, minRatio=min(ratio.values()), splitMaxDiff=mx, splitASDOSS=asd,
                   plain=dict((k, P[k]) for k in ('gens', 'seconds', 'gensPerSecond', 'ess', 'meanTreeLen')),
                   learned=dict((k, L[k]) for k in ('gens', 'seconds', 'gensPerSecond', 'ess', 'meanTreeLen',
                                                     'learnedAcceptance')))
        s = json.dumps(rep, indent=1)
        print(s)
        if a.out:
            with open(a.out, 'w') as fh:
                fh.write(s)
        return 0

    tree = _tree(a.dataSeed)
    m = Mcmc(tree, nChains=a.chains, seed=a.seed, sampleInterval=a.interval, verbose=False)
    L = None
    if a.phase in ('train', 'learned'):
        L = LearnedProposal(tree, seed=a.seed, weights=dict(ridge=a.wRidge, block=a.wBlock, comp=a.wComp))
    if a.phase == 'train':
        L.attach(m)
    if a.phase == 'learned':
        L.load(a.net)
        m.addLearnedProposal(L, a.weight)
    print("phase %s: %d generations, %d chain(s), sample every %d" % (a.phase, a.gens, a.chains, a.interval),
          flush=True)
    t0 = time.perf_counter()
    _progress(m, a.gens, max(1000, a.gens // 20), t0)
    wall = time.perf_counter() - t0
    extra = {}
    if a.phase == 'train':
        nb = int(a.burnin * len(L.records))
        L.records = L.records[nb:]
        t1 = time.perf_counter()
        log = L.train(epochs=a.epochs, verbose=True)
        extra = dict(trainSamples=len(L.records), trainSeconds=time.perf_counter() - t1, trainLoss=log)
        L.save(a.net)
    if a.phase == 'learned':
        extra = dict(learnedAcceptance=L.stats['accepted'] / max(1, L.stats['proposed']),
                     learnedStats=dict(L.stats))
    summ = _summary(m, a.burnin, wall, extra)
    s = json.dumps(summ)
    if a.out:
        with open(a.out, 'w') as fh:
            fh.write(s)
    print(json.dumps(dict((k, summ[k]) for k in summ if k != 'splitFreqs')), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
