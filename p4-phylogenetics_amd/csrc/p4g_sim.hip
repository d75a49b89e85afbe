// p4g_sim.hip -- simulation of sequences on the tree, and the sequence
// summaries of posterior-predictive checks.
//
//  k_simulate   one thread per site (sites are independent given the
//               model): the site's gamma category, root state and
//               invariant flag, then each non-root node in preorder
//               mutates its parent's state through the running sums of its
//               P row (Pf/p4_treeSim.c:14-590; the picker decks of
//               Pf/p4_node.c:545-586, whose last entry is 1.0).  Uniforms
//               come from Philox4x64-10 keyed by the seed and counted by
//               (site, node, part), so a site's draws do not depend on the
//               launch shape.  Node states live in LDS ([node][thread] bytes;
//               a global [node][site] buffer for trees too large for that);
//               leaves also go to the part's [taxon][site] sequences.
//  k_seqStats   one thread per site: the distinct states among the taxa
//               (a bit mask), whether the site is constant, and per-taxon
//               state counts in LDS, added to global counters once per
//               workgroup (Pf/part.c:1419-1571).
#include "p4g_internal.h"

#include <cmath>

namespace p4g {

constexpr uint64_t SIM_CTR = 0x73696dull;       // "sim"
constexpr uint64_t SIM_KEY_ROOT = 0x726f6f74ull; // "root"
constexpr uint64_t SIM_KEY_NODE = 0x6e6f6465ull; // "node"

// LDS: the node states of the workgroup's sites live in LDS, [node][thread]
// bytes, so each node reads its parent's state there instead of from the
// global buffer it wrote a moment before (a dependent round trip per node).
template <bool LDS>
__global__ __launch_bounds__(256) void k_simulate(SimArgs a)
{
    extern __shared__ uint8_t stL[];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nChar) return;
    const int S = a.dim, K = a.K;
    const size_t nC = (size_t)a.nChar;
    const int bd = blockDim.x, tx = threadIdx.x;
    auto put = [&](int n, int v) {
        if (LDS)
            stL[n * bd + tx] = (uint8_t)v;
        else
            a.states[(size_t)n * nC + c] = (uint8_t)v;
    };
    auto get = [&](int n) -> int { return LDS ? (int)stL[n * bd + tx] : (int)a.states[(size_t)n * nC + c]; };
    int cat = 0, state = 0;
    bool inv = false;
    if (a.refDraw) {
        const int *d = a.refDraw + (size_t)c * 4;
        if (d[2]) {
            inv = true;
            state = d[3];
        } else {
            cat = d[1];
            state = d[0];
        }
    } else {
        const Philox4 r = philox4x64(Philox4{{(uint64_t)c, 0, (uint64_t)a.pNum, SIM_CTR}}, a.seed, SIM_KEY_ROOT);
        if (K > 1) cat = min((int)floor((double)K * uniform53(r.v[0])), K - 1);
        // the root composition's running sums, last entry 1.0
        const double u = uniform53(r.v[1]);
        double run = 0.0;
        state = S - 1;
        for (int k = 0; k < S - 1; k++) {
            run += a.rootComp[k];
            if (u < run) {
                state = k;
                break;
            }
        }
        inv = a.pInvar > 0.0 && uniform53(r.v[2]) < a.pInvar;
    }
    put(a.root, state);
    if (a.seqNum[a.root] >= 0) a.seq[(size_t)a.seqNum[a.root] * nC + c] = (int8_t)state;
    a.simCat[c] = inv ? -1 : cat;
    const size_t SS = (size_t)S * S;
    for (int i = 0; i < a.nPre; i++) {
        const int n = a.pre[i];
        const int ps = get(a.parent[n]);
        int s = ps;
        if (!inv) {
            const double u = uniform53(
                philox4x64(Philox4{{(uint64_t)c, (uint64_t)n + 1, (uint64_t)a.pNum, SIM_CTR}}, a.seed, SIM_KEY_NODE)
                    .v[0]);
            const double *PT = a.PT64 + ((size_t)n * K + cat) * SS;   // P[ps][l] = PT[l][ps]
            double run = 0.0;
            s = S - 1;
            for (int l = 0; l < S - 1; l++) {
                run += PT[(size_t)l * S + ps];
                if (u < run) {
                    s = l;
                    break;
                }
            }
        }
        put(n, s);
        if (a.seqNum[n] >= 0) a.seq[(size_t)a.seqNum[n] * nC + c] = (int8_t)s;
    }
}

constexpr int SIM_LDS_MAX = 64 * 1024;

hipError_t launchSimulate(const SimArgs &a, hipStream_t s)
{
    if (a.nChar <= 0) return hipSuccess;
    // the largest workgroup (256, 128 or 64 sites) whose node states fit in 64 KB of LDS
    int bd = 256;
    while (bd > 64 && (size_t)a.nNodes * bd > SIM_LDS_MAX) bd /= 2;
    if ((size_t)a.nNodes * bd <= SIM_LDS_MAX)
        hipLaunchKernelGGL(k_simulate<true>, dim3((a.nChar + bd - 1) / bd), dim3(bd), (size_t)a.nNodes * bd, s, a);
    else
        hipLaunchKernelGGL(k_simulate<false>, dim3((a.nChar + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

constexpr int STATS_LDS_COUNTS = 8192;   // per-taxon counts held in LDS up to this many

__global__ __launch_bounds__(256) void k_seqStats(const int8_t *seq, int nTax, int nChar, int dim, int *counts,
                                                  int *acc)
{
    __shared__ int hist[STATS_LDS_COUNTS];
    __shared__ int red[3];
    const int nh = nTax * dim;
    const bool lds = nh <= STATS_LDS_COUNTS;
    if (lds)
        for (int i = threadIdx.x; i < nh; i += 256) hist[i] = 0;
    if (threadIdx.x < 3) red[threadIdx.x] = 0;
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    int isConst = 0, distinct = 0, other = 0;
    if (c < nChar) {
        uint64_t m = 0;
        for (int t = 0; t < nTax; t++) {
            const int code = seq[(size_t)t * nChar + c];
            if (code >= 0 && code < dim) {
                m |= 1ull << code;
                if (lds)
                    atomicAdd(&hist[t * dim + code], 1);
                else
                    atomicAdd(&counts[t * dim + code], 1);
            } else {
                other++;
            }
        }
        distinct = __popcll(m);
        isConst = distinct == 1;
    }
    if (isConst) atomicAdd(&red[0], 1);
    if (distinct) atomicAdd(&red[1], distinct);
    if (other) atomicAdd(&red[2], other);
    __syncthreads();
    if (lds)
        for (int i = threadIdx.x; i < nh; i += 256)
            if (hist[i]) atomicAdd(&counts[i], hist[i]);
    if (threadIdx.x < 3 && red[threadIdx.x]) atomicAdd(&acc[threadIdx.x], red[threadIdx.x]);
}

hipError_t launchSeqStats(const int8_t *seq, int nTax, int nChar, int dim, int *counts, int *acc, hipStream_t s)
{
    if (nChar <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_seqStats, dim3((nChar + 255) / 256), dim3(256), 0, s, seq, nTax, nChar, dim, counts, acc);
    return hipGetLastError();
}

}  // namespace p4g
