// p4g_deriv.hip -- branch-length derivatives of the log likelihood and the
// up-pass ("cl2") conditional likelihoods they need, in FP64.
//
//  k_pderiv       P' = r Q P and P'' = r^2 Q Q P for every (node, category),
//                 one thread per entry, from Q of the node's (comp, rMatrix)
//                 and the node's P (the reference evaluates the same
//                 derivatives as V diag(r lam e^(lam t)) V^-1,
//                 Pf/eig.c:294-395, called from Pf/p4_node.c:442-540).
//  k_edgeDerivs   ONE launch per part for every branch: each thread owns a
//                 (pattern, category) and walks the non-root nodes in
//                 preorder.  At node n it forms cl2[n] -- the likelihood of
//                 everything outside n's subtree given the state of n's
//                 parent -- from the parent's cl2 through the parent's P
//                 (or the root frequencies) times the siblings' downward
//                 messages (Pf/p4_node.c:858-1018 with the dispatch of
//                 Pf/p4_treeNewt.c:597-620), then the site likelihood across
//                 n's branch and its first and second derivatives with P, P'
//                 and P'' (Pf/p4_treeNewt.c:257-500).  The K category waves
//                 of a workgroup combine each pattern's categories in LDS;
//                 per-workgroup sums go to a buffer.
//  k_derivReduce  per node, the workgroups' sums in a fixed order with
//                 Neumaier compensation.
#include "p4g_internal.h"

#include <cmath>

namespace p4g {

__global__ __launch_bounds__(256) void k_pderiv(DerivPArgs a)
{
    const int S = a.dim, SS = S * S;
    const int e = blockIdx.x * 256 + threadIdx.x;   // entry of (item, cat, s, j)
    if (e >= a.nItems * a.K * SS) return;
    const int it = e / (a.K * SS);
    const int rem = e - it * a.K * SS;
    const int k = rem / SS;
    const int s = (rem - k * SS) / S;
    const int j = rem - k * SS - s * S;
    const int node = a.nodeList[it];
    const double r = a.rateFac[(size_t)it * a.K + k];
    const double *Q = a.Q + (size_t)a.eigSlot[it] * SS;
    const double *PT = a.PT64 + ((size_t)node * a.K + k) * SS;   // PT[j][s] = P[s][j]
    // QP[m][j] = sum_l Q[m][l] P[l][j]; P' = r Q P; P'' = r^2 Q QP
    double p1 = 0.0, p2 = 0.0;
    for (int m = 0; m < S; m++) {
        double qp = 0.0;
        for (int l = 0; l < S; l++) qp += Q[m * S + l] * PT[j * S + l];
        p1 += Q[s * S + m] * PT[j * S + m];
        p2 += Q[s * S + m] * qp;
    }
    const size_t o = ((size_t)node * a.K + k) * SS + (size_t)j * S + s;   // transposed like P^T
    a.P1T[o] = r * p1;
    a.P2T[o] = r * r * p2;
}

hipError_t launchPderiv(const DerivPArgs &a, hipStream_t s)
{
    const int n = a.nItems * a.K * a.dim * a.dim;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pderiv, dim3((n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <int SC>
__global__ __launch_bounds__(64 * DERIV_MAX_K) void k_edgeDerivs(DerivArgs a)
{
    constexpr int SM = SC ? SC : MAX_DIM;
    const int S = SC ? SC : a.dim;
    const int K = a.K;
    const int cat = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * 64 + lane;
    const size_t nPP = (size_t)a.nPatPad;
    const int SS = S * S;
    __shared__ double sL[DERIV_MAX_K][64], sF[DERIV_MAX_K][64], sS[DERIV_MAX_K][64];
    // element (slot, cat, pattern, state) of the down-pass partials
    auto clAt = [&](int slot, int s) -> double {
        const size_t o = a.waveLayout ? (((size_t)slot * K + cat) * nPP + p) * 4 + s
                                      : (((size_t)slot * nPP + p) * K + cat) * S + s;
        return a.cl[o];
    };
    auto cl2At = [&](int node) -> double * { return a.cl2 + (((size_t)node * K + cat) * nPP + p) * S; };
    auto PT = [&](const double *base, int node) { return base + ((size_t)node * K + cat) * SS; };
    // mask of a tip code: 0 for none of S states (never), all bits for TIP_ALL
    auto tipMask = [&](unsigned code) -> uint64_t {
        if (code < (unsigned)S) return 1ull << code;
        if (code == TIP_ALL) return ~0ull;
        return a.ambMask[code - S];
    };
    const double *pi = a.rootPi;
    for (int e = 0; e < a.nOps; e++) {
        const DerivOp op = a.ops[e];
        double av[SM];
        if (op.flags & DOP_PARENT_ROOT) {
            // the root frequencies (Pf/p4_node.c:858-879); a leaf root's own
            // data enters here too
            uint64_t m = ~0ull;
            if (a.rootLeafRow >= 0) m = tipMask(a.tips[(size_t)a.rootLeafRow * nPP + p]);
#pragma unroll
            for (int s = 0; s < SM; s++)
                if (s < S) av[s] = ((m >> s) & 1ull) ? pi[s] : 0.0;
        } else {
            // up through the parent's branch: sum_from cl2[parent][from] P[from][s]
            // (Pf/p4_node.c:883-937)
            const double *c2 = cl2At(op.parent);
            const double *Pp = PT(a.PT64, op.parent);   // Pp[s][from] = P[from][s]
#pragma unroll
            for (int s = 0; s < SM; s++)
                if (s < S) {
                    double sum = 0.0;
                    for (int f = 0; f < S; f++) sum += Pp[s * S + f] * c2[f];
                    av[s] = sum;
                }
        }
        // times every sibling's message down its branch (Pf/p4_node.c:942-1018)
        for (int q = 0; q < op.nSib; q++) {
            const DerivSib sb = a.sibs[op.sibStart + q];
            const double *Ps = PT(a.PT64, sb.pnode);     // Ps[j][s] = P[s][j]
            if (sb.kind == 0) {
                const unsigned code = a.tips[(size_t)sb.idx * nPP + p];
                if (code == TIP_ALL) continue;
                const uint64_t m = tipMask(code);
#pragma unroll
                for (int s = 0; s < SM; s++)
                    if (s < S) {
                        double sum;
                        if (code < (unsigned)S) {
                            sum = Ps[code * S + s];
                        } else {
                            sum = 0.0;
                            for (int j = 0; j < S; j++)
                                if ((m >> j) & 1ull) sum += Ps[j * S + s];
                        }
                        av[s] *= sum;
                    }
            } else {
                double x[SM];
#pragma unroll
                for (int j = 0; j < SM; j++)
                    if (j < S) x[j] = clAt(sb.idx, j);
#pragma unroll
                for (int s = 0; s < SM; s++)
                    if (s < S) {
                        double sum = 0.0;
                        for (int j = 0; j < S; j++) sum += Ps[j * S + s] * x[j];
                        av[s] *= sum;
                    }
            }
        }
        double *c2n = cl2At(op.node);
#pragma unroll
        for (int s = 0; s < SM; s++)
            if (s < S) c2n[s] = av[s];
        // the site likelihood across the node's branch, and d/dt, d2/dt2
        const double *P0 = PT(a.PT64, op.node), *P1 = PT(a.P1T, op.node), *P2 = PT(a.P2T, op.node);
        double like = 0.0, first = 0.0, second = 0.0;
        if (op.flags & DOP_LEAF) {
            const unsigned code = a.tips[(size_t)op.idx * nPP + p];
            if (code < (unsigned)S) {
                for (int f = 0; f < S; f++) {
                    const double t2 = av[f];
                    like += t2 * P0[code * S + f];
                    first += t2 * P1[code * S + f];
                    second += t2 * P2[code * S + f];
                }
            } else if (code == TIP_ALL) {
                for (int f = 0; f < S; f++) {
                    const double t2 = av[f];
                    like += t2;
                    for (int t = 0; t < S; t++) {
                        first += t2 * P1[t * S + f];
                        second += t2 * P2[t * S + f];
                    }
                }
            } else {
                const uint64_t m = a.ambMask[code - S];
                for (int f = 0; f < S; f++) {
                    const double t2 = av[f];
                    for (int t = 0; t < S; t++)
                        if ((m >> t) & 1ull) {
                            like += t2 * P0[t * S + f];
                            first += t2 * P1[t * S + f];
                            second += t2 * P2[t * S + f];
                        }
                }
            }
        } else {
            double x[SM];
#pragma unroll
            for (int t = 0; t < SM; t++)
                if (t < S) x[t] = clAt(op.idx, t);
            for (int f = 0; f < S; f++)
                for (int t = 0; t < S; t++) {
                    const double t2 = av[f] * x[t];
                    like += t2 * P0[t * S + f];
                    first += t2 * P1[t * S + f];
                    second += t2 * P2[t * S + f];
                }
        }
        sL[cat][lane] = like;
        sF[cat][lane] = first;
        sS[cat][lane] = second;
        __syncthreads();
        if (cat == 0) {
            double lnL = 0.0, d1 = 0.0, d2 = 0.0;
            if (p < a.nPat) {
                double likeS = 0.0, firstS = 0.0, secondS = 0.0;
                for (int k = 0; k < K; k++) {
                    likeS += sL[k][lane];
                    firstS += sF[k][lane];
                    secondS += sS[k][lane];
                }
                if (a.pInvar != 0.0) {
                    const double f = (1.0 - a.pInvar) / (double)K;
                    likeS *= f;
                    firstS *= f;
                    secondS *= f;
                    for (int w = 0; w < a.nWords; w++) {
                        uint32_t mw = a.invMask[(size_t)w * nPP + p];
                        while (mw) {
                            const int b = __ffs(mw) - 1;
                            mw &= mw - 1;
                            likeS += pi[w * 32 + b] * a.pInvar;
                        }
                    }
                } else if (K > 1) {
                    likeS /= (double)K;
                    firstS /= (double)K;
                    secondS /= (double)K;
                }
                const double cnt = a.counts[p];
                if (likeS < 1.0e-300) {   // Pf/p4_treeNewt.c's penalty for an impossible site
                    lnL = cnt * -100000.0;
                    d1 = cnt * 1000000.0;
                    d2 = cnt * 10000000.0;
                } else {
                    lnL = cnt * log(likeS);
                    d1 = cnt * (firstS / likeS);
                    d2 = cnt * ((secondS * likeS - firstS * firstS) / (likeS * likeS));
                }
            }
            for (int off = 32; off >= 1; off >>= 1) {
                lnL += __shfl_xor(lnL, off);
                d1 += __shfl_xor(d1, off);
                d2 += __shfl_xor(d2, off);
            }
            if (lane == 0) {
                double *o = a.partial + ((size_t)e * gridDim.x + blockIdx.x) * 3;
                o[0] = lnL;
                o[1] = d1;
                o[2] = d2;
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(64) void k_derivReduce(const double *partial, int nBlk, const DerivOp *ops, int nOps,
                                                    double *out)
{
    const int e = blockIdx.x;
    const int lane = threadIdx.x;
    if (e >= nOps) return;
    double v[3];
    for (int c = 0; c < 3; c++) {
        double s = 0.0, comp = 0.0;
        for (int i = lane; i < nBlk; i += 64) {
            const double x = partial[((size_t)e * nBlk + i) * 3 + c];
            const double t = s + x;   // Neumaier
            if (fabs(s) >= fabs(x))
                comp += (s - t) + x;
            else
                comp += (x - t) + s;
            s = t;
        }
        double w = s + comp;
        for (int off = 32; off >= 1; off >>= 1) w += __shfl_xor(w, off);
        v[c] = w;
    }
    if (lane == 0) {
        const int node = ops[e].node;
        out[node * 3] = v[0];
        out[node * 3 + 1] = v[1];
        out[node * 3 + 2] = v[2];
    }
}

hipError_t launchEdgeDerivs(const DerivArgs &a, int nBlocks, hipStream_t s)
{
    const dim3 block(64 * a.K);
    if (a.dim == 4)
        hipLaunchKernelGGL(k_edgeDerivs<4>, dim3(nBlocks), block, 0, s, a);
    else if (a.dim == 20)
        hipLaunchKernelGGL(k_edgeDerivs<20>, dim3(nBlocks), block, 0, s, a);
    else
        hipLaunchKernelGGL(k_edgeDerivs<0>, dim3(nBlocks), block, 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_derivReduce, dim3(a.nOps), dim3(64), 0, s, a.partial, nBlocks, a.ops, a.nOps, a.out);
    return hipGetLastError();
}

}  // namespace p4g
