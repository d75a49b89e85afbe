// p4g_anc.hip -- consumers of the root-side likelihoods: CPO sums and
// draws of the root's ancestral states.
//
//  k_cpoAccum   sum[p] += 1/like[p] per pattern, straight from the pattern
//               likelihoods the pruning kernel leaves in GPU memory (the
//               per-sample step of p4/mcmc.py:3286-3297, which fetched the
//               site likelihoods to the host and summed them in Python).
//  k_ancDraw    one thread per site: the root's conditional likelihoods for
//               the site's pattern, from its children's partials and P
//               (what the reference keeps in root->cl), then the draw of
//               Pf/p4_treeSim.c:591-845 -- a cumulative table over
//               (category, state), category-major, then the invariant
//               states, searched with u * (table total).
#include "p4g_internal.h"

#include <cmath>

namespace p4g {

__global__ __launch_bounds__(256) void k_cpoAccum(const double *like, double *sum, int nPat)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < nPat) sum[p] += 1.0 / like[p];
}

hipError_t launchCpoAccum(const double *like, double *sum, int nPat, hipStream_t s)
{
    if (nPat <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cpoAccum, dim3((nPat + 255) / 256), dim3(256), 0, s, like, sum, nPat);
    return hipGetLastError();
}

// Partial value and exponent of (slot, cat, pattern, state) in the layout
// the pruning kernel left.
__device__ __forceinline__ double ancClAt(const AncArgs &a, int slot, int k, int p, int j)
{
    const size_t nPP = (size_t)a.nPatPad;
    size_t o;
    if (a.layout == CL_WAVE)
        o = (((size_t)k * nPP) + p) * 4 + j;
    else if (a.layout == CL_TILED)
        o = (size_t)(p / 16) * a.K * 320 + (((size_t)k * 5 + j / 4) * 4 + j % 4) * 16 + p % 16;
    else
        o = ((size_t)p * a.K + k) * a.dim + j;
    const char *base = (const char *)a.cl + a.slotBytes * slot;
    return a.fp32 ? (double)((const float *)base)[o] : ((const double *)base)[o];
}
__device__ __forceinline__ int ancExpAt(const AncArgs &a, int slot, int k, int p)
{
    if (!a.fp32) return 0;
    const size_t o = a.layout == CL_WAVE ? (size_t)k * a.nPatPad + p : (size_t)p;
    return a.scl[(size_t)slot * a.sclPerSlot + o];
}

// The root's conditional likelihoods of category k for pattern p, scaled
// by 2^E (E returned): the product over the root's children of
// sum_j P[s][j] x_child[j].
template <int SM>
__device__ __forceinline__ int rootCl(const AncArgs &a, int k, int p, double (&r)[SM])
{
    const int S = a.dim;
    const size_t SS = (size_t)S * S;
#pragma unroll
    for (int s = 0; s < SM; s++) r[s] = 1.0;
    int E = 0;
    for (int c = 0; c < a.nKid; c++) {
        const DerivSib kd = a.kids[c];
        const double *PT = a.PT64 + ((size_t)kd.pnode * a.K + k) * SS;   // PT[j][s] = P[s][j]
        if (kd.kind == 0) {
            const unsigned code = a.tips[(size_t)kd.idx * a.nPatPad + p];
            if (code == TIP_ALL) continue;
            if (code < (unsigned)S) {
#pragma unroll
                for (int s = 0; s < SM; s++)
                    if (s < S) r[s] *= PT[(size_t)code * S + s];
            } else {
                const uint64_t m = a.ambMask[code - S];
#pragma unroll
                for (int s = 0; s < SM; s++)
                    if (s < S) {
                        double sum = 0.0;
                        for (int j = 0; j < S; j++)
                            if ((m >> j) & 1ull) sum += PT[(size_t)j * S + s];
                        r[s] *= sum;
                    }
            }
        } else {
            double x[SM];
#pragma unroll
            for (int j = 0; j < SM; j++)
                if (j < S) x[j] = ancClAt(a, kd.idx, k, p, j);
            E += ancExpAt(a, kd.idx, k, p);
#pragma unroll
            for (int s = 0; s < SM; s++)
                if (s < S) {
                    double sum = 0.0;
                    for (int j = 0; j < S; j++) sum += PT[(size_t)j * S + s] * x[j];
                    r[s] *= sum;
                }
        }
    }
    return E;
}

template <int SC>
__global__ __launch_bounds__(128) void k_ancDraw(AncArgs a)
{
    constexpr int SM = SC ? SC : MAX_DIM;
    const int i = blockIdx.x * 128 + threadIdx.x;
    if (i >= a.n) return;
    const int site = a.first + i;
    const int p = a.sitePat[site];
    const int S = SC ? SC : a.dim;
    const int K = a.K;
    const double fac = (1.0 - a.pInvar) / (double)K;   // freqsTimesOneMinusPInvar
    const double *pi = a.rootPi;
    const size_t nPP = (size_t)a.nPatPad;
    // invariant weights present?  Without them a common power of two of the
    // variable weights cancels, so exponents are taken relative to the
    // smallest (no underflow where the reference's unscaled values would
    // not underflow either).
    bool anyInv = false;
    if (a.pInvar != 0.0)
        for (int w = 0; w < a.nWords; w++) anyInv |= a.invMask[(size_t)w * nPP + p] != 0u;
    int Emin = 0;
    if (a.fp32 && !anyInv) {
        Emin = 0x7fffffff;
        for (int k = 0; k < K; k++) {
            double r[SM];
            Emin = min(Emin, rootCl<SM>(a, k, p, r));
        }
    }
    // pass 1: the table total
    double sLike = 0.0;
    for (int k = 0; k < K; k++) {
        double r[SM];
        const int E = rootCl<SM>(a, k, p, r) - Emin;
#pragma unroll
        for (int s = 0; s < SM; s++)
            if (s < S) sLike += ldexp(pi[s] * r[s] * fac, -E);
    }
    if (a.pInvar != 0.0)
        for (int s = 0; s < S; s++)
            if ((a.invMask[(size_t)(s >> 5) * nPP + p] >> (s & 31)) & 1u) sLike += pi[s] * a.pInvar;
    double u;
    if (a.u) {
        u = a.u[i];
    } else {
        Philox4 c{{(uint64_t)site, 0x616e63ull /* "anc" */, 0, 0}};
        u = uniform53(philox4x64(c, a.seed, 0x726f6f74ull /* "root" */).v[0]);
    }
    const double target = u * sLike;
    // pass 2: the same running sum until it passes the target
    int d0 = -1, d1 = -1, d2 = 0, d3 = -1;
    double run = 0.0;
    bool got = false;
    for (int k = 0; k < K && !got; k++) {
        double r[SM];
        const int E = rootCl<SM>(a, k, p, r) - Emin;
#pragma unroll
        for (int s = 0; s < SM; s++)
            if (s < S && !got) {
                run += ldexp(pi[s] * r[s] * fac, -E);
                if (target < run) {
                    got = true;
                    d0 = s;
                    d1 = k;
                }
            }
    }
    if (!got && a.pInvar != 0.0)
        for (int s = 0; s < S && !got; s++)
            if ((a.invMask[(size_t)(s >> 5) * nPP + p] >> (s & 31)) & 1u) {
                run += pi[s] * a.pInvar;
                if (target < run) {
                    got = true;
                    d2 = 1;
                    d3 = s;
                }
            }
    int *o = a.draw + (size_t)i * 4;
    o[0] = d0;
    o[1] = d1;
    o[2] = d2;
    o[3] = d3;
}

hipError_t launchAncDraw(const AncArgs &a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    const dim3 grid((a.n + 127) / 128), block(128);
    if (a.dim == 4)
        hipLaunchKernelGGL(k_ancDraw<4>, grid, block, 0, s, a);
    else if (a.dim == 20)
        hipLaunchKernelGGL(k_ancDraw<20>, grid, block, 0, s, a);
    else
        hipLaunchKernelGGL(k_ancDraw<0>, grid, block, 0, s, a);
    return hipGetLastError();
}

}  // namespace p4g
