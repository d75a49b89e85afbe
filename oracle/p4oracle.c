/*
 * p4oracle.c -- CPU FP64 restatement of p4's Pf likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library (tests/,
 * __graft_entry__.smoke(), bench.py cpu_baseline).  The product never links
 * or calls this file.
 *
 * Scalar, single-threaded, no partial-likelihood scaling: the same numeric
 * contract as the reference Pf code.  Each function names the reference
 * function it restates.  The one deliberate difference is the eigensolver:
 * the reference uses EISPACK's general real eigen route (Pf/linalg.c:249-335)
 * while this oracle uses cyclic Jacobi on the symmetrised reversible Q.  P(t)
 * is unique, so both routes agree to rounding; the compiled reference subset
 * (oracle/_ref) pins that.
 *
 * Parity status of each function (checks live in tests/):
 *   or_discrete_gamma & callees : pinned (known answers in Pf/p4_treeSim.c:185-197,
 *                                 golden vectors from compiled Pf/gamma.c, mean rate = 1)
 *   or_bigq                     : pinned (_ref golden Q, row sums 0, detailed balance)
 *   or_eigen_reversible, or_pmat: pinned (_ref golden P, scipy expm, JC/F81 closed forms)
 *   or_part_loglike             : pinned (_ref golden cl + lnL, brute-force enumeration,
 *                                 2-taxon closed form, all-gap site = 1)
 *   or_make_patterns            : pinned (pattern->site expansion round trip)
 *   or_invar_sites              : pinned (hand-built known answers in tests)
 */
#include "p4oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Discrete gamma: Yang (1994) equal-probability categories, mean rate */
/* of each category.  Restates Pf/gamma.c:17-50 (DiscreteGamma) and its */
/* callees, which implement published algorithms:                       */
/*   AS 91 (Best & Roberts 1975) chi-square percentage points,           */
/*   AS 32 (Bhattacharjee 1970) incomplete gamma ratio,                  */
/*   Pike & Hill (1966, CACM alg. 291) log gamma,                        */
/*   Odeh & Evans (1974, AS 70) normal percentage points.                */
/* The constants below are those of the published algorithms; the      */
/* iteration tolerances (0.5e-6 in AS91, 1e-8 in AS32) set how far the  */
/* rates sit from the exact quantiles, so they must be the same.        */
/* ------------------------------------------------------------------ */

/* Pf/gamma.c:232-250 (Pike & Hill log gamma, Stirling with shift to x>=7) */
double or_ln_gamma(double alpha)
{
    double x = alpha, f = 0.0, z;
    if (x < 7) {
        f = 1.0;
        z = x - 1.0;
        for (;;) {
            z += 1.0;
            if (!(z < 7.0)) break;
            f *= z;
        }
        x = z;
        f = -log(f);
    }
    z = 1.0 / (x * x);
    return f + (x - 0.5) * log(x) - x + 0.918938533204673 +
           (((-0.000595238095238 * z + 0.000793650793651) * z - 0.002777777777778) * z +
            0.083333333333333) / x;
}

/* Pf/gamma.c:272-290 (Odeh & Evans normal quantile) */
double or_point_normal(double prob)
{
    const double a0 = -0.322232431088, a1 = -1.0, a2 = -0.342242088547,
                 a3 = -0.0204231210245, a4 = -0.453642210148e-4;
    const double b0 = 0.0993484626060, b1 = 0.588581570495, b2 = 0.531103462366,
                 b3 = 0.103537752850, b4 = 0.0038560700634;
    double p = prob, p1, y, z;
    p1 = (p < 0.5 ? p : 1 - p);
    if (p1 < 1e-20) return -9999;
    y = sqrt(log(1 / (p1 * p1)));
    z = y + ((((y * a4 + a3) * y + a2) * y + a1) * y + a0) /
                ((((y * b4 + b3) * y + b2) * y + b1) * y + b0);
    return (p < 0.5 ? -z : z);
}

/* Pf/gamma.c:149-223 (AS 32 incomplete gamma ratio) */
double or_incomplete_gamma(double x, double alpha, double lnGammaAlpha)
{
    const double accurate = 1e-8, overflow = 1e30;
    double p = alpha, g = lnGammaAlpha, factor, gin, rn, a, b, an, dif, term, pn[6];
    int i;

    if (x == 0.0) return 0.0;
    if (x < 0 || p <= 0) return -1.0;
    factor = exp(p * log(x) - x - g);
    if (!(x > 1 && x >= p)) {
        /* series expansion */
        gin = 1.0;
        term = 1.0;
        rn = p;
        do {
            rn++;
            term *= x / rn;
            gin += term;
        } while (term > accurate);
        gin *= factor / p;
        return gin;
    }
    /* continued fraction */
    a = 1.0 - p;
    b = a + x + 1.0;
    term = 0.0;
    pn[0] = 1.0;
    pn[1] = x;
    pn[2] = x + 1;
    pn[3] = x * b;
    gin = pn[2] / pn[3];
    for (;;) {
        a++;
        b += 2.0;
        term++;
        an = a * term;
        for (i = 0; i < 2; i++) pn[i + 4] = b * pn[i + 2] - an * pn[i];
        if (pn[5] != 0) {
            rn = pn[4] / pn[5];
            dif = fabs(gin - rn);
            if (!(dif > accurate) && dif <= accurate * rn) return 1.0 - factor * gin;
            gin = rn;
        }
        for (i = 0; i < 4; i++) pn[i] = pn[i + 2];
        if (!(fabs(pn[4]) < overflow)) {
            for (i = 0; i < 4; i++) pn[i] /= overflow;
        }
    }
}

/* Pf/gamma.c:66-129 (AS 91 chi-square percentage point) */
double or_point_chi2(double prob, double v)
{
    const double e = 0.5e-6, aa = 0.6931471805;
    double p = prob, g, xx, c, ch, a = 0.0, q = 0.0, p1 = 0.0, p2 = 0.0, t = 0.0, x = 0.0,
           b = 0.0, s1, s2, s3, s4, s5, s6;

    if (p < 0.000002 || p > 0.999998 || v <= 0.0) return -1.0;
    g = or_ln_gamma(v / 2.0);
    xx = v / 2.0;
    c = xx - 1.0;
    if (!(v >= -1.24 * log(p))) {
        /* starting approximation for small df */
        ch = pow((p * xx * exp(g + xx * aa)), 1.0 / xx);
        if (ch - e < 0) return ch;
    } else if (!(v > 0.32)) {
        /* Newton on the df<=0.32 approximation */
        ch = 0.4;
        a = log(1.0 - p);
        do {
            q = ch;
            p1 = 1.0 + ch * (4.67 + ch);
            p2 = ch * (6.73 + ch * (6.66 + ch));
            t = -0.5 + (4.67 + 2.0 * ch) / p1 - (6.73 + ch * (13.32 + 3.0 * ch)) / p2;
            ch -= (1.0 - exp(a + g + 0.5 * ch + c * aa) * p2 / p1) / t;
        } while (!(fabs(q / ch - 1.0) - 0.01 <= 0.0));
    } else {
        /* Wilson-Hilferty start */
        x = or_point_normal(p);
        p1 = 0.222222 / v;
        ch = v * pow((x * sqrt(p1) + 1.0 - p1), 3.0);
        if (ch > 2.2 * v + 6.0) ch = -2.0 * (log(1.0 - p) - c * log(0.5 * ch) + g);
    }
    /* seven-term Taylor refinement */
    do {
        q = ch;
        p1 = 0.5 * ch;
        if ((t = or_incomplete_gamma(p1, xx, g)) < 0.0) return -1.0;
        p2 = p - t;
        t = p2 * exp(xx * aa + g + p1 - c * log(ch));
        b = t / ch;
        a = 0.5 * t - b * c;
        s1 = (210.0 + a * (140.0 + a * (105.0 + a * (84.0 + a * (70.0 + 60.0 * a))))) / 420.0;
        s2 = (420.0 + a * (735.0 + a * (966.0 + a * (1141.0 + 1278.0 * a)))) / 2520.0;
        s3 = (210.0 + a * (462.0 + a * (707.0 + 932.0 * a))) / 2520.0;
        s4 = (252.0 + a * (672.0 + 1182.0 * a) + c * (294.0 + a * (889.0 + 1740.0 * a))) / 5040.0;
        s5 = (84.0 + 264.0 * a + c * (175.0 + 606.0 * a)) / 2520.0;
        s6 = (120.0 + c * (346.0 + 127.0 * c)) / 5040.0;
        ch += t * (1 + 0.5 * t * s1 - b * c * (s1 - b * (s2 - b * (s3 - b * (s4 - b * (s5 - b * s6))))));
    } while (fabs(q / ch - 1.0) > e);
    return ch;
}

/* Pf/gamma.c:17-50 (DiscreteGamma).  rK: mean rate of each of K equal-
 * probability categories of Gamma(alfa, beta); freqK ends as 1/K. */
int or_discrete_gamma(double *freqK, double *rK, double alfa, double beta, int K, int median)
{
    int i;
    double gap05 = 1.0 / (2.0 * K), t, factor = alfa / beta * K, lnga1;

    if (K == 1) {
        freqK[0] = 1.0;
        rK[0] = 1.0;
        return 0;
    }
    if (median) {
        for (i = 0; i < K; i++)
            rK[i] = or_point_chi2((i * 2.0 + 1) * gap05, 2.0 * alfa) / (2.0 * beta);
        for (i = 0, t = 0; i < K; i++) t += rK[i];
        for (i = 0; i < K; i++) rK[i] *= factor / t;
    } else {
        lnga1 = or_ln_gamma(alfa + 1);
        for (i = 0; i < K - 1; i++) freqK[i] = or_point_chi2((i + 1.0) / K, 2.0 * alfa) / (2.0 * beta);
        for (i = 0; i < K - 1; i++) freqK[i] = or_incomplete_gamma(freqK[i] * beta, alfa + 1, lnga1);
        rK[0] = freqK[0] * factor;
        rK[K - 1] = (1 - freqK[K - 2]) * factor;
        for (i = 1; i < K - 1; i++) rK[i] = (freqK[i] - freqK[i - 1]) * factor;
    }
    for (i = 0; i < K; i++) freqK[i] = 1.0 / K;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Rate matrix.  Pf/util.c:44-74 (setBigQFromRMatrixDotCharFreq) and   */
/* Pf/util.c:76-115 (normalizeBigQ).  Returns nonzero where the        */
/* reference would exit(1) (sum of pi outside [0.999, 1.001]).         */
/* ------------------------------------------------------------------ */
int or_bigq(int dim, const double *bigR, const double *pi, double *Q)
{
    int r, c;
    double sum, sumPi = 0.0, sumODE = 0.0;
    for (c = 0; c < dim; c++)
        for (r = 0; r < dim; r++) Q[r * dim + c] = bigR[r * dim + c] * pi[c];
    for (r = 0; r < dim; r++) {
        sum = 0.0;
        for (c = 0; c < dim; c++)
            if (r != c) sum = sum + Q[r * dim + c];
        Q[r * dim + r] = -sum;
    }
    for (r = 0; r < dim; r++) sumPi = sumPi + pi[r];
    if (sumPi < 0.999 || sumPi > 1.001) return 1;
    for (r = 0; r < dim; r++)
        for (c = 0; c < dim; c++)
            if (r != c) sumODE = sumODE + pi[r] * Q[r * dim + c];
    sumODE = 1.0 / sumODE;
    for (r = 0; r < dim; r++)
        for (c = 0; c < dim; c++) Q[r * dim + c] = Q[r * dim + c] * sumODE;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Eigensystem.  Replaces Pf/eig.c:63-161 (eigensystem: EISPACK route + */
/* LU inverse).  Q is reversible (pi_i Q_ij = pi_j Q_ji) so             */
/* A = Pi^{1/2} Q Pi^{-1/2} is symmetric; cyclic Jacobi gives A = U L U^T */
/* and then V = Pi^{-1/2} U, Vinv = U^T Pi^{1/2}.                       */
/* ------------------------------------------------------------------ */
int or_eigen_reversible(int dim, const double *Q, const double *pi, double *V, double *Vinv,
                        double *lam)
{
    int n = dim, i, j, k, p, q, sweep;
    double *A = (double *)malloc(sizeof(double) * n * n);
    double *U = (double *)malloc(sizeof(double) * n * n);
    double *sq = (double *)malloc(sizeof(double) * n);
    if (!A || !U || !sq) {
        free(A);
        free(U);
        free(sq);
        return 1;
    }
    for (i = 0; i < n; i++) sq[i] = sqrt(pi[i]);
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) {
            /* average the two triangles so A is exactly symmetric */
            double aij = sq[i] * Q[i * n + j] / sq[j];
            double aji = sq[j] * Q[j * n + i] / sq[i];
            A[i * n + j] = (i == j) ? Q[i * n + i] : 0.5 * (aij + aji);
            U[i * n + j] = (i == j) ? 1.0 : 0.0;
        }
    for (sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, diag = 0.0;
        for (i = 0; i < n; i++) {
            diag += A[i * n + i] * A[i * n + i];
            for (j = i + 1; j < n; j++) off += A[i * n + j] * A[i * n + j];
        }
        if (off == 0.0 || off < 1e-34 * diag) break;
        for (p = 0; p < n - 1; p++) {
            for (q = p + 1; q < n; q++) {
                double apq = A[p * n + q], theta, t, c, s;
                if (apq == 0.0) continue;
                theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                c = 1.0 / sqrt(t * t + 1.0);
                s = t * c;
                for (k = 0; k < n; k++) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (k = 0; k < n; k++) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (k = 0; k < n; k++) {
                    double ukp = U[k * n + p], ukq = U[k * n + q];
                    U[k * n + p] = c * ukp - s * ukq;
                    U[k * n + q] = s * ukp + c * ukq;
                }
            }
        }
    }
    for (k = 0; k < n; k++) lam[k] = A[k * n + k];
    for (i = 0; i < n; i++)
        for (k = 0; k < n; k++) {
            V[i * n + k] = U[i * n + k] / sq[i];
            Vinv[k * n + i] = U[i * n + k] * sq[i];
        }
    free(A);
    free(U);
    free(sq);
    return 0;
}

/* Pf/eig.c:163-223 (matrixExpTimesBranchLength, real branch):
 * P_ij = sum_k V_ik Vinv_kj exp(lam_k t), same summation order. */
void or_pmat(int dim, const double *V, const double *Vinv, const double *lam, double t,
             double *P)
{
    int i, j, k;
    for (i = 0; i < dim; i++)
        for (j = 0; j < dim; j++) {
            double r = 0.0;
            for (k = 0; k < dim; k++) r = r + (V[i * dim + k] * Vinv[k * dim + j] * exp(lam[k] * t));
            P[i * dim + j] = r;
        }
}

/* Pf/p4_tree.c:218-541 (p4_setPramsPart, Q/eig for every (comp,rMatrix)
 * used by a non-root node) and Pf/p4_node.c:296-346
 * (p4_calculateBigPDecksPart: t = brLen*rate*relRate[/(1-pInvar)]).
 * Gamma rates are inputs (mp->gdasrvRates), computed by or_discrete_gamma.
 * P: [nNodes][nCat][dim][dim]. */
int or_set_prams_part(const or_tree *tr, const or_modelpart *mp, double *P)
{
    int dim = mp->dim, K = mp->nCat, n, c, r, k;
    int nQ = mp->nComps * mp->nRMatrices;
    int *done = (int *)calloc(nQ, sizeof(int));
    double *Q = (double *)malloc(sizeof(double) * dim * dim * nQ);
    double *V = (double *)malloc(sizeof(double) * dim * dim * nQ);
    double *Vi = (double *)malloc(sizeof(double) * dim * dim * nQ);
    double *L = (double *)malloc(sizeof(double) * dim * nQ);
    int ret = 0;
    if (!done || !Q || !V || !Vi || !L) {
        ret = 2;
        goto out;
    }
    /* comp checks, Pf/p4_tree.c:424-452 (PIVEC_MIN is checked by the caller) */
    for (c = 0; c < mp->nComps; c++) {
        double s = 0.0;
        for (k = 0; k < dim; k++) s += mp->comps[c * dim + k];
        if (fabs(s - 1.0) > 1e-14) {
            ret = 3;
            goto out;
        }
    }
    for (n = 0; n < tr->nNodes; n++) {
        if (n == tr->root) continue;
        c = tr->compNums[n];
        r = tr->rMatrixNums[n];
        if (c < 0 || c >= mp->nComps || r < 0 || r >= mp->nRMatrices ||
            (mp->nGdasrvs && (tr->gdasrvNums[n] < 0 || tr->gdasrvNums[n] >= mp->nGdasrvs))) {
            ret = 6; /* node model numbers out of range */
            goto out;
        }
        if (!done[c * mp->nRMatrices + r]) {
            int slot = c * mp->nRMatrices + r;
            if (or_bigq(dim, mp->bigR + (size_t)r * dim * dim, mp->comps + (size_t)c * dim,
                        Q + (size_t)slot * dim * dim)) {
                ret = 4;
                goto out;
            }
            if (or_eigen_reversible(dim, Q + (size_t)slot * dim * dim, mp->comps + (size_t)c * dim,
                                    V + (size_t)slot * dim * dim, Vi + (size_t)slot * dim * dim,
                                    L + (size_t)slot * dim)) {
                ret = 5;
                goto out;
            }
            done[slot] = 1;
        }
    }
    for (n = 0; n < tr->nNodes; n++) {
        int slot;
        const double *rates = NULL;
        if (n == tr->root) continue;
        slot = tr->compNums[n] * mp->nRMatrices + tr->rMatrixNums[n];
        if (mp->nGdasrvs) rates = mp->gdasrvRates + (size_t)tr->gdasrvNums[n] * K;
        for (k = 0; k < K; k++) {
            double t;
            if (mp->pInvar == 0.0) {
                t = mp->nGdasrvs ? (tr->brLen[n] * rates[k] * mp->relRate) : (tr->brLen[n] * mp->relRate);
            } else {
                t = mp->nGdasrvs ? (tr->brLen[n] * rates[k] * mp->relRate) / (1.0 - mp->pInvar)
                                 : (tr->brLen[n] * mp->relRate) / (1.0 - mp->pInvar);
            }
            or_pmat(dim, V + (size_t)slot * dim * dim, Vi + (size_t)slot * dim * dim,
                    L + (size_t)slot * dim, t, P + ((size_t)n * K + k) * dim * dim);
        }
    }
out:
    free(done);
    free(Q);
    free(V);
    free(Vi);
    free(L);
    return ret;
}

static int is_n_like(const or_part *dp, int code)
{
    int j;
    for (j = 0; j < dp->dim; j++)
        if (!dp->equates[(code - OR_EQUATES_BASE) * dp->dim + j]) return 0;
    return 1;
}

/* Pf/p4_node.c:636-857 (p4_setConditionalLikelihoodsOfInternalNodePart).
 * Loop nest pattern -> category -> parent state -> {first child; siblings}. */
static void or_cl_node(const or_tree *tr, const or_modelpart *mp, const or_part *dp,
                       const double *P, double *cl, int node)
{
    int dim = mp->dim, K = mp->nCat, nP = dp->nPatterns, p, k, s, j;
    size_t clNode = (size_t)K * dim * nP;
    double *out = cl + (size_t)node * clNode;
    for (p = 0; p < nP; p++) {
        for (k = 0; k < K; k++) {
            for (s = 0; s < dim; s++) {
                int child = tr->leftChild[node], first = 1;
                double *dst = out + ((size_t)k * dim + s) * nP + p;
                while (child >= 0) {
                    const double *Pc = P + ((size_t)child * K + k) * dim * dim + (size_t)s * dim;
                    double v;
                    int skip = 0;
                    if (tr->isLeaf[child]) {
                        int code = dp->patterns[(size_t)tr->seqNum[child] * nP + p];
                        if (code >= 0) {
                            v = Pc[code];
                        } else if (code == OR_GAP_CODE || code == OR_QMARK_CODE) {
                            v = 1.0;
                            skip = !first;
                        } else if (code >= OR_EQUATES_BASE && code < OR_EQUATES_BASE + dp->nEquates) {
                            if (is_n_like(dp, code)) {
                                v = 1.0;
                                skip = !first;
                            } else {
                                v = 0.0;
                                for (j = 0; j < dp->dim; j++)
                                    if (dp->equates[(code - OR_EQUATES_BASE) * dp->dim + j]) v = v + Pc[j];
                            }
                        } else {
                            v = 0.0; /* unreachable for validated input */
                        }
                    } else {
                        const double *cc = cl + (size_t)child * clNode + (size_t)k * dim * nP + p;
                        v = 0.0;
                        for (j = 0; j < dim; j++) v = v + (Pc[j] * cc[(size_t)j * nP]);
                    }
                    if (first)
                        *dst = v;
                    else if (!skip)
                        *dst *= v;
                    first = 0;
                    child = tr->sibling[child];
                }
            }
        }
    }
}

/* Pf/p4_tree.c:868-922 (p4_treeLogLike postorder) + 924-1027 (p4_partLogLike)
 * + 1029-1197 (p4_partLogLikeLoop) + 1199-1378 (...RootLeaf). */
double or_part_loglike(const or_tree *tr, const or_modelpart *mp, const or_part *dp,
                       const double *P, double *cl, double *patternLikes)
{
    int dim = mp->dim, K = mp->nCat, nP = dp->nPatterns, i, p, k, s;
    size_t clNode = (size_t)K * dim * nP;
    const double *pi;
    const double *rcl;
    double lnL = 0.0, like, fTOMP;
    int root = tr->root;

    for (i = 0; i < tr->nNodes; i++) {
        int n = tr->postOrder[i];
        if (n == OR_NO_ORDER) continue;
        if (!tr->isLeaf[n] || n == root) or_cl_node(tr, mp, dp, P, cl, n);
    }
    pi = mp->comps + (size_t)tr->compNums[root] * dim;
    rcl = cl + (size_t)root * clNode;
    fTOMP = (1.0 - mp->pInvar) / (double)K;
    for (p = 0; p < nP; p++) {
        like = 0.0;
        for (k = 0; k < K; k++) {
            if (tr->isLeaf[root]) {
                int code = dp->patterns[(size_t)tr->seqNum[root] * nP + p];
                if (code >= 0) {
                    like += pi[code] * rcl[((size_t)k * dim + code) * nP + p];
                } else if (code == OR_GAP_CODE || code == OR_QMARK_CODE ||
                           is_n_like(dp, code)) {
                    for (s = 0; s < dim; s++) like += pi[s] * rcl[((size_t)k * dim + s) * nP + p];
                } else {
                    for (s = 0; s < dp->dim; s++)
                        if (dp->equates[(code - OR_EQUATES_BASE) * dp->dim + s])
                            like += pi[s] * rcl[((size_t)k * dim + s) * nP + p];
                }
            } else {
                for (s = 0; s < dim; s++) like += pi[s] * rcl[((size_t)k * dim + s) * nP + p];
            }
        }
        if (mp->pInvar != 0.0) {
            like *= fTOMP;
            if (dp->invarVec[p] > 0) {
                for (s = 0; s < dp->dim; s++)
                    if (dp->invarArray[(size_t)s * nP + p]) like += pi[s] * mp->pInvar;
            }
        } else if (K > 1) {
            like = like / (double)K;
        }
        if (like <= 0.0) return -1.0e99;
        lnL = lnL + (dp->patternCounts[p] * log(like));
        if (patternLikes) patternLikes[p] = like;
    }
    return lnL;
}

double or_tree_loglike_part(const or_tree *tr, const or_modelpart *mp, const or_part *dp,
                            double *patternLikes)
{
    size_t nP = (size_t)mp->nCat * mp->dim * mp->dim * tr->nNodes;
    size_t nC = (size_t)tr->nNodes * mp->nCat * mp->dim * dp->nPatterns;
    double *P = (double *)calloc(nP, sizeof(double));
    double *cl = (double *)calloc(nC, sizeof(double));
    double r;
    if (!P || !cl) {
        free(P);
        free(cl);
        return -2.0e99;
    }
    if (or_set_prams_part(tr, mp, P)) {
        r = -3.0e99;
    } else {
        r = or_part_loglike(tr, mp, dp, P, cl, patternLikes);
    }
    free(P);
    free(cl);
    return r;
}

/* Pf/part.c:317-448 (makePatterns): first-occurrence order, quadratic
 * search as in the reference.  Returns nPatterns.  patterns: [nTax][nChar]. */
int or_make_patterns(int nTax, int nChar, const int *seq, int *patterns, int *patternCounts,
                     int *seqPosPatternIndex)
{
    int i, j, i2, nPat = 0;
    for (i = 0; i < nChar; i++) patternCounts[i] = 0;
    for (i = 0; i < nChar; i++) {
        int already = -1;
        for (i2 = 0; i2 < nPat && already < 0; i2++) {
            for (j = 0; j < nTax; j++)
                if (seq[(size_t)j * nChar + i] != patterns[(size_t)j * nChar + i2]) break;
            if (j == nTax) already = i2;
        }
        if (already >= 0) {
            patternCounts[already]++;
            seqPosPatternIndex[i] = already;
        } else {
            for (j = 0; j < nTax; j++) patterns[(size_t)j * nChar + nPat] = seq[(size_t)j * nChar + i];
            patternCounts[nPat] = 1;
            seqPosPatternIndex[i] = nPat;
            nPat++;
        }
    }
    return nPat;
}

/* Pf/part.c:716-848 (setGlobalInvarSitesVec): per pattern, the states
 * compatible with every taxon.  patterns: [nTax][nPatterns]. */
void or_invar_sites(int nTax, int dim, int nPatterns, const int *patterns, int nEquates,
                    const int *equates, int *invarVec, int *invarArray)
{
    int p, t, j;
    for (p = 0; p < nPatterns; p++) {
        int sum = 0;
        for (j = 0; j < dim; j++) {
            int ok = 1;
            for (t = 0; t < nTax && ok; t++) {
                int code = patterns[(size_t)t * nPatterns + p];
                if (code >= 0)
                    ok = (code == j);
                else if (code == OR_N_LIKE || code == OR_GAP_CODE || code == OR_QMARK_CODE)
                    ok = 1;
                else if (code - OR_EQUATES_BASE >= 0 && code - OR_EQUATES_BASE < nEquates)
                    ok = equates[(code - OR_EQUATES_BASE) * dim + j] != 0;
                else
                    ok = 0;
            }
            invarArray[(size_t)j * nPatterns + p] = ok;
            sum += ok;
        }
        invarVec[p] = sum;
    }
}
