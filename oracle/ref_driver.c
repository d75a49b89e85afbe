/*
 * ref_driver.c -- drives the reference Pf C code, compiled from its own
 * sources under /root/reference/Pf, to produce golden vectors.
 *
 * TEST INFRASTRUCTURE ONLY (builds oracle/_ref/libp4ref.so, which is
 * git-ignored).  This file is our own code: it fills the reference's
 * structs (Pf/pftypes.h) from flat arrays and calls the reference's
 * functions:
 *   DiscreteGamma                       Pf/gamma.c:17-50
 *   setBigQFromRMatrixDotCharFreq       Pf/util.c:44-74
 *   normalizeBigQ                       Pf/util.c:76-115
 *   allocEig / eigensystem              Pf/eig.c:11-161
 *   matrixExpTimesBranchLength          Pf/eig.c:163-223
 *   p4_newModel / p4_newModelPart /
 *   p4_newRMatrix / p4_resetBQET        Pf/p4_model.c:12-51,90-209,345-452,513-543
 *   p4_newNode                          Pf/p4_node.c:10-136
 *   p4_calculateBigPDecksPart           Pf/p4_node.c:296-439
 *   p4_setConditionalLikelihoodsOfInternalNode  Pf/p4_node.c:601-857
 * Pf/p4_tree.c cannot be compiled in this image (it includes
 * <gsl/gsl_rng.h>, and GSL is absent), so the driver restates the
 * per-part orchestration of p4_setPramsPart (Pf/p4_tree.c:218-541) and the
 * root site sum of p4_partLogLikeLoop / ...RootLeaf
 * (Pf/p4_tree.c:1029-1378); the conditional likelihoods, P matrices,
 * eigensystems and gamma rates all come from the reference code itself.
 */
#include "pftypes.h"
#include "p4_node.h"
#include "p4_model.h"
#include "pmatrices.h"
#include "eig.h"
#include "gamma.h"
#include "util.h"
#include "defines.h"
#include "p4oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

int ref_discrete_gamma(double alpha, int K, double *rates, double *freqs)
{
    return DiscreteGamma(freqs, rates, alpha, alpha, K, 0);
}

double ref_point_chi2(double p, double v) { return PointChi2(p, v); }
double ref_incomplete_gamma(double x, double a, double lg) { return IncompleteGamma(x, a, lg); }
double ref_ln_gamma(double a) { return LnGamma(a); }
double ref_point_normal(double p) { return PointNormal(p); }

/* Q (normalised), and P(t) for each t in ts, via the reference eigen route. */
int ref_bigq_pmat(int dim, const double *bigR, const double *pi, int nT, const double *ts,
                  double *Qout, double *Pout)
{
    double **R = psdmatrix(dim), **Q = psdmatrix(dim), **P = psdmatrix(dim);
    double *piCopy = (double *)malloc(sizeof(double) * dim);
    eig *e;
    int i, j, k, ret;
    for (i = 0; i < dim; i++) {
        piCopy[i] = pi[i];
        for (j = 0; j < dim; j++) R[i][j] = bigR[i * dim + j];
    }
    setBigQFromRMatrixDotCharFreq(Q, R, piCopy, dim);
    normalizeBigQ(Q, piCopy, dim);
    for (i = 0; i < dim; i++)
        for (j = 0; j < dim; j++) Qout[i * dim + j] = Q[i][j];
    e = allocEig(dim, Q);
    ret = eigensystem(e);
    if (ret == 0) {
        for (k = 0; k < nT; k++) {
            matrixExpTimesBranchLength(e, ts[k], P);
            for (i = 0; i < dim; i++)
                for (j = 0; j < dim; j++) Pout[(size_t)k * dim * dim + i * dim + j] = P[i][j];
        }
    }
    freeEig(e);
    free_psdmatrix(R);
    free_psdmatrix(Q);
    free_psdmatrix(P);
    free(piCopy);
    return ret;
}

/* gdasrv shape values per gdasrv; rates are computed here by DiscreteGamma
 * exactly as p4_setPramsPart does for a free gdasrv. */
typedef struct {
    const double *alphas; /* [nGdasrvs] */
} ref_extra;

/*
 * Full single-part evaluation through the reference structs.
 *   tr, mp, dp: flat description (same as the oracle's); mp->gdasrvRates is
 *   ignored -- rates come from DiscreteGamma(alphas[g]).
 *   outP   : [nNodes][nCat][dim][dim]  (may be NULL)
 *   outCl  : [nNodes][nCat][dim][nPatterns] for internal nodes (may be NULL)
 *   outRates: [nGdasrvs][nCat]         (may be NULL)
 *   patternLikes: [nPatterns]          (may be NULL)
 * Returns lnL (or -1e99 as the reference does).
 */
typedef struct {
    p4_tree *tree;
    p4_model *model;
    part *aPart;
    int *bQET;
} ref_built;

/* Build the reference structs for one part, run p4_setPramsPart's
 * orchestration and the postorder conditional likelihoods. */
static ref_built ref_build(const or_tree *tr, const or_modelpart *mp, const or_part *dp, const double *alphas,
                           double *outRates);

double ref_part_loglike(const or_tree *tr, const or_modelpart *mp, const or_part *dp,
                        const double *alphas, double *outP, double *outCl, double *outRates,
                        double *patternLikes)
{
    int dim = mp->dim, K = mp->nCat, nP = dp->nPatterns, n, i, j, k, s;
    double lnL = 0.0, like, fTOMP;
    const double *pi;
    p4_node *root;
    ref_built b = ref_build(tr, mp, dp, alphas, outRates);
    p4_tree *aTree = b.tree;
    p4_modelPart *pm = b.model->parts[0];
    part *aPart = b.aPart;
    root = aTree->root;
    if (outP)
        for (n = 0; n < tr->nNodes; n++) {
            if (aTree->nodes[n] == root) continue;
            for (k = 0; k < K; k++)
                for (i = 0; i < dim; i++)
                    for (j = 0; j < dim; j++)
                        outP[(((size_t)n * K + k) * dim + i) * dim + j] = aTree->nodes[n]->bigPDecks[0][k][i][j];
        }
    if (outCl)
        for (n = 0; n < tr->nNodes; n++) {
            if (aTree->nodes[n]->cl == NULL) continue;
            for (k = 0; k < K; k++)
                for (i = 0; i < dim; i++)
                    for (j = 0; j < nP; j++)
                        outCl[(((size_t)n * K + k) * dim + i) * nP + j] = aTree->nodes[n]->cl[0][k][i][j];
        }

    /* root site sum (Pf/p4_tree.c:1029-1197 and 1199-1378) */
    pi = pm->comps[root->compNums[0]]->val;
    fTOMP = (1.0 - pm->pInvar->val[0]) / (double)K;
    for (j = 0; j < nP; j++) {
        like = 0.0;
        for (k = 0; k < K; k++) {
            if (root->isLeaf) {
                int code = aPart->patterns[root->seqNum][j];
                int isN = 0;
                if (code >= EQUATES_BASE && code < EQUATES_BASE + aPart->nEquates) {
                    isN = 1;
                    for (s = 0; s < dim; s++)
                        if (!aPart->equates[code - EQUATES_BASE][s]) isN = 0;
                }
                if (code >= 0)
                    like += pi[code] * root->cl[0][k][code][j];
                else if (code == GAP_CODE || code == QMARK_CODE || isN)
                    for (s = 0; s < dim; s++) like += pi[s] * root->cl[0][k][s][j];
                else
                    for (s = 0; s < dim; s++)
                        if (aPart->equates[code - EQUATES_BASE][s]) like += pi[s] * root->cl[0][k][s][j];
            } else {
                for (s = 0; s < dim; s++) like += pi[s] * root->cl[0][k][s][j];
            }
        }
        if (pm->pInvar->val[0]) {
            like *= fTOMP;
            if (aPart->globalInvarSitesVec[j] > 0)
                for (s = 0; s < dim; s++)
                    if (aPart->globalInvarSitesArray[s][j]) like += pi[s] * pm->pInvar->val[0];
        } else if (K > 1) {
            like = like / (double)K;
        }
        if (like <= 0.0) return -1.0e99;
        lnL = lnL + (aPart->patternCounts[j] * log(like));
        if (patternLikes) patternLikes[j] = like;
    }
    return lnL;
}

static ref_built ref_build(const or_tree *tr, const or_modelpart *mp, const or_part *dp, const double *alphas,
                           double *outRates)
{
    static int iOne = 1;
    static double PINVAR_MIN = 0.0, PINVAR_MAX = 0.99, KAPPA_MIN = 0.000001, KAPPA_MAX = 100.0,
                  GAMMA_SHAPE_MIN = 0.000001, GAMMA_SHAPE_MAX = 300.0, PIVEC_MIN = 1e-13,
                  PIVEC_MAX = 0.999, RATE_MIN = 1e-13, RATE_MAX = 0.9999999, RELRATE_MIN = 1e-8,
                  RELRATE_MAX = 1e8, BRLEN_MIN = 1e-8, BRLEN_MAX = 3.0;
    int dim = mp->dim, K = mp->nCat, nP = dp->nPatterns, n, i, j, k, s, c, r;
    p4_model *model;
    p4_modelPart *pm;
    int *bQET;
    data *aData;
    part *aPart;
    p4_tree *aTree;
    double partLikes[1];
    int nLeaves = 0;
    double lnL = 0.0, like, fTOMP;
    const double *pi;
    p4_node *root;

    model = p4_newModel(1, 0, 0, 0, 1, &iOne, &PINVAR_MIN, &PINVAR_MAX, &KAPPA_MIN, &KAPPA_MAX,
                        &GAMMA_SHAPE_MIN, &GAMMA_SHAPE_MAX, &PIVEC_MIN, &PIVEC_MAX, &RATE_MIN,
                        &RATE_MAX, &RELRATE_MIN, &RELRATE_MAX, &BRLEN_MIN, &BRLEN_MAX);
    bQET = (int *)malloc(sizeof(int) * mp->nComps * mp->nRMatrices);
    for (i = 0; i < mp->nComps * mp->nRMatrices; i++) bQET[i] = 1;
    p4_newModelPart(model, 0, dim, mp->nComps, mp->nRMatrices, mp->nGdasrvs, K, 0, bQET);
    pm = model->parts[0];
    for (c = 0; c < mp->nComps; c++) {
        p4_comp *cm = (p4_comp *)malloc(sizeof(p4_comp));
        cm->free = 1;
        cm->val = (double *)malloc(sizeof(double) * dim);
        memcpy(cm->val, mp->comps + (size_t)c * dim, sizeof(double) * dim);
        pm->comps[c] = cm;
    }
    for (r = 0; r < mp->nRMatrices; r++) {
        p4_newRMatrix(model, 0, r, 1, RMATRIX_SPECIFIED);
        for (i = 0; i < dim; i++)
            for (j = 0; j < dim; j++) pm->rMatrices[r]->bigR[i][j] = mp->bigR[((size_t)r * dim + i) * dim + j];
    }
    for (k = 0; k < mp->nGdasrvs; k++) {
        p4_gdasrv *g = (p4_gdasrv *)malloc(sizeof(p4_gdasrv));
        g->free = 1;
        g->nCat = K;
        g->val = (double *)malloc(sizeof(double));
        g->val[0] = alphas[k];
        g->freqs = (double *)malloc(sizeof(double) * K);
        g->rates = (double *)malloc(sizeof(double) * K);
        pm->gdasrvs[k] = g;
    }
    pm->pInvar->val[0] = mp->pInvar;
    pm->relRate[0] = mp->relRate;

    aData = (data *)calloc(1, sizeof(data));
    aData->nTax = dp->nTax;
    aData->nParts = 1;
    aData->parts = (part **)malloc(sizeof(part *));
    aPart = (part *)calloc(1, sizeof(part));
    aData->parts[0] = aPart;
    aPart->data = aData;
    aPart->dim = dim;
    aPart->nTax = dp->nTax;
    aPart->nChar = nP;
    aPart->nPatterns = nP;
    aPart->patterns = pimatrix(dp->nTax, nP);
    for (i = 0; i < dp->nTax; i++)
        for (j = 0; j < nP; j++) aPart->patterns[i][j] = dp->patterns[(size_t)i * nP + j];
    aPart->patternCounts = (int *)malloc(sizeof(int) * nP);
    memcpy(aPart->patternCounts, dp->patternCounts, sizeof(int) * nP);
    aPart->nEquates = dp->nEquates;
    if (dp->nEquates) {
        aPart->equates = pimatrix(dp->nEquates, dim);
        for (i = 0; i < dp->nEquates; i++)
            for (j = 0; j < dim; j++) aPart->equates[i][j] = dp->equates[i * dim + j];
    }
    aPart->globalInvarSitesVec = (int *)malloc(sizeof(int) * nP);
    memcpy(aPart->globalInvarSitesVec, dp->invarVec, sizeof(int) * nP);
    aPart->globalInvarSitesArray = pimatrix(dim, nP);
    for (i = 0; i < dim; i++)
        for (j = 0; j < nP; j++) aPart->globalInvarSitesArray[i][j] = dp->invarArray[(size_t)i * nP + j];

    for (n = 0; n < tr->nNodes; n++) nLeaves += tr->isLeaf[n] ? 1 : 0;
    aTree = (p4_tree *)calloc(1, sizeof(p4_tree));
    aTree->nNodes = tr->nNodes;
    aTree->nLeaves = nLeaves;
    aTree->nodes = (p4_node **)malloc(sizeof(p4_node *) * tr->nNodes);
    aTree->data = aData;
    aTree->model = model;
    aTree->nParts = 1;
    aTree->postOrder = (int *)tr->postOrder;
    aTree->partLikes = partLikes;
    for (n = 0; n < tr->nNodes; n++) p4_newNode(n, aTree, tr->seqNum[n], tr->isLeaf[n], 1);
    for (n = 0; n < tr->nNodes; n++) {
        p4_node *nd = aTree->nodes[n];
        nd->parent = tr->parent[n] >= 0 ? aTree->nodes[tr->parent[n]] : NULL;
        nd->leftChild = tr->leftChild[n] >= 0 ? aTree->nodes[tr->leftChild[n]] : NULL;
        nd->sibling = tr->sibling[n] >= 0 ? aTree->nodes[tr->sibling[n]] : NULL;
        nd->brLen[0] = tr->brLen[n];
        nd->compNums[0] = tr->compNums[n];
        nd->rMatrixNums[0] = tr->rMatrixNums[n];
        nd->gdasrvNums[0] = tr->gdasrvNums[n];
    }
    aTree->root = aTree->nodes[tr->root];
    root = aTree->root;

    /* p4_setPramsPart orchestration (Pf/p4_tree.c:236-258, 454-528) */
    if (K > 1)
        for (k = 0; k < mp->nGdasrvs; k++)
            DiscreteGamma(pm->gdasrvs[k]->freqs, pm->gdasrvs[k]->rates, pm->gdasrvs[k]->val[0],
                          pm->gdasrvs[k]->val[0], K, 0);
    if (outRates)
        for (k = 0; k < mp->nGdasrvs; k++)
            for (i = 0; i < K; i++) outRates[k * K + i] = pm->gdasrvs[k]->rates[i];
    for (n = 0; n < tr->nNodes; n++) {
        p4_node *nd = aTree->nodes[n];
        if (nd == root) continue;
        c = nd->compNums[0];
        r = nd->rMatrixNums[0];
        if (bQET[c * mp->nRMatrices + r]) p4_resetBQET(model, 0, c, r);
    }
    for (n = 0; n < tr->nNodes; n++)
        if (aTree->nodes[n] != root) p4_calculateBigPDecksPart(aTree->nodes[n], 0);

    /* postorder conditional likelihoods (Pf/p4_tree.c:875-887) */
    for (j = 0; j < tr->nNodes; j++) {
        int ni = tr->postOrder[j];
        p4_node *nd;
        if (ni == NO_ORDER) continue;
        nd = aTree->nodes[ni];
        if (!nd->isLeaf || (nd == root && nd->isLeaf)) p4_setConditionalLikelihoodsOfInternalNode(nd);
    }
    {
        ref_built b;
        b.tree = aTree;
        b.model = model;
        b.aPart = aPart;
        b.bQET = bQET;
        (void)lnL;
        (void)like;
        (void)fTOMP;
        (void)pi;
        (void)partLikes;
        return b;
    }
}


/*
 * Branch-length derivatives through the reference's own functions:
 *   p4_calculateBigPDecks_1stD / _2ndD      Pf/p4_node.c:442-540
 *   p4_initializeCL2ToRootComp, p4_setCL2Up,
 *   p4_setCL2Down                           Pf/p4_node.c:858-1018
 * p4_treeNewt.c cannot be compiled here (it includes p4_tree.h, which
 * includes <gsl/gsl_rng.h>), so the driver restates its p4_setNodeCL2
 * dispatch (Pf/p4_treeNewt.c:597-620) and the per-site derivative sums of
 * p4_newtNode (Pf/p4_treeNewt.c:257-500) around them.
 *   outCl2 : [nNodes][nCat][dim][nPatterns] for the nodes in wantCl2 (may be NULL)
 *   outP1, outP2 : [nNodes][nCat][dim][dim] (may be NULL)
 *   outLnL, outD1, outD2 : [nNodes], per non-root node: the part's log
 *   likelihood computed across that node's branch, and its first and
 *   second derivatives with respect to the node's branch length.
 */
int ref_part_derivs(const or_tree *tr, const or_modelpart *mp, const or_part *dp, const double *alphas,
                    const int *wantCl2, double *outCl2, double *outP1, double *outP2, double *outLnL,
                    double *outD1, double *outD2)
{
    int dim = mp->dim, K = mp->nCat, nP = dp->nPatterns, n, i, j, k, from, to, q;
    ref_built b = ref_build(tr, mp, dp, alphas, NULL);
    p4_tree *aTree = b.tree;
    p4_modelPart *pm = b.model->parts[0];
    part *aPart = b.aPart;
    p4_node *root = aTree->root;
    int *pre = (int *)malloc(sizeof(int) * tr->nNodes);
    int nPre = 0;
    for (n = 0; n < tr->nNodes; n++) {
        p4_node *nd = aTree->nodes[n];
        nd->cl2 = (double ****)malloc(sizeof(double ***));
        nd->cl2[0] = (double ***)malloc(K * sizeof(double **));
        nd->bigPDecks_1stD = (double ****)malloc(sizeof(double ***));
        nd->bigPDecks_1stD[0] = (double ***)malloc(K * sizeof(double **));
        nd->bigPDecks_2ndD = (double ****)malloc(sizeof(double ***));
        nd->bigPDecks_2ndD[0] = (double ***)malloc(K * sizeof(double **));
        for (k = 0; k < K; k++) {
            nd->cl2[0][k] = pdmatrix(dim, nP);
            nd->bigPDecks_1stD[0][k] = psdmatrix(dim);
            nd->bigPDecks_2ndD[0][k] = psdmatrix(dim);
        }
        if (nd != root) {
            p4_calculateBigPDecks_1stD(nd);
            p4_calculateBigPDecks_2ndD(nd);
        }
    }
    /* preorder: the reverse of the postorder */
    for (j = tr->nNodes - 1; j >= 0; j--)
        if (tr->postOrder[j] != NO_ORDER) pre[nPre++] = tr->postOrder[j];
    for (q = 0; q < nPre; q++) {
        p4_node *nd = aTree->nodes[pre[q]], *p;
        if (nd == root) continue;
        /* p4_setNodeCL2 */
        if (nd->parent == root)
            p4_initializeCL2ToRootComp(nd);
        else
            p4_setCL2Up(nd);
        for (p = nd->parent->leftChild; p; p = p->sibling)
            if (p != nd) p4_setCL2Down(nd, p);
    }
    for (n = 0; n < tr->nNodes; n++) {
        p4_node *nd = aTree->nodes[n];
        double lnL = 0.0, firstD = 0.0, secondD = 0.0;
        double fTOMP = (1.0 - pm->pInvar->val[0]) / (double)K;
        if (nd == root) {
            outLnL[n] = outD1[n] = outD2[n] = 0.0;
            continue;
        }
        for (q = 0; q < nP; q++) {
            double likeS = 0.0, firstS = 0.0, secondS = 0.0;
            for (k = 0; k < K; k++) {
                double like = 0.0, first = 0.0, second = 0.0, t2;
                double **P = nd->bigPDecks[0][k], **P1 = nd->bigPDecks_1stD[0][k], **P2 = nd->bigPDecks_2ndD[0][k];
                double **c2 = nd->cl2[0][k];
                if (nd->isLeaf) {
                    int code = aPart->patterns[nd->seqNum][q];
                    if (code >= 0) {
                        for (from = 0; from < dim; from++) {
                            t2 = c2[from][q];
                            like += t2 * P[from][code];
                            first += t2 * P1[from][code];
                            second += t2 * P2[from][code];
                        }
                    } else if (code == GAP_CODE || code == QMARK_CODE) {
                        for (from = 0; from < dim; from++) {
                            t2 = c2[from][q];
                            like += t2;
                            for (to = 0; to < dim; to++) {
                                first += t2 * P1[from][to];
                                second += t2 * P2[from][to];
                            }
                        }
                    } else {
                        int e = code - EQUATES_BASE, isN = 1;
                        for (from = 0; from < dim; from++)
                            if (!aPart->equates[e][from]) isN = 0;
                        for (from = 0; from < dim; from++) {
                            t2 = c2[from][q];
                            for (to = 0; to < dim; to++)
                                if (isN || aPart->equates[e][to]) {
                                    like += t2 * P[from][to];
                                    first += t2 * P1[from][to];
                                    second += t2 * P2[from][to];
                                }
                        }
                    }
                } else {
                    for (from = 0; from < dim; from++)
                        for (to = 0; to < dim; to++) {
                            t2 = c2[from][q] * nd->cl[0][k][to][q];
                            like += t2 * P[from][to];
                            first += t2 * P1[from][to];
                            second += t2 * P2[from][to];
                        }
                }
                likeS += like;
                firstS += first;
                secondS += second;
            }
            if (pm->pInvar->val[0]) {
                likeS *= fTOMP;
                firstS *= fTOMP;
                secondS *= fTOMP;
                if (aPart->globalInvarSitesVec[q] > 0)
                    for (from = 0; from < dim; from++)
                        if (aPart->globalInvarSitesArray[from][q])
                            likeS += pm->comps[root->compNums[0]]->val[from] * pm->pInvar->val[0];
            } else if (K > 1) {
                likeS /= (double)K;
                firstS /= (double)K;
                secondS /= (double)K;
            }
            if (likeS < 1.0e-300) {
                lnL += aPart->patternCounts[q] * -100000;
                firstD += aPart->patternCounts[q] * 1000000;
                secondD += aPart->patternCounts[q] * 10000000;
            } else {
                lnL += aPart->patternCounts[q] * log(likeS);
                firstD += aPart->patternCounts[q] * (firstS / likeS);
                secondD += aPart->patternCounts[q] * ((secondS * likeS - firstS * firstS) / (likeS * likeS));
            }
        }
        outLnL[n] = lnL;
        outD1[n] = firstD;
        outD2[n] = secondD;
    }
    for (n = 0; n < tr->nNodes; n++) {
        p4_node *nd = aTree->nodes[n];
        for (k = 0; k < K; k++) {
            if (outP1 && nd != root)
                for (i = 0; i < dim; i++)
                    for (j = 0; j < dim; j++) {
                        outP1[(((size_t)n * K + k) * dim + i) * dim + j] = nd->bigPDecks_1stD[0][k][i][j];
                        outP2[(((size_t)n * K + k) * dim + i) * dim + j] = nd->bigPDecks_2ndD[0][k][i][j];
                    }
            if (outCl2 && wantCl2 && wantCl2[n] && nd != root)
                for (i = 0; i < dim; i++)
                    for (j = 0; j < nP; j++)
                        outCl2[(((size_t)n * K + k) * dim + i) * nP + j] = nd->cl2[0][k][i][j];
        }
    }
    free(pre);
    return 0;
}
