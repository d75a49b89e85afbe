/*
 * p4oracle.h -- CPU FP64 restatement of p4's Pf likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the GPU
 * library; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  Nothing in the product links or calls it.
 *
 * The data model is flat (plain arrays, no pointer-chasing structs) but the
 * arithmetic and loop order follow the reference functions cited on each
 * declaration in p4oracle.c.
 */
#ifndef P4ORACLE_H
#define P4ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_GAP_CODE (-1)
#define OR_QMARK_CODE (-2)
#define OR_N_LIKE (-3)
#define OR_EQUATES_BASE (-64)
#define OR_NO_ORDER (-10000)

/* One data partition: coded site patterns (Pf/pftypes.h:31-53). */
typedef struct {
    int nTax;
    int dim;
    int nPatterns;
    const int *patterns;       /* [nTax][nPatterns], codes as Pf/defines.h:33-36 */
    const int *patternCounts;  /* [nPatterns] */
    int nEquates;
    const int *equates;        /* [nEquates][dim] 0/1 */
    const int *invarVec;       /* [nPatterns]   (Pf/part.c:716-848) */
    const int *invarArray;     /* [dim][nPatterns] */
} or_part;

/* One model partition (Pf/pftypes.h:169-186). */
typedef struct {
    int dim;
    int nCat;
    int nComps;
    const double *comps;       /* [nComps][dim] */
    int nRMatrices;
    const double *bigR;        /* [nRMatrices][dim][dim], symmetric */
    int nGdasrvs;
    const double *gdasrvRates; /* [nGdasrvs][nCat] mean category rates */
    double pInvar;
    double relRate;
} or_modelpart;

/* Tree topology + per-node model assignment (Pf/pftypes.h:116-147). */
typedef struct {
    int nNodes;
    int root;
    const int *parent;       /* -1 for none */
    const int *leftChild;
    const int *sibling;
    const int *isLeaf;
    const int *seqNum;
    const double *brLen;
    const int *postOrder;    /* [nNodes], NO_ORDER padded */
    const int *compNums;     /* [nNodes] for this part */
    const int *rMatrixNums;  /* [nNodes] */
    const int *gdasrvNums;   /* [nNodes] */
} or_tree;

/* Discrete gamma mean-category rates, Yang 1994. */
int or_discrete_gamma(double *freqK, double *rK, double alfa, double beta, int K, int median);
double or_point_chi2(double prob, double v);
double or_incomplete_gamma(double x, double alpha, double lnGammaAlpha);
double or_ln_gamma(double alpha);
double or_point_normal(double prob);

/* Q = R.diag(pi), normalised to mean rate 1. */
int or_bigq(int dim, const double *bigR, const double *pi, double *Q);

/* Eigensystem of a reversible Q: Q = V diag(lam) Vinv. */
int or_eigen_reversible(int dim, const double *Q, const double *pi,
                        double *V, double *Vinv, double *lam);

/* P(t) = V diag(exp(lam t)) Vinv. */
void or_pmat(int dim, const double *V, const double *Vinv, const double *lam,
             double t, double *P);

/* All P matrices for one part: P[node][cat][dim][dim]; root's slot untouched. */
int or_set_prams_part(const or_tree *tr, const or_modelpart *mp, double *P);

/* Conditional likelihoods of every internal node in postorder and the
 * site sum.  cl: [nNodes][nCat][dim][nPatterns] scratch supplied by caller.
 * patternLikes may be NULL.  Returns lnL or -1e99. */
double or_part_loglike(const or_tree *tr, const or_modelpart *mp, const or_part *dp,
                       const double *P, double *cl, double *patternLikes);

/* Convenience: set prams + log like with internal allocation. */
double or_tree_loglike_part(const or_tree *tr, const or_modelpart *mp, const or_part *dp,
                            double *patternLikes);

/* Data preparation (Pf/part.c). */
int or_make_patterns(int nTax, int nChar, const int *sequences,
                     int *patterns, int *patternCounts, int *seqPosPatternIndex);
void or_invar_sites(int nTax, int dim, int nPatterns, const int *patterns,
                    int nEquates, const int *equates, int *invarVec, int *invarArray);

#ifdef __cplusplus
}
#endif
#endif
