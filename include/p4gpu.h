/*
 * p4gpu.h -- C-ABI of libp4gpu.so, the MI355X likelihood engine behind
 * Tree.calcLogLike() and the Chain/Mcmc proposal loop.
 *
 * Each entry point replaces one function of p4's CPython extension `p4.pf`
 * (Pf/pfmodule.c, method table 2873-3031); the citation names the
 * pfmodule.c wrapper and the Pf C function it calls.  Handles are opaque
 * pointers (the reference passes C pointers as Python ints).  Arrays that the
 * reference aliases from numpy (comp values, gamma val/freqs/rates,
 * bQETneedsReset, partLikes, pre/postOrder, the var bounds) are passed here
 * as plain host pointers and are likewise kept, not copied: the library
 * re-reads them at every p4g_setPrams / p4g_treeLogLike / p4g_partLogLike /
 * p4g_calculateBigPDecks call, as the reference does.
 *
 * Error behaviour: where the reference prints and calls exit(1), these
 * entry points return a nonzero status (or NULL / NaN for value returns) and
 * leave a message for p4g_lastError().  Site likelihoods <= 0 give a part
 * log likelihood of -1.0e99, as in the reference (Pf/p4_tree.c:1134-1183).
 *
 * Threading: one handle is used by one host thread at a time.  All GPU work
 * of a tree is ordered on that tree's HIP stream.
 */
#ifndef P4GPU_H
#define P4GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct p4g_data p4g_data;
typedef struct p4g_part p4g_part;
typedef struct p4g_model p4g_model;
typedef struct p4g_tree p4g_tree;
typedef struct p4g_node p4g_node;
typedef struct p4g_gdasrv p4g_gdasrv;
typedef struct p4g_batch p4g_batch;

/* ---- status ---------------------------------------------------------- */
/* Last error message of the calling thread ("" if none).  No counterpart in
 * the reference, which exits instead of returning errors. */
const char *p4g_lastError(void);
/* 1 if a usable gfx9 GPU is present and the kernels load; 0 otherwise. */
int p4g_deviceAvailable(void);

/* ---- data (Pf/part.c, Pf/data.c) -------------------------------------- */
/* pf.newData -> newData (Pf/pfmodule.c:29-40, Pf/data.c:7) */
p4g_data *p4g_newData(int nTax, int nParts);
/* pf.freeData (Pf/pfmodule.c:43-55) */
void p4g_freeData(p4g_data *d);
/* pf.pokePartInData (Pf/pfmodule.c:73-87) */
int p4g_pokePartInData(p4g_part *p, p4g_data *d, int partNum);
/* pf.newPart -> newPart (Pf/pfmodule.c:112-131, Pf/part.c:15-78) */
p4g_part *p4g_newPart(int nTax, int nChar, const char *equateSymbols, int nEquates,
                      const char *symbols, int dim);
/* pf.freePart (Pf/pfmodule.c:134-148) */
void p4g_freePart(p4g_part *p);
/* pf.pokeEquatesTable -> pokeEquatesTable (Pf/pfmodule.c:151-166, Pf/part.c:279-315):
 * nEquates*dim characters '0'/'1', row per equate symbol. */
int p4g_pokeEquatesTable(p4g_part *p, const char *table);
/* pf.pokeSequences -> pokeSequences (Pf/pfmodule.c:169-183, Pf/part.c:127-275):
 * nTax*nChar characters, taxon-major. */
int p4g_pokeSequences(p4g_part *p, const char *seqs);
/* pf.makePatterns -> makePatterns (Pf/pfmodule.c:186-197, Pf/part.c:317-448).
 * Same first-occurrence order and counts; hashed instead of quadratic. */
int p4g_makePatterns(p4g_part *p);
/* makePatterns on the GPU (no separate counterpart: the device form of
 * Pf/part.c:317-448): the same patterns, counts, first-occurrence order and
 * site map as p4g_makePatterns, by a hash of every site column, a radix
 * sort, a check of every column against its run's first column (a collision
 * retries with another hash seed) and scans.  Reads the part's device
 * sequences when a simulation left them there. */
int p4g_makePatternsDevice(p4g_part *p);
/* pf.setGlobalInvarSitesVec (Pf/pfmodule.c:310-320, Pf/part.c:716-848) */
int p4g_setGlobalInvarSitesVec(p4g_part *p);
/* Accessors of part fields the reference reads through its struct
 * (part->nPatterns, patternCounts, sequencePositionPatternIndex, patterns,
 * globalInvarSitesVec/Array; Pf/pftypes.h:31-53). */
int p4g_partNPatterns(const p4g_part *p);
int p4g_partNChar(const p4g_part *p);
int p4g_partNTax(const p4g_part *p);
int p4g_partDim(const p4g_part *p);
/* Copies out: patterns [nTax][nPatterns], counts [nPatterns],
 * seqPosPatternIndex [nChar], invarVec [nPatterns], invarArray [dim][nPatterns].
 * Any pointer may be NULL. */
int p4g_partGetPatterns(const p4g_part *p, int *patterns, int *counts, int *seqPosPatternIndex,
                        int *invarVec, int *invarArray);
/* Directly set coded patterns (codes as Pf/defines.h:33-36) instead of
 * pokeSequences+makePatterns; used for pre-compressed input. */
int p4g_partSetPatterns(p4g_part *p, int nPatterns, const int *patterns, const int *counts);
/* pf.getSiteLikes (Pf/pfmodule.c:378-394): part->siteLikes after a
 * p4g_treeLogLike(tree, getSiteLikes=1).  out: [nChar]. */
int p4g_getSiteLikes(const p4g_part *p, double *out);

/* ---- model (Pf/p4_model.c) -------------------------------------------- */
/* pf.p4_newModel (Pf/pfmodule.c:1484-1550, Pf/p4_model.c:12-51).  The 15
 * pointers are the var bounds arrays, aliased. */
p4g_model *p4g_newModel(int nParts, int doRelRates, int relRatesAreFree, int nFreePrams,
                        int isHet, int *rMatrixNormalizeTo1, double *PINVAR_MIN,
                        double *PINVAR_MAX, double *KAPPA_MIN, double *KAPPA_MAX,
                        double *GAMMA_SHAPE_MIN, double *GAMMA_SHAPE_MAX, double *PIVEC_MIN,
                        double *PIVEC_MAX, double *RATE_MIN, double *RATE_MAX,
                        double *RELRATE_MIN, double *RELRATE_MAX, double *BRLEN_MIN,
                        double *BRLEN_MAX);
/* pf.p4_freeModel (Pf/pfmodule.c:1554-1571) */
void p4g_freeModel(p4g_model *m);
/* pf.p4_newModelPart (Pf/pfmodule.c:1590-1610, Pf/p4_model.c:90-209);
 * bQETneedsReset: aliased int32 [nComps][nRMatrices]. */
int p4g_newModelPart(p4g_model *m, int pNum, int dim, int nComps, int nRMatrices,
                     int nGdasrvs, int nCat, int pInvarFree, int *bQETneedsReset);
/* pf.p4_newComp (Pf/pfmodule.c:1634-1650, Pf/p4_model.c:328-342); val aliased [dim]. */
int p4g_newComp(p4g_model *m, int pNum, int mNum, int free, double *val);
/* pf.p4_newRMatrix (Pf/pfmodule.c:1653-1669, Pf/p4_model.c:345-452).
 * spec: 100 ones, 20 specified, 5 two-parameter; empirical protein specs
 * (101-120) are not built in and are rejected. */
int p4g_newRMatrix(p4g_model *m, int pNum, int mNum, int free, int spec);
/* pf.p4_newGdasrv (Pf/pfmodule.c:1672-1701, Pf/p4_model.c:490-511); arrays aliased. */
p4g_gdasrv *p4g_newGdasrv(p4g_model *m, int pNum, int mNum, int nCat, int free, double *val,
                          double *freqs, double *rates);
/* pf.p4_setRMatrixBigR (Pf/pfmodule.c:1720-1738): bigR[i][j] = bigR[j][i] = val */
int p4g_setRMatrixBigR(p4g_model *m, int pNum, int mNum, int i, int j, double val);
/* pf.p4_setKappa (Pf/pfmodule.c:1741-1757) */
int p4g_setKappa(p4g_model *m, int pNum, int mNum, double val);
/* pf.gdasrvCalcRates (Pf/pfmodule.c:1813-1833): DiscreteGamma into the
 * aliased freqs/rates of the gdasrv. */
int p4g_gdasrvCalcRates(p4g_gdasrv *g);
/* pf.gdasrvCalcRates_np (Pf/pfmodule.c:1836-1865) */
int p4g_gdasrvCalcRates_np(int nCat, double val, double *freqs, double *rates);
/* pf.p4_setPInvarVal (Pf/pfmodule.c:1871-1884) */
int p4g_setPInvarVal(p4g_model *m, int pNum, double val);
/* pf.p4_setRelRateVal (Pf/pfmodule.c:1887-1903) */
int p4g_setRelRateVal(p4g_model *m, int pNum, double val);
/* pf.p4_resetBQET (Pf/pfmodule.c:1613-1631, Pf/p4_model.c:513-543): marks the
 * (comp, rMatrix) eigensystem for recomputation at the next P build. */
int p4g_resetBQET(p4g_model *m, int pNum, int compNum, int rMatrixNum);

/* ---- tree and nodes (Pf/p4_tree.c, Pf/p4_node.c) ---------------------- */
/* pf.p4_newTree (Pf/pfmodule.c:1383-1406, Pf/p4_tree.c:15-95).  preOrder,
 * postOrder (int32 [nNodes], NO_ORDER = -10000 padded) and partLikes
 * (double [nParts]) are aliased.  The reference also takes
 * newtAndBrentPowellOptPassLimit, used only by ML optimisation. */
p4g_tree *p4g_newTree(int nNodes, int nLeaves, int *preOrder, int *postOrder,
                      double *partLikes, p4g_data *data, p4g_model *model);
/* pf.p4_freeTree (Pf/pfmodule.c:1410-1421) */
void p4g_freeTree(p4g_tree *t);
/* pf.p4_newNode (Pf/pfmodule.c:1443-1466, Pf/p4_node.c:10-136) */
p4g_node *p4g_newNode(int nodeNum, p4g_tree *t, int seqNum, int isLeaf, int inTree);
/* pf.p4_freeNode (Pf/pfmodule.c:1460-1481) */
void p4g_freeNode(p4g_node *n);
/* pf.p4_setNodeRelation (Pf/pfmodule.c:1907-1940): relation 0 parent,
 * 1 leftChild, 2 sibling; relNum -1 for none. */
int p4g_setNodeRelation(p4g_node *n, int relation, int relNum);
/* pf.p4_setTreeRoot (Pf/pfmodule.c:1944-1957) */
int p4g_setTreeRoot(p4g_tree *t, p4g_node *n);
/* pf.p4_setBrLen (Pf/pfmodule.c:1960-1975) */
int p4g_setBrLen(p4g_node *n, double len);
/* pf.p4_getBrLen / p4_getTreeLen (Pf/pfmodule.c:1978-2010) */
double p4g_getBrLen(const p4g_node *n);
double p4g_getTreeLen(const p4g_tree *t);
/* pf.p4_setCompNum / p4_setRMatrixNum / p4_setGdasrvNum (Pf/pfmodule.c:2013-2058) */
int p4g_setCompNum(p4g_node *n, int pNum, int val);
int p4g_setRMatrixNum(p4g_node *n, int pNum, int val);
int p4g_setGdasrvNum(p4g_node *n, int pNum, int val);

/* Precision of the partials of one tree: 0 = FP64 without scaling (the
 * reference's numerics), 1 = FP32 with per-node per-pattern power-of-two
 * scale exponents.  No counterpart in the reference (FP64 only). */
int p4g_setPrecision(p4g_tree *t, int fp32);
int p4g_getPrecision(const p4g_tree *t);

/* ---- evaluation ------------------------------------------------------- */
/* pf.p4_setPrams (Pf/pfmodule.c:2086-2104, Pf/p4_tree.c:207-541):
 * gamma rates of free gdasrvs, Q + eigensystem for every (comp, rMatrix)
 * used, then P for every non-root node, all on the GPU.  pNum -1 = all. */
int p4g_setPrams(p4g_tree *t, int pNum);
/* pf.p4_treeLogLike (Pf/pfmodule.c:2132-2144, Pf/p4_tree.c:868-922). */
int p4g_treeLogLike(p4g_tree *t, int getSiteLikes, double *lnL);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving or MCMC loops,
 * harnesses that forbid hipMalloc): p4g_treeLogLike(t, 0, lnL) computed with
 * no allocation.  The pruning kernels' reduction scratch (per-workgroup sums,
 * bad-site flags, the ticket counter) comes from one library-wide device
 * array of 512 KiB per GPU instead of the tree's own buffers; a part needing
 * more (more than about 43,000 pattern tiles, some 700,000 patterns) is
 * rejected with an error.  The tree must already hold device state for its
 * current precision (one earlier p4g_treeLogLike, which allocates);
 * otherwise the call fails and allocates nothing.  Calls that overlap on
 * one GPU race on the shared array. */
int p4g_treeLogLikeNoAlloc(p4g_tree *t, double *lnL);
/* pf.p4_partLogLike (Pf/pfmodule.c:2147-2160, Pf/p4_tree.c:924-1027).
 * Recomputes the partials of nodes marked by
 * p4g_setConditionalLikelihoodsOfInternalNodePart (and of all internal
 * nodes on the first call), then the site sum of part pNum. */
int p4g_partLogLike(p4g_tree *t, p4g_part *p, int pNum, int getSiteLikes, double *lnL);
/* pf.p4_calculateBigPDecks (Pf/pfmodule.c:2165-2177, Pf/p4_node.c:288-294) */
int p4g_calculateBigPDecks(p4g_node *n);
/* pf.p4_calculateBigPDecksPart (Pf/p4_node.c:296-439) */
int p4g_calculateBigPDecksPart(p4g_node *n, int pNum);
/* pf.p4_calculateAllBigPDecksAllParts (Pf/pfmodule.c:2180-2192, Pf/p4_tree.c:815-830) */
int p4g_calculateAllBigPDecksAllParts(p4g_tree *t);
/* pf.p4_setConditionalLikelihoodsOfInternalNodePart (Pf/pfmodule.c:2195-2210,
 * Pf/p4_node.c:636-857).  Deferred: marks the node; the partials of all
 * marked nodes of the part are recomputed, in postorder, in one launch at
 * the next p4g_partLogLike.  Results equal eager evaluation. */
int p4g_setConditionalLikelihoodsOfInternalNodePart(p4g_node *n, int pNum);
/* pf.p4_setConditionalLikelihoodsOfInternalNode (Pf/p4_node.c:601-634) */
int p4g_setConditionalLikelihoodsOfInternalNode(p4g_node *n);

/* ---- MCMC state sync (Pf/p4_treeCopyVerify.c) ------------------------- */
/* pf.p4_copyCondLikes (Pf/pfmodule.c:2445-2459, Pf/p4_treeCopyVerify.c:7-31):
 * device-to-device copy of the internal partials (and scale exponents)
 * of tree a into tree b; doAll=0 copies only nodes with clNeedsUpdating. */
int p4g_copyCondLikes(p4g_tree *a, p4g_tree *b, int doAll);
/* pf.p4_copyBigPDecks (Pf/pfmodule.c:2462-2476, Pf/p4_treeCopyVerify.c:33-61) */
int p4g_copyBigPDecks(p4g_tree *a, p4g_tree *b, int doAll);
/* pf.p4_copyModelPrams (Pf/pfmodule.c:2479-2492, Pf/p4_treeCopyVerify.c:63-180) */
int p4g_copyModelPrams(p4g_tree *a, p4g_tree *b);
/* pf.p4_verifyIdentityOfTwoTrees (Pf/pfmodule.c:2431-2442,
 * Pf/p4_treeCopyVerify.c:187-676): 0 = same (partials and P within 1e-15),
 * 1 = different. */
int p4g_verifyIdentityOfTwoTrees(p4g_tree *a, p4g_tree *b);

/* ---- inspection (tests) ------------------------------------------------ */
/* Copy out P of node n, part pNum: [nCat][dim][dim] FP64.  The reference
 * exposes these through the node struct (Pf/pftypes.h:133). */
int p4g_getBigPDeck(const p4g_node *n, int pNum, double *out);
/* Copy out the partials of internal node n, part pNum, as
 * [nCat][dim][nPatterns] FP64 (the reference cl layout, Pf/p4_node.c:97-106),
 * unscaled: FP32 values are multiplied by 2^-exponent. */
int p4g_getCondLikes(const p4g_node *n, int pNum, double *out);

/* ---- branch-length derivatives (FP64 trees) ---------------------------- */
/* For every non-root node n: the log likelihood computed across n's branch
 * (equal to the tree's for every n), and its first and second derivatives
 * with respect to n's branch length, summed over part pNum or all parts
 * (pNum -1); the root's entries are 0.  Arrays [nNodes].  This is what the
 * reference's Newton optimiser computes one node at a time:
 * p4_calculateBigPDecks_1stD / _2ndD (Pf/p4_node.c:442-540, reached from
 * pf.p4_newtSetup / p4_newtAndBrentPowellOpt, Pf/pfmodule.c:2259-2285),
 * p4_setNodeCL2 with p4_initializeCL2ToRootComp / p4_setCL2Up /
 * p4_setCL2Down (Pf/p4_treeNewt.c:597-620, Pf/p4_node.c:858-1018) and the
 * per-site sums of p4_newtNode (Pf/p4_treeNewt.c:257-500); here for every
 * branch in one launch per part.  Runs a full evaluation first. */
int p4g_branchDerivatives(p4g_tree *t, int pNum, double *lnLEdge, double *d1, double *d2);
/* The up-pass partials of non-root node n (node->cl2 of the reference,
 * Pf/pftypes.h:134) as [nCat][dim][nPatterns], after p4g_branchDerivatives. */
int p4g_getCondLikes2(const p4g_node *n, int pNum, double *out);
/* dP/dt and d2P/dt2 of node n's branch (node->bigPDecks_1stD / _2ndD,
 * Pf/pftypes.h:131-132) as [nCat][dim][dim], after p4g_branchDerivatives. */
int p4g_getBigPDeck1stD(const p4g_node *n, int pNum, double *out);
int p4g_getBigPDeck2ndD(const p4g_node *n, int pNum, double *out);

/* ---- MCMC state sync without whole-tree copies (no single counterpart:
 * replaces the per-node marking loops of p4/chain.py's proposals, through
 * p4_setConditionalLikelihoodsOfInternalNodePart, and the doAll copies of
 * Pf/p4_treeCopyVerify.c:7-61 in Chain's accept/reject step,
 * p4/chain.py:1351-1531).
 * p4g_markPath: node nodeNum and its ancestors (internal nodes and a leaf
 * root) marked for recomputation in part pNum, or every part (-1).
 * p4g_markAll: every internal node (and a leaf root) of part pNum or all. */
int p4g_markPath(p4g_tree *t, int nodeNum, int pNum);
int p4g_markAll(p4g_tree *t, int pNum);
/* dst takes src's partial slots, scale exponents and P matrices wherever
 * either tree changed them since the two were last synced (slots written
 * by an evaluation, P blocks of marked nodes; everything after buffers were
 * made or re-laid out), for every part in one copy launch; dirty marks,
 * part likelihoods and lnL follow.  Model values and eigensystems are
 * p4g_copyModelPrams' job.  The copy is queued on the library's stream. */
int p4g_syncChanged(p4g_tree *src, p4g_tree *dst);

/* ---- CPO / LPML (the doCpo step of Mcmc's sampling, p4/mcmc.py:3286-3297,
 * and Mcmc.lpml, p4/mcmc.py:3730-3748).  p4g_cpoAccumulate evaluates tree t
 * fully, as p4_treeLogLike(t, 1) (Pf/p4_tree.c:868-1021) does, keeps each
 * pattern's likelihood in GPU memory and adds 1/like to every part's
 * device-resident per-pattern sum; each part's sample count goes up by
 * one.  lnL: the tree's log likelihood.  An error if a site likelihood is
 * zero (the reference's 1./sL raises there).  The sums belong to the data
 * parts, so every chain's tree adds to the same sums. */
int p4g_cpoAccumulate(p4g_tree *t, double *lnL);
/* cpo_sumsOfInverseSiteLikes for one part, by site: out [p4g_partNSites];
 * nSamples: the number of accumulations. */
int p4g_cpoGetSums(const p4g_part *p, double *out, int *nSamples);
int p4g_cpoReset(p4g_part *p);
/* Sites of a part: nChar, or the pattern count when the patterns were set
 * without a site map. */
int p4g_partNSites(const p4g_part *p);

/* ---- ancestral states of the root: pf.p4_drawAncState (Pf/pfmodule.c:
 * 2353-2367 -> p4_drawAncStateP, Pf/p4_treeSim.c:591-857) for the sites
 * [first, first + n) of part pNum in one launch.  Each site draws a
 * (gamma category, state) or an invariant state from the table the
 * reference builds from root->cl: comp[s] * cl[cat][s] * (1 - pInvar)/nCat,
 * category-major, then comp[s] * pInvar for the site's invariant states;
 * the pick is the first entry whose running sum exceeds u * total.
 * u: n uniforms in [0, 1) (the reference's random() / (RAND_MAX + 1.0)),
 * or NULL: u from Philox4x64-10 keyed by (seed, site).
 * draw: [n][4] = chStNum, catNum, isInvar, invarChNum (Pf/p4_treeSim.c:
 * 821-840).  The partials are brought up to date first. */
int p4g_drawAncStates(p4g_tree *t, int pNum, int first, int n, const double *u, unsigned long long seed,
                      int *draw);
/* ---- simulation and posterior-predictive summaries.
 * p4g_simulate: pf.p4_simulate (Pf/p4_treeSim.c:14-590) -- new sequences
 * for every part of t's data, one GPU thread per site: the site's gamma
 * category (floor(nCat * u)), root state (running sums of the root comp,
 * last entry 1.0) and invariant flag (u < pInvar), or with refTree those
 * three from the root draws of refTree's data (p4g_drawAncStates); then
 * each non-root node in preorder takes its parent's state through the
 * running sums of its P row (the picker decks of Pf/p4_node.c:545-586);
 * invariant sites keep the root state.  P comes from the current model and
 * branch lengths.  Uniforms: Philox4x64-10 keyed by seed, counted by (site,
 * node, part) -- the same draws for a seed whatever the launch shape (the
 * reference's GSL stream is not reproduced).  Afterwards each part holds
 * the leaves' sequences and no patterns, as after the reference's call
 * (make them again with p4g_makePatterns). */
int p4g_simulate(p4g_tree *t, p4g_tree *refTree, unsigned long long seed);
/* pf.symbolSequences (Pf/part.c:604-680): the part's sequences as symbols,
 * out [nTax * nChar] (no terminator). */
int p4g_symbolSequences(const p4g_part *p, char *out);
/* The last simulation's gamma category of each site, -1 where the site was
 * invariant (part->simCats and globalInvarSitesVec of the reference). */
int p4g_partSimCats(const p4g_part *p, int *cats);
/* Sequence summaries on the GPU (Pf/part.c:1419-1571): compCounts
 * [nTax][dim] per-taxon state counts, nConstant (partSimpleConstantSites-
 * Count), meanNChars (partMeanNCharsPerSite), nNonState: codes that are
 * gaps or ambiguities. */
int p4g_partSeqStats(p4g_part *p, int *compCounts, int *nConstant, double *meanNChars, int *nNonState);
/* pf.partBigXSquared (Pf/part.c:1490-1571): -2.0 if the part has gaps or
 * ambiguities; -3.0 on an error. */
double p4g_partBigXSquared(p4g_part *p);
/* pf.getUnconstrainedLogLike (Pf/part.c:682-714) from the part's
 * patterns. */
double p4g_getUnconstrainedLogLike(const p4g_part *p);
/* The counter-based generator of the GPU draws (no counterpart in the
 * reference, whose draws use random() and GSL): Philox4x64-10 of counter[4]
 * under key[2], on the host, from the same source as the device code. */
void p4g_philox4x64(const unsigned long long *counter, const unsigned long long *key, unsigned long long *out);

/* ---- batched evaluation (no single counterpart: replaces the sequential
 * per-chain, per-part p4_partLogLike calls of Mcmc.run, p4/mcmc.py:2861-2950) */
/* A batch evaluates the pending partials and site sums of several trees
 * (all parts of each) in one launch.  lnL: [nTrees] (sum over parts);
 * each tree's partLikes array is written too. */
p4g_batch *p4g_newBatch(int nTrees, p4g_tree **trees);
void p4g_freeBatch(p4g_batch *b);
int p4g_batchLogLike(p4g_batch *b, double *lnL);
/* Kernel timing (no counterpart in the reference): when on, every pruning
 * launch is bracketed by HIP events on the library's stream, and
 * p4g_lastKernelMicroseconds returns the device time of the last
 * evaluation's pruning launches of tree t (summed over parts), -1 if none. */
void p4g_setKernelTiming(int on);
/* Tuning of the 4-state pruning kernel (no counterpart in the reference;
 * set by the autotuner, p4gpu/autotune.py), for FP64 (fp32 = 0) or FP32
 * partials: opsPerGroup 0, -1 or -2 selects the wave-per-category kernel
 * (the default) with its P^T read by vector loads from a program-order
 * stream (0), by scalar loads (-1), or by scalar loads with the tree's
 * subtrees split between two sets of waves (-2); 1 or 4 the grouped kernel
 * with that many independent
 * binary ops per group, with the program's P rows staged in LDS if
 * stagePInLds.  Partials already computed are re-laid out (recomputed) at
 * the next evaluation when the choice changes the layout. */
void p4g_setPruneTuning(int fp32, int opsPerGroup, int stagePInLds);
/* How P matrices are computed (no counterpart in the reference, which
 * always goes through the eigensystem, Pf/eig.c:163-223): direct = 1 (the
 * default) evaluates exp(Q t) per branch and category from the (pi, R) of
 * its eigen slot by a scaled Taylor series of expm1, with no eigensystem,
 * for 4-state parts always and for 20-state parts when at least 16 P
 * matrices are pending (fewer reuse cached eigensystems); direct = 0 uses
 * the eigensystem (computed when a slot changes) and
 * P = I + V diag(expm1(lam t)) V^-1.  Other state counts always use the
 * eigensystem.  Takes effect at the next flush of pending P matrices;
 * eigensystems skipped meanwhile are recomputed when needed. */
void p4g_setPmatRoute(int direct);
/* Layout of the 20-state FP32 (MFMA) pruning launch (no counterpart in the
 * reference): with categoryWaves = 1 (the default) a program of at least
 * minOps node ops (default 32: full evaluations, not short dirty paths)
 * runs one wavefront per (16-pattern tile, gamma category), with one
 * barrier per op; otherwise one wavefront per tile holds every category.
 * Partials are the same either way. */
void p4g_setAaTuning(int categoryWaves, int minOps);
/* LDS cap for the 4-state program's preloaded clean children (no
 * counterpart in the reference): a dirty-path program preloads the clean
 * children of its ops into LDS slots at launch; with bytes > 0 the
 * workgroup's LDS stays at or below bytes (further clean children are read
 * from HBM when used), keeping the occupancy of full programs; 0 (the
 * default) lets preloads fill the LDS budget. */
void p4g_setWavePreloadLds(int bytes);
/* LDS slot file of the 20-state category-wave kernel (no counterpart in the
 * reference): its programs run in Sethi-Ullman order and forward up to
 * maxSlots children per workgroup through LDS instead of re-reading them
 * from HBM (default 5; 0 turns forwarding off).  Same partials either way. */
void p4g_setAaLdsSlots(int maxSlots);
double p4g_lastKernelMicroseconds(const p4g_tree *t);

/* ---- learned proposal network (no counterpart in the reference, whose
 * proposals are all hand-designed, p4/mcmc.py:1600-2200).  A 3-layer MLP
 * nIn -> nHidden -> nHidden -> nOut (ReLU) with BF16 weights on the MFMA
 * units and FP32 accumulation / master weights / Adam.  nOut = 2D: columns
 * [0, D) are Gaussian means, [D, 2D) log standard deviations.
 * x: [n][nIn] row-major; y, mask: [n][D]; out: [n][nOut]; grads: the
 * parameter vector layout W1 [H][F], b1 [H], W2 [H][H], b2 [H], W3 [O][H],
 * b3 [O].  loss = sum_i sum_d mask (s + (y - mu)^2 e^{-2s} / 2) over the n
 * rows (mu, s: the row's mean and log sd); grads = d(loss / n)/dparams. */
typedef struct p4g_mlp p4g_mlp;
p4g_mlp *p4g_mlpNew(int nIn, int nHidden, int nOut, unsigned long long seed);
void p4g_mlpFree(p4g_mlp *m);
int p4g_mlpNumParams(const p4g_mlp *m);
int p4g_mlpGetParams(p4g_mlp *m, float *out);
int p4g_mlpSetParams(p4g_mlp *m, const float *in);
int p4g_mlpForward(p4g_mlp *m, const float *x, int n, float *out);
/* Loss and gradients without an update. */
int p4g_mlpGradients(p4g_mlp *m, const float *x, const float *y, const float *mask, int n, float *grads,
                     float *loss);
/* One Adam step (b1 0.9, b2 0.999, eps 1e-8) on grads at learning rate lr;
 * loss is the batch loss before the step. */
int p4g_mlpTrainStep(p4g_mlp *m, const float *x, const float *y, const float *mask, int n, float lr,
                     float *loss);

#ifdef __cplusplus
}
#endif
#endif
