// p4g_internal.h -- host-side state of libp4gpu and the argument blocks of
// its HIP kernels.  Not part of the public C-ABI (include/p4gpu.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace p4g {

// Character codes of coded sequences (Pf/defines.h:33-36).
constexpr int GAP_CODE = -1;
constexpr int QMARK_CODE = -2;
constexpr int N_LIKE = -3;
constexpr int EQUATES_BASE = -64;
constexpr int NO_ORDER = -10000;

// rMatrix spec ids (Pf/defines.h:1-30).
constexpr int RMATRIX_2P = 5;
constexpr int RMATRIX_SPECIFIED = 20;
constexpr int RMATRIX_POISSON = 100;

// Device tip-code alphabet: 0..dim-1 a state, TIP_ALL every state (gap, '?',
// or an equate allowing every state), dim..TIP_ALL-1 a partial ambiguity.
constexpr unsigned TIP_ALL = 255;
constexpr int MAX_DIM = 64;
constexpr int MAX_CHILDREN = 64;

// Tree precision modes.
enum Precision { FP64 = 0, FP32 = 1 };

void setError(const std::string &msg);
void clearError();
bool hipOk(hipError_t e, const char *what);
hipStream_t stream();
bool deviceReady();

// ------------------------------------------------------------------ data
struct Part;
struct Data {
    int nTax = 0;
    int nParts = 0;
    std::vector<Part *> parts;
};

// Device copy of one data part's patterns, laid out for the pruning kernel.
struct DevPart {
    bool valid = false;
    int nPatPad = 0;       // padded to a multiple of 64 patterns
    uint8_t *tips = nullptr;      // [nTax][nPatPad] device tip codes
    double *counts = nullptr;     // [nPatPad] pattern weights (0 for padding)
    uint32_t *invMask = nullptr;  // [nWords][nPatPad] states compatible with every taxon
    uint64_t *ambMask = nullptr;  // [nAmb] state masks of partial ambiguities
    int *sitePat = nullptr;       // [nSites] pattern of each site
    int nSites = 0;
    int nAmb = 0;
    int nWords = 0;
    std::vector<uint64_t> hostAmb;
};

struct Part {
    Data *data = nullptr;
    int dim = 0, nTax = 0, nChar = 0, nEquates = 0;
    std::string symbols, equateSymbols;
    std::vector<int> equates;       // [nEquates][dim]
    std::vector<int> sequences;     // [nTax][nChar]
    int nPatterns = 0;
    std::vector<int> patterns;      // [nTax][nPatterns]
    std::vector<int> patternCounts; // [nPatterns]
    std::vector<int> seqPosPatternIndex; // [nChar]
    std::vector<int> invarVec;      // [nPatterns]
    std::vector<int> invarArray;    // [dim][nPatterns]
    bool invarSet = false;
    std::vector<double> siteLikes;  // [nChar]
    DevPart dev;
    // CPO: per pattern, the sum over samples of 1/like, in GPU memory
    double *cpoSum = nullptr;
    int cpoPatterns = 0, cpoSamples = 0;
    // device copy of `sequences` ([nTax][nChar] codes), written by the
    // simulation kernel or uploaded for the sequence statistics
    int8_t *dSeq = nullptr;
    size_t dSeqCap = 0;             // bytes, a multiple of 4
    bool dSeqValid = false;
    std::vector<int> simCats;       // [nChar] after a simulation: category, -1 invariant
    ~Part();
};

// ----------------------------------------------------------------- model
struct Comp {
    int free = 0;
    double *val = nullptr;  // aliased host array [dim]
};
struct RMatrix {
    int free = 0, spec = RMATRIX_POISSON;
    std::vector<double> bigR;  // [dim][dim]
    double kappa = 2.0;
};
struct Gdasrv {
    int free = 0, nCat = 1;
    double *val = nullptr, *freqs = nullptr, *rates = nullptr;  // aliased
};
struct ModelPart {
    int dim = 0, nComps = 0, nRMatrices = 0, nGdasrvs = 0, nCat = 1, pInvarFree = 0;
    std::vector<Comp> comps;
    std::vector<RMatrix> rMatrices;
    std::vector<Gdasrv *> gdasrvs;
    double pInvar = 0.0, relRate = 1.0;
    int *bQETneedsReset = nullptr;  // aliased [nComps][nRMatrices]
    ~ModelPart();
};
struct Model {
    int nParts = 0, doRelRates = 0, relRatesAreFree = 0, nFreePrams = 0, isHet = 0;
    int *rMatrixNormalizeTo1 = nullptr;
    double *PINVAR_MIN = nullptr, *PINVAR_MAX = nullptr, *KAPPA_MIN = nullptr,
           *KAPPA_MAX = nullptr, *GAMMA_SHAPE_MIN = nullptr, *GAMMA_SHAPE_MAX = nullptr,
           *PIVEC_MIN = nullptr, *PIVEC_MAX = nullptr, *RATE_MIN = nullptr, *RATE_MAX = nullptr,
           *RELRATE_MIN = nullptr, *RELRATE_MAX = nullptr, *BRLEN_MIN = nullptr,
           *BRLEN_MAX = nullptr;
    std::vector<ModelPart *> parts;
    ~Model();
};

// ------------------------------------------------------------------ tree
struct Tree;
struct Node {
    int nodeNum = 0;
    Tree *tree = nullptr;
    int parent = -1, leftChild = -1, sibling = -1;
    int seqNum = -1, isLeaf = 0, inTree = 1;
    double brLen = -1.0;
    std::vector<int> compNums, rMatrixNums, gdasrvNums;
};

// One 4-int program record: an internal node to combine, and its children.
struct OpRec {
    int dst;     // cl slot written (root: unused)
    int flags;   // bit0: root (fused site sum)
    int nKids;
    int kidStart;
};
struct KidRec {
    int kind;    // 0 leaf, 1 internal
    int idx;     // leaf: tip row (seqNum); internal: cl slot
    int pnode;   // node whose P matrices (branch above it) are used
    int pad;
};
constexpr int OP_ROOT = 1;

// The 20-state category kernel's fused op record: the op and its first two
// children in one 64-byte line, [dst, flags, nKids, kidStart, then per
// child (kind, idx, pnode, pad)], four words spare.  Written after the kid
// records of that program; the kernel reads it one op ahead.
constexpr int AA_REC_WORDS = 16;
struct AaRec {
    int dst, flags, nKids, kidStart;
    int kind[2], idx[2], pnode[2], slot[2];
};

// Program of the wave-per-category DNA kernel (k_pruneWave): one 16-int
// record per node, in execution order, children inline (up to 3; more in
// a side array of 4-int records).  A child is a tip row, a slot of the
// wave's LDS slot file (a partial computed earlier in the same launch or
// preloaded at its start), or a partials slot in global memory.
enum { WK_TIP = 0, WK_LDS = 1, WK_GLOBAL = 2 };
struct WaveOp {
    int dst;       // partials slot written (the root: -1)
    int ldsDst;    // LDS slot that also receives the result, or -1
    int flags;     // OP_ROOT
    int nKids;
    int kidStart;  // index of the 4th child in the side array
    int kind[3], idx[3], pnode[3];
    int pad;
    int kidBase;   // program-order index of the first child (staged P)
};
static_assert(sizeof(WaveOp) == 64, "WaveOp is one 64-byte scalar load");
struct WaveKid {
    int kind, idx, pnode, pad;
};
// The lean kernel's op records (one per WaveOp, same order): binary ops
// whose children are staged tips or LDS slots carry precomputed offsets, so
// the kernel does no index arithmetic; every other op (the root,
// multifurcations, children in global memory, unstaged tips) is PK_GEN and
// runs from its WaveOp (its record is all zeros but the kind).  Tip
// children come first.  FP32 records are 32 bytes: the current and the next
// record take 16 scalar registers, leaving room for both children's P^T
// blocks; FP64 records are padded to 64 bytes (leanRecInts), which its
// register allocation handles better.
enum { PK_TT = 0, PK_TL = 1, PK_LL = 2, PK_GEN = 3 };
struct LeanOp {
    int kindLd;          // PK_* | (LDS slot of the result + 1) << 8 (0: none)
    int a0, a1;          // tip: code offset row * 64; LDS child: slot index
    int t0, t1;          // tip: byte offset of its staged P^T block in the category's region;
                         // LDS child: exponent index slot * 64
    int p0, p1;          // byte offset of the child's P^T block in global memory, node * K * 16 * element
    int dst;             // partials slot of the result
};
static_assert(sizeof(LeanOp) == 32, "LeanOp is one 32-byte scalar load");
inline int leanRecInts(bool fp32) { return fp32 ? 8 : 16; }   // record stride in ints
constexpr size_t LEAN_PT_WARM_MAX = 64 * 1024;   // P^T tables up to this size are pulled into L2
constexpr int WAVE_MAX_K = 8;             // categories (waves per workgroup)
constexpr int WAVE_MAX_LDS_SLOTS = 16;
constexpr int WAVE_TIP_LDS_MAX_TAXA = 256;
// Group of ops processed together: kind 0 = up to G independent binary,
// non-root ops; kind 1 = one op with any number of children (or the root).
struct GroupRec {
    int kind;
    int nOps;
    int opStart;
    int pad;
};
// P rows of a program's child branches are staged in LDS when they fit.
constexpr int LDS_P_BUDGET = 48 * 1024;

// Per-(tree, part) device state.
struct TreePart {
    bool alloc = false;
    int dim = 0, K = 0, nPat = 0, nPatPad = 0, nSlots = 0;
    int precision = FP64;
    // Eigen slots: [nComps*nRMat]; input pi[dim] + R[dim*dim], output
    // V[dim*dim], Vinv[dim*dim], lam[dim].
    int nEig = 0;
    double *eigIn = nullptr, *eigV = nullptr, *eigVi = nullptr, *eigLam = nullptr;
    // P^T per node and category: [nNodes][K][dim(j)][dim(s)] = P[s][j]
    double *PT64 = nullptr;
    float *PT32 = nullptr;
    // Partials: [nSlots][nPatPad][K][dim] of float or double, scale
    // exponents [nSlots][nPatPad] (FP32 mode); the wave kernel's layout is
    // [nSlots][K][nPatPad][dim] with exponents [nSlots][K][nPatPad].
    void *cl = nullptr;
    int *scl = nullptr;
    size_t clBytes = 0;
    // Reduction state.
    double *blockSums = nullptr;
    int *blockBad = nullptr;
    unsigned *counter = nullptr;
    double *patLikes = nullptr;    // [nPatPad]
    int maxBlocks = 0;
    // Staged program (device pointers into the arena).
    int nGroups = 0, nKids = 0, progWords = 0;
    // Wave-kernel program: records, side kids, preloads, LDS slots per
    // wave, tips in LDS; per-slot exponent count (K*nPatPad for the wave
    // layout, else nPatPad).
    bool wave = false;
    int nWOps = 0, nWKids = 0, nPre = 0, nLds = 0, tipLds = 0, nKidTot = 0;
    int nStream = 1, nS0 = 0;
    int popsOff = 0;               // LeanOp records, in ints from the program start
    int aaLds = 0;                 // LDS slots of the 20-state category program
    int aaRecOff = -1;             // word offset of its fused op records in the program, or -1
    int tipPnOff = 0, nTipKid = 0; // staged tip children's P^T nodes
    // The staged program is reused while everything it depends on is
    // unchanged (tree structure, postorder, marks, kernel choice, layout).
    bool progValid = false;
    uint64_t progKey = 0;
    int progNOps = 0, progNKids = 0;
    unsigned layoutGen = 0;
    void *PP = nullptr;            // [nNodes][K][16] of float or double
    size_t waveLds = 0;
    size_t sclPerSlot = 0;
    // Host-side flags.
    std::vector<char> dirty;       // per node
    std::vector<char> pValid;      // per node
    bool everComputed = false;
    // Deferred parameter work.  setPrams / calculateBigPDecks snapshot their
    // inputs here (pinned host memory) and mark them pending; the pending
    // eigensystems and P matrices are built by one launch each right before
    // the next pruning launch, on the same stream.
    // All per-evaluation host->device staging lives in one pinned arena with
    // a device mirror at the same offsets, uploaded by one copy.
    char *hArena = nullptr, *dArena = nullptr;
    size_t arenaBytes = 0, offEigList = 0, offPack = 0, offPi = 0, offProg = 0;
    int progCapWords = 0;
    double *hEigIn = nullptr;      // [nEig][dim + dim*dim]  pi, R   (arena)
    int *hEigList = nullptr;       // [nEig] pending slots           (arena)
    char *hPack = nullptr;         // packed pending P items          (arena)
    double *hRootPi = nullptr;     // [MAX_DIM]                       (arena)
    int *hProg = nullptr;          // groups, ops, kids               (arena)
    std::vector<char> eigPending;  // per eigen slot
    std::vector<char> eigSnap;     // per eigen slot: hEigIn holds a snapshot
    bool flushDirect = false;      // the staged P matrices come straight from exp(Qt)
    bool eigStale = false;         // snapshots taken while P came straight from
                                   // (pi, R), so eigV/eigVi/eigLam lag them
    double *hT = nullptr;          // [nNodes][K] effective branch lengths
    int *hEs = nullptr;            // [nNodes] eigen slot of each node
    size_t packBytes = 0;
    std::vector<char> pPending;    // per node
    int pendE = 0, pendI = 0;      // staged counts for the next launch
    double *hResult = nullptr;     // pinned [2]
    double *dResultHost = nullptr; // device address of hResult: the kernel writes lnL there
    const char *dArenaHost = nullptr;  // device address of hArena: the 4-state P kernel
                                       // reads its staged inputs there, no upload
    bool progOnDevice = false;     // dArena holds the staged program
    bool rootPiOnDevice = false;   // dArena holds rootPiUp as the root frequencies
    double rootPiUp[MAX_DIM] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double lastKernelUs = -1.0;
    // branch-length derivatives (allocated by their first use)
    double *cl2 = nullptr, *P1T = nullptr, *P2T = nullptr, *dQ = nullptr, *dRate = nullptr;
    double *dPartial = nullptr, *dOut = nullptr;
    int *dItems = nullptr;
    void *dProg = nullptr;
    size_t dPartialCap = 0, dProgCap = 0;
    bool derivValid = false;   // cl2, P1T and P2T hold the last derivative call's values
    // ancestral draws: the root's children, uniforms and draws
    char *ancBuf = nullptr;
    size_t ancCap = 0;
    // Changed since the last p4g_syncChanged with its partner tree: partial
    // slots written and nodes whose P was marked (chAll: everything, after
    // the buffers were made or re-laid out).
    std::vector<char> chSlot, chP;
    bool chAll = true;
    std::vector<int> progSlots;    // slots the staged program writes
};

struct Tree {
    int nNodes = 0, nLeaves = 0;
    std::vector<Node *> nodes;
    int root = -1;
    int *preOrder = nullptr, *postOrder = nullptr;  // aliased
    double *partLikes = nullptr;                    // aliased
    Data *data = nullptr;
    Model *model = nullptr;
    int nParts = 0;
    double logLike = 0.0;
    int precision = FP64;
    std::vector<TreePart> tp;
    std::vector<int> slotOf;  // node -> cl slot (-1 none)
    char *simBuf = nullptr;   // simulation: node lists, root comp, states [nNodes][nChar], categories
    size_t simCap = 0;
    // p4g_syncChanged: copy descriptors in pinned memory (read by the copy
    // kernel through their device address) and the event of the last copy
    struct CopyRange *hRanges = nullptr, *dRanges = nullptr;
    int rangeCap = 0;
    hipEvent_t syncEv = nullptr;
    ~Tree();
};

// ------------------------------------------------------- kernel arguments
struct EigArgs {
    int dim;
    int nSlots;
    const int *slotList;  // which eigen slots to compute
    const double *in;     // per slot: pi[dim], R[dim*dim]
    double *V, *Vi, *lam;
};

struct PmatArgs {
    int dim, K, nItems;     // items = listed nodes
    const int *nodeList;    // [nItems] node numbers
    const int *eigSlot;     // [nItems]
    const double *t;        // [nItems][K] effective branch lengths
    const double *V, *Vi, *lam;
    double *PT64;
    float *PT32;
};

struct PruneArgs {
    const GroupRec *groups;
    const OpRec *ops;
    const KidRec *kids;
    int nGroups, nOps, nKids;
    int dim, K, log2Kp;
    int nPat, nPatPad;
    const uint8_t *tips;
    const void *PT;          // float or double by precision
    void *cl;
    int *scl;
    const double *counts;
    const uint32_t *invMask;
    int nWords;
    const uint64_t *ambMask;
    const double *rootPi;
    double pInvar;
    int rootLeafRow;         // tip row of a leaf root, else -1
    double *blockSums;
    int *blockBad;
    unsigned *counter;
    double *result;
    double *patLikes;        // may be null
    // wave-per-category DNA kernel
    const WaveOp *wops;
    const WaveKid *wkids;
    const int *wpre;         // [nPre][3]: partials slot, LDS slot, stream
    int nWOps, nPre, nLds;   // nLds: LDS slots per wave
    int tipLds;              // 1: the tile's tip codes are staged in LDS
    int nTax;
    int nStream, nS0;        // wave sets per category; stream 0 runs ops [0, nS0),
                             // stream 1 [nS0, nWOps - 1), the root is last
    // P^T of every child in program order, [nKidTot][K][16] (k_pgather)
    const void *PP;
    int nKidTot;
    // lean kernel: one LeanOp per WaveOp; the P^T nodes of the staged tip
    // children (nTipKid, in the order of their staging)
    const LeanOp *pops;
    int nTipKid;
    const int *tipPn;
    int nAmb;                // entries of ambMask
    int ptWarmBytes;         // lean kernel: bytes of P^T to pull into L2 at the start (0: none)
    const int *aaRec;        // 20-state category kernel: fused op records, [nOps][AA_REC_WORDS]
};

// ---------------------------------------------- branch-length derivatives
// One entry per non-root node of the up pass (preorder), and the node's
// siblings (the other children of its parent).
constexpr int DOP_LEAF = 1;          // the node is a leaf (idx: its tip row)
constexpr int DOP_PARENT_ROOT = 2;   // its parent is the root
struct DerivOp {
    int node, flags, idx, parent, sibStart, nSib, pad0, pad1;   // idx: tip row or partials slot
};
struct DerivSib {
    int kind, idx, pnode, pad;   // kind 0 tip (idx: tip row), 1 internal (idx: partials slot)
};
constexpr int DERIV_MAX_K = 16;
struct DerivPArgs {
    int dim, K, nItems;
    const int *nodeList;         // [nItems]
    const int *eigSlot;          // [nItems] (comp, rMatrix) slot of each node
    const double *rateFac;       // [nItems][K]: d(effective length)/d(branch length)
    const double *Q;             // [nEig][dim][dim] normalised rate matrices
    const double *PT64;          // P^T of the tree part
    double *P1T, *P2T;           // P'^T and P''^T, same layout
};
struct DerivArgs {
    int dim, K, nPat, nPatPad, nOps;
    int waveLayout;              // the down-pass partials are [slot][cat][pattern][4]
    const DerivOp *ops;
    const DerivSib *sibs;
    const uint8_t *tips;
    const uint64_t *ambMask;
    const double *cl;            // FP64 down-pass partials
    double *cl2;                 // [nNodes][K][nPatPad][dim]
    const double *PT64, *P1T, *P2T;
    const double *rootPi;
    int rootLeafRow;
    double pInvar;
    const double *counts;
    const uint32_t *invMask;
    int nWords;
    double *partial;             // [nOps][nBlocks][3]
    double *out;                 // [nNodes][3]: lnL across the node's branch, d1, d2
};
hipError_t launchPderiv(const DerivPArgs &a, hipStream_t s);
hipError_t launchEdgeDerivs(const DerivArgs &a, int nBlocks, hipStream_t s);

// Kernel launchers (p4g_kernels.hip).
hipError_t launchEig(const EigArgs &a, hipStream_t s);
hipError_t launchPmat(const PmatArgs &a, hipStream_t s);
// The eigen route's P for several items in one launch (first: nItems + 1
// prefix sums of each item's nItems * K * dim * dim entries).
hipError_t launchPmatMulti(const PmatArgs *args, const int *first, int nItems, int total, hipStream_t s);
// P = exp(Q t) per (branch, category) straight from the (pi, R) snapshots
// eigIn (4-state parts only); pmatDirect(dim) says whether that route is on.
hipError_t launchPexp(const PmatArgs &a, const double *eigIn, hipStream_t s);
bool pmatDirect(int dim);
// Direct P (dim 4 or 20) for several items in one launch: args/ins/first
// are device-readable arrays of nItems entries (first: nItems + 1 prefix
// sums of each item's nItems * K matrices, first[nItems] == total).
hipError_t launchPexpMulti(int dim, const PmatArgs *args, const double *const *ins, const int *first, int nItems,
                           int total, hipStream_t s);
// 20-state flushes with fewer pending P matrices use the eigen route.
constexpr int PEXP_WAVE_MIN_ITEMS = 16;
void setPmatRoute(int direct);
// 20-state FP32 pruning: one wavefront per (tile, category) for programs
// of at least minOps ops when catWaves is on (the default), else one per
// tile with every category.
constexpr int AA_CAT_MIN_OPS = 32;
void setAaTuning(int catWaves, int minOps);
bool pruneAaCategoryWaves(int nOps);
hipError_t launchPrune(const PruneArgs &a, int precision, int nBlocks, hipStream_t s);
hipError_t launchPruneBatch(const PruneArgs *dArgs, const int *dBlockItem, int nBlocks,
                            int dim, int precision, size_t lds, hipStream_t s);
int pruneBlockPatterns(int K);  // patterns per 256-thread block
int pruneGroupWidthFor(int dim, int precision);  // ops per binary group
void setPruneTuning(int precision, int G, int lds);
size_t pruneLdsBytes(int nKids, int K, int dim, int precision);  // 0 if P stays in global
// 20-state FP32 parts use the MFMA kernel and a tile-blocked partial layout
bool pruneUsesMfma(int dim, int K, int precision);
int pruneBlockPatternsFor(int dim, int K, int precision, int nOps);
hipError_t launchPruneAa(const PruneArgs &a, int nBlocks, hipStream_t s);
hipError_t launchPruneAaBatch(const PruneArgs *d, const int *bi, int nBlocks, int K, bool catWaves, size_t ldsBytes,
                              bool tipLds, hipStream_t s);
// the category kernel's LDS slot file: the most slots a program may use
// (0: none), and the bytes of n slots
int pruneAaLdsSlots(int K);
size_t pruneAaLdsBytes(int K, int nLds, int tipTaxa);
bool pruneAaTipLds(int nTax);
bool pruneAaCatItemOk(const PruneArgs &a);
void setAaLdsSlots(int n);
// 4-state parts with at most WAVE_MAX_K categories use the wave-per-category
// kernel: partials [slot][cat][pattern][state], FP32 exponents
// [slot][cat][pattern]; 64 patterns per workgroup of 64*K threads.
bool pruneUsesWave(int dim, int K, int precision);
bool pruneWaveLean(int precision);      // the lean kernel runs the wave program
int pruneWaveTilePatterns(int precision);  // patterns per workgroup of the wave program's kernel
bool pruneWaveVectorP(int precision);   // the wave kernel reads the k_pgather P stream
bool pruneWaveStreams(int precision);   // the wave kernel splits the tree into two streams
size_t pruneWaveLdsBytes(int K, int nLds, int tipLds, int nTax, int nTipKid, int precision);
int pruneWaveMaxSlots(int K, int tipLds, int nTax, int nTipKid, int precision);
// slots a program's preloads may reach (a cap below pruneWaveMaxSlots)
int pruneWavePreloadSlots(int K, int tipLds, int nTax, int nTipKid, int precision);
void setWavePreloadLds(int bytes);
hipError_t launchPruneWave(const PruneArgs &a, int precision, int nBlocks, size_t lds, hipStream_t s);
// P^T of the program's children in program order: PP[k][j] = PT[kidP[j]][k]
hipError_t launchPgather(const void *PT, const int *kidP, int nKidTot, int K, void *PP, int precision,
                         hipStream_t s);
hipError_t launchPruneWaveBatch(const PruneArgs *d, const int *bi, int nBlocks, int K, int nStream, int precision,
                                size_t lds, hipStream_t s);

// n words of p set to v by a small kernel (usable under stream capture,
// unlike hipMemsetAsync)
hipError_t launchFillU32(unsigned *p, unsigned v, int n, hipStream_t s);
// Device-to-device copies of n ranges in one launch (bytes: multiples of 4).
struct CopyRange {
    const void *src;
    void *dst;
    unsigned long long bytes;
};
hipError_t launchCopyRanges(const CopyRange *r, int n, hipStream_t s);

// ------------------------------------------- root-side consumers (p4g_anc.hip)
// Philox4x64-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC 2011): four 64-bit words from a 256-bit counter
// and a 128-bit key.  Counter-based, so each (stream, site, draw) has its
// own numbers whatever the launch shape.
struct Philox4 {
    uint64_t v[4];
};
__host__ __device__ inline Philox4 philox4x64(Philox4 c, uint64_t k0, uint64_t k1)
{
    const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
    for (int r = 0; r < 10; r++) {
#if defined(__HIP_DEVICE_COMPILE__)
        const uint64_t hi0 = __umul64hi(M0, c.v[0]), hi1 = __umul64hi(M1, c.v[2]);
#else
        const uint64_t hi0 = (uint64_t)(((unsigned __int128)M0 * c.v[0]) >> 64);
        const uint64_t hi1 = (uint64_t)(((unsigned __int128)M1 * c.v[2]) >> 64);
#endif
        const uint64_t lo0 = M0 * c.v[0], lo1 = M1 * c.v[2];
        Philox4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}
// The top 53 bits as a double in [0, 1).
__host__ __device__ inline double uniform53(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

constexpr int ANC_MAX_K = 16;
// Where the partials of a slot live (the layouts of the pruning kernels).
constexpr int CL_GROUPED = 0;   // [slot][pattern][cat][state], exponents [slot][pattern]
constexpr int CL_WAVE = 1;      // [slot][cat][pattern][4],     exponents [slot][cat][pattern]
constexpr int CL_TILED = 2;     // 20 states: [slot][pattern/16][cat][state/4][state%4][pattern%16],
                                // exponents [slot][pattern]
struct AncArgs {
    int dim, K, nPatPad, layout, fp32;
    const void *cl;              // the tree part's partials
    size_t slotBytes;
    const int *scl;              // FP32 exponents (fp32 only)
    size_t sclPerSlot;
    int nKid;
    const DerivSib *kids;        // the root's children: tip row or partials slot, P node
    const uint8_t *tips;
    const uint64_t *ambMask;
    const double *PT64;
    const double *rootPi;
    double pInvar;
    const uint32_t *invMask;
    int nWords;
    const int *sitePat;          // [nChar] pattern of each site
    int first, n;                // sites [first, first + n)
    const double *u;             // [n] uniforms, or null: Philox stream (seed, site)
    uint64_t seed;
    int *draw;                   // [n][4] chStNum, catNum, isInvar, invarChNum
};
hipError_t launchAncDraw(const AncArgs &a, hipStream_t s);

// ------------------------------------------------ simulation (p4g_sim.hip)
struct SimArgs {
    int dim, K, nChar, nNodes, pNum, root;
    int nPre;
    const int *pre;              // [nPre] non-root nodes in preorder
    const int *parent;           // [nNodes]
    const int *seqNum;           // [nNodes] tip row of a leaf, else -1
    const double *PT64;          // P^T of the tree part
    const double *rootComp;      // [dim]
    double pInvar;
    const int *refDraw;          // [nChar][4] root draws of a reference tree, or null
    uint64_t seed;
    uint8_t *states;             // [nNodes][nChar]
    int8_t *seq;                 // [nTax][nChar] the leaves' states
    int *simCat;                 // [nChar] category, -1 for an invariant site
};
hipError_t launchSimulate(const SimArgs &a, hipStream_t s);
// Over the sites of seq [nTax][nChar]: per-taxon counts of each state
// (counts [nTax][dim], added to), and acc[0] += constant sites, acc[1] +=
// distinct states summed over sites, acc[2] += codes outside 0..dim-1.
hipError_t launchSeqStats(const int8_t *seq, int nTax, int nChar, int dim, int *counts, int *acc, hipStream_t s);

// ------------------------------------------- pattern compression (p4g_compress.hip)
// The distinct columns of seq [nTax][nChar] in order of first occurrence:
// pats [nTax][nPat] (codes), counts [nPat], sitePat [nChar].  collisions:
// hash seeds that collided (the pipeline retries up to four).
hipError_t compressPatterns(const int8_t *dSeq, int nTax, int nChar, std::vector<int8_t> &pats,
                            std::vector<int> &counts, std::vector<int> &sitePat, int *nPatOut, int *collisions);
// sum[p] += 1 / like[p] for p < nPat (the CPO sums)
hipError_t launchCpoAccum(const double *like, double *sum, int nPat, hipStream_t s);

// Host numerics (p4g_host.hip).
int discreteGamma(double *freqK, double *rK, double alfa, double beta, int K, int median);

}  // namespace p4g
