// p4g_api.hip -- host side of the C-ABI declared in include/p4gpu.h.
//
// Object model mirrors p4's Pf structs (Pf/pftypes.h:24-240): data ->
// parts, model -> model parts -> comps / rMatrices / gdasrvs, tree ->
// nodes.  Host-owned arrays that the reference aliases from numpy are kept
// as pointers and re-read at every evaluation call.  Device state lives per
// (tree, part) in TreePart (p4g_internal.h).  Every GPU operation is ordered
// on one library stream; results come back through pinned host memory.
#include "../../include/p4gpu.h"
#include "p4g_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_map>

struct p4g_data : p4g::Data {};
struct p4g_part : p4g::Part {};
struct p4g_model : p4g::Model {};
struct p4g_tree : p4g::Tree {};
struct p4g_node : p4g::Node {};
struct p4g_gdasrv : p4g::Gdasrv {};

namespace p4g {

// ------------------------------------------------------------- utilities
static thread_local std::string g_err;
static bool g_timing = false;

void setError(const std::string &msg) { g_err = msg; }
void clearError() { g_err.clear(); }

bool hipOk(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    setError(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

static hipStream_t g_stream = nullptr;
static int g_devState = 0;  // 0 unknown, 1 ready, -1 unavailable

// Scratch memory of the entry point that allocates nothing
// (p4g_treeLogLikeNoAlloc): the pruning kernels' per-workgroup partial sums,
// bad-site flags and ticket counter live here instead of in the tree's own
// buffers.  Every GPU holds its own copy of the array until the process
// ends; host code uses the device address that hipGetSymbolAddress returns
// for the current device (g_scratch), never the array's name.
constexpr size_t NOALLOC_SCRATCH_BYTES = 512 * 1024;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static unsigned char *g_scratch[MAX_DEVICES] = {};

bool deviceReady()
{
    if (g_devState == 1) return true;
    if (g_devState == -1) {
        setError("no usable AMD GPU: libp4gpu runs its likelihood kernels on the GPU only");
        return false;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        g_devState = -1;
        setError("no usable AMD GPU: libp4gpu runs its likelihood kernels on the GPU only");
        return false;
    }
    if (!hipOk(hipSetDevice(0), "hipSetDevice")) {
        g_devState = -1;
        return false;
    }
    if (!hipOk(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking), "hipStreamCreate")) {
        g_devState = -1;
        return false;
    }
    // the no-allocation entry point's scratch array, for this device
    int dev = 0;
    void *ptr = nullptr;
    if (!hipOk(hipGetDevice(&dev), "hipGetDevice") ||
        !hipOk(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)") || dev < 0 ||
        dev >= MAX_DEVICES) {
        g_devState = -1;
        return false;
    }
    g_scratch[dev] = static_cast<unsigned char *>(ptr);
    g_devState = 1;
    return true;
}

hipStream_t stream() { return g_stream; }

template <typename T>
static bool devAlloc(T **p, size_t n, const char *what)
{
    if (*p) return true;
    void *q = nullptr;
    if (!hipOk(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)), what)) return false;
    *p = (T *)q;
    return true;
}

template <typename T>
static void devFree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

template <typename T>
static bool hostAlloc(T **p, size_t n, const char *what)
{
    void *q = nullptr;
    if (!hipOk(hipHostMalloc(&q, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault), what))
        return false;
    *p = (T *)q;
    return true;
}

// Device address of pinned host memory: kernels write results there
// directly, and the host reads them after the stream synchronises.
template <typename T>
static bool hostDevPtr(T *h, T **d, const char *what)
{
    void *q = nullptr;
    if (!hipOk(hipHostGetDevicePointer(&q, h, 0), what)) return false;
    *d = (T *)q;
    return true;
}

template <typename T>
static void hostFree(T *&p)
{
    if (p) (void)hipHostFree((void *)p);
    p = nullptr;
}

Part::~Part()
{
    devFree(dev.tips);
    devFree(dev.counts);
    devFree(dev.invMask);
    devFree(dev.ambMask);
    devFree(dev.sitePat);
    devFree(cpoSum);
    devFree(dSeq);
}

ModelPart::~ModelPart()
{
    for (auto *g : gdasrvs) delete static_cast<p4g_gdasrv *>(g);
}

Model::~Model()
{
    for (auto *mp : parts) delete mp;
}

static void freeTreePart(TreePart &t)
{
    devFree(t.eigV);
    devFree(t.eigVi);
    devFree(t.eigLam);
    devFree(t.PT64);
    devFree(t.PT32);
    if (t.cl) (void)hipFree(t.cl);
    t.cl = nullptr;
    if (t.PP) (void)hipFree(t.PP);
    t.PP = nullptr;
    devFree(t.scl);
    devFree(t.blockSums);
    devFree(t.blockBad);
    devFree(t.counter);
    devFree(t.patLikes);
    devFree(t.dArena);
    hostFree(t.hArena);
    hostFree(t.hT);
    hostFree(t.hEs);
    hostFree(t.hResult);
    if (t.ev0) (void)hipEventDestroy(t.ev0);
    if (t.ev1) (void)hipEventDestroy(t.ev1);
    t.ev0 = t.ev1 = nullptr;
    devFree(t.cl2);
    devFree(t.P1T);
    devFree(t.P2T);
    devFree(t.dQ);
    devFree(t.dRate);
    devFree(t.dPartial);
    devFree(t.dOut);
    devFree(t.dItems);
    if (t.dProg) (void)hipFree(t.dProg);
    t.dProg = nullptr;
    t.dPartialCap = t.dProgCap = 0;
    if (t.ancBuf) (void)hipFree(t.ancBuf);
    t.ancBuf = nullptr;
    t.ancCap = 0;
    t.derivValid = false;
    t.alloc = false;
}

Tree::~Tree()
{
    if (g_devState == 1) (void)hipStreamSynchronize(g_stream);
    for (auto &t : tp) freeTreePart(t);
    if (simBuf) (void)hipFree(simBuf);
    simBuf = nullptr;
    hostFree(hRanges);
    if (syncEv) (void)hipEventDestroy(syncEv);
    syncEv = nullptr;
}

// ------------------------------------------------------------- data side
static bool equateIsNLike(const Part *p, int e)
{
    for (int j = 0; j < p->dim; j++)
        if (!p->equates[(size_t)e * p->dim + j]) return false;
    return true;
}

// Pf/part.c:716-848: states compatible with every taxon, per pattern.
static void computeInvar(Part *p)
{
    const int nP = p->nPatterns, dim = p->dim;
    p->invarVec.assign(nP, 0);
    p->invarArray.assign((size_t)dim * nP, 0);
    for (int k = 0; k < nP; k++) {
        int sum = 0;
        for (int j = 0; j < dim; j++) {
            bool ok = true;
            for (int t = 0; t < p->nTax && ok; t++) {
                const int code = p->patterns[(size_t)t * nP + k];
                if (code >= 0)
                    ok = (code == j);
                else if (code == N_LIKE || code == GAP_CODE || code == QMARK_CODE)
                    ok = true;
                else
                    ok = p->equates[(size_t)(code - EQUATES_BASE) * dim + j] != 0;
            }
            p->invarArray[(size_t)j * nP + k] = ok;
            sum += ok;
        }
        p->invarVec[k] = sum;
    }
    p->invarSet = true;
}

static bool checkCodes(const Part *p)
{
    for (size_t i = 0; i < p->patterns.size(); i++) {
        const int c = p->patterns[i];
        if (c >= p->dim) {
            setError("bad character code " + std::to_string(c) + " (>= dim)");
            return false;
        }
        if (c < 0 && c != GAP_CODE && c != QMARK_CODE && c != N_LIKE &&
            !(c >= EQUATES_BASE && c < EQUATES_BASE + p->nEquates)) {
            setError("bad character code " + std::to_string(c));
            return false;
        }
    }
    return true;
}

static bool partSeqBuffer(Part *p);

// Upload one data part in the device layout of the pruning kernel.
static bool uploadPart(Part *p)
{
    if (p->dev.valid) return true;
    if (p->nPatterns < 1) {
        setError("part has no patterns (call makePatterns first)");
        return false;
    }
    if (p->dim > MAX_DIM) {
        setError("dim " + std::to_string(p->dim) + " exceeds the supported maximum of 64");
        return false;
    }
    if (!checkCodes(p)) return false;
    if (!p->invarSet) computeInvar(p);
    DevPart &d = p->dev;
    if (d.tips || d.counts || d.invMask || d.ambMask || d.sitePat) {
        // re-upload after the part changed: the sizes may differ
        if (!hipOk(hipStreamSynchronize(stream()), "sync before re-upload")) return false;
        devFree(d.tips);
        devFree(d.counts);
        devFree(d.invMask);
        devFree(d.ambMask);
        devFree(d.sitePat);
    }
    const int nP = p->nPatterns;
    const int nPatPad = (nP + 255) / 256 * 256;
    d.nPatPad = nPatPad;
    d.nWords = (p->dim + 31) / 32;
    // ambiguity table: each partial (not all-state) equate gets an index
    std::vector<int> ambIndex(p->nEquates, -1);
    d.hostAmb.clear();
    for (int e = 0; e < p->nEquates; e++) {
        if (equateIsNLike(p, e)) continue;
        uint64_t m = 0;
        for (int j = 0; j < p->dim; j++)
            if (p->equates[(size_t)e * p->dim + j]) m |= (1ull << j);
        ambIndex[e] = (int)d.hostAmb.size();
        d.hostAmb.push_back(m);
    }
    d.nAmb = (int)d.hostAmb.size();
    if (p->dim + d.nAmb > (int)TIP_ALL) {
        setError("too many ambiguity codes for the 8-bit tip alphabet");
        return false;
    }
    std::vector<uint8_t> tips((size_t)p->nTax * nPatPad, (uint8_t)TIP_ALL);
    for (int t = 0; t < p->nTax; t++)
        for (int k = 0; k < nP; k++) {
            const int c = p->patterns[(size_t)t * nP + k];
            uint8_t v;
            if (c >= 0)
                v = (uint8_t)c;
            else if (c == GAP_CODE || c == QMARK_CODE || c == N_LIKE)
                v = (uint8_t)TIP_ALL;
            else {
                const int e = c - EQUATES_BASE;
                v = ambIndex[e] < 0 ? (uint8_t)TIP_ALL : (uint8_t)(p->dim + ambIndex[e]);
            }
            tips[(size_t)t * nPatPad + k] = v;
        }
    std::vector<double> counts(nPatPad, 0.0);
    for (int k = 0; k < nP; k++) counts[k] = (double)p->patternCounts[k];
    std::vector<uint32_t> inv((size_t)d.nWords * nPatPad, 0u);
    for (int j = 0; j < p->dim; j++)
        for (int k = 0; k < nP; k++)
            if (p->invarArray[(size_t)j * nP + k]) inv[(size_t)(j / 32) * nPatPad + k] |= (1u << (j % 32));
    // the pattern of each site (sites are patterns when no site map is set)
    std::vector<int> sitePat;
    if (p->nChar > 0 && (int)p->seqPosPatternIndex.size() == p->nChar)
        sitePat = p->seqPosPatternIndex;
    else
        for (int k = 0; k < nP; k++) sitePat.push_back(k);
    d.nSites = (int)sitePat.size();
    for (int v : sitePat)
        if (v < 0 || v >= nP) {
            setError("a site maps to an out-of-range pattern");
            return false;
        }
    if (!devAlloc(&d.tips, tips.size(), "hipMalloc tips") ||
        !devAlloc(&d.counts, counts.size(), "hipMalloc counts") ||
        !devAlloc(&d.invMask, inv.size(), "hipMalloc invMask") ||
        !devAlloc(&d.ambMask, std::max<size_t>(d.hostAmb.size(), 1), "hipMalloc ambMask") ||
        !devAlloc(&d.sitePat, sitePat.size(), "hipMalloc site patterns"))
        return false;
    hipStream_t s = stream();
    if (!hipOk(hipMemcpyAsync(d.sitePat, sitePat.data(), sitePat.size() * 4, hipMemcpyHostToDevice, s),
               "copy site patterns") ||
        !hipOk(hipMemcpyAsync(d.tips, tips.data(), tips.size(), hipMemcpyHostToDevice, s), "copy tips") ||
        !hipOk(hipMemcpyAsync(d.counts, counts.data(), counts.size() * 8, hipMemcpyHostToDevice, s),
               "copy counts") ||
        !hipOk(hipMemcpyAsync(d.invMask, inv.data(), inv.size() * 4, hipMemcpyHostToDevice, s),
               "copy invMask"))
        return false;
    if (!d.hostAmb.empty() &&
        !hipOk(hipMemcpyAsync(d.ambMask, d.hostAmb.data(), d.hostAmb.size() * 8,
                              hipMemcpyHostToDevice, s),
               "copy ambMask"))
        return false;
    if (!hipOk(hipStreamSynchronize(s), "upload part")) return false;
    d.valid = true;
    return true;
}

// ------------------------------------------------------------ tree side
static int log2ceil(int K)
{
    int l = 0;
    while ((1 << l) < K) l++;
    return l;
}

static bool allocTreePart(Tree *t, int pNum)
{
    TreePart &tp = t->tp[pNum];
    ModelPart *mp = t->model->parts[pNum];
    Part *dp = t->data->parts[pNum];
    if (!uploadPart(dp)) return false;
    const bool wave = pruneUsesWave(mp->dim, mp->nCat, t->precision);
    // the data part's patterns changed (a simulation, new sequences): every
    // buffer sized by the pattern count is made again
    const bool patsChanged = tp.alloc && (tp.nPat != dp->nPatterns || tp.nPatPad != dp->dev.nPatPad);
    if (tp.alloc && !patsChanged && tp.precision == t->precision && tp.wave == wave) return true;
    if (tp.alloc) {
        // precision, kernel or pattern change: the partial buffers are re-laid out
        if (!hipOk(hipStreamSynchronize(stream()), "sync before re-layout")) return false;
        if (patsChanged) {
            devFree(tp.patLikes);
            devFree(tp.cl2);
            devFree(tp.dPartial);
            tp.dPartialCap = 0;
            tp.derivValid = false;
        }
        if (tp.cl) (void)hipFree(tp.cl);
        tp.cl = nullptr;
        devFree(tp.scl);
        // the block count depends on the kernel
        devFree(tp.blockSums);
        devFree(tp.blockBad);
        tp.alloc = false;
        tp.progValid = false;
        tp.layoutGen++;
        std::fill(tp.dirty.begin(), tp.dirty.end(), 1);
        tp.everComputed = false;
    }
    const int S = mp->dim, K = mp->nCat;
    if (S != dp->dim) {
        setError("model part dim differs from data part dim");
        return false;
    }
    tp.dim = S;
    tp.K = K;
    tp.nPat = dp->nPatterns;
    tp.nPatPad = dp->dev.nPatPad;
    tp.precision = t->precision;
    tp.wave = wave;
    tp.sclPerSlot = (size_t)(wave ? K : 1) * tp.nPatPad;
    // slots: every non-leaf node, plus one for a leaf root
    t->slotOf.assign(t->nNodes, -1);
    int ns = 0;
    for (int n = 0; n < t->nNodes; n++)
        if (t->nodes[n] && !t->nodes[n]->isLeaf) t->slotOf[n] = ns++;
    tp.nSlots = ns + 1;
    tp.chSlot.assign(tp.nSlots, 1);
    tp.chP.assign(t->nNodes, 1);
    tp.chAll = true;
    tp.nEig = mp->nComps * mp->nRMatrices;
    const size_t nPS = (size_t)t->nNodes * K * S * S;
    // blocks of the launch: the largest count any pruning kernel may use
    // enough block sums for any kernel choice (the smallest tile: 16 patterns)
    tp.maxBlocks = tp.nPatPad / std::min({pruneBlockPatternsFor(S, K, t->precision, 0), pruneBlockPatterns(K), 16});
    const size_t elt = t->precision == FP32 ? sizeof(float) : sizeof(double);
    tp.clBytes = (size_t)tp.nPatPad * K * S * elt;
    if (!devAlloc(&tp.eigV, (size_t)tp.nEig * S * S, "hipMalloc eigV") ||
        !devAlloc(&tp.eigVi, (size_t)tp.nEig * S * S, "hipMalloc eigVi") ||
        !devAlloc(&tp.eigLam, (size_t)tp.nEig * S, "hipMalloc eigLam") ||
        !devAlloc(&tp.PT64, nPS, "hipMalloc P64") || !devAlloc(&tp.PT32, nPS, "hipMalloc P32") ||
        !devAlloc(&tp.blockSums, tp.maxBlocks, "hipMalloc blockSums") ||
        !devAlloc(&tp.blockBad, tp.maxBlocks, "hipMalloc blockBad") ||
        !devAlloc(&tp.counter, 1, "hipMalloc counter") ||
        !devAlloc(&tp.patLikes, tp.nPatPad, "hipMalloc patLikes"))
        return false;
    if (!tp.cl && !hipOk(hipMalloc(&tp.cl, tp.clBytes * tp.nSlots), "hipMalloc partials")) return false;
    if (!tp.PP && !hipOk(hipMalloc(&tp.PP, (size_t)K * t->nNodes * 16 * sizeof(double)), "hipMalloc staged P"))
        return false;
    if (t->precision == FP32 && !devAlloc(&tp.scl, (size_t)tp.nSlots * tp.sclPerSlot, "hipMalloc scale"))
        return false;
    if (!tp.hArena) {
        auto up16 = [](size_t b) { return (b + 15) & ~(size_t)15; };
        const size_t eigBytes = up16(sizeof(double) * (size_t)tp.nEig * (S + S * S));
        tp.packBytes = (size_t)t->nNodes * (K * sizeof(double) + 2 * sizeof(int));
        tp.offEigList = eigBytes;
        tp.offPack = tp.offEigList + up16(sizeof(int) * tp.nEig);
        tp.offPi = tp.offPack + up16(tp.packBytes);
        tp.offProg = tp.offPi + up16(sizeof(double) * MAX_DIM);
        tp.progCapWords = 48 * t->nNodes + 128;
        tp.arenaBytes = tp.offProg + sizeof(int) * tp.progCapWords;
        if (!hostAlloc(&tp.hArena, tp.arenaBytes, "hipHostMalloc staging arena") ||
            !devAlloc(&tp.dArena, tp.arenaBytes, "hipMalloc staging arena") ||
            !hostAlloc(&tp.hT, (size_t)t->nNodes * K, "hipHostMalloc branch lengths") ||
            !hostAlloc(&tp.hEs, t->nNodes, "hipHostMalloc eigen slots") ||
            !hostAlloc(&tp.hResult, 2, "hipHostMalloc result") ||
            !hostDevPtr(tp.hResult, &tp.dResultHost, "device address of result"))
            return false;
        char *dah = nullptr;
        if (!hostDevPtr(tp.hArena, &dah, "device address of staging arena")) return false;
        tp.dArenaHost = dah;
        tp.eigSnap.assign(tp.nEig, 0);
        tp.eigStale = false;
        tp.progOnDevice = false;
        tp.rootPiOnDevice = false;
        tp.hEigIn = reinterpret_cast<double *>(tp.hArena);
        tp.hEigList = reinterpret_cast<int *>(tp.hArena + tp.offEigList);
        tp.hPack = tp.hArena + tp.offPack;
        tp.hRootPi = reinterpret_cast<double *>(tp.hArena + tp.offPi);
        tp.hProg = reinterpret_cast<int *>(tp.hArena + tp.offProg);
        if (!hipOk(hipEventCreate(&tp.ev0), "hipEventCreate") ||
            !hipOk(hipEventCreate(&tp.ev1), "hipEventCreate"))
            return false;
    }
    tp.eigPending.assign(tp.nEig, 0);
    tp.pPending.assign(t->nNodes, 0);
    if (!hipOk(hipMemsetAsync(tp.counter, 0, sizeof(unsigned), stream()), "zero counter")) return false;
    tp.dirty.assign(t->nNodes, 1);
    tp.pValid.assign(t->nNodes, 0);
    tp.alloc = true;
    return true;
}

// needPatterns = false: a part whose patterns a simulation has cleared
// keeps the tree state it has (parameters and simulation do not read the
// patterns).
static bool treeReady(Tree *t, bool needPatterns = true)
{
    if (!deviceReady()) return false;
    if (!t->data || !t->model) {
        setError("tree has no data or model");
        return false;
    }
    if (t->root < 0 || !t->nodes[t->root]) {
        setError("tree root is not set");
        return false;
    }
    if (t->data->nParts != t->model->nParts || (int)t->tp.size() != t->model->nParts) {
        setError("data and model part counts differ");
        return false;
    }
    for (int n = 0; n < t->nNodes; n++)
        if (!t->nodes[n]) {
            setError("node " + std::to_string(n) + " was not created");
            return false;
        }
    for (int p = 0; p < t->nParts; p++) {
        if (!t->model->parts[p] || !t->data->parts[p]) {
            setError("model or data part missing");
            return false;
        }
        if ((needPatterns || !t->tp[p].alloc || t->data->parts[p]->nPatterns > 0) && !allocTreePart(t, p))
            return false;
    }
    return true;
}

// Effective branch length of Pf/p4_node.c:321-346.
static double branchT(const Tree *t, const ModelPart *mp, const Node *nd, int pNum, int k)
{
    const double rate = mp->nGdasrvs ? mp->gdasrvs[nd->gdasrvNums[pNum]]->rates[k] : 1.0;
    const double base = mp->nGdasrvs ? (nd->brLen * rate * mp->relRate) : (nd->brLen * mp->relRate);
    (void)t;
    return mp->pInvar == 0.0 ? base : base / (1.0 - mp->pInvar);
}

static bool checkNodeModel(const Tree *t, const ModelPart *mp, const Node *nd, int pNum)
{
    const int c = nd->compNums[pNum], r = nd->rMatrixNums[pNum], g = nd->gdasrvNums[pNum];
    if (c < 0 || c >= mp->nComps || (nd->nodeNum != t->root && (r < 0 || r >= mp->nRMatrices)) ||
        (mp->nGdasrvs && nd->nodeNum != t->root && (g < 0 || g >= mp->nGdasrvs))) {
        setError("node " + std::to_string(nd->nodeNum) + " has an out-of-range model component");
        return false;
    }
    if (mp->comps[c].val == nullptr) {
        setError("comp " + std::to_string(c) + " has no values");
        return false;
    }
    return true;
}

// Snapshot the inputs of eigensystem `slot` (pi of its comp, R of its
// rMatrix) and mark it for the next flush.
static void markEig(Tree *t, int pNum, int slot)
{
    TreePart &tp = t->tp[pNum];
    ModelPart *mp = t->model->parts[pNum];
    const int S = mp->dim;
    const int c = slot / mp->nRMatrices, r = slot % mp->nRMatrices;
    double *dst = tp.hEigIn + (size_t)slot * (S + S * S);
    std::memcpy(dst, mp->comps[c].val, sizeof(double) * S);
    std::memcpy(dst + S, mp->rMatrices[r].bigR.data(), sizeof(double) * S * S);
    tp.eigPending[slot] = 1;
    tp.eigSnap[slot] = 1;
}

// Snapshot the effective branch lengths of node n (Pf/p4_node.c:321-346)
// and mark its P matrices for the next flush.
static void markP(Tree *t, int pNum, const Node *nd)
{
    TreePart &tp = t->tp[pNum];
    ModelPart *mp = t->model->parts[pNum];
    const int K = mp->nCat, n = nd->nodeNum;
    for (int k = 0; k < K; k++) tp.hT[(size_t)n * K + k] = branchT(t, mp, nd, pNum, k);
    tp.hEs[n] = nd->compNums[pNum] * mp->nRMatrices + nd->rMatrixNums[pNum];
    tp.pPending[n] = 1;
    tp.pValid[n] = 1;
    if ((size_t)n < tp.chP.size()) tp.chP[n] = 1;
}

// Pack the pending eigensystem slots and P items of one part into the
// staging arena (no copies, no launches).  When P comes straight from the
// (pi, R) snapshots no eigensystem is computed; the first flush on the
// eigen route after that recomputes every slot that holds a snapshot.
static void stageParams(Tree *t, int pNum)
{
    TreePart &tp = t->tp[pNum];
    const int K = t->model->parts[pNum]->nCat;
    const int dim = t->model->parts[pNum]->dim;
    int nI = 0;
    for (int n = 0; n < t->nNodes; n++) nI += tp.pPending[n] ? 1 : 0;
    // 20-state P straight from exp(Qt) is one latency-bound wavefront per
    // matrix: worth it for many matrices (a Q change, a full evaluation),
    // not for the few of a branch move, which reuse cached eigensystems
    const bool direct = pmatDirect(dim) && (dim == 4 || nI >= PEXP_WAVE_MIN_ITEMS);
    tp.flushDirect = direct;
    if (!direct && tp.eigStale) {
        for (int i = 0; i < tp.nEig; i++)
            if (tp.eigSnap[i]) tp.eigPending[i] = 1;
        tp.eigStale = false;
    }
    int nE = 0;
    for (int i = 0; i < tp.nEig; i++)
        if (tp.eigPending[i]) {
            if (direct)
                tp.eigStale = true;
            else
                tp.hEigList[nE++] = i;
            tp.eigPending[i] = 0;
        }
    double *ht = reinterpret_cast<double *>(tp.hPack);
    int *hn = reinterpret_cast<int *>(ht + (size_t)nI * K);
    int *he = hn + nI;
    int i = 0;
    for (int n = 0; n < t->nNodes && nI; n++) {
        if (!tp.pPending[n]) continue;
        for (int k = 0; k < K; k++) ht[(size_t)i * K + k] = tp.hT[(size_t)n * K + k];
        hn[i] = n;
        he[i] = tp.hEs[n];
        tp.pPending[n] = 0;
        i++;
    }
    tp.pendE = nE;
    tp.pendI = nI;
}


// Launch the staged eigensystems, then the staged P matrices.  Both read
// their inputs (slot list, (pi, R) snapshots, branch lengths, nodes, slots)
// from the pinned arena itself, through its device address: no upload.
// A direct-P flush (4 or 20 states) whose launch a batched call makes
// together with the other items' of the same state count.
struct DeferredP {
    PmatArgs a;
    const double *in;      // (pi, R) snapshots of the direct route (pinned arena); null: eigen route
    const double *devIn;   // their device copy once the call's uploads have run, or null
};

// Error path of a batched call: the P flushes deferred so far are launched
// one by one, so the items already staged (their marks cleared) keep
// consistent P matrices; the first error stays the reported one.
static void launchDeferredEach(const std::vector<DeferredP> &defer)
{
    for (const DeferredP &dp : defer) {
        if (dp.in)
            (void)launchPexp(dp.a, dp.in, stream());
        else
            (void)launchPmat(dp.a, stream());
    }
    // the kernels read the pinned staging arena, which the next staging call
    // overwrites: they must have run before this error path returns
    (void)hipStreamSynchronize(stream());
}

// devEig: the (pi, R) snapshots and the pack were copied to the device
// arena ahead of the P kernel on the stream (20 states: one wavefront per
// (branch, category) would otherwise read its 3.4 KB across PCIe, 1,012
// times for prot128, after two dependent reads of the pack)
static bool launchParams(Tree *t, int pNum, std::vector<DeferredP> *defer = nullptr, bool devEig = false)
{
    TreePart &tp = t->tp[pNum];
    ModelPart *mp = t->model->parts[pNum];
    const int S = mp->dim, K = mp->nCat;
    hipStream_t s = stream();
    const char *base = tp.dArenaHost;
    if (tp.pendE) {
        EigArgs a{S, tp.pendE, reinterpret_cast<const int *>(base + tp.offEigList),
                  reinterpret_cast<const double *>(base), tp.eigV, tp.eigVi, tp.eigLam};
        if (!hipOk(launchEig(a, s), "launch eigensystem kernel")) return false;
    }
    if (tp.pendI) {
        const int nI = tp.pendI;
        const bool direct = tp.flushDirect;
        // the pack (t, nodes, eigen slots) from the device copy too when it went up
        const char *d = (devEig ? (const char *)tp.dArena : base) + tp.offPack;
        PmatArgs a{S, K, nI, reinterpret_cast<const int *>(d + sizeof(double) * nI * K),
                   reinterpret_cast<const int *>(d + sizeof(double) * nI * K + sizeof(int) * nI),
                   reinterpret_cast<const double *>(d), tp.eigV, tp.eigVi, tp.eigLam, tp.PT64, tp.PT32};
        const double *inDev = devEig ? reinterpret_cast<const double *>(tp.dArena) : nullptr;
        if (direct && defer) {
            defer->push_back({a, reinterpret_cast<const double *>(base), inDev});
        } else if (direct) {
            if (!hipOk(launchPexp(a, inDev ? inDev : reinterpret_cast<const double *>(base), s), "launch P kernel"))
                return false;
        } else if (defer) {
            defer->push_back({a, nullptr, nullptr});
        } else if (!hipOk(launchPmat(a, s), "launch P kernel")) {
            return false;
        }
    }
    tp.pendE = tp.pendI = 0;
    return true;
}

static bool flushParams(Tree *t, int pNum)
{
    stageParams(t, pNum);
    TreePart &tp = t->tp[pNum];
    if (!tp.pendE && !tp.pendI) return true;
    return launchParams(t, pNum);
}

// Flush every part and wait (for inspection and copy entry points).
static bool flushAll(Tree *t)
{
    for (int p = 0; p < t->nParts; p++)
        if (t->tp[p].alloc && !flushParams(t, p)) return false;
    return hipOk(hipStreamSynchronize(stream()), "parameter kernels");
}

// Pf/p4_tree.c:218-541 for one part.
static bool setPramsPart(Tree *t, int pNum)
{
    ModelPart *mp = t->model->parts[pNum];
    const int S = mp->dim;
    if (mp->nCat > 1)
        for (auto *g : mp->gdasrvs)
            if (g->free) discreteGamma(g->freqs, g->rates, g->val[0], g->val[0], mp->nCat, 0);
    for (auto &r : mp->rMatrices)
        if (r.free && r.spec == RMATRIX_2P && S == 4) {
            const double al = 1.0 / 3.0, be = al * r.kappa;
            const double R[16] = {0, al, be, al, al, 0, al, be, be, al, 0, al, al, be, al, 0};
            r.bigR.assign(R, R + 16);
        }
    const double pivecMin = t->model->PIVEC_MIN ? t->model->PIVEC_MIN[0] : 1e-13;
    for (int c = 0; c < mp->nComps; c++) {
        if (!mp->comps[c].val) {
            setError("comp " + std::to_string(c) + " of part " + std::to_string(pNum) + " has no values");
            return false;
        }
        double sum = 0.0;
        for (int j = 0; j < S; j++) {
            if (mp->comps[c].val[j] < pivecMin) {
                char b[160];
                snprintf(b, sizeof b, "p4_setPramsPart() part %d, comp %d, value %d is %g   Bad.", pNum, c, j,
                         mp->comps[c].val[j]);
                setError(b);
                return false;
            }
            sum += mp->comps[c].val[j];
        }
        if (std::fabs(sum - 1.0) > 1e-14) {
            char b[160];
            snprintf(b, sizeof b, "p4_setPramsPart() part %d, comp %d, values do not sum to 1.0.  sum=%.17g",
                     pNum, c, sum);
            setError(b);
            return false;
        }
    }
    if (mp->bQETneedsReset)
        for (int i = 0; i < mp->nComps * mp->nRMatrices; i++) mp->bQETneedsReset[i] = 1;
    std::vector<char> used(mp->nComps * mp->nRMatrices, 0);
    std::vector<int> slots, nodes;
    for (int n = 0; n < t->nNodes; n++) {
        const Node *nd = t->nodes[n];
        if (n == t->root) continue;
        if (!checkNodeModel(t, mp, nd, pNum)) return false;
        const int slot = nd->compNums[pNum] * mp->nRMatrices + nd->rMatrixNums[pNum];
        if (!used[slot]) {
            used[slot] = 1;
            slots.push_back(slot);
        }
        nodes.push_back(n);
    }
    for (int slot : slots) markEig(t, pNum, slot);
    if (mp->bQETneedsReset)
        for (int slot : slots) mp->bQETneedsReset[slot] = 0;
    for (int n : nodes) markP(t, pNum, t->nodes[n]);
    return true;
}

// P of one node (Pf/p4_node.c:296-346), recomputing a stale eigensystem
// first as the reference does (Pf/p4_node.c:310-315).
static bool bigPDecksPart(Tree *t, Node *nd, int pNum)
{
    if (nd->nodeNum == t->root) return true;
    ModelPart *mp = t->model->parts[pNum];
    if (!checkNodeModel(t, mp, nd, pNum)) return false;
    const int slot = nd->compNums[pNum] * mp->nRMatrices + nd->rMatrixNums[pNum];
    if (mp->bQETneedsReset && mp->bQETneedsReset[slot]) {
        markEig(t, pNum, slot);
        mp->bQETneedsReset[slot] = 0;
    }
    markP(t, pNum, nd);
    return true;
}

// Build the node program for part pNum: all internal nodes in postorder
// (full) or only the marked ones, always ending with the root.
// The nodes to (re)compute, in postorder, with their children: every
// non-leaf node when `full`, else the dirty ones; the root last.
static bool collectOps(Tree *t, int pNum, bool full, std::vector<OpRec> &ops, std::vector<KidRec> &kids)
{
    TreePart &tp = t->tp[pNum];
    std::vector<char> have(t->nNodes, 0);
    for (int n = 0; n < t->nNodes; n++) have[n] = (!t->nodes[n]->isLeaf && tp.everComputed && !tp.dirty[n]);
    for (int j = 0; j < t->nNodes; j++) {
        const int n = t->postOrder[j];
        if (n == NO_ORDER) continue;
        if (n < 0 || n >= t->nNodes) {
            setError("postOrder holds an out-of-range node number");
            return false;
        }
        const Node *nd = t->nodes[n];
        const bool isRoot = (n == t->root);
        if (nd->isLeaf && !isRoot) continue;
        if (!isRoot && !full && tp.everComputed && !tp.dirty[n]) continue;
        OpRec op{isRoot ? -1 : t->slotOf[n], isRoot ? OP_ROOT : 0, 0, (int)kids.size()};
        if (isRoot && nd->isLeaf) op.dst = tp.nSlots - 1;
        int c = nd->leftChild;
        while (c >= 0) {
            if (c >= t->nNodes) {
                setError("node relation out of range");
                return false;
            }
            const Node *cn = t->nodes[c];
            KidRec k{};
            if (cn->isLeaf) {
                if (cn->seqNum < 0 || cn->seqNum >= t->data->nTax) {
                    setError("leaf node " + std::to_string(c) + " has no sequence");
                    return false;
                }
                k.kind = 0;
                k.idx = cn->seqNum;
            } else {
                if (!have[c]) {
                    setError("partials of node " + std::to_string(c) +
                             " are needed but were never computed (postorder or marks are incomplete)");
                    return false;
                }
                k.kind = 1;
                k.idx = t->slotOf[c];
            }
            if (!tp.pValid[c]) {
                setError("P matrices of node " + std::to_string(c) + " were never computed (call setPrams)");
                return false;
            }
            k.pnode = c;
            kids.push_back(k);
            op.nKids++;
            if (op.nKids > MAX_CHILDREN) {
                setError("too many children");
                return false;
            }
            c = cn->sibling;
        }
        if (op.nKids == 0) {
            setError("internal node " + std::to_string(n) + " has no children");
            return false;
        }
        if (!isRoot) have[n] = 1;
        ops.push_back(op);
        if (isRoot) break;
    }
    if (ops.empty() || !(ops.back().flags & OP_ROOT)) {
        setError("postOrder does not end at the root");
        return false;
    }
    return true;
}

static bool buildWaveProgram(Tree *t, int pNum, const std::vector<OpRec> &ops, const std::vector<KidRec> &kids);
static bool buildGroupedProgram(Tree *t, int pNum, const std::vector<OpRec> &ops, const std::vector<KidRec> &kids,
                                int *nOpsOut, int *nKidsOut);
static bool buildAaProgram(Tree *t, int pNum, const std::vector<OpRec> &ops, const std::vector<KidRec> &kids,
                           int *nOpsOut, int *nKidsOut);

// A 64-bit hash over everything the staged program depends on: one
// multiply-xorshift step per value (it runs for every part of every tree
// at every evaluation, so it has to be cheap).
static uint64_t programKey(const Tree *t, int pNum, bool full)
{
    const TreePart &tp = t->tp[pNum];
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](int64_t v) {
        h = (h ^ (uint64_t)v) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 29;
    };
    mix(full);
    mix(tp.everComputed);
    mix(tp.wave);
    mix(t->precision);
    mix(pruneWaveStreams(t->precision));
    mix(pruneWaveVectorP(t->precision));
    mix(pruneWaveLean(t->precision));
    mix(pruneGroupWidthFor(tp.dim, t->precision));
    mix(pruneAaLdsSlots(tp.K));
    mix(pruneWavePreloadSlots(tp.K, 1, t->data->nTax, 0, t->precision));
    mix(tp.layoutGen);
    mix(t->root);
    mix(t->nNodes);
    for (int j = 0; j < t->nNodes; j++) mix(t->postOrder[j]);
    for (int n = 0; n < t->nNodes; n++) {
        const Node *nd = t->nodes[n];
        mix(((int64_t)nd->parent << 32) ^ (int64_t)(uint32_t)nd->leftChild);
        mix(((int64_t)nd->sibling << 32) ^ (int64_t)(uint32_t)nd->seqNum);
        mix(((int64_t)nd->isLeaf << 2) | ((int64_t)(tp.dirty[n] != 0) << 1) | (int64_t)(tp.pValid[n] != 0));
    }
    return h;
}

static bool buildProgram(Tree *t, int pNum, bool full, int *nOpsOut, int *nKidsOut)
{
    TreePart &tp = t->tp[pNum];
    const uint64_t key = programKey(t, pNum, full);
    if (tp.progValid && tp.progKey == key) {
        // the staging area still holds this program
        *nOpsOut = tp.progNOps;
        *nKidsOut = tp.progNKids;
        return true;
    }
    tp.progValid = false;
    tp.progOnDevice = false;
    std::vector<OpRec> ops;
    std::vector<KidRec> kids;
    if (!collectOps(t, pNum, full, ops, kids)) return false;
    tp.progSlots.clear();
    for (const OpRec &op : ops)
        if (op.dst >= 0) tp.progSlots.push_back(op.dst);
    bool ok;
    tp.aaLds = 0;
    tp.aaRecOff = -1;
    if (tp.wave) {
        *nOpsOut = (int)ops.size();
        *nKidsOut = (int)kids.size();
        ok = buildWaveProgram(t, pNum, ops, kids);
    } else if (pruneUsesMfma(tp.dim, tp.K, t->precision) && pruneAaCategoryWaves((int)ops.size()) &&
               pruneAaLdsSlots(tp.K) > 0) {
        ok = buildAaProgram(t, pNum, ops, kids, nOpsOut, nKidsOut);
    } else {
        ok = buildGroupedProgram(t, pNum, ops, kids, nOpsOut, nKidsOut);
    }
    if (ok) {
        tp.progValid = true;
        tp.progKey = key;
        tp.progNOps = *nOpsOut;
        tp.progNKids = *nKidsOut;
    }
    return ok;
}

// Postorder from the root, children with the larger slot need first
// (Sethi-Ullman), so few results are live at once; nOps entries of
// progKids name each op's children computed in the same program.
static bool sethiUllmanOrder(const std::vector<OpRec> &ops, const std::vector<KidRec> &kids, int nSlots,
                             std::vector<int> &order, std::vector<int> &parentOp, std::vector<std::vector<int>> &progKids)
{
    const int nOps = (int)ops.size();
    std::vector<int> opOfSlot(nSlots, -1), need(nOps, 1);
    parentOp.assign(nOps, -1);
    progKids.assign(nOps, {});
    for (int i = 0; i < nOps; i++)
        if (!(ops[i].flags & OP_ROOT)) opOfSlot[ops[i].dst] = i;
    for (int i = 0; i < nOps; i++)
        for (int c = 0; c < ops[i].nKids; c++) {
            const KidRec &k = kids[ops[i].kidStart + c];
            if (k.kind == 1 && opOfSlot[k.idx] >= 0 && opOfSlot[k.idx] < i) {
                progKids[i].push_back(opOfSlot[k.idx]);
                parentOp[opOfSlot[k.idx]] = i;
            }
        }
    for (int i = 0; i < nOps; i++) {
        auto &pk = progKids[i];
        std::stable_sort(pk.begin(), pk.end(), [&](int x, int y) { return need[x] > need[y]; });
        int nd = 1;
        for (size_t q = 0; q < pk.size(); q++) nd = std::max(nd, need[pk[q]] + (int)q);
        need[i] = nd;
    }
    order.clear();
    std::vector<std::pair<int, size_t>> st;
    st.push_back({nOps - 1, 0});
    while (!st.empty()) {
        auto &top = st.back();
        if (top.second < progKids[top.first].size()) {
            const int c = progKids[top.first][top.second++];
            st.push_back({c, 0});
        } else {
            order.push_back(top.first);
            st.pop_back();
        }
    }
    if ((int)order.size() != nOps) {
        // ops whose parent is not in the program run first, in postorder
        std::vector<char> seen(nOps, 0);
        for (int i : order) seen[i] = 1;
        std::vector<int> rest;
        for (int i = 0; i < nOps; i++)
            if (!seen[i]) rest.push_back(i);
        order.insert(order.begin(), rest.begin(), rest.end());
        std::vector<int> pos(nOps);
        for (int q = 0; q < nOps; q++) pos[order[q]] = q;
        for (int i = 0; i < nOps; i++)
            for (int c : progKids[i])
                if (pos[c] > pos[i]) {
                    setError("internal: program order");
                    return false;
                }
    }
    return true;
}

// Fused op records of a 20-state program (AaRec), appended after its kid
// records: the category kernel reads one per op.
static bool appendAaRecs(TreePart &tp, const std::vector<OpRec> &ops2, const std::vector<KidRec> &kids2,
                         size_t *words)
{
    const size_t need = *words + ops2.size() * AA_REC_WORDS;
    if ((int)need > tp.progCapWords) {
        setError("internal: node program larger than its staging area");
        return false;
    }
    int *r = tp.hProg + *words;
    for (size_t i = 0; i < ops2.size(); i++, r += AA_REC_WORDS) {
        const OpRec &o = ops2[i];
        std::memset(r, 0, sizeof(int) * AA_REC_WORDS);
        r[0] = o.dst;
        r[1] = o.flags;
        r[2] = o.nKids;
        r[3] = o.kidStart;
        for (int c = 0; c < 2 && c < o.nKids; c++) {
            const KidRec &k = kids2[o.kidStart + c];
            r[4 + 4 * c] = k.kind;
            r[5 + 4 * c] = k.idx;
            r[6 + 4 * c] = k.pnode;
            r[7 + 4 * c] = k.pad;
        }
    }
    tp.aaRecOff = (int)*words;
    *words = need;
    return true;
}

// The 20-state category kernel's program: the ops in Sethi-Ullman order,
// one group each; a result whose parent is in the program also goes to an
// LDS slot (op flags bits 8.. = slot + 1), which the parent reads as kid
// kind 2 (KidRec.pad = the slot; idx still names the partials slot).
// Results that find no free slot are read back from HBM.
static bool buildAaProgram(Tree *t, int pNum, const std::vector<OpRec> &ops, const std::vector<KidRec> &kids,
                           int *nOpsOut, int *nKidsOut)
{
    TreePart &tp = t->tp[pNum];
    const int nOps = (int)ops.size();
    std::vector<int> order, parentOp;
    std::vector<std::vector<int>> progKids;
    if (!sethiUllmanOrder(ops, kids, tp.nSlots, order, parentOp, progKids)) return false;
    const int maxSlots = pruneAaLdsSlots(tp.K);
    std::vector<int> ldsOf(nOps, -1), freeList;
    for (int sl = maxSlots - 1; sl >= 0; sl--) freeList.push_back(sl);
    int used = 0, peak = 0;
    for (int i : order) {
        for (int c : progKids[i])
            if (ldsOf[c] >= 0) {
                freeList.push_back(ldsOf[c]);
                used--;
            }
        if (parentOp[i] >= 0 && !freeList.empty()) {
            ldsOf[i] = freeList.back();
            freeList.pop_back();
            used++;
            peak = std::max(peak, used);
        }
    }
    std::vector<int> opOfSlot(tp.nSlots, -1);
    for (int i = 0; i < nOps; i++)
        if (!(ops[i].flags & OP_ROOT)) opOfSlot[ops[i].dst] = i;
    std::vector<GroupRec> groups;
    std::vector<OpRec> ops2;
    std::vector<KidRec> kids2;
    for (int i : order) {
        groups.push_back(GroupRec{1, 1, (int)ops2.size(), 0});
        OpRec o = ops[i];
        o.flags = (o.flags & 0xff) | ((ldsOf[i] + 1) << 8);
        o.kidStart = (int)kids2.size();
        for (int c = 0; c < ops[i].nKids; c++) {
            KidRec k = kids[ops[i].kidStart + c];
            if (k.kind == 1) {
                const int j = opOfSlot[k.idx];
                if (j >= 0 && parentOp[j] == i && ldsOf[j] >= 0) {
                    k.kind = 2;
                    k.pad = ldsOf[j];
                }
            }
            kids2.push_back(k);
        }
        ops2.push_back(o);
    }
    const size_t words = groups.size() * 4 + ops2.size() * 4 + kids2.size() * 4;
    if ((int)words > tp.progCapWords) {
        setError("internal: node program larger than its staging area");
        return false;
    }
    std::memcpy(tp.hProg, groups.data(), groups.size() * sizeof(GroupRec));
    std::memcpy(tp.hProg + groups.size() * 4, ops2.data(), ops2.size() * sizeof(OpRec));
    std::memcpy(tp.hProg + (groups.size() + ops2.size()) * 4, kids2.data(), kids2.size() * sizeof(KidRec));
    size_t allWords = words;
    if (!appendAaRecs(tp, ops2, kids2, &allWords)) return false;
    tp.nGroups = (int)groups.size();
    tp.nKids = (int)kids2.size();
    tp.progWords = (int)allWords;
    tp.aaLds = peak;
    *nOpsOut = (int)ops2.size();
    *nKidsOut = (int)kids2.size();
    return true;
}

static bool buildGroupedProgram(Tree *t, int pNum, const std::vector<OpRec> &ops, const std::vector<KidRec> &kids,
                                int *nOpsOut, int *nKidsOut)
{
    TreePart &tp = t->tp[pNum];
    // Group ops by height above the nodes already computed: ops of equal
    // height are independent, so binary ones are packed G to a group.
    const int G = pruneGroupWidthFor(t->model->parts[pNum]->dim, t->precision);
    const int nOps = (int)ops.size();
    std::vector<int> opOfSlot(tp.nSlots, -1), height(nOps, 0);
    for (int i = 0; i < nOps; i++) {
        int h = 0;
        for (int c = 0; c < ops[i].nKids; c++) {
            const KidRec &k = kids[ops[i].kidStart + c];
            if (k.kind == 1 && opOfSlot[k.idx] >= 0) h = std::max(h, height[opOfSlot[k.idx]] + 1);
        }
        height[i] = h;
        if (!(ops[i].flags & OP_ROOT)) opOfSlot[ops[i].dst] = i;
    }
    std::vector<int> order(nOps - 1);
    for (int i = 0; i < nOps - 1; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return height[x] < height[y]; });
    order.push_back(nOps - 1);  // the root last
    std::vector<OpRec> ops2;
    std::vector<KidRec> kids2;
    std::vector<GroupRec> groups;
    for (size_t q = 0; q < order.size();) {
        const int i = order[q];
        const bool binary = ops[i].nKids == 2 && !(ops[i].flags & OP_ROOT);
        GroupRec gr{binary ? 0 : 1, 0, (int)ops2.size(), 0};
        size_t r = q;
        while (r < order.size()) {
            const int j = order[r];
            const bool bj = ops[j].nKids == 2 && !(ops[j].flags & OP_ROOT);
            if (gr.nOps > 0 && (!binary || !bj || height[j] != height[i] || gr.nOps == G)) break;
            OpRec o = ops[j];
            o.kidStart = (int)kids2.size();
            for (int c = 0; c < ops[j].nKids; c++) kids2.push_back(kids[ops[j].kidStart + c]);
            ops2.push_back(o);
            gr.nOps++;
            r++;
            if (!binary) break;
        }
        groups.push_back(gr);
        q = r;
    }
    const size_t words = groups.size() * 4 + ops2.size() * 4 + kids2.size() * 4;
    if ((int)words > tp.progCapWords) {
        setError("internal: node program larger than its staging area");
        return false;
    }
    std::memcpy(tp.hProg, groups.data(), groups.size() * sizeof(GroupRec));
    std::memcpy(tp.hProg + groups.size() * 4, ops2.data(), ops2.size() * sizeof(OpRec));
    std::memcpy(tp.hProg + (groups.size() + ops2.size()) * 4, kids2.data(), kids2.size() * sizeof(KidRec));
    size_t allWords = words;
    if (pruneUsesMfma(tp.dim, tp.K, t->precision) && !appendAaRecs(tp, ops2, kids2, &allWords)) return false;
    tp.nGroups = (int)groups.size();
    tp.nKids = (int)kids2.size();
    tp.progWords = (int)allWords;
    *nOpsOut = (int)ops2.size();
    *nKidsOut = (int)kids2.size();
    return true;
}

// Program of the wave kernel.  Order: from the root, children with the
// larger slot need first (Sethi-Ullman), so few results are live at once.
// A result goes to an LDS slot when its parent is in the program; the slot
// is free again once the parent has read it.  Clean children are preloaded
// into the slots the results never need; the rest are read from HBM.
static bool buildWaveProgram(Tree *t, int pNum, const std::vector<OpRec> &ops, const std::vector<KidRec> &kids)
{
    TreePart &tp = t->tp[pNum];
    const Data *data = t->data;
    const int nOps = (int)ops.size();
    std::vector<int> opOfSlot(tp.nSlots, -1), parentOp(nOps, -1), need(nOps, 1);
    for (int i = 0; i < nOps; i++)
        if (!(ops[i].flags & OP_ROOT)) opOfSlot[ops[i].dst] = i;
    std::vector<std::vector<int>> progKids(nOps);
    for (int i = 0; i < nOps; i++)
        for (int c = 0; c < ops[i].nKids; c++) {
            const KidRec &k = kids[ops[i].kidStart + c];
            if (k.kind == 1 && opOfSlot[k.idx] >= 0 && opOfSlot[k.idx] < i) {
                progKids[i].push_back(opOfSlot[k.idx]);
                parentOp[opOfSlot[k.idx]] = i;
            }
        }
    for (int i = 0; i < nOps; i++) {  // postorder: children first
        auto &pk = progKids[i];
        std::stable_sort(pk.begin(), pk.end(), [&](int x, int y) { return need[x] > need[y]; });
        int nd = 1;
        for (size_t q = 0; q < pk.size(); q++) nd = std::max(nd, need[pk[q]] + (int)q);
        need[i] = nd;
    }
    // depth-first emission from the root (iterative)
    std::vector<int> order;
    order.reserve(nOps);
    {
        std::vector<std::pair<int, size_t>> st;
        st.push_back({nOps - 1, 0});
        while (!st.empty()) {
            auto &top = st.back();
            if (top.second < progKids[top.first].size()) {
                const int c = progKids[top.first][top.second++];
                st.push_back({c, 0});
            } else {
                order.push_back(top.first);
                st.pop_back();
            }
        }
    }
    // ops whose parent is not in the program (possible when a node is
    // marked without its ancestors) run first, in postorder
    if ((int)order.size() != nOps) {
        std::vector<char> seen(nOps, 0);
        for (int i : order) seen[i] = 1;
        std::vector<int> rest;
        for (int i = 0; i < nOps; i++)
            if (!seen[i]) rest.push_back(i);
        // a stray op may itself have program children: keep postorder
        order.insert(order.begin(), rest.begin(), rest.end());
        std::vector<int> pos(nOps);
        for (int q = 0; q < nOps; q++) pos[order[q]] = q;
        for (int i = 0; i < nOps; i++)
            for (int c : progKids[i])
                if (pos[c] > pos[i]) {
                    setError("internal: wave program order");
                    return false;
                }
    }
    const int tipLds = data->nTax <= WAVE_TIP_LDS_MAX_TAXA ? 1 : 0;
    std::vector<int> pos(nOps);
    for (int q = 0; q < nOps; q++) pos[order[q]] = q;
    // Streams: with two, the subtrees under the root are split between two
    // sets of category waves that run at the same time (half the node chain
    // per wave); the root runs after a barrier on stream 0, reading stream
    // 1's root children from junction slots in stream 0's slot file.
    const int rootOp = order.back();
    const int nStream = (pruneWaveStreams(t->precision) && tp.K * 2 <= WAVE_MAX_K && nOps > 3) ? 2 : 1;
    // The lean kernel stages the P^T blocks of every tip child of the
    // program in LDS when they fit beside a few slots; otherwise its ops
    // with tip children take the generic path.
    int nTipKid = 0;
    if (pruneWaveLean(t->precision) && tipLds && nStream == 1) {
        for (int i = 0; i < nOps; i++)
            for (int c = 0; c < ops[i].nKids; c++)
                if (kids[ops[i].kidStart + c].kind == 0) nTipKid++;
        if (pruneWaveMaxSlots(tp.K, tipLds, data->nTax, nTipKid, t->precision) < 4) nTipKid = 0;
    }
    const int nStaged = nTipKid;
    // LDS budget shared by all slot files of the workgroup
    const int maxSlotsS = pruneWaveMaxSlots(tp.K * nStream, tipLds, data->nTax, nStaged, t->precision);
    std::vector<int> top(nOps, -1);
    for (int q = nOps - 1; q >= 0; q--) {
        const int i = order[q];
        if (i == rootOp) continue;
        const int pa = parentOp[i];
        top[i] = (pa < 0 || pa == rootOp) ? i : top[pa];
    }
    std::vector<int> streamOf(nOps, 0);
    if (nStream == 2) {
        std::vector<int> size(nOps, 0), tops;
        for (int i = 0; i < nOps; i++)
            if (i != rootOp) {
                if (size[top[i]]++ == 0) tops.push_back(top[i]);
            }
        std::stable_sort(tops.begin(), tops.end(), [&](int x, int y) { return size[x] > size[y]; });
        int load[2] = {0, 0};
        std::vector<int> streamOfTop(nOps, 0);
        for (int tp_ : tops) {
            const int st = load[1] < load[0] ? 1 : 0;
            streamOfTop[tp_] = st;
            load[st] += size[tp_];
        }
        for (int i = 0; i < nOps; i++)
            if (i != rootOp) streamOf[i] = streamOfTop[top[i]];
    }
    std::vector<int> lists[2];
    for (int q = 0; q < nOps; q++) {
        const int i = order[q];
        if (i != rootOp) lists[streamOf[i]].push_back(i);
    }
    // slot use of each stream's results; stream 1's root children are left
    // for the junction
    std::vector<int> ldsOf(nOps, -1);
    int peak[2] = {0, 0};
    for (int st = 0; st < nStream; st++) {
        std::vector<int> freeList;
        for (int sl = maxSlotsS - 1; sl >= 0; sl--) freeList.push_back(sl);
        int used = 0;
        for (int i : lists[st]) {
            for (int c : progKids[i])
                if (ldsOf[c] >= 0) {
                    freeList.push_back(ldsOf[c]);
                    used--;
                }
            const bool junction = st == 1 && parentOp[i] == rootOp;
            if (parentOp[i] >= 0 && !junction && !freeList.empty()) {
                ldsOf[i] = freeList.back();
                freeList.pop_back();
                used++;
                peak[st] = std::max(peak[st], used);
            }
        }
    }
    // junction slots follow stream 0's results: [peak0, peak0 + J)
    int nJ = 0;
    if (nStream == 2)
        for (int i : lists[1])
            if (parentOp[i] == rootOp && peak[0] + nJ < maxSlotsS) {
                ldsOf[i] = -(2 + peak[0] + nJ);   // written into stream 0's file
                nJ++;
            }
    // clean internal children, by consumer, per stream (the root's: stream 0)
    struct Pre {
        int st, op, c, cl;
    };
    std::vector<Pre> cleans;
    for (int st = 0; st < nStream; st++) {
        std::vector<int> lst = lists[st];
        if (st == 0) lst.push_back(rootOp);
        for (int i : lst)
            for (int c = 0; c < ops[i].nKids; c++) {
                const KidRec &k = kids[ops[i].kidStart + c];
                if (k.kind == 1 && !(opOfSlot[k.idx] >= 0 && opOfSlot[k.idx] < i)) cleans.push_back({st, i, c, k.idx});
            }
    }
    // every preload is loaded at the start, so each takes a slot of its own
    // above the results (and, in stream 0, the junction)
    int nextPre[2] = {peak[0] + nJ, peak[1]};
    std::vector<int> preSlot(cleans.size(), -1);
    int nPre = 0;
    const int preCap = pruneWavePreloadSlots(tp.K * nStream, tipLds, data->nTax, nStaged, t->precision);
    for (size_t j = 0; j < cleans.size(); j++) {
        const int st = cleans[j].st;
        if (nextPre[st] < std::min(maxSlotsS, preCap)) {
            preSlot[j] = nextPre[st]++;
            nPre++;
        }
    }
    const int nLds = std::max(nextPre[0], nStream == 2 ? nextPre[1] : 0);
    std::vector<std::vector<int>> preOf(nOps);
    for (int i = 0; i < nOps; i++) preOf[i].assign(ops[i].nKids, -1);
    for (size_t j = 0; j < cleans.size(); j++) preOf[cleans[j].op][cleans[j].c] = preSlot[j];
    // emission order: stream 0, stream 1, the root
    std::vector<int> emit = lists[0];
    emit.insert(emit.end(), lists[1].begin(), lists[1].end());
    emit.push_back(rootOp);
    for (int q = 0; q < nOps; q++) pos[emit[q]] = q;
    int nSide = 0;
    for (int i = 0; i < nOps; i++) nSide += std::max(0, ops[i].nKids - 3);
    int nKidTot = 0;
    for (int i = 0; i < nOps; i++) nKidTot += ops[i].nKids;
    const size_t wordsW = (size_t)16 * nOps + 4 * (size_t)nSide + 3 * (size_t)nPre + (size_t)nKidTot;
    const size_t popsOff = (wordsW + 15) & ~(size_t)15;   // LeanOps, 64-byte aligned
    const int recInts = leanRecInts(t->precision == FP32);
    const size_t tipPnOff = popsOff + (size_t)recInts * nOps;
    const size_t words = tipPnOff + (size_t)nStaged;
    if ((int)words > tp.progCapWords) {
        setError("internal: node program larger than its staging area");
        return false;
    }
    WaveOp *wo = reinterpret_cast<WaveOp *>(tp.hProg);
    WaveKid *wk = reinterpret_cast<WaveKid *>(tp.hProg + 16 * nOps);
    int *wp = tp.hProg + 16 * nOps + 4 * nSide;
    int *kidP = wp + 3 * nPre;
    int side = 0, kb = 0;
    for (int q = 0; q < nOps; q++) {
        const int i = emit[q];
        WaveOp w{};
        w.dst = (ops[i].flags & OP_ROOT) ? -1 : ops[i].dst;
        w.ldsDst = ldsOf[i];
        w.flags = ops[i].flags;
        w.nKids = ops[i].nKids;
        w.kidStart = side;
        w.kidBase = kb;
        for (int c = 0; c < ops[i].nKids; c++) {
            const KidRec &k = kids[ops[i].kidStart + c];
            int kind, idx;
            if (k.kind == 0) {
                kind = WK_TIP;
                idx = k.idx;
            } else {
                const int j = opOfSlot[k.idx];
                if (j >= 0 && j < i && pos[j] < q) {
                    if (ldsOf[j] >= 0) {
                        kind = WK_LDS;
                        idx = ldsOf[j];
                    } else if (ldsOf[j] <= -2) {   // junction: the root reads its own file
                        kind = WK_LDS;
                        idx = -ldsOf[j] - 2;
                    } else {
                        kind = WK_GLOBAL;
                        idx = k.idx;
                    }
                } else if (preOf[i][c] >= 0) {
                    kind = WK_LDS;
                    idx = preOf[i][c];
                } else {
                    kind = WK_GLOBAL;
                    idx = k.idx;
                }
            }
            kidP[kb++] = k.pnode;
            if (c < 3) {
                w.kind[c] = kind;
                w.idx[c] = idx;
                w.pnode[c] = k.pnode;
            } else {
                wk[side++] = WaveKid{kind, idx, k.pnode, 0};
            }
        }
        wo[q] = w;
    }
    int np = 0;
    for (size_t j = 0; j < cleans.size(); j++)
        if (preSlot[j] >= 0) {
            wp[3 * np] = cleans[j].cl;
            wp[3 * np + 1] = preSlot[j];
            wp[3 * np + 2] = cleans[j].st;
            np++;
        }
    // the lean kernel's records: precomputed offsets for binary ops whose
    // children are staged tips or LDS slots; the staged tips' P^T nodes
    {
        const size_t elt = t->precision == FP32 ? sizeof(float) : sizeof(double);
        const size_t blkB = 16 * elt;
        int *lo = tp.hProg + popsOff;
        int *tipPn = tp.hProg + tipPnOff;
        int nt = 0;
        for (int q = 0; q < nOps; q++) {
            const WaveOp &w = wo[q];
            LeanOp r{};
            r.kindLd = PK_GEN;
            // tip children: staged P^T rows in LDS, or (no room to stage
            // them: FP64) scalar P^T and the select tree; both need the
            // tile's tip codes in LDS
            const bool tipsOk = tipLds != 0;
            const bool fast = w.nKids == 2 && !(w.flags & OP_ROOT) && w.ldsDst >= -1 &&
                              ((w.kind[0] == WK_TIP && tipsOk) || w.kind[0] == WK_LDS) &&
                              ((w.kind[1] == WK_TIP && tipsOk) || w.kind[1] == WK_LDS);
            if (fast) {
                int c0 = 0, c1 = 1;
                if (w.kind[0] != WK_TIP && w.kind[1] == WK_TIP) std::swap(c0, c1);   // tips first
                const int nTip = (w.kind[0] == WK_TIP) + (w.kind[1] == WK_TIP);
                const int kind = nTip == 2 ? PK_TT : nTip == 1 ? PK_TL : PK_LL;
                const int cs[2] = {c0, c1};
                int *aa[2] = {&r.a0, &r.a1}, *tt[2] = {&r.t0, &r.t1}, *pp[2] = {&r.p0, &r.p1};
                for (int u = 0; u < 2; u++) {
                    const int c = cs[u];
                    if (w.kind[c] == WK_TIP) {
                        *aa[u] = w.idx[c] * 64;
                        const size_t pb = (size_t)w.pnode[c] * tp.K * blkB;
                        if (pb > (size_t)INT32_MAX) {
                            setError("internal: P table too large for 32-bit offsets");
                            return false;
                        }
                        *pp[u] = (int)pb;
                        if (nTipKid > 0) {
                            if (nt >= nTipKid) {
                                setError("internal: staged tip count");
                                return false;
                            }
                            *tt[u] = (int)(nt * blkB);
                            tipPn[nt++] = w.pnode[c];
                        }
                    } else {
                        *aa[u] = w.idx[c];
                        *tt[u] = w.idx[c] * 64;
                        const size_t pb = (size_t)w.pnode[c] * tp.K * blkB;
                        if (pb > (size_t)INT32_MAX) {
                            setError("internal: P table too large for 32-bit offsets");
                            return false;
                        }
                        *pp[u] = (int)pb;
                    }
                }
                r.dst = w.dst;
                r.kindLd = kind | ((w.ldsDst >= 0 ? w.ldsDst + 1 : 0) << 8);
            }
            int *rq = lo + (size_t)recInts * q;
            std::memcpy(rq, &r, sizeof r);
            for (int z = (int)(sizeof r / sizeof(int)); z < recInts; z++) rq[z] = 0;
        }
        // tips that only generic ops read are staged too (unused, harmless)
        for (; nt < nTipKid; nt++) tipPn[nt] = 0;
    }
    tp.popsOff = (int)popsOff;
    tp.tipPnOff = (int)tipPnOff;
    tp.nTipKid = nStaged;
    tp.nWOps = nOps;
    tp.nWKids = nSide;
    tp.nPre = nPre;
    tp.nLds = nLds;
    tp.tipLds = tipLds;
    tp.nKidTot = nKidTot;
    tp.nStream = nStream;
    tp.nS0 = (int)lists[0].size();
    tp.waveLds = pruneWaveLdsBytes(tp.K * nStream, nLds, tipLds, data->nTax, nStaged, t->precision);
    tp.progWords = (int)words;
    tp.nGroups = 0;
    tp.nKids = (int)kids.size();
    return true;
}

static void fillPruneArgs(Tree *t, int pNum, int nOps, PruneArgs &a, bool siteLikes)
{
    TreePart &tp = t->tp[pNum];
    ModelPart *mp = t->model->parts[pNum];
    Part *dp = t->data->parts[pNum];
    const Node *root = t->nodes[t->root];
    const int *prog = reinterpret_cast<const int *>(tp.dArena + tp.offProg);
    a.ptWarmBytes = 0;
    a.groups = reinterpret_cast<const GroupRec *>(prog);
    a.ops = reinterpret_cast<const OpRec *>(prog + tp.nGroups * 4);
    a.kids = reinterpret_cast<const KidRec *>(prog + (tp.nGroups + nOps) * 4);
    a.nGroups = tp.nGroups;
    a.nOps = nOps;
    a.nKids = tp.nKids;
    a.dim = mp->dim;
    a.K = mp->nCat;
    a.log2Kp = log2ceil(mp->nCat);
    a.nPat = tp.nPat;
    a.nPatPad = tp.nPatPad;
    a.tips = dp->dev.tips;
    a.PT = t->precision == FP32 ? (const void *)tp.PT32 : (const void *)tp.PT64;
    a.cl = tp.cl;
    a.scl = tp.scl;
    a.counts = dp->dev.counts;
    a.invMask = dp->dev.invMask;
    a.nWords = dp->dev.nWords;
    a.ambMask = dp->dev.ambMask;
    a.nAmb = dp->dev.nAmb;
    a.rootPi = reinterpret_cast<const double *>(tp.dArena + tp.offPi);
    a.pInvar = mp->pInvar;
    a.rootLeafRow = root->isLeaf ? root->seqNum : -1;
    a.blockSums = tp.blockSums;
    a.blockBad = tp.blockBad;
    a.counter = tp.counter;
    a.result = tp.dResultHost;
    a.patLikes = siteLikes ? tp.patLikes : nullptr;
    a.nLds = tp.aaLds;   // the 20-state category program's slot file (the wave kernels set theirs below)
    a.nTax = dp->nTax;
    a.tipLds = pruneAaTipLds(dp->nTax) ? 1 : 0;   // the 20-state category kernel's tip codes in LDS
    a.aaRec = tp.aaRecOff >= 0 ? prog + tp.aaRecOff : nullptr;
    if (tp.wave) {
        a.wops = reinterpret_cast<const WaveOp *>(prog);
        a.wkids = reinterpret_cast<const WaveKid *>(prog + 16 * tp.nWOps);
        a.wpre = prog + 16 * tp.nWOps + 4 * tp.nWKids;
        a.nStream = tp.nStream;
        a.nS0 = tp.nS0;
        a.nWOps = tp.nWOps;
        a.nPre = tp.nPre;
        a.nLds = tp.nLds;
        a.tipLds = tp.tipLds;
        a.nTax = dp->nTax;
        a.PP = tp.PP;
        a.nKidTot = tp.nKidTot;
        a.pops = reinterpret_cast<const LeanOp *>(prog + tp.popsOff);
        a.nTipKid = tp.nTipKid;
        a.tipPn = prog + tp.tipPnOff;
        // the P^T table, pulled into the L2 of every XCD the launch runs on
        // while the tiles stage their tips (the scalar loads of the op chain
        // then hit L2); skipped for tables that would crowd it
        const size_t ptB = (size_t)t->nNodes * tp.K * 16 * (t->precision == FP32 ? sizeof(float) : sizeof(double));
        a.ptWarmBytes = ptB <= LEAN_PT_WARM_MAX ? (int)ptB : 0;
    }
}

// Stage the program and root frequencies of one part (async).
// With `uploads`, the arena range the device lacks is not copied here but
// listed for one copy launch that covers every item of a batched call.
static bool stagePart(Tree *t, int pNum, bool full, int *nOps, std::vector<DeferredP> *defer = nullptr,
                      std::vector<CopyRange> *uploads = nullptr)
{
    TreePart &tp = t->tp[pNum];
    ModelPart *mp = t->model->parts[pNum];
    const Node *root = t->nodes[t->root];
    int nKids = 0;
    if (!buildProgram(t, pNum, full, nOps, &nKids)) return false;
    for (int sl : tp.progSlots)
        if ((size_t)sl < tp.chSlot.size()) tp.chSlot[sl] = 1;
    const int rc = root->compNums[pNum];
    if (rc < 0 || rc >= mp->nComps || !mp->comps[rc].val) {
        setError("root comp out of range");
        return false;
    }
    std::memcpy(tp.hRootPi, mp->comps[rc].val, sizeof(double) * mp->dim);
    // what the device copy of the arena lacks -- changed root frequencies, a
    // rebuilt program -- goes up in one copy of the covering range (none when
    // nothing changed); then the parameter kernels (which read the pinned
    // arena directly); the pruning launch follows on the stream
    stageParams(t, pNum);
    (void)nKids;
    size_t lo = tp.arenaBytes, hi = 0;
    auto need = [&](size_t a, size_t b) {
        lo = std::min(lo, a);
        hi = std::max(hi, b);
    };
    const size_t piBytes = sizeof(double) * mp->dim;
    if (!tp.rootPiOnDevice || std::memcmp(tp.rootPiUp, tp.hRootPi, piBytes) != 0) need(tp.offPi, tp.offPi + piBytes);
    if (!tp.progOnDevice) need(tp.offProg, tp.offProg + sizeof(int) * tp.progWords);
    // 20-state P straight from exp(Qt) for many matrices: its (pi, R)
    // snapshots and the pack (the arena's first bytes) go up in the same
    // copy (a few matrices read them across PCIe sooner than a copy runs)
    const bool devEig = tp.flushDirect && mp->dim == 20 && tp.pendI * mp->nCat >= 64;
    if (devEig) need(0, tp.offPack + tp.packBytes);   // snapshots, eigen list and the pack
    if (hi > lo) {
        if (uploads)
            uploads->push_back({tp.dArenaHost + lo, tp.dArena + lo, (unsigned long long)(hi - lo)});
        else if (!hipOk(hipMemcpyAsync(tp.dArena + lo, tp.hArena + lo, hi - lo, hipMemcpyHostToDevice, stream()),
                        "stage inputs"))
            return false;
        std::memcpy(tp.rootPiUp, tp.hRootPi, piBytes);
        tp.rootPiOnDevice = true;
        tp.progOnDevice = true;
    }
    return launchParams(t, pNum, defer, devEig);
}

static void finishPart(Tree *t, int pNum, bool siteLikes, double lnL, bool bad)
{
    TreePart &tp = t->tp[pNum];
    std::fill(tp.dirty.begin(), tp.dirty.end(), 0);
    tp.everComputed = true;
    if (t->partLikes) t->partLikes[pNum] = lnL;
    (void)siteLikes;
    (void)bad;
}

// Evaluate pending partials + site sum of one part (one launch).  With
// `scr` (a device address of the no-allocation scratch array) the pruning
// kernel's reduction scratch is taken from there, the counter is cleared by
// a fill kernel first, and the wave kernel reads P^T by node (no staged
// program-order P stream).
// With `cpoSum` (device, one double per pattern) 1/like of every pattern is
// added to it after the pruning kernel, on the same stream.
static bool evalPart(Tree *t, int pNum, bool full, bool siteLikes, double *lnL, unsigned char *scr = nullptr,
                     double *cpoSum = nullptr)
{
    TreePart &tp = t->tp[pNum];
    Part *dp = t->data->parts[pNum];
    int nOps = 0;
    if (!stagePart(t, pNum, full, &nOps)) return false;
    PruneArgs a{};
    fillPruneArgs(t, pNum, nOps, a, siteLikes || cpoSum);
    hipStream_t s = stream();
    const int nBlocks = tp.wave ? tp.nPatPad / pruneWaveTilePatterns(t->precision)
                                : tp.nPatPad / pruneBlockPatternsFor(a.dim, a.K, t->precision, nOps);
    if (scr) {
        // [counter, 256 B][flags int[nBlocks]][sums double[nBlocks]]
        const size_t offBad = 256, offSums = offBad + (((size_t)nBlocks * sizeof(int) + 255) & ~(size_t)255);
        const size_t need = offSums + (size_t)nBlocks * sizeof(double);
        if (need > NOALLOC_SCRATCH_BYTES) {
            (void)hipStreamSynchronize(s);  // queued P kernels read the staging arena
            setError("treeLogLikeNoAlloc: part " + std::to_string(pNum) + " needs " + std::to_string(need) +
                     " bytes of scratch, more than the " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                     " the library holds (" + std::to_string(tp.nPat) + " patterns); use p4g_treeLogLike");
            return false;
        }
        a.counter = reinterpret_cast<unsigned *>(scr);
        a.blockBad = reinterpret_cast<int *>(scr + offBad);
        a.blockSums = reinterpret_cast<double *>(scr + offSums);
        a.PP = nullptr;
        if (!hipOk(launchFillU32(a.counter, 0u, 1, s), "clear scratch counter")) {
            (void)hipStreamSynchronize(s);
            return false;
        }
    }
    if (tp.wave && pruneWaveVectorP(t->precision) && a.PP) {
        const int *kidP = a.wpre + 3 * tp.nPre;
        if (!hipOk(launchPgather(a.PT, kidP, tp.nKidTot, tp.K, tp.PP, t->precision, s), "launch P gather"))
            return false;
    }
    if (g_timing) (void)hipEventRecord(tp.ev0, s);
    const hipError_t le = tp.wave ? launchPruneWave(a, t->precision, nBlocks, tp.waveLds, s)
                          : pruneUsesMfma(a.dim, a.K, t->precision) ? launchPruneAa(a, nBlocks, s)
                                                                    : launchPrune(a, t->precision, nBlocks, s);
    if (!hipOk(le, "launch pruning kernel")) {
        (void)hipStreamSynchronize(s);  // queued P kernels read the staging arena
        return false;
    }
    if (g_timing) (void)hipEventRecord(tp.ev1, s);
    if (cpoSum && !hipOk(launchCpoAccum(tp.patLikes, cpoSum, tp.nPat, s), "launch CPO sums")) {
        (void)hipStreamSynchronize(s);
        return false;
    }
    std::vector<double> pl;
    if (siteLikes) {
        pl.resize(tp.nPat);
        if (!hipOk(hipMemcpyAsync(pl.data(), tp.patLikes, sizeof(double) * tp.nPat,
                                  hipMemcpyDeviceToHost, s),
                   "copy site likes")) {
            (void)hipStreamSynchronize(s);
            return false;
        }
    }
    if (!hipOk(hipStreamSynchronize(s), "pruning kernel")) return false;
    if (g_timing) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, tp.ev0, tp.ev1) == hipSuccess) tp.lastKernelUs = ms * 1000.0;
    }
    const bool bad = tp.hResult[1] != 0.0;
    *lnL = tp.hResult[0];
    if (siteLikes && !bad) {
        dp->siteLikes.assign(dp->nChar, 0.0);
        if ((int)dp->seqPosPatternIndex.size() == dp->nChar)
            for (int i = 0; i < dp->nChar; i++) dp->siteLikes[i] = pl[dp->seqPosPatternIndex[i]];
        else {
            dp->siteLikes.assign(pl.begin(), pl.end());
        }
    }
    finishPart(t, pNum, siteLikes, *lnL, bad);
    return true;
}

// Q of Pf/util.c:44-115 (setBigQFromRMatrixDotCharFreq, normalizeBigQ), in
// the reference's operation order.
static void hostBigQ(int S, const double *pi, const double *R, double *Q)
{
    for (int c = 0; c < S; c++)
        for (int r = 0; r < S; r++) Q[r * S + c] = R[r * S + c] * pi[c];
    for (int r = 0; r < S; r++) {
        double sum = 0.0;
        for (int c = 0; c < S; c++)
            if (r != c) sum = sum + Q[r * S + c];
        Q[r * S + r] = -sum;
    }
    double sumODE = 0.0;
    for (int r = 0; r < S; r++)
        for (int c = 0; c < S; c++)
            if (r != c) sumODE = sumODE + pi[r] * Q[r * S + c];
    sumODE = 1.0 / sumODE;
    for (int i = 0; i < S * S; i++) Q[i] = Q[i] * sumODE;
}

// Per non-root node of part pNum: lnL across its branch and its first and
// second derivatives with respect to the branch length (out [nNodes][3]).
// The partials and P matrices must be current (a full evaluation just ran).
static bool derivPart(Tree *t, int pNum, std::vector<double> &out)
{
    TreePart &tp = t->tp[pNum];
    ModelPart *mp = t->model->parts[pNum];
    Part *dp = t->data->parts[pNum];
    const int S = tp.dim, K = tp.K, N = t->nNodes;
    hipStream_t s = stream();
    if (K > DERIV_MAX_K) {
        setError("branch derivatives: at most 16 gamma categories");
        return false;
    }
    const size_t SS = (size_t)S * S;
    if (!devAlloc(&tp.cl2, (size_t)N * K * tp.nPatPad * S, "hipMalloc cl2") ||
        !devAlloc(&tp.P1T, (size_t)N * K * SS, "hipMalloc P'") || !devAlloc(&tp.P2T, (size_t)N * K * SS, "hipMalloc P''") ||
        !devAlloc(&tp.dQ, (size_t)std::max(tp.nEig, 1) * SS, "hipMalloc Q") ||
        !devAlloc(&tp.dRate, (size_t)N * K, "hipMalloc rate factors") ||
        !devAlloc(&tp.dItems, (size_t)2 * N, "hipMalloc derivative items") ||
        !devAlloc(&tp.dOut, (size_t)3 * N, "hipMalloc derivative output"))
        return false;
    // Q of every (comp, rMatrix) slot
    std::vector<double> Q((size_t)std::max(tp.nEig, 1) * SS, 0.0);
    for (int c = 0; c < mp->nComps; c++)
        for (int r = 0; r < mp->nRMatrices; r++)
            hostBigQ(S, mp->comps[c].val, mp->rMatrices[r].bigR.data(), Q.data() + (size_t)(c * mp->nRMatrices + r) * SS);
    // every non-root node: its slot and d(effective length)/d(branch length)
    // (Pf/p4_node.c:442-540 without the branch length)
    std::vector<int> items(2 * N);
    std::vector<double> rate((size_t)N * K);
    int nI = 0;
    for (int n = 0; n < N; n++) {
        if (n == t->root) continue;
        const Node *nd = t->nodes[n];
        items[nI] = n;
        items[N + nI] = nd->compNums[pNum] * mp->nRMatrices + nd->rMatrixNums[pNum];
        for (int k = 0; k < K; k++) {
            const double g = mp->nGdasrvs ? mp->gdasrvs[nd->gdasrvNums[pNum]]->rates[k] : 1.0;
            const double r = g * mp->relRate;
            rate[(size_t)nI * K + k] = mp->pInvar == 0.0 ? r : r / (1.0 - mp->pInvar);
        }
        nI++;
    }
    // the up-pass program: non-root nodes in preorder with their siblings
    std::vector<DerivOp> ops;
    std::vector<DerivSib> sibs;
    for (int j = N - 1; j >= 0; j--) {
        const int n = t->postOrder[j];
        if (n == NO_ORDER || n == t->root) continue;
        if (n < 0 || n >= N) {
            setError("postOrder holds an out-of-range node number");
            return false;
        }
        const Node *nd = t->nodes[n];
        DerivOp op{};
        op.node = n;
        op.flags = (nd->isLeaf ? DOP_LEAF : 0) | (nd->parent == t->root ? DOP_PARENT_ROOT : 0);
        op.idx = nd->isLeaf ? nd->seqNum : t->slotOf[n];
        op.parent = nd->parent;
        op.sibStart = (int)sibs.size();
        for (int c = t->nodes[nd->parent]->leftChild; c >= 0; c = t->nodes[c]->sibling) {
            if (c == n) continue;
            const Node *cn = t->nodes[c];
            sibs.push_back(DerivSib{cn->isLeaf ? 0 : 1, cn->isLeaf ? cn->seqNum : t->slotOf[c], c, 0});
        }
        op.nSib = (int)sibs.size() - op.sibStart;
        ops.push_back(op);
    }
    const int nOps = (int)ops.size();
    const int nBlocks = tp.nPatPad / 64;
    const size_t progBytes = ops.size() * sizeof(DerivOp) + std::max<size_t>(sibs.size(), 1) * sizeof(DerivSib);
    if (progBytes > tp.dProgCap) {
        if (tp.dProg) (void)hipFree(tp.dProg);
        tp.dProg = nullptr;
        if (!hipOk(hipMalloc(&tp.dProg, progBytes), "hipMalloc derivative program")) return false;
        tp.dProgCap = progBytes;
    }
    const size_t partN = (size_t)std::max(nOps, 1) * nBlocks * 3;
    if (partN > tp.dPartialCap) {
        devFree(tp.dPartial);
        if (!devAlloc(&tp.dPartial, partN, "hipMalloc derivative sums")) return false;
        tp.dPartialCap = partN;
    }
    DerivOp *dOps = reinterpret_cast<DerivOp *>(tp.dProg);
    DerivSib *dSibs = reinterpret_cast<DerivSib *>(dOps + ops.size());
    if (!hipOk(hipMemcpyAsync(tp.dQ, Q.data(), Q.size() * 8, hipMemcpyHostToDevice, s), "stage Q") ||
        !hipOk(hipMemcpyAsync(tp.dRate, rate.data(), rate.size() * 8, hipMemcpyHostToDevice, s), "stage rates") ||
        !hipOk(hipMemcpyAsync(tp.dItems, items.data(), items.size() * 4, hipMemcpyHostToDevice, s), "stage items") ||
        !hipOk(hipMemcpyAsync(dOps, ops.data(), ops.size() * sizeof(DerivOp), hipMemcpyHostToDevice, s), "stage program") ||
        (!sibs.empty() &&
         !hipOk(hipMemcpyAsync(dSibs, sibs.data(), sibs.size() * sizeof(DerivSib), hipMemcpyHostToDevice, s),
                "stage siblings"))) {
        (void)hipStreamSynchronize(s);
        return false;
    }
    DerivPArgs pa{S, K, nI, tp.dItems, tp.dItems + N, tp.dRate, tp.dQ, tp.PT64, tp.P1T, tp.P2T};
    if (!hipOk(launchPderiv(pa, s), "launch P derivative kernel")) {
        (void)hipStreamSynchronize(s);
        return false;
    }
    DerivArgs a{};
    a.dim = S;
    a.K = K;
    a.nPat = tp.nPat;
    a.nPatPad = tp.nPatPad;
    a.nOps = nOps;
    a.waveLayout = tp.wave ? 1 : 0;
    a.ops = dOps;
    a.sibs = dSibs;
    a.tips = dp->dev.tips;
    a.ambMask = dp->dev.ambMask;
    a.cl = reinterpret_cast<const double *>(tp.cl);
    a.cl2 = tp.cl2;
    a.PT64 = tp.PT64;
    a.P1T = tp.P1T;
    a.P2T = tp.P2T;
    a.rootPi = reinterpret_cast<const double *>(tp.dArena + tp.offPi);   // staged by the evaluation just run
    a.rootLeafRow = t->nodes[t->root]->isLeaf ? t->nodes[t->root]->seqNum : -1;
    a.pInvar = mp->pInvar;
    a.counts = dp->dev.counts;
    a.invMask = dp->dev.invMask;
    a.nWords = dp->dev.nWords;
    a.partial = tp.dPartial;
    a.out = tp.dOut;
    if (!hipOk(launchEdgeDerivs(a, nBlocks, s), "launch derivative kernels")) {
        (void)hipStreamSynchronize(s);
        return false;
    }
    out.assign((size_t)3 * N, 0.0);
    if (!hipOk(hipMemcpyAsync(out.data(), tp.dOut, (size_t)3 * N * 8, hipMemcpyDeviceToHost, s), "copy derivatives") ||
        !hipOk(hipStreamSynchronize(s), "derivative kernels"))
        return false;
    for (int k = 0; k < 3; k++) out[(size_t)t->root * 3 + k] = 0.0;
    tp.derivValid = true;
    return true;
}

}  // namespace p4g

using namespace p4g;

// ===================================================================== C-ABI
extern "C" {

const char *p4g_lastError(void) { return g_err.c_str(); }

int p4g_deviceAvailable(void) { return deviceReady() ? 1 : 0; }

void p4g_setKernelTiming(int on) { g_timing = on != 0; }

void p4g_setPruneTuning(int fp32, int opsPerGroup, int stagePInLds)
{
    setPruneTuning(fp32 ? FP32 : FP64, opsPerGroup, stagePInLds);
}

void p4g_setPmatRoute(int direct)
{
    setPmatRoute(direct);
}

void p4g_setAaTuning(int categoryWaves, int minOps)
{
    setAaTuning(categoryWaves, minOps);
}

void p4g_setWavePreloadLds(int bytes) { setWavePreloadLds(bytes); }

void p4g_setAaLdsSlots(int maxSlots) { setAaLdsSlots(maxSlots); }

double p4g_lastKernelMicroseconds(const p4g_tree *t)
{
    if (!t || t->tp.empty()) return -1.0;
    double s = 0.0;
    int n = 0;
    for (auto &tp : t->tp)
        if (tp.lastKernelUs >= 0) {
            s += tp.lastKernelUs;
            n++;
        }
    return n ? s : -1.0;
}

// ---------------------------------------------------------------- data
p4g_data *p4g_newData(int nTax, int nParts)
{
    clearError();
    if (nTax < 1 || nParts < 1) {
        setError("newData: nTax and nParts must be positive");
        return nullptr;
    }
    auto *d = new p4g_data();
    d->nTax = nTax;
    d->nParts = nParts;
    d->parts.assign(nParts, nullptr);
    return d;
}

void p4g_freeData(p4g_data *d) { delete d; }

int p4g_pokePartInData(p4g_part *p, p4g_data *d, int partNum)
{
    clearError();
    if (!p || !d || partNum < 0 || partNum >= d->nParts) {
        setError("pokePartInData: bad arguments");
        return 1;
    }
    d->parts[partNum] = p;
    p->data = d;
    return 0;
}

p4g_part *p4g_newPart(int nTax, int nChar, const char *equateSymbols, int nEquates,
                      const char *symbols, int dim)
{
    clearError();
    if (nTax < 1 || nChar < 1 || dim < 2 || nEquates < 0) {
        setError("newPart: bad sizes");
        return nullptr;
    }
    auto *p = new p4g_part();
    p->nTax = nTax;
    p->nChar = nChar;
    p->dim = dim;
    p->nEquates = nEquates;
    p->symbols = symbols ? std::string(symbols, strnlen(symbols, dim)) : std::string();
    p->equateSymbols = equateSymbols ? std::string(equateSymbols, strnlen(equateSymbols, nEquates)) : std::string();
    p->equates.assign((size_t)nEquates * dim, 0);
    p->sequences.assign((size_t)nTax * nChar, 0);
    return p;
}

void p4g_freePart(p4g_part *p) { delete p; }

int p4g_pokeEquatesTable(p4g_part *p, const char *table)
{
    clearError();
    if (!p || !table) return 1;
    for (int i = 0; i < p->nEquates * p->dim; i++) p->equates[i] = (table[i] == '1') ? 1 : 0;
    p->dev.valid = false;
    return 0;
}

// Pf/part.c:127-275: symbols -> 0..dim-1, '-' gap, '?' qmark, equates.
int p4g_pokeSequences(p4g_part *p, const char *seqs)
{
    clearError();
    if (!p || !seqs) return 1;
    int lut[256];
    for (int i = 0; i < 256; i++) lut[i] = INT32_MIN;
    for (int e = p->nEquates - 1; e >= 0; e--) lut[(unsigned char)p->equateSymbols[e]] = EQUATES_BASE + e;
    lut[(unsigned char)'-'] = GAP_CODE;
    lut[(unsigned char)'?'] = QMARK_CODE;
    for (int m = p->dim - 1; m >= 0; m--)
        if (m < (int)p->symbols.size()) lut[(unsigned char)p->symbols[m]] = m;
    const size_t n = (size_t)p->nTax * p->nChar;
    for (size_t i = 0; i < n; i++) {
        const int c = lut[(unsigned char)seqs[i]];
        if (c == INT32_MIN) {
            char b[128];
            snprintf(b, sizeof b, "pokeSequences: got character '%c', neither a symbol nor an equate", seqs[i]);
            setError(b);
            return 1;
        }
        p->sequences[i] = c;
    }
    p->dev.valid = false;
    p->dSeqValid = false;
    p->simCats.clear();
    return 0;
}

// Pf/part.c:317-448: patterns in first-occurrence order with counts.  A hash
// of each column replaces the reference's quadratic search.
int p4g_makePatterns(p4g_part *p)
{
    clearError();
    if (!p) return 1;
    const int nTax = p->nTax, nChar = p->nChar;
    std::unordered_map<std::string, int> idx;
    idx.reserve((size_t)nChar * 2);
    std::string key((size_t)nTax * 2, '\0');
    std::vector<int> firstSite;
    p->patternCounts.clear();
    p->seqPosPatternIndex.assign(nChar, 0);
    for (int i = 0; i < nChar; i++) {
        for (int t = 0; t < nTax; t++) {
            const int c = p->sequences[(size_t)t * nChar + i] + 128;
            key[2 * t] = (char)(c & 0xff);
            key[2 * t + 1] = (char)((c >> 8) & 0xff);
        }
        auto it = idx.find(key);
        if (it != idx.end()) {
            p->patternCounts[it->second]++;
            p->seqPosPatternIndex[i] = it->second;
        } else {
            const int k = (int)firstSite.size();
            idx.emplace(key, k);
            firstSite.push_back(i);
            p->patternCounts.push_back(1);
            p->seqPosPatternIndex[i] = k;
        }
    }
    const int nP = (int)firstSite.size();
    p->nPatterns = nP;
    p->patterns.assign((size_t)nTax * nP, 0);
    for (int t = 0; t < nTax; t++)
        for (int k = 0; k < nP; k++) p->patterns[(size_t)t * nP + k] = p->sequences[(size_t)t * nChar + firstSite[k]];
    p->invarSet = false;
    p->dev.valid = false;
    if (!checkCodes(p)) return 1;
    return 0;
}

// makePatterns on the GPU: the same patterns, counts and site map as
// p4g_makePatterns (Pf/part.c:317-448), from the part's device sequences
// (a simulation leaves them there; otherwise the host sequences go up).
int p4g_makePatternsDevice(p4g_part *p)
{
    clearError();
    if (!p) return 1;
    if (!deviceReady()) return 1;
    const size_t n = (size_t)p->nTax * p->nChar;
    if (p->nChar < 1 || p->sequences.size() != n) {
        setError("makePatternsDevice: the part has no sequences");
        return 1;
    }
    hipStream_t s = stream();
    if (!p->dSeqValid) {
        std::vector<int8_t> h(n);
        for (size_t i = 0; i < n; i++) h[i] = (int8_t)p->sequences[i];   // codes fit: -64..63
        if (!partSeqBuffer(p) ||
            !hipOk(hipMemcpyAsync(p->dSeq, h.data(), n, hipMemcpyHostToDevice, s), "copy sequences") ||
            !hipOk(hipStreamSynchronize(s), "copy sequences"))
            return 1;
        p->dSeqValid = true;
    }
    std::vector<int8_t> pats;
    std::vector<int> counts, sitePat;
    int nP = 0, coll = 0;
    const hipError_t e = compressPatterns(p->dSeq, p->nTax, p->nChar, pats, counts, sitePat, &nP, &coll);
    if (e != hipSuccess) {
        setError(coll >= 4 ? "makePatternsDevice: the column hash collided under four seeds"
                           : std::string("makePatternsDevice: ") + hipGetErrorString(e));
        return 1;
    }
    p->nPatterns = nP;
    p->patterns.assign(pats.begin(), pats.end());
    p->patternCounts = counts;
    p->seqPosPatternIndex = sitePat;
    p->invarSet = false;
    p->dev.valid = false;
    return checkCodes(p) ? 0 : 1;
}

int p4g_setGlobalInvarSitesVec(p4g_part *p)
{
    clearError();
    if (!p || p->nPatterns < 1) {
        setError("setGlobalInvarSitesVec: no patterns");
        return 1;
    }
    computeInvar(p);
    p->dev.valid = false;
    return 0;
}

int p4g_partNPatterns(const p4g_part *p) { return p ? p->nPatterns : -1; }
int p4g_partNChar(const p4g_part *p) { return p ? p->nChar : -1; }
int p4g_partNTax(const p4g_part *p) { return p ? p->nTax : -1; }
int p4g_partDim(const p4g_part *p) { return p ? p->dim : -1; }

int p4g_partGetPatterns(const p4g_part *p, int *patterns, int *counts, int *seqPosPatternIndex,
                        int *invarVec, int *invarArray)
{
    if (!p) return 1;
    if (patterns) std::copy(p->patterns.begin(), p->patterns.end(), patterns);
    if (counts) std::copy(p->patternCounts.begin(), p->patternCounts.end(), counts);
    if (seqPosPatternIndex) std::copy(p->seqPosPatternIndex.begin(), p->seqPosPatternIndex.end(), seqPosPatternIndex);
    if (invarVec || invarArray) {
        if (!p->invarSet) computeInvar(const_cast<p4g_part *>(p));
        if (invarVec) std::copy(p->invarVec.begin(), p->invarVec.end(), invarVec);
        if (invarArray) std::copy(p->invarArray.begin(), p->invarArray.end(), invarArray);
    }
    return 0;
}

int p4g_partSetPatterns(p4g_part *p, int nPatterns, const int *patterns, const int *counts)
{
    clearError();
    if (!p || nPatterns < 1 || !patterns || !counts) {
        setError("partSetPatterns: bad arguments");
        return 1;
    }
    p->nPatterns = nPatterns;
    p->patterns.assign(patterns, patterns + (size_t)p->nTax * nPatterns);
    p->patternCounts.assign(counts, counts + nPatterns);
    // sites are taken in pattern order, each pattern repeated count times
    p->seqPosPatternIndex.clear();
    for (int k = 0; k < nPatterns; k++)
        for (int c = 0; c < counts[k]; c++) p->seqPosPatternIndex.push_back(k);
    p->nChar = (int)p->seqPosPatternIndex.size();
    p->invarSet = false;
    p->dev.valid = false;
    return checkCodes(p) ? 0 : 1;
}

int p4g_getSiteLikes(const p4g_part *p, double *out)
{
    if (!p || p->siteLikes.empty()) {
        setError("getSiteLikes: no site likelihoods (evaluate with getSiteLikes=1 first)");
        return 1;
    }
    std::copy(p->siteLikes.begin(), p->siteLikes.end(), out);
    return 0;
}

// --------------------------------------------------------------- model
p4g_model *p4g_newModel(int nParts, int doRelRates, int relRatesAreFree, int nFreePrams, int isHet,
                        int *rMatrixNormalizeTo1, double *PINVAR_MIN, double *PINVAR_MAX,
                        double *KAPPA_MIN, double *KAPPA_MAX, double *GAMMA_SHAPE_MIN,
                        double *GAMMA_SHAPE_MAX, double *PIVEC_MIN, double *PIVEC_MAX,
                        double *RATE_MIN, double *RATE_MAX, double *RELRATE_MIN,
                        double *RELRATE_MAX, double *BRLEN_MIN, double *BRLEN_MAX)
{
    clearError();
    if (nParts < 1) {
        setError("newModel: nParts must be positive");
        return nullptr;
    }
    auto *m = new p4g_model();
    m->nParts = nParts;
    m->doRelRates = doRelRates;
    m->relRatesAreFree = relRatesAreFree;
    m->nFreePrams = nFreePrams;
    m->isHet = isHet;
    m->rMatrixNormalizeTo1 = rMatrixNormalizeTo1;
    m->PINVAR_MIN = PINVAR_MIN;
    m->PINVAR_MAX = PINVAR_MAX;
    m->KAPPA_MIN = KAPPA_MIN;
    m->KAPPA_MAX = KAPPA_MAX;
    m->GAMMA_SHAPE_MIN = GAMMA_SHAPE_MIN;
    m->GAMMA_SHAPE_MAX = GAMMA_SHAPE_MAX;
    m->PIVEC_MIN = PIVEC_MIN;
    m->PIVEC_MAX = PIVEC_MAX;
    m->RATE_MIN = RATE_MIN;
    m->RATE_MAX = RATE_MAX;
    m->RELRATE_MIN = RELRATE_MIN;
    m->RELRATE_MAX = RELRATE_MAX;
    m->BRLEN_MIN = BRLEN_MIN;
    m->BRLEN_MAX = BRLEN_MAX;
    m->parts.assign(nParts, nullptr);
    return m;
}

void p4g_freeModel(p4g_model *m) { delete m; }

static ModelPart *modelPart(p4g_model *m, int pNum)
{
    if (!m || pNum < 0 || pNum >= m->nParts || !m->parts[pNum]) {
        setError("model part " + std::to_string(pNum) + " does not exist");
        return nullptr;
    }
    return m->parts[pNum];
}

int p4g_newModelPart(p4g_model *m, int pNum, int dim, int nComps, int nRMatrices, int nGdasrvs,
                     int nCat, int pInvarFree, int *bQETneedsReset)
{
    clearError();
    if (!m || pNum < 0 || pNum >= m->nParts || dim < 2 || dim > MAX_DIM || nComps < 1 ||
        nRMatrices < 1 || nGdasrvs < 0 || nCat < 1 || (nCat > 1 && nGdasrvs < 1)) {
        setError("newModelPart: bad arguments");
        return 1;
    }
    auto *mp = new ModelPart();
    mp->dim = dim;
    mp->nComps = nComps;
    mp->nRMatrices = nRMatrices;
    mp->nGdasrvs = nGdasrvs;
    mp->nCat = nCat;
    mp->pInvarFree = pInvarFree;
    mp->comps.resize(nComps);
    mp->rMatrices.resize(nRMatrices);
    mp->gdasrvs.assign(nGdasrvs, nullptr);
    mp->pInvar = -1.0;
    mp->relRate = -1.0;
    mp->bQETneedsReset = bQETneedsReset;
    delete m->parts[pNum];
    m->parts[pNum] = mp;
    return 0;
}

int p4g_newComp(p4g_model *m, int pNum, int mNum, int free, double *val)
{
    clearError();
    ModelPart *mp = modelPart(m, pNum);
    if (!mp || mNum < 0 || mNum >= mp->nComps || !val) {
        if (mp) setError("newComp: bad arguments");
        return 1;
    }
    mp->comps[mNum].free = free;
    mp->comps[mNum].val = val;
    return 0;
}

int p4g_newRMatrix(p4g_model *m, int pNum, int mNum, int free, int spec)
{
    clearError();
    ModelPart *mp = modelPart(m, pNum);
    if (!mp || mNum < 0 || mNum >= mp->nRMatrices) {
        if (mp) setError("newRMatrix: bad arguments");
        return 1;
    }
    if (spec > RMATRIX_POISSON) {
        setError("newRMatrix: empirical protein rate matrices (spec " + std::to_string(spec) +
                 ") are not built in; give the 190 exchangeabilities as a 'specified' rMatrix");
        return 1;
    }
    if (spec == RMATRIX_2P && mp->dim != 4) {
        setError("newRMatrix: the 2p matrix is for 4-state data");
        return 1;
    }
    RMatrix &r = mp->rMatrices[mNum];
    r.free = free;
    r.spec = spec;
    r.kappa = 2.0;
    r.bigR.assign((size_t)mp->dim * mp->dim, 1.0);
    if (spec == RMATRIX_2P) {
        const double al = 1.0 / 3.0, be = al * r.kappa;
        const double R[16] = {0, al, be, al, al, 0, al, be, be, al, 0, al, al, be, al, 0};
        r.bigR.assign(R, R + 16);
    }
    return 0;
}

p4g_gdasrv *p4g_newGdasrv(p4g_model *m, int pNum, int mNum, int nCat, int free, double *val,
                          double *freqs, double *rates)
{
    clearError();
    ModelPart *mp = modelPart(m, pNum);
    if (!mp || mNum < 0 || mNum >= mp->nGdasrvs || !val || !freqs || !rates || nCat != mp->nCat) {
        if (mp) setError("newGdasrv: bad arguments");
        return nullptr;
    }
    auto *g = new p4g_gdasrv();
    g->free = free;
    g->nCat = nCat;
    g->val = val;
    g->freqs = freqs;
    g->rates = rates;
    delete static_cast<p4g_gdasrv *>(mp->gdasrvs[mNum]);
    mp->gdasrvs[mNum] = g;
    return g;
}

int p4g_setRMatrixBigR(p4g_model *m, int pNum, int mNum, int i, int j, double val)
{
    clearError();
    ModelPart *mp = modelPart(m, pNum);
    if (!mp || mNum < 0 || mNum >= mp->nRMatrices || i < 0 || j < 0 || i >= mp->dim || j >= mp->dim) {
        if (mp) setError("setRMatrixBigR: bad arguments");
        return 1;
    }
    mp->rMatrices[mNum].bigR[(size_t)i * mp->dim + j] = val;
    mp->rMatrices[mNum].bigR[(size_t)j * mp->dim + i] = val;
    return 0;
}

int p4g_setKappa(p4g_model *m, int pNum, int mNum, double val)
{
    clearError();
    ModelPart *mp = modelPart(m, pNum);
    if (!mp || mNum < 0 || mNum >= mp->nRMatrices) return 1;
    mp->rMatrices[mNum].kappa = val;
    return 0;
}

int p4g_gdasrvCalcRates(p4g_gdasrv *g)
{
    clearError();
    if (!g) return 1;
    return discreteGamma(g->freqs, g->rates, g->val[0], g->val[0], g->nCat, 0);
}

int p4g_gdasrvCalcRates_np(int nCat, double val, double *freqs, double *rates)
{
    clearError();
    if (nCat < 1 || !freqs || !rates) return 1;
    return discreteGamma(freqs, rates, val, val, nCat, 0);
}

int p4g_setPInvarVal(p4g_model *m, int pNum, double val)
{
    clearError();
    ModelPart *mp = modelPart(m, pNum);
    if (!mp) return 1;
    if (!(val >= 0.0 && val < 1.0)) {
        setError("setPInvarVal: pInvar must be in [0, 1)");
        return 1;
    }
    mp->pInvar = val;
    return 0;
}

int p4g_setRelRateVal(p4g_model *m, int pNum, double val)
{
    clearError();
    ModelPart *mp = modelPart(m, pNum);
    if (!mp) return 1;
    mp->relRate = val;
    return 0;
}

int p4g_resetBQET(p4g_model *m, int pNum, int compNum, int rMatrixNum)
{
    clearError();
    ModelPart *mp = modelPart(m, pNum);
    if (!mp || compNum < 0 || compNum >= mp->nComps || rMatrixNum < 0 || rMatrixNum >= mp->nRMatrices)
        return 1;
    if (mp->bQETneedsReset) mp->bQETneedsReset[compNum * mp->nRMatrices + rMatrixNum] = 1;
    return 0;
}

// ---------------------------------------------------------------- tree
p4g_tree *p4g_newTree(int nNodes, int nLeaves, int *preOrder, int *postOrder, double *partLikes,
                      p4g_data *data, p4g_model *model)
{
    clearError();
    if (nNodes < 1 || !preOrder || !postOrder || !data || !model) {
        setError("newTree: bad arguments");
        return nullptr;
    }
    if (data->nParts != model->nParts) {
        setError("newTree: data and model part counts differ");
        return nullptr;
    }
    auto *t = new p4g_tree();
    t->nNodes = nNodes;
    t->nLeaves = nLeaves;
    t->nodes.assign(nNodes, nullptr);
    t->preOrder = preOrder;
    t->postOrder = postOrder;
    t->partLikes = partLikes;
    t->data = data;
    t->model = model;
    t->nParts = model->nParts;
    t->tp.resize(t->nParts);
    return t;
}

void p4g_freeTree(p4g_tree *t)
{
    if (!t) return;
    for (auto *n : t->nodes) delete static_cast<p4g_node *>(n);
    delete t;
}

p4g_node *p4g_newNode(int nodeNum, p4g_tree *t, int seqNum, int isLeaf, int inTree)
{
    clearError();
    if (!t || nodeNum < 0 || nodeNum >= t->nNodes) {
        setError("newNode: bad node number");
        return nullptr;
    }
    auto *n = new p4g_node();
    n->nodeNum = nodeNum;
    n->tree = t;
    n->seqNum = seqNum;
    n->isLeaf = isLeaf;
    n->inTree = inTree;
    n->compNums.assign(t->nParts, 0);
    n->rMatrixNums.assign(t->nParts, 0);
    n->gdasrvNums.assign(t->nParts, 0);
    delete static_cast<p4g_node *>(t->nodes[nodeNum]);
    t->nodes[nodeNum] = n;
    for (auto &tp : t->tp) {
        if (tp.alloc) freeTreePart(tp);
    }
    return n;
}

void p4g_freeNode(p4g_node *n)
{
    if (!n) return;
    if (n->tree && n->nodeNum >= 0 && n->nodeNum < n->tree->nNodes && n->tree->nodes[n->nodeNum] == n)
        n->tree->nodes[n->nodeNum] = nullptr;
    delete n;
}

int p4g_setNodeRelation(p4g_node *n, int relation, int relNum)
{
    clearError();
    if (!n || relNum >= n->tree->nNodes || relNum < -1) {
        setError("setNodeRelation: bad relative");
        return 1;
    }
    if (relation == 0)
        n->parent = relNum;
    else if (relation == 1)
        n->leftChild = relNum;
    else if (relation == 2)
        n->sibling = relNum;
    else {
        setError("Error in p4_setNodeRelation: \"relation\" is out of range");
        return 1;
    }
    return 0;
}

int p4g_setTreeRoot(p4g_tree *t, p4g_node *n)
{
    clearError();
    if (!t || !n) return 1;
    t->root = n->nodeNum;
    return 0;
}

int p4g_setBrLen(p4g_node *n, double len)
{
    if (!n) return 1;
    n->brLen = len;
    return 0;
}

double p4g_getBrLen(const p4g_node *n) { return n ? n->brLen : NAN; }

double p4g_getTreeLen(const p4g_tree *t)
{
    if (!t) return NAN;
    double s = 0.0;
    for (int j = 0; j < t->nNodes; j++) {
        const int i = t->preOrder[j];
        if (i == NO_ORDER || i == t->root) continue;
        s += t->nodes[i]->brLen;
    }
    return s;
}

static int setNodeNum(p4g_node *n, int pNum, int val, int which)
{
    clearError();
    if (!n || pNum < 0 || pNum >= n->tree->nParts) {
        setError("set model number: bad part");
        return 1;
    }
    (which == 0 ? n->compNums : which == 1 ? n->rMatrixNums : n->gdasrvNums)[pNum] = val;
    return 0;
}

int p4g_setCompNum(p4g_node *n, int pNum, int val) { return setNodeNum(n, pNum, val, 0); }
int p4g_setRMatrixNum(p4g_node *n, int pNum, int val) { return setNodeNum(n, pNum, val, 1); }
int p4g_setGdasrvNum(p4g_node *n, int pNum, int val) { return setNodeNum(n, pNum, val, 2); }

int p4g_setPrecision(p4g_tree *t, int fp32)
{
    clearError();
    if (!t) return 1;
    t->precision = fp32 ? FP32 : FP64;
    return 0;
}

int p4g_getPrecision(const p4g_tree *t) { return t ? t->precision : -1; }

// ----------------------------------------------------------- evaluation
int p4g_setPrams(p4g_tree *t, int pNum)
{
    clearError();
    if (!t) return 1;
    if (!treeReady(t, false)) return 1;
    if (pNum >= t->nParts) {
        setError("setPrams: part out of range");
        return 1;
    }
    for (int p = 0; p < t->nParts; p++) {
        if (pNum >= 0 && p != pNum) continue;
        if (!setPramsPart(t, p)) return 1;
    }
    return 0;
}

int p4g_treeLogLike(p4g_tree *t, int getSiteLikes, double *lnL)
{
    clearError();
    if (!t || !lnL) return 1;
    if (!treeReady(t)) return 1;
    double sum = 0.0;
    for (int p = 0; p < t->nParts; p++) {
        double v = 0.0;
        if (!evalPart(t, p, true, getSiteLikes != 0, &v)) return 1;
        sum += v;
    }
    t->logLike = sum;
    *lnL = sum;
    return 0;
}

int p4g_treeLogLikeNoAlloc(p4g_tree *t, double *lnL)
{
    clearError();
    if (!t || !lnL) {
        setError("treeLogLikeNoAlloc: bad arguments");
        return 1;
    }
    if (!deviceReady()) return 1;
    int dev = 0;
    if (!hipOk(hipGetDevice(&dev), "hipGetDevice")) return 1;
    unsigned char *scr = (dev >= 0 && dev < MAX_DEVICES) ? g_scratch[dev] : nullptr;
    if (!scr) {
        setError("treeLogLikeNoAlloc: no scratch address for the current device");
        return 1;
    }
    // the tree's own device state must exist already: this call sets up
    // nothing (treeReady would allocate)
    if (!t->data || !t->model || t->root < 0 || t->data->nParts != t->model->nParts ||
        (int)t->tp.size() != t->model->nParts) {
        setError("treeLogLikeNoAlloc: the tree is not set up");
        return 1;
    }
    for (int p = 0; p < t->nParts; p++) {
        const TreePart &tp = t->tp[p];
        const ModelPart *mp = t->model->parts[p];
        const Part *dp = t->data->parts[p];
        if (!mp || !dp || !dp->dev.valid || !tp.alloc || tp.precision != t->precision ||
            tp.wave != pruneUsesWave(mp->dim, mp->nCat, t->precision)) {
            setError("treeLogLikeNoAlloc: part " + std::to_string(p) +
                     " has no device state for the tree's current precision and kernel; evaluate it once with "
                     "p4g_treeLogLike first (that call allocates)");
            return 1;
        }
    }
    double sum = 0.0;
    for (int p = 0; p < t->nParts; p++) {
        double v = 0.0;
        if (!evalPart(t, p, true, false, &v, scr)) return 1;
        sum += v;
    }
    t->logLike = sum;
    *lnL = sum;
    return 0;
}

int p4g_partLogLike(p4g_tree *t, p4g_part *p, int pNum, int getSiteLikes, double *lnL)
{
    clearError();
    if (!t || !lnL || pNum < 0 || pNum >= t->nParts) {
        setError("partLogLike: bad arguments");
        return 1;
    }
    if (p && t->data->parts[pNum] != p) {
        setError("partLogLike: part is not part pNum of the tree's data");
        return 1;
    }
    if (!treeReady(t)) return 1;
    return evalPart(t, pNum, false, getSiteLikes != 0, lnL) ? 0 : 1;
}

int p4g_calculateBigPDecksPart(p4g_node *n, int pNum)
{
    clearError();
    if (!n) return 1;
    Tree *t = n->tree;
    if (!treeReady(t)) return 1;
    return bigPDecksPart(t, n, pNum) ? 0 : 1;
}

int p4g_calculateBigPDecks(p4g_node *n)
{
    clearError();
    if (!n) return 1;
    for (int p = 0; p < n->tree->nParts; p++)
        if (p4g_calculateBigPDecksPart(n, p)) return 1;
    return 0;
}

int p4g_calculateAllBigPDecksAllParts(p4g_tree *t)
{
    clearError();
    if (!t || !treeReady(t)) return 1;
    for (int p = 0; p < t->nParts; p++) {
        std::vector<int> nodes;
        ModelPart *mp = t->model->parts[p];
        for (int n = 0; n < t->nNodes; n++) {
            if (n == t->root) continue;
            Node *nd = t->nodes[n];
            if (!checkNodeModel(t, mp, nd, p)) return 1;
            const int slot = nd->compNums[p] * mp->nRMatrices + nd->rMatrixNums[p];
            if (mp->bQETneedsReset && mp->bQETneedsReset[slot]) {
                markEig(t, p, slot);
                mp->bQETneedsReset[slot] = 0;
            }
            markP(t, p, nd);
        }
    }
    return 0;
}

int p4g_setConditionalLikelihoodsOfInternalNodePart(p4g_node *n, int pNum)
{
    clearError();
    if (!n || pNum < 0 || pNum >= n->tree->nParts) return 1;
    Tree *t = n->tree;
    TreePart &tp = t->tp[pNum];
    if (tp.dirty.size() != (size_t)t->nNodes) tp.dirty.assign(t->nNodes, 1);
    tp.dirty[n->nodeNum] = 1;
    return 0;
}

int p4g_setConditionalLikelihoodsOfInternalNode(p4g_node *n)
{
    if (!n) return 1;
    for (int p = 0; p < n->tree->nParts; p++)
        if (p4g_setConditionalLikelihoodsOfInternalNodePart(n, p)) return 1;
    return 0;
}

// Node n and its ancestors (internal nodes, and a leaf root) marked for
// recomputation in part pNum or every part (-1): the walk Chain's proposals
// make after a change at n, as one call.
int p4g_markPath(p4g_tree *t, int nodeNum, int pNum)
{
    clearError();
    if (!t || nodeNum < 0 || nodeNum >= t->nNodes || pNum < -1 || pNum >= t->nParts) {
        setError("markPath: bad arguments");
        return 1;
    }
    int guard = 0;
    for (int n = nodeNum; n >= 0; n = t->nodes[n]->parent) {
        if (!t->nodes[n] || ++guard > t->nNodes) {
            setError("markPath: the parent chain does not end at the root");
            return 1;
        }
        if (t->nodes[n]->isLeaf && n != t->root) continue;
        for (int p = 0; p < t->nParts; p++) {
            if (pNum >= 0 && p != pNum) continue;
            TreePart &tp = t->tp[p];
            if (tp.dirty.size() != (size_t)t->nNodes) tp.dirty.assign(t->nNodes, 1);
            tp.dirty[n] = 1;
        }
    }
    return 0;
}

// Every internal node (and a leaf root) of part pNum, or every part (-1).
int p4g_markAll(p4g_tree *t, int pNum)
{
    clearError();
    if (!t || pNum < -1 || pNum >= t->nParts) {
        setError("markAll: bad arguments");
        return 1;
    }
    for (int p = 0; p < t->nParts; p++) {
        if (pNum >= 0 && p != pNum) continue;
        TreePart &tp = t->tp[p];
        if (tp.dirty.size() != (size_t)t->nNodes) tp.dirty.assign(t->nNodes, 1);
        for (int n = 0; n < t->nNodes; n++)
            if (t->nodes[n] && (!t->nodes[n]->isLeaf || n == t->root)) tp.dirty[n] = 1;
    }
    return 0;
}

// ---------------------------------------------------------- state sync
static bool sameShape(Tree *a, Tree *b)
{
    if (a->nNodes != b->nNodes || a->nParts != b->nParts) {
        setError("trees differ in shape");
        return false;
    }
    return true;
}

int p4g_copyCondLikes(p4g_tree *a, p4g_tree *b, int doAll)
{
    clearError();
    if (!a || !b || !sameShape(a, b) || !treeReady(a) || !treeReady(b)) return 1;
    if (a->precision != b->precision) {
        setError("copyCondLikes: trees have different precision");
        return 1;
    }
    hipStream_t s = stream();
    for (int p = 0; p < a->nParts; p++) {
        TreePart &A = a->tp[p], &B = b->tp[p];
        if (A.nSlots != B.nSlots || A.clBytes != B.clBytes || A.wave != B.wave || A.sclPerSlot != B.sclPerSlot) {
            setError("copyCondLikes: partial buffers differ");
            return 1;
        }
        if (doAll) {
            if (!hipOk(hipMemcpyAsync(B.cl, A.cl, A.clBytes * A.nSlots, hipMemcpyDeviceToDevice, s), "copy partials"))
                return 1;
            if (A.scl && !hipOk(hipMemcpyAsync(B.scl, A.scl, sizeof(int) * (size_t)A.nSlots * A.sclPerSlot,
                                               hipMemcpyDeviceToDevice, s),
                                "copy scale exponents"))
                return 1;
            B.dirty = A.dirty;
        } else {
            for (int j = 0; j < a->nNodes; j++) {
                const int i = a->preOrder[j];
                if (i == NO_ORDER || a->nodes[i]->isLeaf) continue;
                if (!(A.dirty[i] || B.dirty[i])) continue;
                const int sl = a->slotOf[i];
                if (!hipOk(hipMemcpyAsync((char *)B.cl + A.clBytes * sl, (char *)A.cl + A.clBytes * sl, A.clBytes,
                                          hipMemcpyDeviceToDevice, s),
                           "copy partials"))
                    return 1;
                if (A.scl && !hipOk(hipMemcpyAsync(B.scl + (size_t)sl * A.sclPerSlot, A.scl + (size_t)sl * A.sclPerSlot,
                                                   sizeof(int) * A.sclPerSlot, hipMemcpyDeviceToDevice, s),
                                    "copy scale exponents"))
                    return 1;
                B.dirty[i] = A.dirty[i];
            }
        }
        B.everComputed = A.everComputed;
    }
    return 0;
}

int p4g_copyBigPDecks(p4g_tree *a, p4g_tree *b, int doAll)
{
    clearError();
    if (!doAll) {
        setError("p4_copyBigPDecks() doAll is not set. Programming error?");
        return 1;
    }
    if (!a || !b || !sameShape(a, b) || !treeReady(a) || !treeReady(b)) return 1;
    if (!flushAll(a) || !flushAll(b)) return 1;
    hipStream_t s = stream();
    for (int p = 0; p < a->nParts; p++) {
        TreePart &A = a->tp[p], &B = b->tp[p];
        const size_t n = (size_t)a->nNodes * A.K * A.dim * A.dim;
        if (!hipOk(hipMemcpyAsync(B.PT64, A.PT64, n * 8, hipMemcpyDeviceToDevice, s), "copy P") ||
            !hipOk(hipMemcpyAsync(B.PT32, A.PT32, n * 4, hipMemcpyDeviceToDevice, s), "copy P"))
            return 1;
        B.pValid = A.pValid;
    }
    return 0;
}

// The MCMC state sync without whole-tree copies: dst takes src's partial
// slots, scale exponents and P blocks wherever either tree changed them
// since their last sync (the recomputed path of a proposal; everything after
// a re-layout), all parts in one copy launch.  Replaces p4_copyCondLikes and
// p4_copyBigPDecks with doAll (Pf/p4_treeCopyVerify.c:7-61) in Chain's
// accept/reject step (p4/chain.py:1351-1531).
int p4g_syncChanged(p4g_tree *a, p4g_tree *b)
{
    clearError();
    if (!a || !b || a == b || !sameShape(a, b) || !treeReady(a, false) || !treeReady(b, false)) return 1;
    if (a->precision != b->precision) {
        setError("syncChanged: trees have different precision");
        return 1;
    }
    hipStream_t s = stream();
    // P matrices still pending in src are computed first
    for (int p = 0; p < a->nParts; p++)
        if (!flushParams(a, p)) {
            (void)hipStreamSynchronize(s);
            return 1;
        }
    std::vector<CopyRange> r;
    for (int p = 0; p < a->nParts; p++) {
        TreePart &A = a->tp[p], &B = b->tp[p];
        if (A.nSlots != B.nSlots || A.clBytes != B.clBytes || A.wave != B.wave || A.sclPerSlot != B.sclPerSlot ||
            A.K != B.K || A.dim != B.dim) {
            setError("syncChanged: partial buffers differ");
            return 1;
        }
        const bool all = A.chAll || B.chAll;
        for (int sl = 0; sl < A.nSlots; sl++) {
            if (!all && !A.chSlot[sl] && !B.chSlot[sl]) continue;
            r.push_back({(const char *)A.cl + A.clBytes * sl, (char *)B.cl + A.clBytes * sl, A.clBytes});
            if (A.scl)
                r.push_back({A.scl + (size_t)sl * A.sclPerSlot, B.scl + (size_t)sl * A.sclPerSlot,
                             sizeof(int) * A.sclPerSlot});
        }
        const size_t pe = (size_t)A.K * A.dim * A.dim;
        for (int n = 0; n < a->nNodes; n++) {
            if (!all && !A.chP[n] && !B.chP[n]) continue;
            r.push_back({A.PT64 + (size_t)n * pe, B.PT64 + (size_t)n * pe, pe * 8});
            r.push_back({A.PT32 + (size_t)n * pe, B.PT32 + (size_t)n * pe, pe * 4});
        }
    }
    if (!r.empty()) {
        // the descriptors live in pinned memory that the previous copy
        // launch of this tree may still be reading
        if (b->syncEv && !hipOk(hipEventSynchronize(b->syncEv), "sync")) return 1;
        if ((int)r.size() > b->rangeCap) {
            hostFree(b->hRanges);
            b->rangeCap = 0;
            const int cap = std::max<int>((int)r.size(), 64);
            if (!hostAlloc(&b->hRanges, (size_t)cap, "hipHostMalloc copy ranges") ||
                !hostDevPtr(b->hRanges, &b->dRanges, "copy ranges device address"))
                return 1;
            b->rangeCap = cap;
        }
        std::copy(r.begin(), r.end(), b->hRanges);
        if (!b->syncEv && !hipOk(hipEventCreateWithFlags(&b->syncEv, hipEventDisableTiming), "event")) return 1;
        if (!hipOk(launchCopyRanges(b->dRanges, (int)r.size(), s), "launch copy kernel") ||
            !hipOk(hipEventRecord(b->syncEv, s), "record copy event")) {
            (void)hipStreamSynchronize(s);
            return 1;
        }
    }
    for (int p = 0; p < a->nParts; p++) {
        TreePart &A = a->tp[p], &B = b->tp[p];
        B.dirty = A.dirty;
        B.pValid = A.pValid;
        B.everComputed = A.everComputed;
        std::fill(A.chSlot.begin(), A.chSlot.end(), 0);
        std::fill(B.chSlot.begin(), B.chSlot.end(), 0);
        std::fill(A.chP.begin(), A.chP.end(), 0);
        std::fill(B.chP.begin(), B.chP.end(), 0);
        A.chAll = B.chAll = false;
        if (a->partLikes && b->partLikes) b->partLikes[p] = a->partLikes[p];
    }
    b->logLike = a->logLike;
    return 0;
}

// Pf/p4_treeCopyVerify.c:63-180
int p4g_copyModelPrams(p4g_tree *a, p4g_tree *b)
{
    clearError();
    if (!a || !b || a->model->nParts != b->model->nParts) return 1;
    if ((a->tp[0].alloc && !flushAll(a)) || (b->tp[0].alloc && !flushAll(b))) return 1;
    hipStream_t s = stream();
    for (int p = 0; p < a->model->nParts; p++) {
        ModelPart *A = a->model->parts[p], *B = b->model->parts[p];
        const int dim = A->dim;
        for (int c = 0; c < A->nComps; c++)
            if (A->comps[c].free && A->comps[c].val != B->comps[c].val)
                std::memcpy(B->comps[c].val, A->comps[c].val, sizeof(double) * dim);
        for (int r = 0; r < A->nRMatrices; r++)
            if (A->rMatrices[r].free) {
                B->rMatrices[r].bigR = A->rMatrices[r].bigR;
                B->rMatrices[r].kappa = A->rMatrices[r].kappa;
            }
        TreePart &TA = a->tp[p], &TB = b->tp[p];
        if (TA.alloc && TB.alloc && TA.nEig == TB.nEig) {
            const size_t ne = (size_t)TA.nEig;
            if (!hipOk(hipMemcpyAsync(TB.eigV, TA.eigV, ne * dim * dim * 8, hipMemcpyDeviceToDevice, s), "copy eig") ||
                !hipOk(hipMemcpyAsync(TB.eigVi, TA.eigVi, ne * dim * dim * 8, hipMemcpyDeviceToDevice, s), "copy eig") ||
                !hipOk(hipMemcpyAsync(TB.eigLam, TA.eigLam, ne * dim * 8, hipMemcpyDeviceToDevice, s), "copy eig"))
                return 1;
            // the (pi, R) snapshots go with the eigensystems: P of tree b
            // may be computed from them directly
            std::memcpy(TB.hEigIn, TA.hEigIn, ne * (dim + dim * dim) * sizeof(double));
            TB.eigSnap = TA.eigSnap;
            TB.eigStale = TA.eigStale;
        }
        for (int g = 0; g < A->nGdasrvs; g++)
            if (A->gdasrvs[g]->free && A->gdasrvs[g] != B->gdasrvs[g]) {
                B->gdasrvs[g]->val[0] = A->gdasrvs[g]->val[0];
                for (int k = 0; k < A->nCat; k++) B->gdasrvs[g]->rates[k] = A->gdasrvs[g]->rates[k];
            }
        if (A->pInvarFree) B->pInvar = A->pInvar;
        if (a->model->doRelRates && a->model->relRatesAreFree) B->relRate = A->relRate;
    }
    return 0;
}

int p4g_verifyIdentityOfTwoTrees(p4g_tree *a, p4g_tree *b)
{
    clearError();
    if (!a || !b || !sameShape(a, b) || !treeReady(a) || !treeReady(b)) return 1;
    if (!flushAll(a) || !flushAll(b)) return 1;
    for (int p = 0; p < a->nParts; p++) {
        TreePart &A = a->tp[p], &B = b->tp[p];
        const size_t n = (size_t)a->nNodes * A.K * A.dim * A.dim;
        std::vector<double> pa(n), pb(n);
        if (!hipOk(hipMemcpy(pa.data(), A.PT64, n * 8, hipMemcpyDeviceToHost), "copy P") ||
            !hipOk(hipMemcpy(pb.data(), B.PT64, n * 8, hipMemcpyDeviceToHost), "copy P"))
            return 1;
        for (int node = 0; node < a->nNodes; node++) {
            if (node == a->root || !A.pValid[node] || !B.pValid[node]) continue;
            for (size_t i = (size_t)node * A.K * A.dim * A.dim; i < (size_t)(node + 1) * A.K * A.dim * A.dim; i++)
                if (std::fabs(pa[i] - pb[i]) > 1e-15) return 1;
        }
        if (A.clBytes != B.clBytes) return 1;
        const size_t nb = A.clBytes * A.nSlots;
        std::vector<char> ca(nb), cb(nb);
        if (!hipOk(hipMemcpy(ca.data(), A.cl, nb, hipMemcpyDeviceToHost), "copy partials") ||
            !hipOk(hipMemcpy(cb.data(), B.cl, nb, hipMemcpyDeviceToHost), "copy partials"))
            return 1;
        const size_t per = (size_t)A.nPatPad * A.K * A.dim;
        for (int node = 0; node < a->nNodes; node++) {
            const int sl = a->slotOf[node];
            if (sl < 0 || node == a->root) continue;
            if (a->precision == FP32) {
                const float *fa = (const float *)ca.data() + per * sl, *fb = (const float *)cb.data() + per * sl;
                for (size_t i = 0; i < per; i++)
                    if (std::fabs((double)fa[i] - (double)fb[i]) > 1e-15) return 1;
            } else {
                const double *da = (const double *)ca.data() + per * sl, *db = (const double *)cb.data() + per * sl;
                for (size_t i = 0; i < per; i++)
                    if (std::fabs(da[i] - db[i]) > 1e-15) return 1;
            }
        }
    }
    return 0;
}

// ------------------------------------------------------------ inspection
int p4g_getBigPDeck(const p4g_node *n, int pNum, double *out)
{
    clearError();
    if (!n || !out || pNum < 0 || pNum >= n->tree->nParts) return 1;
    const TreePart &tp = n->tree->tp[pNum];
    if (!tp.alloc || !tp.pValid[n->nodeNum]) {
        setError("getBigPDeck: P not computed");
        return 1;
    }
    const int K = tp.K, S = tp.dim;
    std::vector<double> h((size_t)K * S * S);
    if (!flushAll(n->tree) ||
        !hipOk(hipMemcpy(h.data(), tp.PT64 + (size_t)n->nodeNum * K * S * S, h.size() * 8, hipMemcpyDeviceToHost),
               "copy P"))
        return 1;
    for (int k = 0; k < K; k++)
        for (int i = 0; i < S; i++)
            for (int j = 0; j < S; j++) out[((size_t)k * S + i) * S + j] = h[((size_t)k * S + j) * S + i];
    return 0;
}

int p4g_getCondLikes(const p4g_node *n, int pNum, double *out)
{
    clearError();
    if (!n || !out || pNum < 0 || pNum >= n->tree->nParts) return 1;
    const Tree *t = n->tree;
    const TreePart &tp = t->tp[pNum];
    const int sl = t->slotOf.empty() ? -1 : t->slotOf[n->nodeNum];
    if (!tp.alloc || sl < 0 || n->nodeNum == t->root) {
        setError("getCondLikes: no stored partials for this node (leaf or root)");
        return 1;
    }
    const int K = tp.K, S = tp.dim, nP = tp.nPat;
    const size_t per = (size_t)tp.nPatPad * K * S;
    std::vector<int> e;
    if (!hipOk(hipStreamSynchronize(stream()), "sync")) return 1;
    std::vector<double> v(per);
    if (t->precision == FP32) {
        std::vector<float> f(per);
        e.resize(tp.sclPerSlot);
        if (!hipOk(hipMemcpy(f.data(), (const char *)tp.cl + tp.clBytes * sl, per * 4, hipMemcpyDeviceToHost), "copy") ||
            !hipOk(hipMemcpy(e.data(), tp.scl + (size_t)sl * tp.sclPerSlot, e.size() * 4, hipMemcpyDeviceToHost), "copy"))
            return 1;
        for (size_t i = 0; i < per; i++) v[i] = f[i];
    } else if (!hipOk(hipMemcpy(v.data(), (const char *)tp.cl + tp.clBytes * sl, per * 8, hipMemcpyDeviceToHost), "copy"))
        return 1;
    const bool tiled = pruneUsesMfma(S, K, t->precision);
    for (int k = 0; k < K; k++)
        for (int s = 0; s < S; s++)
            for (int p = 0; p < nP; p++) {
                // wave layout [k][p][s]; tile-blocked layout of the MFMA
                // kernel [tile][k][s/4][s%4][p%16]; else [p][k][s]
                double x = tp.wave ? v[((size_t)k * tp.nPatPad + p) * S + s]
                           : tiled ? v[(size_t)(p / 16) * K * S * 16 + ((size_t)(k * 5 + s / 4) * 4 + s % 4) * 16 + p % 16]
                                   : v[((size_t)p * K + k) * S + s];
                if (!e.empty()) x = std::ldexp(x, -e[tp.wave ? (size_t)k * tp.nPatPad + p : (size_t)p]);
                out[((size_t)k * S + s) * nP + p] = x;
            }
    return 0;
}

// ------------------------------------------- branch-length derivatives
int p4g_branchDerivatives(p4g_tree *t, int pNum, double *lnLEdge, double *d1, double *d2)
{
    clearError();
    if (!t || !lnLEdge || !d1 || !d2 || pNum >= t->nParts) {
        setError("branchDerivatives: bad arguments");
        return 1;
    }
    if (!treeReady(t)) return 1;
    if (t->precision != FP64) {
        setError("branchDerivatives: the tree's partials must be FP64 (setPrecision 0)");
        return 1;
    }
    const int N = t->nNodes;
    for (int n = 0; n < N; n++) lnLEdge[n] = d1[n] = d2[n] = 0.0;
    std::vector<double> out;
    double sum = 0.0;
    for (int p = 0; p < t->nParts; p++) {
        if (pNum >= 0 && p != pNum) continue;
        double v = 0.0;
        if (!evalPart(t, p, true, false, &v)) return 1;
        sum += v;
        if (!derivPart(t, p, out)) return 1;
        for (int n = 0; n < N; n++) {
            lnLEdge[n] += out[(size_t)n * 3];
            d1[n] += out[(size_t)n * 3 + 1];
            d2[n] += out[(size_t)n * 3 + 2];
        }
    }
    if (pNum < 0) t->logLike = sum;
    return 0;
}

// Up-pass partials of a non-root node as [nCat][dim][nPatterns] (the
// reference cl2 layout, Pf/p4_node.c:858-1018), after p4g_branchDerivatives.
int p4g_getCondLikes2(const p4g_node *n, int pNum, double *out)
{
    clearError();
    if (!n || !out || pNum < 0 || pNum >= n->tree->nParts) return 1;
    const Tree *t = n->tree;
    const TreePart &tp = t->tp[pNum];
    if (!tp.derivValid || n->nodeNum == t->root) {
        setError("getCondLikes2: no up-pass partials (call branchDerivatives; the root has none)");
        return 1;
    }
    const int K = tp.K, S = tp.dim, nP = tp.nPat;
    const size_t per = (size_t)K * tp.nPatPad * S;
    std::vector<double> v(per);
    if (!hipOk(hipStreamSynchronize(stream()), "sync") ||
        !hipOk(hipMemcpy(v.data(), tp.cl2 + (size_t)n->nodeNum * per, per * 8, hipMemcpyDeviceToHost), "copy cl2"))
        return 1;
    for (int k = 0; k < K; k++)
        for (int s = 0; s < S; s++)
            for (int p = 0; p < nP; p++) out[((size_t)k * S + s) * nP + p] = v[((size_t)k * tp.nPatPad + p) * S + s];
    return 0;
}

static int getPDeriv(const p4g_node *n, int pNum, double *out, int which)
{
    clearError();
    if (!n || !out || pNum < 0 || pNum >= n->tree->nParts) return 1;
    const Tree *t = n->tree;
    const TreePart &tp = t->tp[pNum];
    if (!tp.derivValid || n->nodeNum == t->root) {
        setError("getBigPDeck derivative: none stored (call branchDerivatives; the root has none)");
        return 1;
    }
    const int K = tp.K, S = tp.dim;
    std::vector<double> v((size_t)K * S * S);
    const double *src = (which == 1 ? tp.P1T : tp.P2T) + (size_t)n->nodeNum * K * S * S;
    if (!hipOk(hipStreamSynchronize(stream()), "sync") ||
        !hipOk(hipMemcpy(v.data(), src, v.size() * 8, hipMemcpyDeviceToHost), "copy P derivative"))
        return 1;
    for (int k = 0; k < K; k++)
        for (int i = 0; i < S; i++)
            for (int j = 0; j < S; j++) out[((size_t)k * S + i) * S + j] = v[((size_t)k * S + j) * S + i];
    return 0;
}

int p4g_getBigPDeck1stD(const p4g_node *n, int pNum, double *out) { return getPDeriv(n, pNum, out, 1); }
int p4g_getBigPDeck2ndD(const p4g_node *n, int pNum, double *out) { return getPDeriv(n, pNum, out, 2); }

// ------------------------------------------------------------ CPO / LPML
int p4g_cpoAccumulate(p4g_tree *t, double *lnL)
{
    clearError();
    if (!t || !lnL) {
        setError("cpoAccumulate: bad arguments");
        return 1;
    }
    if (!treeReady(t)) return 1;
    hipStream_t s = stream();
    for (int p = 0; p < t->nParts; p++) {
        Part *dp = t->data->parts[p];
        if (dp->cpoSum && dp->cpoPatterns != dp->nPatterns) {
            setError("cpoAccumulate: part " + std::to_string(p) + " changed its patterns since its CPO sums began; "
                     "call p4g_cpoReset");
            return 1;
        }
        if (!dp->cpoSum) {
            if (!devAlloc(&dp->cpoSum, (size_t)dp->nPatterns, "hipMalloc CPO sums") ||
                !hipOk(launchFillU32(reinterpret_cast<unsigned *>(dp->cpoSum), 0u, 2 * dp->nPatterns, s),
                       "clear CPO sums"))
                return 1;
            dp->cpoPatterns = dp->nPatterns;
            dp->cpoSamples = 0;
        }
    }
    double sum = 0.0;
    bool bad = false;
    for (int p = 0; p < t->nParts; p++) {
        double v = 0.0;
        if (!evalPart(t, p, true, false, &v, nullptr, t->data->parts[p]->cpoSum)) return 1;
        bad |= t->tp[p].hResult[1] != 0.0;
        sum += v;
    }
    for (int p = 0; p < t->nParts; p++) t->data->parts[p]->cpoSamples++;
    t->logLike = sum;
    *lnL = sum;
    if (bad) {
        setError("cpoAccumulate: a site likelihood is zero, so its inverse (and the CPO sums) are infinite");
        return 1;
    }
    return 0;
}

int p4g_cpoGetSums(const p4g_part *p, double *out, int *nSamples)
{
    clearError();
    if (!p || !out || !nSamples) {
        setError("cpoGetSums: bad arguments");
        return 1;
    }
    *nSamples = p->cpoSamples;
    const int nSites = ((int)p->seqPosPatternIndex.size() == p->nChar && p->nChar > 0) ? p->nChar : p->nPatterns;
    if (!p->cpoSum) {
        std::fill(out, out + nSites, 0.0);
        return 0;
    }
    std::vector<double> h(p->cpoPatterns);
    if (!hipOk(hipStreamSynchronize(stream()), "sync") ||
        !hipOk(hipMemcpy(h.data(), p->cpoSum, h.size() * 8, hipMemcpyDeviceToHost), "copy CPO sums"))
        return 1;
    if (nSites == p->nChar && (int)p->seqPosPatternIndex.size() == p->nChar)
        for (int i = 0; i < nSites; i++) out[i] = h[p->seqPosPatternIndex[i]];
    else
        std::copy(h.begin(), h.end(), out);
    return 0;
}

int p4g_cpoReset(p4g_part *p)
{
    clearError();
    if (!p) return 1;
    if (p->cpoSum) {
        if (!hipOk(hipStreamSynchronize(stream()), "sync")) return 1;
        devFree(p->cpoSum);
    }
    p->cpoPatterns = 0;
    p->cpoSamples = 0;
    return 0;
}

int p4g_partNSites(const p4g_part *p)
{
    if (!p) return -1;
    return ((int)p->seqPosPatternIndex.size() == p->nChar && p->nChar > 0) ? p->nChar : p->nPatterns;
}

void p4g_philox4x64(const unsigned long long *counter, const unsigned long long *key, unsigned long long *out)
{
    Philox4 c{{counter[0], counter[1], counter[2], counter[3]}};
    const Philox4 r = philox4x64(c, key[0], key[1]);
    for (int i = 0; i < 4; i++) out[i] = r.v[i];
}

// ------------------------------------------------------ ancestral draws
}  // extern "C"

namespace p4g {
// Launch the root draws of sites [first, first + n) of part pNum; the draws
// land in the tree part's buffer (*dDraw, device, [n][4]), valid until the
// next draw on that part.  u: host uniforms or null.
static bool ancLaunch(Tree *t, int pNum, int first, int n, const double *u, uint64_t seed, int **dDrawOut)
{
    Part *dp = t->data->parts[pNum];
    TreePart &tp = t->tp[pNum];
    ModelPart *mp = t->model->parts[pNum];
    if ((long long)first + n > dp->dev.nSites) {
        setError("drawAncStates: sites [" + std::to_string(first) + ", " + std::to_string(first + n) +
                 ") run past the part's " + std::to_string(dp->dev.nSites) + " sites");
        return false;
    }
    if (tp.K > ANC_MAX_K) {
        setError("drawAncStates: at most 16 gamma categories");
        return false;
    }
    // bring the partials up to date (Tree.ancestralStateDraw evaluates first)
    double v = 0.0;
    if (!evalPart(t, pNum, !tp.everComputed, false, &v)) return false;
    std::vector<DerivSib> kids;
    for (int c = t->nodes[t->root]->leftChild; c >= 0; c = t->nodes[c]->sibling) {
        const Node *cn = t->nodes[c];
        kids.push_back(DerivSib{cn->isLeaf ? 0 : 1, cn->isLeaf ? cn->seqNum : t->slotOf[c], c, 0});
    }
    const size_t kidBytes = (std::max<size_t>(kids.size(), 1) * sizeof(DerivSib) + 255) & ~(size_t)255;
    const size_t uBytes = u ? (((size_t)n * 8 + 255) & ~(size_t)255) : 0;
    const size_t need = kidBytes + uBytes + (size_t)n * 16;
    hipStream_t s = stream();
    if (need > tp.ancCap) {
        if (tp.ancBuf) {
            if (!hipOk(hipStreamSynchronize(s), "sync")) return false;
            (void)hipFree(tp.ancBuf);
            tp.ancBuf = nullptr;
            tp.ancCap = 0;
        }
        if (!hipOk(hipMalloc(&tp.ancBuf, need), "hipMalloc ancestral draw buffer")) return false;
        tp.ancCap = need;
    }
    DerivSib *dKids = reinterpret_cast<DerivSib *>(tp.ancBuf);
    double *dU = u ? reinterpret_cast<double *>(tp.ancBuf + kidBytes) : nullptr;
    int *dDraw = reinterpret_cast<int *>(tp.ancBuf + kidBytes + uBytes);
    AncArgs a{};
    a.dim = tp.dim;
    a.K = tp.K;
    a.nPatPad = tp.nPatPad;
    a.layout = tp.wave ? CL_WAVE : pruneUsesMfma(tp.dim, tp.K, t->precision) ? CL_TILED : CL_GROUPED;
    a.fp32 = t->precision == FP32 ? 1 : 0;
    a.cl = tp.cl;
    a.slotBytes = tp.clBytes;
    a.scl = tp.scl;
    a.sclPerSlot = tp.sclPerSlot;
    a.nKid = (int)kids.size();
    a.kids = dKids;
    a.tips = dp->dev.tips;
    a.ambMask = dp->dev.ambMask;
    a.PT64 = tp.PT64;
    a.rootPi = reinterpret_cast<const double *>(tp.dArena + tp.offPi);   // staged by the evaluation above
    a.pInvar = mp->pInvar;
    a.invMask = dp->dev.invMask;
    a.nWords = dp->dev.nWords;
    a.sitePat = dp->dev.sitePat;
    a.first = first;
    a.n = n;
    a.u = dU;
    a.seed = seed;
    a.draw = dDraw;
    if ((!kids.empty() &&
         !hipOk(hipMemcpyAsync(dKids, kids.data(), kids.size() * sizeof(DerivSib), hipMemcpyHostToDevice, s),
                "stage root children")) ||
        (u && !hipOk(hipMemcpyAsync(dU, u, (size_t)n * 8, hipMemcpyHostToDevice, s), "stage uniforms")) ||
        !hipOk(launchAncDraw(a, s), "launch ancestral draw kernel")) {
        (void)hipStreamSynchronize(s);
        return false;
    }
    *dDrawOut = dDraw;
    return true;
}
}  // namespace p4g

extern "C" {

int p4g_drawAncStates(p4g_tree *t, int pNum, int first, int n, const double *u, unsigned long long seed, int *draw)
{
    clearError();
    if (!t || !draw || pNum < 0 || pNum >= t->nParts || first < 0 || n < 0) {
        setError("drawAncStates: bad arguments");
        return 1;
    }
    if (n == 0) return 0;
    if (!treeReady(t)) return 1;
    int *dDraw = nullptr;
    hipStream_t s = stream();
    if (!ancLaunch(t, pNum, first, n, u, (uint64_t)seed, &dDraw)) return 1;
    if (!hipOk(hipMemcpyAsync(draw, dDraw, (size_t)n * 16, hipMemcpyDeviceToHost, s), "copy draws") ||
        !hipOk(hipStreamSynchronize(s), "ancestral draw kernel")) {
        (void)hipStreamSynchronize(s);
        return 1;
    }
    for (int i = 0; i < n; i++)
        if (draw[4 * i + 1] < 0 && draw[4 * i + 2] == 0) {
            setError("drawAncStates: site " + std::to_string(first + i) +
                     " has a zero (or non-finite) root likelihood; nothing to draw from");
            return 1;
        }
    return 0;
}

// ------------------------------------------------------------ simulation
}  // extern "C"

namespace p4g {
// The part's device sequence buffer, [nTax][nChar] bytes rounded up to 4.
static bool partSeqBuffer(Part *p)
{
    const size_t need = ((size_t)p->nTax * p->nChar + 3) & ~(size_t)3;
    if (p->dSeq && p->dSeqCap >= need) return true;
    if (p->dSeq) {
        if (!hipOk(hipStreamSynchronize(stream()), "sync")) return false;
        devFree(p->dSeq);
    }
    p->dSeqCap = 0;
    p->dSeqValid = false;
    if (!hipOk(hipMalloc(&p->dSeq, std::max<size_t>(need, 4)), "hipMalloc sequences")) return false;
    p->dSeqCap = std::max<size_t>(need, 4);
    return true;
}
}  // namespace p4g

extern "C" {

int p4g_simulate(p4g_tree *t, p4g_tree *refTree, unsigned long long seed)
{
    clearError();
    if (!t) {
        setError("simulate: bad arguments");
        return 1;
    }
    if (refTree == t) {
        setError("simulate: the tree and the reference tree are the same tree");
        return 1;
    }
    if (!treeReady(t, false)) return 1;
    if (refTree) {
        if (!treeReady(refTree)) return 1;
        if (refTree->nParts != t->nParts) {
            setError("simulate: the reference tree has a different number of parts");
            return 1;
        }
        for (int p = 0; p < t->nParts; p++)
            if (refTree->data->parts[p]->dev.nSites != t->data->parts[p]->nChar) {
                setError("simulate: part " + std::to_string(p) + " of the reference tree has " +
                         std::to_string(refTree->data->parts[p]->dev.nSites) + " sites, the tree's " +
                         std::to_string(t->data->parts[p]->nChar));
                return 1;
            }
    }
    hipStream_t s = stream();
    const int N = t->nNodes;
    // preorder non-root nodes, parents, tip rows
    std::vector<int> host;
    host.reserve(3 * N);
    for (int j = 0; j < N; j++) {
        const int n = t->preOrder[j];
        if (n == NO_ORDER || n == t->root) continue;
        if (n < 0 || n >= N) {
            setError("preOrder holds an out-of-range node number");
            return 1;
        }
        host.push_back(n);
    }
    const int nPre = (int)host.size();
    host.resize((size_t)nPre + 2 * N);
    for (int n = 0; n < N; n++) {
        host[nPre + n] = t->nodes[n]->parent;
        host[nPre + N + n] = t->nodes[n]->isLeaf ? t->nodes[n]->seqNum : -1;
    }
    for (int p = 0; p < t->nParts; p++) {
        Part *dp = t->data->parts[p];
        ModelPart *mp = t->model->parts[p];
        const int nChar = dp->nChar, S = mp->dim, nTax = dp->nTax;
        for (int n = 0; n < N; n++)
            if (t->nodes[n]->isLeaf && (t->nodes[n]->seqNum < 0 || t->nodes[n]->seqNum >= nTax)) {
                setError("simulate: leaf " + std::to_string(n) + " has no sequence row");
                return 1;
            }
        const int rc = t->nodes[t->root]->compNums[p];
        if (rc < 0 || rc >= mp->nComps || !mp->comps[rc].val) {
            setError("simulate: root comp out of range");
            return 1;
        }
        // P of every non-root node from the current model and branch lengths
        // (Pf/p4_treeSim.c:362-368)
        for (int n = 0; n < N; n++)
            if (n != t->root && !bigPDecksPart(t, t->nodes[n], p)) return 1;
        if (!flushParams(t, p)) {
            (void)hipStreamSynchronize(s);
            return 1;
        }
        int *dRef = nullptr;
        if (refTree && !ancLaunch(refTree, p, 0, nChar, nullptr, (uint64_t)seed ^ 0x9E3779B97F4A7C15ull, &dRef))
            return 1;
        auto up256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
        const size_t offComp = up256(host.size() * 4), offStates = offComp + up256((size_t)S * 8),
                     offCat = offStates + up256((size_t)N * nChar), need = offCat + (size_t)nChar * 4;
        if (need > t->simCap) {
            if (!hipOk(hipStreamSynchronize(s), "sync")) return 1;
            if (t->simBuf) (void)hipFree(t->simBuf);
            t->simBuf = nullptr;
            t->simCap = 0;
            if (!hipOk(hipMalloc(&t->simBuf, need), "hipMalloc simulation buffer")) return 1;
            t->simCap = need;
        }
        if (!partSeqBuffer(dp)) return 1;
        SimArgs a{};
        a.dim = S;
        a.K = mp->nCat;
        a.nChar = nChar;
        a.nNodes = N;
        a.pNum = p;
        a.root = t->root;
        a.nPre = nPre;
        a.pre = reinterpret_cast<const int *>(t->simBuf);
        a.parent = a.pre + nPre;
        a.seqNum = a.pre + nPre + N;
        a.PT64 = t->tp[p].PT64;
        a.rootComp = reinterpret_cast<const double *>(t->simBuf + offComp);
        a.pInvar = mp->pInvar;
        a.refDraw = dRef;
        a.seed = (uint64_t)seed;
        a.states = reinterpret_cast<uint8_t *>(t->simBuf + offStates);
        a.seq = dp->dSeq;
        a.simCat = reinterpret_cast<int *>(t->simBuf + offCat);
        std::vector<int8_t> seq((size_t)nTax * nChar, 0);
        dp->simCats.assign(nChar, 0);
        if (!hipOk(hipMemcpyAsync(t->simBuf, host.data(), host.size() * 4, hipMemcpyHostToDevice, s), "stage nodes") ||
            !hipOk(hipMemcpyAsync(t->simBuf + offComp, mp->comps[rc].val, (size_t)S * 8, hipMemcpyHostToDevice, s),
                   "stage root comp") ||
            !hipOk(launchFillU32(reinterpret_cast<unsigned *>(dp->dSeq), 0u, (int)(dp->dSeqCap / 4), s),
                   "clear sequences") ||
            !hipOk(launchSimulate(a, s), "launch simulation kernel") ||
            !hipOk(hipMemcpyAsync(seq.data(), dp->dSeq, seq.size(), hipMemcpyDeviceToHost, s), "copy sequences") ||
            !hipOk(hipMemcpyAsync(dp->simCats.data(), a.simCat, (size_t)nChar * 4, hipMemcpyDeviceToHost, s),
                   "copy categories") ||
            !hipOk(hipStreamSynchronize(s), "simulation kernel")) {
            (void)hipStreamSynchronize(s);
            return 1;
        }
        // the part now holds the simulated sequences; its patterns are
        // stale until they are made again (Pf/p4_treeSim.c:64-69)
        dp->sequences.assign(seq.begin(), seq.end());
        dp->dSeqValid = true;
        dp->nPatterns = 0;
        dp->patterns.clear();
        dp->patternCounts.clear();
        dp->invarSet = false;
        dp->siteLikes.clear();
        dp->dev.valid = false;
    }
    return 0;
}

// Pf/part.c:604-680
int p4g_symbolSequences(const p4g_part *p, char *out)
{
    clearError();
    if (!p || !out) return 1;
    const size_t n = (size_t)p->nTax * p->nChar;
    if (p->nChar < 1 || p->sequences.size() != n) {
        setError("symbolSequences: no sequences");
        return 1;
    }
    for (size_t i = 0; i < n; i++) {
        const int c = p->sequences[i];
        if (c >= 0 && c < (int)p->symbols.size())
            out[i] = p->symbols[c];
        else if (c == GAP_CODE)
            out[i] = '-';
        else if (c == QMARK_CODE)
            out[i] = '?';
        else if (c >= EQUATES_BASE && c < EQUATES_BASE + p->nEquates)
            out[i] = p->equateSymbols[c - EQUATES_BASE];
        else {
            setError("symbolSequences: character number " + std::to_string(c) + " is not recognized");
            return 1;
        }
    }
    return 0;
}

int p4g_partSimCats(const p4g_part *p, int *cats)
{
    clearError();
    if (!p || !cats) return 1;
    if ((int)p->simCats.size() != p->nChar) {
        setError("partSimCats: the part holds no simulation");
        return 1;
    }
    std::copy(p->simCats.begin(), p->simCats.end(), cats);
    return 0;
}

int p4g_partSeqStats(p4g_part *p, int *compCounts, int *nConstant, double *meanNChars, int *nNonState)
{
    clearError();
    if (!p || !compCounts || !nConstant || !meanNChars || !nNonState) {
        setError("partSeqStats: bad arguments");
        return 1;
    }
    if (!deviceReady()) return 1;
    hipStream_t s = stream();
    const size_t n = (size_t)p->nTax * p->nChar;
    if (n == 0 || p->sequences.size() != n) {
        setError("partSeqStats: the part has no sequences");
        return 1;
    }
    if (!p->dSeqValid) {
        std::vector<int8_t> h(n);
        for (size_t i = 0; i < n; i++) h[i] = (int8_t)p->sequences[i];   // exact codes: -64..63
        if (!partSeqBuffer(p)) return 1;
        if (!hipOk(hipMemcpyAsync(p->dSeq, h.data(), n, hipMemcpyHostToDevice, s), "copy sequences") ||
            !hipOk(hipStreamSynchronize(s), "copy sequences"))
            return 1;
        p->dSeqValid = true;
    }
    const size_t nCounts = (size_t)p->nTax * p->dim;
    int *d = nullptr;
    if (!hipOk(hipMalloc(&d, (nCounts + 4) * 4), "hipMalloc statistics")) return 1;
    std::vector<int> h(nCounts + 4);
    const bool ok = hipOk(launchFillU32(reinterpret_cast<unsigned *>(d), 0u, (int)(nCounts + 4), s), "clear") &&
                    hipOk(launchSeqStats(p->dSeq, p->nTax, p->nChar, p->dim, d + 4, d, s), "launch statistics") &&
                    hipOk(hipMemcpyAsync(h.data(), d, h.size() * 4, hipMemcpyDeviceToHost, s), "copy statistics") &&
                    hipOk(hipStreamSynchronize(s), "statistics kernel");
    (void)hipFree(d);
    if (!ok) return 1;
    std::copy(h.begin() + 4, h.end(), compCounts);
    *nConstant = h[0];
    *meanNChars = (double)h[1] / (double)p->nChar;
    *nNonState = h[2];
    return 0;
}

double p4g_partBigXSquared(p4g_part *p)
{
    clearError();
    if (!p) return -3.0;
    std::vector<int> c((size_t)p->nTax * p->dim);
    int nConst = 0, nOther = 0;
    double mean = 0.0;
    if (p4g_partSeqStats(p, c.data(), &nConst, &mean, &nOther)) return -3.0;
    if (nOther) return -2.0;   // Pf/part.c:1519-1525: no gaps or ambiguities
    // Pf/part.c:1530-1560, in its order of operations
    const double oneOverNTax = 1.0 / p->nTax;
    double xSq = 0.0;
    std::vector<double> ex(p->dim);
    for (int j = 0; j < p->dim; j++) {
        double soc = 0.0;
        for (int t = 0; t < p->nTax; t++) soc += c[(size_t)t * p->dim + j];
        ex[j] = oneOverNTax * soc;
    }
    for (int t = 0; t < p->nTax; t++)
        for (int j = 0; j < p->dim; j++)
            if (ex[j] != 0.0) {
                const double d = c[(size_t)t * p->dim + j] - ex[j];
                xSq += (d * d) / ex[j];
            }
    return xSq;
}

double p4g_getUnconstrainedLogLike(const p4g_part *p)
{
    clearError();
    if (!p || p->nPatterns < 1) {
        setError("getUnconstrainedLogLike: no patterns");
        return 0.0;
    }
    double dsum = 0.0;   // Pf/part.c:682-714
    for (int i = 0; i < p->nPatterns; i++) dsum += p->patternCounts[i] * std::log((double)p->patternCounts[i]);
    return dsum - p->nChar * std::log((double)p->nChar);
}

// -------------------------------------------------------------- batches
}  // extern "C"

struct p4g_batch {
    std::vector<Tree *> trees;
    PruneArgs *dArgs = nullptr;
    int *dBlockItem = nullptr;
    PruneArgs *hArgs = nullptr;
    int *hBlockItem = nullptr;
    int capItems = 0, capBlocks = 0;
    // deferred 20-state direct-P flushes, read by the kernel from pinned memory
    PmatArgs *hPm = nullptr;
    const double **hPmIn = nullptr;
    int *hPmFirst = nullptr;
    int capPm = 0;
    double *hRes = nullptr;
    CopyRange *hRanges = nullptr, *dRanges = nullptr;   // program uploads (pinned)
    int capRanges = 0;
    ~p4g_batch()
    {
        hostFree(hRanges);
        devFree(dArgs);
        devFree(dBlockItem);
        hostFree(hArgs);
        hostFree(hBlockItem);
        hostFree(hRes);
        hostFree(hPm);
        hostFree(hPmIn);
        hostFree(hPmFirst);
    }
};

extern "C" {

p4g_batch *p4g_newBatch(int nTrees, p4g_tree **trees)
{
    clearError();
    if (nTrees < 1 || !trees) {
        setError("newBatch: bad arguments");
        return nullptr;
    }
    auto *b = new p4g_batch();
    for (int i = 0; i < nTrees; i++) b->trees.push_back(trees[i]);
    return b;
}

void p4g_freeBatch(p4g_batch *b) { delete b; }

// All parts of all trees: pending partials + site sums, one launch per
// (state count, precision) group.
int p4g_batchLogLike(p4g_batch *b, double *lnL)
{
    clearError();
    if (!b || !lnL) return 1;
    struct Item {
        Tree *t;
        int p, nOps, tree;
    };
    std::vector<Item> items;
    std::vector<DeferredP> defer;
    std::vector<CopyRange> uploads;
    // the items' program and root-frequency uploads, in one copy launch
    auto launchUploads = [&]() -> bool {
        if (uploads.empty()) return true;
        if ((int)uploads.size() > b->capRanges) {
            hostFree(b->hRanges);
            b->capRanges = 0;
            const int cap = std::max<int>(2 * (int)uploads.size(), 64);
            if (!hostAlloc(&b->hRanges, (size_t)cap, "hipHostMalloc upload ranges") ||
                !hostDevPtr(b->hRanges, &b->dRanges, "upload ranges device address"))
                return false;
            b->capRanges = cap;
        }
        std::copy(uploads.begin(), uploads.end(), b->hRanges);
        return hipOk(launchCopyRanges(b->dRanges, (int)uploads.size(), stream()), "launch upload kernel");
    };
    for (size_t i = 0; i < b->trees.size(); i++) {
        Tree *t = b->trees[i];
        if (!treeReady(t)) {
            launchDeferredEach(defer);
            (void)launchUploads();
            (void)hipStreamSynchronize(stream());
            return 1;
        }
        for (int p = 0; p < t->nParts; p++) {
            int nOps = 0;
            if (!stagePart(t, p, false, &nOps, &defer, &uploads)) {
                launchDeferredEach(defer);
                (void)launchUploads();
                (void)hipStreamSynchronize(stream());
                return 1;
            }
            items.push_back({t, p, nOps, (int)i});
        }
    }
    // error paths once kernels are queued: they read pinned staging memory
    // (arena, batch argument arrays) that the next call rewrites, so wait
    auto failQueued = [&]() {
        (void)hipStreamSynchronize(stream());
        return 1;
    };
    if (!launchUploads()) return failQueued();
    if (!defer.empty()) {
        // the deferred direct-P flushes in one launch per state count, before
        // any pruning; the arrays hold the 20-state items, then the 4-state
        const int nD = (int)defer.size();
        if (nD > b->capPm) {
            hostFree(b->hPm);
            hostFree(b->hPmIn);
            hostFree(b->hPmFirst);
            b->capPm = 2 * nD;
            if (!hostAlloc(&b->hPm, b->capPm, "hipHostMalloc batch P args") ||
                !hostAlloc(&b->hPmIn, b->capPm, "hipHostMalloc batch P inputs") ||
                !hostAlloc(&b->hPmFirst, b->capPm + 3, "hipHostMalloc batch P map")) {
                b->capPm = 0;
                launchDeferredEach(defer);
                return 1;
            }
        }
        PmatArgs *dPm = nullptr;
        const double **dPmIn = nullptr;
        int *dPmFirst = nullptr;
        if (!hostDevPtr(b->hPm, &dPm, "device address of batch P args") ||
            !hostDevPtr(b->hPmIn, &dPmIn, "device address of batch P inputs") ||
            !hostDevPtr(b->hPmFirst, &dPmFirst, "device address of batch P map")) {
            launchDeferredEach(defer);
            return 1;
        }
        int q = 0, f = 0;
        for (int dim : {20, 4}) {
            const int q0 = q, f0 = f;
            int total = 0;
            for (const DeferredP &dp : defer)
                if (dp.in && dp.a.dim == dim) {
                    b->hPm[q] = dp.a;
                    b->hPmIn[q] = dp.devIn ? dp.devIn : dp.in;   // the uploads ran before this launch
                    b->hPmFirst[f++] = total;
                    total += dp.a.nItems * dp.a.K;
                    q++;
                }
            b->hPmFirst[f++] = total;
            if (!hipOk(launchPexpMulti(dim, dPm + q0, dPmIn + q0, dPmFirst + f0, q - q0, total, stream()),
                       "launch batched P kernel"))
                return failQueued();
        }
        // the eigen route's items (after their eigensystem launches)
        const int q0 = q, f0 = f;
        int total = 0;
        for (const DeferredP &dp : defer)
            if (!dp.in) {
                b->hPm[q] = dp.a;
                b->hPmIn[q] = nullptr;
                b->hPmFirst[f++] = total;
                total += dp.a.nItems * dp.a.K * dp.a.dim * dp.a.dim;
                q++;
            }
        b->hPmFirst[f++] = total;
        if (!hipOk(launchPmatMulti(dPm + q0, dPmFirst + f0, q - q0, total, stream()), "launch batched P kernel"))
            return failQueued();
    }
    const int nItems = (int)items.size();
    int nBlocksTot = 0;
    auto itemBlocks = [](const Item &it) {
        const ModelPart *mp = it.t->model->parts[it.p];
        const TreePart &tp = it.t->tp[it.p];
        return tp.wave ? tp.nPatPad / pruneWaveTilePatterns(it.t->precision)
                       : tp.nPatPad / pruneBlockPatternsFor(mp->dim, mp->nCat, it.t->precision, it.nOps);
    };
    for (auto &it : items) nBlocksTot += itemBlocks(it);
    if (nItems > b->capItems) {
        devFree(b->dArgs);
        hostFree(b->hArgs);
        hostFree(b->hRes);
        b->capItems = nItems * 2;
        if (!devAlloc(&b->dArgs, b->capItems, "hipMalloc batch args") ||
            !hostAlloc(&b->hArgs, b->capItems, "hipHostMalloc batch args") ||
            !hostAlloc(&b->hRes, (size_t)b->capItems * 2, "hipHostMalloc batch results"))
            return failQueued();
    }
    if (nBlocksTot > b->capBlocks) {
        devFree(b->dBlockItem);
        hostFree(b->hBlockItem);
        b->capBlocks = nBlocksTot * 2;
        if (!devAlloc(&b->dBlockItem, (size_t)b->capBlocks * 2, "hipMalloc batch map") ||
            !hostAlloc(&b->hBlockItem, (size_t)b->capBlocks * 2, "hipHostMalloc batch map"))
            return failQueued();
    }
    // group by (dim, precision): items of a group are contiguous in hArgs
    std::vector<int> order(nItems);
    for (int i = 0; i < nItems; i++) order[i] = i;
    auto key = [&](int i) {
        const ModelPart *mp = items[i].t->model->parts[items[i].p];
        const int prec = items[i].t->precision;
        // MFMA kernels: one launch per (category count, kernel choice) (the
        // category-split kernel's workgroup is 64 * K threads)
        const int kc = pruneUsesMfma(mp->dim, mp->nCat, prec)
                           ? 2 * mp->nCat + (pruneAaCategoryWaves(items[i].nOps) ? 1 : 0)
                           : 0;
        // the wave kernel's workgroup size is 64 * K: one launch per K
        const TreePart &tq = items[i].t->tp[items[i].p];
        const int wk = tq.wave ? 16 * tq.nStream + mp->nCat : 0;
        return ((mp->dim * 2 + prec) * 64 + kc) * 32 + wk;
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key(x) < key(y); });
    double *dRes = nullptr;
    if (!hostDevPtr(b->hRes, &dRes, "device address of batch results")) return failQueued();
    for (int j = 0; j < nItems; j++) {
        const Item &it = items[order[j]];
        fillPruneArgs(it.t, it.p, it.nOps, b->hArgs[j], false);
        b->hArgs[j].result = dRes + 2 * j;  // lnL and bad flag straight into pinned memory
    }
    hipStream_t s = stream();
    if (!hipOk(hipMemcpyAsync(b->dArgs, b->hArgs, sizeof(PruneArgs) * nItems, hipMemcpyHostToDevice, s),
               "stage batch args"))
        return failQueued();
    int j = 0, blk = 0;
    std::vector<std::pair<int, int>> launches;  // (first block, count) per group
    std::vector<int> groupDim, groupPrec;
    std::vector<size_t> groupLds;
    std::vector<int> groupK, groupWave, groupStreams, groupAaCat, groupAaTip;
    std::vector<size_t> groupAaLds;
    while (j < nItems) {
        const int k0 = key(order[j]);
        const int first = blk;
        int jj = j;
        while (jj < nItems && key(order[jj]) == k0) {
            const Item &it = items[order[jj]];
            const int nb = itemBlocks(it);
            for (int q = 0; q < nb; q++) {
                b->hBlockItem[2 * blk] = jj;
                b->hBlockItem[2 * blk + 1] = q;
                blk++;
            }
            jj++;
        }
        launches.push_back({first, blk - first});
        groupDim.push_back(items[order[j]].t->model->parts[items[order[j]].p]->dim);
        groupPrec.push_back(items[order[j]].t->precision);
        groupK.push_back(items[order[j]].t->model->parts[items[order[j]].p]->nCat);
        groupAaCat.push_back(pruneAaCategoryWaves(items[order[j]].nOps) ? 1 : 0);
        {
            // the largest slot file and tip codes of the launch's 20-state category items
            size_t nl = 0;
            int allTip = 1;
            const bool aaCat = groupAaCat.back() && pruneUsesMfma(groupDim.back(), groupK.back(), groupPrec.back());
            for (int q = j; q < jj && aaCat; q++) {
                const PruneArgs &pa = b->hArgs[q];
                if (!pruneAaCatItemOk(pa)) {
                    setError("batch: bad 20-state category program");
                    return failQueued();
                }
                nl = std::max(nl, pruneAaLdsBytes(pa.K, pa.nLds, pa.tipLds ? pa.nTax : 0));
                allTip &= pa.tipLds ? 1 : 0;
            }
            groupAaLds.push_back(nl);
            groupAaTip.push_back(allTip);
        }
        groupWave.push_back(items[order[j]].t->tp[items[order[j]].p].wave ? 1 : 0);
        groupStreams.push_back(items[order[j]].t->tp[items[order[j]].p].nStream);
        size_t lds = 0;
        if (groupWave.back()) {
            // the largest slot file of the launch's items
            for (int q = j; q < jj; q++) lds = std::max(lds, items[order[q]].t->tp[items[order[q]].p].waveLds);
            groupLds.push_back(lds);
        } else {
            // P in LDS for the whole launch only if every item's P rows fit
            bool fits = true;
            for (int q = j; q < jj; q++) {
                const PruneArgs &pa = b->hArgs[q];
                const size_t l = pruneLdsBytes(pa.nKids, pa.K, pa.dim, groupPrec.back());
                if (!l) fits = false;
                lds = std::max(lds, l);
            }
            groupLds.push_back(fits ? lds : 0);
        }
        j = jj;
    }
    if (!hipOk(hipMemcpyAsync(b->dBlockItem, b->hBlockItem, sizeof(int) * 2 * blk, hipMemcpyHostToDevice, s),
               "stage batch map"))
        return failQueued();
    // P^T of each wave item's children in program order
    for (int q = 0; q < nItems; q++) {
        const Item &it = items[order[q]];
        const TreePart &tp = it.t->tp[it.p];
        if (!tp.wave || !pruneWaveVectorP(it.t->precision)) continue;
        const PruneArgs &pa = b->hArgs[q];
        if (!hipOk(launchPgather(pa.PT, pa.wpre + 3 * pa.nPre, tp.nKidTot, tp.K, tp.PP, it.t->precision, s),
                   "launch P gather"))
            return failQueued();
    }
    for (size_t g = 0; g < launches.size(); g++) {
        const int *bi = b->dBlockItem + 2 * launches[g].first;
        const hipError_t le = groupWave[g] ? launchPruneWaveBatch(b->dArgs, bi, launches[g].second, groupK[g],
                                                                 groupStreams[g], groupPrec[g], groupLds[g], s)
                              : pruneUsesMfma(groupDim[g], groupK[g], groupPrec[g])
                                  ? launchPruneAaBatch(b->dArgs, bi, launches[g].second, groupK[g], groupAaCat[g] != 0,
                                                       groupAaLds[g], groupAaTip[g] != 0, s)
                                  : launchPruneBatch(b->dArgs, bi, launches[g].second, groupDim[g], groupPrec[g],
                                                     groupLds[g], s);
        if (!hipOk(le, "launch batched pruning kernel")) return failQueued();
    }
    if (!hipOk(hipStreamSynchronize(s), "batched pruning kernel")) return 1;
    for (size_t i = 0; i < b->trees.size(); i++) lnL[i] = 0.0;
    for (int q = 0; q < nItems; q++) {
        const Item &it = items[order[q]];
        const double v = b->hRes[2 * q];
        lnL[it.tree] += v;
        finishPart(it.t, it.p, false, v, b->hRes[2 * q + 1] != 0.0);
    }
    for (size_t i = 0; i < b->trees.size(); i++) b->trees[i]->logLike = lnL[i];
    return 0;
}

}  // extern "C"
