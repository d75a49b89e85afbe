// p4g_compress.hip -- site-pattern compression on the GPU (the device form
// of makePatterns, Pf/part.c:317-448): the distinct columns of a part's
// [taxon][site] codes, in order of first occurrence, with their counts and
// the pattern of every site.
//
//  1. k_colHash     one thread per site: a 64-bit hash of the site's column
//                   (FNV-1a over the codes, then a splitmix finaliser).
//  2. radix sort    (hash, site) pairs by hash (hipCUB; stable, so equal
//                   hashes keep ascending site order).
//  3. k_segStart    sorted position i -> i where a run of equal hashes
//                   starts, else 0; an inclusive max-scan then gives every
//                   position the start of its run, whose site is the run's
//                   first occurrence.
//  4. k_verify      every column compared with its run's first column; a
//                   difference (a hash collision) is flagged and the host
//                   runs the pipeline again with another hash seed.
//  5. exclusive scan of the first-occurrence flags in site order numbers
//                   the patterns in order of first occurrence.
//  6. k_assign      the pattern of every site, the counts, and the site of
//                   each pattern; k_gather writes the pattern columns.
#include "p4g_internal.h"

#include <hipcub/hipcub.hpp>

namespace p4g {

__global__ __launch_bounds__(256) void k_colHash(const int8_t *seq, int nTax, int nChar, uint64_t seed,
                                                 uint64_t *keys, int *vals)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= nChar) return;
    uint64_t h = 1469598103934665603ull ^ seed;
    for (int t = 0; t < nTax; t++) {
        h ^= (uint64_t)(uint8_t)seq[(size_t)t * nChar + c];
        h *= 1099511628211ull;
    }
    // splitmix64 finaliser
    h ^= h >> 30;
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 27;
    h *= 0x94d049bb133111ebull;
    h ^= h >> 31;
    keys[c] = h;
    vals[c] = c;
}

__global__ __launch_bounds__(256) void k_segStart(const uint64_t *keys, int n, int *start)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    start[i] = (i == 0 || keys[i] != keys[i - 1]) ? i : 0;
}

// flag[site] = 1 for the first site of every run; *bad = 1 when a column
// differs from its run's first column
__global__ __launch_bounds__(256) void k_verify(const int8_t *seq, int nTax, int nChar, const int *sites,
                                                const int *start, int *flag, int *bad)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nChar) return;
    const int s0 = start[i];
    const int site = sites[i], head = sites[s0];
    flag[site] = (s0 == i) ? 1 : 0;
    if (s0 != i)
        for (int t = 0; t < nTax; t++)
            if (seq[(size_t)t * nChar + site] != seq[(size_t)t * nChar + head]) {
                atomicOr(bad, 1);
                break;
            }
}

__global__ __launch_bounds__(256) void k_assign(int nChar, const int *sites, const int *start, const int *pidOfSite,
                                                int *sitePat, int *counts, int *headOfPat)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nChar) return;
    const int site = sites[i], head = sites[start[i]];
    const int pid = pidOfSite[head];
    sitePat[site] = pid;
    atomicAdd(&counts[pid], 1);
    if (site == head) headOfPat[pid] = site;
}

__global__ __launch_bounds__(256) void k_gather(const int8_t *seq, int nTax, int nChar, int nPat, const int *headOfPat,
                                                int8_t *pats)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nPat * nTax) return;
    const int t = i / nPat, k = i - t * nPat;
    pats[i] = seq[(size_t)t * nChar + headOfPat[k]];
}

struct Max {
    __device__ __forceinline__ int operator()(int a, int b) const { return a > b ? a : b; }
};

// Device scratch of one compression, freed by its destructor after the
// stream has drained.
struct CompressBufs {
    char *base = nullptr;
    ~CompressBufs()
    {
        if (base) {
            (void)hipStreamSynchronize(stream());
            (void)hipFree(base);
        }
    }
};

hipError_t compressPatterns(const int8_t *dSeq, int nTax, int nChar, std::vector<int8_t> &pats,
                            std::vector<int> &counts, std::vector<int> &sitePat, int *nPatOut, int *collisions)
{
    hipStream_t s = stream();
    *collisions = 0;
    if (nChar <= 0 || nTax <= 0) return hipErrorInvalidValue;
    const size_t n = (size_t)nChar;
    // temporary storage of the sort and the scans
    size_t sortBytes = 0, scanMaxBytes = 0, scanSumBytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                                      (int *)nullptr, (int *)nullptr, nChar, 0, 64, s);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::InclusiveScan(nullptr, scanMaxBytes, (int *)nullptr, (int *)nullptr, Max(), nChar, s);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, scanSumBytes, (int *)nullptr, (int *)nullptr, nChar, s);
    if (e != hipSuccess) return e;
    const size_t tmpBytes = std::max(sortBytes, std::max(scanMaxBytes, scanSumBytes));
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    // keys 2x, sites 2x, start, flag, pid, sitePat, counts, headOfPat, bad, pats, temp
    const size_t offK0 = 0, offK1 = offK0 + up(n * 8), offV0 = offK1 + up(n * 8), offV1 = offV0 + up(n * 4),
                 offSt = offV1 + up(n * 4), offFl = offSt + up(n * 4), offPid = offFl + up(n * 4),
                 offSP = offPid + up(n * 4), offCnt = offSP + up(n * 4), offHd = offCnt + up(n * 4),
                 offBad = offHd + up(n * 4), offPat = offBad + 256, offTmp = offPat + up(n * nTax),
                 total = offTmp + up(tmpBytes);
    CompressBufs B;
    if ((e = hipMalloc(&B.base, total)) != hipSuccess) return e;
    uint64_t *k0 = (uint64_t *)(B.base + offK0), *k1 = (uint64_t *)(B.base + offK1);
    int *v0 = (int *)(B.base + offV0), *v1 = (int *)(B.base + offV1), *st = (int *)(B.base + offSt),
        *fl = (int *)(B.base + offFl), *pid = (int *)(B.base + offPid), *sp = (int *)(B.base + offSP),
        *cnt = (int *)(B.base + offCnt), *hd = (int *)(B.base + offHd), *bad = (int *)(B.base + offBad);
    int8_t *pt = (int8_t *)(B.base + offPat);
    void *tmp = B.base + offTmp;
    const dim3 grid((nChar + 255) / 256), blk(256);
    for (int attempt = 0; attempt < 4; attempt++) {
        const uint64_t seed = 0x9E3779B97F4A7C15ull * (uint64_t)(attempt + 1);
        hipLaunchKernelGGL(k_colHash, grid, blk, 0, s, dSeq, nTax, nChar, seed, k0, v0);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        size_t tb = tmpBytes;
        if ((e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, nChar, 0, 64, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_segStart, grid, blk, 0, s, k1, nChar, fl);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        tb = tmpBytes;
        if ((e = hipcub::DeviceScan::InclusiveScan(tmp, tb, fl, st, Max(), nChar, s)) != hipSuccess) return e;
        if ((e = launchFillU32(reinterpret_cast<unsigned *>(bad), 0u, 1, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_verify, grid, blk, 0, s, dSeq, nTax, nChar, v1, st, fl, bad);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        int hBad = 0;
        if ((e = hipMemcpyAsync(&hBad, bad, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return e;
        if (hBad) {
            (*collisions)++;
            continue;
        }
        tb = tmpBytes;
        if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, fl, pid, nChar, s)) != hipSuccess) return e;
        int lastPid = 0, lastFlag = 0;
        if ((e = hipMemcpyAsync(&lastPid, pid + nChar - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipMemcpyAsync(&lastFlag, fl + nChar - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return e;
        const int nPat = lastPid + lastFlag;
        if ((e = launchFillU32(reinterpret_cast<unsigned *>(cnt), 0u, nPat, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_assign, grid, blk, 0, s, nChar, v1, st, pid, sp, cnt, hd);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_gather, dim3(((size_t)nPat * nTax + 255) / 256), blk, 0, s, dSeq, nTax, nChar, nPat, hd,
                           pt);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        pats.resize((size_t)nPat * nTax);
        counts.resize(nPat);
        sitePat.resize(nChar);
        if ((e = hipMemcpyAsync(pats.data(), pt, pats.size(), hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipMemcpyAsync(counts.data(), cnt, (size_t)nPat * 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipMemcpyAsync(sitePat.data(), sp, n * 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return e;
        *nPatOut = nPat;
        return hipSuccess;
    }
    return hipErrorUnknown;   // four hash seeds collided: the caller reports it
}

}  // namespace p4g
