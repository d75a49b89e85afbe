// p4g_kernels.hip -- CDNA4 (gfx950) kernels of the likelihood engine.
//
//  k_eig    one 64-lane workgroup per (comp, rMatrix) eigensystem: builds Q
//           (Pf/util.c:44-115), symmetrises it with pi^(1/2) and runs a
//           parallel cyclic Jacobi sweep (round-robin pairing, all disjoint
//           rotations of a round applied at once) in FP64.  Replaces the
//           EISPACK route of Pf/eig.c:63-161.
//  k_pmat   one thread per (branch, category, i, j) entry:
//           P = I + V diag(expm1(lam t)) Vinv in FP64 (Pf/eig.c:163-223
//           computes V diag(exp(lam t)) Vinv; the expm1 form keeps short
//           branches accurate), stored transposed in FP64 and FP32.
//  k_prune  the whole postorder in ONE launch: every workgroup owns a tile
//           of site patterns and walks the node program for that tile, so
//           no inter-workgroup dependency (patterns are independent) and no
//           per-level launches.  Lane = (pattern, gamma category); partials
//           are [slot][pattern][cat][state], so a wavefront reads and writes
//           contiguous runs (1 KiB per instruction for FP32 DNA).  The root
//           node is fused with the weighted site-log-likelihood sum
//           (Pf/p4_tree.c:1029-1378) and a last-workgroup compensated FP64
//           reduction.  FP32 mode rescales a (node, pattern) vector by an
//           exact power of two when its max over categories and states
//           drops below 2^-24, accumulating integer exponents per node.
#include "p4g_internal.h"

#include <cmath>

namespace p4g {

// ------------------------------------------------------------------- eigen
__device__ __forceinline__ int rrPlayer(int pos, int r, int n)
{
    return pos == 0 ? 0 : ((pos - 1 + r) % (n - 1)) + 1;
}

// 1/x and 1/sqrt(x) from the hardware estimates plus two Newton steps
__device__ __forceinline__ double rcpN(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

__device__ __forceinline__ double rsqN(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * x * y, y, 1.5);
    y = y * fma(-0.5 * x * y, y, 1.5);
    return y;
}

__device__ __forceinline__ double waveSum(double v)
{
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__global__ __launch_bounds__(64) void k_eig(EigArgs a)
{
    extern __shared__ double sm[];
    const int S = a.dim;
    const int n = (S + 1) & ~1;
    const int h = n / 2;
    double *A = sm;             // [n][n]
    double *U = A + n * n;      // [n][n]
    double *sq = U + n * n;     // [n]
    double *cs = sq + n;        // [h][2]
    int *pq = (int *)(cs + n);  // [h][2]
    const int slot = a.slotList[blockIdx.x];
    const double *pi = a.in + (size_t)slot * (S + S * S);
    const double *R = pi + S;
    const int lane = threadIdx.x;

    // Q = R . diag(pi), diagonal = -row sum, scaled to mean rate 1
    // (Pf/util.c:44-115), built in parallel in U's storage, then
    // symmetrised into A = diag(pi)^(1/2) Q diag(pi)^(-1/2).
    double *Q = U;
    double *isq = cs;          // [n], free until the sweeps start
    double *rowS = A;          // row sums, in A's first row before A is built
    for (int i = lane; i < n; i += 64) {
        const double v = (i < S) ? sqrt(pi[i]) : 1.0;
        sq[i] = v;
        isq[i] = 1.0 / v;
    }
    for (int i = lane; i < n * n; i += 64) {
        const int r = i / n, c = i % n;
        Q[i] = (r < S && c < S && r != c) ? R[r * S + c] * pi[c] : 0.0;
    }
    __syncthreads();
    for (int r = lane; r < S; r += 64) {
        double sum = 0.0;
        for (int c = 0; c < S; c++)
            if (r != c) sum += Q[r * n + c];
        Q[r * n + r] = -sum;
        rowS[r] = pi[r] * sum;
    }
    __syncthreads();
    double sumODE = 0.0;     // sum_r pi_r sum_{c != r} Q_rc, the same on every lane
    for (int r = 0; r < S; r++) sumODE += rowS[r];
    const double inv = 1.0 / sumODE;
    __syncthreads();
    for (int i = lane; i < n * n; i += 64) {
        const int r = i / n, c = i % n;
        double v = 0.0;
        if (r < S && c < S)
            v = (r == c) ? Q[i] * inv : 0.5 * inv * (sq[r] * Q[i] * isq[c] + sq[c] * Q[c * n + r] * isq[r]);
        A[i] = v;
    }
    __syncthreads();
    for (int i = lane; i < n * n; i += 64) U[i] = ((i / n) == (i % n)) ? 1.0 : 0.0;
    __syncthreads();

    for (int sweep = 0; sweep < 30; sweep++) {
        double off = 0.0, dg = 0.0;
        for (int idx = lane; idx < S * S; idx += 64) {
            int i = idx / S, j = idx % S;
            double v = A[i * n + j];
            if (i < j) off += v * v;
            if (i == j) dg += v * v;
        }
        off = waveSum(off);
        dg = waveSum(dg);
        if (off == 0.0 || off < 1e-30 * dg) break;  // off-diagonal norm below 1e-15 of the diagonal
        for (int r = 0; r < n - 1; r++) {
            if (lane < h) {
                int p = rrPlayer(lane, r, n), q = rrPlayer(n - 1 - lane, r, n);
                if (p > q) {
                    int t = p;
                    p = q;
                    q = t;
                }
                double c = 1.0, s = 0.0;
                if (q < S) {
                    double apq = A[p * n + q];
                    if (apq != 0.0) {
                        // the rotation angle needs no correctly rounded divide:
                        // reciprocal and reciprocal square root estimates,
                        // each refined by two Newton steps (full precision),
                        // keep the dependent FP64 chain short
                        const double theta = (A[q * n + q] - A[p * n + p]) * rcpN(2.0 * apq);
                        double t;
                        if (fabs(theta) > 1e150) {
                            t = 0.5 * rcpN(theta);
                        } else {
                            const double q2 = theta * theta + 1.0;
                            const double root = q2 * rsqN(q2);
                            t = (theta >= 0 ? 1.0 : -1.0) * rcpN(fabs(theta) + root);
                        }
                        c = rsqN(t * t + 1.0);
                        s = t * c;
                    }
                }
                pq[2 * lane] = p;
                pq[2 * lane + 1] = q;
                cs[2 * lane] = c;
                cs[2 * lane + 1] = s;
            }
            __syncthreads();
            // rows: A <- J^T A
            for (int it = lane; it < h * S; it += 64) {
                int pr = it / S, k = it % S;
                int p = pq[2 * pr], q = pq[2 * pr + 1];
                double c = cs[2 * pr], s = cs[2 * pr + 1];
                if (s == 0.0 || q >= S) continue;
                double apk = A[p * n + k], aqk = A[q * n + k];
                A[p * n + k] = c * apk - s * aqk;
                A[q * n + k] = s * apk + c * aqk;
            }
            __syncthreads();
            // columns: A <- A J, U <- U J
            for (int it = lane; it < h * S; it += 64) {
                int pr = it / S, k = it % S;
                int p = pq[2 * pr], q = pq[2 * pr + 1];
                double c = cs[2 * pr], s = cs[2 * pr + 1];
                if (s == 0.0 || q >= S) continue;
                double akp = A[k * n + p], akq = A[k * n + q];
                A[k * n + p] = c * akp - s * akq;
                A[k * n + q] = s * akp + c * akq;
                double ukp = U[k * n + p], ukq = U[k * n + q];
                U[k * n + p] = c * ukp - s * ukq;
                U[k * n + q] = s * ukp + c * ukq;
            }
            __syncthreads();
        }
    }
    double *V = a.V + (size_t)slot * S * S;
    double *Vi = a.Vi + (size_t)slot * S * S;
    double *lam = a.lam + (size_t)slot * S;
    for (int idx = lane; idx < S * S; idx += 64) {
        int i = idx / S, k = idx % S;
        double u = U[i * n + k];
        V[i * S + k] = u / sq[i];
        Vi[k * S + i] = u * sq[i];
    }
    for (int k = lane; k < S; k += 64) lam[k] = A[k * n + k];
}

hipError_t launchEig(const EigArgs &a, hipStream_t s)
{
    const int n = (a.dim + 1) & ~1;
    size_t bytes = sizeof(double) * (2 * n * n + 2 * n) + sizeof(int) * n;
    if (bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_eig,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_eig, dim3(a.nSlots), dim3(64), bytes, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------- P matrices
__device__ __forceinline__ void pmatBody(const PmatArgs &a, long gid)
{
    const int S = a.dim, K = a.K;
    const int j = gid % S;
    const int i = (gid / S) % S;
    const int k = (gid / (S * S)) % K;
    const int it = gid / ((long)S * S * K);
    const int node = a.nodeList[it];
    const int es = a.eigSlot[it];
    const double t = a.t[(size_t)it * K + k];
    const double *V = a.V + (size_t)es * S * S;
    const double *Vi = a.Vi + (size_t)es * S * S;
    const double *lam = a.lam + (size_t)es * S;
    double acc = 0.0;
    for (int m = 0; m < S; m++) acc += V[i * S + m] * Vi[m * S + j] * expm1(lam[m] * t);
    const double pij = (i == j ? 1.0 : 0.0) + acc;
    const size_t o = (((size_t)node * K + k) * S + j) * S + i;
    a.PT64[o] = pij;
    a.PT32[o] = (float)pij;
}

__global__ __launch_bounds__(256) void k_pmat(PmatArgs a)
{
    const long total = (long)a.nItems * a.K * a.dim * a.dim;
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    pmatBody(a, gid);
}

// The eigen route's P matrices of all items of a batched call in one
// launch: item i owns entries [first[i], first[i + 1]), one thread each;
// the item map is read once into LDS.
__global__ __launch_bounds__(256) void k_pmatMulti(const PmatArgs *args, const int *first, int nItems, int total)
{
    extern __shared__ int sFirst[];
    for (int j = threadIdx.x; j <= nItems; j += blockDim.x) sFirst[j] = first[j];
    __syncthreads();
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    int i = 0;
    while (i + 1 < nItems && sFirst[i + 1] <= gid) i++;
    const PmatArgs a = args[i];
    const int local = gid - sFirst[i];
    if ((long)local >= (long)a.nItems * a.K * a.dim * a.dim) return;
    pmatBody(a, local);
}

hipError_t launchPmatMulti(const PmatArgs *args, const int *first, int nItems, int total, hipStream_t s)
{
    if (nItems == 0 || total == 0) return hipSuccess;
    const size_t lds = sizeof(int) * ((size_t)nItems + 1);
    if (lds > 32 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pmatMulti, dim3((total + 255) / 256), dim3(256), lds, s, args, first, nItems, total);
    return hipGetLastError();
}

hipError_t launchPmat(const PmatArgs &a, hipStream_t s)
{
    const long total = (long)a.nItems * a.K * a.dim * a.dim;
    if (total == 0) return hipSuccess;
    const int blocks = (int)((total + 255) / 256);
    hipLaunchKernelGGL(k_pmat, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------- P by direct exponentiation
// 4-state parts skip the eigensystem: one thread per (branch, category)
// builds Q from the snapshot of (pi, R) (the normalisation of Pf/util.c:
// 44-115, as in k_eig) and evaluates E = expm1(Q t) in registers:
// B = Q t / 2^s with ||B||_inf <= 1/4, the Taylor series of expm1(B) to
// degree 12 by Horner (first omitted term below 2.5e-18 ||B||), then s
// doublings E <- 2E + E E, the expm1 form of squaring (I + E)^2 = I + 2E + E^2,
// so that short branches keep their small off-diagonal entries to full
// relative precision.  P = I + E is stored transposed, as k_pmat stores it.
constexpr int PEXP_DEGREE = 12;

template <int S>
__device__ __forceinline__ void matMul(const double (&X)[S][S], const double (&Y)[S][S], double (&Z)[S][S])
{
#pragma unroll
    for (int i = 0; i < S; i++)
#pragma unroll
        for (int j = 0; j < S; j++) {
            double acc = X[i][0] * Y[0][j];
#pragma unroll
            for (int m = 1; m < S; m++) acc = fma(X[i][m], Y[m][j], acc);
            Z[i][j] = acc;
        }
}

template <int S>
__device__ __forceinline__ void pexpBody(const PmatArgs &a, const double *in, int it, int k)
{
    const int node = a.nodeList[it];
    const double t = a.t[(size_t)it * a.K + k];
    const double *pi = in + (size_t)a.eigSlot[it] * (S + S * S);
    const double *R = pi + S;
    double B[S][S];
    double sumODE = 0.0;
#pragma unroll
    for (int r = 0; r < S; r++) {
        double row = 0.0;
#pragma unroll
        for (int c = 0; c < S; c++) {
            const double q = (r == c) ? 0.0 : R[r * S + c] * pi[c];
            B[r][c] = q;
            row += q;
        }
        B[r][r] = -row;
        sumODE += pi[r] * row;
    }
    // B = Q t / sumODE, then its infinity norm picks the number of doublings
    const double f = t / sumODE;
    double nrm = 0.0;
#pragma unroll
    for (int r = 0; r < S; r++) {
        double row = 0.0;
#pragma unroll
        for (int c = 0; c < S; c++) {
            B[r][c] *= f;
            row += fabs(B[r][c]);
        }
        nrm = fmax(nrm, row);
    }
    int sq = 0;
    if (nrm > 0.25) {
        (void)frexp(nrm * 4.0, &sq);  // nrm * 4 <= 2^sq
    }
#pragma unroll
    for (int r = 0; r < S; r++)
#pragma unroll
        for (int c = 0; c < S; c++) B[r][c] = ldexp(B[r][c], -sq);
    // Z = I + B/2 + B^2/3! + ... + B^(d-1)/d!  (Horner), E = B Z
    double Z[S][S], W[S][S];
#pragma unroll
    for (int r = 0; r < S; r++)
#pragma unroll
        for (int c = 0; c < S; c++) Z[r][c] = (r == c ? 1.0 : 0.0) + B[r][c] * (1.0 / PEXP_DEGREE);
#pragma unroll
    for (int j = PEXP_DEGREE - 1; j >= 2; j--) {
        matMul<S>(B, Z, W);
        const double inv = 1.0 / j;
#pragma unroll
        for (int r = 0; r < S; r++)
#pragma unroll
            for (int c = 0; c < S; c++) Z[r][c] = (r == c ? 1.0 : 0.0) + W[r][c] * inv;
    }
    matMul<S>(B, Z, W);  // W = E
    for (int i = 0; i < sq; i++) {
        matMul<S>(W, W, Z);
#pragma unroll
        for (int r = 0; r < S; r++)
#pragma unroll
            for (int c = 0; c < S; c++) W[r][c] = fma(2.0, W[r][c], Z[r][c]);
    }
    const size_t o = ((size_t)node * a.K + k) * S * S;
#pragma unroll
    for (int j = 0; j < S; j++)
#pragma unroll
        for (int i = 0; i < S; i++) {
            const double pij = (i == j ? 1.0 : 0.0) + W[i][j];
            a.PT64[o + j * S + i] = pij;
            a.PT32[o + j * S + i] = (float)pij;
        }
}

// The same P with 16 lanes per matrix, lane e holding entry (e / 4, e % 4):
// a 4-state matrix's series is a chain of some 15 dependent products, and one
// lane per matrix ran it at the latency of 64 serial FP64 FMAs per product.
// Every lane forms Q, the scale and the doublings count for the whole matrix
// exactly as pexpBody does (so all 16 agree bit for bit), keeps its row of
// B, and fetches the columns (and, when doubling, the row) of the other
// factor from the lanes of its matrix; each entry sums over m in matMul's
// order, so P is pexpBody's bit for bit.  base: the matrix's first lane.
__device__ __forceinline__ void pexpLanes4(const PmatArgs &a, const double *in, int it, int k, int e, int base)
{
    constexpr int S = 4;
    const int r = e >> 2, c = e & 3;
    const int node = a.nodeList[it];
    const double t = a.t[(size_t)it * a.K + k];
    const double *pi = in + (size_t)a.eigSlot[it] * (S + S * S);
    const double *R = pi + S;
    double B[S][S];
    double sumODE = 0.0;
#pragma unroll
    for (int rr = 0; rr < S; rr++) {
        double row = 0.0;
#pragma unroll
        for (int cc = 0; cc < S; cc++) {
            const double q = (rr == cc) ? 0.0 : R[rr * S + cc] * pi[cc];
            B[rr][cc] = q;
            row += q;
        }
        B[rr][rr] = -row;
        sumODE += pi[rr] * row;
    }
    const double f = t / sumODE;
    double nrm = 0.0;
#pragma unroll
    for (int rr = 0; rr < S; rr++) {
        double row = 0.0;
#pragma unroll
        for (int cc = 0; cc < S; cc++) {
            B[rr][cc] *= f;
            row += fabs(B[rr][cc]);
        }
        nrm = fmax(nrm, row);
    }
    int sq = 0;
    if (nrm > 0.25) {
        (void)frexp(nrm * 4.0, &sq);
    }
    double Br[S];   // this lane's row of B
#pragma unroll
    for (int m = 0; m < S; m++) Br[m] = ldexp(B[r][m], -sq);
    const double diag = r == c ? 1.0 : 0.0;
    double z = diag + Br[c] * (1.0 / PEXP_DEGREE);
    auto col = [&](double v, double (&o)[S]) {
#pragma unroll
        for (int m = 0; m < S; m++) o[m] = __shfl(v, base + m * S + c, 64);
    };
#pragma unroll
    for (int j = PEXP_DEGREE - 1; j >= 2; j--) {
        double zc[S];
        col(z, zc);
        double acc = Br[0] * zc[0];
#pragma unroll
        for (int m = 1; m < S; m++) acc = fma(Br[m], zc[m], acc);
        z = diag + acc * (1.0 / j);
    }
    double w;
    {
        double zc[S];
        col(z, zc);
        double acc = Br[0] * zc[0];
#pragma unroll
        for (int m = 1; m < S; m++) acc = fma(Br[m], zc[m], acc);
        w = acc;   // E
    }
    for (int i = 0; i < sq; i++) {   // every lane of the matrix has the same sq
        double wr[S], wc[S];
#pragma unroll
        for (int m = 0; m < S; m++) wr[m] = __shfl(w, base + r * S + m, 64);
        col(w, wc);
        double acc = wr[0] * wc[0];
#pragma unroll
        for (int m = 1; m < S; m++) acc = fma(wr[m], wc[m], acc);
        w = fma(2.0, w, acc);
    }
    const size_t o = ((size_t)node * a.K + k) * S * S + (size_t)c * S + r;
    const double pij = diag + w;
    a.PT64[o] = pij;
    a.PT32[o] = (float)pij;
}

template <int S>
__global__ __launch_bounds__(64) void k_pexp(PmatArgs a, const double *in)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= a.nItems * a.K) return;
    pexpBody<S>(a, in, gid / a.K, gid % a.K);
}

// 16 lanes per (branch, category): pexpLanes4.  A matrix's lanes share
// one 16-lane row of the wave and leave together.
__global__ __launch_bounds__(64) void k_pexpLanes4(PmatArgs a, const double *in)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int mi = gid >> 4;
    if (mi >= a.nItems * a.K) return;
    pexpLanes4(a, in, mi / a.K, mi % a.K, gid & 15, (int)(threadIdx.x & 48));
}

// All items of a batched call in one launch: item i owns threads
// [first[i], first[i + 1]), one per (branch, category).
template <int S>
__global__ __launch_bounds__(64) void k_pexpMulti(const PmatArgs *args, const double *const *ins, const int *first,
                                                  int nItems, int total)
{
    // the item map (pinned host memory) is read once, in parallel, into LDS
    extern __shared__ int sFirst[];
    for (int j = threadIdx.x; j <= nItems; j += blockDim.x) sFirst[j] = first[j];
    __syncthreads();
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (S == 4) {
        // 16 lanes per matrix (pexpLanes4)
        const int mi = gid >> 4;
        if (mi >= total) return;
        int i = 0;
        while (i + 1 < nItems && sFirst[i + 1] <= mi) i++;
        const PmatArgs a = args[i];
        const int local = mi - sFirst[i];
        if (local >= a.nItems * a.K) return;
        pexpLanes4(a, ins[i], local / a.K, local % a.K, gid & 15, (int)(threadIdx.x & 48));
        return;
    }
    if (gid >= total) return;
    int i = 0;
    while (i + 1 < nItems && sFirst[i + 1] <= gid) i++;
    const PmatArgs a = args[i];
    const int local = gid - sFirst[i];
    if (local >= a.nItems * a.K) return;
    pexpBody<S>(a, ins[i], local / a.K, local % a.K);
}

// The same series and doublings for larger state counts (20: proteins),
// one wavefront per (branch, category): B, Z and W live in LDS and lane l
// owns entries l, l + 64, ...; every product sums over m in the same order
// as matMul.
template <int S>
__device__ __forceinline__ void pexpWaveBody(const PmatArgs &a, const double *in, int it, int k)
{
    constexpr int NE = S * S;
    constexpr int PER = (NE + 63) / 64;
    __shared__ double B[NE], Z[NE], W[NE], T[NE];
    __shared__ double red[S];
    __shared__ double sumODE;
    const int lane = threadIdx.x;
    const int node = a.nodeList[it];
    const double t = a.t[(size_t)it * a.K + k];
    const double *pi = in + (size_t)a.eigSlot[it] * (S + NE);
    const double *R = pi + S;
    for (int e = lane; e < NE; e += 64) {
        const int r = e / S, c = e - (e / S) * S;
        B[e] = (r == c) ? 0.0 : R[e] * pi[c];
    }
    __syncthreads();
    if (lane < S) {
        double row = 0.0;
        for (int c = 0; c < S; c++) row += B[lane * S + c];
        B[lane * S + lane] = -row;
        red[lane] = pi[lane] * row;
    }
    __syncthreads();
    if (lane == 0) {
        double sOde = 0.0;
        for (int r = 0; r < S; r++) sOde += red[r];
        sumODE = sOde;
    }
    __syncthreads();
    const double f = t / sumODE;
    for (int e = lane; e < NE; e += 64) B[e] *= f;
    __syncthreads();
    if (lane < S) {
        double row = 0.0;
        for (int c = 0; c < S; c++) row += fabs(B[lane * S + c]);
        red[lane] = row;
    }
    __syncthreads();
    double nrm = 0.0;
    for (int r = 0; r < S; r++) nrm = fmax(nrm, red[r]);
    int sq = 0;
    if (nrm > 0.25) {
        (void)frexp(nrm * 4.0, &sq);
    }
    for (int e = lane; e < NE; e += 64) {
        const int r = e / S, c = e - (e / S) * S;
        const double b = ldexp(B[e], -sq);
        B[e] = b;
        Z[e] = (r == c ? 1.0 : 0.0) + b * (1.0 / PEXP_DEGREE);
    }
    __syncthreads();
    // acc[i] = (X Y)[entry lane + 64 i], on the FP64 matrix cores: the
    // product padded to 32 x 32 is four 16 x 16 blocks, each five
    // v_mfma_f64_16x16x4f64 steps over k (lane l: A[l % 16][l / 16],
    // B[l / 16][l % 16]; result register r holds row l / 16 + 4 r of column
    // l % 16), written to T and read back in the lane-owns-entries layout.
    static_assert(S > 16 && S <= 32, "two 16-row blocks");
    typedef double d4v __attribute__((ext_vector_type(4)));
    const int q = lane >> 4, pl = lane & 15;
    auto product = [&](const double *X, const double *Y, double (&acc)[PER]) {
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++) {
                const int i = rb * 16 + pl, j = cb * 16 + pl;
                d4v d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < (S + 3) / 4; ks++) {
                    const int kk = ks * 4 + q;
                    const double av = (i < S && kk < S) ? X[i * S + kk] : 0.0;
                    const double bv = (j < S && kk < S) ? Y[kk * S + j] : 0.0;
                    d = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, d, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = rb * 16 + q + 4 * r;   // the FP64 result layout interleaves rows
                    if (row < S && j < S) T[row * S + j] = d[r];
                }
            }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int e = lane + 64 * i;
            acc[i] = e < NE ? T[e] : 0.0;
        }
    };
    double acc[PER];
    for (int j = PEXP_DEGREE - 1; j >= 2; j--) {
        product(B, Z, acc);
        __syncthreads();
        const double inv = 1.0 / j;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int e = lane + 64 * i;
            if (e < NE) {
                const int r = e / S, c = e - (e / S) * S;
                Z[e] = (r == c ? 1.0 : 0.0) + acc[i] * inv;
            }
        }
        __syncthreads();
    }
    product(B, Z, acc);  // E
#pragma unroll
    for (int i = 0; i < PER; i++)
        if (lane + 64 * i < NE) W[lane + 64 * i] = acc[i];
    __syncthreads();
    for (int d = 0; d < sq; d++) {
        product(W, W, acc);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int e = lane + 64 * i;
            if (e < NE) W[e] = fma(2.0, W[e], acc[i]);
        }
        __syncthreads();
    }
    const size_t o = ((size_t)node * a.K + k) * NE;
    for (int e = lane; e < NE; e += 64) {
        const int i = e / S, j = e - (e / S) * S;
        const double pij = (i == j ? 1.0 : 0.0) + W[e];
        a.PT64[o + j * S + i] = pij;
        a.PT32[o + j * S + i] = (float)pij;
    }
}

template <int S>
__global__ __launch_bounds__(64) void k_pexpWave(PmatArgs a, const double *in)
{
    const int it = blockIdx.x / a.K, k = blockIdx.x - (blockIdx.x / a.K) * a.K;
    if (it >= a.nItems) return;   // the whole workgroup
    pexpWaveBody<S>(a, in, it, k);
}

// All (tree, part) items of a batched call in one launch: item i owns
// workgroups [first[i], first[i + 1]), one per (branch, category).
template <int S>
__global__ __launch_bounds__(64) void k_pexpWaveMulti(const PmatArgs *args, const double *const *ins,
                                                      const int *first, int nItems)
{
    extern __shared__ int sFirst[];
    for (int j = threadIdx.x; j <= nItems; j += blockDim.x) sFirst[j] = first[j];
    __syncthreads();
    const int b = blockIdx.x;
    int i = 0;
    while (i + 1 < nItems && sFirst[i + 1] <= b) i++;
    const PmatArgs a = args[i];
    const int local = b - sFirst[i];
    const int it = local / a.K, k = local - (local / a.K) * a.K;
    if (it >= a.nItems) return;   // the whole workgroup
    pexpWaveBody<S>(a, ins[i], it, k);
}

hipError_t launchPexpMulti(int dim, const PmatArgs *args, const double *const *ins, const int *first, int nItems,
                           int total, hipStream_t s)
{
    if (nItems == 0 || total == 0) return hipSuccess;
    const size_t lds = sizeof(int) * ((size_t)nItems + 1);   // the item map
    if (lds > 32 * 1024) return hipErrorInvalidValue;
    if (dim == 20)
        hipLaunchKernelGGL(k_pexpWaveMulti<20>, dim3(total), dim3(64), lds, s, args, ins, first, nItems);
    else if (dim == 4)
        hipLaunchKernelGGL(k_pexpMulti<4>, dim3(((size_t)total * 16 + 63) / 64), dim3(64), lds, s, args, ins, first,
                           nItems, total);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

static int g_pexpDirect = 1;

void setPmatRoute(int direct) { g_pexpDirect = direct ? 1 : 0; }

bool pmatDirect(int dim) { return g_pexpDirect && (dim == 4 || dim == 20); }

hipError_t launchPexp(const PmatArgs &a, const double *eigIn, hipStream_t s)
{
    const int total = a.nItems * a.K;
    if (total == 0) return hipSuccess;
    if (a.dim == 4) {
        hipLaunchKernelGGL(k_pexpLanes4, dim3(((size_t)total * 16 + 63) / 64), dim3(64), 0, s, a, eigIn);
    } else if (a.dim == 20) {
        hipLaunchKernelGGL(k_pexpWave<20>, dim3(total), dim3(64), 0, s, a, eigIn);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ----------------------------------------------------------------- pruning
constexpr int PRUNE_THREADS = 256;
constexpr float SCALE_THRESH = 5.9604645e-08f;  // 2^-24

int pruneBlockPatterns(int K)
{
    int Kp = 1;
    while (Kp < K) Kp <<= 1;
    return PRUNE_THREADS / Kp;
}


template <typename T>
__device__ __forceinline__ T groupMax(T v, int Kp)
{
    for (int off = 1; off < Kp; off <<= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// Contribution of a leaf child (Pf/p4_node.c:660-725, 753-809): column
// `code` of P, all ones for gap / '?' / an all-state equate, or the sum of
// the allowed columns for a partial ambiguity.  Pk is P^T: Pk[j*S + s].
template <int SC, typename T>
__device__ __forceinline__ void leafVec(const T *Pk, unsigned code, int S, const uint64_t *amb, T *v)
{
    constexpr int SM = SC ? SC : MAX_DIM;
    constexpr int UR = SC ? SM : 1;  // the generic path keeps its loops rolled
    if (code < (unsigned)S) {
#pragma unroll UR
        for (int s = 0; s < SM; s++)
            if (s < S) v[s] = Pk[code * S + s];
    } else if (code == TIP_ALL) {
#pragma unroll UR
        for (int s = 0; s < SM; s++) v[s] = T(1);
    } else {
        const uint64_t m = amb[code - S];
#pragma unroll UR
        for (int s = 0; s < SM; s++) v[s] = T(0);
        for (int j = 0; j < S; j++)
            if ((m >> j) & 1ull) {
#pragma unroll UR
                for (int s = 0; s < SM; s++)
                    if (s < S) v[s] += Pk[j * S + s];
            }
    }
}

// Contribution of an internal child (Pf/p4_node.c:726-744, 810-834):
// v[s] = sum_j P[s][j] x[j], summed over j in increasing order.
template <int SC, typename T>
__device__ __forceinline__ void internalVec(const T *Pk, const T *x, int S, T *v)
{
    constexpr int SM = SC ? SC : MAX_DIM;
    constexpr int UR = SC ? SM : 1;  // the generic path keeps its loops rolled
#pragma unroll UR
    for (int s = 0; s < SM; s++) v[s] = T(0);
#pragma unroll UR
    for (int j = 0; j < SM; j++) {
        if (j < S) {
#pragma unroll UR
            for (int s = 0; s < SM; s++)
                if (s < S) v[s] += x[j] * Pk[j * S + s];
        }
    }
}

// FP32 rescaling: when the largest entry of the (pattern, all categories)
// vector is below 2^-24, multiply by the power of two that brings it to
// [0.5, 1) and add that power to the node's exponent.  Exact.
template <int SC, typename T>
__device__ __forceinline__ void rescale(T *acc, int &e, int S, int Kp)
{
    constexpr int SM = SC ? SC : MAX_DIM;
    constexpr int UR = SC ? SM : 1;  // the generic path keeps its loops rolled
    T m = T(0);
#pragma unroll UR
    for (int s = 0; s < SM; s++)
        if (s < S) m = fmax(m, acc[s]);
    m = groupMax(m, Kp);
    if (m < T(SCALE_THRESH) && m > T(0)) {
        int ex;
        frexp(m, &ex);
        const int E = -ex;
#pragma unroll UR
        for (int s = 0; s < SM; s++)
            if (s < S) acc[s] = ldexp(acc[s], E);
        e += E;
    }
}

// Site likelihood of pattern p from the root's category-summed
// sum_s pi_s cl_s (Lk) and the pattern's scale exponent e, as
// Pf/p4_tree.c:1029-1197: (1 - pInvar)/K * Lk plus the invariant-site term,
// or Lk/K without pInvar; adds count * log(like) to contrib.  In FP32 mode
// the partials carry 2^-e, so the variable-site part is unscaled in FP64
// before the invariant part is added; with no invariant part the log is
// taken in the scaled domain.
template <bool SCALE>
__device__ __forceinline__ void siteLogLike(const PruneArgs &a, double Lk, int e, int p, int K,
                                            const double *pi, double &contrib, int &bad)
{
    const size_t nPatPad = (size_t)a.nPatPad;
    const double LN2 = 0.69314718055994530942;
    double like, lnl;
    bool ok;
    if (a.pInvar != 0.0) {
        const double varS = Lk * ((1.0 - a.pInvar) / (double)K);
        double inv = 0.0;
        for (int w = 0; w < a.nWords; w++) {
            uint32_t mw = a.invMask[(size_t)w * nPatPad + p];
            while (mw) {
                const int b = __ffs(mw) - 1;
                mw &= mw - 1;
                inv += pi[w * 32 + b] * a.pInvar;
            }
        }
        if (SCALE && e != 0 && inv == 0.0) {
            ok = varS > 0.0;
            lnl = log(varS) - (double)e * LN2;
            like = ldexp(varS, -e);
        } else {
            like = ((SCALE && e != 0) ? ldexp(varS, -e) : varS) + inv;
            ok = like > 0.0;
            lnl = log(like);
        }
    } else {
        const double likeS = (K > 1) ? Lk / (double)K : Lk;
        ok = likeS > 0.0;
        lnl = log(likeS);
        like = likeS;
        if (SCALE && e != 0) {
            lnl -= (double)e * LN2;
            like = ldexp(likeS, -e);
        }
    }
    if (p < a.nPat) {
        if (!ok) bad = 1;
        contrib = a.counts[p] * lnl;
        if (a.patLikes) a.patLikes[p] = like;
    }
}

// Workgroup reduction of the per-lane weighted site log likelihoods, then
// the last workgroup of the item (atomic ticket) sums the per-workgroup
// partials in a fixed order with Neumaier compensation and resets the
// ticket for the next launch.
__device__ __forceinline__ void finishSiteSum(const PruneArgs &a, double contrib, int bad, int blk,
                                              int nBlk)
{
    const int tid = threadIdx.x;
    for (int off = 32; off >= 1; off >>= 1) {
        contrib += __shfl_xor(contrib, off);
        bad |= __shfl_xor(bad, off);
    }
    __shared__ double wsum[PRUNE_THREADS / 64];
    __shared__ int wbad[PRUNE_THREADS / 64];
    __shared__ int isLast;
    const int wave = tid >> 6, lane = tid & 63;
    if (lane == 0) {
        wsum[wave] = contrib;
        wbad[wave] = bad;
    }
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        int b = 0;
        for (int w = 0; w < PRUNE_THREADS / 64; w++) {
            s += wsum[w];
            b |= wbad[w];
        }
        a.blockSums[blk] = s;
        a.blockBad[blk] = b;
        __threadfence();
        unsigned prev = __hip_atomic_fetch_add(a.counter, 1u, __ATOMIC_ACQ_REL,
                                               __HIP_MEMORY_SCOPE_AGENT);
        isLast = (prev == (unsigned)(nBlk - 1));
    }
    __syncthreads();
    if (isLast) {
        __threadfence();
        double s = 0.0, comp = 0.0;
        int b = 0;
        for (int i = tid; i < nBlk; i += PRUNE_THREADS) {
            const double x = __hip_atomic_load(a.blockSums + i, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
            b |= __hip_atomic_load(a.blockBad + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const double t = s + x;  // Neumaier
            if (fabs(s) >= fabs(x))
                comp += (s - t) + x;
            else
                comp += (x - t) + s;
            s = t;
        }
        double v = s + comp;
        for (int off = 32; off >= 1; off >>= 1) {
            v += __shfl_xor(v, off);
            b |= __shfl_xor(b, off);
        }
        __syncthreads();
        if (lane == 0) {
            wsum[wave] = v;
            wbad[wave] = b;
        }
        __syncthreads();
        if (tid == 0) {
            double tot = 0.0;
            int bb = 0;
            for (int w = 0; w < PRUNE_THREADS / 64; w++) {
                tot += wsum[w];
                bb |= wbad[w];
            }
            a.result[0] = bb ? -1.0e99 : tot;
            a.result[1] = (double)bb;
            *a.counter = 0u;
        }
    }
}


template <int SC, typename T, bool SCALE, bool PLDS, int G>
__device__ __forceinline__ void pruneBody(const PruneArgs &a, int blk, int nBlk, T *sP)
{
    constexpr int SM = SC ? SC : MAX_DIM;
    constexpr int UR = SC ? SM : 1;  // the generic path keeps its loops rolled
    const int S = SC ? SC : a.dim;
    const int K = a.K;
    const int lk = a.log2Kp;
    const int Kp = 1 << lk;
    const int tid = threadIdx.x;
    const int kk0 = tid & (Kp - 1);
    const bool catOn = kk0 < K;
    const int kk = catOn ? kk0 : K - 1;
    const int p = blk * (PRUNE_THREADS >> lk) + (tid >> lk);
    const size_t nPatPad = (size_t)a.nPatPad;
    const size_t SS = (size_t)S * S;
    const T *PTg = (const T *)a.PT;
    T *cl = (T *)a.cl;

    if (PLDS) {
        // P^T of every child branch the program uses, [kid][K][S][S], copied
        // in 16-byte chunks with 8 independent loads in flight per thread
        constexpr int CH = 16 / sizeof(T);           // elements per chunk
        const int per = K * S * S;                   // a multiple of CH for S = 4, 20
        const int cpk = per / CH;                    // chunks per kid
        const int total = a.nKids * cpk;
        for (int c0 = tid; c0 < total; c0 += PRUNE_THREADS * 8) {
            float4 buf[8];
#pragma unroll UR
            for (int u = 0; u < 8; u++) {
                const int c = c0 + u * PRUNE_THREADS;
                if (c < total) {
                    const int kid = c / cpk;
                    const float4 *src = reinterpret_cast<const float4 *>(PTg + (size_t)a.kids[kid].pnode * per);
                    buf[u] = src[c - kid * cpk];
                }
            }
#pragma unroll UR
            for (int u = 0; u < 8; u++) {
                const int c = c0 + u * PRUNE_THREADS;
                if (c < total) reinterpret_cast<float4 *>(sP)[c] = buf[u];
            }
        }
        __syncthreads();
    }
    auto Pof = [&](int kidIdx, int pnode) -> const T * {
        return PLDS ? sP + ((size_t)kidIdx * K + kk) * SS : PTg + ((size_t)pnode * K + kk) * SS;
    };

    double contrib = 0.0;
    int bad = 0;
    for (int g = 0; g < a.nGroups; ++g) {
        const GroupRec gr = a.groups[g];
        if (gr.kind == 0) {
            // ---- up to G independent binary, non-root ops
            T xv[G][2][SM];
            unsigned code[G][2];
            int ek[G];
#pragma unroll UR
            for (int j = 0; j < G; j++) {
                ek[j] = 0;
                if (j < gr.nOps) {
                    const OpRec op = a.ops[gr.opStart + j];
#pragma unroll UR
                    for (int c = 0; c < 2; c++) {
                        const KidRec kd = a.kids[op.kidStart + c];
                        if (kd.kind == 0) {
                            code[j][c] = a.tips[(size_t)kd.idx * nPatPad + p];
                        } else {
                            code[j][c] = 0u;
                            const T *x = cl + (((size_t)kd.idx * nPatPad + p) * K + kk) * S;
#pragma unroll UR
                            for (int s = 0; s < SM; s++)
                                if (s < S) xv[j][c][s] = x[s];
                            if (SCALE) ek[j] += a.scl[(size_t)kd.idx * nPatPad + p];
                        }
                    }
                }
            }
#pragma unroll UR
            for (int j = 0; j < G; j++) {
                if (j < gr.nOps) {
                    const OpRec op = a.ops[gr.opStart + j];
                    T acc[SM], v[SM];
#pragma unroll UR
                    for (int c = 0; c < 2; c++) {
                        const KidRec kd = a.kids[op.kidStart + c];
                        const T *Pk = Pof(op.kidStart + c, kd.pnode);
                        if (kd.kind == 0)
                            leafVec<SC, T>(Pk, code[j][c], S, a.ambMask, v);
                        else
                            internalVec<SC, T>(Pk, xv[j][c], S, v);
                        if (c == 0) {
#pragma unroll UR
                            for (int s = 0; s < SM; s++) acc[s] = v[s];
                        } else {
#pragma unroll UR
                            for (int s = 0; s < SM; s++) acc[s] *= v[s];
                        }
                    }
                    int e = ek[j];
                    if (SCALE) rescale<SC, T>(acc, e, S, Kp);
                    if (catOn) {
                        T *dst = cl + (((size_t)op.dst * nPatPad + p) * K + kk) * S;
#pragma unroll UR
                        for (int s = 0; s < SM; s++)
                            if (s < S) dst[s] = acc[s];
                    }
                    if (SCALE && kk0 == 0) a.scl[(size_t)op.dst * nPatPad + p] = e;
                }
            }
            continue;
        }
        // ---- one op with any number of children, possibly the root
        const OpRec op = a.ops[gr.opStart];
        T acc[SM];
        int e = 0;
        for (int c = 0; c < op.nKids; ++c) {
            const KidRec kd = a.kids[op.kidStart + c];
            const T *Pk = Pof(op.kidStart + c, kd.pnode);
            T v[SM];
            bool skip = false;
            if (kd.kind == 0) {
                const unsigned code = a.tips[(size_t)kd.idx * nPatPad + p];
                leafVec<SC, T>(Pk, code, S, a.ambMask, v);
                skip = (code == TIP_ALL) && c != 0;
            } else {
                internalVec<SC, T>(Pk, cl + (((size_t)kd.idx * nPatPad + p) * K + kk) * S, S, v);
                if (SCALE) e += a.scl[(size_t)kd.idx * nPatPad + p];
            }
            if (c == 0) {
#pragma unroll UR
                for (int s = 0; s < SM; s++) acc[s] = v[s];
            } else if (!skip) {
#pragma unroll UR
                for (int s = 0; s < SM; s++) acc[s] *= v[s];
            }
        }
        if (SCALE) rescale<SC, T>(acc, e, S, Kp);
        if (!(op.flags & OP_ROOT)) {
            if (catOn) {
                T *dst = cl + (((size_t)op.dst * nPatPad + p) * K + kk) * S;
#pragma unroll UR
                for (int s = 0; s < SM; s++)
                    if (s < S) dst[s] = acc[s];
            }
            if (SCALE && kk0 == 0) a.scl[(size_t)op.dst * nPatPad + p] = e;
            continue;
        }
        // ---- root: site likelihood and weighted log (Pf/p4_tree.c:1029-1378)
        const double *pi = a.rootPi;
        double Lk = 0.0;
        if (a.rootLeafRow >= 0) {
            const unsigned code = a.tips[(size_t)a.rootLeafRow * nPatPad + p];
            uint64_t m;
            if (code < (unsigned)S)
                m = 1ull << code;
            else if (code == TIP_ALL)
                m = ~0ull;
            else
                m = a.ambMask[code - S];
#pragma unroll UR
            for (int s = 0; s < SM; s++)
                if (s < S && ((m >> s) & 1ull)) Lk += pi[s] * (double)acc[s];
        } else {
#pragma unroll UR
            for (int s = 0; s < SM; s++)
                if (s < S) Lk += pi[s] * (double)acc[s];
        }
        if (!catOn) Lk = 0.0;
        for (int off = 1; off < Kp; off <<= 1) Lk += __shfl_xor(Lk, off);
        if (kk0 == 0) siteLogLike<SCALE>(a, Lk, e, p, K, pi, contrib, bad);
    }
    finishSiteSum(a, contrib, bad, blk, nBlk);
}

template <int SC, typename T, bool SCALE, bool PLDS, int G>
__global__ __launch_bounds__(PRUNE_THREADS) void k_prune(PruneArgs a)
{
    extern __shared__ __align__(16) unsigned char smemRaw[];
    pruneBody<SC, T, SCALE, PLDS, G>(a, blockIdx.x, gridDim.x, reinterpret_cast<T *>(smemRaw));
}

// Batched variant: several (tree, part) items of the same state count and
// precision in one launch; blockItem[2b] = item, blockItem[2b+1] = the
// block's index within the item.
template <int SC, typename T, bool SCALE, bool PLDS, int G>
__global__ __launch_bounds__(PRUNE_THREADS) void k_pruneBatch(const PruneArgs *args,
                                                              const int *blockItem)
{
    extern __shared__ __align__(16) unsigned char smemRaw[];
    const int item = blockItem[2 * blockIdx.x];
    const int local = blockItem[2 * blockIdx.x + 1];
    const PruneArgs a = args[item];
    const int nBlk = a.nPatPad / (PRUNE_THREADS >> a.log2Kp);
    pruneBody<SC, T, SCALE, PLDS, G>(a, local, nBlk, reinterpret_cast<T *>(smemRaw));
}

// Tuning of the 4-state kernel: ops per binary group (1, 2, 4, 8) and
// whether P rows are staged in LDS.  Other state counts use 1 op per group.
static int g_tuneG[2] = {4, 4};      // [precision]
static int g_tuneLds[2] = {1, 1};
static int g_wave[2] = {1, 1};       // 1: 4-state parts use k_pruneWave
static int g_waveVecP[2] = {0, 0};   // 1: its vector-prefetched P stream
static int g_waveStreams[2] = {0, 0};  // 1: two streams of category waves (scalar P)
static int g_waveLean[2] = {1, 1};   // 1: the lean kernel (LeanOp records, tip P^T in LDS)

// G = -3 selects the lean kernel (the default: the wave kernel's layout with
// precomputed records and tip P^T staged in LDS);
// G = 0 selects the wave-per-category kernel for 4-state parts with its
// vector-prefetched P stream, G = -1 the same kernel with P by scalar
// loads, G = -2 that with the tree split into two streams of waves; G = 1
// or 4 the grouped kernel with that many ops per group (other values: 4).
void setPruneTuning(int precision, int G, int lds)
{
    if (precision < 0 || precision > 1) return;
    if (G == 0 || G == -1 || G == -2 || G == -3) {
        g_wave[precision] = 1;
        g_waveLean[precision] = G == -3 ? 1 : 0;
        g_waveVecP[precision] = G == 0 ? 1 : 0;
        g_waveStreams[precision] = G == -2 ? 1 : 0;
        return;
    }
    g_wave[precision] = 0;
    g_tuneG[precision] = G == 1 ? 1 : 4;   // the grouped kernel is built for 1 and 4
    g_tuneLds[precision] = lds ? 1 : 0;
}

int pruneGroupWidthFor(int dim, int precision) { return dim == 4 ? g_tuneG[precision & 1] : 1; }

template <int SC, typename T, bool SCALE, int G>
static hipError_t launchPruneG(const PruneArgs &a, size_t lds, int nBlocks, hipStream_t s)
{
    // lds <= LDS_P_BUDGET (48 KiB), within the default per-workgroup limit
    if (lds)
        hipLaunchKernelGGL((k_prune<SC, T, SCALE, true, G>), dim3(nBlocks), dim3(PRUNE_THREADS), lds, s, a);
    else
        hipLaunchKernelGGL((k_prune<SC, T, SCALE, false, G>), dim3(nBlocks), dim3(PRUNE_THREADS), 0, s, a);
    return hipGetLastError();
}

template <int SC, typename T, bool SCALE>
static hipError_t launchPruneP(const PruneArgs &a, size_t lds, int nBlocks, hipStream_t s)
{
    if (SC != 4) return launchPruneG<SC, T, SCALE, 1>(a, lds, nBlocks, s);
    switch (pruneGroupWidthFor(4, SCALE ? FP32 : FP64)) {
    case 1: return launchPruneG<SC, T, SCALE, 1>(a, lds, nBlocks, s);
    default: return launchPruneG<SC, T, SCALE, 4>(a, lds, nBlocks, s);
    }
}

template <int SC>
static hipError_t launchPruneS(const PruneArgs &a, int precision, size_t lds, int nBlocks, hipStream_t s)
{
    if (precision == FP32) return launchPruneP<SC, float, true>(a, lds, nBlocks, s);
    return launchPruneP<SC, double, false>(a, lds, nBlocks, s);
}

size_t pruneLdsBytes(int nKids, int K, int dim, int precision)
{
    if (!g_tuneLds[precision & 1]) return 0;
    const size_t per = (size_t)K * dim * dim * (precision == FP32 ? 4 : 8);
    if (per % 16) return 0;  // staged in 16-byte chunks
    const size_t b = (size_t)nKids * per;
    return b <= (size_t)LDS_P_BUDGET ? b : 0;
}

hipError_t launchPrune(const PruneArgs &a, int precision, int nBlocks, hipStream_t s)
{
    const size_t lds = pruneLdsBytes(a.nKids, a.K, a.dim, precision);
    switch (a.dim) {
    case 2: return launchPruneS<2>(a, precision, lds, nBlocks, s);
    case 4: return launchPruneS<4>(a, precision, lds, nBlocks, s);
    case 20: return launchPruneS<20>(a, precision, lds, nBlocks, s);
    default: return launchPruneS<0>(a, precision, lds, nBlocks, s);
    }
}

template <int SC, typename T, bool SCALE, int G>
static hipError_t launchBatchG(const PruneArgs *d, const int *bi, int nBlocks, size_t lds, hipStream_t s)
{
    if (lds)
        hipLaunchKernelGGL((k_pruneBatch<SC, T, SCALE, true, G>), dim3(nBlocks), dim3(PRUNE_THREADS), lds, s, d, bi);
    else
        hipLaunchKernelGGL((k_pruneBatch<SC, T, SCALE, false, G>), dim3(nBlocks), dim3(PRUNE_THREADS), 0, s, d, bi);
    return hipGetLastError();
}

template <int SC, typename T, bool SCALE>
static hipError_t launchBatchP(const PruneArgs *d, const int *bi, int nBlocks, size_t lds, hipStream_t s)
{
    if (SC != 4) return launchBatchG<SC, T, SCALE, 1>(d, bi, nBlocks, lds, s);
    switch (pruneGroupWidthFor(4, SCALE ? FP32 : FP64)) {
    case 1: return launchBatchG<SC, T, SCALE, 1>(d, bi, nBlocks, lds, s);
    default: return launchBatchG<SC, T, SCALE, 4>(d, bi, nBlocks, lds, s);
    }
}

template <int SC>
static hipError_t launchBatchS(const PruneArgs *d, const int *bi, int nBlocks, int precision, size_t lds,
                               hipStream_t s)
{
    if (precision == FP32) return launchBatchP<SC, float, true>(d, bi, nBlocks, lds, s);
    return launchBatchP<SC, double, false>(d, bi, nBlocks, lds, s);
}

hipError_t launchPruneBatch(const PruneArgs *dArgs, const int *dBlockItem, int nBlocks, int dim,
                            int precision, size_t lds, hipStream_t s)
{
    switch (dim) {
    case 2: return launchBatchS<2>(dArgs, dBlockItem, nBlocks, precision, lds, s);
    case 4: return launchBatchS<4>(dArgs, dBlockItem, nBlocks, precision, lds, s);
    case 20: return launchBatchS<20>(dArgs, dBlockItem, nBlocks, precision, lds, s);
    default: return launchBatchS<0>(dArgs, dBlockItem, nBlocks, precision, lds, s);
    }
}

// ------------------------------------------------ 20-state pruning on MFMA
// One wavefront owns a tile of 16 site patterns (all gamma categories); a
// 256-thread workgroup owns 4 tiles.  Partials of a 20-state FP32 part are
// tile-blocked, [slot][tile][cat][t = state/4][state%4][pattern%16], so the
// B operand of v_mfma_f32_16x16x4f32 for k-step t is 64 consecutive floats.
// Per child and category: out[s][p] = sum_j P[s][j] x[j][p] as 5 k-steps
// of 4 on two 16-row tiles (states 0-15 and 16-19; rows 20-31 are zero
// padding), with A[s][j] = P^T[j][s] read from the P^T table.
// Lane l holds A[l%16][l/16], B[l/16][l%16] and D rows 4*(l/16)+r, r<4,
// of column l%16.
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int AA_S = 20;
constexpr int AA_TILE = 16;

__device__ __forceinline__ f4v mfma4(float a, float b, f4v c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int KC>
__device__ __forceinline__ void pruneAaBody(const PruneArgs &a, int blk, int nBlk)
{
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int q = lane >> 4, pl = lane & 15;
    const int K = a.K;
    const int tile = blk * (PRUNE_THREADS / 64) + wave;
    const int p = tile * AA_TILE + pl;
    const size_t nPatPad = (size_t)a.nPatPad;
    const size_t tileFloats = (size_t)K * AA_S * AA_TILE;
    const size_t slotFloats = nPatPad * K * AA_S;
    const float *PT = (const float *)a.PT;
    float *cl = (float *)a.cl;
    const f4v zero = {0.f, 0.f, 0.f, 0.f};
    const f4v ones = {1.f, 1.f, 1.f, 1.f};

    double contrib = 0.0;
    int bad = 0;
    for (int g = 0; g < a.nGroups; ++g) {
        const GroupRec gr = a.groups[g];
        for (int o = 0; o < gr.nOps; ++o) {
            const OpRec op = a.ops[gr.opStart + o];
            f4v pr0[KC], pr1[KC];
            int e = 0;
            for (int c = 0; c < op.nKids; ++c) {
                const KidRec kd = a.kids[op.kidStart + c];
                const float *Pn = PT + (size_t)kd.pnode * K * AA_S * AA_S;
                f4v v0[KC], v1[KC];
                bool skip = false;
                if (kd.kind == 0) {
                    const unsigned code = a.tips[(size_t)kd.idx * nPatPad + p];
                    skip = (code == TIP_ALL) && c != 0;
#pragma unroll
                    for (int k = 0; k < KC; k++) {
                        if (k >= K) break;
                        const float *Pk = Pn + k * AA_S * AA_S;
                        if (code < (unsigned)AA_S) {
                            v0[k] = *reinterpret_cast<const f4v *>(Pk + code * AA_S + 4 * q);
                            v1[k] = q == 0 ? *reinterpret_cast<const f4v *>(Pk + code * AA_S + 16) : zero;
                        } else if (code == TIP_ALL) {
                            v0[k] = ones;
                            v1[k] = q == 0 ? ones : zero;
                        } else {
                            const uint64_t m = a.ambMask[code - AA_S];
                            v0[k] = zero;
                            v1[k] = zero;
                            for (int j = 0; j < AA_S; j++)
                                if ((m >> j) & 1ull) {
                                    v0[k] += *reinterpret_cast<const f4v *>(Pk + j * AA_S + 4 * q);
                                    if (q == 0) v1[k] += *reinterpret_cast<const f4v *>(Pk + j * AA_S + 16);
                                }
                        }
                    }
                } else {
                    const float *X = cl + (size_t)kd.idx * slotFloats + (size_t)tile * tileFloats;
#pragma unroll
                    for (int k = 0; k < KC; k++) {
                        if (k >= K) break;
                        const float *Pk = Pn + k * AA_S * AA_S;
                        f4v acc0 = zero, acc1 = zero;
#pragma unroll
                        for (int t = 0; t < 5; t++) {
                            const float b = X[(k * 5 + t) * 64 + lane];
                            const float a0 = Pk[(4 * t + q) * AA_S + pl];
                            const float a1 = pl < 4 ? Pk[(4 * t + q) * AA_S + 16 + pl] : 0.f;
                            acc0 = mfma4(a0, b, acc0);
                            acc1 = mfma4(a1, b, acc1);
                        }
                        v0[k] = acc0;
                        v1[k] = acc1;
                    }
                    e += a.scl[(size_t)kd.idx * nPatPad + p];
                }
#pragma unroll
                for (int k = 0; k < KC; k++) {
                    if (k >= K) break;
                    if (c == 0) {
                        pr0[k] = v0[k];
                        pr1[k] = v1[k];
                    } else if (!skip) {
                        pr0[k] *= v0[k];
                        pr1[k] *= v1[k];
                    }
                }
            }
            // rescale: max over categories and states of this pattern
            float m = 0.f;
#pragma unroll
            for (int k = 0; k < KC; k++) {
                if (k >= K) break;
                for (int r = 0; r < 4; r++) m = fmaxf(m, fmaxf(pr0[k][r], pr1[k][r]));
            }
            m = fmaxf(m, __shfl_xor(m, 16));
            m = fmaxf(m, __shfl_xor(m, 32));
            if (m < SCALE_THRESH && m > 0.f) {
                int ex;
                frexpf(m, &ex);
                const int E = -ex;
#pragma unroll
                for (int k = 0; k < KC; k++) {
                    if (k >= K) break;
                    for (int r = 0; r < 4; r++) {
                        pr0[k][r] = ldexpf(pr0[k][r], E);
                        pr1[k][r] = ldexpf(pr1[k][r], E);
                    }
                }
                e += E;
            }
            if (!(op.flags & OP_ROOT)) {
                float *D = cl + (size_t)op.dst * slotFloats + (size_t)tile * tileFloats;
#pragma unroll
                for (int k = 0; k < KC; k++) {
                    if (k >= K) break;
                    for (int r = 0; r < 4; r++) D[((k * 5 + q) * 4 + r) * AA_TILE + pl] = pr0[k][r];
                    if (q == 0)
                        for (int r = 0; r < 4; r++) D[((k * 5 + 4) * 4 + r) * AA_TILE + pl] = pr1[k][r];
                }
                if (q == 0) a.scl[(size_t)op.dst * nPatPad + p] = e;
                continue;
            }
            // root: sum_k sum_s pi_s cl_s over this lane's states, then
            // across the four lanes of the pattern
            const double *pi = a.rootPi;
            uint64_t msk = ~0ull;
            if (a.rootLeafRow >= 0) {
                const unsigned code = a.tips[(size_t)a.rootLeafRow * nPatPad + p];
                msk = code < (unsigned)AA_S ? (1ull << code) : code == TIP_ALL ? ~0ull : a.ambMask[code - AA_S];
            }
            double Lk = 0.0;
#pragma unroll
            for (int k = 0; k < KC; k++) {
                if (k >= K) break;
                for (int r = 0; r < 4; r++) {
                    const int s0 = 4 * q + r;
                    if ((msk >> s0) & 1ull) Lk += pi[s0] * (double)pr0[k][r];
                    if (q == 0 && ((msk >> (16 + r)) & 1ull)) Lk += pi[16 + r] * (double)pr1[k][r];
                }
            }
            Lk += __shfl_xor(Lk, 16);
            Lk += __shfl_xor(Lk, 32);
            if (q == 0) siteLogLike<true>(a, Lk, e, p, K, pi, contrib, bad);
        }
    }
    finishSiteSum(a, contrib, bad, blk, nBlk);
}

template <int KC>
__global__ __launch_bounds__(PRUNE_THREADS) void k_pruneAa(PruneArgs a)
{
    pruneAaBody<KC>(a, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void finishSiteSumWave(const PruneArgs &a, double contrib, int bad, int blk, int nBlk);

// Category-split variant: one wavefront per (16-pattern tile, category),
// the K category waves of a tile in one workgroup, so a launch has K times
// the wavefronts and each walks the program with a K-th of the work.  The
// scale exponent stays per (node, pattern): every wave posts its
// per-pattern maximum in LDS (double-buffered, one barrier per op) and all
// take the maximum over categories, so partials and exponents are those of
// k_pruneAa bit for bit; only the FP64 root sum over categories is added
// in another order.
//
// Children computed earlier in the same launch come from an LDS slot file
// (a.nLds slots; op flags bits 8.. = the result's slot + 1, kid kind 2 =
// read slot KidRec.pad): per slot and category wave the tile's 20 x 16
// values in the global tile layout, and the wave's own copy of the 16
// exponents (every wave computes the same exponents, so no wave reads
// another's).  Only the wave that wrote a slot region reads it, in program
// order, so no barrier is needed for it.  With a.tipLds the tile's tip
// codes ([nTax][16] bytes) are staged in LDS once.
//
// Stores are deferred by one op.  The vector memory counter retires loads
// and stores in issue order, so a load issued after an op's HBM stores
// waits for their acknowledgement as well.  Here the P rows (and tip rows)
// of an op's first two children, which depend on the program only, are
// loaded while the op before it finishes; the op then issues the previous
// op's HBM stores (from LDS, where each result is staged), reads its
// children's partials (LDS, or HBM after those stores) and computes, so no
// load queues behind a store.  Every wave writes the (identical) exponents
// of a result, so a later op reads exponents its own wave wrote.
template <bool TIPLDS>
__device__ __forceinline__ void pruneAaCatBody(const PruneArgs &a, int tile, int nTiles)
{
    __shared__ float mx[2][WAVE_MAX_K][AA_TILE];
    __shared__ double rootL[WAVE_MAX_K][AA_TILE];
    __shared__ int rootE[AA_TILE];
    // [nLds + 1][K][320] values, [nLds + 1][K][16] exponents (region nLds:
    // the staged result of an op without a slot), [nTax][16] tip codes
    extern __shared__ float aaSlots[];
    const int tid = threadIdx.x;
    const int k = __builtin_amdgcn_readfirstlane(tid >> 6);   // category
    const int lane = tid & 63;
    const int q = lane >> 4, pl = lane & 15;
    const int K = a.K;
    const int p = tile * AA_TILE + pl;
    const size_t nPatPad = (size_t)a.nPatPad;
    const size_t tileFloats = (size_t)K * AA_S * AA_TILE;
    const size_t slotFloats = nPatPad * K * AA_S;
    const float *PT = (const float *)a.PT;
    float *cl = (float *)a.cl;
    const f4v zero = {0.f, 0.f, 0.f, 0.f};
    const f4v ones = {1.f, 1.f, 1.f, 1.f};

    int *slotEx = reinterpret_cast<int *>(aaSlots + (size_t)(a.nLds + 1) * K * 320);
    uint8_t *tipL = reinterpret_cast<uint8_t *>(slotEx + (size_t)(a.nLds + 1) * K * AA_TILE);
    if (TIPLDS) {
        for (int t = tid; t < a.nTax; t += blockDim.x)
            *reinterpret_cast<uint4 *>(tipL + t * AA_TILE) =
                *reinterpret_cast<const uint4 *>(a.tips + (size_t)t * nPatPad + (size_t)tile * AA_TILE);
        __syncthreads();
    }
    // (two separate loads, not a select of pointers: a generic pointer
    // would make every tip code a flat load waiting on both counters)
    auto tipCode = [&](int row) -> unsigned {
        if (TIPLDS) return (unsigned)tipL[row * AA_TILE + pl];
        return (unsigned)a.tips[(size_t)row * nPatPad + p];
    };
    // a tip row of P: this lane's four states, and states 16..19 (used on
    // q == 0 only; pr1 of the other lanes is cleared before it is used)
    auto tipRow = [&](const float *Pk, unsigned code, f4v &v0, f4v &v1) {
        if (code < (unsigned)AA_S) {
            v0 = *reinterpret_cast<const f4v *>(Pk + code * AA_S + 4 * q);
            v1 = *reinterpret_cast<const f4v *>(Pk + code * AA_S + 16);
        } else if (code == TIP_ALL) {
            v0 = ones;
            v1 = ones;
        } else {
            const uint64_t m = a.ambMask[code - AA_S];
            v0 = zero;
            v1 = zero;
            for (int j = 0; j < AA_S; j++)
                if ((m >> j) & 1ull) {
                    v0 += *reinterpret_cast<const f4v *>(Pk + j * AA_S + 4 * q);
                    v1 += *reinterpret_cast<const f4v *>(Pk + j * AA_S + 16);
                }
        }
    };
    // The pending result: an op leaves its values and exponents in LDS (its
    // slot, or region nLds) and the next op copies them to HBM after its
    // own loads, so nothing of it occupies registers meanwhile.  Addresses
    // are a wave-uniform base plus a 32-bit unsigned lane byte offset (the
    // scalar-base + vector-offset form of the global instructions).
    bool pend = false;
    int pDst = 0, pSrc = 0;
    const unsigned laneB = (unsigned)(q * 64 + pl) * 4u;   // state 4q, pattern pl of a [20][16] tile
    const unsigned laneW = (unsigned)lane * 4u;            // state q (of four), pattern pl
    const unsigned plB = (unsigned)pl * 4u;
    // Every lane stores the same number of words (no branch on q): states
    // 0..15 in four stores, states 16 + q in one, and the pattern's
    // exponent (four lanes write each, the same value).  The stores are
    // non-temporal: the launch's partials stream through L2 once, and
    // would otherwise evict the P matrices every op gathers from there.
    auto flush = [&]() {
        const float *L = aaSlots + ((size_t)pSrc * K + k) * 320;
        char *D = (char *)(cl + (size_t)pDst * slotFloats + (size_t)tile * tileFloats + (size_t)k * 320);
        float v[5];
        for (int r = 0; r < 4; r++) v[r] = L[(q * 4 + r) * AA_TILE + pl];
        v[4] = L[256 + q * AA_TILE + pl];
        const int ex = slotEx[((size_t)pSrc * K + k) * AA_TILE + pl];
        for (int r = 0; r < 4; r++) __builtin_nontemporal_store(v[r], (float *)(D + 64 * r + laneB));
        __builtin_nontemporal_store(v[4], (float *)(D + 1024 + laneW));
        __builtin_nontemporal_store(ex, (int *)((char *)(a.scl + (size_t)pDst * nPatPad + (size_t)tile * AA_TILE) + plB));
        pend = false;
    };

    // The program: op i's fused record (AaRec, 16 words) is loaded one op
    // ahead, lanes 0..11 one word each (a single vector load into one
    // register), and taken apart with readlane when the op before it
    // finishes: a scalar load there would wait out a scalar-cache miss on
    // most ops.  Children past the second (a root with three or more) come
    // from the kid records.
    typedef int i4v __attribute__((ext_vector_type(4)));
    typedef const __attribute__((address_space(4))) i4v ci4v;
    ci4v *kidsC = (ci4v *)a.kids;
    auto kidAt = [&](int i) -> KidRec {
        const i4v r = kidsC[i];
        return KidRec{r.x, r.y, r.z, r.w};
    };
    const int *recG = a.aaRec;
    auto recLoad = [&](int i) -> int { return lane < 12 ? recG[(size_t)i * AA_REC_WORDS + lane] : 0; };
    auto recTake = [&](int v) -> AaRec {
        AaRec r;
        r.dst = __builtin_amdgcn_readlane(v, 0);
        r.flags = __builtin_amdgcn_readlane(v, 1);
        r.nKids = __builtin_amdgcn_readlane(v, 2);
        r.kidStart = __builtin_amdgcn_readlane(v, 3);
        for (int c = 0; c < 2; c++) {
            r.kind[c] = __builtin_amdgcn_readlane(v, 4 + 4 * c);
            r.idx[c] = __builtin_amdgcn_readlane(v, 5 + 4 * c);
            r.pnode[c] = __builtin_amdgcn_readlane(v, 6 + 4 * c);
            r.slot[c] = __builtin_amdgcn_readlane(v, 7 + 4 * c);
        }
        return r;
    };

    // The first two children of an op.  A partials child loads its MFMA A
    // operands: rows (4t + q) of P at column pl, and at column 16 + (pl & 3)
    // (states 16..19; lanes pl >= 4 repeat them, which only fills output
    // rows 4..15 of acc1, on lanes q > 0, cleared before use).  A tip child
    // loads the P row of its code: entries 4q..4q+3 in R[0..3] and 16..19
    // in R[6..9].  These loads depend on the program only, so an op's are
    // issued while the op before it finishes (after its products, before
    // its barrier); the op's HBM stores of the previous result always
    // follow them, so no wait on them reaches back past those stores.
    float R[2][10];
    int kind[2];
    unsigned code[2];
    // An op with one child loads that child twice: every op issues the same
    // twenty loads, so the wait counts the compiler derives do not depend
    // on the path (a shorter path would lower them for all).
    auto issueLoads = [&](const AaRec &op) {
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int cc = c < op.nKids ? c : 0;
            const KidRec kd{op.kind[cc], op.idx[cc], op.pnode[cc], op.slot[cc]};
            const bool tip = kd.kind == 0;
            kind[c] = kd.kind;
            const unsigned cd = tipCode(tip ? kd.idx : 0);
            code[c] = cd;
            const char *Pb = (const char *)PT + (size_t)(unsigned)(kd.pnode * K + k) * (AA_S * AA_S * 4);
            if (tip) {
                // two 16-byte loads per lane: a tip row touches one or two
                // lines per pattern, and ten dword gathers would touch them
                // five times over
                const unsigned row = (cd < (unsigned)AA_S ? cd : 0u) * (AA_S * 4u);
                const f4v v0 = *(const f4v *)(Pb + row + 16u * q);
                const f4v v1 = *(const f4v *)(Pb + row + 64u);
                for (int r = 0; r < 4; r++) {
                    R[c][r] = v0[r];
                    R[c][6 + r] = v1[r];
                }
            } else {
                const unsigned o0 = (unsigned)(q * AA_S + pl) * 4u;
                const unsigned o1 = (unsigned)(q * AA_S + 16 + (pl & 3)) * 4u;
#pragma unroll
                for (int t = 0; t < 5; t++) {
                    R[c][t] = *(const float *)(Pb + 320 * t + o0);
                    R[c][5 + t] = *(const float *)(Pb + 320 * t + o1);
                }
            }
        }
    };

    double contrib = 0.0;
    int bad = 0;
    int par = 0;
    bool sawRoot = false;
    // the ops of all groups are contiguous, in program order (root last)
    // Every op flushes a pending result (a fixed number of stores on every
    // path, for the same reason).  The first op has none: it flushes the
    // staging region (unwritten) to its own destination, which its real
    // result overwrites one op later, in order, before anything reads it.
    // Programs here have two or more ops, the root last.
    const int nOps = a.nOps;
    int recNext = recLoad(1);
    AaRec op = recTake(recLoad(0));
    pend = true;
    pDst = op.dst;
    pSrc = a.nLds;
    issueLoads(op);
    for (int i = 0; i < nOps; ++i) {
        {
            // the previous op's HBM stores, behind this op's loads
            flush();
            // products
            f4v pr0 = zero, pr1 = zero;
            int e = 0;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                if (c >= op.nKids) break;
                f4v v0, v1;
                if (kind[c] == 0) {
                    // a one-state code: its P row; all states: 1 (a product
                    // with 1.0 leaves the other factor's bits unchanged, as
                    // skipping it would); no branch on the lane's code
                    const unsigned cd = code[c];
                    const bool all = cd == TIP_ALL;
                    for (int r = 0; r < 4; r++) {
                        v0[r] = all ? 1.f : R[c][r];
                        v1[r] = all ? 1.f : R[c][6 + r];
                    }
                    // partial ambiguity (rare): the sum of its states' P rows
                    const bool amb = cd >= (unsigned)AA_S && !all;
                    if (__builtin_amdgcn_ballot_w64(amb)) {
                        if (amb) {
                            const float *Pk = PT + ((size_t)op.pnode[c] * K + k) * AA_S * AA_S;
                            const uint64_t m = a.ambMask[cd - AA_S];
                            v0 = zero;
                            v1 = zero;
                            for (int j = 0; j < AA_S; j++)
                                if ((m >> j) & 1ull) {
                                    v0 += *reinterpret_cast<const f4v *>(Pk + j * AA_S + 4 * q);
                                    v1 += *reinterpret_cast<const f4v *>(Pk + j * AA_S + 16);
                                }
                        }
                        // settle the rare path's loads here (see the HBM branch below)
                        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
                    }
                } else {
                    // (each branch reads through its own pointer: a pointer
                    // chosen between LDS and HBM would become a flat access)
                    const KidRec kd{op.kind[c], op.idx[c], op.pnode[c], op.slot[c]};
                    f4v acc0 = zero, acc1 = zero;
                    // (the products sit in each branch, so the HBM branch's
                    // loads are consumed before the join and leave nothing
                    // for the LDS branch to wait on)
                    float b[5];
                    if (kd.kind == 2) {
                        const unsigned sr = (unsigned)(kd.pad * K + k);   // slot region
                        const float *X = aaSlots + sr * 320u;
                        e += slotEx[sr * AA_TILE + pl];
#pragma unroll
                        for (int t = 0; t < 5; t++) b[t] = X[t * 64 + lane];
#pragma unroll
                        for (int t = 0; t < 5; t++) {
                            acc0 = mfma4(R[c][t], b[t], acc0);
                            acc1 = mfma4(R[c][5 + t], b[t], acc1);
                        }
                    } else {
                        const char *X =
                            (const char *)(cl + (size_t)kd.idx * slotFloats + (size_t)tile * tileFloats + (size_t)k * 320);
                        e += *(const int *)((const char *)(a.scl + (size_t)kd.idx * nPatPad + (size_t)tile * AA_TILE) + plB);
#pragma unroll
                        for (int t = 0; t < 5; t++) b[t] = *(const float *)(X + 256 * t + laneW);
#pragma unroll
                        for (int t = 0; t < 5; t++) {
                            acc0 = mfma4(R[c][t], b[t], acc0);
                            acc1 = mfma4(R[c][5 + t], b[t], acc1);
                        }
                        // settle this branch's loads here: left pending past the
                        // join, they make the LDS branch wait on every
                        // outstanding store as well (the wait counter is
                        // merged across paths conservatively, so every rare
                        // path that loads ends with such a wait)
                        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
                    }
                    v0 = acc0;
                    v1 = acc1;
                }
                if (c == 0) {
                    pr0 = v0;
                    pr1 = v1;
                } else {
                    pr0 *= v0;
                    pr1 *= v1;
                }
            }
            // further children (a root with three or more) load and multiply in turn
            for (int c = 2; c < op.nKids; ++c) {
                const KidRec kd = kidAt(op.kidStart + c);
                const float *Pk = PT + ((size_t)kd.pnode * K + k) * AA_S * AA_S;
                f4v v0, v1;
                bool skip = false;
                if (kd.kind == 0) {
                    const unsigned code = tipCode(kd.idx);
                    skip = code == TIP_ALL;
                    tipRow(Pk, code, v0, v1);
                } else {
                    f4v acc0 = zero, acc1 = zero;
                    float b[5];
                    if (kd.kind == 2) {
                        const float *X = aaSlots + ((size_t)kd.pad * K + k) * 320;
                        e += slotEx[((size_t)kd.pad * K + k) * AA_TILE + pl];
#pragma unroll
                        for (int t = 0; t < 5; t++) b[t] = X[t * 64 + lane];
                    } else {
                        const float *X = cl + (size_t)kd.idx * slotFloats + (size_t)tile * tileFloats + (size_t)k * 320;
                        e += a.scl[(size_t)kd.idx * nPatPad + p];
#pragma unroll
                        for (int t = 0; t < 5; t++) b[t] = X[t * 64 + lane];
                    }
#pragma unroll
                    for (int t = 0; t < 5; t++) {
                        acc0 = mfma4(Pk[(4 * t + q) * AA_S + pl], b[t], acc0);
                        acc1 = mfma4(Pk[(4 * t + q) * AA_S + 16 + (pl & 3)], b[t], acc1);
                    }
                    v0 = acc0;
                    v1 = acc1;
                }
                if (!skip) {
                    pr0 *= v0;
                    pr1 *= v1;
                }
                __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): a rare path, settled here
            }
            if (q != 0) pr1 = zero;   // rows 4..15 of the states-16..19 products
            // the next op's loads go out now (R is free), so their latency
            // overlaps this op's barrier, rescale and result write
            // (unconditional, past the end the last op again: every pass
            // through the loop issues the same loads)
            const AaRec cur = op;
            op = recTake(recNext);
            issueLoads(op);
            recNext = recLoad(min(i + 2, nOps - 1));
            // rescale: max over this wave's states, then over the categories
            float m = 0.f;
            for (int r = 0; r < 4; r++) m = fmaxf(m, fmaxf(pr0[r], pr1[r]));
            m = fmaxf(m, __shfl_xor(m, 16));
            m = fmaxf(m, __shfl_xor(m, 32));
            if (q == 0) mx[par][k][pl] = m;
            __syncthreads();
            m = 0.f;
            for (int kk = 0; kk < K; kk++) m = fmaxf(m, mx[par][kk][pl]);
            par ^= 1;   // the next op posts into the other buffer: no second barrier
            if (m < SCALE_THRESH && m > 0.f) {
                int ex;
                frexpf(m, &ex);
                const int E = -ex;
                for (int r = 0; r < 4; r++) {
                    pr0[r] = ldexpf(pr0[r], E);
                    pr1[r] = ldexpf(pr1[r], E);
                }
                e += E;
            }
            if (!(cur.flags & OP_ROOT)) {
                // into the op's slot (the parent reads it there), or region
                // nLds; the next op copies it to HBM
                const int ld = (cur.flags >> 8) - 1;
                const int dl = ld >= 0 ? ld : a.nLds;
                float *L = aaSlots + ((size_t)dl * K + k) * 320;
                for (int r = 0; r < 4; r++) L[(q * 4 + r) * AA_TILE + pl] = pr0[r];
                if (q == 0) {
                    for (int r = 0; r < 4; r++) L[(4 * 4 + r) * AA_TILE + pl] = pr1[r];
                    slotEx[((size_t)dl * K + k) * AA_TILE + pl] = e;
                }
                pend = true;
                pDst = cur.dst;
                pSrc = dl;
                continue;
            }
            // root: this category's sum_s pi_s cl_s across the four lanes of
            // the pattern, then wave 0 adds the categories in order
            const double *pi = a.rootPi;
            uint64_t msk = ~0ull;
            if (a.rootLeafRow >= 0) {
                const unsigned code = tipCode(a.rootLeafRow);
                msk = code < (unsigned)AA_S ? (1ull << code) : code == TIP_ALL ? ~0ull : a.ambMask[code - AA_S];
            }
            double Lk = 0.0;
            for (int r = 0; r < 4; r++) {
                const int s0i = 4 * q + r;
                if ((msk >> s0i) & 1ull) Lk += pi[s0i] * (double)pr0[r];
                if (q == 0 && ((msk >> (16 + r)) & 1ull)) Lk += pi[16 + r] * (double)pr1[r];
            }
            Lk += __shfl_xor(Lk, 16);
            Lk += __shfl_xor(Lk, 32);
            if (q == 0) rootL[k][pl] = Lk;
            __syncthreads();
            if (k == 0 && q == 0) {
                double L = 0.0;
                for (int kk = 0; kk < K; kk++) L += rootL[kk][pl];
                rootL[0][pl] = L;   // the site term is taken after the loop, where its
                rootE[pl] = e;      // constants do not occupy registers across it
            }
            sawRoot = true;
            __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the root's loads, settled here
        }
    }
    if (pend) flush();
    if (sawRoot && k == 0 && q == 0) siteLogLike<true>(a, rootL[0][pl], rootE[pl], p, K, a.rootPi, contrib, bad);
    if (k == 0) finishSiteSumWave(a, contrib, bad, tile, nTiles);
}

// at most 102 VGPRs: five waves per SIMD, so prot128's 4,976 category
// waves are resident in one round on 1,024 SIMDs (four per SIMD took two)
template <bool TIPLDS>
__global__ __launch_bounds__(64 * WAVE_MAX_K) __attribute__((amdgpu_waves_per_eu(5))) void k_pruneAaCat(PruneArgs a)
{
    pruneAaCatBody<TIPLDS>(a, blockIdx.x, gridDim.x);
}

// TIPLDS: every item of the launch stages its tip codes (items that could
// are read from global memory by the other instance, which ignores tipLds)
template <bool TIPLDS>
__global__ __launch_bounds__(64 * WAVE_MAX_K) __attribute__((amdgpu_waves_per_eu(5))) void k_pruneAaCatBatch(
    const PruneArgs *args, const int *blockItem)
{
    const int item = blockItem[2 * blockIdx.x];
    const int local = blockItem[2 * blockIdx.x + 1];
    const PruneArgs a = args[item];
    pruneAaCatBody<TIPLDS>(a, local, a.nPatPad / AA_TILE);
}

// Category waves pay one barrier per op; short programs (dirty paths)
// run faster on the tile kernel, long ones (full evaluations) on these.
static int g_aaCat = 1;
static int g_aaCatMinOps = AA_CAT_MIN_OPS;

void setAaTuning(int catWaves, int minOps)
{
    g_aaCat = catWaves ? 1 : 0;
    g_aaCatMinOps = minOps < 0 ? 0 : minOps;
}

bool pruneAaCategoryWaves(int nOps) { return g_aaCat != 0 && nOps >= std::max(2, g_aaCatMinOps); }

// LDS slot file of the category kernel: per slot, K x (320 values + 16
// exponents), and one more such region for a result without a slot.  Four
// slots and that region for K = 4 take 26 KB, and the tip codes of up to
// 128 taxa 2 KB more, so five workgroups (the five waves per SIMD the
// kernel is built for) fit a CU's 160 KB.
static int g_aaLdsSlots = 5;
constexpr size_t AA_LDS_BUDGET = 28 * 1024;   // slots and the staging region
void setAaLdsSlots(int n) { g_aaLdsSlots = n < 0 ? 0 : n; }
int pruneAaLdsSlots(int K)
{
    const size_t per = (size_t)K * (320 + 16) * 4;
    const size_t fit = AA_LDS_BUDGET / per;
    return (int)std::min<size_t>((size_t)g_aaLdsSlots, fit > 0 ? fit - 1 : 0);
}
// the tile's tip codes go to LDS up to this many taxa (4 KB)
constexpr int AA_TIP_LDS_MAX_TAXA = 256;
bool pruneAaTipLds(int nTax) { return nTax <= AA_TIP_LDS_MAX_TAXA; }
size_t pruneAaLdsBytes(int K, int nLds, int tipTaxa)
{
    return (size_t)(nLds + 1) * K * (320 + 16) * 4 + (size_t)tipTaxa * AA_TILE;
}
static bool aaLdsArgsOk(const PruneArgs &a)
{
    return a.nLds >= 0 && a.nLds <= pruneAaLdsSlots(a.K) && (!a.tipLds || pruneAaTipLds(a.nTax)) &&
           a.aaRec != nullptr && a.nOps >= 2;
}

template <int KC>
__global__ __launch_bounds__(PRUNE_THREADS) void k_pruneAaBatch(const PruneArgs *args, const int *blockItem)
{
    const int item = blockItem[2 * blockIdx.x];
    const int local = blockItem[2 * blockIdx.x + 1];
    const PruneArgs a = args[item];
    pruneAaBody<KC>(a, local, a.nPatPad / (AA_TILE * PRUNE_THREADS / 64));
}

bool pruneUsesMfma(int dim, int K, int precision) { return dim == AA_S && precision == FP32 && K <= 8; }

int pruneBlockPatternsFor(int dim, int K, int precision, int nOps)
{
    if (pruneUsesMfma(dim, K, precision))
        return pruneAaCategoryWaves(nOps) ? AA_TILE : AA_TILE * PRUNE_THREADS / 64;
    return pruneBlockPatterns(K);
}

static int aaKC(int K) { return K <= 1 ? 1 : K <= 2 ? 2 : K <= 4 ? 4 : 8; }

hipError_t launchPruneAa(const PruneArgs &a, int nBlocks, hipStream_t s)
{
    if (pruneAaCategoryWaves(a.nOps)) {
        if (a.K < 1 || a.K > WAVE_MAX_K || nBlocks * AA_TILE > a.nPatPad || !aaLdsArgsOk(a))
            return hipErrorInvalidValue;
        const size_t lds = pruneAaLdsBytes(a.K, a.nLds, a.tipLds ? a.nTax : 0);
        if (a.tipLds)
            hipLaunchKernelGGL(k_pruneAaCat<true>, dim3(nBlocks), dim3(64 * a.K), lds, s, a);
        else
            hipLaunchKernelGGL(k_pruneAaCat<false>, dim3(nBlocks), dim3(64 * a.K), lds, s, a);
        return hipGetLastError();
    }
    if (nBlocks * AA_TILE * (PRUNE_THREADS / 64) > a.nPatPad) return hipErrorInvalidValue;
    switch (aaKC(a.K)) {
    case 1: hipLaunchKernelGGL(k_pruneAa<1>, dim3(nBlocks), dim3(PRUNE_THREADS), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_pruneAa<2>, dim3(nBlocks), dim3(PRUNE_THREADS), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_pruneAa<4>, dim3(nBlocks), dim3(PRUNE_THREADS), 0, s, a); break;
    default: hipLaunchKernelGGL(k_pruneAa<8>, dim3(nBlocks), dim3(PRUNE_THREADS), 0, s, a); break;
    }
    return hipGetLastError();
}

bool pruneAaCatItemOk(const PruneArgs &a) { return aaLdsArgsOk(a); }

hipError_t launchPruneAaBatch(const PruneArgs *d, const int *bi, int nBlocks, int K, bool catWaves, size_t ldsBytes,
                              bool tipLds, hipStream_t s)
{
    if (catWaves) {
        // every item of the launch has exactly K categories (the batch groups
        // by K); ldsBytes: the largest slot file and tip codes of its items;
        // tipLds: every item stages its tip codes
        if (K < 1 || K > WAVE_MAX_K ||
            ldsBytes > pruneAaLdsBytes(K, pruneAaLdsSlots(K), AA_TIP_LDS_MAX_TAXA))
            return hipErrorInvalidValue;
        if (tipLds)
            hipLaunchKernelGGL(k_pruneAaCatBatch<true>, dim3(nBlocks), dim3(64 * K), ldsBytes, s, d, bi);
        else
            hipLaunchKernelGGL(k_pruneAaCatBatch<false>, dim3(nBlocks), dim3(64 * K), ldsBytes, s, d, bi);
        return hipGetLastError();
    }
    switch (aaKC(K)) {
    case 1: hipLaunchKernelGGL(k_pruneAaBatch<1>, dim3(nBlocks), dim3(PRUNE_THREADS), 0, s, d, bi); break;
    case 2: hipLaunchKernelGGL(k_pruneAaBatch<2>, dim3(nBlocks), dim3(PRUNE_THREADS), 0, s, d, bi); break;
    case 4: hipLaunchKernelGGL(k_pruneAaBatch<4>, dim3(nBlocks), dim3(PRUNE_THREADS), 0, s, d, bi); break;
    default: hipLaunchKernelGGL(k_pruneAaBatch<8>, dim3(nBlocks), dim3(PRUNE_THREADS), 0, s, d, bi); break;
    }
    return hipGetLastError();
}

// ------------------------------------------------ wave-per-category DNA
// One wavefront per (64-pattern tile, gamma category); a workgroup holds
// the K category waves of one tile and they share nothing but the staged
// tip codes and the root combine.  Partials [slot][cat][pattern][4].
// A child computed earlier in the same launch, or a clean child listed
// for preload, is read from the wave's LDS slot file, so the node program
// never waits on HBM between nodes; every computed partial is still
// written to HBM, where later incremental evaluations read it.  P^T and
// the program records are wave-uniform and come through scalar loads.
// FP32 exponents are per (node, pattern, category), so no lane waits on
// another until the root, where one barrier combines the categories.
typedef const __attribute__((address_space(4))) int cint_t;
template <typename T> struct ConstT;
template <> struct ConstT<float> { typedef const __attribute__((address_space(4))) float type; };
template <> struct ConstT<double> { typedef const __attribute__((address_space(4))) double type; };
template <typename T> struct Vec4T;
template <> struct Vec4T<float> { typedef float4 type; };
template <> struct Vec4T<double> { typedef double4 type; };

constexpr int WAVE_LDS_BUDGET = 48 * 1024;   // dynamic LDS per workgroup
// s_waitcnt vmcnt(0) with the other counters left alone (gfx9 encoding:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
constexpr int VMCNT0 = 0x0F70;

// leafVec for 4 states with P^T in registers.  A single-state code (the
// common case) picks its row with a 2-bit select tree; gaps, '?' and
// partial ambiguities are fixed up on a divergent path that is skipped
// when no lane of the wave holds one.  Same values and summation order as
// leafVec.
template <typename T>
__device__ __forceinline__ void leafVec4(const T *Pl, unsigned code, const uint64_t *amb, T *v)
{
    const bool b0 = code & 1u, b1 = code & 2u;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const T lo = b0 ? Pl[4 + s] : Pl[s];
        const T hi = b0 ? Pl[12 + s] : Pl[8 + s];
        v[s] = b1 ? hi : lo;
    }
    if (code >= 4u) {
        if (code == TIP_ALL) {
#pragma unroll
            for (int s = 0; s < 4; s++) v[s] = T(1);
        } else {
            const unsigned m = (unsigned)amb[code - 4] & 15u;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                T x = T(0);
#pragma unroll
                for (int j = 0; j < 4; j++) x += ((m >> j) & 1u) ? Pl[j * 4 + s] : T(0);
                v[s] = x;
            }
        }
    }
}

// The last workgroup of the item sums the per-workgroup partials (wave 0
// of each workgroup reports) in a fixed order with Neumaier compensation.
__device__ __forceinline__ void finishSiteSumWave(const PruneArgs &a, double contrib, int bad, int blk, int nBlk)
{
    const int lane = threadIdx.x & 63;
    for (int off = 32; off >= 1; off >>= 1) {
        contrib += __shfl_xor(contrib, off);
        bad |= __shfl_xor(bad, off);
    }
    unsigned prev = 0;
    if (lane == 0) {
        a.blockSums[blk] = contrib;
        a.blockBad[blk] = bad;
        __threadfence();
        prev = __hip_atomic_fetch_add(a.counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    }
    prev = __shfl(prev, 0);
    if (prev != (unsigned)(nBlk - 1)) return;
    __threadfence();
    double s = 0.0, comp = 0.0;
    int b = 0;
    for (int i = lane; i < nBlk; i += 64) {
        const double x = __hip_atomic_load(a.blockSums + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b |= __hip_atomic_load(a.blockBad + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const double t = s + x;  // Neumaier
        if (fabs(s) >= fabs(x))
            comp += (s - t) + x;
        else
            comp += (x - t) + s;
        s = t;
    }
    double v = s + comp;
    for (int off = 32; off >= 1; off >>= 1) {
        v += __shfl_xor(v, off);
        b |= __shfl_xor(b, off);
    }
    if (lane == 0) {
        a.result[0] = b ? -1.0e99 : v;
        a.result[1] = (double)b;
        *a.counter = 0u;
    }
}

template <typename T, bool SCALE, bool PV>
__device__ __forceinline__ void pruneWaveBody(const PruneArgs &a, int tile, int nTiles)
{
    typedef typename Vec4T<T>::type V4;
    typedef typename ConstT<T>::type cT;
    extern __shared__ __align__(16) unsigned char smemW[];
    __shared__ double rootL[WAVE_MAX_K][64];
    __shared__ int rootE[WAVE_MAX_K][64];
    const int K = a.K;
    // wave wv = (stream, category); a stream is one set of category waves
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nS = a.nStream;
    const int stream = wv / K;
    const int cat = wv - stream * K;
    const int lane = threadIdx.x & 63;
    const int p = tile * 64 + lane;
    const size_t nPP = (size_t)a.nPatPad;
    const int nL = a.nLds;
    const int tipBytes = a.tipLds ? ((a.nTax * 64 + 31) & ~31) : 0;   // 32-byte aligned: FP64 rows
    // explicit address spaces keep LDS and HBM reads of a child as separate
    // ds_ / global_ instructions (one merged generic load would wait on
    // both memory counters at every child)
    typedef __attribute__((address_space(3))) T lT;
    typedef __attribute__((address_space(3))) int lint;
    typedef __attribute__((address_space(3))) const unsigned char lu8;
    typedef __attribute__((address_space(1))) const T gT;
    typedef __attribute__((address_space(1))) const int gint;
    typedef __attribute__((address_space(1))) const unsigned char gu8;
    lu8 *sTip = (lu8 *)(smemW);
    lT *slotT = (lT *)(smemW + tipBytes) + ((size_t)wv * nL * 64 + lane) * 4;   // slot s: slotT[s * 256 + 0..3]
    lT *slot0T = (lT *)(smemW + tipBytes) + ((size_t)cat * nL * 64 + lane) * 4;  // stream 0's file, same category
    auto ldS = [&](int sl) {
        V4 x;
        x.x = slotT[sl * 256];
        x.y = slotT[sl * 256 + 1];
        x.z = slotT[sl * 256 + 2];
        x.w = slotT[sl * 256 + 3];
        return x;
    };
    auto stS = [&](int sl, const V4 &x) {
        slotT[sl * 256] = x.x;
        slotT[sl * 256 + 1] = x.y;
        slotT[sl * 256 + 2] = x.z;
        slotT[sl * 256 + 3] = x.w;
    };
    lint *sexp = (lint *)(smemW + tipBytes + sizeof(V4) * (size_t)nS * K * nL * 64) + (size_t)wv * nL * 64 + lane;
    lint *sexp0 = (lint *)(smemW + tipBytes + sizeof(V4) * (size_t)nS * K * nL * 64) + (size_t)cat * nL * 64 + lane;
    gu8 *gTips = (gu8 *)(a.tips);
    gint *gScl = (gint *)(a.scl);
    if (a.tipLds) {
        // the tile's tip codes, [taxon][64 patterns], 4 bytes per thread
        const int words = a.nTax * 16;
        for (int w = threadIdx.x; w < words; w += blockDim.x) {
            const int r = w >> 4, q = w & 15;
            reinterpret_cast<uint32_t *>(smemW)[w] =
                *reinterpret_cast<const uint32_t *>(a.tips + (size_t)r * nPP + (size_t)tile * 64 + q * 4);
        }
        __syncthreads();
    }
    gT *clr = (gT *)(a.cl);
    V4 *clw = reinterpret_cast<V4 *>(a.cl);
    auto ldG = [&](size_t o) {
        V4 x;
        x.x = clr[4 * o];
        x.y = clr[4 * o + 1];
        x.z = clr[4 * o + 2];
        x.w = clr[4 * o + 3];
        return x;
    };
    const size_t catOff = (size_t)cat * nPP + p;   // element of slot s: s*K*nPP + catOff
    const size_t slotStride = (size_t)K * nPP;

    // clean children read by this stream's ops, loaded into its slot file
    const int *pre = a.wpre;
    for (int i = 0; i < a.nPre; i++) {
        if (__builtin_amdgcn_readfirstlane(pre[3 * i + 2]) != stream) continue;
        const size_t o = (size_t)__builtin_amdgcn_readfirstlane(pre[3 * i]) * slotStride + catOff;
        const V4 x = ldG(o);
        const int ex = SCALE ? gScl[o] : 0;
        const int l = __builtin_amdgcn_readfirstlane(pre[3 * i + 1]);
        stS(l, x);
        if (SCALE) sexp[l * 64] = ex;
    }

    // Program records and P^T.  PV = 1: records and the program-order P^T
    // stream ([kid][K][16], k_pgather) come through vector loads from
    // wave-uniform addresses, one op ahead of use (scalar loads share their
    // wait counter with the slot file's LDS traffic).  PV = 0: records and
    // P^T (indexed by the child's node) through scalar loads when used.
    // Both compute the same values; the autotuner picks the faster.
    const int4 *ops4 = reinterpret_cast<const int4 *>(a.wops);
    cint_t *opsS = (cint_t *)(a.wops);
    const int *side = reinterpret_cast<const int *>(a.wkids);
    const T *PPg = reinterpret_cast<const T *>(a.PP);
    cT *PTs = (cT *)(a.PT);
    const int nW = a.nWOps;
    const int nKT = a.nKidTot;
    double contrib = 0.0;
    int bad = 0;
    int kb = 0;
    int4 rn0 = make_int4(0, 0, 0, 0), rn1 = rn0, rn2 = rn0;
    T Pn0[16], Pn1[16];
    if (PV) {
        rn0 = ops4[0];
        rn1 = ops4[1];
        rn2 = ops4[2];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            Pn0[j] = PPg[(size_t)cat * 16 + j];
            Pn1[j] = nKT > 1 ? PPg[((size_t)K + cat) * 16 + j] : T(0);
        }
    }
    // stream 0 runs ops [0, nS0), stream 1 [nS0, nW - 1) (one stream: all
    // but the root); then stream 0 runs the root after a barrier
    const int sBeg = stream == 0 ? 0 : a.nS0;
    const int sEnd = (nS == 1 || stream == 1) ? nW - 1 : a.nS0;
    for (int it = sBeg; it <= sEnd; it++) {
        const bool rootIter = it == sEnd;
        if (rootIter) {
            __syncthreads();   // stream 1's root children are in stream 0's slot file
            if (stream != 0) {
                __syncthreads();   // joins the category combine of the root
                break;
            }
        }
        const int i = rootIter ? nW - 1 : it;
        int nk, opDst, opLds, opFlags, kidStart;
        int4 r1, r2;
        T P0[16], P1[16];
        int kbC = kb;
        if (PV) {
            const int4 r0 = rn0;
            r1 = rn1;
            r2 = rn2;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                P0[j] = Pn0[j];
                P1[j] = Pn1[j];
            }
            // record: [dst, ldsDst, flags, nKids] [kidStart, kind0..2] [idx0..2, pnode0] ...
            nk = __builtin_amdgcn_readfirstlane(r0.w);
            kb += nk;
            if (i + 1 < nW) {
                rn0 = ops4[4 * (i + 1)];
                rn1 = ops4[4 * (i + 1) + 1];
                rn2 = ops4[4 * (i + 1) + 2];
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    Pn0[j] = PPg[((size_t)kb * K + cat) * 16 + j];
                    Pn1[j] = kb + 1 < nKT ? PPg[((size_t)(kb + 1) * K + cat) * 16 + j] : T(0);
                }
            }
            opDst = __builtin_amdgcn_readfirstlane(r0.x);
            opLds = __builtin_amdgcn_readfirstlane(r0.y);
            opFlags = __builtin_amdgcn_readfirstlane(r0.z);
            kidStart = __builtin_amdgcn_readfirstlane(r1.x);
        } else {
            cint_t *op = opsS + 16 * i;
            nk = op[3];
            opDst = op[0];
            opLds = op[1];
            opFlags = op[2];
            kidStart = op[4];
        }
        T acc[4];
        int e = 0;
        // one child: its contribution v, folded into acc (first child sets it)
        auto child = [&](int c) {
            int kind, idx;
            T Pl[16];
            if (PV) {
                if (c == 0) {
                    kind = __builtin_amdgcn_readfirstlane(r1.y);
                    idx = __builtin_amdgcn_readfirstlane(r2.x);
                } else if (c == 1) {
                    kind = __builtin_amdgcn_readfirstlane(r1.z);
                    idx = __builtin_amdgcn_readfirstlane(r2.y);
                } else if (c == 2) {
                    kind = __builtin_amdgcn_readfirstlane(r1.w);
                    idx = __builtin_amdgcn_readfirstlane(r2.z);
                } else {
                    const int *kr = side + 4 * (kidStart + c - 3);
                    kind = __builtin_amdgcn_readfirstlane(kr[0]);
                    idx = __builtin_amdgcn_readfirstlane(kr[1]);
                }
                if (c == 0) {
#pragma unroll
                    for (int j = 0; j < 16; j++) Pl[j] = P0[j];
                } else if (c == 1) {
#pragma unroll
                    for (int j = 0; j < 16; j++) Pl[j] = P1[j];
                } else {
#pragma unroll
                    for (int j = 0; j < 16; j++) Pl[j] = PPg[((size_t)(kbC + c) * K + cat) * 16 + j];
                }
            } else {
                cint_t *op = opsS + 16 * i;
                int pn;
                if (c < 3) {
                    kind = op[5 + c];
                    idx = op[8 + c];
                    pn = op[11 + c];
                } else {
                    cint_t *kr = (cint_t *)(side) + 4 * (kidStart + c - 3);
                    kind = kr[0];
                    idx = kr[1];
                    pn = kr[2];
                }
                cT *Pk = PTs + ((size_t)pn * K + cat) * 16;
#pragma unroll
                for (int j = 0; j < 16; j++) Pl[j] = Pk[j];
            }
            T v[4];
            bool skip = false;
            if (kind == WK_TIP) {
                unsigned code;
                if (a.tipLds)
                    code = sTip[idx * 64 + lane];
                else
                    code = gTips[(size_t)idx * nPP + p];
                leafVec4<T>(Pl, code, a.ambMask, v);
                skip = (code == TIP_ALL) && c != 0;
            } else {
                V4 x;
                int ex = 0;
                if (kind == WK_LDS) {
                    x = ldS(idx);
                    if (SCALE) ex = sexp[idx * 64];
                } else {
                    const size_t o = (size_t)idx * slotStride + catOff;
                    x = ldG(o);
                    if (SCALE) ex = gScl[o];
                }
                const T xv[4] = {x.x, x.y, x.z, x.w};
                internalVec<4, T>(Pl, xv, 4, v);
                e += ex;
            }
            if (c == 0) {
#pragma unroll
                for (int s = 0; s < 4; s++) acc[s] = v[s];
            } else if (!skip) {
#pragma unroll
                for (int s = 0; s < 4; s++) acc[s] *= v[s];
            }
        };
        if (nk == 2) {
            // binary op (the common case): child indices known at compile time
            child(0);
            child(1);
        } else {
            for (int c = 0; c < nk; c++) child(c);
        }
        if (SCALE) {
            const T m = fmax(fmax(acc[0], acc[1]), fmax(acc[2], acc[3]));
            if (m < T(SCALE_THRESH) && m > T(0)) {
                int ex;
                frexp(m, &ex);
                const int E = -ex;
#pragma unroll
                for (int s = 0; s < 4; s++) acc[s] = ldexp(acc[s], E);
                e += E;
            }
        }
        if (!(opFlags & OP_ROOT)) {
            V4 r;
            r.x = acc[0];
            r.y = acc[1];
            r.z = acc[2];
            r.w = acc[3];
            const size_t o = (size_t)opDst * slotStride + catOff;
            clw[o] = r;
            if (SCALE) a.scl[o] = e;
            const int ld = opLds;
            if (ld >= 0) {
                stS(ld, r);
                if (SCALE) sexp[ld * 64] = e;
            } else if (ld <= -2) {   // a root child of stream 1: junction slot in stream 0's file
                const int j = -ld - 2;
                slot0T[j * 256] = r.x;
                slot0T[j * 256 + 1] = r.y;
                slot0T[j * 256 + 2] = r.z;
                slot0T[j * 256 + 3] = r.w;
                if (SCALE) sexp0[j * 64] = e;
            }
            continue;
        }
        // ---- root: category-summed site likelihood (Pf/p4_tree.c:1029-1378)
        const double *pi = a.rootPi;
        double Lk = 0.0;
        if (a.rootLeafRow >= 0) {
            const unsigned code = a.tips[(size_t)a.rootLeafRow * nPP + p];
            uint64_t m;
            if (code < 4u)
                m = 1ull << code;
            else if (code == TIP_ALL)
                m = ~0ull;
            else
                m = a.ambMask[code - 4];
#pragma unroll
            for (int s = 0; s < 4; s++)
                if ((m >> s) & 1ull) Lk += pi[s] * (double)acc[s];
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++) Lk += pi[s] * (double)acc[s];
        }
        rootL[cat][lane] = Lk;
        rootE[cat][lane] = e;
        __syncthreads();
        if (wv == 0) {
            int mn = rootE[0][lane];
            for (int k = 1; k < K; k++) mn = min(mn, rootE[k][lane]);
            double L = 0.0;
            for (int k = 0; k < K; k++) L += SCALE ? ldexp(rootL[k][lane], mn - rootE[k][lane]) : rootL[k][lane];
            siteLogLike<SCALE>(a, L, mn, p, K, pi, contrib, bad);
        }
    }
    if (wv == 0) finishSiteSumWave(a, contrib, bad, tile, nTiles);
}

template <typename T, bool SCALE, bool PV>
__global__ __launch_bounds__(64 * WAVE_MAX_K) void k_pruneWave(PruneArgs a)
{
    pruneWaveBody<T, SCALE, PV>(a, blockIdx.x, gridDim.x);
}

template <typename T, bool SCALE>
__global__ __launch_bounds__(64 * WAVE_MAX_K) void k_pruneWaveBatch(const PruneArgs *args, const int *blockItem)
{
    const int item = blockItem[2 * blockIdx.x];
    const int local = blockItem[2 * blockIdx.x + 1];
    const PruneArgs a = args[item];
    pruneWaveBody<T, SCALE, false>(a, local, a.nPatPad / 64);
}

// ------------------------------------------- lean DNA kernel (default)
// The wave kernel's lane mapping (lane = pattern, wavefront = gamma
// category, workgroup = the K category waves of a 64-pattern tile) and
// program, with three changes that cut the instructions per node op:
//  * binary ops whose children are tips or LDS slots run from LeanOp
//    records of precomputed byte offsets (one 32-byte scalar load);
//  * the P^T blocks of the program's tip children are staged in LDS at the
//    start, so a tip child is one 16-byte LDS read of row `code` instead of
//    a select tree over P held in scalar registers (16 moves + 12 selects);
//  * an internal child's P^T comes by scalar loads (one 64-byte load per
//    child in FP32, two in FP64) and feeds the 16 FMAs directly as scalar
//    operands.
// The root, multifurcations and children in global memory take the generic
// path of the wave kernel.
template <typename T, bool SCALE>
__device__ __forceinline__ void pruneLeanBody(const PruneArgs &a, int tile, int nTiles)
{
    typedef typename Vec4T<T>::type V4;
    typedef typename ConstT<T>::type cT;
    typedef __attribute__((address_space(3))) T lT;
    typedef __attribute__((address_space(3))) int lint;
    typedef __attribute__((address_space(3))) const unsigned char lu8;
    typedef __attribute__((address_space(3))) char lchar;
    typedef __attribute__((address_space(1))) const T gT;
    typedef __attribute__((address_space(1))) const int gint;
    typedef __attribute__((address_space(1))) const unsigned char gu8;
    typedef __attribute__((address_space(4))) const char cchar;
    // op records: 32 bytes in FP32, padded to 64 in FP64 (same fields)
    typedef int LeanRecV __attribute__((ext_vector_type(sizeof(T) == 4 ? 8 : 16)));
    typedef const __attribute__((address_space(4))) LeanRecV cLeanRecV;
    extern __shared__ __align__(16) unsigned char smemW[];
    __shared__ double rootL[WAVE_MAX_K][64];
    __shared__ int rootE[WAVE_MAX_K][64];
    __shared__ unsigned char ambL[256];   // state masks of partial-ambiguity codes
    const int K = a.K;
    const int cat = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int p = tile * 64 + lane;
    const size_t nPP = (size_t)a.nPatPad;
    const int nL = a.nLds;
    const int tipBytes = a.tipLds ? ((a.nTax * 64 + 31) & ~31) : 0;   // 32-byte aligned: FP64 rows
    const int blkB = 16 * (int)sizeof(T);                   // one P^T block
    for (int j = threadIdx.x; j < a.nAmb && j < 251; j += blockDim.x) ambL[j] = (unsigned char)(a.ambMask[j] & 15u);
    const int tipPBytes = K * a.nTipKid * blkB;
    lchar *lds = (lchar *)smemW;
    lu8 *tipCodes = (lu8 *)lds + lane;                      // + row * 64
    lchar *tipP = lds + tipBytes + cat * a.nTipKid * blkB;  // + t * blkB + code * 4 * sizeof(T)
    // slot l of this wave: slotV[l * 256 elements], exponent slotE[l * 64]
    lT *slotV = (lT *)(lds + tipBytes + tipPBytes) + ((size_t)cat * nL * 64 + lane) * 4;
    lint *slotE = (lint *)(lds + tipBytes + tipPBytes + (size_t)K * nL * 64 * 4 * sizeof(T)) + (size_t)cat * nL * 64 +
                  lane;
    gu8 *gTips = (gu8 *)(a.tips);
    gint *gScl = (gint *)(a.scl);
    // pull the P^T table and the op records into this XCD's L2 while the
    // tile stages its tips: the op chain's scalar loads then hit L2 (the
    // table was written by other XCDs' workgroups)
    unsigned warm = 0;
    for (int w = threadIdx.x * 128; w < a.ptWarmBytes; w += blockDim.x * 128)
        warm ^= *reinterpret_cast<const unsigned *>(static_cast<const char *>(a.PT) + w);
    for (int w = threadIdx.x * 128; w < a.nWOps * (int)sizeof(LeanRecV); w += blockDim.x * 128)
        warm ^= *reinterpret_cast<const unsigned *>(reinterpret_cast<const char *>(a.pops) + w);
    if (a.tipLds) {
        // the tile's tip codes, [taxon][64 patterns], 4 bytes per thread
        const int words = a.nTax * 16;
        for (int w = threadIdx.x; w < words; w += blockDim.x) {
            const int r = w >> 4, q = w & 15;
            reinterpret_cast<uint32_t *>(smemW)[w] =
                *reinterpret_cast<const uint32_t *>(a.tips + (size_t)r * nPP + (size_t)tile * 64 + q * 4);
        }
        // the P^T blocks of the staged tip children, [cat][t][16], 16 bytes per thread
        const int chunks = K * a.nTipKid * blkB / 16;
        const int perBlk = blkB / 16;
        for (int w = threadIdx.x; w < chunks; w += blockDim.x) {
            const int blk = w / perBlk, part = w - blk * perBlk;
            const int k = blk / a.nTipKid, t = blk - k * a.nTipKid;
            const char *src = (const char *)a.PT + ((size_t)a.tipPn[t] * K + k) * blkB + part * 16;
            *reinterpret_cast<uint4 *>(smemW + tipBytes + (size_t)blk * blkB + part * 16) =
                *reinterpret_cast<const uint4 *>(src);
        }
        __syncthreads();
    }
    gT *clr = (gT *)(a.cl);
    V4 *clw = reinterpret_cast<V4 *>(a.cl);
    auto ldG = [&](size_t o) {
        V4 x;
        x.x = clr[4 * o];
        x.y = clr[4 * o + 1];
        x.z = clr[4 * o + 2];
        x.w = clr[4 * o + 3];
        return x;
    };
    auto ldS = [&](int sl) {
        V4 x;
        x.x = slotV[sl * 256];
        x.y = slotV[sl * 256 + 1];
        x.z = slotV[sl * 256 + 2];
        x.w = slotV[sl * 256 + 3];
        return x;
    };
    auto stS = [&](int sl, const V4 &x) {
        slotV[sl * 256] = x.x;
        slotV[sl * 256 + 1] = x.y;
        slotV[sl * 256 + 2] = x.z;
        slotV[sl * 256 + 3] = x.w;
    };
    const size_t catOff = (size_t)cat * nPP + p;   // element of slot s: s*K*nPP + catOff
    const size_t slotStride = (size_t)K * nPP;
    // a staged tip child: row `code` of its P^T block (Pf/p4_node.c:660-725)
    typedef T tv4 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(3))) const tv4 ltv4;
    auto tipRow = [&](int tOff, unsigned code, T *v) {
        if (code < 4u) {
            const tv4 row = *(ltv4 *)(tipP + tOff + code * 4 * (int)sizeof(T));
            v[0] = row.x;
            v[1] = row.y;
            v[2] = row.z;
            v[3] = row.w;
        } else if (code == TIP_ALL) {
#pragma unroll
            for (int s = 0; s < 4; s++) v[s] = T(1);
        } else {
            const unsigned m = ambL[code - 4];
            const lT *blk = (const lT *)(tipP + tOff);
#pragma unroll
            for (int s = 0; s < 4; s++) v[s] = T(0);
            for (int j = 0; j < 4; j++)
                if ((m >> j) & 1u) {
#pragma unroll
                    for (int s = 0; s < 4; s++) v[s] += blk[j * 4 + s];
                }
        }
    };
    // this category's P^T blocks; a record's P^T byte offsets are added as
    // unsigned values, which the scalar loads take as a register offset
    cchar *PTc = (cchar *)(a.PT) + (size_t)cat * blkB;

    const int *pre = a.wpre;
    for (int i = 0; i < a.nPre; i++) {
        const size_t o = (size_t)__builtin_amdgcn_readfirstlane(pre[3 * i]) * slotStride + catOff;
        const int l = __builtin_amdgcn_readfirstlane(pre[3 * i + 1]);
        stS(l, ldG(o));
        if (SCALE) slotE[l * 64] = gScl[o];
    }
    __asm__ volatile("" ::"v"(warm));   // the warming loads are kept and done here
    // no global load is left in flight past this point or past a generic
    // op (the fast ops issue none), so the fast path never waits on the
    // vector-memory counter, which also counts the partials' stores
    __builtin_amdgcn_s_waitcnt(VMCNT0);

    cLeanRecV *lops = (cLeanRecV *)(a.pops);
    // the program's tip children read P^T rows staged in LDS, or (none
    // staged) scalar P^T through leafVec4's select tree, with the
    // partial-ambiguity masks from LDS (no global load in the fast loop)
    const bool stagedTips = a.nTipKid > 0;
    auto leafS = [&](const T *Pl, unsigned code, T *v) {
        const bool b0 = code & 1u, b1 = code & 2u;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const T lo = b0 ? Pl[4 + s] : Pl[s];
            const T hi = b0 ? Pl[12 + s] : Pl[8 + s];
            v[s] = b1 ? hi : lo;
        }
        if (code >= 4u) {
            if (code == TIP_ALL) {
#pragma unroll
                for (int s = 0; s < 4; s++) v[s] = T(1);
            } else {
                const unsigned m = ambL[code - 4];
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    T x = T(0);
#pragma unroll
                    for (int j = 0; j < 4; j++) x += ((m >> j) & 1u) ? Pl[j * 4 + s] : T(0);
                    v[s] = x;
                }
            }
        }
    };
    cint_t *opsS = (cint_t *)(a.wops);
    cint_t *sideS = (cint_t *)(a.wkids);
    const int nW = a.nWOps;
    double contrib = 0.0;
    int bad = 0;
    // FP32 rescaling by the pattern's largest entry, an exact power of two
    auto rescale = [&](T *acc, int &e) {
        if (SCALE) {
            const T m = fmax(fmax(acc[0], acc[1]), fmax(acc[2], acc[3]));
            if (m < T(SCALE_THRESH) && m > T(0)) {
                int ex;
                frexp(m, &ex);
#pragma unroll
                for (int s = 0; s < 4; s++) acc[s] = ldexp(acc[s], -ex);
                e -= ex;
            }
        }
    };
    // the result to HBM and, when its parent is in the program, to LDS.
    // FP32: one 16-byte store per lane, so a wave's store covers whole
    // lines; non-temporal, so the 190 MB of partials of a full evaluation
    // stream past L2 instead of evicting the P^T table and op records the
    // chain reads from it (het64: 67 us from 73).  FP64 results are two
    // 16-byte stores per lane, each covering half of every line; streamed
    // past L2 those cost twice the kernel's time, so they are cached.
    auto store = [&](const T *acc, int e, size_t dstE, int ld) {
        V4 rv;
        rv.x = acc[0];
        rv.y = acc[1];
        rv.z = acc[2];
        rv.w = acc[3];
        if constexpr (sizeof(T) == 4) {
            const tv4 nv = {acc[0], acc[1], acc[2], acc[3]};
            __builtin_nontemporal_store(nv, reinterpret_cast<tv4 *>(a.cl) + dstE + catOff);
            if (SCALE) __builtin_nontemporal_store(e, &a.scl[dstE + catOff]);
        } else {
            clw[dstE + catOff] = rv;
            if (SCALE) a.scl[dstE + catOff] = e;
        }
        if (ld >= 0) {
            stS(ld, rv);
            if (SCALE) slotE[ld * 64] = e;
        }
    };
    // Fast ops run in an inner loop of their own: no global load is ever in
    // flight there, so the compiler's wait-count analysis never makes a fast
    // op wait on the vector-memory counter, which also counts the partials'
    // stores.  Generic ops (and the root) run between, and wait for their
    // own loads before the fast loop resumes.
    int i = 0;
    while (true) {
        // FP64: the op counter held in a scalar register (het64 100 us from
        // 103-109; in FP32 the same costs 2 us)
        if constexpr (sizeof(T) == 8) i = __builtin_amdgcn_readfirstlane(i);
        LeanRecV rn = lops[i];   // each record is one scalar load, one op ahead
        for (; i < nW; i++) {
            const LeanRecV r = rn;
            const int kind = r[0] & 0xff;
            if (kind == PK_GEN) break;
            if (i + 1 < nW) rn = lops[i + 1];
            T u[4], w[4], acc[4];
            int e = 0;
            if (kind == PK_LL) {
                const V4 x = ldS(r[1]), y = ldS(r[2]);
                if (SCALE) e = slotE[r[3]] + slotE[r[4]];
                cT *P0 = (cT *)(PTc + (unsigned)r[5]);
                cT *P1 = (cT *)(PTc + (unsigned)r[6]);
                T Pa[16], Pb[16];
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    Pa[j] = P0[j];
                    Pb[j] = P1[j];
                }
                const T xv[4] = {x.x, x.y, x.z, x.w}, yv[4] = {y.x, y.y, y.z, y.w};
                internalVec<4, T>(Pa, xv, 4, u);
                internalVec<4, T>(Pb, yv, 4, w);
            } else {
                if (stagedTips) {
                    tipRow(r[3], (unsigned)tipCodes[r[1]], u);
                } else {
                    cT *Pt = (cT *)(PTc + (unsigned)r[5]);
                    T Pl[16];
#pragma unroll
                    for (int j = 0; j < 16; j++) Pl[j] = Pt[j];
                    leafS(Pl, (unsigned)tipCodes[r[1]], u);
                }
                if (kind == PK_TT) {
                    if (stagedTips) {
                        tipRow(r[4], (unsigned)tipCodes[r[2]], w);
                    } else {
                        cT *Pt = (cT *)(PTc + (unsigned)r[6]);
                        T Pl[16];
#pragma unroll
                        for (int j = 0; j < 16; j++) Pl[j] = Pt[j];
                        leafS(Pl, (unsigned)tipCodes[r[2]], w);
                    }
                } else {
                    const V4 y = ldS(r[2]);
                    if (SCALE) e = slotE[r[4]];
                    cT *P1 = (cT *)(PTc + (unsigned)r[6]);
                    T Pb[16];
#pragma unroll
                    for (int j = 0; j < 16; j++) Pb[j] = P1[j];
                    const T yv[4] = {y.x, y.y, y.z, y.w};
                    internalVec<4, T>(Pb, yv, 4, w);
                }
            }
#pragma unroll
            for (int s = 0; s < 4; s++) acc[s] = u[s] * w[s];
            rescale(acc, e);
            store(acc, e, (size_t)(unsigned)r[7] * slotStride, (r[0] >> 8) - 1);
        }
        if (i >= nW) break;
        // ---- a generic op: any children, from anywhere; maybe the root
        cint_t *op = opsS + 16 * i;
        const int nk = op[3];
        const int kidStart = op[4];
        const bool isRoot = (op[2] & OP_ROOT) != 0;
        T acc[4];
        int e = 0;
        for (int c = 0; c < nk; c++) {
            int ck, idx, pn;
            if (c < 3) {
                ck = op[5 + c];
                idx = op[8 + c];
                pn = op[11 + c];
            } else {
                cint_t *kr = sideS + 4 * (kidStart + c - 3);
                ck = kr[0];
                idx = kr[1];
                pn = kr[2];
            }
            cT *Pk = (cT *)(PTc + (size_t)pn * K * blkB);
            T v[4];
            T Pl[16];
#pragma unroll
            for (int j = 0; j < 16; j++) Pl[j] = Pk[j];
            if (ck == WK_TIP) {
                const unsigned code = a.tipLds ? (unsigned)tipCodes[idx * 64] : (unsigned)gTips[(size_t)idx * nPP + p];
                leafVec4<T>(Pl, code, a.ambMask, v);
            } else {
                V4 x;
                int ex = 0;
                if (ck == WK_LDS) {
                    x = ldS(idx);
                    if (SCALE) ex = slotE[idx * 64];
                } else {
                    const size_t o = (size_t)idx * slotStride + catOff;
                    x = ldG(o);
                    if (SCALE) ex = gScl[o];
                }
                const T xv[4] = {x.x, x.y, x.z, x.w};
                internalVec<4, T>(Pl, xv, 4, v);
                e += ex;
            }
            if (c == 0) {
#pragma unroll
                for (int s = 0; s < 4; s++) acc[s] = v[s];
            } else {
#pragma unroll
                for (int s = 0; s < 4; s++) acc[s] *= v[s];
            }
        }
        rescale(acc, e);
        if (!isRoot) {
            store(acc, e, (size_t)op[0] * slotStride, op[1]);
        } else {
            // ---- root: category-summed site likelihood (Pf/p4_tree.c:1029-1378)
            const double *pi = a.rootPi;
            double Lk = 0.0;
            if (a.rootLeafRow >= 0) {
                const unsigned code = a.tips[(size_t)a.rootLeafRow * nPP + p];
                const uint64_t m = code < 4u ? (1ull << code) : code == TIP_ALL ? ~0ull : a.ambMask[code - 4];
#pragma unroll
                for (int s = 0; s < 4; s++)
                    if ((m >> s) & 1ull) Lk += pi[s] * (double)acc[s];
            } else {
#pragma unroll
                for (int s = 0; s < 4; s++) Lk += pi[s] * (double)acc[s];
            }
            rootL[cat][lane] = Lk;
            rootE[cat][lane] = e;
            __syncthreads();
            if (cat == 0) {
                int mn = rootE[0][lane];
                for (int k = 1; k < K; k++) mn = min(mn, rootE[k][lane]);
                double L = 0.0;
                for (int k = 0; k < K; k++) L += SCALE ? ldexp(rootL[k][lane], mn - rootE[k][lane]) : rootL[k][lane];
                siteLogLike<SCALE>(a, L, mn, p, K, pi, contrib, bad);
            }
        }
        __builtin_amdgcn_s_waitcnt(VMCNT0);
        i++;
    }
    if (cat == 0) finishSiteSumWave(a, contrib, bad, tile, nTiles);
}

template <typename T, bool SCALE>
__global__ __launch_bounds__(64 * WAVE_MAX_K) void k_pruneLean(PruneArgs a)
{
    pruneLeanBody<T, SCALE>(a, blockIdx.x, gridDim.x);
}

template <typename T, bool SCALE>
__global__ __launch_bounds__(64 * WAVE_MAX_K) void k_pruneLeanBatch(const PruneArgs *args, const int *blockItem)
{
    const int item = blockItem[2 * blockIdx.x];
    const int local = blockItem[2 * blockIdx.x + 1];
    const PruneArgs a = args[item];
    pruneLeanBody<T, SCALE>(a, local, a.nPatPad / 64);
}

template <typename T>
__global__ __launch_bounds__(256) void k_pgather(const T *PT, const int *kidP, int nKidTot, int K, T *PP)
{
    // PP[j][k][e] = PT[kidP[j]][k][e]: P^T of the program's children in
    // program order
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int per = K * 16;
    if (i >= nKidTot * per) return;
    const int j = i / per, r = i - j * per;
    PP[i] = PT[(size_t)kidP[j] * per + r];
}

hipError_t launchPgather(const void *PT, const int *kidP, int nKidTot, int K, void *PP, int precision,
                         hipStream_t s)
{
    const int n = K * nKidTot * 16;
    const dim3 grid((n + 255) / 256);
    if (precision == FP32)
        hipLaunchKernelGGL(k_pgather<float>, grid, dim3(256), 0, s, (const float *)PT, kidP, nKidTot, K, (float *)PP);
    else
        hipLaunchKernelGGL(k_pgather<double>, grid, dim3(256), 0, s, (const double *)PT, kidP, nKidTot, K,
                           (double *)PP);
    return hipGetLastError();
}

bool pruneWaveLean(int precision) { return g_waveLean[precision & 1] != 0; }
bool pruneWaveVectorP(int precision) { return g_waveVecP[precision & 1] != 0 && !g_waveLean[precision & 1]; }
bool pruneWaveStreams(int precision)
{
    return g_waveStreams[precision & 1] != 0 && !g_waveVecP[precision & 1] && !g_waveLean[precision & 1];
}
int pruneWaveTilePatterns(int precision) { (void)precision; return 64; }

bool pruneUsesWave(int dim, int K, int precision)
{
    return dim == 4 && K >= 1 && K <= WAVE_MAX_K && g_wave[precision & 1];
}

// dynamic LDS of the lean kernel's workgroup (plus 6 KB static for the root
// combine; gfx950 allows 160 KB per workgroup).  A het64-sized tile needs
// ~40 KB in FP32, so four workgroups share a CU.  In FP64 its tip blocks
// alone take 32 KB: within this budget they are not staged (the generic
// path), and a larger budget (two workgroups per CU) measured slower still.
constexpr int LEAN_LDS_BUDGET32 = 56 * 1024;
constexpr int LEAN_LDS_BUDGET64 = 56 * 1024;

static size_t waveTipBytes(int tipLds, int nTax) { return tipLds ? (((size_t)nTax * 64 + 31) & ~(size_t)31) : 0; }

static size_t waveSlotBytes(int K, int precision)
{
    const size_t elt = precision == FP32 ? sizeof(float) : sizeof(double);
    return (size_t)K * 64 * (4 * elt + (precision == FP32 ? sizeof(int) : 0));
}

// LDS of a workgroup: tip codes, the lean kernel's staged tip P^T blocks
// (nTipKid per category; 0 for the wave kernel), the K slot files
size_t pruneWaveLdsBytes(int K, int nLds, int tipLds, int nTax, int nTipKid, int precision)
{
    const size_t elt = precision == FP32 ? sizeof(float) : sizeof(double);
    return waveTipBytes(tipLds, nTax) + (size_t)K * nTipKid * 16 * elt + (size_t)nLds * waveSlotBytes(K, precision);
}

// Preloads of clean children (dirty-path programs) may raise the
// workgroup's LDS to at most this many bytes (0: the slot budget), so that
// short programs keep the occupancy of full ones.
static size_t g_preloadLds = 0;
void setWavePreloadLds(int bytes) { g_preloadLds = bytes > 0 ? (size_t)bytes : 0; }
int pruneWavePreloadSlots(int K, int tipLds, int nTax, int nTipKid, int precision)
{
    const int maxS = pruneWaveMaxSlots(K, tipLds, nTax, nTipKid, precision);
    if (!g_preloadLds) return maxS;
    const size_t fixed = pruneWaveLdsBytes(K, 0, tipLds, nTax, nTipKid, precision);
    if (fixed >= g_preloadLds) return 0;
    return std::min<int>(maxS, (int)((g_preloadLds - fixed) / waveSlotBytes(K, precision)));
}

int pruneWaveMaxSlots(int K, int tipLds, int nTax, int nTipKid, int precision)
{
    const size_t fixed = pruneWaveLdsBytes(K, 0, tipLds, nTax, nTipKid, precision);
    const size_t budget = !pruneWaveLean(precision) ? WAVE_LDS_BUDGET
                          : precision == FP32          ? LEAN_LDS_BUDGET32
                                                       : LEAN_LDS_BUDGET64;
    if (fixed >= budget) return 0;
    const size_t n = (budget - fixed) / waveSlotBytes(K, precision);
    return (int)std::min<size_t>(n, WAVE_MAX_LDS_SLOTS);
}

hipError_t launchPruneWave(const PruneArgs &a, int precision, int nBlocks, size_t lds, hipStream_t s)
{
    // lds <= WAVE_LDS_BUDGET, within the default per-workgroup limit
    const dim3 block(64 * a.K * a.nStream);
    if (pruneWaveLean(precision)) {
        if (precision == FP32)
            hipLaunchKernelGGL((k_pruneLean<float, true>), dim3(nBlocks), block, lds, s, a);
        else
            hipLaunchKernelGGL((k_pruneLean<double, false>), dim3(nBlocks), block, lds, s, a);
        return hipGetLastError();
    }
    // the vector-P variant needs the program-order P stream (a.PP)
    const bool pv = pruneWaveVectorP(precision) && a.nStream == 1 && a.PP != nullptr;
    if (precision == FP32) {
        if (pv)
            hipLaunchKernelGGL((k_pruneWave<float, true, true>), dim3(nBlocks), block, lds, s, a);
        else
            hipLaunchKernelGGL((k_pruneWave<float, true, false>), dim3(nBlocks), block, lds, s, a);
    } else {
        if (pv)
            hipLaunchKernelGGL((k_pruneWave<double, false, true>), dim3(nBlocks), block, lds, s, a);
        else
            hipLaunchKernelGGL((k_pruneWave<double, false, false>), dim3(nBlocks), block, lds, s, a);
    }
    return hipGetLastError();
}

hipError_t launchPruneWaveBatch(const PruneArgs *d, const int *bi, int nBlocks, int K, int nStream, int precision,
                                size_t lds, hipStream_t s)
{
    const dim3 block(64 * K * nStream);
    if (pruneWaveLean(precision)) {
        if (precision == FP32)
            hipLaunchKernelGGL((k_pruneLeanBatch<float, true>), dim3(nBlocks), block, lds, s, d, bi);
        else
            hipLaunchKernelGGL((k_pruneLeanBatch<double, false>), dim3(nBlocks), block, lds, s, d, bi);
        return hipGetLastError();
    }
    if (precision == FP32)
        hipLaunchKernelGGL((k_pruneWaveBatch<float, true>), dim3(nBlocks), block, lds, s, d, bi);
    else
        hipLaunchKernelGGL((k_pruneWaveBatch<double, false>), dim3(nBlocks), block, lds, s, d, bi);
    return hipGetLastError();
}

__global__ __launch_bounds__(64) void k_fillU32(unsigned *p, unsigned v, int n)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launchFillU32(unsigned *p, unsigned v, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fillU32, dim3((n + 63) / 64), dim3(64), 0, s, p, v, n);
    return hipGetLastError();
}

// blockIdx.x: the range; blockIdx.y strides over it.  16-byte moves when
// both addresses and the size allow, else 4-byte words.
__global__ __launch_bounds__(256) void k_copyRanges(const CopyRange *r)
{
    const CopyRange c = r[blockIdx.x];
    const size_t step = (size_t)gridDim.y * 256, first = (size_t)blockIdx.y * 256 + threadIdx.x;
    if ((((uintptr_t)c.src | (uintptr_t)c.dst | (uintptr_t)c.bytes) & 15) == 0) {
        const uint4 *src = static_cast<const uint4 *>(c.src);
        uint4 *dst = static_cast<uint4 *>(c.dst);
        const size_t n = c.bytes / 16;
        for (size_t i = first; i < n; i += step) dst[i] = src[i];
    } else {
        const unsigned *src = static_cast<const unsigned *>(c.src);
        unsigned *dst = static_cast<unsigned *>(c.dst);
        const size_t n = c.bytes / 4;
        for (size_t i = first; i < n; i += step) dst[i] = src[i];
    }
}

hipError_t launchCopyRanges(const CopyRange *r, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copyRanges, dim3(n, 16), dim3(256), 0, s, r);
    return hipGetLastError();
}

}  // namespace p4g
