// Work groups for the nonbonded kernel.
//
// The kernel gives one workgroup G consecutive i-clusters (z-neighbours in
// a grid column) at a time and accumulates every j-force of their cluster
// pairs in an LDS window, one slot per distinct j-cluster, before a single
// global atomic per atom component flushes the window.  This pass builds
// what the kernel needs, from the search's per-cluster lists:
//   * the union of the group's j-clusters (sorted, unique; the group's own
//     clusters take the first G slots) -> unionList,
//   * masked j-entries first within every i-entry (the reference's
//     sort_cj_excl, pairlist.cpp:2002-2067), written back to the list, so
//     unmasked tiles run the kernel's check-free path,
//   * the work list: every i-entry's j-entries as {cj, slot << 16 | mask}
//     (slot = the j-cluster's window slot; the exclusion mask uses bits
//     0..15, bit 4*i + j), padded to a whole number of kTileJ-entry tiles
//     with a dummy entry (cluster ncluster, far-away fillers) whose slot is
//     the group's spare slot G + U, which the kernel never flushes, so the
//     kernel reads whole tiles without bounds checks and no two lanes of a
//     tile share a live slot,
//   * work chunks: i-entries cut into pieces of at most kChunkEntries
//     j-entries, {ci, shift | self << 8 | special << 9, work start, tiles};
//     self marks the piece that carries the cluster's Ewald/RF self energy
//     (first piece of the central entry, kernel_ref_outer.h:286-326),
//     special counts its leading tiles that hold masked entries.
// The union comes from a bitmap over the range of the group's j-cluster
// indices (one bit per cluster, set with LDS atomics; the per-word bit
// counts scanned give each cluster's rank, so a j-entry's slot is one
// popcount away), or from a bitonic sort of the keys when the range is wider
// than the bitmap.
// Like the search it runs as a count pass, a device scan, and a fill pass.
#include <algorithm>
#include <climits>

#include "nbx_internal.h"

namespace nbx
{

namespace
{

constexpr int kGroupThreads = 256;

__device__ void bitonicSort(int* keys, int n)
{
    for (int k = 2; k <= n; k <<= 1)
    {
        for (int j = k >> 1; j > 0; j >>= 1)
        {
            for (int t = threadIdx.x; t < n; t += kGroupThreads)
            {
                const int p = t ^ j;
                if (p > t)
                {
                    const bool up = (t & k) == 0;
                    const int  a  = keys[t];
                    const int  b  = keys[p];
                    if ((a > b) == up)
                    {
                        keys[t] = b;
                        keys[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Exclusive scan of v[0, n) in LDS by the whole workgroup (kGroupThreads
// threads, every one calling it): each thread scans a contiguous run, the
// run totals are scanned across the lanes with shuffles and across the
// wavefronts through waveTot[kGroupThreads / 64].  Returns the total.
__device__ int blockExclusiveScan(int* v, int n, int* waveTot)
{
    const int per  = (n + kGroupThreads - 1) / kGroupThreads;
    const int t0   = min(static_cast<int>(threadIdx.x) * per, n);
    const int t1   = min(t0 + per, n);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int       sum  = 0;
    for (int t = t0; t < t1; t++)
    {
        sum += v[t];
    }
    int incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const int x = __shfl_up(incl, off, 64);
        if (lane >= off)
        {
            incl += x;
        }
    }
    if (lane == 63)
    {
        waveTot[wave] = incl;
    }
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kGroupThreads / 64; w++)
    {
        const int x = waveTot[w];
        base += (w < wave) ? x : 0;
        total += x;
    }
    int run = base + incl - sum;
    for (int t = t0; t < t1; t++)
    {
        const int x = v[t];
        v[t]        = run;
        run += x;
    }
    __syncthreads(); // v is read and waveTot rewritten after the call
    return total;
}

template<bool kFill>
__global__ __launch_bounds__(kGroupThreads) void buildGroupsKernel(GroupArgs a)
{
    __shared__ int sKeys[kGroupMaxJ];
    __shared__ int sFlag[kGroupMaxJ];
    __shared__ int sUnion;
    __shared__ int sChunks;
    __shared__ int sWork;
    __shared__ int sChunkBase[kGroupThreads];
    __shared__ int sWorkBase[kGroupThreads];
    __shared__ int sWaveTot[kGroupThreads / 64];

    const int g  = blockIdx.x;
    const int c0 = g * a.groupSize;
    const int c1 = min(c0 + a.groupSize, a.ncluster);
    const int jb = a.offJ[c0];
    const int n  = a.offJ[c1] - jb;
    const int ib = a.offI[c0];
    const int ni = a.offI[c1] - ib;
    // range of the group's other j-clusters: a bitmap over it (in sKeys)
    // gives the sorted union and each cluster's rank directly; a range
    // wider than the bitmap takes a bitonic sort of the keys instead
    __shared__ int sMin[kGroupThreads / 64], sMax[kGroupThreads / 64];
    int lo = INT_MAX, hi = -1;
    for (int t = threadIdx.x; t < n; t += kGroupThreads)
    {
        const int cj = static_cast<int>(a.jentryIn[jb + t].x);
        if (cj < c0 || cj >= c1)
        {
            lo = min(lo, cj);
            hi = max(hi, cj);
        }
    }
#pragma unroll
    for (int m = 1; m < 64; m *= 2)
    {
        lo = min(lo, __shfl_xor(lo, m, 64));
        hi = max(hi, __shfl_xor(hi, m, 64));
    }
    if ((threadIdx.x & 63) == 0)
    {
        sMin[threadIdx.x >> 6] = lo;
        sMax[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kGroupThreads / 64; w++)
    {
        lo = min(lo, sMin[w]);
        hi = max(hi, sMax[w]);
    }
    const bool bitmap = !a.unionBySort && (hi < lo || hi - lo < kGroupMaxJ * 32);
    unsigned*  bm     = reinterpret_cast<unsigned*>(sKeys);
    const int  nwords = hi < lo ? 0 : ((hi - lo) >> 5) + 1;
    int        p      = 1;
    if (bitmap)
    {
        for (int w = threadIdx.x; w < nwords; w += kGroupThreads)
        {
            bm[w] = 0u;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < n; t += kGroupThreads)
        {
            const int cj = static_cast<int>(a.jentryIn[jb + t].x);
            if (cj < c0 || cj >= c1)
            {
                atomicOr(bm + ((cj - lo) >> 5), 1u << ((cj - lo) & 31));
            }
        }
        __syncthreads();
        // sFlag: per word its set-bit count, scanned below into the rank of
        // the word's first cluster
        for (int w = threadIdx.x; w < nwords; w += kGroupThreads)
        {
            sFlag[w] = __popc(bm[w]);
        }
        __syncthreads();
    }
    else
    {
        while (p < n)
        {
            p <<= 1;
        }
        for (int t = threadIdx.x; t < p; t += kGroupThreads)
        {
            int key = INT_MAX;
            if (t < n)
            {
                const int cj = static_cast<int>(a.jentryIn[jb + t].x);
                key          = (cj >= c0 && cj < c1) ? INT_MAX : cj;
            }
            sKeys[t] = key;
        }
        __syncthreads();
        bitonicSort(sKeys, p);
        // unique: rank of each first occurrence
        for (int t = threadIdx.x; t < p; t += kGroupThreads)
        {
            const int k = sKeys[t];
            sFlag[t]    = (k != INT_MAX && (t == 0 || sKeys[t - 1] != k)) ? 1 : 0;
        }
        __syncthreads();
    }
    // the group's i-entries (at most kGroupThreads; more is flagged as
    // overflow below): their chunk counts and padded work lengths
    {
        const int e = threadIdx.x;
        int       nc = 0, nw = 0;
        if (e < ni)
        {
            const int4 ie = a.ientry[ib + e];
            nc            = (ie.w - ie.z + a.chunkEntries - 1) / a.chunkEntries;
            nw            = (ie.w - ie.z + kTileJ - 1) / kTileJ * kTileJ;
        }
        sChunkBase[e] = nc;
        sWorkBase[e]  = nw;
    }
    // exclusive scans: ranks of the unique keys, and the chunk and padded
    // work offsets of the i-entries
    const int unionTotal = blockExclusiveScan(sFlag, bitmap ? nwords : p, sWaveTot);
    const int chunkTotal = blockExclusiveScan(sChunkBase, kGroupThreads, sWaveTot);
    const int workTotal  = blockExclusiveScan(sWorkBase, kGroupThreads, sWaveTot);
    if (threadIdx.x == 0)
    {
        sUnion  = unionTotal;
        sChunks = chunkTotal;
        sWork   = workTotal;
    }
    __syncthreads();
    const int U      = sUnion;
    // union clusters that get a window slot (the rest overflow)
    const int Uw     = kFill ? min(U, a.windowCap - 1 - a.groupSize) : U;
    const int nchunk = sChunks;
    if (!kFill || a.upperBound)
    {
        if (threadIdx.x == 0)
        {
            a.unionCount[g] = U;
            a.chunkCount[g] = nchunk;
            a.workCount[g]  = sWork;
            a.overflow[g]   = (ni > kGroupThreads) ? 1 : 0;
        }
        if (!kFill || ni > kGroupThreads)
        {
            return;
        }
    }
    // where the group's union, chunks and work go: the scanned offsets, or
    // (upperBound) bounds that need no scan of other groups: the union has
    // at most n entries, the chunks (one per i-entry) at most ni, the padded
    // work at most n + (kTileJ - 1) ni, so bases from the cluster offsets
    // never overlap; groupCompactKernel moves them together afterwards
    const int uoff = a.upperBound ? jb : a.unionOff[g];
    if (bitmap)
    {
        for (int w = threadIdx.x; w < nwords; w += kGroupThreads)
        {
            int r = sFlag[w];
            for (unsigned bits = bm[w]; bits; bits &= bits - 1u)
            {
                a.unionList[uoff + r++] = lo + 32 * w + (__ffs(bits) - 1);
            }
        }
    }
    else
    {
        for (int t = threadIdx.x; t < p; t += kGroupThreads)
        {
            const int k = sKeys[t];
            if (k != INT_MAX && (t == 0 || sKeys[t - 1] != k))
            {
                a.unionList[uoff + sFlag[t]] = k;
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < U; t += kGroupThreads)
        {
            sKeys[t] = a.unionList[uoff + t];
        }
    }
    __syncthreads();
    const int coff = a.upperBound ? ib : a.chunkOff[g];
    if (threadIdx.x == 0)
    {
        a.groups[g]    = make_int4(coff, coff + nchunk, uoff, Uw);
        a.groupInfo[g] = make_int4(c0, (c1 - c0) | (a.groupSize << 16), 0, g);
    }
    // per i-entry: masked-first partition, slots, work list, chunks (one
    // wave per entry)
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int woff = a.upperBound ? jb + (kTileJ - 1) * ib : a.workOff[g];
    for (int e = wave; e < ni; e += kGroupThreads / 64)
    {
        const int4 ie      = a.ientry[ib + e];
        const int  ci      = ie.x;
        const int  shift   = ie.y;
        const int  js      = ie.z;
        const int  je      = ie.w;
        const int  wb      = woff + sWorkBase[e];
        int        written = 0;
        int        nmasked = 0;
        for (int pass = 0; pass < 2; pass++)
        {
            for (int t0 = js; t0 < je; t0 += 64)
            {
                const int  t    = t0 + lane;
                uint2      ent  = make_uint2(0u, 0u);
                bool       take = false;
                if (t < je)
                {
                    ent               = a.jentryIn[t];
                    const bool masked = (ent.y & 0xffffu) != 0xffffu;
                    take              = (pass == 0) ? masked : !masked;
                }
                const unsigned long long ballot = __ballot(take);
                const int before = __popcll(ballot & ((1ull << lane) - 1ull));
                if (take)
                {
                    const int cj = static_cast<int>(ent.x);
                    int       slot;
                    if (cj >= c0 && cj < c1)
                    {
                        slot = cj - c0;
                    }
                    else
                    {
                        // the cluster's rank in the sorted union
                        int l;
                        if (bitmap)
                        {
                            const int d = cj - lo;
                            l           = sFlag[d >> 5] + __popc(bm[d >> 5] & ((1u << (d & 31)) - 1u));
                        }
                        else
                        {
                            l     = 0;
                            int r = U;
                            while (l < r)
                            {
                                const int m = (l + r) >> 1;
                                if (sKeys[m] < cj)
                                {
                                    l = m + 1;
                                }
                                else
                                {
                                    r = m;
                                }
                            }
                        }
                        slot = l < Uw ? a.groupSize + l : static_cast<int>(kSlotOverflow);
                    }
                    a.jentryOut[js + written + before] = ent;
                    a.work[wb + written + before] =
                            make_uint2(ent.x, (static_cast<unsigned>(slot) << 16) | (ent.y & 0xffffu));
                }
                written += __popcll(ballot);
            }
            if (pass == 0)
            {
                nmasked = written;
            }
        }
        // pad to whole tiles with the dummy cluster, in the spare slot
        const int n    = je - js;
        const int npad = (n + kTileJ - 1) / kTileJ * kTileJ;
        if (n + lane < npad)
        {
            a.work[wb + n + lane] = make_uint2(static_cast<unsigned>(a.ncluster),
                                               (static_cast<unsigned>(a.groupSize + Uw) << 16) | 0xffffu);
        }
        if (lane == 0)
        {
            const int base = coff + sChunkBase[e];
            int       k    = 0;
            for (int s = 0; s < n; s += a.chunkEntries, k++)
            {
                const int self  = (shift == kCentral && k == 0) ? 1 : 0;
                const int tiles = (min(a.chunkEntries, n - s) + kTileJ - 1) / kTileJ;
                // tiles of this chunk holding masked entries (they lead the list)
                const int left    = max(0, nmasked - s);
                const int special = min((left + kTileJ - 1) / kTileJ, tiles);
                a.chunks[base + k] = make_int4(ci, shift | (self << 8) | (special << 9), wb + s, tiles);
            }
        }
    }
}

// The layout written at the upper-bound bases moved to its scanned
// offsets (one workgroup per group): union, work entries and chunks, the
// chunks' work starts and the group record rebased.  With a window of
// a.windowCap slots, union clusters beyond it take kSlotOverflow and the
// padding entries the spare slot after the last windowed one.
__global__ __launch_bounds__(kGroupThreads) void groupCompactKernel(GroupArgs a, const int* __restrict__ unionTmp,
                                                                    const uint2* __restrict__ workTmp,
                                                                    const int4* __restrict__ chunksTmp)
{
    const int g  = blockIdx.x;
    const int c0 = g * a.groupSize;
    const int jb = a.offJ[c0];
    const int ib = a.offI[c0];
    const int wb = jb + (kTileJ - 1) * ib;
    const int U = a.unionCount[g], nch = a.chunkCount[g], nw = a.workCount[g];
    const int uo = a.unionOff[g], co = a.chunkOff[g], wo = a.workOff[g];
    const int Uw = min(U, a.windowCap - 1 - a.groupSize);
    const unsigned firstOver = static_cast<unsigned>(a.groupSize + Uw);
    for (int t = threadIdx.x; t < U; t += kGroupThreads)
    {
        a.unionList[uo + t] = unionTmp[jb + t];
    }
    for (int t = threadIdx.x; t < nw; t += kGroupThreads)
    {
        uint2          e    = workTmp[wb + t];
        const unsigned slot = e.y >> 16;
        if (static_cast<int>(e.x) == a.ncluster)
        {
            e.y = (firstOver << 16) | (e.y & 0xffffu);
        }
        else if (slot >= firstOver && slot != kSlotOverflow)
        {
            e.y = (kSlotOverflow << 16) | (e.y & 0xffffu);
        }
        a.work[wo + t] = e;
    }
    for (int t = threadIdx.x; t < nch; t += kGroupThreads)
    {
        int4 c         = chunksTmp[ib + t];
        c.z            = wo + (c.z - wb);
        a.chunks[co + t] = c;
    }
    if (threadIdx.x == 0)
    {
        a.groups[g] = make_int4(co, co + nch, uo, Uw);
    }
}

// Per-workgroup maxima for the host's buffer sizing, reduced by the host
// (no zeroed accumulator, no atomics): workgroup b writes the maxima of the
// a.nv arrays, then with a.offJ given those of the j-entry spans
// offJ[min(c + G, n)] - offJ[c] of the groups of every candidate size G, to
// partial[b * stride, (b + 1) * stride) with stride = partialMaxStride.
__global__ __launch_bounds__(kPartialMaxThreads) void partialMaxKernel(PartialMaxArgs a)
{
    constexpr int kVals = 2 + kGroupSizeCount;
    int           m[kVals] = {};
    for (int c = blockIdx.x * kPartialMaxThreads + threadIdx.x; c < a.n; c += gridDim.x * kPartialMaxThreads)
    {
#pragma unroll
        for (int k = 0; k < 2; k++)
        {
            if (k < a.nv)
            {
                m[k] = max(m[k], a.v[k][c]);
            }
        }
        if (a.offJ)
        {
            const int o = a.offJ[c];
#pragma unroll
            for (int k = 0; k < kGroupSizeCount; k++)
            {
                const int G = kGroupSizeMax >> k;
                if (c % G == 0)
                {
                    m[2 + k] = max(m[2 + k], a.offJ[min(c + G, a.n)] - o);
                }
            }
        }
    }
    __shared__ int sWave[kPartialMaxThreads / 64][kVals];
#pragma unroll
    for (int k = 0; k < kVals; k++)
    {
        int v = m[k];
#pragma unroll
        for (int o = 1; o < 64; o *= 2)
        {
            v = max(v, __shfl_xor(v, o, 64));
        }
        if ((threadIdx.x & 63) == 0)
        {
            sWave[threadIdx.x >> 6][k] = v;
        }
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k < kVals && (k < a.nv || (k >= 2 && a.offJ)))
    {
        int v = 0;
#pragma unroll
        for (int w = 0; w < kPartialMaxThreads / 64; w++)
        {
            v = max(v, sWave[w][k]);
        }
        const int slot = k < 2 ? k : a.nv + (k - 2);
        a.partial[blockIdx.x * partialMaxStride(a.nv, a.offJ != nullptr) + slot] = v;
    }
}

// Occurrences of every cluster in the groups' windows: its own slot in its
// group and its slots in other groups' windowed unions.  One block per group.
template<bool kFill>
__global__ void occurrenceKernel(int ngroups, const int4* __restrict__ groups, const int4* __restrict__ groupInfo,
                                 const int* __restrict__ unionList, int* __restrict__ count,
                                 const int* __restrict__ occOff, int2* __restrict__ occ, int ncluster)
{
    const int  g    = blockIdx.x;
    if (kFill && g == 0 && threadIdx.x == 0)
    {
        count[ncluster] = 0; // the sort's long-list counter
    }
    const int4 grp  = groups[g];
    const int4 gi   = groupInfo[g];
    const int  nOwn = gi.y & 0xffff;
    const int  G    = gi.y >> 16;
    for (int t = threadIdx.x; t < nOwn + grp.w; t += blockDim.x)
    {
        const int c = t < nOwn ? gi.x + t : unionList[grp.z + t - nOwn];
        const int k = atomicAdd(&count[c], 1);
        if (kFill)
        {
            occ[occOff[c] + k] = make_int2(g, t < nOwn ? t : G + (t - nOwn));
        }
    }
}

// Each cluster's occurrences in (group, slot) order, so that the force
// gather adds them in a fixed order; a (group, slot) pair is one 64-bit key,
// and a cluster occurs at most once per group, so the keys are distinct.
// A wavefront per cluster sorts up to 256 entries in registers (R = 1, 2 or
// 4 keys per lane, key e = 64 r + lane in register r) with a bitonic
// network: partners 64 or more apart are in the lane's own registers, nearer
// ones in another lane (shuffles).  Longer lists (not expected at supported
// cut-offs) are listed and sorted afterwards by a workgroup each, in LDS.
__device__ __forceinline__ unsigned long long shflXorU64(unsigned long long v, int j)
{
    const unsigned lo = __shfl_xor(static_cast<unsigned>(v), j, 64);
    const unsigned hi = __shfl_xor(static_cast<unsigned>(v >> 32), j, 64);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

template<int R>
__device__ void waveBitonicSort(unsigned long long (&k)[R], int lane)
{
#pragma unroll
    for (int s = 2; s <= 64 * R; s <<= 1)
    {
#pragma unroll
        for (int j = s >> 1; j > 0; j >>= 1)
        {
            if (j >= 64)
            {
                const int jr = j / 64;
#pragma unroll
                for (int r = 0; r < R; r++)
                {
                    const int p = r ^ jr;
                    if (p > r)
                    {
                        // r holds the lower position: the smaller key when ascending
                        const bool               asc = ((r * 64 + lane) & s) == 0;
                        const unsigned long long x = k[r], y = k[p];
                        if ((x > y) == asc)
                        {
                            k[r] = y;
                            k[p] = x;
                        }
                    }
                }
            }
            else
            {
#pragma unroll
                for (int r = 0; r < R; r++)
                {
                    const unsigned long long o     = shflXorU64(k[r], j);
                    const bool               asc   = ((r * 64 + lane) & s) == 0;
                    const bool               lower = (lane & j) == 0;
                    // the lower position keeps the smaller key when ascending,
                    // the larger when descending; the upper one the other
                    const bool take = (lower == asc) ? (o < k[r]) : (o > k[r]);
                    if (take)
                    {
                        k[r] = o;
                    }
                }
            }
        }
    }
}

template<int R>
__device__ void sortOccurrencesWave(int2* __restrict__ occ, int n, int lane)
{
    unsigned long long k[R];
#pragma unroll
    for (int r = 0; r < R; r++)
    {
        const int e = r * 64 + lane;
        k[r]        = ~0ull; // past the end: sorts last
        if (e < n)
        {
            const int2 v = occ[e];
            k[r]         = (static_cast<unsigned long long>(static_cast<unsigned>(v.x)) << 32)
                   | static_cast<unsigned>(v.y);
        }
    }
    waveBitonicSort<R>(k, lane);
#pragma unroll
    for (int r = 0; r < R; r++)
    {
        const int e = r * 64 + lane;
        if (e < n)
        {
            occ[e] = make_int2(static_cast<int>(k[r] >> 32), static_cast<int>(k[r] & 0xffffffffu));
        }
    }
}

constexpr int kSortWaveMax = 256;
constexpr int kSortMax     = 1024;
constexpr int kSortLongBlocks = 64;

// A wavefront per cluster; lists longer than kSortWaveMax go to
// longList[*longCount] (longCount zeroed by the occurrence fill).
__global__ __launch_bounds__(256) void occurrenceSortKernel(int ncluster, const int* __restrict__ occOff,
                                                            int2* __restrict__ occ, int* __restrict__ longList,
                                                            int* __restrict__ longCount)
{
    const int lane = threadIdx.x & 63;
    const int c    = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= ncluster)
    {
        return;
    }
    const int b = occOff[c], n = occOff[c + 1] - b;
    if (n <= 1)
    {
        return;
    }
    if (n <= 64)
    {
        sortOccurrencesWave<1>(occ + b, n, lane);
    }
    else if (n <= 128)
    {
        sortOccurrencesWave<2>(occ + b, n, lane);
    }
    else if (n <= kSortWaveMax)
    {
        sortOccurrencesWave<4>(occ + b, n, lane);
    }
    else if (lane == 0)
    {
        longList[atomicAdd(longCount, 1)] = c;
    }
}

// The listed long lists, a workgroup each (every thread of a workgroup runs
// the same iterations): up to kSortMax entries with a bitonic network in
// LDS, longer ones by one thread.
__global__ __launch_bounds__(64) void occurrenceSortLongKernel(const int* __restrict__ occOff, int2* __restrict__ occ,
                                                               const int* __restrict__ longList,
                                                               const int* __restrict__ longCount)
{
    __shared__ unsigned long long key[kSortMax];
    const int numLong = *longCount;
    for (int t = blockIdx.x; t < numLong; t += gridDim.x)
    {
        const int c = longList[t];
        const int b = occOff[c], n = occOff[c + 1] - b;
        if (n > kSortMax)
        {
            if (threadIdx.x == 0)
            {
                for (int i = b + 1; i < b + n; i++)
                {
                    const int2 v = occ[i];
                    int        j = i - 1;
                    while (j >= b && (occ[j].x > v.x || (occ[j].x == v.x && occ[j].y > v.y)))
                    {
                        occ[j + 1] = occ[j];
                        j--;
                    }
                    occ[j + 1] = v;
                }
            }
            continue;
        }
        int m = 2 * kSortWaveMax;
        while (m < n)
        {
            m <<= 1;
        }
        for (int i = threadIdx.x; i < m; i += blockDim.x)
        {
            if (i < n)
            {
                const int2 v = occ[b + i];
                key[i] = (static_cast<unsigned long long>(static_cast<unsigned>(v.x)) << 32)
                         | static_cast<unsigned>(v.y);
            }
            else
            {
                key[i] = ~0ull;
            }
        }
        __syncthreads();
        for (int k = 2; k <= m; k <<= 1)
        {
            for (int j = k >> 1; j > 0; j >>= 1)
            {
                for (int i = threadIdx.x; i < m; i += blockDim.x)
                {
                    const int l = i ^ j;
                    if (l > i)
                    {
                        const unsigned long long x = key[i], y = key[l];
                        if ((x > y) == ((i & k) == 0))
                        {
                            key[i] = y;
                            key[l] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int i = threadIdx.x; i < n; i += blockDim.x)
        {
            occ[b + i] = make_int2(static_cast<int>(key[i] >> 32), static_cast<int>(key[i] & 0xffffffffu));
        }
        __syncthreads(); // key is rewritten by the next list
    }
}

// Per cluster, the staging positions (float4 index of its window slot) of
// all its occurrences, one per wavefront segment of the occurrence's group,
// in occurrence order.
template<bool kFill>
__global__ __launch_bounds__(256) void gatherListKernel(int ncluster, const int* __restrict__ occOff,
                                                        const int2* __restrict__ occ, const int4* __restrict__ segInfo,
                                                        int slotF4, int* __restrict__ count,
                                                        const int* __restrict__ gposOff, int* __restrict__ gpos)
{
    // one wavefront per cluster: lane l takes occurrences l, l + 64, ...;
    // each occurrence stands for one staged copy per wavefront that ran its
    // group, placed in occurrence order after a wave prefix sum
    const int lane = threadIdx.x & 63;
    const int c    = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= ncluster)
    {
        return;
    }
    const int o1 = occOff[c + 1];
    int       n  = kFill ? gposOff[c] : 0;
    for (int o0 = occOff[c]; o0 < o1; o0 += 64)
    {
        const int o  = o0 + lane;
        int2      e  = make_int2(0, 0);
        int4      si = make_int4(0, 0, 0, 0);
        if (o < o1)
        {
            e  = occ[o];
            si = segInfo[e.x];
        }
        int incl = si.z;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1)
        {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off)
            {
                incl += t;
            }
        }
        if (kFill)
        {
            int pos = n + incl - si.z;
            for (int w = 0; w < si.z; w++, pos++)
            {
                gpos[pos] = si.x + w * si.w + e.y * slotF4;
            }
        }
        n += __shfl(incl, 63, 64);
    }
    if (!kFill && lane == 0)
    {
        count[c] = n;
    }
}

// the last wavefront whose run starts at or before chunk x
__device__ int waveOfChunk(int numWaves, const int2* __restrict__ waveStart, int x)
{
    int lo = 0, hi = numWaves - 1;
    while (lo < hi)
    {
        const int m = (lo + hi + 1) / 2;
        if (waveStart[m].x <= x)
        {
            lo = m;
        }
        else
        {
            hi = m - 1;
        }
    }
    return lo;
}

// Per group: the wavefronts that run its chunks (first, count) and its
// window size in float4 (slotF4 per cluster slot: 3 for the forces, 4 for
// the adjoint's {dL/dx, dL/dq}); size = their product.
__global__ void segmentKernel(int ngroups, const int4* __restrict__ groups, const int4* __restrict__ groupInfo,
                              int numWaves, const int2* __restrict__ waveStart, int slotF4, int* __restrict__ size,
                              int4* __restrict__ segInfo)
{
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= ngroups)
    {
        return;
    }
    const int4 grp = groups[h];
    const int  win = ((groupInfo[h].y >> 16) + grp.w) * slotF4;
    if (grp.x >= grp.y)
    {
        size[h]    = 0;
        segInfo[h] = make_int4(0, 0, 0, win);
        return;
    }
    const int wf = waveOfChunk(numWaves, waveStart, grp.x);
    const int wl = waveOfChunk(numWaves, waveStart, grp.y - 1);
    size[h]      = (wl - wf + 1) * win;
    segInfo[h]   = make_int4(0, wf, wl - wf + 1, win);
}

__global__ void segmentBaseKernel(int ngroups, const int* __restrict__ base, int4* __restrict__ segInfo)
{
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h < ngroups)
    {
        segInfo[h].x = base[h];
    }
}

__global__ void chunkShiftKernel(int nchunk, const int4* __restrict__ chunks, const float4* __restrict__ shiftVec,
                                 float4* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nchunk)
    {
        out[k] = shiftVec[chunks[k].y & 0xff];
    }
}

// Measured cost of a chunk in the pair kernel, in 1/32 tile: 32 per tile,
// plus 1.5 per tile for each i-atom whose type has LJ (the kernel skips the
// LJ terms of the others), plus 27 per chunk (its set-up and i-force
// reduction).  Fitted to the run times of individual wavefronts on the
// 288,000-atom water box, where it leaves 2.3 us of 205 us unexplained.
__global__ void chunkCostKernel(int nchunk, const int4* __restrict__ chunks, const int* __restrict__ type,
                                const int* __restrict__ typeLJ, int* __restrict__ cost)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < nchunk)
    {
        const int4 cd  = chunks[m];
        int        nlj = 0;
#pragma unroll
        for (int i = 0; i < kClusterSize; i++)
        {
            nlj += typeLJ[type[cd.x * kClusterSize + i]];
        }
        cost[m] = cd.w * (kCostTile + (3 * nlj) / 2) + kCostChunk;
    }
}

// Share of the work given to each resident generation of workgroups.  The
// dispatcher places workgroup b on CU b mod numCU, so with two workgroups
// per CU the first numCU workgroups share every SIMD with the second ones;
// the SIMD's arbiter issues for the older wavefront first, and at equal
// work the second generation measured 9% slower, so the first generation
// takes (1 + skew) and the second (1 - skew) of the average.  Other numbers
// of generations are not weighted (not measured).  Returns the cumulative
// weight of wavefronts [0, w) in 1/1000.
__device__ long long generationWeight(int w, int numWaves, int genWaves)
{
    if (genWaves <= 0 || numWaves != 2 * genWaves)
    {
        return 1000LL * w;
    }
    const long long first = 1000 + kGenerationSkewMilli, second = 1000 - kGenerationSkewMilli;
    return w <= genWaves ? first * w : first * genWaves + second * (w - genWaves);
}

// Smallest m in [0, n) with key(m) >= target (n if none) for a
// non-decreasing key, from an LDS sample s[i] = key(i * stride), i < ns =
// ceil(n / stride): a search of the sample, then of the at most stride
// entries it leaves, so the dependent global loads number log2(stride).
template<typename T, typename Key>
__device__ int sampledLowerBound(const T* s, int ns, int stride, int n, T target, Key key)
{
    int a = 0, b = ns;
    while (a < b)
    {
        const int m = (a + b) / 2;
        if (s[m] < target)
        {
            a = m + 1;
        }
        else
        {
            b = m;
        }
    }
    // key((a - 1) * stride) < target <= key(a * stride) (or a == ns)
    int lo = a > 0 ? (a - 1) * stride + 1 : 0;
    int hi = min(a * stride, n);
    while (lo < hi)
    {
        const int m = (lo + hi) / 2;
        if (key(m) < target)
        {
            lo = m + 1;
        }
        else
        {
            hi = m;
        }
    }
    return lo;
}

constexpr int kSchedThreads    = 1024;
constexpr int kSchedSample     = 4096; // LDS sample entries of each key
constexpr int kSchedCostStride = 64;   // least sample stride over the chunks
constexpr int kSchedGroupStride = 16;  // and over the groups

// One workgroup.  Wavefront w's target start is the first chunk whose cost
// offset reaches the wavefront's cumulative share of the total; equal
// targets are merged, so the runs are non-empty and in order, and the
// wavefronts left over (more wavefronts than distinct starts) get empty
// runs at the end.  A group's chunks are then run by consecutive
// wavefronts, each of which writes its staging segment.  The cost offset of
// chunk m is costOff[m] (total costTotal) when given, else its work offset
// plus chunkWeight per earlier chunk.  The searches (a start's chunk, the
// group holding it) go through LDS samples of the cost offsets and of the
// groups' end chunks.
__global__ __launch_bounds__(kSchedThreads) void waveScheduleKernel(int numWaves, int nchunk,
                                                                   const int4* __restrict__ chunks, int nwork,
                                                                   int chunkWeight, int ngroups,
                                                                   const int4* __restrict__ groups,
                                                                   const int* __restrict__ costOff,
                                                                   const int* __restrict__ costTotal, int genWaves,
                                                                   int2* __restrict__ out)
{
    __shared__ long long sCost[kSchedSample];
    __shared__ int       sGroupEnd[kSchedSample];
    __shared__ int       sLo[kSchedThreads];
    __shared__ int       sWave[kSchedThreads / 64];
    auto costKey = [&](int m) -> long long {
        return costOff ? static_cast<long long>(costOff[m]) : chunks[m].z + static_cast<long long>(chunkWeight) * m;
    };
    auto endKey = [&](int g) -> int { return groups[g].y; };
    const int cStride = max(kSchedCostStride, (nchunk + kSchedSample - 1) / kSchedSample);
    const int cn      = (nchunk + cStride - 1) / cStride;
    const int gStride = max(kSchedGroupStride, (ngroups + kSchedSample - 1) / kSchedSample);
    const int gn      = (ngroups + gStride - 1) / gStride;
    for (int i = threadIdx.x; i < cn; i += kSchedThreads)
    {
        sCost[i] = costKey(i * cStride);
    }
    for (int i = threadIdx.x; i < gn; i += kSchedThreads)
    {
        sGroupEnd[i] = endKey(i * gStride);
    }
    __syncthreads();
    const long long total  = costOff ? static_cast<long long>(*costTotal)
                                     : static_cast<long long>(nwork) + static_cast<long long>(chunkWeight) * nchunk;
    const long long wTotal = generationWeight(numWaves, numWaves, genWaves);
    const int       lane   = threadIdx.x & 63;
    const int       wave   = threadIdx.x >> 6;
    int             carry = 0, prevLo = -1;
    for (int base = 0; base < numWaves; base += kSchedThreads)
    {
        const int w  = base + threadIdx.x;
        int       lo = nchunk;
        if (w < numWaves)
        {
            const long long target = total * generationWeight(w, numWaves, genWaves) / wTotal;
            lo                     = sampledLowerBound(sCost, cn, cStride, nchunk, target, costKey);
        }
        sLo[threadIdx.x] = lo;
        __syncthreads();
        const int                prev = threadIdx.x > 0 ? sLo[threadIdx.x - 1] : prevLo;
        const bool               flag = w < numWaves && lo < nchunk && lo != prev;
        const unsigned long long bal  = __ballot(flag);
        if (lane == 0)
        {
            sWave[wave] = __popcll(bal);
        }
        __syncthreads();
        int before = 0, roundTotal = 0;
#pragma unroll
        for (int v = 0; v < kSchedThreads / 64; v++)
        {
            const int t = sWave[v];
            before += (v < wave) ? t : 0;
            roundTotal += t;
        }
        if (flag)
        {
            const int r = carry + before + __popcll(bal & ((1ull << lane) - 1ull));
            // the group holding chunk lo: the first whose end exceeds it
            out[r] = make_int2(lo, sampledLowerBound(sGroupEnd, gn, gStride, ngroups, lo + 1, endKey));
        }
        carry += roundTotal;
        prevLo = sLo[kSchedThreads - 1];
        __syncthreads();
    }
    for (int r = carry + threadIdx.x; r <= numWaves; r += blockDim.x)
    {
        out[r] = make_int2(nchunk, ngroups);
    }
}

__global__ void systemSlotKernel(int numSystems, const int* __restrict__ groupOff, int ngroups,
                                 const int4* __restrict__ groups, int numWaves, const int2* __restrict__ waveStart,
                                 int* __restrict__ slotOff)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > numSystems)
    {
        return;
    }
    const int ga = groupOff[k];
    if (ga >= ngroups)
    {
        slotOff[k] = ngroups + numWaves;
        return;
    }
    // the wavefront whose run holds the group's first chunk position: the
    // last one starting at or before it
    const int x  = groups[ga].x;
    int       lo = 0, hi = numWaves - 1;
    while (lo < hi)
    {
        const int m = (lo + hi + 1) / 2;
        if (waveStart[m].x <= x)
        {
            lo = m;
        }
        else
        {
            hi = m - 1;
        }
    }
    slotOff[k] = ga + lo;
}

} // namespace

void launchOccurrences(hipStream_t s, int ngroups, const int4* groups, const int4* groupInfo, const int* unionList,
                       int ncluster, int* count, int* occOff, int2* occ, bool fill)
{
    if (ngroups == 0)
    {
        return;
    }
    if (!fill)
    {
        occurrenceKernel<false><<<ngroups, 256, 0, s>>>(ngroups, groups, groupInfo, unionList, count, nullptr, nullptr,
                                                        ncluster);
        return;
    }
    occurrenceKernel<true><<<ngroups, 256, 0, s>>>(ngroups, groups, groupInfo, unionList, count, occOff, occ,
                                                   ncluster);
    // the per-cluster cursors in count are spent: count[0, ncluster) takes
    // the long lists, count[ncluster] their number
    if (ncluster > 0)
    {
        occurrenceSortKernel<<<(ncluster + 3) / 4, 256, 0, s>>>(ncluster, occOff, occ, count, count + ncluster);
        occurrenceSortLongKernel<<<kSortLongBlocks, 64, 0, s>>>(occOff, occ, count, count + ncluster);
    }
}

void launchSegments(hipStream_t s, int ngroups, const int4* groups, const int4* groupInfo, int numWaves,
                    const int2* waveStart, int slotF4, int* size, int* base, int* total, int4* segInfo)
{
    if (ngroups == 0)
    {
        return;
    }
    segmentKernel<<<(ngroups + 255) / 256, 256, 0, s>>>(ngroups, groups, groupInfo, numWaves, waveStart, slotF4, size,
                                                        segInfo);
    launchExclusiveScan(s, ngroups, size, base, total);
    segmentBaseKernel<<<(ngroups + 255) / 256, 256, 0, s>>>(ngroups, base, segInfo);
}

void launchGatherList(hipStream_t s, int ncluster, const int* occOff, const int2* occ, const int4* segInfo,
                      int slotF4, int* count, const int* gposOff, int* gpos, bool fill)
{
    if (ncluster == 0)
    {
        return;
    }
    if (fill)
    {
        gatherListKernel<true><<<(ncluster + 3) / 4, 256, 0, s>>>(ncluster, occOff, occ, segInfo, slotF4, count,
                                                                  gposOff, gpos);
    }
    else
    {
        gatherListKernel<false><<<(ncluster + 3) / 4, 256, 0, s>>>(ncluster, occOff, occ, segInfo, slotF4, count,
                                                                   gposOff, gpos);
    }
}

void launchChunkShifts(hipStream_t s, int nchunk, const int4* chunks, const float4* shiftVec, float4* out)
{
    if (nchunk > 0)
    {
        chunkShiftKernel<<<(nchunk + 255) / 256, 256, 0, s>>>(nchunk, chunks, shiftVec, out);
    }
}

void launchWaveSchedule(hipStream_t s, int numWaves, int nchunk, const int4* chunks, int nwork, int chunkWeight,
                        int ngroups, const int4* groups, int2* out, const int* costOff, const int* costTotal,
                        int genWaves)
{
    waveScheduleKernel<<<1, kSchedThreads, 0, s>>>(numWaves, nchunk, chunks, nwork, chunkWeight, ngroups, groups,
                                                   costOff, costTotal, genWaves, out);
}

void launchChunkCosts(hipStream_t s, int nchunk, const int4* chunks, const int* type, const int* typeLJ, int* cost)
{
    if (nchunk > 0)
    {
        chunkCostKernel<<<(nchunk + 255) / 256, 256, 0, s>>>(nchunk, chunks, type, typeLJ, cost);
    }
}

void launchSystemSlots(hipStream_t s, int numSystems, const int* groupOff, int ngroups, const int4* groups,
                       int numWaves, const int2* waveStart, int* slotOff)
{
    systemSlotKernel<<<(numSystems + 256) / 256, 256, 0, s>>>(numSystems, groupOff, ngroups, groups, numWaves,
                                                             waveStart, slotOff);
}

void launchGroupCompact(hipStream_t s, const GroupArgs& a, const int* unionTmp, const uint2* workTmp,
                        const int4* chunksTmp)
{
    if (a.ngroups > 0)
    {
        groupCompactKernel<<<a.ngroups, kGroupThreads, 0, s>>>(a, unionTmp, workTmp, chunksTmp);
    }
}

void launchBuildGroups(hipStream_t s, const GroupArgs& a, bool fill)
{
    if (a.ngroups == 0)
    {
        return;
    }
    if (fill)
    {
        buildGroupsKernel<true><<<a.ngroups, kGroupThreads, 0, s>>>(a);
    }
    else
    {
        buildGroupsKernel<false><<<a.ngroups, kGroupThreads, 0, s>>>(a);
    }
}

int launchPartialMax(hipStream_t s, const PartialMaxArgs& a)
{
    if (a.n <= 0)
    {
        return 0;
    }
    const int blocks = std::min(kPartialMaxBlocks, (a.n + kPartialMaxThreads - 1) / kPartialMaxThreads);
    partialMaxKernel<<<blocks, kPartialMaxThreads, 0, s>>>(a);
    return blocks;
}

void reducePartialMax(const int* partial, int blocks, int stride, int* out)
{
    std::fill(out, out + stride, 0);
    for (int b = 0; b < blocks; b++)
    {
        for (int j = 0; j < stride; j++)
        {
            out[j] = std::max(out[j], partial[b * stride + j]);
        }
    }
}

} // namespace nbx
