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

    const int g  = blockIdx.x;
    const int c0 = g * a.groupSize;
    const int c1 = min(c0 + a.groupSize, a.ncluster);
    const int jb = a.offJ[c0];
    const int n  = a.offJ[c1] - jb;
    const int ib = a.offI[c0];
    const int ni = a.offI[c1] - ib;
    int       p  = 1;
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
    // exclusive scan of the flags (ranks of the unique keys), and the chunk
    // and padded work offsets of the group's i-entries, all serial: a group
    // holds at most kGroupMaxJ j-entries and a few dozen i-entries
    if (threadIdx.x == 0)
    {
        int run = 0;
        for (int t = 0; t < p; t++)
        {
            const int f = sFlag[t];
            sFlag[t]    = run;
            run += f;
        }
        sUnion = run;
    }
    else if (threadIdx.x == 64)
    {
        int run = 0, work = 0;
        for (int e = 0; e < ni && e < kGroupThreads; e++)
        {
            const int4 ie = a.ientry[ib + e];
            sChunkBase[e] = run;
            sWorkBase[e]  = work;
            run += (ie.w - ie.z + a.chunkEntries - 1) / a.chunkEntries;
            work += (ie.w - ie.z + kTileJ - 1) / kTileJ * kTileJ;
        }
        sChunks = run;
        sWork   = work;
    }
    __syncthreads();
    const int U      = sUnion;
    // union clusters that get a window slot (the rest overflow)
    const int Uw     = kFill ? min(U, a.windowCap - 1 - a.groupSize) : U;
    const int nchunk = sChunks;
    if (!kFill)
    {
        if (threadIdx.x == 0)
        {
            a.unionCount[g] = U;
            a.chunkCount[g] = nchunk;
            a.workCount[g]  = sWork;
            a.overflow[g]   = (ni > kGroupThreads) ? 1 : 0;
        }
        return;
    }
    // compact the unique keys in place at the front of sKeys (rank order)
    const int uoff = a.unionOff[g];
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
    __syncthreads();
    const int coff = a.chunkOff[g];
    if (threadIdx.x == 0)
    {
        a.groups[g]    = make_int4(coff, coff + nchunk, uoff, Uw);
        a.groupInfo[g] = make_int4(c0, (c1 - c0) | (a.groupSize << 16), 0, g);
    }
    // per i-entry: masked-first partition, slots, work list, chunks (one
    // wave per entry)
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int woff = a.workOff[g];
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
                        int l = 0, r = U;
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

// max over groups of (offJ[c0 + G] - offJ[c0]) for the candidate sizes
__global__ void groupSpanKernel(int ncluster, const int* __restrict__ offJ, int* __restrict__ maxSpan)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncluster)
    {
        return;
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
        const int G = kGroupSizeMax >> k;
        if (c % G == 0)
        {
            const int e = min(c + G, ncluster);
            atomicMax(&maxSpan[k], offJ[e] - offJ[c]);
        }
    }
}

__global__ void maxReduceKernel(int n, const int* __restrict__ v, int* __restrict__ out)
{
    int m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        m = max(m, v[i]);
    }
    for (int o = 1; o < 64; o *= 2)
    {
        m = max(m, __shfl_xor(m, o, 64));
    }
    if ((threadIdx.x & 63) == 0)
    {
        atomicMax(out, m);
    }
}

__global__ void waveScheduleKernel(int numWaves, int nchunk, const int4* __restrict__ chunks, int nwork,
                                   int chunkWeight, int ngroups, const int4* __restrict__ groups, int2* __restrict__ out)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w > numWaves)
    {
        return;
    }
    // cost before chunk k: its work offset plus chunkWeight per earlier chunk
    const long long total  = static_cast<long long>(nwork) + static_cast<long long>(chunkWeight) * nchunk;
    const long long target = total * w / numWaves;
    int             lo = 0, hi = nchunk;
    while (lo < hi)
    {
        const int m = (lo + hi) / 2;
        if (chunks[m].z + static_cast<long long>(chunkWeight) * m < target)
        {
            lo = m + 1;
        }
        else
        {
            hi = m;
        }
    }
    if (w == numWaves)
    {
        lo = nchunk;
    }
    // the group holding chunk lo: the first whose chunk range ends after it
    int glo = 0, ghi = ngroups;
    while (glo < ghi)
    {
        const int m = (glo + ghi) / 2;
        if (groups[m].y <= lo)
        {
            glo = m + 1;
        }
        else
        {
            ghi = m;
        }
    }
    out[w] = make_int2(lo, glo);
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

void launchWaveSchedule(hipStream_t s, int numWaves, int nchunk, const int4* chunks, int nwork, int chunkWeight,
                        int ngroups, const int4* groups, int2* out)
{
    waveScheduleKernel<<<(numWaves + 256) / 256, 256, 0, s>>>(numWaves, nchunk, chunks, nwork, chunkWeight, ngroups,
                                                             groups, out);
}

void launchSystemSlots(hipStream_t s, int numSystems, const int* groupOff, int ngroups, const int4* groups,
                       int numWaves, const int2* waveStart, int* slotOff)
{
    systemSlotKernel<<<(numSystems + 256) / 256, 256, 0, s>>>(numSystems, groupOff, ngroups, groups, numWaves,
                                                             waveStart, slotOff);
}

void launchGroupSpans(hipStream_t s, int ncluster, const int* offJ, int* maxSpan)
{
    if (ncluster > 0)
    {
        groupSpanKernel<<<(ncluster + 255) / 256, 256, 0, s>>>(ncluster, offJ, maxSpan);
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

void launchMaxReduce(hipStream_t s, int n, const int* v, int* out)
{
    if (n > 0)
    {
        maxReduceKernel<<<std::min(1024, (n + 255) / 256), 256, 0, s>>>(n, v, out);
    }
}

} // namespace nbx
